"""The DSL's HIP target end to end on the GPU: the host programs that
pairs_amd.dsl emitted for the reference's own example scripts
(tests/golden/dsl_*.cpp, written by tests/golden/make_dsl_fixtures.py from
examples/md.py, examples/dem.py and examples/lj_onetype.py), built against
libpairs_hip.so by the package build, run here; their printed thermo lines and
final states are compared with the reference program's own output (the golden
fixtures) or with the oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
BIN = os.path.join(ROOT, "build", "dsl")


def _program(name):
    exe = os.path.join(BIN, "dsl_" + name)
    if not os.path.exists(exe):  # not built by the package build: build it now
        sys.path.insert(0, ROOT)
        from pairs_amd import build
        build.build_host_program(os.path.join(GOLDEN, f"dsl_{name}.cpp"), exe)
    return exe


def _run(name, tmp_path, *args):
    out = str(tmp_path / f"{name}.bin")
    print(f"running the emitted program dsl_{name}", flush=True)
    r = subprocess.run([_program(name), out, *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout, open(out, "rb").read()


def _unpack(raw, doubles_per_row, ints_per_row=0):
    n = int(np.frombuffer(raw, np.int32, 1)[0])
    d = np.frombuffer(raw, np.float64, doubles_per_row * n, offset=4).reshape(-1, n, 3)
    i = np.frombuffer(raw, np.int32, ints_per_row * n, offset=4 + 8 * doubles_per_row * n) if ints_per_row else None
    return n, d, i


@pytest.mark.parametrize("name,golden", [("md_fcc4_full", "md_fcc4_full"), ("md_fcc4_half", "md_fcc4_half")])
def test_emitted_md_program_equals_reference_run(tmp_path, name, golden):
    """examples/md.py (4x4x4, 100 steps; full or half lists): the emitted
    program's printed thermo lines are the reference program's printed lines
    (6 significant digits), and its final positions and velocities (after
    iteration 100) are the reference's, bit for bit."""
    z = np.load(os.path.join(GOLDEN, golden + ".npz"))
    stdout, raw = _run(name, tmp_path)
    rows = np.array([[float(x) for x in line.split("\t")] for line in stdout.strip().splitlines()])
    ref = z["thermo_printed"]
    assert rows.shape == ref.shape == (2, 2)
    assert np.all(np.abs(rows - ref) <= 5e-6 * np.abs(ref))
    n, d, _ = _unpack(raw, 6)
    assert n == 256
    assert np.array_equal(d[0], z["s100_position"])
    assert np.array_equal(d[1], z["s100_velocity"])


def test_emitted_dem_program_equals_reference_run(tmp_path):
    """examples/dem.py (0.04 x 0.015 x 0.04 m, 500 steps, the reference's
    planes inlined by the emitter): particle order identical to the
    reference program's, positions and velocities within 1e-10 (the
    quaternion update's sin / cos may differ in the last bit)."""
    z = np.load(os.path.join(GOLDEN, "dem_sc168.npz"))
    _, raw = _run("dem_sc168", tmp_path)
    n, d, uid = _unpack(raw, 6, 1)
    assert n == len(z["s500_uid"])
    assert np.array_equal(uid, z["s500_uid"])
    for k, name in enumerate(("position", "linear_velocity")):
        ref = z[f"s500_{name}"]
        assert np.abs(d[k] - ref).max() <= 1e-10 * np.abs(ref).max(), name


def test_emitted_older_api_program_equals_oracle(tmp_path):
    """examples/lj_onetype.py (the older API: from_file, its own pair
    expression, Euler, rebuilt every step) with the reference's 4x4x4
    particle file and a box of 4 lattice constants: 101 iterations,
    bit-identical to the oracle's restatement."""
    data = os.path.join(GOLDEN, "minimd_setup_4x4x4_onetype.input")
    _, raw = _run("lj_onetype", tmp_path, data)
    n, d, _ = _unpack(raw, 9)
    L = 4 * 1.6796
    rows = np.loadtxt(data, delimiter=",")
    o = po.OracleMD([(0.0, L)] * 3, 4096, ncap=100, reneighbor_every=1, thermo_every=0)
    o.set_particles(rows[:, 1:4], rows[:, 4:7], m=rows[:, 0])
    o.setup_cells()
    for _ in range(101):
        o.rebuild()
        o.lj_older()
        o.euler()
    assert n == o.nlocal == 256
    assert np.array_equal(d[0], o.positions())
    assert np.array_equal(d[1], o.velocities())
    assert np.array_equal(d[2], o.forces())
