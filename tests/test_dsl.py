"""The DSL front end (pairs_amd/dsl.py): module recognition by traced
expression graphs, lowering to the C ABI, and the emitted host C++ program
(on the host; the GPU runs are in tests/test_gpu_dsl.py)."""
import os
import shutil
import subprocess
import textwrap

import pytest

from pairs_amd import dsl

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _module(tmp_path, src, name):
    p = tmp_path / f"m_{name}.py"
    p.write_text(textwrap.dedent(src))
    import importlib.util
    spec = importlib.util.spec_from_file_location(f"m_{name}", str(p))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return getattr(m, name)


LJ = """
def lennard_jones(i, j):
    sr2 = 1.0 / squared_distance(i, j)
    sr6 = sr2 * sr2 * sr2 * sigma6[i, j]
    apply(force, delta(i, j) * (48.0 * sr6 * (sr6 - 0.5) * sr2 * epsilon[i, j]))
"""


def test_recognises_the_example_modules():
    sims = dsl.run_script(os.path.join(HERE, "dsl", "lj_fcc.py"), ["gpu"], run=False)
    assert [dsl.recognise(m["fn"], m["symbols"]) for m in sims[0].modules] == \
        ["initial_integrate", "lennard_jones", "final_integrate"]
    sims = dsl.run_script(os.path.join(HERE, "dsl", "dem_settle.py"), ["gpu"], run=False)
    s = sims[0]
    assert [dsl.recognise(m["fn"], m["symbols"]) for m in s.modules] == ["gravity", "linear_spring_dashpot",
                                                                           "euler"]
    assert [dsl.recognise(f, sy) for f, sy in s.setup_modules] == ["update_mass_and_inertia"]


def test_operation_order_is_part_of_the_module(tmp_path):
    assert dsl.recognise(_module(tmp_path, LJ, "lennard_jones")) == "lennard_jones"
    # the same physics with another rounding order: not the kernel's module
    swapped = LJ.replace("48.0 * sr6 * (sr6 - 0.5) * sr2", "48.0 * sr6 * sr2 * (sr6 - 0.5)")
    with pytest.raises(dsl.UnsupportedModule):
        dsl.recognise(_module(tmp_path, swapped, "lennard_jones"))
    # a dropped factor
    with pytest.raises(dsl.UnsupportedModule):
        dsl.recognise(_module(tmp_path, LJ.replace(" * epsilon[i, j]", ""), "lennard_jones"))
    # an unknown keyword or property
    with pytest.raises(dsl.UnsupportedModule):
        dsl.recognise(_module(tmp_path, LJ.replace("sigma6[i, j]", "sigma12[i, j]"), "lennard_jones"))


def test_rejects_unsupported_configurations(tmp_path):
    sim = dsl.simulation("x", [dsl.point_mass()], timesteps=1, double_prec=True)
    sim.compute(_module(tmp_path, LJ, "lennard_jones"), 2.5)
    with pytest.raises(dsl.UnsupportedModule):
        sim.generate(run=False)  # no integrators: not the md.py program
    with pytest.raises(dsl.UnsupportedModule):
        dsl.simulation("y", [dsl.point_mass()], double_prec=False)
    with pytest.raises(dsl.UnsupportedModule):  # target_cpu(): this build has the GPU target only
        dsl.run_script(os.path.join(HERE, "dsl", "lj_fcc.py"), ["cpu"], run=False)


@pytest.mark.parametrize("script", ["lj_fcc.py", "dem_settle.py"])
def test_emitted_host_program_compiles(tmp_path, script):
    out = str(tmp_path / (script[:-3] + ".cpp"))
    dsl.run_script(os.path.join(HERE, "dsl", script), ["gpu"], run=False, emit_cpp=out)
    text = open(out).read()
    assert "int main()" in text and ("pairs_hip_run(" in text or "pairs_hip_dem_run(" in text)
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host compiler")
    r = subprocess.run([cxx, "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), out],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
