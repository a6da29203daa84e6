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
    with pytest.raises(dsl.UnsupportedModule):  # target_cpu(): pairs_amd has the GPU target only
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


def test_older_api_aliases():
    """The older calls configure a simulation (properties, file input, cells,
    neighbour lists, periodicity); its modules are traced with rsq / delta as
    bare names and recognised as the older LJ program (its own pair
    expression, Euler), which needs the box the older API leaves implicit."""
    script = os.path.join(HERE, "dsl", "older_api.py")
    data = os.path.join(HERE, "golden", "minimd_setup_4x4x4_onetype.input")
    with pytest.raises(dsl.UnsupportedModule, match="set_domain"):
        dsl.run_script(script, [data], run=False)
    assert [dsl.recognise(m["fn"], m["symbols"]) for m in dsl._CREATED[-1].modules] == ["lj_older", "euler_md"]
    sim = dsl._CREATED[-1]
    assert set(sim.props) == {"mass", "position", "velocity", "force"}
    assert sim.props["force"][2] is True and sim.position_name == "position"
    assert sim.setups == [("read", data, ["mass", "position", "velocity"], dsl.point_mass())]
    assert sim.cell_spacing == 2.8 and sim.neighbor_radius == 2.8 and sim.pbc_flags == [True, True, True]
    # the pair module traced: 1/rsq, the pair's separation, an update of force[i]
    form = dsl.trace(sim.modules[0]["fn"], ("force", "mass", "position", "velocity"), ("sigma6", "epsilon"))
    ids = dict(line.split("=", 1) for line in form.splitlines() if "=" in line and not line.startswith("set"))
    rsq = next(k for k, v in ids.items() if v == "squared_distance(i,j)")
    dlt = next(k for k, v in ids.items() if v == "delta(i,j)")
    assert f"div(const:1.0,{rsq})" in ids.values()
    assert f"mul({dlt},const:48.0)" in ids.values()  # the older order: delta * 48.0 first
    assert form.splitlines()[-1].startswith("set[](name:force,i,")
    with pytest.raises(dsl.UnsupportedModule, match="loop-style kernels"):
        sim.particle_pairs(2.5, None)
