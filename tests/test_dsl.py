"""The DSL front end (pairs_amd/dsl.py): module recognition by traced
expression graphs, lowering to the C ABI, and the emitted host C++ program
(on the host; the GPU runs are in tests/test_gpu_dsl.py)."""
import json
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


def _fixture(name):
    with open(os.path.join(HERE, "golden", f"dsl_{name}.json")) as f:
        return json.load(f)


def test_example_scripts_were_recognised():
    """The reference's example scripts, run through the front end by
    tests/golden/make_dsl_fixtures.py: their modules were recognised as the
    kernels' modules and lowered with the scripts' own configuration."""
    md = _fixture("md_fcc4_full")
    assert md["script"] == "examples/md.py" and md["plan"] == "md"
    assert [m["name"] for m in md["modules"]] == ["initial_integrate", "lennard_jones", "final_integrate"]
    ii, lj, fi = md["modules"]
    assert ii["pre_step"] and ii["skip_first"] and fi["skip_first"] and not fi["pre_step"]
    assert lj["cutoff"] == 2.5 and md["neighbor_radius"] == 2.8 and ii["symbols"]["dt"] == 0.005
    assert md["reneighbor_every"] == 20 and md["thermo_every"] == 100 and not md["half"]
    assert _fixture("md_fcc4_half")["half"]
    dem = _fixture("dem_sc168")
    assert dem["script"] == "examples/dem.py" and dem["plan"] == "dem"
    assert [m["name"] for m in dem["modules"]] == ["gravity", "linear_spring_dashpot", "euler"]
    assert dem["setup_modules"] == ["update_mass_and_inertia"]
    assert dem["pbc"] == [True, True, False] and dem["timesteps"] == 500
    old = _fixture("lj_onetype")
    assert old["script"] == "examples/lj_onetype.py" and old["plan"] == "md_euler"
    assert [m["name"] for m in old["modules"]] == ["lj_older", "euler_md"]
    assert old["reneighbor_every"] == 1 and old["neighbor_radius"] == 2.8


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
    cpu = dsl.simulation("z", [dsl.point_mass()], timesteps=1, double_prec=True)
    cpu.target(dsl.target_cpu())
    with pytest.raises(dsl.UnsupportedModule, match="GPU target only"):
        cpu.generate(run=False)


@pytest.mark.parametrize("name", ["md_fcc4_full", "md_fcc4_half", "dem_sc168", "lj_onetype"])
def test_emitted_host_program_compiles(name):
    """The committed emitted programs compile against the C ABI headers (the
    GPU tests build, run and check them)."""
    src = os.path.join(HERE, "golden", f"dsl_{name}.cpp")
    text = open(src).read()
    assert "int main(int argc, char **argv)" in text
    assert any(call in text for call in ("pairs_hip_run(", "pairs_hip_dem_run(", "pairs_hip_lj_older("))
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host compiler")
    r = subprocess.run([cxx, "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


# the older API's modules: bare rsq / delta in the pair module, whose
# expression tree must be the older kernel's (examples/lj_onetype.py:5-13)
OLDER_PAIR = """
def pair(i, j):
    inv = 1.0 / rsq
    s6 = inv * inv * inv * sigma6
    force[i] += delta * 48.0 * s6 * (s6 - 0.5) * inv * epsilon
"""
OLDER_STEP = """
def step(i):
    velocity[i] += dt * force[i] / mass[i]
    position[i] += dt * velocity[i]
"""


def test_older_api_aliases(tmp_path):
    """The older calls configure a simulation (properties, file input, cells,
    neighbour lists, periodicity); its modules are traced with rsq / delta as
    bare names and recognised as the older LJ program (its own pair
    expression, Euler), which needs the box the older API leaves implicit."""
    data = os.path.join(HERE, "golden", "minimd_setup_4x4x4_onetype.input")
    sim = dsl.simulation("older_api")
    sim.add_real_property("mass", 1.0)
    sim.add_position("position")
    sim.add_vector_property("velocity")
    sim.add_vector_property("force", vol=True)
    sim.from_file(data, ["mass", "position", "velocity"])
    sim.create_cell_lists(2.8, 2.5)
    sim.create_neighbor_lists()
    sim.periodic(2.8)
    sim.compute(_module(tmp_path, OLDER_PAIR, "pair"), 2.5, {"sigma6": 1.0, "epsilon": 1.0})
    sim.compute(_module(tmp_path, OLDER_STEP, "step"), symbols={"dt": 0.005})
    sim.target(dsl.target_gpu())
    with pytest.raises(dsl.UnsupportedModule, match="set_domain"):
        sim.generate(run=False)
    assert [dsl.recognise(m["fn"], m["symbols"]) for m in sim.modules] == ["lj_older", "euler_md"]
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
    # another rounding order of the same pair force is not the older kernel
    with pytest.raises(dsl.UnsupportedModule):
        dsl.recognise(_module(tmp_path, OLDER_PAIR.replace("delta * 48.0", "48.0 * delta"), "pair"),
                      {"sigma6": 1.0, "epsilon": 1.0})
    with pytest.raises(dsl.UnsupportedModule, match="loop-style kernels"):
        sim.particle_pairs(2.5, None)
