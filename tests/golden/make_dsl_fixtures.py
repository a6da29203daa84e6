"""Write the DSL fixtures: the reference's own example programs run through
the front end pairs_amd.dsl (installed as ``pairs``), and the host C++
programs its HIP target emits for them.

This script runs ONLY in the development container: it reads the example
scripts and data files of the reference tree.  What it writes under
tests/golden/ is what travels to the GPU host:

  dsl_<name>.json  the recognised modules, the lowering plan and the lowered
                   configuration (cutoff, neighbour radius, dt, capacities,
                   schedule, domain) of the specialised script;
  dsl_<name>.cpp   the emitted host program (include/pairs_hip*.h); the
                   package build compiles it against libpairs_hip.so, and the
                   GPU tests run it and compare its printed thermo lines and
                   its final state with the reference program's own output
                   (tests/golden/md_fcc4_*.npz, dem_sc168.npz) or the oracle.

Specialisations (string substitutions, as make_golden.py does, so that the
programs match the golden runs):
  md_fcc4_full  examples/md.py with a 4x4x4 lattice, one type, 100 steps;
  md_fcc4_half  the same with its compute_half() line enabled;
  dem_sc168     examples/dem.py on a 0.04 x 0.015 x 0.04 m domain, 500 steps;
  lj_onetype    examples/lj_onetype.py reading the 4x4x4 particle file of the
                reference's data/, with a set_domain() call of 4 lattice
                constants added (the older API leaves the box implicit).

Usage:  python tests/golden/make_dsl_fixtures.py
"""
import json
import os
import re
import shutil
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from make_golden import specialise_dem, specialise_md  # noqa: E402

from pairs_amd import dsl  # noqa: E402

LJ_ONETYPE_BOX = 4 * 1.6796  # 4 lattice constants of the 4x4x4 file (rho 0.8442)


def specialise_lj_onetype(src):
    src, n = re.subn(r'"data/minimd_setup_32x32x32\.input"', '"data/minimd_setup_4x4x4_onetype.input"', src)
    assert n == 1, "particle file of lj_onetype.py not found"
    L = repr(LJ_ONETYPE_BOX)
    src, n = re.subn(r'^(psim = pairs\.simulation\("lj", debug=True\))$',
                     r"\1" + f"\npsim.set_domain([0.0, 0.0, 0.0, {L}, {L}, {L}])", src, flags=re.M)
    assert n == 1, "simulation() call of lj_onetype.py not found"
    return src


def describe(sim, kind, script, subs):
    mods = [dict(name=dsl.recognise(m["fn"], m["symbols"]), cutoff=m["cutoff"], pre_step=m["pre_step"],
                 skip_first=m["skip_first"], symbols={k: float(v) for k, v in m["symbols"].items()})
            for m in sim.modules]
    return dict(script=script, specialised=subs, plan=kind, modules=mods,
                setup_modules=[dsl.recognise(f, s) for f, s in sim.setup_modules],
                timesteps=sim.timesteps, half=bool(sim.half), reneighbor_every=sim.reneighbor,
                thermo_every=sim.thermo_every, neighbor_radius=sim.neighbor_radius, cell_spacing=sim.cell_spacing,
                particle_capacity=sim.particle_capacity, neighbor_capacity=sim.neighbor_capacity,
                domain=sim.domain, pbc=[bool(x) for x in sim.pbc_flags])


def emit(name, script_rel, text, subs, data_files=()):
    work = tempfile.mkdtemp(prefix="pairs_dsl_")
    cwd = os.getcwd()
    try:
        os.makedirs(os.path.join(work, "data"))
        os.makedirs(os.path.join(work, "output"))
        for f in data_files:
            shutil.copy(os.path.join(REF, "data", f), os.path.join(work, "data", f))
        path = os.path.join(work, os.path.basename(script_rel))
        with open(path, "w") as f:
            f.write(text)
        out_cpp = os.path.join(HERE, f"dsl_{name}.cpp")
        os.chdir(work)
        sims = dsl.run_script(path, ["gpu"], run=False, emit_cpp=out_cpp, quiet=True)
        kind, _ = sims[0]._plan()
        meta = describe(sims[0], kind, script_rel, subs)
    finally:
        os.chdir(cwd)
        shutil.rmtree(work, ignore_errors=True)
    with open(os.path.join(HERE, f"dsl_{name}.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote dsl_{name}.cpp / .json ({meta['plan']}: {[m['name'] for m in meta['modules']]})")


def main():
    md = open(os.path.join(REF, "examples/md.py")).read()
    emit("md_fcc4_full", "examples/md.py", specialise_md(md, 4, 1, 100, False),
         dict(nx=4, ntypes=1, timesteps=100))
    emit("md_fcc4_half", "examples/md.py", specialise_md(md, 4, 1, 100, True),
         dict(nx=4, ntypes=1, timesteps=100, compute_half=True))
    dem = open(os.path.join(REF, "examples/dem.py")).read()
    emit("dem_sc168", "examples/dem.py", specialise_dem(dem, 500),
         dict(domain=[0.04, 0.015, 0.04], timesteps=500), data_files=("planes.input",))
    old = open(os.path.join(REF, "examples/lj_onetype.py")).read()
    emit("lj_onetype", "examples/lj_onetype.py", specialise_lj_onetype(old),
         dict(particle_file="data/minimd_setup_4x4x4_onetype.input", box=LJ_ONETYPE_BOX),
         data_files=("minimd_setup_4x4x4_onetype.input",))


if __name__ == "__main__":
    main()
