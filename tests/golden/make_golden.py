"""Write the golden fixtures of the LJ molecular-dynamics path from the
reference's own generated CPU program.

This script runs ONLY in the development container (it needs the reference
tree, its Python code generator and MPICH's ``mpic++``); the fixtures it writes
(``tests/golden/*.npz``) are what travels.  It is the one place where the
reference is executed: the combined trajectory semantics of the generated
program -- stale multi-hop ghost refresh (``src/pairs/sim/comm.py:45-54,369-442``),
the rebuild schedule (``src/pairs/sim/timestep.py:13,46-56``), the exchange
reorder (``src/pairs/sim/comm.py:445-520``) and the atomic-insertion order of
cells and ghosts on the CPU -- cannot be pinned by closed forms alone.

Steps:
  1. read ``examples/md.py`` from the reference, specialise its configuration
     by string substitution (lattice 4x4x4, one type, 100 steps, CPU target,
     optionally ``compute_half``) and run it with the reference's ``pairs``
     package on ``sys.path``; it writes ``md.cpp`` into a scratch directory;
  2. splice ``golden_dump.hpp`` calls into the generated ``main`` (after the
     initial lattice, and at the end of every timestep);
  3. build with ``g++ -O3 -ffp-contract=off`` and MPICH against the reference runtime
     (``runtime/pairs.cpp``, ``runtime/domain/regular_6d_stencil.cpp``,
     ``runtime/devices/dummy.cpp``), run it, and pack selected steps into npz.

Usage:  python tests/golden/make_golden.py [--keep DIR]
"""
import argparse
import glob
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
MPI_INC = "/opt/conda/include"  # MPICH 3.3.2
MPI_LIB = "/opt/conda/lib"
CXXFLAGS = ["-O3", "-ffp-contract=off", "-std=c++17", "-w"]

# steps whose full local state (x, v, f) is kept
STATE_STEPS = [0, 1, 2, 18, 19, 20, 21, 38, 39, 40, 59, 60, 79, 80, 98, 99, 100]
# steps whose ghost positions are kept (rebuild and stale-refresh steps)
GHOST_STEPS = [0, 1, 2, 19, 20, 21, 100]
# rebuild steps whose neighbour lists / cells / send maps are kept
LIST_STEPS = [0, 19, 99]


def specialise_md(src, nx, ntypes, steps, half):
    subs = [
        (r"^nx = \d+$", f"nx = {nx}"),
        (r"^ny = \d+$", f"ny = {nx}"),
        (r"^nz = \d+$", f"nz = {nx}"),
        (r"^ntypes = \d+$", f"ntypes = {ntypes}"),
        (r"timesteps=\d+", f"timesteps={steps}"),
    ]
    for pat, rep in subs:
        src, n = re.subn(pat, rep, src, flags=re.M)
        assert n == 1, f"pattern not found in md.py: {pat}"
    if half:
        src, n = re.subn(r"^#psim\.compute_half\(\)$", "psim.compute_half()", src, flags=re.M)
        assert n == 1
    return src


def generate(workdir, script_text):
    path = os.path.join(workdir, "md_golden.py")
    with open(path, "w") as f:
        f.write(script_text)
    env = dict(os.environ)
    env["PYTHONPATH"] = os.path.join(REF, "src")
    subprocess.run([sys.executable, path, "cpu"], cwd=workdir, env=env, check=True,
                   stdout=subprocess.DEVNULL)
    return os.path.join(workdir, "md.cpp")


def instrument(cpp_path, dump_dir, steps):
    src = open(cpp_path).read()
    src = src.replace('#include "runtime/vtk.hpp"',
                      '#include "runtime/vtk.hpp"\n#include "golden_dump.hpp"', 1)
    main_at = src.index("int main(int argc, char **argv) {")
    m = re.compile(r"for\(int (i\d+) = 0; \1 < %d; \1\+\+\) \{" % (steps + 1)).search(src, main_at)
    assert m, "timestep loop not found"
    ts_var = m.group(1)
    args = ("nlocal, nghost, nsend_all, ncells, neighbor_capacity, cell_capacity, "
            "position, linear_velocity, force, type, numneighs, neighborlists, particle_cell, "
            "send_map, send_mult, nsend, cell_sizes, cell_particles")
    # initial state (ts = -1), right before the timestep loop
    init_call = (f'golden::dump_md("{dump_dir}", -1, nlocal, 0, 0, 0, neighbor_capacity, cell_capacity, '
                 "position, linear_velocity, force, type, nullptr, nullptr, particle_cell, "
                 "send_map, send_mult, nsend, cell_sizes, cell_particles);\n    ")
    src = src[:m.start()] + init_call + src[m.start():]
    # end of each iteration: the loop's closing brace precedes LIKWID_MARKER_CLOSE
    end = src.index("    LIKWID_MARKER_CLOSE;")
    close = src.rindex("    }\n", 0, end)
    step_call = f'        golden::dump_md("{dump_dir}", {ts_var}, {args});\n'
    src = src[:close] + step_call + src[close:]
    with open(cpp_path, "w") as f:
        f.write(src)


def build_and_run(workdir):
    """Compile the generated program with g++ against MPICH's C library.

    The conda ``mpic++`` wrapper names a cross compiler this image lacks and its
    ``-L`` puts conda's older libstdc++ first, so the MPI library and the two
    Fortran runtime libraries it needs are linked through a private directory of
    symlinks instead.
    """
    exe = os.path.join(workdir, "md_ref")
    libdir = os.path.join(workdir, "mpilib")
    os.makedirs(libdir, exist_ok=True)
    for lib in ("libmpi.so.12", "libgfortran.so.4", "libquadmath.so.0"):
        dst = os.path.join(libdir, lib)
        if not os.path.exists(dst):
            os.symlink(os.path.join(MPI_LIB, lib), dst)
    cmd = ["g++"] + CXXFLAGS + ["-I" + MPI_INC, "-I" + REF, "-I" + HERE, "-o", exe,
                                os.path.join(workdir, "md.cpp"),
                                os.path.join(REF, "runtime/pairs.cpp"),
                                os.path.join(REF, "runtime/domain/regular_6d_stencil.cpp"),
                                os.path.join(REF, "runtime/devices/dummy.cpp"),
                                os.path.join(libdir, "libmpi.so.12"),
                                "-Wl,-rpath," + libdir, "-Wl,-rpath-link," + libdir]
    subprocess.run(cmd, check=True, cwd=workdir)
    out = subprocess.run([exe], cwd=workdir, check=True, capture_output=True, text=True)
    return out.stdout


def read_dump(path):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw, dtype=np.int32, count=12)
    assert hdr[0] == 0x50414952
    _, ts, nlocal, nghost, nsend_all, ncells, ncap, ccap, has_lists = [int(x) for x in hdr[:9]]
    off = 48
    nall = nlocal + nghost

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=off).copy()
        off += a.nbytes
        return a

    d = dict(ts=ts, nlocal=nlocal, nghost=nghost, nsend_all=nsend_all, ncells=ncells)
    d["position"] = take(np.float64, nall * 3).reshape(nall, 3)
    d["velocity"] = take(np.float64, nall * 3).reshape(nall, 3)
    d["force"] = take(np.float64, nlocal * 3).reshape(nlocal, 3)
    d["type"] = take(np.int32, nall)
    if has_lists:
        d["numneighs"] = take(np.int32, nlocal)
        d["neighborlists"] = take(np.int32, nlocal * ncap).reshape(nlocal, ncap)
    d["particle_cell"] = take(np.int32, nall)
    d["send_map"] = take(np.int32, nsend_all)
    d["send_mult"] = take(np.int32, nsend_all * 3).reshape(nsend_all, 3)
    d["nsend"] = take(np.int32, 6)
    d["cell_sizes"] = take(np.int32, ncells)
    d["cell_particles"] = take(np.int32, ncells * ccap).reshape(ncells, ccap)
    return d


def pack(dump_dir, stdout, out_path, meta):
    dumps = {}
    for p in sorted(glob.glob(os.path.join(dump_dir, "dump_*.bin"))):
        d = read_dump(p)
        dumps[d["ts"]] = d
    arrays = {}
    init = dumps[-1]
    n0 = init["nlocal"]
    arrays["init_position"] = init["position"][:n0]
    arrays["init_velocity"] = init["velocity"][:n0]
    arrays["init_type"] = init["type"][:n0]
    nsteps = max(dumps)
    arrays["nlocal_all"] = np.array([dumps[t]["nlocal"] for t in range(nsteps + 1)], np.int32)
    arrays["nghost_all"] = np.array([dumps[t]["nghost"] for t in range(nsteps + 1)], np.int32)
    for t in STATE_STEPS:
        d = dumps[t]
        n = d["nlocal"]
        arrays[f"s{t}_position"] = d["position"][:n]
        arrays[f"s{t}_velocity"] = d["velocity"][:n]
        arrays[f"s{t}_force"] = d["force"]
    for t in GHOST_STEPS:
        d = dumps[t]
        n = d["nlocal"]
        arrays[f"s{t}_ghost_position"] = d["position"][n:]
    for t in LIST_STEPS:
        d = dumps[t]
        nn = d["numneighs"]
        arrays[f"s{t}_numneighs"] = nn
        arrays[f"s{t}_neighbors"] = np.concatenate(
            [d["neighborlists"][i, :nn[i]] for i in range(d["nlocal"])]).astype(np.int32)
        arrays[f"s{t}_particle_cell"] = d["particle_cell"]
        arrays[f"s{t}_send_map"] = d["send_map"]
        arrays[f"s{t}_send_mult"] = d["send_mult"].astype(np.int8)
        arrays[f"s{t}_nsend"] = d["nsend"]
        cs = d["cell_sizes"]
        arrays[f"s{t}_cell_sizes"] = cs
        arrays[f"s{t}_cell_particles"] = np.concatenate(
            [d["cell_particles"][c, :cs[c]] for c in range(d["ncells"])]).astype(np.int32)
    thermo = [[float(x) for x in line.split("\t")] for line in stdout.splitlines()
              if re.fullmatch(r"[-+0-9.eE]+\t[-+0-9.eE]+", line.strip())]
    arrays["thermo_printed"] = np.array(thermo, np.float64)
    for k, v in meta.items():
        arrays["meta_" + k] = np.array(v)
    np.savez_compressed(out_path, **arrays)
    return arrays


def make_md(name, nx, ntypes, steps, half, keep=None):
    src = open(os.path.join(REF, "examples/md.py")).read()
    script = specialise_md(src, nx, ntypes, steps, half)
    workdir = keep or tempfile.mkdtemp(prefix="pairs_golden_")
    os.makedirs(workdir, exist_ok=True)
    dump_dir = os.path.join(workdir, "dumps_" + name)
    shutil.rmtree(dump_dir, ignore_errors=True)
    os.makedirs(dump_dir)
    cpp = generate(workdir, script)
    instrument(cpp, dump_dir, steps)
    stdout = build_and_run(workdir)
    meta = dict(nx=nx, ntypes=ntypes, steps=steps, half=int(half), rho=0.8442, temp=1.44,
                cutoff=2.5, skin=0.3, dt=0.005, reneighbor=20, thermo_every=100,
                compiler="g++ 11.4 + MPICH 3.3.2 " + " ".join(CXXFLAGS))
    out = os.path.join(HERE, f"{name}.npz")
    pack(dump_dir, stdout, out, meta)
    if keep is None:
        shutil.rmtree(workdir, ignore_errors=True)
    print(f"wrote {out} ({os.path.getsize(out)} bytes)")


# ---------------------------------------------------------------- DEM
DEM_STEPS = 500


def specialise_dem(src, steps):
    subs = [
        (r"^domainSize_SI = \[[^\]]*\]$", "domainSize_SI = [0.04, 0.015, 0.04]"),
        (r"^timeSteps = \d+$", f"timeSteps = {steps}"),
        (r"^visSpacing = \d+$", "visSpacing = 1000000"),
    ]
    for pat, rep in subs:
        src, n = re.subn(pat, rep, src, count=1, flags=re.M)
        assert n == 1, f"pattern not found in dem.py: {pat}"
    return src


def generate_dem(workdir, steps):
    src = open(os.path.join(REF, "examples/dem.py")).read()
    script = specialise_dem(src, steps)
    os.makedirs(os.path.join(workdir, "data"), exist_ok=True)
    os.makedirs(os.path.join(workdir, "output"), exist_ok=True)
    shutil.copy(os.path.join(REF, "data/planes.input"), os.path.join(workdir, "data/planes.input"))
    path = os.path.join(workdir, "dem_golden.py")
    with open(path, "w") as f:
        f.write(script)
    env = dict(os.environ)
    env["PYTHONPATH"] = os.path.join(REF, "src")
    subprocess.run([sys.executable, path, "cpu"], cwd=workdir, env=env, check=True, stdout=subprocess.DEVNULL)
    return os.path.join(workdir, "dem.cpp")


DEM_ARGS = ("nlocal, nghost, neighbor_capacity, position, linear_velocity, angular_velocity, force, torque, "
            "radius, mass, rotation_quat, rotation_matrix, uid, flags, type, num_contacts, contact_lists, "
            "is_sticking, tangential_spring_displacement, impact_velocity_magnitude")


def instrument_dem(cpp_path, dump_dir, steps):
    src = open(cpp_path).read()
    src = src.replace('#include "runtime/vtk.hpp"',
                      '#include "runtime/vtk.hpp"\n#include "golden_dump.hpp"', 1)
    main_at = src.index("int main(int argc, char **argv) {")
    m = re.compile(r"for\(int (i\d+) = 0; \1 < %d; \1\+\+\) \{" % (steps + 1)).search(src, main_at)
    assert m, "timestep loop not found"
    ts_var = m.group(1)
    init_call = f'golden::dump_dem("{dump_dir}", -1, {DEM_ARGS});\n    '
    src = src[:m.start()] + init_call + src[m.start():]
    end = src.index("    LIKWID_MARKER_CLOSE;")
    close = src.rindex("    }\n", 0, end)
    src = src[:close] + f'        golden::dump_dem("{dump_dir}", {ts_var}, {DEM_ARGS});\n' + src[close:]
    with open(cpp_path, "w") as f:
        f.write(src)


def read_dem_dump(path):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw, dtype=np.int32, count=8)
    assert hdr[0] == 0x4D454400
    ts, nlocal, nghost, cap = [int(x) for x in hdr[1:5]]
    off = 32
    nall = nlocal + nghost

    def take(dtype, n, shape=None):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=off).copy()
        off += a.nbytes
        return a.reshape(shape) if shape else a

    d = dict(ts=ts, nlocal=nlocal, nghost=nghost, cap=cap)
    d["position"] = take(np.float64, nall * 3, (nall, 3))
    d["linear_velocity"] = take(np.float64, nall * 3, (nall, 3))
    d["angular_velocity"] = take(np.float64, nall * 3, (nall, 3))
    d["force"] = take(np.float64, nlocal * 3, (nlocal, 3))
    d["torque"] = take(np.float64, nlocal * 3, (nlocal, 3))
    d["radius"] = take(np.float64, nall)
    d["mass"] = take(np.float64, nall)
    d["rotation_quat"] = take(np.float64, nlocal * 4, (nlocal, 4))
    d["rotation_matrix"] = take(np.float64, nlocal * 9, (nlocal, 9))
    d["uid"] = take(np.int32, nall)
    d["flags"] = take(np.int32, nall)
    d["type"] = take(np.int32, nall)
    d["num_contacts"] = take(np.int32, nlocal)
    d["contact_lists"] = take(np.int32, nlocal * cap, (nlocal, cap))
    d["is_sticking"] = take(np.int32, nlocal * cap, (nlocal, cap))
    d["tangential_spring_displacement"] = take(np.float64, nlocal * cap * 3, (nlocal, cap, 3))
    d["impact_velocity_magnitude"] = take(np.float64, nlocal * cap, (nlocal, cap))
    return d


# DEM steps whose full state is kept (every step's contact count is kept)
DEM_STATE_STEPS = [0, 1, 2, 5, 10, 50, 100, 150, 200, 250, 300, 350, 400, 450, 499, 500]


def pack_dem(dump_dir, out_path, meta):
    dumps = {}
    for p in sorted(glob.glob(os.path.join(dump_dir, "dem_*.bin"))):
        d = read_dem_dump(p)
        dumps[d["ts"]] = d
    arrays = {}
    init = dumps[-1]
    n0 = init["nlocal"]
    for k in ("position", "linear_velocity", "angular_velocity", "radius", "mass", "uid", "flags", "type",
              "rotation_quat", "rotation_matrix"):
        arrays["init_" + k] = init[k][:n0]
    nsteps = max(dumps)
    arrays["nlocal_all"] = np.array([dumps[t]["nlocal"] for t in range(nsteps + 1)], np.int32)
    arrays["nghost_all"] = np.array([dumps[t]["nghost"] for t in range(nsteps + 1)], np.int32)
    arrays["ncontacts_all"] = np.array([int(dumps[t]["num_contacts"].sum()) for t in range(nsteps + 1)], np.int32)
    for t in DEM_STATE_STEPS:
        if t not in dumps:
            continue
        d = dumps[t]
        n = d["nlocal"]
        for k in ("position", "linear_velocity", "angular_velocity", "uid", "rotation_quat"):
            arrays[f"s{t}_{k}"] = d[k][:n]
        arrays[f"s{t}_force"] = d["force"]
        arrays[f"s{t}_torque"] = d["torque"]
        arrays[f"s{t}_ghost_uid"] = d["uid"][n:]
        arrays[f"s{t}_ghost_position"] = d["position"][n:]
        nc = d["num_contacts"]
        arrays[f"s{t}_num_contacts"] = nc
        rows = [np.arange(nc[i]) for i in range(n)]
        arrays[f"s{t}_contact_uid"] = np.concatenate(
            [d["contact_lists"][i, r] for i, r in enumerate(rows)] or [np.zeros(0, np.int32)]).astype(np.int32)
        arrays[f"s{t}_contact_sticking"] = np.concatenate(
            [d["is_sticking"][i, r] for i, r in enumerate(rows)] or [np.zeros(0, np.int32)]).astype(np.int32)
        arrays[f"s{t}_contact_tsd"] = np.concatenate(
            [d["tangential_spring_displacement"][i, r] for i, r in enumerate(rows)] or [np.zeros((0, 3))])
        arrays[f"s{t}_contact_impact"] = np.concatenate(
            [d["impact_velocity_magnitude"][i, r] for i, r in enumerate(rows)] or [np.zeros(0)])
    for k, v in meta.items():
        arrays["meta_" + k] = np.array(v)
    np.savez_compressed(out_path, **arrays)
    return arrays


def build_and_run_dem(workdir, cpp):
    exe = os.path.join(workdir, "dem_ref")
    libdir = os.path.join(workdir, "mpilib")
    os.makedirs(libdir, exist_ok=True)
    for lib in ("libmpi.so.12", "libgfortran.so.4", "libquadmath.so.0"):
        dst = os.path.join(libdir, lib)
        if not os.path.exists(dst):
            os.symlink(os.path.join(MPI_LIB, lib), dst)
    cmd = ["g++"] + CXXFLAGS + ["-I" + MPI_INC, "-I" + REF, "-I" + HERE, "-o", exe, cpp,
                                os.path.join(REF, "runtime/pairs.cpp"),
                                os.path.join(REF, "runtime/domain/regular_6d_stencil.cpp"),
                                os.path.join(REF, "runtime/devices/dummy.cpp"),
                                os.path.join(libdir, "libmpi.so.12"),
                                "-Wl,-rpath," + libdir, "-Wl,-rpath-link," + libdir]
    subprocess.run(cmd, check=True, cwd=workdir)
    subprocess.run([exe], cwd=workdir, check=True, capture_output=True, text=True)


def make_dem(name, steps, keep=None):
    workdir = keep or tempfile.mkdtemp(prefix="pairs_golden_dem_")
    os.makedirs(workdir, exist_ok=True)
    dump_dir = os.path.join(workdir, "dumps_" + name)
    shutil.rmtree(dump_dir, ignore_errors=True)
    os.makedirs(dump_dir)
    cpp = generate_dem(workdir, steps)
    instrument_dem(cpp, dump_dir, steps)
    build_and_run_dem(workdir, cpp)
    meta = dict(domain=[0.04, 0.015, 0.04], steps=steps, spacing=0.005, diameter=0.0029, min_diameter=0.00261,
                max_diameter=0.00319, initial_velocity=1.0, density=2550.0, dt=5e-5,
                compiler="g++ 11.4 + MPICH 3.3.2 " + " ".join(CXXFLAGS))
    out = os.path.join(HERE, f"{name}.npz")
    pack_dem(dump_dir, out, meta)
    if keep is None:
        shutil.rmtree(workdir, ignore_errors=True)
    print(f"wrote {out} ({os.path.getsize(out)} bytes)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--keep", default=None, help="keep the scratch build directory here")
    ap.add_argument("--only", default=None, choices=["md", "dem"])
    args = ap.parse_args()
    if args.only in (None, "md"):
        make_md("md_fcc4_full", 4, 1, 100, False, args.keep)
        make_md("md_fcc4_half", 4, 1, 100, True, args.keep)
    if args.only in (None, "dem"):
        make_dem("dem_sc168", DEM_STEPS, args.keep)


if __name__ == "__main__":
    main()
