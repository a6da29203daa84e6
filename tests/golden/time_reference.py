"""Calibrate the CPU baseline: time the reference's generated CPU program for
examples/md.py against the oracle's restatement (oracle/md_oracle.c) on the
same cores, for the 4x4x4 (256 atoms, 100 steps) and 32x32x32 (131072 atoms,
a 200-step sample) FCC workloads, one thread and all threads.

Development container only (like make_golden.py, whose generation step it
reuses): the reference tree, its Python code generator and MPICH are needed.
The generated program is built with the reference's own CPU flags
(Makefile: -O3 -mavx2 -mfma -fopenmp) and timed end to end; its own `all`
timer (src/pairs/sim/timestep.py:67-69, runtime/timers.hpp:31-68) covers
the timestep loop only, which is what the oracle's run() covers too.

Usage:  python tests/golden/time_reference.py [--threads N]
"""
import argparse
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import make_golden as mg  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

REF_FLAGS = ["-O3", "-mavx2", "-mfma", "-fopenmp", "-std=c++17", "-w"]


def build(workdir, cpp, exe):
    libdir = os.path.join(workdir, "mpilib")
    os.makedirs(libdir, exist_ok=True)
    for lib in ("libmpi.so.12", "libgfortran.so.4", "libquadmath.so.0"):
        dst = os.path.join(libdir, lib)
        if not os.path.exists(dst):
            os.symlink(os.path.join(mg.MPI_LIB, lib), dst)
    cmd = ["g++"] + REF_FLAGS + ["-I" + mg.MPI_INC, "-I" + mg.REF, "-o", exe, cpp,
                                 os.path.join(mg.REF, "runtime/pairs.cpp"),
                                 os.path.join(mg.REF, "runtime/domain/regular_6d_stencil.cpp"),
                                 os.path.join(mg.REF, "runtime/devices/dummy.cpp"),
                                 os.path.join(libdir, "libmpi.so.12"),
                                 "-Wl,-rpath," + libdir, "-Wl,-rpath-link," + libdir]
    subprocess.run(cmd, check=True, cwd=workdir)


def time_reference(nx, steps, threads, workdir):
    src = open(os.path.join(mg.REF, "examples/md.py")).read()
    script = mg.specialise_md(src, nx, 1, steps, False)
    sub = os.path.join(workdir, f"nx{nx}")
    os.makedirs(sub, exist_ok=True)
    cpp = mg.generate(sub, script)
    exe = os.path.join(sub, "md_ref")
    build(sub, cpp, exe)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    t0 = time.perf_counter()
    out = subprocess.run([exe], cwd=sub, env=env, check=True, capture_output=True, text=True).stdout
    wall = time.perf_counter() - t0
    m = re.search(r"^all: ([0-9.eE+-]+)", out, re.M)
    loop = float(m.group(1)) if m else None  # the program's own timer (ms)
    return wall, loop


def time_oracle(nx, steps, threads):
    po.set_threads(threads)
    s = po.fcc_system(nx)
    t0 = time.perf_counter()
    s.run(0, steps)
    return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    workdir = tempfile.mkdtemp(prefix="pairs_time_")
    try:
        print("| workload | threads | reference loop (s) | reference wall (s) | oracle (s) | oracle / reference |")
        print("|---|---:|---:|---:|---:|---:|")
        for nx, steps in ((4, 100), (32, 200)):
            for th in sorted({1, args.threads}):
                wall, loop = time_reference(nx, steps, th, workdir)
                ref = loop / 1e3 if loop is not None else wall
                orc = time_oracle(nx, steps, th)
                print(f"| FCC {nx}^3 = {4 * nx ** 3} atoms, {steps} steps | {th} | "
                      f"{'%.3f' % ref if loop is not None else '-'} | {wall:.3f} | {orc:.3f} | {orc / ref:.2f} |",
                      flush=True)
    finally:
        shutil.rmtree(workdir, ignore_errors=True)


if __name__ == "__main__":
    main()
