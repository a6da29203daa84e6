"""Benchmark of the LJ molecular-dynamics hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): FCC lattice
32x32x32 unit cells = 131072 atoms, one type, reduced density 0.8442, T 1.44,
cutoff 2.5, skin 0.3, dt 0.005, fully periodic, neighbour rebuild every 20
steps, FP64.  One "step" is one iteration of the reference's timestep loop
(src/pairs/sim/timestep.py:9-72): integrate, rebuild (every 20) or ghost
refresh, LJ force, integrate, thermo (every 100).

value = atoms x steps / wall time of the timed steps (inputs resident in HBM,
set-up excluded), full neighbour list; the half-list form, the force kernel's
HBM roofline and the CPU baseline are reported beside it.

  python bench.py [--gpus 1] [--steps 1000] [--warmup 100] [--no-half] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
NX = 32
NATOMS = 4 * NX ** 3


def force_kernel_bytes(n, g, s, fused):
    """Algorithmic HBM bytes of one force launch.

    LJ only (the neighbour-list force): 24(N+G) positions + 4N counts + 4S list
    indices + 24N force write.  The fused kernel of the timed loop also does
    final_integrate + initial_integrate of its own atoms: + 8N mass + 48N
    velocity read/write + 24N position write (its forces stay in registers).
    """
    b = 24 * (n + g) + 4 * n + 4 * s + 24 * n
    if fused:
        b += 8 * n + 48 * n + 24 * n
    return b


PMC_SUMMARY = os.path.join(ROOT, "profiles", "lj_fcc32_force_pmc.md")


def profiled_traffic():
    """HBM bytes per launch of the fused force kernel, as measured by the
    FETCH_SIZE / WRITE_SIZE counter runs (scaled by a copy of known size)
    recorded in the profile summary; None when the summary is absent."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith("TRAFFIC_BYTES_PER_LAUNCH"):
                    return int(line.split("=")[1])
    except OSError:
        pass
    return None


def run_gpu(steps, warmup, half, time_every):
    from pairs_amd import md

    sim = md.fcc_system(NX, half=half)
    nsteps = warmup + steps
    sim.run(0, warmup, nsteps)
    sim.sync()
    sim.set_force_timing(time_every)
    t0 = time.perf_counter()
    thermo = sim.run(warmup, nsteps, nsteps)
    sim.sync()
    t1 = time.perf_counter()
    ms, launches = sim.force_kernel_time()
    n = sim.nlocal
    g = sim.nghost
    s_total = sim.count(7)
    del sim
    return dict(seconds=t1 - t0, force_ms=ms, force_launches=launches, nlocal=n, nghost=g, nneigh=s_total,
                thermo=thermo.tolist())


def run_cpu_baseline(sample_steps):
    """The oracle (single-threaded C restatement of the reference's generated
    CPU program) on a bounded sample of the same workload."""
    from oracle import pyoracle as po

    s = po.fcc_system(NX)
    t0 = time.perf_counter()
    s.run(0, sample_steps)
    t1 = time.perf_counter()
    return NATOMS * sample_steps / (t1 - t0), t1 - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-half", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--time-every", type=int, default=10)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device once: a kernel of known HBM traffic that "
                         "calibrates FETCH_SIZE / WRITE_SIZE counter runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the single-domain hot path; sharding is a design note)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe > 0:
        import torch
        n = args.copy_probe * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    full = run_gpu(args.steps, args.warmup, False, args.time_every)
    value = NATOMS * args.steps / full["seconds"]
    fused_bytes = force_kernel_bytes(full["nlocal"], full["nghost"], full["nneigh"], True)
    lj_bytes = force_kernel_bytes(full["nlocal"], full["nghost"], full["nneigh"], False)
    achieved = fused_bytes / (full["force_ms"] * 1e-3) / 1e9 if full["force_ms"] > 0 else None
    rec = {
        "metric": "atom-steps/sec + force-kernel HBM GB/s, LJ 131k atoms FP64; batched/MLP vs plain",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": full["seconds"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: FCC lattice with Park-Miller velocities (the reference's copper_fcc_lattice + adjust_thermo)",
        "config": {"workload": "plain LJ MD, FCC 32^3 = 131072 atoms, rho 0.8442, T 1.44, rc 2.5, skin 0.3, "
                               "dt 0.005, periodic, rebuild every 20, full neighbour list",
                   "atoms": NATOMS, "ghosts": full["nghost"], "neighbour_entries": full["nneigh"],
                   "list": "full"},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": profiled_traffic(),
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this workload, scaled by a "
                              "512 MiB copy (profiles/lj_fcc32_force_pmc.md)",
            "kernel": "k_lj_full<8,true,false> (LJ force + final/initial integrate, fused)",
            "bytes_per_launch": fused_bytes,
            "lj_only_bytes_per_launch": lj_bytes,
            "avg_launch_ms": full["force_ms"],
            "launches_timed": full["force_launches"],
        },
    }
    if not args.no_half:
        half = run_gpu(args.steps, args.warmup, True, args.time_every)
        hb = force_kernel_bytes(half["nlocal"], half["nghost"], half["nneigh"], False)
        rec["half_list"] = {
            "value": NATOMS * args.steps / half["seconds"],
            "ms_per_step": half["seconds"] * 1e3 / args.steps,
            "force_kernel_ms": half["force_ms"],
            "force_kernel_GBps": hb / (half["force_ms"] * 1e-3) / 1e9 if half["force_ms"] > 0 else None,
            "neighbour_entries": half["nneigh"],
        }
    if not args.no_cpu:
        v, secs = run_cpu_baseline(args.cpu_steps)
        rec["cpu_baseline"] = {"value": v, "unit": "atom-steps/s", "cores": 1, "kind": "port",
                               "sample": f"same workload, iterations ts=0..{args.cpu_steps - 1} "
                                         f"({secs:.1f} s, includes {args.cpu_steps // 20} rebuilds), one thread"}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
