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

Beside it: the MLP potential at 131072 atoms (training, force check, MD
throughput) and the DEM path on a larger domain, each as its own object of the
same JSON line; with --batched also the batched ensemble (512 uneven
replicas).  The batched line is off by default: one run of it ended in a GPU
memory fault whose cause has not been found (the small batched ensembles of
the GPU tests run).

  python bench.py [--gpus 1] [--steps 1000] [--warmup 100] [--no-half] [--no-cpu]
                  [--batched] [--no-mlp] [--no-dem]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
FP64_VECTOR_PEAK_TFLOPS = 78.6  # MI355X FP64 vector, vendor figure
FP64_OPS_PER_ENTRY = 21
NX = 32
NATOMS = 4 * NX ** 3


def force_kernel_bytes(n, g, s, fused):
    """Algorithmic HBM bytes of one force launch.

    LJ only (the neighbour-list force): 24(N+G) positions + 4N counts + 4S list
    indices + 24N force write.  The fused kernel of the timed loop also does
    final_integrate + initial_integrate of its own atoms and writes the next
    positions of its atoms and of their ghost images: + 8N mass + 48N velocity
    read/write + 24(N+G) position write + 32(N+G) packed (x, y, z, 0) write.
    """
    b = 24 * (n + g) + 4 * n + 4 * s + 24 * n
    if fused:
        b += 8 * n + 48 * n + 24 * (n + g) + 32 * (n + g)
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


def run_gpu(steps, warmup, half, time_every, time_steps):
    """warmup steps, graph capture (set-up), exactly `steps` timed steps with
    nothing else in the window, then a separate pass of `time_steps` steps
    that times every `time_every`-th force launch with HIP events"""
    from pairs_amd import md

    sim = md.fcc_system(NX, half=half)
    nsteps = warmup + steps + time_steps
    if warmup > 0:
        sim.run(0, warmup, nsteps)
        sim.prepare_graphs()  # set-up: both buffers' step and rebuild graphs
    sim.sync()
    t0 = time.perf_counter()
    thermo = sim.run(warmup, warmup + steps, nsteps)
    sim.sync()
    t1 = time.perf_counter()
    sim.set_force_timing(time_every)
    sim.run(warmup + steps, nsteps, nsteps)
    sim.sync()
    ms, launches = sim.force_kernel_time()
    n = sim.nlocal
    g = sim.nghost
    s_total = sim.count(7)
    del sim
    return dict(seconds=t1 - t0, force_ms=ms, force_launches=launches, nlocal=n, nghost=g, nneigh=s_total,
                thermo=thermo.tolist())


def run_batched(steps, warmup, nreplicas):
    """The uneven replica ensemble: batched (one launch per step, graphs,
    autotuned segment size) against a per-replica launch loop over the same
    replicas and steps."""
    from pairs_amd import batch

    t0 = time.perf_counter()
    B = batch.uneven_ensemble(nreplicas)
    setup_s = time.perf_counter() - t0
    natoms = B.natoms
    nsteps = 100000  # schedule length: longer than every range run here
    B.run(0, warmup, nsteps)
    chosen, cand_ms, ts = B.autotune(warmup, 10, nsteps)
    B.run(ts, ts + 20, nsteps)  # re-enter the graph path after tuning
    ts += 20
    t0 = time.perf_counter()
    B.run(ts, ts + steps, nsteps)
    t_batched = time.perf_counter() - t0
    ts += steps
    t0 = time.perf_counter()
    B.run_each(ts, ts + steps, nsteps)
    t_each = time.perf_counter() - t0
    B.close()
    return {
        "value": natoms * steps / t_batched,
        "unit": "atom-steps/s",
        "per_replica_loop_value": natoms * steps / t_each,
        "speedup_vs_per_replica_loop": t_each / t_batched,
        "ms_per_step": t_batched * 1e3 / steps,
        "replicas": nreplicas,
        "atoms": natoms,
        "steps": steps,
        "autotune": {"chosen_home_per_segment": chosen, "ms_per_step": cand_ms},
        "setup_s": setup_s,
        "workload": f"{nreplicas} periodic LJ replicas, N ~ U[500, 20000], rho ~ U[0.70, 0.90], "
                    f"T ~ U[0.8, 2.0], full lists, rebuild every 20, FP64",
    }


MLP_TEMPS = (0.8, 1.1, 1.4, 1.7, 2.0)
MLP_CONFIG_NX = 8  # force-training systems: FCC 8^3 = 2048 atoms each


def run_mlp(steps, warmup, train_steps, force_steps):
    """The MLP potential at 131072 atoms: GPU training on LJ data, the force
    check that decides MLP or LJ fallback, and the MLP MD throughput.

    Training, all on the GPU: (1) energy rows -- every atom's descriptor and
    per-atom LJ energy from 10 snapshots of the 131072-atom system (five
    temperatures in [0.8, 2.0], two snapshots 50 steps apart each), Adam on
    minibatches of 4096 rows; (2) configurations with forces -- five plain-LJ
    systems of 2048 atoms at the same temperatures each advance 50 steps
    between Adam steps (velocities rescaled to their temperature every 500
    steps), and every step takes the gradient of 0.01 x the energy term plus
    the relative squared force error of all five configurations
    (force_steps x 5 snapshots sampled every 50 steps; Adam from a fresh
    state, learning rate 1e-2 decaying linearly to 5e-4)."""
    from pairs_amd import md, mlp

    m = mlp.MLPModel(seed=20241202, dataset_capacity=12 * NATOMS)
    t0 = time.perf_counter()
    rows = mlp.collect_lj_snapshots(m, lambda T: md.fcc_system(NX, temp=T), MLP_TEMPS,
                                    equilibrate=200, every=50, per_temp=2)
    t_collect = time.perf_counter() - t0
    m.fit_scaling()
    rmse0 = m.rmse()
    batch = 4096
    t0 = time.perf_counter()
    loss = m.train(train_steps, batch=batch, lr=1e-2, seed=11)
    t_train = time.perf_counter() - t0
    rmse1 = m.rmse()
    # held-out state: plain LJ at T = 1.44 after 300 steps (the last a rebuild)
    sim = md.fcc_system(NX)
    sim.run(0, 300, 100000)
    rel_energy_only = mlp.force_error(sim, m)
    t0 = time.perf_counter()
    hist = []
    if force_steps > 0:
        sims = []
        for T in MLP_TEMPS:
            s = md.fcc_system(MLP_CONFIG_NX, temp=T, particle_capacity=16384, send_capacity=16384)
            s.run(0, 200, 10 ** 9)
            sims.append(s)
        # a fresh optimiser state for the new loss (the energy phase's Adam
        # moments belong to another objective), the parameters copied back
        p0, s0 = m.params()
        mf = mlp.MLPModel(seed=1, dataset_capacity=16)
        mf.set_params(p0, s0)
        _, hist = mlp.train_on_configurations(mf, sims, force_steps, every=50, ts0=200, lr=1e-2, w_energy=0.01,
                                              w_force=1.0, total_steps=10 ** 9,
                                              record_every=max(force_steps // 10, 1), temps=MLP_TEMPS)
        m.set_params(*mf.params())
        mf.close()
        for s in sims:
            s.close()
    t_force = time.perf_counter() - t0
    rmse2 = m.rmse()
    rel = mlp.force_error(sim, m)
    attached, _ = mlp.attach(sim, m, 0.02)
    # MLP MD throughput (attached regardless of the check, to time the path)
    if not attached:
        mlp.attach(sim, m, -1.0)
    sim.run(300, 300 + warmup, 100000)
    sim.sync()
    sim.set_force_timing(10)
    t0 = time.perf_counter()
    sim.run(300 + warmup, 300 + warmup + steps, 100000)
    sim.sync()
    t_md = time.perf_counter() - t0
    ms, launches = sim.force_kernel_time()
    sim.close()
    return {
        "value": NATOMS * steps / t_md,
        "unit": "atom-steps/s",
        "ms_per_step": t_md * 1e3 / steps,
        "force_eval_ms": ms,
        "force_rel_rms_vs_lj": rel,
        "force_rel_rms_vs_lj_energy_training_only": rel_energy_only,
        "fallback_to_lj": not attached,
        "fallback_rule": "MLP used only if relative RMS force error vs LJ <= 0.02",
        "energy_rmse_train_before": rmse0,
        "energy_rmse_train_after": rmse1,
        "energy_rmse_train_after_force_training": rmse2,
        "training": {"energy_rows": rows, "energy_steps": train_steps, "batch": batch,
                     "samples_per_s": train_steps * batch / t_train, "seconds": t_train,
                     "final_loss": float(loss[-50:].mean()) if len(loss) else None,
                     "snapshot_collection_s": t_collect,
                     "force_steps": force_steps,
                     "force_configurations": force_steps * len(MLP_TEMPS),
                     "force_configuration_atoms": 4 * MLP_CONFIG_NX ** 3,
                     "force_seconds": t_force,
                     "force_loss_history": [{"step": k, "energy_term": a, "force_term": b} for k, a, b in hist]},
        "workload": "FCC 32^3 = 131072 atoms, rho 0.8442, T 1.44, descriptor 16 Gaussians rc 2.5, "
                    "MLP 16-64-64-1 SiLU BF16 MFMA / FP32, FP64 integration; trained on LJ energies (10 snapshots "
                    "of 131072 atoms) and on energies + forces of 2048-atom LJ configurations sampled every 50 "
                    "steps at T 0.8..2.0",
    }


def run_dem(steps):
    """The DEM path (examples/dem.py physics) on a larger domain: about 191k
    spheres in 0.35 x 0.35 x 0.2 m, periodic x/y, planes at the bottom and top,
    timed over the whole run from the initial grid (free fall, impacts and
    settling on the bottom plane)."""
    from pairs_amd import dem

    s = dem.reference_system((0.35, 0.35, 0.2), particle_capacity=400000, send_capacity=80000)
    n = s.nlocal
    s.set_timing(50)
    s.sync()
    t0 = time.perf_counter()
    s.run(0, steps)
    s.sync()
    t = time.perf_counter() - t0
    ms, launches = s.contact_kernel_time()
    phases = s.phase_times()
    nc = s.get("num_contacts")
    s.close()
    return {
        "value": n * steps / t,
        "unit": "particle-steps/s",
        "ms_per_step": t * 1e3 / steps,
        "contact_kernel_ms": ms,
        "phase_ms": dict(zip(("prepare", "contact_search", "history_slots", "spring_dashpot", "sums", "euler"),
                             phases)),
        "spheres": n,
        "steps": steps,
        "contacts_at_end": int(nc.sum()),
        "workload": "examples/dem.py parameters (d 2.9 mm, rho 2550, dt 5e-5, e 0.1, mu 0.5, collision time "
                    "5e-4) on a 0.35 x 0.35 x 0.2 m domain from the SC grid, rebuild every step, FP64",
    }


def run_dem_settling(steps):
    """The DEM settling workload: 200 000 spheres, radii uniform in
    [0.9, 1.1] mm, 2500 kg/m3, at rest at seeded random non-overlapping
    positions (a 40 x 40 x 125 grid of pitch 2.5 x 2.5 x 2.4 mm, each
    coordinate jittered uniformly by +-0.1 mm, MT19937-64 seed 7) in a
    0.1 x 0.1 x 0.3 m box, periodic in x and y, a floor plane at z = 0,
    gravity -9.81 z, linear spring-dashpot with fixed k_n = 1e4 N/m and the
    damping of restitution 0.6, Coulomb friction 0.5 with per-contact history,
    dt 1e-6 s."""
    from pairs_amd import dem

    params = dict(diameter=0.002, gravity=9.81, density_fluid=0.0, density_particle=2500.0, dt=1e-6,
                  friction=0.5, restitution=0.6, friction_static=0.0, stiffness_norm=1e4)
    box = (0.1, 0.1, 0.3)
    s = dem.DEMSystem(box, periodic=(True, True, False), particle_capacity=320000, send_capacity=80000,
                      params=params)
    s.jittered_grid((40, 40, 125), (0.00125, 0.00125, 0.0012), (0.0025, 0.0025, 0.0024), 1e-4, 0.9e-3, 1.1e-3,
                    seed=7)
    s.add_halfspace(1000000, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    s.setup()
    n = s.nlocal
    s.set_timing(50)
    s.sync()
    t0 = time.perf_counter()
    s.run(0, steps)
    s.sync()
    t = time.perf_counter() - t0
    phases = s.phase_times()
    nc = s.get("num_contacts")
    s.close()
    return {
        "value": n * steps / t,
        "unit": "particle-steps/s",
        "ms_per_step": t * 1e3 / steps,
        "phase_ms": dict(zip(("prepare", "contact_search", "history_slots", "spring_dashpot", "sums", "euler"),
                             phases)),
        "particles": n,
        "steps": steps,
        "contacts_at_end": int(nc.sum()),
        "workload": "DEM settling: 200000 spheres r ~ U[0.9, 1.1] mm, 2500 kg/m3, jittered grid (seed 7), "
                    "0.1 x 0.1 x 0.3 m, periodic x/y, floor plane, k_n 1e4 N/m, e 0.6, mu 0.5 with history, "
                    "dt 1e-6 s, FP64",
    }


def measured_copy_gbps(mib=1024, reps=10):
    """A device-to-device copy of mib MiB timed with HIP events: the HBM rate
    this device reaches on plain streaming (bytes read + bytes written)."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * n * 8 / (ms * 1e-3) / 1e9


def run_cpu_baseline(sample_steps, threads):
    """The oracle (C restatement of the reference's generated CPU program) on a
    bounded sample of the same workload, with the given number of OpenMP
    threads (1 = the reference's serial target_cpu())."""
    from oracle import pyoracle as po

    po.set_threads(threads)
    s = po.fcc_system(NX)
    t0 = time.perf_counter()
    s.run(0, sample_steps)
    t1 = time.perf_counter()
    return NATOMS * sample_steps / (t1 - t0), t1 - t0


def host_threads():
    """Cores for the parallel CPU baseline: OMP_NUM_THREADS when set, else the
    cores this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def progress(msg):
    """one line per finished section on stderr (the JSON line stays alone on stdout)"""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-half", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--time-every", type=int, default=5,
                    help="in the force-timing pass after the timed steps, time every n-th force launch")
    ap.add_argument("--time-steps", type=int, default=200, help="steps of the force-timing pass")
    ap.add_argument("--batched", action="store_true", help="add the 512-replica batched ensemble line")
    ap.add_argument("--no-batched", action="store_true", help="(default)")
    ap.add_argument("--no-mlp", action="store_true")
    ap.add_argument("--no-dem", action="store_true")
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--side-steps", type=int, default=200,
                    help="timed steps of the batched, MLP and DEM lines")
    ap.add_argument("--mlp-train-steps", type=int, default=12000)
    ap.add_argument("--mlp-force-steps", type=int, default=10000,
                    help="Adam steps on configurations with forces (5 configurations each)")
    ap.add_argument("--dem-steps", type=int, default=10000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device once: a kernel of known HBM traffic that "
                         "calibrates FETCH_SIZE / WRITE_SIZE counter runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: pairs_amd has no multi-GPU path "
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
    copy_gbps = measured_copy_gbps()
    full = run_gpu(args.steps, args.warmup, False, args.time_every, args.time_steps)
    progress(f"full list: {full['seconds'] * 1e3 / args.steps:.4f} ms/step, force kernel {full['force_ms']:.4f} ms")
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
            "measured_copy_GBps": copy_gbps,
            "frac_of_measured_copy": (achieved / copy_gbps) if (achieved and copy_gbps) else None,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this workload, scaled by a "
                              "1 GiB device copy (profiles/lj_fcc32_force_pmc.md)",
            "kernel": "k_lj2<1,1,false,false,true> (LJ force + final/initial integrate + next packed positions, "
                      "fused; two lanes per local, 32 locals per wave)",
            "timing": "HIP events around every 5th force launch of a separate pass of 200 steps after the timed "
                      "steps (eager launches on the context stream)",
            "bytes_per_launch": fused_bytes,
            "lj_only_bytes_per_launch": lj_bytes,
            "avg_launch_ms": full["force_ms"],
            "launches_timed": full["force_launches"],
            # the same kernel against the FP64 vector peak: every list entry
            # is evaluated (branch-free, masked past the cutoff), 21 FP64
            # operations as the reference writes the pair (3 differences, 3
            # multiplies + 2 adds for rsq, 1 division, 2 multiplies for sr6,
            # 4 for the scalar factor, 3 multiplies + 3 adds for the force)
            "fp64": {
                "flops_per_launch": FP64_OPS_PER_ENTRY * full["nneigh"],
                "achieved_TFLOPs": (FP64_OPS_PER_ENTRY * full["nneigh"] / (full["force_ms"] * 1e-3) / 1e12
                                    if full["force_ms"] > 0 else None),
                "peak_TFLOPs": FP64_VECTOR_PEAK_TFLOPS,
                "frac": (FP64_OPS_PER_ENTRY * full["nneigh"] / (full["force_ms"] * 1e-3) / 1e12
                         / FP64_VECTOR_PEAK_TFLOPS if full["force_ms"] > 0 else None),
            },
        },
    }
    if not args.no_half:
        half = run_gpu(args.steps, args.warmup, True, args.time_every, args.time_steps)
        progress(f"half list: {half['seconds'] * 1e3 / args.steps:.4f} ms/step")
        hb = force_kernel_bytes(half["nlocal"], half["nghost"], half["nneigh"], False)
        rec["half_list"] = {
            "value": NATOMS * args.steps / half["seconds"],
            "ms_per_step": half["seconds"] * 1e3 / args.steps,
            "force_kernel_ms": half["force_ms"],
            "force_kernel_GBps": hb / (half["force_ms"] * 1e-3) / 1e9 if half["force_ms"] > 0 else None,
            "neighbour_entries": half["nneigh"],
        }
    if args.batched and not args.no_batched:
        rec["batched"] = run_batched(args.side_steps, 40, args.replicas)
        progress(f"batched: {rec['batched']['ms_per_step']:.4f} ms/step")
    if not args.no_mlp:
        rec["mlp"] = run_mlp(args.side_steps, 20, args.mlp_train_steps, args.mlp_force_steps)
        rec["mlp"]["vs_plain_lj"] = rec["mlp"]["value"] / value
        progress(f"mlp: force error {rec['mlp']['force_rel_rms_vs_lj']:.4f}")
    if not args.no_dem:
        rec["dem"] = run_dem(args.dem_steps)
        progress(f"dem: {rec['dem']['ms_per_step']:.4f} ms/step")
        rec["dem_settling"] = run_dem_settling(args.dem_steps)
        progress(f"dem settling: {rec['dem_settling']['ms_per_step']:.4f} ms/step")
    if not args.no_cpu:
        v1, secs1 = run_cpu_baseline(args.cpu_steps, 1)
        progress(f"cpu baseline, 1 thread: {secs1:.1f} s")
        nt = host_threads()
        vp, secsp = run_cpu_baseline(args.cpu_steps, nt) if nt > 1 else (v1, secs1)
        sample = f"same workload, iterations ts=0..{args.cpu_steps - 1} (includes {args.cpu_steps // 20} rebuilds)"
        rec["cpu_baseline"] = {"value": vp, "unit": "atom-steps/s", "cores": nt, "kind": "port",
                               "sample": f"{sample}: {secsp:.1f} s on {nt} OpenMP threads "
                                         "(neighbour build, force and integration loops)",
                               "serial": {"value": v1, "cores": 1, "seconds": secs1}}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
