"""MLP potential on the GPU (include/pairs_hip_mlp.h) against the FP64 CPU
restatement (oracle/mlp_oracle.py, pinned by finite differences and closed
forms in tests/test_oracle_mlp.py).

Tolerances: descriptors are FP32 sums of at most ~80 terms of size <= 1,
each Gaussian from a two-anchor recurrence with relative error below 1e-5
(rtol 3e-5, atol 1e-5).  The network runs in BF16 on MFMA; against the
oracle with the device's rounding points emulated (forward_bf16) the energy
and gradients agree to float32-accumulation noise (rel 5e-3 RMS; a transposed
operand or wrong lane layout gives errors of order 1), and the forces to
rel 1e-2 RMS; against the exact FP64 network the BF16 budget is 2 %."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _mods():
    from pairs_amd import md, mlp
    return md, mlp


def _scale(seed=3):
    rng = np.random.default_rng(seed)
    sc = np.zeros(mo.NSCALE)
    sc[0:16] = rng.uniform(0.0, 3.0, 16)
    sc[16:32] = rng.uniform(0.3, 2.0, 16)
    sc[32], sc[33] = -5.5, 0.7
    return sc


def _state(n=4, steps=40, temp=1.44):
    """an FCC n^3 system after `steps` iterations; the last one (ts = steps - 1)
    is a rebuild step, so every ghost is an exact periodic image"""
    md, _ = _mods()
    s = md.fcc_system(n, temp=temp)
    s.run(0, steps, 1000)
    L = s.box[0][1] - s.box[0][0]
    return s, L


def _rel_rms(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2) / np.mean(np.asarray(b) ** 2)))


def test_descriptors_vs_oracle():
    _, mlp = _mods()
    s, L = _state()
    G = mlp.descriptors(s)
    Gr = mo.descriptors(s.positions(), L)
    assert G.shape == (256, 16)
    assert np.allclose(G, Gr, rtol=3e-5, atol=1e-5), np.abs(G - Gr).max()
    assert Gr.max() > 1.0  # the shells are populated


def test_network_vs_oracle():
    _, mlp = _mods()
    m = mlp.MLPModel(seed=1)
    p, sc = mo.xavier(7).astype(np.float32), _scale().astype(np.float32)
    m.set_params(p, sc)
    rng = np.random.default_rng(4)
    G = rng.uniform(0.0, 4.0, (1000, 16)).astype(np.float32)  # 62.5 tiles: a ragged last tile
    E, D = m.eval(G)
    Eb, Db = mo.forward_bf16(p, sc, G)
    Ex, Dx = mo.forward(p, sc, G)
    dev = Eb - Eb.mean()
    assert np.sqrt(np.mean((E - Eb) ** 2)) < 5e-3 * np.sqrt(np.mean(dev ** 2)), "energy vs emulated rounding"
    assert _rel_rms(D, Db) < 5e-3, "dE/dG vs emulated rounding"
    assert np.sqrt(np.mean((E - Ex) ** 2)) < 2e-2 * Ex.std(), "BF16 budget (energy)"
    assert _rel_rms(D, Dx) < 2e-2, "BF16 budget (gradient)"


def test_forces_energy_virial_vs_oracle():
    _, mlp = _mods()
    s, L = _state()
    m = mlp.MLPModel(seed=2)
    G = mlp.descriptors(s)
    sc = np.zeros(mo.NSCALE, np.float32)
    sc[0:16] = G.mean(0)
    sc[16:32] = 1.0 / (G.std(0) + 1e-3)
    sc[32], sc[33] = -6.0, 0.5
    p = mo.xavier(9).astype(np.float32)
    m.set_params(p, sc)
    e, w = mlp.forces(s, m)
    F = s.forces()
    x = s.positions()
    Fb, Eb = mo.forces(p, sc, x, L, device_rounding=True)
    Fx, Ex = mo.forces(p, sc, x, L)
    assert _rel_rms(F, Fb) < 1e-2, _rel_rms(F, Fb)
    assert _rel_rms(F, Fx) < 3e-2, _rel_rms(F, Fx)
    assert abs(e - Eb.sum()) < 1e-3 * abs(Eb.sum())
    assert abs(w - mo.virial(p, sc, x, L)) < 3e-2 * abs(mo.virial(p, sc, x, L)) + 1e-3
    assert np.abs(F.sum(0)).max() < 1e-3 * np.abs(F).max() * len(F) ** 0.5


def test_fallback_keeps_lj_trajectory_bit_identical():
    md, mlp = _mods()
    a = md.fcc_system(4)
    b = md.fcc_system(4)
    m = mlp.MLPModel(seed=3)  # untrained: its forces are nowhere near LJ
    a.run(0, 20, 100)
    b.run(0, 20, 100)
    used, rel = mlp.attach(a, m, 0.02)
    assert not used and rel > 0.02 and not mlp.active(a)
    a.run(20, 81, 100)
    b.run(20, 81, 100)
    assert np.array_equal(a.positions(True), b.positions(True))
    assert np.array_equal(a.velocities(), b.velocities())
    assert np.array_equal(a.forces(), b.forces())


def test_mlp_md_graph_and_module_paths_agree():
    """the MLP step replayed from graphs == launched kernel by kernel"""
    md, mlp = _mods()
    m = mlp.MLPModel(seed=4)
    s0, _ = _state(n=5, steps=20)
    G = mlp.descriptors(s0)
    sc = np.zeros(mo.NSCALE, np.float32)
    sc[0:16] = G.mean(0)
    sc[16:32] = 1.0 / (G.std(0) + 1e-3)
    sc[32], sc[33] = -6.0, 0.05  # a gentle potential keeps the run stable
    m.set_params(mo.xavier(5).astype(np.float32), sc)
    runs = []
    for graphs in (True, False):
        s, _ = _state(n=5, steps=20)
        s.set_graphs(graphs)
        used, _ = mlp.attach(s, m, -1.0)
        assert used and mlp.active(s)
        s.run(20, 90, 1000)
        runs.append((s.positions(True), s.velocities(), s.forces()))
    for u, v in zip(runs[0], runs[1]):
        assert np.array_equal(u, v)
    assert np.all(np.isfinite(runs[0][0]))


def _lj_set(md, mlp, temps, n, per_temp, seed_steps=120):
    """descriptors and per-atom LJ energies of snapshots (host copies)"""
    Gs, Es = [], []
    for T in temps:
        s = md.fcc_system(n, temp=T)
        total = seed_steps + 40 * per_temp
        s.run(0, seed_steps, total)
        ts = seed_steps
        for _ in range(per_temp):
            s.run(ts, ts + 40, total)
            ts += 40
            Gs.append(mlp.descriptors(s))
            Es.append(s.atom_energies())
        s.close()
    return np.concatenate(Gs), np.concatenate(Es)


def test_training_and_ablation():
    md, mlp = _mods()
    m = mlp.MLPModel(seed=5, dataset_capacity=200000)
    added = mlp.collect_lj_snapshots(m, lambda T: md.fcc_system(8, temp=T), (0.8, 1.2, 1.6, 2.0),
                                     equilibrate=120, every=40, per_temp=4)
    assert added == 16 * 2048 and m.dataset_size == added
    m.fit_scaling()
    before = m.rmse()
    loss = m.train(12000, batch=4096, lr=1e-2, seed=11)
    after = m.rmse()
    assert np.all(np.isfinite(loss))
    assert loss[-100:].mean() < 0.2 * loss[:20].mean()
    assert after < 0.3 * before
    # held-out snapshots (other temperatures), against a constant and a
    # least-squares linear model on the same descriptors
    Gt, Et = _lj_set(md, mlp, (1.0, 1.44, 1.9), 8, 2)
    Gl, El = _lj_set(md, mlp, (0.8, 1.2, 1.6, 2.0), 8, 2, seed_steps=200)
    E_mlp, _ = m.eval(Gt)
    A = np.c_[Gl, np.ones(len(Gl))]
    coef, *_ = np.linalg.lstsq(A, El, rcond=None)
    E_lin = np.c_[Gt, np.ones(len(Gt))] @ coef
    r_mlp = float(np.sqrt(np.mean((E_mlp - Et) ** 2)))
    r_lin = float(np.sqrt(np.mean((E_lin - Et) ** 2)))
    r_const = float(np.sqrt(np.mean((El.mean() - Et) ** 2)))
    print(f"held-out energy RMSE: mlp {r_mlp:.4f}  linear {r_lin:.4f}  constant {r_const:.4f}")
    assert r_mlp < 0.1 * r_const      # the descriptors carry the energy
    assert r_mlp < r_lin              # and the network beats a linear fit on them
    # the force check decides MLP or LJ by the 2 % rule, whatever the error is
    s = md.fcc_system(8, temp=1.44)
    s.run(0, 200, 1000)
    rel = mlp.force_error(s, m)
    used, rel2 = mlp.attach(s, m, 0.02)
    print(f"force error vs LJ: {rel:.4f} -> {'MLP' if used else 'LJ fallback'}")
    assert rel == rel2 and used == (rel <= 0.02) and mlp.active(s) == used
