"""The batched ensemble (include/pairs_hip_batch.h): every replica's
trajectory is the one its own single-system run (and the CPU oracle) gives --
bit-identical positions, velocities and forces -- whether the replicas advance
in one fused launch per step, in a per-replica launch loop, or after the
autotuner changed their segment size."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

SIZES = [500, 1372, 2100, 3001, 864, 4000]
RHO = [0.70, 0.90, 0.8442, 0.75, 0.82, 0.88]
TEMP = [0.8, 2.0, 1.44, 1.1, 1.7, 1.0]


def _batch():
    from pairs_amd import batch
    return batch


def _single(b, r):
    """replica r as its own LJSystem with the same initial state"""
    from pairs_amd import md
    n = b.sizes[r]
    L = b.boxes[r]
    s = md.LJSystem([(0.0, L)] * 3, particle_capacity=8 * n + 4096, send_capacity=8 * n + 4096,
                    neighbor_capacity=128)
    x, v = b.initial[r]
    s.upload(x, v)
    s._call("pairs_hip_setup_cells")
    return s


def _same(a, b, tag):
    assert np.array_equal(a.positions(), b.positions()), tag
    assert np.array_equal(a.velocities(), b.velocities()), tag
    assert np.array_equal(a.forces(), b.forces()), tag


def test_batched_equals_single_systems():
    B = _batch().BatchEnsemble(SIZES, RHO, TEMP)
    singles = [_single(B, r) for r in range(len(SIZES))]
    rows = []
    for a, e in ((0, 1), (1, 45), (45, 122)):
        B.run(a, e, 300)
        rows = [s.run(a, e, 300) for s in singles]
        for r, s in enumerate(singles):
            _same(B.replicas[r], s, (r, e))
    # the batched thermo reduction (one workgroup per replica) gives each
    # replica's own row of ts = 99
    th = B.thermo()
    assert np.all(th[:, 0] > 0)
    for r, rr in enumerate(rows):
        t, p = rr[-1][0], rr[-1][1]
        assert abs(th[r, 0] - t) <= 1e-12 * t and abs(th[r, 1] - p) <= 1e-12 * abs(p), r


def test_batched_equals_oracle():
    B = _batch().BatchEnsemble(SIZES[:3], RHO[:3], TEMP[:3])
    B.run(0, 81, 300)
    for r in range(3):
        L = B.boxes[r]
        o = po.OracleMD([(0.0, L)] * 3, 8 * B.sizes[r] + 4096, ncap=128, send_cap=8 * B.sizes[r] + 4096)
        x, v = B.initial[r]
        o.set_particles(x, v)
        o.setup_cells()
        o.run(0, 81)
        g = B.replicas[r]
        assert np.array_equal(g.positions(True), o.positions(True)), r
        assert np.array_equal(g.velocities(), o.velocities()), r
        assert np.array_equal(g.forces(), o.forces()), r


def test_persistent_launch_grid_keeps_trajectories():
    """The persistent force launch with one, seven or all resident workgroups
    (each walking a contiguous run of the wave table): the same trajectories,
    bit for bit, as the replicas' own single-system runs."""
    bm = _batch()
    ens = []
    for wg in (1, 7, 0):
        B = bm.BatchEnsemble(SIZES, RHO, TEMP)
        B.set_workgroups(wg)
        B.run(0, 1, 300)
        B.run(1, 61, 300)
        ens.append(B)
    singles = [_single(ens[0], r) for r in range(len(SIZES))]
    for s in singles:
        s.run(0, 61, 300)
    for B, wg in zip(ens, (1, 7, 0)):
        for r, s in enumerate(singles):
            _same(B.replicas[r], s, (wg, r))


def test_run_each_and_autotune_keep_trajectories():
    bm = _batch()
    A = bm.BatchEnsemble(SIZES, RHO, TEMP)
    C = bm.BatchEnsemble(SIZES, RHO, TEMP)
    A.run(0, 1, 400)
    chosen, times, ts = A.autotune(1, 10, 400, candidates=(48, 112, 200))
    assert chosen in (48, 112, 200) and all(t > 0 for t in times.values())
    assert ts == 70  # rebuilds at 19, 39, 59, ten timed steps after each
    A.run(ts, 121, 400)
    C.run_each(0, 121, 400)
    for r in range(len(SIZES)):
        _same(A.replicas[r], C.replicas[r], r)


def _oracle_replica(B, r, ts_end):
    L = B.boxes[r]
    n = B.sizes[r]
    o = po.OracleMD([(0.0, L)] * 3, 8 * n + 4096, ncap=128, send_cap=8 * n + 4096)
    x, v = B.initial[r]
    o.set_particles(x, v)
    o.setup_cells()
    o.run(0, ts_end)
    return o


def test_batched_overflow_retry_and_tile_growth():
    """Replicas whose send map (replica 0) or particle arrays (replica 1)
    overflow at the rebuild of ts = 19, while the other buffer holds ts = 18,
    retry on their own from the current state; a larger segment size set
    mid-run makes every replica's tiles outgrow their staged capacity at the
    rebuild of ts = 59, so the batched launch runs with grown tile_caps.  All
    three replicas stay bit-identical to the oracle."""
    from pairs_amd import native as nat
    sizes, rho, temp = [500, 864, 1372], [0.55, 0.6, 0.8442], [1.44, 2.0, 1.44]
    bm = _batch()
    probe = bm.BatchEnsemble(sizes, rho, temp)
    need = []
    for r in range(3):
        o = _oracle_replica(probe, r, 1)
        need.append((o.nsend_all, o.nlocal + o.nghost))
        o.run(1, 20)
        assert r == 2 or (o.nsend_all > need[r][0] and o.nlocal + o.nghost > need[r][1])
    probe.close()
    send = [need[0][0], 8 * sizes[1] + 4096, 8 * sizes[2] + 4096]
    part = [8 * sizes[0] + 4096, need[1][1], 8 * sizes[2] + 4096]
    B = bm.BatchEnsemble(sizes, rho, temp, send_capacity=send, particle_capacity=part)
    B.run(0, 41, 300)
    assert B.replicas[0].count(nat.SEND_CAPACITY) > need[0][0]
    assert B.replicas[1].count(nat.PARTICLE_CAPACITY) > need[1][1]
    for r in range(3):
        o = _oracle_replica(B, r, 41)
        g = B.replicas[r]
        assert np.array_equal(g.positions(True), o.positions(True)), r
        assert np.array_equal(g.velocities(), o.velocities()), r
        assert np.array_equal(g.forces(), o.forces()), r
    caps = [B.replicas[r].count(nat.TILE_CAPACITY) for r in range(3)]
    B.set_home(400)
    B.run(41, 101, 300)
    grown = [B.replicas[r].count(nat.TILE_CAPACITY) for r in range(3)]
    assert any(gc > c for gc, c in zip(grown, caps)), (caps, grown)
    for r in range(3):
        o = _oracle_replica(B, r, 101)
        g = B.replicas[r]
        assert np.array_equal(g.positions(True), o.positions(True)), r
        assert np.array_equal(g.velocities(), o.velocities()), r
        assert np.array_equal(g.forces(), o.forces()), r
