"""Parity of the HIP LJ path (libpairs_hip.so through its C ABI) with the CPU
oracle and with the golden fixtures of the reference's generated program.

Full and half neighbour lists: every value the reference computes is
bit-identical (positions, velocities, forces, ghosts, neighbour lists, cells,
send maps).  For half lists the mirror forces are summed from the receiving
particle's side in the reference's loop order, so no FP64 atomics reorder
them.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _md():
    from pairs_amd import md
    return md


def _rel(a, b):
    """max |a - b| relative to the array's scale, floored at 1 (the LJ unit of
    length, velocity and force): forces of a perfect lattice are rounding noise
    around zero, where a pointwise relative error means nothing."""
    scale = max(np.abs(b).max(), 1.0)
    return np.abs(a - b).max() / scale


def _compare_state(g, o, tag):
    """bit-identical positions (ghosts included), velocities and forces; for
    half lists too: the mirror forces are summed in the reference's order"""
    assert g.nlocal == o.nlocal and g.nghost == o.nghost, tag
    for name, a, b in (("position", g.positions(True), o.positions(True)), ("velocity", g.velocities(), o.velocities()),
                       ("force", g.forces(), o.forces())):
        assert np.array_equal(a, b), f"{tag} {name}: max |diff| {np.abs(a - b).max()} (rel {_rel(a, b):.3g})"


def _compare_lists(g, o, tag):
    assert np.array_equal(g.numneighs(), o.numneighs()), tag
    for a, b in zip(g.neighbor_lists(), o.neighbor_lists()):
        assert np.array_equal(a, b), tag
    assert np.array_equal(g.particle_cell(), o.particle_cell()), tag
    gs, gc = g.cells()
    os_, oc = o.cells()
    assert np.array_equal(gs, os_), tag
    assert all(np.array_equal(a, b) for a, b in zip(gc, oc)), tag
    gm, gmult = g.send_map()
    om, omult = o.send_map()
    assert np.array_equal(gm, om) and np.array_equal(gmult, omult), tag


@pytest.mark.parametrize("half", [False, True])
def test_modules_step_by_step_vs_oracle_fcc4(half):
    md = _md()
    g = md.fcc_system(4, half=half)
    o = po.fcc_system(4, half=half)
    assert np.array_equal(g.positions(), o.positions())
    assert np.array_equal(g.velocities(), o.velocities())
    for ts in range(101):
        g.timestep(ts, 100)
        o.step(ts)
        _compare_state(g, o, tag=f"ts={ts}")
        if (ts + 1) % 20 == 0 or ts == 0:
            _compare_lists(g, o, tag=f"ts={ts}")


@pytest.mark.parametrize("name,half", [("md_fcc4_full", False), ("md_fcc4_half", True)])
def test_fused_run_vs_golden_fixture(name, half):
    """pairs_hip_run (fused kernels, double-buffered positions, hipGraph replay)
    against the reference program's own trajectory."""
    md = _md()
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    g = md.fcc_system(4, half=half)
    stops = [0, 1, 2, 19, 20, 21, 40, 60, 99, 100]
    ts = 0
    for stop in stops:
        thermo = g.run(ts, stop + 1, 100)
        ts = stop + 1
        x, v, f = g.positions(), g.velocities(), g.forces()
        if f"s{stop}_position" in z.files:
            assert np.array_equal(x, z[f"s{stop}_position"]), stop
            assert np.array_equal(v, z[f"s{stop}_velocity"]), stop
            assert np.array_equal(f, z[f"s{stop}_force"]), stop
        if f"s{stop}_ghost_position" in z.files:
            gp = g.positions(True)[g.nlocal:]
            assert np.array_equal(gp, z[f"s{stop}_ghost_position"]), stop
        if stop in (0, 99):
            row = z["thermo_printed"][0 if stop == 0 else 1]
            assert len(thermo) >= 1
            assert abs(thermo[-1][0] - row[0]) <= 5e-6 * row[0] and abs(thermo[-1][1] - row[1]) <= 5e-6 * row[1]


@pytest.mark.parametrize("graphs", [True, False])
def test_fused_run_vs_oracle_fcc8_300_steps(graphs):
    md = _md()
    g = md.fcc_system(8)
    g.set_graphs(graphs)
    o = po.fcc_system(8)
    th = g.run(0, 301, 300)
    o.run(0, 301)
    _compare_state(g, o, tag="fcc8")
    _compare_lists(g, o, tag="fcc8")
    # thermo records at ts = 0, 99, 199, 299; the last one against the oracle's
    assert len(th) == 4
    t, p = o.last_thermo()
    assert abs(th[-1][0] - t) <= 1e-12 * t and abs(th[-1][1] - p) <= 1e-12 * p


def test_half_list_fcc8_vs_oracle():
    md = _md()
    g = md.fcc_system(8, half=True)
    o = po.fcc_system(8, half=True)
    g.run(0, 201, 200)
    o.run(0, 201)
    _compare_state(g, o, tag="fcc8-half")
    _compare_lists(g, o, tag="fcc8-half")


def test_resize_protocol_small_capacities():
    """Tiny neighbour/send capacities overflow; the library grows them (x2 the
    need, as the generated resize loop) and the results are unchanged."""
    md = _md()
    g = md.fcc_system(4, neighbor_capacity=8, send_capacity=16, particle_capacity=300)
    o = po.fcc_system(4)
    g.run(0, 41, 40)
    o.run(0, 41)
    _compare_state(g, o, tag="resize")
    _compare_lists(g, o, tag="resize")
    assert g.count(6) >= g.nsend_all and g.count(4) > 8


@pytest.mark.parametrize("which", ["send", "particles"])
def test_overflow_at_a_later_rebuild(which):
    """A send-map or particle overflow at the rebuild of ts = 19, when the
    other position buffer holds the positions of ts = 18: the retry must start
    from the current state.  At density 0.55 the FCC 4^3 lattice sends 843
    particles at ts = 0 and more at ts = 19 (the oracle counts them), so a
    capacity that fits ts = 0 exactly overflows at ts = 19."""
    md = _md()
    o = po.fcc_system(4, rho=0.55)
    o.step(0)
    nsend0, nall0 = o.nsend_all, o.nlocal + o.nghost
    for ts in range(1, 20):
        o.step(ts)
    assert o.nsend_all > nsend0 and o.nlocal + o.nghost > nall0
    kw = dict(send_capacity=nsend0) if which == "send" else dict(particle_capacity=nall0)
    g = md.fcc_system(4, rho=0.55, **kw)
    g.run(0, 1, 60)
    from pairs_amd import native as nat
    kind = nat.SEND_CAPACITY if which == "send" else nat.PARTICLE_CAPACITY
    cap0 = g.count(kind)
    assert cap0 == (nsend0 if which == "send" else nall0)  # ts = 0 fitted
    g.run(1, 61, 60)
    assert g.count(kind) > cap0  # grown at a later rebuild
    # run() checks a rebuild's counters only at the next rebuild or at the
    # end: the overflowing rebuild ran ahead and was repeated from its
    # checkpoint
    assert g.count(nat.REBUILDS_DEFERRED) >= 2 and g.count(nat.REBUILD_ROLLBACKS) >= 1
    o = po.fcc_system(4, rho=0.55)
    o.run(0, 61)
    _compare_state(g, o, tag=f"overflow {which}")
    _compare_lists(g, o, tag=f"overflow {which}")


@pytest.mark.parametrize("half", [False, True])
def test_deferred_rebuild_checks_match_immediate_ones(half):
    """Rebuild counters checked at the next rebuild (default) or right after
    each rebuild: bit-identical trajectories, lists and maps over runs split
    at arbitrary steps, with and without an overflow that forces a repeat."""
    md = _md()
    from pairs_amd import native as nat
    res = []
    for spec in (True, False):
        g = md.fcc_system(6, half=half, neighbor_capacity=40)
        g.set_speculative_rebuilds(spec)
        for a, b in ((0, 7), (7, 45), (45, 46), (46, 130)):
            g.run(a, b, 129)
        res.append(g)
    a, b = res
    assert a.count(nat.REBUILDS_DEFERRED) > 0 and b.count(nat.REBUILDS_DEFERRED) == 0
    assert a.count(nat.NEIGHBOR_CAPACITY) == b.count(nat.NEIGHBOR_CAPACITY) > 40
    _compare_state(a, b, tag=f"deferred half={half}")
    _compare_lists(a, b, tag=f"deferred half={half}")


@pytest.mark.parametrize("nx", [6, 3])
@pytest.mark.parametrize("half", [False, True])
def test_graph_chains_and_run_starts_vs_oracle(half, nx):
    """Runs split at arbitrary steps, rebuilds every 60 steps: each run's
    first step replays from its own graph, each stretch of steady steps from
    one chain graph (stretches longer than the longest chain, 24 steps, are
    cut into several), one-step runs and last steps run eagerly.  The state
    after 200 steps is the oracle's, bit for bit.  FCC 6^3 (box 10.1): every
    ghost is one of at most 8 images of a local, so a run starts by
    integrating into the other buffer with the ghost images written there;
    FCC 3^3 (box 5.04, under two cell spacings): locals have up to 17 images,
    so a run starts with the in-place ghost refresh."""
    md = _md()
    g = md.fcc_system(nx, half=half, reneighbor_every=60, thermo_every=1000)
    o = po.fcc_system(nx, half=half, reneighbor_every=60, thermo_every=1000)
    g.run(0, 3, 200)
    g.prepare_graphs()
    for a, b in ((3, 31), (31, 32), (32, 90), (90, 91), (91, 157), (157, 201)):
        g.run(a, b, 200)
    o.run(0, 201)
    _compare_state(g, o, tag=f"chains half={half} nx={nx}")
    _compare_lists(g, o, tag=f"chains half={half} nx={nx}")


def test_deferred_check_repeat_keeps_thermo_rows():
    """A send-map overflow at a later rebuild, found a rebuild later: the run
    repeats from the checkpoint, and the thermo rows it returns (every 10
    steps, some of them computed on the faulty rebuild before the repeat) are
    those of checks after every rebuild; the state equals the oracle's."""
    md = _md()
    from pairs_amd import native as nat
    o = po.fcc_system(4, rho=0.55)
    o.step(0)
    nsend0 = o.nsend_all
    rows = []
    for spec in (True, False):
        g = md.fcc_system(4, rho=0.55, send_capacity=nsend0, thermo_every=10)
        g.set_speculative_rebuilds(spec)
        th = [g.run(0, 1, 80)]
        th.append(g.run(1, 81, 80))
        rows.append((g, th))
    (a, ta), (b, tb) = rows
    assert a.count(nat.REBUILD_ROLLBACKS) >= 1 and b.count(nat.REBUILD_ROLLBACKS) == 0
    for x, y in zip(ta, tb):
        assert x.shape == y.shape and np.array_equal(x, y)
    assert len(tb[1]) == 8
    o = po.fcc_system(4, rho=0.55)
    o.run(0, 81)
    _compare_state(a, o, tag="deferred repeat")
    _compare_lists(a, o, tag="deferred repeat")


@pytest.mark.parametrize("half", [False, True])
def test_energy_and_virial_vs_oracle(half):
    md = _md()
    g = md.fcc_system(6, half=half)
    o = po.fcc_system(6, half=half)
    g.run(0, 51, 50)
    o.run(0, 51)
    pe_g, w_g = g.energy()
    pe_o, w_o, e_o = o.energy()
    assert abs(pe_g - pe_o) <= 1e-12 * abs(pe_o)
    assert abs(w_g - w_o) <= 1e-12 * abs(w_o)
    e_g = g.atom_energies()
    assert np.allclose(e_g, e_o, rtol=1e-12, atol=1e-14)
    th = g.compute_thermo()
    assert abs(th[0] - o.temperature()) <= 1e-12 * th[0]
    V = po.fcc_box_length(6, 0.8442) ** 3
    n = o.nlocal
    assert abs(th[4] - (th[0] * (3 * n - 3) / 3.0 + w_o / 3.0) / V) <= 1e-10 * abs(th[4])


def test_c2_size_properties_and_oracle_100_steps():
    """131072 atoms (FCC 32^3): the perfect lattice has exactly 78 neighbours per
    atom inside 2.8; 100 fused steps stay bit-identical to the oracle."""
    md = _md()
    g = md.fcc_system(32)
    g.run(0, 1, 1000)
    nn = g.numneighs()
    assert len(nn) == 131072 and np.all(nn == 78)
    o = po.fcc_system(32)
    o.step(0)
    assert np.array_equal(g.forces(), o.forces())
    g.run(1, 101, 1000)
    o.run(1, 101)
    assert np.array_equal(g.positions(), o.positions())
    assert np.array_equal(g.velocities(), o.velocities())
    assert np.array_equal(g.forces(), o.forces())
    assert g.nghost == o.nghost


def test_c2_half_lists_oracle_100_steps():
    """131072 atoms with half lists: the lists equal the oracle's (a local
    keeps its later-index neighbours and every ghost neighbour, so the counts
    of the perfect lattice add up to 78 N / 2 plus half the local-ghost
    pairs); 100 fused steps, mirror parts included, bit-identical to the
    oracle's sequential half-list loop."""
    md = _md()
    g = md.fcc_system(32, half=True)
    g.run(0, 1, 1000)
    o = po.fcc_system(32, half=True)
    o.step(0)
    _compare_lists(g, o, tag="c2-half ts=0")
    nn = g.numneighs()
    n_ghost_pairs = 2 * int(nn.sum()) - 78 * len(nn)
    assert nn.max() == 78 and 0 < n_ghost_pairs < 78 * len(nn)
    assert np.array_equal(g.forces(), o.forces())
    g.run(1, 101, 1000)
    o.run(1, 101)
    _compare_state(g, o, tag="c2-half")


def test_vtk_output_of_lj_system(tmp_path):
    from pairs_amd import md, io
    s = md.fcc_system(4)
    s.run(0, 10, 100)
    assert io.vtk_write(s, str(tmp_path / "md"), 10, 5)
    lines = open(str(tmp_path / "md_10.vtk")).read().split("\n")
    assert lines[4] == "POINTS 256 double"
    x = s.positions()
    assert lines[5] == f"{x[0, 0]:.8f} {x[0, 1]:.8f} {x[0, 2]:.8f}"
