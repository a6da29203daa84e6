"""Pin the CPU oracle of the LJ path (oracle/md_oracle.c).

1. Against the golden fixtures written by tests/golden/make_golden.py from the
   reference's own generated CPU program (bit-exact: positions, velocities,
   forces, ghost positions, neighbour lists, cell membership, send maps).
2. Against what the mathematics fixes, independently of the reference:
   closed-form two-body LJ force, brute-force neighbour sets over explicit
   periodic images, Newton's third law, -dU/dx == F, full == half forces,
   lattice/thermo closed forms.
"""
import itertools
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


@pytest.mark.parametrize("name,half", [("md_fcc4_full", False), ("md_fcc4_half", True)])
def test_oracle_matches_reference_program_bit_exact(name, half):
    z = _golden(name)
    s = po.fcc_system(4, half=half)
    assert np.array_equal(s.positions(), z["init_position"])
    assert np.array_equal(s.velocities(), z["init_velocity"])
    checked = 0
    for ts in range(int(z["meta_steps"]) + 1):
        s.step(ts)
        assert s.nlocal == z["nlocal_all"][ts] and s.nghost == z["nghost_all"][ts], ts
        if f"s{ts}_position" in z.files:
            assert np.array_equal(s.positions(), z[f"s{ts}_position"]), ts
            assert np.array_equal(s.velocities(), z[f"s{ts}_velocity"]), ts
            assert np.array_equal(s.forces(), z[f"s{ts}_force"]), ts
            checked += 1
        if f"s{ts}_ghost_position" in z.files:
            assert np.array_equal(s.positions(True)[s.nlocal:], z[f"s{ts}_ghost_position"]), ts
        if f"s{ts}_numneighs" in z.files:
            assert np.array_equal(s.numneighs(), z[f"s{ts}_numneighs"])
            assert np.array_equal(np.concatenate(s.neighbor_lists()), z[f"s{ts}_neighbors"])
            assert np.array_equal(s.particle_cell(), z[f"s{ts}_particle_cell"])
            sm, mult = s.send_map()
            assert np.array_equal(sm, z[f"s{ts}_send_map"])
            assert np.array_equal(mult, z[f"s{ts}_send_mult"])
            assert np.array_equal(s.nsend(), z[f"s{ts}_nsend"])
            sizes, cells = s.cells()
            assert np.array_equal(sizes, z[f"s{ts}_cell_sizes"])
            assert np.array_equal(np.concatenate(cells), z[f"s{ts}_cell_particles"])
        if ts in (0, 99):
            t, p = s.last_thermo()
            row = z["thermo_printed"][0 if ts == 0 else 1]
            # the program prints 6 significant digits
            assert abs(t - row[0]) <= 5e-6 * abs(row[0]) and abs(p - row[1]) <= 5e-6 * abs(row[1])
    assert checked == 17


def test_fcc_lattice_closed_form():
    for n in (2, 3, 4):
        s = po.fcc_system(n)
        x = s.positions()
        assert len(x) == 4 * n ** 3
        a = math.pow(4.0 / 0.8442, 1.0 / 3.0)
        # every site is 0.5*a*(i,j,k) with i+j+k even
        ijk = np.rint(x / (0.5 * a)).astype(int)
        assert np.allclose(ijk * (0.5 * a), x, rtol=0, atol=1e-12)
        assert np.all(ijk.sum(axis=1) % 2 == 0)
        assert len({tuple(r) for r in ijk}) == len(x)
        v = s.velocities()
        assert np.allclose(v.sum(axis=0), 0.0, atol=1e-12)
        assert abs(s.temperature() - 1.44) < 1e-12


def _lj_force(d, eps=1.0, s6=1.0):
    r2 = d @ d
    sr2 = 1.0 / r2
    sr6 = sr2 ** 3 * s6
    return d * (48.0 * sr6 * (sr6 - 0.5) * sr2 * eps)


def test_two_body_force_closed_form():
    L = 20.0
    for r in (0.95, 2.0 ** (1.0 / 6.0), 1.3, 2.2, 2.49, 2.6):
        s = po.OracleMD([(0.0, L)] * 3, 64, reneighbor_every=1)
        x = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
        s.set_particles(x, np.zeros_like(x))
        s.setup_cells()
        s.rebuild()
        s.force()
        f = s.forces()
        # analytic: F = 24 eps (2 (s/r)^12 - (s/r)^6) / r along the separation
        fa = 24.0 * (2.0 / r ** 13 - 1.0 / r ** 7) if r < 2.5 else 0.0
        assert abs(f[0, 0] + fa) <= 1e-12 * max(1.0, abs(fa))
        assert np.allclose(f[0] + f[1], 0.0, atol=1e-14)
        if abs(r - 2.0 ** (1.0 / 6.0)) < 1e-15:
            assert abs(f[0, 0]) < 1e-12


def test_older_api_modules():
    """The older API's pair force (examples/lj_onetype.py:5-8) is the same
    LJ force as md.py's with another rounding order: both agree to a few
    ulps, each equal to the two-body closed form; its Euler update moves a
    particle beyond the cutoff (zero force) by exactly dt * v per step."""
    L = 20.0
    for r in (0.95, 1.3, 2.2):
        s = po.OracleMD([(0.0, L)] * 3, 64, reneighbor_every=1)
        x = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
        s.set_particles(x, np.zeros_like(x))
        s.setup_cells()
        s.rebuild()
        s.force()
        f_md = s.forces().copy()
        s.lj_older()
        f_old = s.forces()
        fa = 24.0 * (2.0 / r ** 13 - 1.0 / r ** 7)
        assert abs(f_old[0, 0] + fa) <= 1e-12 * max(1.0, abs(fa))
        assert np.allclose(f_old, f_md, rtol=1e-14, atol=1e-300)
        assert np.all(f_old[0] + f_old[1] == 0.0)
    # a many-body configuration: still the same forces to rounding
    o = po.fcc_system(4)
    o.run(0, 30)
    o.force()
    f_md = o.forces().copy()
    o.lj_older()
    assert np.allclose(o.forces(), f_md, rtol=0, atol=1e-12 * np.abs(f_md).max())
    # Euler on two particles 5 apart (cutoff 2.5): x_k = x_{k-1} + dt v, v fixed
    s = po.OracleMD([(0.0, L)] * 3, 64, reneighbor_every=1, dt=0.005)
    x = np.array([[5.0, 5.0, 5.0], [10.0, 5.0, 5.0]])
    v = np.array([[0.3, -0.2, 0.1], [-0.1, 0.25, 0.05]])
    s.set_particles(x, v)
    s.setup_cells()
    xr = x.copy()
    for _ in range(7):
        s.rebuild()
        s.lj_older()
        s.euler()
        xr = xr + 0.005 * v
    assert np.array_equal(s.velocities(), v)
    assert np.array_equal(s.positions(), xr)


def _brute_images(x, L, rlist):
    """All (i, j, image) pairs with |x_i - (x_j + image L)| < rlist, images in {-1,0,1}^3."""
    n = len(x)
    out = set()
    r2 = rlist * rlist
    for sh in itertools.product((-1, 0, 1), repeat=3):
        shift = np.array(sh, float) * L
        d = x[:, None, :] - (x[None, :, :] + shift)
        rsq = (d * d).sum(axis=2)
        ii, jj = np.nonzero(rsq < r2)
        for i, j in zip(ii, jj):
            if sh == (0, 0, 0) and i == j:
                continue
            out.add((int(i), int(j), sh))
    return out


@pytest.mark.parametrize("half", [False, True])
def test_neighbor_sets_brute_force(half):
    """Lists built at ts=19 (liquid) equal the brute-force set of (i, root j,
    periodic image) within the list radius; a half list holds each local pair
    once (i < j) and each cross-boundary pair from both sides."""
    s = po.fcc_system(4, half=half)
    for ts in range(20):
        s.step(ts)
    x = s.positions(True)
    nl = s.nlocal
    L = po.fcc_box_length(4, 0.8442)
    sm, mult = s.send_map()
    root = np.arange(len(x))
    img = np.zeros((len(x), 3), int)
    for g in range(s.nghost):
        src = sm[g]
        root[nl + g] = root[src]
        img[nl + g] = img[src] + mult[g]
    got = [(i, int(root[j]), tuple(int(k) for k in img[j])) for i, lst in enumerate(s.neighbor_lists()) for j in lst]
    assert len(got) == len(set(got))
    got = set(got)
    # ts=19 is a rebuild step; the lists use the positions of that step, which
    # stay in the locals until the force kernel -- the step has not integrated since
    want = _brute_images(x[:nl], L, 2.8)
    if not half:
        assert got == want
    else:
        local = {(i, j, sh) for (i, j, sh) in want if sh == (0, 0, 0) and i < j}
        cross = {(i, j, sh) for (i, j, sh) in want if sh != (0, 0, 0)}
        assert got == local | cross


def test_neighbor_sets_brute_force_exact_images():
    """Full list at ts=0 (lattice, rebuild) equals the brute-force image set exactly."""
    s = po.fcc_system(4)
    s.step(0)
    x = s.positions(True)
    nl = s.nlocal
    sm, mult = s.send_map()
    root = np.arange(len(x))
    for g in range(s.nghost):
        root[nl + g] = root[sm[g]]
    L = po.fcc_box_length(4, 0.8442)
    got = set()
    for i, lst in enumerate(s.neighbor_lists()):
        for j in lst:
            d = x[j] - x[root[j]]
            sh = tuple(int(round(c / L)) for c in d)
            got.add((i, int(root[j]), sh))
    want = _brute_images(x[:nl], L, 2.8)
    assert got == want
    assert all(n == 78 for n in s.numneighs())  # FCC shells 12+6+24+12+24 inside 2.8


def test_borders_small_box_both_faces():
    """FCC 3^3 (box 5.04, shorter than two cell spacings of 2.8): the send map
    and ghost positions of the first rebuild equal the borders stated from the
    reference's selection rule (src/pairs/sim/comm.py:57-99,
    domain_partitioning.py:31-65): per dimension, every particle so far (locals,
    then the earlier dimensions' ghosts) with x_d < min_d + 2.8 in index order,
    then every one with x_d > max_d - 2.8, each appended as a ghost shifted by
    +L / -L -- a particle near both faces is sent to both sides."""
    s = po.fcc_system(3)
    x = s.positions().copy()
    L = po.fcc_box_length(3, 0.8442)
    pts, smap, mults = x.copy(), [], []
    for d in range(3):
        lo = [i for i in range(len(pts)) if pts[i, d] < 0.0 + 2.8]
        hi = [i for i in range(len(pts)) if pts[i, d] > L - 2.8]
        new = []
        for side, sel in ((1, lo), (-1, hi)):
            for i in sel:
                q = pts[i].copy()
                q[d] += side * L
                new.append(q)
                smap.append(i)
                mults.append([side if k == d else 0 for k in range(3)])
        pts = np.vstack([pts, np.array(new)])
    both = sum(1 for i in range(len(x)) if 2.8 > x[i, 0] > L - 2.8)
    assert both > 0  # the case under test: particles selected on both sides
    s.step(0)
    sm, mult = s.send_map()
    assert s.nghost == len(pts) - len(x) == 990
    assert np.array_equal(sm, smap) and np.array_equal(mult, mults)
    assert np.array_equal(s.positions(True), pts)


def test_full_and_half_forces_agree():
    a = po.fcc_system(4)
    b = po.fcc_system(4, half=True)
    for ts in range(40):
        a.step(ts)
        b.step(ts)
    fa, fb = a.forces(), b.forces()
    assert np.allclose(fa, fb, rtol=0, atol=1e-11 * np.abs(fa).max())
    # Newton's third law over a periodic system
    assert np.allclose(fa.sum(axis=0), 0.0, atol=1e-10)


def test_energy_gradient_is_force():
    s = po.fcc_system(3)
    for ts in range(30):
        s.step(ts)
    x = s.positions()
    v = s.velocities()
    L = po.fcc_box_length(3, 0.8442)

    def pe_at(xx):
        t = po.OracleMD([(0.0, L)] * 3, 4096, reneighbor_every=1)
        t.set_particles(xx, v)
        t.setup_cells()
        t.rebuild()
        t.force()
        return t.energy()[0], t.forces()

    _, f0 = pe_at(x)
    h = 1e-6
    for (i, d) in [(0, 0), (5, 1), (17, 2), (60, 0)]:
        xp = x.copy(); xp[i, d] += h
        xm = x.copy(); xm[i, d] -= h
        num = -(pe_at(xp)[0] - pe_at(xm)[0]) / (2 * h)
        assert abs(num - f0[i, d]) < 1e-6 * max(1.0, abs(f0[i, d]))


def test_energy_two_body_closed_form():
    s = po.OracleMD([(0.0, 20.0)] * 3, 64, reneighbor_every=1)
    r = 1.2
    x = np.array([[5.0, 5.0, 5.0], [5.0, 5.0 + r, 5.0]])
    s.set_particles(x, np.zeros_like(x))
    s.setup_cells()
    s.rebuild()
    s.force()
    pe, w, e = s.energy()
    u = 4.0 * (r ** -12 - r ** -6)
    assert abs(pe - u) < 1e-13 and abs(e.sum() - u) < 1e-13
    fr = 24.0 * (2.0 / r ** 13 - 1.0 / r ** 7)
    assert abs(w - r * fr) < 1e-12


def _energy_of(x, v, L, half):
    t = po.OracleMD([(0.0, L)] * 3, 8192, half=half, reneighbor_every=1)
    t.set_particles(x, v)
    t.setup_cells()
    t.rebuild()
    t.force()
    return t.energy()


def test_half_list_energy_and_virial_equal_full_list():
    """The half-list weighting (a local-local pair counted once with weight 1,
    a local-ghost pair seen from the local side with weight 1/2) gives the
    full list's PE and virial (every ordered pair with weight 1/2) on the same
    state, to rounding; the per-atom energies sum to the PE on both."""
    s = po.fcc_system(5)
    s.run(0, 60)
    x, v = s.positions(), s.velocities()
    L = po.fcc_box_length(5, 0.8442)
    pe_f, w_f, e_f = _energy_of(x, v, L, False)
    pe_h, w_h, e_h = _energy_of(x, v, L, True)
    assert abs(pe_h - pe_f) <= 1e-12 * abs(pe_f)
    assert abs(w_h - w_f) <= 1e-12 * abs(w_f)
    assert abs(e_f.sum() - pe_f) <= 1e-12 * abs(pe_f)
    # a dropped local-ghost pair or a doubled local-local pair moves the sum
    # by far more than rounding: one pair's energy is about 1e-2 of |pe| / N
    assert abs(pe_f) > 100.0


@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("across", [False, True])
def test_energy_two_body_closed_form_full_and_half(half, across):
    """Two particles at distance 1.2, either inside the box (a local-local
    pair) or across the periodic boundary in x (each local sees the other's
    ghost: two local-ghost entries): U = 4 (r^-12 - r^-6) and
    W = r F(r) with F = 24 (2 r^-13 - r^-7), for full and half lists."""
    r, L = 1.2, 10.0
    if across:
        x = np.array([[0.5, 5.0, 5.0], [L - (r - 0.5), 5.0, 5.0]])
    else:
        x = np.array([[5.0, 5.0, 5.0], [5.0, 5.0 + r, 5.0]])
    pe, w, e = _energy_of(x, np.zeros_like(x), L, half)
    u = 4.0 * (r ** -12 - r ** -6)
    fr = 24.0 * (2.0 / r ** 13 - 1.0 / r ** 7)
    assert abs(pe - u) < 1e-13, (pe, u)
    assert abs(w - r * fr) < 1e-12, (w, r * fr)
    if not half:
        assert abs(e.sum() - u) < 1e-13
