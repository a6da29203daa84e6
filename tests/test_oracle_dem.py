"""The CPU oracle's DEM restatement pinned against the reference.

- bit-exact against tests/golden/dem_sc168.npz, the reference's generated
  program for examples/dem.py on a 0.04 x 0.015 x 0.04 m domain (168 spheres,
  two half-spaces, 500 steps: free fall, impacts on the floor, sphere-sphere
  contacts with tangential history, particles crossing the periodic x/y
  boundaries with their contact history);
- closed forms: the C++ standard's check value of std::mt19937, free fall,
  the weight carried by a sphere resting on a plane, the restitution
  coefficient of a head-on collision.
"""
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dem_sc168.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def test_mt19937_check_value():
    # [rand.predef]: the 10000th consecutive output of a default-constructed
    # std::mt19937 is 4123659995
    o = po.OracleDEM((1.0, 1.0, 1.0))
    v = 0
    for _ in range(10000):
        v = o.mt_next()
    assert v == 4123659995


def test_initial_state_bit_exact(golden):
    o = po.dem_reference_system()
    assert o.nlocal == len(golden["init_uid"]) == 170
    for k in ("position", "linear_velocity", "radius", "mass", "uid", "flags", "type", "rotation_quat",
              "rotation_matrix"):
        assert np.array_equal(o.get(k), golden["init_" + k]), k
    # closed forms: sphere mass 4/3 pi r^3 rho, inverse inertia 1/(0.4 m r^2)
    r, m = o.get("radius")[:168], o.get("mass")[:168]
    assert np.allclose(m, 4.0 / 3.0 * math.pi * r ** 3 * 2550.0, rtol=1e-15)
    inv = o.get("inv_inertia")[:168]
    assert np.allclose(inv[:, 0], 1.0 / (0.4 * m * r * r), rtol=1e-14)
    assert np.all(np.isinf(o.get("mass")[168:]))


def _contacts_flat(o):
    nc, cl, st, tsd, iv = o.contacts()
    rows = [np.arange(nc[i]) for i in range(len(nc))]
    cat = lambda a: np.concatenate([a[i, r] for i, r in enumerate(rows)]) if nc.sum() else a[:0, 0]
    return nc, cat(cl), cat(st), (np.concatenate([tsd[i, r] for i, r in enumerate(rows)]) if nc.sum()
                                  else np.zeros((0, 3))), cat(iv)


def test_trajectory_bit_exact(golden):
    o = po.dem_reference_system()
    steps = sorted({int(k[1:].split("_")[0]) for k in golden.files
                    if k.startswith("s") and k.endswith("_linear_velocity")})
    assert steps[-1] == 500
    ts = 0
    contacts_seen = 0
    for st in steps:
        o.run(ts, st + 1)
        ts = st + 1
        assert o.nlocal == golden["nlocal_all"][st] and o.nghost == golden["nghost_all"][st], st
        for k in ("position", "linear_velocity", "angular_velocity", "force", "torque", "rotation_quat", "uid"):
            assert np.array_equal(o.get(k), golden[f"s{st}_{k}"]), (st, k)
        gp = o.get("position", ghosts=True)[o.nlocal:]
        assert np.array_equal(gp, golden[f"s{st}_ghost_position"]), st
        nc, uid, stick, tsd, imp = _contacts_flat(o)
        assert np.array_equal(nc, golden[f"s{st}_num_contacts"]), st
        assert np.array_equal(uid, golden[f"s{st}_contact_uid"]), st
        assert np.array_equal(stick, golden[f"s{st}_contact_sticking"]), st
        assert np.array_equal(tsd, golden[f"s{st}_contact_tsd"]), st
        assert np.array_equal(imp, golden[f"s{st}_contact_impact"]), st
        contacts_seen = max(contacts_seen, int(nc.sum()))
    assert contacts_seen > 100
    # particles crossed the periodic boundaries (the local order changed)
    assert (np.diff(golden["s500_uid"][:168]) < 0).any()


def test_free_fall_closed_form():
    cfg = dict(po.DEM_REF)
    o = po.OracleDEM((0.1, 0.1, 0.1), periodic=(1, 1, 0), cfg=cfg)
    x0, v0 = np.array([0.05, 0.05, 0.08]), np.array([0.01, -0.02, -0.5])
    o.add_sphere(1, 0.0014, x0, v0)
    o.setup()
    n = 300
    o.run(0, n)
    t = n * cfg["dt"]
    a = np.array([0.0, 0.0, -(1.0 - cfg["rho_f"] / cfg["rho_p"]) * cfg["gravity"]])
    x = o.get("position")[0]
    assert np.allclose(x, x0 + v0 * t + 0.5 * a * t * t, rtol=1e-12, atol=1e-15)
    assert np.allclose(o.get("linear_velocity")[0], v0 + a * t, rtol=1e-12, atol=1e-15)
    assert np.array_equal(o.get("angular_velocity")[0], np.zeros(3))


def test_sphere_resting_on_plane_carries_its_weight():
    cfg = dict(po.DEM_REF)
    o = po.OracleDEM((0.1, 0.1, 0.1), periodic=(1, 1, 0), cfg=cfg)
    r = 0.0014
    o.add_sphere(1, r, (0.05, 0.05, r), (0.0, 0.0, 0.0))
    o.add_halfspace(100000, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    o.setup()
    o.run(0, 4000)  # 40 collision times: the damped oscillation has settled
    m = o.get("mass")[0]
    w = (1.0 - cfg["rho_f"] / cfg["rho_p"]) * cfg["gravity"] * m
    # the contact force balances the weight; penetration = W / k_n with
    # k_n = m (pi^2 + ln^2 e) / t_c^2 (examples/dem.py:22-24; meff = m on a plane)
    assert abs(o.get("force")[0][2]) < 1e-9 * w
    kn = m * (math.pi ** 2 + math.log(cfg["restitution"]) ** 2) / cfg["t_c"] ** 2
    delta = r - o.get("position")[0][2]
    assert abs(delta - w / kn) < 1e-6 * (w / kn)
    nc, uid, _, _, _ = _contacts_flat(o)
    assert nc[0] == 1 and uid[0] == 100000


def test_head_on_collision_restitution():
    # dt = t_c / 200: the explicit integrator resolves the contact
    cfg = dict(po.DEM_REF, gravity=0.0, restitution=0.5, dt=2.5e-6)
    o = po.OracleDEM((0.1, 0.1, 0.1), periodic=(1, 1, 0), cfg=cfg)
    r, v = 0.0014, 0.2
    o.add_sphere(1, r, (0.05 - 1.5 * r, 0.05, 0.05), (v, 0.0, 0.0))
    o.add_sphere(2, r, (0.05 + 1.5 * r, 0.05, 0.05), (-v, 0.0, 0.0))
    o.setup()
    o.run(0, 2000)
    vel = o.get("linear_velocity")
    uid = o.get("uid")
    v1 = vel[np.where(uid == 1)[0][0]][0]
    v2 = vel[np.where(uid == 2)[0][0]][0]
    e = (v2 - v1) / (2 * v)
    # the spring-dashpot of examples/dem.py is built from e: a linear
    # oscillator with damping ratio -ln e / sqrt(pi^2 + ln^2 e) that separates
    # after t_c with relative speed e (continuous time)
    assert abs(e - 0.5) < 0.01
    assert abs(v1 + v2) < 1e-12  # momentum of two equal spheres is conserved


# ---------------------------------------------------------------- the settling workload (fixed k_n, jittered grid)

C3 = dict(po.DEM_REF, gravity=9.81, rho_f=0.0, rho_p=2500.0, dt=1e-6, mu=0.5, restitution=0.6)


def test_mt19937_64_check_value():
    # [rand.predef]: the 10000th consecutive output of a default-constructed
    # std::mt19937_64 is 9981545732273789042
    assert po.mt64_nth(5489, 10000) == 9981545732273789042


def test_jittered_grid_places_non_overlapping_spheres():
    o = po.OracleDEM((0.05, 0.05, 0.04), cap=8192, cfg=C3, stiffness_norm=1e4)
    n = o.jittered_grid((20, 20, 15), (0.00125, 0.00125, 0.0012), (0.0025, 0.0025, 0.0024), 1e-4, 0.9e-3, 1.1e-3)
    assert n == 6000
    x, r = o.get("position"), o.get("radius")
    assert r.min() >= 0.9e-3 and r.max() <= 1.1e-3 and r.std() > 0.05e-3
    site = np.stack(np.meshgrid(np.arange(15), np.arange(20), np.arange(20), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    grid = np.array([0.00125, 0.00125, 0.0012]) + site * np.array([0.0025, 0.0025, 0.0024])
    off = x - grid
    assert np.all(np.abs(off) <= 1e-4) and np.abs(off).max() > 0.9e-4  # jitter used, within bounds
    assert np.array_equal(o.get("uid"), np.arange(1, 6001))
    # no two spheres overlap, across the periodic x/y boundaries too
    d = x[:, None, :] - x[None, :, :]
    d[..., :2] -= 0.05 * np.round(d[..., :2] / 0.05)
    dist = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(dist, np.inf)
    assert np.all(dist > r[:, None] + r[None, :])


def test_fixed_stiffness_equals_collision_time_form():
    """A fixed k_n with the restitution-preserving damping is the collision-time
    model at the t_c that gives the same k_n for the pair's effective mass:
    two equal spheres colliding head-on follow the same trajectory (to
    rounding), and separate with the relative speed e (dt = t_c / 400: the
    explicit integrator resolves the contact)."""
    r, v, kn = 1.0e-3, 0.2, 1e4
    m = 4.0 / 3.0 * math.pi * r ** 3 * 2500.0
    meff = 0.5 * m
    ln_e = math.log(0.6)
    tc = math.sqrt(meff * (math.pi ** 2 + ln_e ** 2) / kn)
    runs = []
    for fixed in (True, False):
        cfg = dict(C3, gravity=0.0, t_c=tc, dt=tc / 400)
        o = po.OracleDEM((0.1, 0.1, 0.1), periodic=(1, 1, 0), cfg=cfg, stiffness_norm=kn if fixed else 0.0)
        o.add_sphere(1, r, (0.05 - 1.02 * r, 0.05, 0.05), (v, 0.0, 0.0))
        o.add_sphere(2, r, (0.05 + 1.02 * r, 0.05, 0.05), (-v, 0.0, 0.0))
        o.setup()
        o.run(0, 1600)
        uid = o.get("uid")
        runs.append(o.get("linear_velocity")[np.argsort(uid)])
    assert np.allclose(runs[0], runs[1], rtol=1e-12, atol=1e-15)
    e = (runs[0][1][0] - runs[0][0][0]) / (2 * v)
    assert abs(e - 0.6) < 0.01


def test_fixed_stiffness_resting_penetration():
    o = po.OracleDEM((0.1, 0.1, 0.1), periodic=(1, 1, 0), cfg=C3, stiffness_norm=1e4)
    r = 1.0e-3
    o.add_sphere(1, r, (0.05, 0.05, r), (0.0, 0.0, 0.0))
    o.add_halfspace(100000, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    o.setup()
    o.run(0, 20000)  # about 40 oscillation periods of sqrt(m / k_n)
    m = o.get("mass")[0]
    w = 9.81 * m
    assert abs(o.get("force")[0][2]) < 1e-8 * w
    delta = r - o.get("position")[0][2]
    assert abs(delta - w / 1e4) < 1e-6 * (w / 1e4)
