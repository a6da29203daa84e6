"""Pins the FP64 MLP restatement (oracle/mlp_oracle.py) by mathematics: the
model has no reference counterpart, so its checks are closed forms and finite
differences, each chosen so that a dropped term, a wrong sign or a transposed
weight fails."""
import numpy as np

from oracle import mlp_oracle as mo


def _scale(seed=3):
    rng = np.random.default_rng(seed)
    sc = np.zeros(mo.NSCALE)
    sc[0:16] = rng.uniform(0.0, 3.0, 16)
    sc[16:32] = rng.uniform(0.3, 2.0, 16)
    sc[32], sc[33] = -5.5, 0.7
    return sc


def _cluster(n=40, L=7.0, seed=5):
    """particles on a jittered grid (no pair closer than about 0.9)"""
    rng = np.random.default_rng(seed)
    k = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    return (g + 0.5) * (L / k) + rng.uniform(-0.12, 0.12, (n, 3)), L


def test_pair_descriptor_closed_form():
    r = 1.37
    x = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
    G = mo.descriptors(x, None)
    t = r - (0.8 + np.arange(16) * 1.7 / 15)
    expect = np.exp(-t * t / (2 * 0.15 ** 2)) * 0.5 * (np.cos(np.pi * r / 2.5) + 1)
    assert np.allclose(G[0], expect, rtol=1e-14, atol=0) and np.allclose(G[1], expect, rtol=1e-14, atol=0)
    assert mo.fc(2.5) < 1e-15 and abs(mo.fc(0.0) - 1.0) < 1e-15
    phi, _ = mo.basis(mo.MU[7])
    assert abs(phi[7] - mo.fc(mo.MU[7])) < 1e-15
    # beyond the cutoff nothing counts
    x[1, 0] = 1.0 + 2.5001
    assert np.all(mo.descriptors(x, None) == 0.0)


def test_basis_derivative_matches_finite_difference():
    r = np.linspace(0.85, 2.45, 33)
    h = 1e-6
    _, d = mo.basis(r)
    fp, _ = mo.basis(r + h)
    fm, _ = mo.basis(r - h)
    assert np.allclose(d, (fp - fm) / (2 * h), rtol=1e-7, atol=1e-9)


def test_backward_matches_finite_difference_on_G():
    p, sc = mo.xavier(11), _scale()
    rng = np.random.default_rng(2)
    G = rng.uniform(0.0, 4.0, (5, 16))
    E, D = mo.forward(p, sc, G)
    h = 1e-6
    for k in range(16):
        Gp, Gm = G.copy(), G.copy()
        Gp[:, k] += h
        Gm[:, k] -= h
        fd = (mo.forward(p, sc, Gp)[0] - mo.forward(p, sc, Gm)[0]) / (2 * h)
        assert np.allclose(D[:, k], fd, rtol=1e-6, atol=1e-9), k


def test_zero_output_layer():
    p, sc = mo.xavier(1), _scale()
    p[5248:5312] = 0.0
    E, D = mo.forward(p, sc, np.ones((3, 16)))
    assert np.allclose(E, sc[32] + sc[33] * p[5312]) and np.all(D == 0.0)


def test_forces_are_minus_energy_gradient():
    x, L = _cluster()
    p, sc = mo.xavier(7), _scale()
    F, _ = mo.forces(p, sc, x, L)
    h = 1e-5
    for i in (0, 3, 17, 39):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            fd = -(mo.energy(p, sc, xp, L) - mo.energy(p, sc, xm, L)) / (2 * h)
            assert abs(F[i, d] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, d, F[i, d], fd)
    # translation invariance: the forces sum to zero
    assert np.allclose(F.sum(0), 0.0, atol=1e-9)


def test_lj_atom_energy_pair():
    r = 1.2
    x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    E = mo.lj_atom_energy(x, None)
    u = 4 * (r ** -12 - r ** -6)
    assert np.allclose(E, [0.5 * u, 0.5 * u], rtol=1e-15)


def test_bf16_rounding_and_budget():
    # nearest-even at the tie, exact values kept
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.14159])
    assert np.array_equal(mo.bf16(x), [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 3.140625])
    # the device's rounding points cost well under 2 % of the signal
    p, sc = mo.xavier(7), _scale()
    G = np.random.default_rng(1).uniform(0, 4, (500, 16))
    E, D = mo.forward(p, sc, G)
    Eb, Db = mo.forward_bf16(p, sc, G)
    assert np.sqrt(np.mean((E - Eb) ** 2)) < 0.01 * E.std()
    assert np.sqrt(np.mean((D - Db) ** 2) / np.mean(D ** 2)) < 0.01


def test_virial_is_minus_scaling_derivative():
    """W = sum over pairs of r_ij . F_ij equals -dE/dlambda at lambda = 1 for
    the uniformly scaled system (x -> lambda x, box -> lambda L): the energy
    depends on the pair distances only, so dE/dlambda = sum_pairs (dE/dr) r.
    A wrong sign, a doubled pair or a dropped c_j term fails it."""
    x, L = _cluster()
    p, sc = mo.xavier(7), _scale()
    W = mo.virial(p, sc, x, L)
    h = 1e-6
    ep = mo.energy(p, sc, x * (1 + h), L * (1 + h))
    em = mo.energy(p, sc, x * (1 - h), L * (1 - h))
    dEdl = (ep - em) / (2 * h)
    assert abs(W + dEdl) <= 1e-6 * max(1.0, abs(W)), (W, -dEdl)
    assert abs(W) > 1e-3  # a non-trivial check


def test_silu_derivatives_match_finite_difference():
    z = np.linspace(-6.0, 6.0, 49)
    h = 1e-6
    _, d1, d2 = mo.silu2(z)
    hp, dp, _ = mo.silu2(z + h)
    hm, dm, _ = mo.silu2(z - h)
    assert np.allclose(d1, (hp - hm) / (2 * h), rtol=1e-8, atol=1e-10)
    assert np.allclose(d2, (dp - dm) / (2 * h), rtol=1e-7, atol=1e-9)


# one parameter of each block at its ends and inside (W1, b1, W2, b2, w3, b3):
# a transposed weight block or a dropped bias term fails one of them
_PARAM_PROBES = [0, 17, 1023, 1024, 1050, 1087, 1088, 1100, 3000, 5183, 5184, 5230, 5247, 5248, 5290, 5311, 5312]


def _fd(loss, p, idx, h=1e-6):
    out = []
    for k in idx:
        pp = p.copy()
        pp[k] += h
        a = loss(pp)
        pp[k] -= 2 * h
        out.append((a - loss(pp)) / (2 * h))
    return np.array(out)


def test_energy_loss_gradient_matches_finite_difference():
    p, sc = mo.xavier(5), _scale()
    rng = np.random.default_rng(4)
    G = rng.uniform(0.0, 4.0, (37, 16))
    lab = rng.normal(-5.0, 0.8, 37)
    L0, g = mo.energy_loss_grad(p, sc, G, lab)
    assert g.shape == (mo.NPARAMS,)
    fd = _fd(lambda q: mo.energy_loss_grad(q, sc, G, lab)[0], p, _PARAM_PROBES)
    assert np.allclose(g[_PARAM_PROBES], fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())
    v = rng.standard_normal(mo.NPARAMS)
    h = 1e-6
    dd = (mo.energy_loss_grad(p + h * v, sc, G, lab)[0] - mo.energy_loss_grad(p - h * v, sc, G, lab)[0]) / (2 * h)
    assert abs(dd - g @ v) <= 1e-6 * abs(dd)


def test_lj_force_labels_are_minus_energy_gradient():
    x, L = _cluster(30, 6.2)
    F = mo.lj_forces(x, L)
    h = 1e-6
    for a in (0, 7, 29):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, c] += h
            xm[a, c] -= h
            dE = (mo.lj_atom_energy(xp, L).sum() - mo.lj_atom_energy(xm, L).sum()) / (2 * h)
            assert abs(F[a, c] + dE) <= 1e-6 * max(1.0, abs(F[a, c]))
    assert np.abs(F.sum(0)).max() < 1e-10  # Newton's third law


def test_config_loss_gradient_matches_finite_difference():
    """Energy term, force term and both: the force term's gradient goes
    through the input gradient of the network (a second derivative) and the
    pair back-projection of the force residuals, including through particles'
    periodic images."""
    x, L = _cluster(40, 5.6)
    p, sc = mo.xavier(5), _scale()
    G = mo.descriptors(x, L)
    sc[0:16], sc[16:32] = G.mean(0), 1.0 / G.std(0)
    El, Fl = mo.lj_atom_energy(x, L), mo.lj_forces(x, L)
    rng = np.random.default_rng(9)
    for wE, wF in ((1.0, 0.0), (0.0, 1.0), (0.7, 1.3)):
        Lo, le, lf, g = mo.config_loss_grad(p, sc, x, L, El, Fl, wE, wF)
        assert abs(Lo - mo.config_loss(p, sc, x, L, El, Fl, wE, wF)[0]) <= 1e-12 * Lo
        fd = _fd(lambda q: mo.config_loss(q, sc, x, L, El, Fl, wE, wF)[0], p, _PARAM_PROBES)
        assert np.allclose(g[_PARAM_PROBES], fd, rtol=1e-5, atol=1e-8 * np.abs(fd).max()), (wE, wF)
        v = rng.standard_normal(mo.NPARAMS)
        h = 1e-6
        dd = (mo.config_loss(p + h * v, sc, x, L, El, Fl, wE, wF)[0]
              - mo.config_loss(p - h * v, sc, x, L, El, Fl, wE, wF)[0]) / (2 * h)
        assert abs(dd - g @ v) <= 1e-6 * abs(dd), (wE, wF)
