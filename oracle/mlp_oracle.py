"""FP64 CPU restatement of the radial-descriptor MLP potential -- test
infrastructure.

Only tests/ (and bench.py's cpu_baseline leg, if it ever times the MLP) use
this module; the product path (pairs_amd/) never imports it.

The reference has no machine-learned potential, so there is no reference
function to follow: this is the build-defined model of include/pairs_hip_mlp.h
written out in plain numpy, in float64, with its own backpropagation.  It is
pinned by mathematics, not by golden vectors (tests/test_oracle_mlp.py):
  * dE/dG from the hand-written backward pass == central finite differences of
    the forward pass;
  * forces from the pair formula == -(E(x + h) - E(x - h)) / 2h of the total
    energy, per particle and component;
  * descriptors of a two-particle system == the closed form g_k(r) fc(r), with
    fc(2.5) = 0 and g_k(mu_k) = 1;
  * a network whose output layer is zero returns b3 (scaled) and zero gradient.

Descriptor: G_k(i) = sum_{j != i, r_ij < 2.5} exp(-(r_ij - mu_k)^2 / (2 w^2)) fc(r_ij),
mu_k = 0.8 + k * 1.7 / 15, w = 0.15, fc(r) = (cos(pi r / 2.5) + 1) / 2.
Network: y = w3 . silu(W2 silu(W1 s + b1) + b2) + b3 on s = (G - g_mean) * g_istd;
E_i = e_mean + e_std * y.
"""
import numpy as np

NDESC, HIDDEN = 16, 64
MU = 0.8 + np.arange(NDESC) * (1.7 / 15.0)
WIDTH = 0.15
RCUT = 2.5
NPARAMS = HIDDEN * NDESC + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1  # 5313
NSCALE = 34


def unpack(p):
    """Flat parameter block -> (W1[64,16], b1, W2[64,64], b2, w3, b3)."""
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + HIDDEN * NDESC].reshape(HIDDEN, NDESC); o += HIDDEN * NDESC
    b1 = p[o:o + HIDDEN]; o += HIDDEN
    W2 = p[o:o + HIDDEN * HIDDEN].reshape(HIDDEN, HIDDEN); o += HIDDEN * HIDDEN
    b2 = p[o:o + HIDDEN]; o += HIDDEN
    w3 = p[o:o + HIDDEN]; o += HIDDEN
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def xavier(seed):
    """Random parameters of the same shapes (numpy generator, not the library's)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS)
    a1, a2, a3 = np.sqrt(6 / 80), np.sqrt(6 / 128), np.sqrt(6 / 65)
    p[:1024] = rng.uniform(-a1, a1, 1024)
    p[1024:1088] = rng.uniform(-0.1, 0.1, 64)
    p[1088:5184] = rng.uniform(-a2, a2, 4096)
    p[5184:5248] = rng.uniform(-0.1, 0.1, 64)
    p[5248:5312] = rng.uniform(-a3, a3, 64)
    p[5312] = 0.3
    return p


def fc(r):
    return 0.5 * (np.cos(np.pi * r / RCUT) + 1.0)


def dfc(r):
    return -0.5 * (np.pi / RCUT) * np.sin(np.pi * r / RCUT)


def basis(r):
    """phi_k(r) = gaussian_k(r) fc(r) and dphi_k/dr, shape [..., 16]."""
    r = np.asarray(r, np.float64)[..., None]
    t = r - MU
    g = np.exp(-t * t / (2 * WIDTH * WIDTH))
    f, df = fc(r), dfc(r)
    return g * f, g * (-t / (WIDTH * WIDTH) * f + df)


def silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s * (1.0 + z * (1.0 - s))


def forward(p, sc, G):
    """E [n] and dE/dG [n, 16] (raw descriptor units) of descriptor rows G."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    sc = np.asarray(sc, np.float64)
    gm, gi, em, es = sc[0:16], sc[16:32], sc[32], sc[33]
    s = (np.asarray(G, np.float64) - gm) * gi
    h1, d1 = silu(s @ W1.T + b1)
    h2, d2 = silu(h1 @ W2.T + b2)
    y = h2 @ w3 + b3
    g2 = d2 * w3                 # dy/dz2
    g1 = (g2 @ W2) * d1          # dy/dz1
    gs = g1 @ W1                 # dy/ds
    return em + es * y, es * gs * gi


def bf16(a):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split(a):
    """a ~ hi + lo, both bfloat16 (the device's two-part operands)."""
    hi = bf16(a)
    return hi, bf16(np.asarray(a, np.float32).astype(np.float64) - hi)


def forward_bf16(p, sc, G):
    """forward() with the device's rounding points: the first layer's input and
    weights as hi + lo bfloat16 pairs (the lo * lo term dropped), the second
    layer's weights and input activations as single bfloat16, the backward
    signals likewise (products exact, sums in float64 here, float32 on the
    device)."""
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, np.float32).astype(np.float64))
    sc = np.asarray(sc, np.float32).astype(np.float64)
    gm, gi, em, es = sc[0:16], sc[16:32], sc[32], sc[33]
    W1h, W1l = split(W1)
    W2b = bf16(W2)
    sh, sl = split((np.asarray(G, np.float32).astype(np.float64) - gm) * gi)
    h1, d1 = silu(sh @ W1h.T + sl @ W1h.T + sh @ W1l.T + b1)
    h1b = bf16(h1)
    h2, d2 = silu(h1b @ W2b.T + b2)
    y = h2 @ w3 + b3
    g2 = bf16(d2 * w3)
    g1h, g1l = split((g2 @ W2b) * d1)
    gs = g1h @ W1h + g1l @ W1h + g1h @ W1l
    return em + es * y, es * gs * gi


def pairs_within(x, box_len, rc=RCUT):
    """All ordered pairs (i, j), i != j, closer than rc under the minimum image
    (box edge > 2 rc): indices, separation vectors x_i - x_j and distances."""
    x = np.asarray(x, np.float64)
    d = x[:, None, :] - x[None, :, :]
    if box_len is not None:
        L = np.asarray(box_len, np.float64)
        d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(r, np.inf)
    i, j = np.nonzero(r < rc)
    return i, j, d[i, j], r[i, j]


def descriptors(x, box_len):
    n = len(x)
    i, j, _, r = pairs_within(x, box_len)
    phi, _ = basis(r)
    G = np.zeros((n, NDESC))
    np.add.at(G, i, phi)
    return G


def energy(p, sc, x, box_len):
    E, _ = forward(p, sc, descriptors(x, box_len))
    return E.sum()


def forces(p, sc, x, box_len, device_rounding=False):
    """F_i = -sum_j (c_i(r) + c_j(r)) (x_i - x_j) / r, c_a(r) = sum_k D_ak dphi_k(r);
    with device_rounding the network is forward_bf16."""
    n = len(x)
    G = descriptors(x, box_len)
    E, D = (forward_bf16 if device_rounding else forward)(p, sc, G)
    i, j, d, r = pairs_within(x, box_len)
    _, dphi = basis(r)
    c = ((D[i] + D[j]) * dphi).sum(-1)
    F = np.zeros((n, 3))
    np.add.at(F, i, -(c / r)[:, None] * d)
    return F, E


def virial(p, sc, x, box_len):
    """Sum over pairs of r . F_pair (each unordered pair once)."""
    G = descriptors(x, box_len)
    _, D = forward(p, sc, G)
    i, j, d, r = pairs_within(x, box_len)
    _, dphi = basis(r)
    c = ((D[i] + D[j]) * dphi).sum(-1)
    return 0.5 * (-(c * r)).sum()


def lj_atom_energy(x, box_len, rc=RCUT):
    """Per-atom LJ energy 0.5 * sum_j 4((1/r)^12 - (1/r)^6), r < rc (the labels)."""
    i, _, _, r = pairs_within(x, box_len, rc)
    u = 4.0 * (r ** -12 - r ** -6)
    E = np.zeros(len(x))
    np.add.at(E, i, 0.5 * u)
    return E


def lj_forces(x, box_len, rc=RCUT):
    """LJ forces sum_j 48 (r^-14 - r^-7 / 2) (x_i - x_j) over r < rc (the force labels)."""
    i, _, d, r = pairs_within(x, box_len, rc)
    s = 48.0 * (r ** -14 - 0.5 * r ** -8)
    F = np.zeros((len(x), 3))
    np.add.at(F, i, s[:, None] * d)
    return F


# ---------------------------------------------------------------- training gradients

def silu2(z):
    """SiLU h(z) = z sigma(z), h'(z) and h''(z) = sigma (1 - sigma) (2 + z (1 - 2 sigma))."""
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s * (1.0 + z * (1.0 - s)), s * (1.0 - s) * (2.0 + z * (1.0 - 2.0 * s))


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]])


def param_grad(p, sc, G, eps, u):
    """Gradient over the parameters of sum_a [eps_a y_a + (dy_a/ds) . u_a]
    with eps [n] and u [n, 16] (directions in standardised-descriptor space)
    held fixed: forward, tangent along u, then reverse mode through both."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    sc = np.asarray(sc, np.float64)
    gm, gi = sc[0:16], sc[16:32]
    s = (np.asarray(G, np.float64) - gm) * gi
    u = np.asarray(u, np.float64)
    eps = np.asarray(eps, np.float64)
    z1 = s @ W1.T + b1
    h1, d1, e1 = silu2(z1)
    z2 = h1 @ W2.T + b2
    h2, d2, e2 = silu2(z2)
    # tangent: dz1 = W1 u, dh1 = h1' dz1, dz2 = W2 dh1, dh2 = h2' dz2 (dy . u = w3 . dh2)
    t1 = u @ W1.T
    th1 = d1 * t1
    t2 = th1 @ W2.T
    th2 = d2 * t2
    # reverse
    bt2 = w3 * d2                               # d/d(dz2)
    bz2 = eps[:, None] * w3 * d2 + w3 * e2 * t2  # d/dz2
    bh1 = bz2 @ W2                              # d/dh1
    bth1 = bt2 @ W2                             # d/d(dh1)
    bt1 = bth1 * d1                             # d/d(dz1)
    bz1 = bh1 * d1 + bth1 * e1 * t1             # d/dz1
    gW2 = bz2.T @ h1 + bt2.T @ th1
    gW1 = bz1.T @ s + bt1.T @ u
    return pack(gW1, bz1.sum(0), gW2, bz2.sum(0), (eps[:, None] * h2 + th2).sum(0), eps.sum())


def energy_loss_grad(p, sc, G, label):
    """L = mean_b (y_b - (label_b - e_mean) / e_std)^2 over rows G and dL/dp."""
    sc = np.asarray(sc, np.float64)
    em, es = sc[32], sc[33]
    E, _ = forward(p, sc, G)
    err = (E - em) / es - (np.asarray(label, np.float64) - em) / es
    n = len(err)
    return (err * err).mean(), param_grad(p, sc, G, 2.0 * err / n, np.zeros((n, NDESC)))


def config_loss(p, sc, x, box_len, E_label, F_label, w_energy, w_force):
    """Loss of one periodic configuration:
    w_energy * mean_a (y_a - (E*_a - e_mean) / e_std)^2
      + w_force * sum_i |F_i - F*_i|^2 / sum_i |F*_i|^2."""
    sc = np.asarray(sc, np.float64)
    em, es = sc[32], sc[33]
    F, E = forces(p, sc, x, box_len)
    t = (E - np.asarray(E_label, np.float64)) / es
    Fl = np.asarray(F_label, np.float64)
    le = (t * t).mean()
    lf = ((F - Fl) ** 2).sum() / (Fl ** 2).sum()
    return w_energy * le + w_force * lf, le, lf


def config_loss_grad(p, sc, x, box_len, E_label, F_label, w_energy, w_force):
    """config_loss and its gradient over the parameters.  With rho_i = dL/dF_i
    and w_aj,k = dphi_k(r_aj) (x_a - x_j) / r_aj, dL/dD_a,k =
    -sum_j w_aj,k . (rho_a - rho_j) (F_i = -sum_j w_ij . (D_i + D_j)); the
    network then sees the direction u_a = e_std g_istd dL/dD_a in standardised
    space (D_a = e_std g_istd dy_a/ds)."""
    sc = np.asarray(sc, np.float64)
    em, es, gi = sc[32], sc[33], sc[16:32]
    n = len(x)
    G = descriptors(x, box_len)
    F, E = forces(p, sc, x, box_len)
    Fl = np.asarray(F_label, np.float64)
    S = (Fl ** 2).sum()
    t = (E - np.asarray(E_label, np.float64)) / es
    le = (t * t).mean()
    lf = ((F - Fl) ** 2).sum() / S
    rho = 2.0 * w_force * (F - Fl) / S
    i, j, d, r = pairs_within(x, box_len)
    _, dphi = basis(r)
    proj = ((d / r[:, None]) * (rho[i] - rho[j])).sum(-1)
    gD = np.zeros((n, NDESC))
    np.add.at(gD, i, -dphi * proj[:, None])
    u = es * gi * gD
    eps = 2.0 * w_energy * t / n
    return w_energy * le + w_force * lf, le, lf, param_grad(p, sc, G, eps, u)
