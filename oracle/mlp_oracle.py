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
