"""Batched LJ ensemble: many independent periodic replicas of uneven sizes
advanced together on one GPU (include/pairs_hip_batch.h).

The reference runs one simulation per process; here R replicas, each a full
LJ system with the reference's schedule and semantics, share one fused launch
per timestep (replayed from a hipGraph), and the segment size that sets the
neighbour-tile shape is chosen by an autotuner.  ``run_each`` advances the same
replicas with a per-replica loop of launches for comparison.

Synthetic replicas (``uneven_ensemble``): replica r has N_r atoms on the first
N_r sites of the smallest FCC n^3 lattice that holds them, scaled into a cube
of density rho_r, with zero-momentum Gaussian velocities at temperature T_r
(build-defined inputs; numpy's PCG64 seeded with seed + r).
"""
import ctypes
import math

import numpy as np

from . import native as nat
from .md import LJSystem

_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_DP, _IP = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)

nat.register({
    "pairs_hip_batch_create": (_I, [ctypes.POINTER(_P), _I, ctypes.POINTER(nat.MDConfig)]),
    "pairs_hip_batch_destroy": (_I, [_P]),
    "pairs_hip_batch_last_error": (ctypes.c_char_p, [_P]),
    "pairs_hip_batch_replica": (_P, [_P, _I]),
    "pairs_hip_batch_run": (_I, [_P, _I, _I, _I]),
    "pairs_hip_batch_run_each": (_I, [_P, _I, _I, _I]),
    "pairs_hip_batch_set_home": (_I, [_P, _I]),
    "pairs_hip_batch_set_workgroups": (_I, [_P, _I]),
    "pairs_hip_batch_autotune": (_I, [_P, _I, _I, _I, _I, _IP, _DP, _IP, _IP]),
    "pairs_hip_batch_thermo": (_I, [_P, _DP]),
})


def fcc_sites(natoms, box_length):
    """First natoms sites of the smallest FCC n^3 lattice holding them (cells in
    i, j, k order, 4-site basis), scaled into [0, L)^3."""
    n = 1
    while 4 * n ** 3 < natoms:
        n += 1
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    idx = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    sites = (idx[:, None, :] + basis[None, :, :]).reshape(-1, 3)[:natoms]
    return sites * (box_length / n)


def replica_state(natoms, rho, temp, seed):
    L = (natoms / rho) ** (1.0 / 3.0)
    x = fcc_sites(natoms, L)
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((natoms, 3))
    v -= v.mean(axis=0)
    t = (v * v).sum() / (3 * natoms - 3)
    v *= math.sqrt(temp / t)
    return L, x, v


class BatchEnsemble:
    def __init__(self, sizes, densities, temps, seed=20241202, cutoff=2.5, skin=0.3, dt=0.005, reneighbor_every=20,
                 thermo_every=100, neighbor_capacity=128, particle_capacity=None, send_capacity=None):
        """particle_capacity / send_capacity: None (8 N_r + 4096 each), one
        value for every replica, or one per replica; a replica that outgrows
        one grows it at its next rebuild (x2 the need)."""
        self.L = nat.lib()
        R = len(sizes)
        self.sizes = [int(n) for n in sizes]

        def per_replica(v, r):
            if v is None:
                return 8 * self.sizes[r] + 4096
            return int(v[r]) if hasattr(v, "__len__") else int(v)
        cfgs = (nat.MDConfig * R)()
        self._keep = []
        boxes, states = [], []
        s = cutoff + skin
        for r in range(R):
            L, x, v = replica_state(self.sizes[r], densities[r], temps[r], seed + r)
            boxes.append(L)
            states.append((x, v))
            c = cfgs[r]
            n = self.sizes[r]
            dims = math.ceil((L + 2 * s) / s) + 1
            c.particle_capacity = per_replica(particle_capacity, r)
            c.neighbor_capacity = neighbor_capacity
            c.send_capacity = per_replica(send_capacity, r)
            c.ncells_capacity = dims ** 3 + 64
            for d in range(3):
                c.box_lo[d] = 0.0
                c.box_hi[d] = L
                c.periodic[d] = 1
            c.cell_spacing = s
            c.cutoff = cutoff
            c.dt = dt
            c.half = 0
            c.ntypes = 1
            eps = np.ones(1)
            sg6 = np.ones(1)
            self._keep += [eps, sg6]
            c.epsilon = eps.ctypes.data_as(_DP)
            c.sigma6 = sg6.ctypes.data_as(_DP)
            c.reneighbor_every = reneighbor_every
            c.thermo_every = thermo_every
            c.stream = None
        self.cfgs = cfgs
        self.h = _P()
        code = self.L.pairs_hip_batch_create(ctypes.byref(self.h), R, cfgs)
        if code != 0:
            raise nat.PairsHipError(f"pairs_hip_batch_create failed ({code})")
        self.replicas = []
        for r in range(R):
            hr = self.L.pairs_hip_batch_replica(self.h, r)
            sim = LJSystem.attach(ctypes.c_void_p(hr), cfgs[r], [(0.0, boxes[r])] * 3, self)
            x, v = states[r]
            sim.upload(x, v)
            sim._call("pairs_hip_setup_cells")
            self.replicas.append(sim)
        self.boxes = boxes
        self.initial = states

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            for s in self.replicas:
                s.h = None
            self.L.pairs_hip_batch_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code, what):
        if code < 0 or code > 0:
            msg = self.L.pairs_hip_batch_last_error(self.h)
            raise nat.PairsHipError(f"{what} failed ({code}): {msg.decode() if msg else ''}")

    @property
    def natoms(self):
        return sum(self.sizes)

    def run(self, ts_begin, ts_end, nsteps):
        self._check(self.L.pairs_hip_batch_run(self.h, ts_begin, ts_end, nsteps), "batch_run")

    def run_each(self, ts_begin, ts_end, nsteps):
        self._check(self.L.pairs_hip_batch_run_each(self.h, ts_begin, ts_end, nsteps), "batch_run_each")

    def set_home(self, home):
        self._check(self.L.pairs_hip_batch_set_home(self.h, home), "batch_set_home")

    def set_workgroups(self, workgroups):
        """at most this many workgroups in the persistent force launch (0: all
        the device holds resident)"""
        self._check(self.L.pairs_hip_batch_set_workgroups(self.h, workgroups), "batch_set_workgroups")

    def autotune(self, ts_begin, steps, nsteps, candidates=(32, 48, 64, 96, 160)):
        """-> (chosen home size, {candidate: ms per step}, next timestep)"""
        cand = np.ascontiguousarray(candidates, np.int32)
        ms = np.zeros(len(cand))
        chosen, ts_next = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.pairs_hip_batch_autotune(self.h, ts_begin, steps, nsteps, len(cand),
                                                    cand.ctypes.data_as(_IP), ms.ctypes.data_as(_DP),
                                                    ctypes.byref(chosen), ctypes.byref(ts_next)), "batch_autotune")
        return chosen.value, dict(zip(cand.tolist(), ms.tolist())), ts_next.value

    def thermo(self):
        out = np.zeros((len(self.sizes), 5))
        self._check(self.L.pairs_hip_batch_thermo(self.h, out.ctypes.data_as(_DP)), "batch_thermo")
        return out


def uneven_ensemble(nreplicas=512, nmin=500, nmax=20000, rho=(0.70, 0.90), temp=(0.8, 2.0), seed=20241202, **kw):
    """R replicas with N ~ U[nmin, nmax], rho ~ U[rho], T ~ U[temp]."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, nreplicas)
    dens = rng.uniform(rho[0], rho[1], nreplicas)
    temps = rng.uniform(temp[0], temp[1], nreplicas)
    return BatchEnsemble(sizes, dens, temps, seed=seed, **kw)
