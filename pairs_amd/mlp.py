"""Radial-descriptor MLP potential for the LJ system (include/pairs_hip_mlp.h).

A build-defined force field with no reference counterpart: 16 radial Gaussian
descriptors under a cosine cutoff at 2.5, a 16-64-64-1 SiLU network on MFMA
(BF16 weights and activations, FP32 accumulation), forces by a hand-written
input-gradient kernel, FP64 integration.  It runs inside the same timestep as
``lennard_jones`` (same lists, ghosts, integrator, graphs) once attached to a
system, and is attached only if its forces on the system's current state are
within a relative RMS tolerance (default 2 %) of the LJ forces -- otherwise
the system keeps LJ.

Training data come from LJ runs, in two forms.  Energy rows: each snapshot
appends every atom's descriptor and its per-atom LJ energy to the model's
device-resident training set, and ``MLPModel.train`` runs Adam on minibatches
of rows.  Configurations with forces: ``train_on_configurations`` advances
plain-LJ systems and, every ``every`` steps, takes the gradient of each
system's configuration loss (standardised energy error plus relative squared
force error against LJ, the force term differentiated through the network's
input gradient) and one Adam step on their mean -- all on the GPU.
"""
import ctypes

import numpy as np

from . import native as nat

NDESC, HIDDEN, NPARAMS, NSCALE = 16, 64, 5313, 34

_P, _I, _D, _U64, _F = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, ctypes.c_float
_DP, _FP = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)

nat.register({
    "pairs_hip_mlp_create": (_I, [ctypes.POINTER(_P), _U64, _I]),
    "pairs_hip_mlp_destroy": (_I, [_P]),
    "pairs_hip_mlp_last_error": (ctypes.c_char_p, [_P]),
    "pairs_hip_mlp_get_params": (_I, [_P, _FP, _FP]),
    "pairs_hip_mlp_set_params": (_I, [_P, _FP, _FP]),
    "pairs_hip_mlp_eval": (_I, [_P, _I, _FP, _FP, _FP]),
    "pairs_hip_mlp_descriptors": (_I, [_P, _FP, _I]),
    "pairs_hip_mlp_add_snapshot": (_I, [_P, _P]),
    "pairs_hip_mlp_dataset_size": (_I, [_P]),
    "pairs_hip_mlp_fit_scaling": (_I, [_P, _I]),
    "pairs_hip_mlp_train": (_I, [_P, _I, _I, _I, _F, _U64, _FP]),
    "pairs_hip_mlp_rmse": (_I, [_P, _I, _I, _DP]),
    "pairs_hip_mlp_add_rows": (_I, [_P, _I, _FP, _FP]),
    "pairs_hip_mlp_energy_grad": (_I, [_P, _I, _I, _U64, _I, _FP, _FP]),
    "pairs_hip_mlp_config_grad": (_I, [_P, _P, _D, _D, _FP, _DP]),
    "pairs_hip_mlp_apply_grad": (_I, [_P, _F]),
    "pairs_hip_mlp_forces": (_I, [_P, _P, _DP, _DP]),
    "pairs_hip_mlp_force_error": (_I, [_P, _P, _DP]),
    "pairs_hip_mlp_attach": (_I, [_P, _P, _D, _DP]),
    "pairs_hip_mlp_detach": (_I, [_P]),
    "pairs_hip_mlp_active": (_I, [_P]),
})


def _fp(a):
    return a.ctypes.data_as(_FP)


class MLPModel:
    """Parameters, Adam state and training set of one model (device-resident)."""

    def __init__(self, seed=20241202, dataset_capacity=3_000_000):
        self.L = nat.lib()
        h = ctypes.c_void_p()
        code = self.L.pairs_hip_mlp_create(ctypes.byref(h), seed, dataset_capacity)
        self.h = h
        self._check(code, "pairs_hip_mlp_create")

    def _check(self, code, what):
        if code < 0:
            msg = self.L.pairs_hip_mlp_last_error(self.h) if self.h else b""
            raise nat.PairsHipError(f"{what} failed ({code}): {msg.decode() if msg else ''}")
        return code

    def close(self):
        if getattr(self, "h", None):
            self.L.pairs_hip_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(NPARAMS, np.float32)
        s = np.zeros(NSCALE, np.float32)
        self._check(self.L.pairs_hip_mlp_get_params(self.h, _fp(p), _fp(s)), "get_params")
        return p, s

    def set_params(self, params=None, scale=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        s = None if scale is None else np.ascontiguousarray(scale, np.float32)
        self._check(self.L.pairs_hip_mlp_set_params(self.h, None if p is None else _fp(p),
                                                    None if s is None else _fp(s)), "set_params")

    def eval(self, G):
        """E [n] and dE/dG [n, 16] of descriptor rows."""
        G = np.ascontiguousarray(G, np.float32)
        n = G.shape[0]
        E = np.zeros(n, np.float32)
        D = np.zeros((n, NDESC), np.float32)
        self._check(self.L.pairs_hip_mlp_eval(self.h, n, _fp(G), _fp(E), _fp(D)), "eval")
        return E, D

    def add_snapshot(self, sim):
        code = self.L.pairs_hip_mlp_add_snapshot(self.h, sim.h)
        return nat.check(sim.h, code, "pairs_hip_mlp_add_snapshot")

    @property
    def dataset_size(self):
        return self.L.pairs_hip_mlp_dataset_size(self.h)

    def fit_scaling(self, rows=None):
        rows = self.dataset_size if rows is None else rows
        self._check(self.L.pairs_hip_mlp_fit_scaling(self.h, rows), "fit_scaling")

    def train(self, steps, *, batch=4096, lr=2e-3, seed=7, rows=None):
        """Adam steps on minibatches of training rows [0, rows); returns the loss
        (mean squared error of the standardised energy) of every step."""
        rows = self.dataset_size if rows is None else rows
        loss = np.zeros(max(steps, 1), np.float32)
        self._check(self.L.pairs_hip_mlp_train(self.h, rows, steps, batch, lr, seed, _fp(loss)), "train")
        return loss[:steps]

    def add_rows(self, G, labels):
        """Append descriptor rows [n, 16] with per-atom energy labels [n]."""
        G = np.ascontiguousarray(G, np.float32)
        lab = np.ascontiguousarray(labels, np.float32)
        if G.ndim != 2 or G.shape[1] != NDESC or lab.shape != (G.shape[0],):
            raise ValueError("rows are [n, 16] descriptors and [n] labels")
        return self._check(self.L.pairs_hip_mlp_add_rows(self.h, G.shape[0], _fp(G), _fp(lab)), "add_rows")

    def energy_grad(self, step, *, batch=4096, seed=7, rows=None):
        """Gradient [NPARAMS] and loss of train()'s minibatch `step` (no update)."""
        rows = self.dataset_size if rows is None else rows
        g = np.zeros(NPARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.L.pairs_hip_mlp_energy_grad(self.h, rows, batch, seed, step, _fp(g), _fp(loss)),
                    "energy_grad")
        return g, float(loss[0])

    def config_grad(self, sim, w_energy=1.0, w_force=1.0, *, want_grad=True):
        """Gradient of the loss of the system's current configuration (added
        to the accumulator of apply_grad); returns (gradient or None,
        (energy term, force term))."""
        g = np.zeros(NPARAMS, np.float32) if want_grad else None
        loss = np.zeros(2)
        nat.check(sim.h, self.L.pairs_hip_mlp_config_grad(sim.h, self.h, w_energy, w_force,
                                                          None if g is None else _fp(g),
                                                          loss.ctypes.data_as(_DP)), "pairs_hip_mlp_config_grad")
        return g, (float(loss[0]), float(loss[1]))

    def apply_grad(self, lr):
        """One Adam step on the mean of the accumulated configuration gradients."""
        self._check(self.L.pairs_hip_mlp_apply_grad(self.h, lr), "apply_grad")

    def rmse(self, begin=0, end=None):
        end = self.dataset_size if end is None else end
        out = ctypes.c_double()
        self._check(self.L.pairs_hip_mlp_rmse(self.h, begin, end, ctypes.byref(out)), "rmse")
        return out.value


def descriptors(sim):
    """[nlocal, 16] descriptors of the system's current state (reference order)."""
    n = sim.nlocal
    G = np.zeros((n, NDESC), np.float32)
    nat.check(sim.h, sim.L.pairs_hip_mlp_descriptors(sim.h, _fp(G), n), "pairs_hip_mlp_descriptors")
    return G


def forces(sim, model):
    """One MLP force evaluation into the system's force array; (energy, virial)."""
    e, w = ctypes.c_double(), ctypes.c_double()
    nat.check(sim.h, sim.L.pairs_hip_mlp_forces(sim.h, model.h, ctypes.byref(e), ctypes.byref(w)),
              "pairs_hip_mlp_forces")
    return e.value, w.value


def force_error(sim, model):
    """Relative RMS difference of MLP and LJ forces on the current state."""
    r = ctypes.c_double()
    nat.check(sim.h, sim.L.pairs_hip_mlp_force_error(sim.h, model.h, ctypes.byref(r)), "pairs_hip_mlp_force_error")
    return r.value


def attach(sim, model, max_rel_rms=0.02):
    """Use the model for the system's later steps if its force error is within
    max_rel_rms (negative: no check); returns (attached, measured error)."""
    r = ctypes.c_double()
    code = nat.check(sim.h, sim.L.pairs_hip_mlp_attach(sim.h, model.h, max_rel_rms, ctypes.byref(r)),
                     "pairs_hip_mlp_attach")
    return code == 0, r.value


def detach(sim):
    nat.check(sim.h, sim.L.pairs_hip_mlp_detach(sim.h), "pairs_hip_mlp_detach")


def active(sim):
    return bool(sim.L.pairs_hip_mlp_active(sim.h))


def collect_lj_snapshots(model, make_system, temps, *, equilibrate=200, every=50, per_temp=4):
    """Append snapshots of plain-LJ runs to the model's training set: for each
    temperature a fresh system (make_system(T)) runs `equilibrate` steps, then
    a snapshot is taken every `every` steps, `per_temp` times."""
    added = 0
    for T in temps:
        sim = make_system(T)
        total = equilibrate + every * per_temp
        ts = 0
        sim.run(ts, equilibrate, total)
        ts = equilibrate
        for _ in range(per_temp):
            sim.run(ts, ts + every, total)
            ts += every
            added += model.add_snapshot(sim)
        sim.close()
    return added


def train_on_configurations(model, sims, steps, *, every=50, ts0=0, lr=2e-3, lr_end_frac=0.05, w_energy=1.0,
                            w_force=1.0, total_steps=None, record_every=0, temps=None, rescale_every=10):
    """Adam on configurations of plain-LJ runs: before each Adam step every
    system advances `every` timesteps of its own LJ run, then the gradient of
    its configuration loss is accumulated; the step uses their mean.  The
    learning rate decays linearly from lr to lr * lr_end_frac.  With temps
    (one per system), every `rescale_every` Adam steps each system's
    velocities are rescaled to its temperature (adjust_thermo), which keeps
    long NVE runs sampling at that temperature.  Returns the next timestep of
    the systems and, every `record_every` steps (0: never), (step, mean energy
    term, mean force term) over the systems."""
    ts = ts0
    total = total_steps if total_steps is not None else ts0 + every * steps
    hist = []
    for k in range(steps):
        rec = record_every > 0 and k % record_every == 0
        le = lf = 0.0
        for i, sim in enumerate(sims):
            sim.run(ts, ts + every, total)
            if temps is not None and k % rescale_every == rescale_every - 1:
                sim.adjust_thermo(temps[i])
            _, (a, b) = model.config_grad(sim, w_energy, w_force, want_grad=False) if rec else \
                (None, (0.0, 0.0))
            if not rec:
                nat.check(sim.h, model.L.pairs_hip_mlp_config_grad(sim.h, model.h, w_energy, w_force, None, None),
                          "pairs_hip_mlp_config_grad")
            le += a
            lf += b
        ts += every
        frac = k / (steps - 1) if steps > 1 else 0.0
        model.apply_grad(lr * (1.0 - (1.0 - lr_end_frac) * frac))
        if rec:
            hist.append((k, le / len(sims), lf / len(sims)))
    return ts, hist
