"""Host-side driver of the LJ molecular-dynamics hot path on one MI355X.

``LJSystem`` owns one C-ABI context (include/pairs_hip.h).  Its methods are the
modules of the reference's generated program for examples/md.py, under the
reference's own names, in the order of its timestep
(src/pairs/sim/simulation.py:382-452, src/pairs/sim/timestep.py:9-72):

    initial_integrate        (pre-step, skipped at ts == 0)
    exchange                 (rebuild steps)
    borders | synchronize    (rebuild | other steps)
    build_cell_lists         (rebuild steps; includes partition_cell_lists)
    build_neighbor_lists     (rebuild steps)
    lennard_jones            (every step; includes reset_volatile_properties)
    final_integrate          (skipped at ts == 0)
    compute_thermo           (thermo steps)

``timestep(ts)`` runs exactly one iteration; ``run`` hands whole ranges to the
fused, graph-replayed form.
"""
import ctypes
import math

import numpy as np

from . import native as nat


def fcc_box_length(n, rho):
    """Box edge as src/pairs/sim/copper_fcc_lattice.py:19-22 computes it."""
    return n * math.pow((4.0 / rho), (1.0 / 3.0))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class LJSystem:
    """One periodic LJ system (one domain) resident on the GPU."""

    def __init__(self, box, *, particle_capacity=800000, neighbor_capacity=100, send_capacity=200000,
                 ncells_capacity=100000, cutoff=2.5, skin=0.3, dt=0.005, half=False, ntypes=1, epsilon=None,
                 sigma6=None, reneighbor_every=20, thermo_every=100, periodic=(True, True, True)):
        L = nat.lib()
        self.ntypes = ntypes
        eps = np.ones(ntypes * ntypes) if epsilon is None else np.asarray(epsilon, np.float64)
        sg6 = np.ones(ntypes * ntypes) if sigma6 is None else np.asarray(sigma6, np.float64)
        self._tabs = (np.ascontiguousarray(eps, np.float64), np.ascontiguousarray(sg6, np.float64))
        cfg = nat.MDConfig()
        cfg.particle_capacity = particle_capacity
        cfg.neighbor_capacity = neighbor_capacity
        cfg.send_capacity = send_capacity
        cfg.ncells_capacity = ncells_capacity
        for d in range(3):
            cfg.box_lo[d] = box[d][0]
            cfg.box_hi[d] = box[d][1]
            cfg.periodic[d] = int(bool(periodic[d]))
        # the reference builds lists and cells at cutoff + skin (examples/md.py:59)
        cfg.cell_spacing = cutoff + skin
        cfg.cutoff = cutoff
        cfg.dt = dt
        cfg.half = int(bool(half))
        cfg.ntypes = ntypes
        cfg.epsilon = _dp(self._tabs[0])
        cfg.sigma6 = _dp(self._tabs[1])
        cfg.reneighbor_every = reneighbor_every
        cfg.thermo_every = thermo_every
        cfg.stream = None
        self.cfg = cfg
        self.half = bool(half)
        self.box = [tuple(b) for b in box]
        h = ctypes.c_void_p()
        code = L.pairs_hip_create(ctypes.byref(h), ctypes.byref(cfg))
        if code != 0:
            raise nat.PairsHipError(f"pairs_hip_create failed ({code})")
        self.h = h
        self.L = L

    # ------------------------------------------------------------ lifetime
    @classmethod
    def attach(cls, handle, cfg, box, owner):
        """A view of a context owned elsewhere (a batch replica); never destroys it."""
        sim = cls.__new__(cls)
        sim.L = nat.lib()
        sim.h = handle
        sim.cfg = cfg
        sim.ntypes = cfg.ntypes
        sim.half = bool(cfg.half)
        sim.box = [tuple(b) for b in box]
        sim._owner = owner
        return sim

    def close(self):
        if getattr(self, "_owner", None) is not None:
            self.h = None
            return
        if getattr(self, "h", None):
            self.L.pairs_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        code = getattr(self.L, name)(self.h, *args)
        return nat.check(self.h, code, name)

    def _retry(self, name, *args):
        """Run a module; on overflow grow the capacity and run it again
        (the generated resize loop, src/pairs/transformations/modules.py:159-205)."""
        for _ in range(8):
            code = self._call(name, *args)
            if code == 0:
                return 0
            need = self.L.pairs_hip_last_required(self.h, code)
            grow = [0, 0, 0]
            grow[{nat.OVERFLOW_PARTICLES: 0, nat.OVERFLOW_NEIGHBORS: 1, nat.OVERFLOW_SEND: 2}[code]] = 2 * need
            self._call("pairs_hip_reserve", *grow)
        raise nat.PairsHipError(f"{name}: capacity growth did not converge")

    # ------------------------------------------------------------ set-up
    def copper_fcc_lattice(self, nx, ny, nz, rho, ntypes=1):
        n = self._call("pairs_hip_copper_fcc_lattice", nx, ny, nz, rho, ntypes)
        return n

    def adjust_thermo(self, temperature):
        self._call("pairs_hip_adjust_thermo", temperature)

    def upload(self, x, v, m=None, types=None, flags=None, uid=None):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        n = len(x)
        m = None if m is None else np.ascontiguousarray(m, np.float64)
        t = None if types is None else np.ascontiguousarray(types, np.int32)
        f = None if flags is None else np.ascontiguousarray(flags, np.int32)
        u = None if uid is None else np.ascontiguousarray(uid, np.int32)
        code = self.L.pairs_hip_upload(self.h, n, _dp(x), _dp(v), _dp(m) if m is not None else None,
                                       _ip(t) if t is not None else None, _ip(f) if f is not None else None,
                                       _ip(u) if u is not None else None)
        if code > 0:
            self._call("pairs_hip_reserve", 2 * n, 0, 0)
            code = self.L.pairs_hip_upload(self.h, n, _dp(x), _dp(v), _dp(m) if m is not None else None,
                                           _ip(t) if t is not None else None, _ip(f) if f is not None else None,
                                           _ip(u) if u is not None else None)
        nat.check(self.h, code, "pairs_hip_upload")

    # ------------------------------------------------------------ modules
    def initial_integrate(self):
        self._call("pairs_hip_vv_initial")

    def exchange(self):
        self._retry("pairs_hip_exchange")

    def borders(self):
        self._retry("pairs_hip_borders")

    def synchronize(self):
        self._call("pairs_hip_sync_ghosts")

    def build_cell_lists(self):
        self._retry("pairs_hip_bin_cells")

    def build_neighbor_lists(self):
        self._retry("pairs_hip_build_neighbors")

    def lennard_jones(self):
        self._call("pairs_hip_lj")

    def final_integrate(self):
        self._call("pairs_hip_vv_final")

    # the older API's modules (examples/lj_onetype.py:5-13)
    def lennard_jones_older(self):
        self._call("pairs_hip_lj_older")

    def euler(self):
        self._call("pairs_hip_euler")

    def compute_thermo(self):
        """(T, P_kinetic, U, W, P) -- the first two as runtime/thermo.hpp:11-51."""
        out = np.zeros(5)
        self._call("pairs_hip_thermo", _dp(out))
        return out

    def energy(self):
        pe = ctypes.c_double()
        w = ctypes.c_double()
        self._call("pairs_hip_energy", ctypes.byref(pe), ctypes.byref(w))
        return pe.value, w.value

    # ------------------------------------------------------------ loop
    def timestep(self, ts, nsteps):
        self._call("pairs_hip_step", ts, nsteps)

    def run(self, ts_begin, ts_end, nsteps, max_thermo=4096):
        out = np.zeros((max_thermo, 5))
        n = self._call("pairs_hip_run", ts_begin, ts_end, nsteps, _dp(out), max_thermo)
        return out[:n]

    def set_graphs(self, enable):
        self._call("pairs_hip_set_graphs", int(bool(enable)))

    def set_speculative_rebuilds(self, enable):
        """Deferred counter checks of run()'s rebuilds (default on)."""
        self._call("pairs_hip_set_speculative_rebuilds", int(bool(enable)))

    def prepare_graphs(self):
        """Capture the step and rebuild graphs of both buffers now (set-up)."""
        self._call("pairs_hip_prepare_graphs")

    def set_force_timing(self, every_n):
        self._call("pairs_hip_set_force_timing", every_n)

    def force_kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        self._call("pairs_hip_force_kernel_time", ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def sync(self):
        self._call("pairs_hip_sync")

    # ------------------------------------------------------------ views
    def count(self, what):
        return self._call("pairs_hip_count", what)

    @property
    def nlocal(self):
        return self.count(nat.NLOCAL)

    @property
    def nghost(self):
        return self.count(nat.NGHOST)

    @property
    def nsend_all(self):
        return self.count(nat.NSEND_ALL)

    @property
    def ncells(self):
        return self.count(nat.NCELLS)

    def _dl(self, what, shape, dtype, n):
        a = np.zeros(shape, dtype)
        if n == 0:
            return a
        got = self._call("pairs_hip_download", what, a.ctypes.data_as(ctypes.c_void_p), n)
        return a[:got]

    def positions(self, ghosts=False):
        n = self.nlocal + (self.nghost if ghosts else 0)
        return self._dl(nat.POSITION, (n, 3), np.float64, n)

    def velocities(self):
        n = self.nlocal
        return self._dl(nat.VELOCITY, (n, 3), np.float64, n)

    def forces(self):
        n = self.nlocal
        return self._dl(nat.FORCE, (n, 3), np.float64, n)

    def types(self, ghosts=False):
        n = self.nlocal + (self.nghost if ghosts else 0)
        return self._dl(nat.TYPE, (n,), np.int32, n)

    def numneighs(self):
        n = self.nlocal
        return self._dl(nat.NUMNEIGHS, (n,), np.int32, n)

    def neighbor_lists(self):
        n = self.nlocal
        ncap = self.count(nat.NEIGHBOR_CAPACITY)
        flat = self._dl(nat.NEIGHBORLISTS, (n, ncap), np.int32, n)
        nn = self.numneighs()
        return [flat[i, :nn[i]].copy() for i in range(n)]

    def particle_cell(self):
        n = self.nlocal + self.nghost
        return self._dl(nat.PARTICLE_CELL, (n,), np.int32, n)

    def cells(self):
        nc = self.ncells
        sizes = self._dl(nat.CELL_SIZES, (nc,), np.int32, nc)
        n = self.nlocal + self.nghost
        flat = self._dl(nat.CELL_PARTICLES, (n,), np.int32, n)
        starts = np.concatenate([[0], np.cumsum(sizes)])
        return sizes, [flat[starts[c]:starts[c + 1]] for c in range(nc)]

    def send_map(self):
        n = self.nsend_all
        return (self._dl(nat.SEND_MAP, (n,), np.int32, n), self._dl(nat.SEND_MULT, (n, 3), np.int32, n))

    def nsend(self):
        return self._dl(nat.NSEND, (6,), np.int32, 6)

    def atom_energies(self):
        self.energy()
        n = self.nlocal
        return self._dl(nat.ATOM_ENERGY, (n,), np.float64, n)


def fcc_system(n, rho=0.8442, temp=1.44, *, half=False, ntypes=1, **kw):
    """The set-up of examples/md.py's generated main(): domain, copper_fcc_lattice,
    adjust_thermo (runtime/copper_fcc_lattice.hpp, runtime/thermo.hpp)."""
    L = fcc_box_length(n, rho)
    natoms = 4 * n ** 3
    kw.setdefault("particle_capacity", max(800000, 4 * natoms))
    kw.setdefault("send_capacity", max(200000, 2 * natoms))
    sim = LJSystem([(0.0, L)] * 3, half=half, ntypes=ntypes, **kw)
    sim.copper_fcc_lattice(n, n, n, rho, ntypes)
    sim.adjust_thermo(temp)
    return sim
