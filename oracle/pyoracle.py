"""ctypes view of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The product path (pairs_amd/) never does.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)

X, V, F, M, TYPE, FLAGS, NUMNEIGHS, NEIGHBORLISTS, PARTICLE_CELL, CELL_SIZES, CELL_PARTICLES, \
    SEND_MAP, SEND_MULT, NSEND, STENCIL, DIM_CELLS, UID = range(17)
NLOCAL, NGHOST, NCELLS, NSEND_ALL, OVERFLOW, NCAP, CCAP, CAP = range(8)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.omd_create.restype = ctypes.c_void_p
        L.omd_create.argtypes = [ctypes.c_int] * 5 + [_D, _D, _I, ctypes.c_double, ctypes.c_double,
                                                       ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                       _D, _D, ctypes.c_int, ctypes.c_int]
        L.omd_destroy.argtypes = [ctypes.c_void_p]
        L.omd_fcc_lattice.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_int]
        L.omd_fcc_lattice.restype = ctypes.c_int
        L.omd_adjust_thermo.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.omd_temperature.argtypes = [ctypes.c_void_p]
        L.omd_temperature.restype = ctypes.c_double
        L.omd_set_particles.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _I, _I]
        for fn in ("omd_setup_cells", "omd_rebuild", "omd_force", "omd_sync"):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
        L.omd_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.omd_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.omd_energy.argtypes = [ctypes.c_void_p, _D, _D, _D]
        L.omd_ptr.restype = ctypes.c_void_p
        L.omd_ptr.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.omd_int.restype = ctypes.c_int
        L.omd_int.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.omd_last_temperature.restype = ctypes.c_double
        L.omd_last_temperature.argtypes = [ctypes.c_void_p]
        L.omd_last_pressure.restype = ctypes.c_double
        L.omd_last_pressure.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def fcc_box_length(n, rho):
    """Box edge exactly as src/pairs/sim/copper_fcc_lattice.py:19-22 computes it."""
    return n * math.pow((4.0 / rho), (1.0 / 3.0))


class OracleMD:
    """One LJ system in the oracle; mirrors the generated program's state."""

    def __init__(self, box, cap, *, ncap=100, ccap=64, send_cap=200000, ncells_cap=100000,
                 spacing=2.8, cutoff=2.5, dt=0.005, half=False, ntypes=1, eps=None, sig6=None,
                 reneighbor_every=20, thermo_every=100, periodic=(1, 1, 1)):
        L = lib()
        lo = np.array([b[0] for b in box], np.float64)
        hi = np.array([b[1] for b in box], np.float64)
        per = np.array(periodic, np.int32)
        eps = np.ones(ntypes * ntypes) if eps is None else np.asarray(eps, np.float64)
        sig6 = np.ones(ntypes * ntypes) if sig6 is None else np.asarray(sig6, np.float64)
        self._keep = (lo, hi, per, eps, sig6)
        self.h = L.omd_create(cap, ncap, ccap, send_cap, ncells_cap, _dp(lo), _dp(hi), _ip(per),
                              spacing, cutoff, dt, int(half), ntypes, _dp(eps), _dp(sig6),
                              reneighbor_every, thermo_every)
        self.cap, self.ncap, self.ccap = cap, ncap, ccap
        self.L = L

    def __del__(self):
        if getattr(self, "h", None):
            self.L.omd_destroy(self.h)
            self.h = None

    # --- set-up
    def fcc_lattice(self, n, rho, ntypes=1):
        return self.L.omd_fcc_lattice(self.h, n, n, n, rho, ntypes)

    def adjust_thermo(self, temp):
        self.L.omd_adjust_thermo(self.h, temp)

    def set_particles(self, x, v, m=None, types=None, flags=None):
        n = len(x)
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        m = np.ones(n) if m is None else np.ascontiguousarray(m, np.float64)
        t = np.zeros(n, np.int32) if types is None else np.ascontiguousarray(types, np.int32)
        fl = np.zeros(n, np.int32) if flags is None else np.ascontiguousarray(flags, np.int32)
        self.L.omd_set_particles(self.h, n, _dp(x), _dp(v), _dp(m), _ip(t), _ip(fl))

    def setup_cells(self):
        self.L.omd_setup_cells(self.h)

    # --- stepping
    def step(self, ts):
        self.L.omd_step(self.h, ts)

    def run(self, ts_begin, ts_end):
        self.L.omd_run(self.h, ts_begin, ts_end)

    def rebuild(self):
        self.L.omd_rebuild(self.h)

    def force(self):
        self.L.omd_force(self.h)

    def sync(self):
        self.L.omd_sync(self.h)

    def temperature(self):
        return self.L.omd_temperature(self.h)

    def last_thermo(self):
        return self.L.omd_last_temperature(self.h), self.L.omd_last_pressure(self.h)

    def energy(self):
        pe = ctypes.c_double()
        w = ctypes.c_double()
        e = np.zeros(self.nlocal)
        self.L.omd_energy(self.h, ctypes.byref(pe), ctypes.byref(w), _dp(e))
        return pe.value, w.value, e

    # --- views (copies)
    def _i(self, which):
        return self.L.omd_int(self.h, which)

    @property
    def nlocal(self):
        return self._i(NLOCAL)

    @property
    def nghost(self):
        return self._i(NGHOST)

    @property
    def ncells(self):
        return self._i(NCELLS)

    @property
    def nsend_all(self):
        return self._i(NSEND_ALL)

    @property
    def overflow(self):
        return self._i(OVERFLOW)

    def _arr(self, which, ctype, n):
        p = self.L.omd_ptr(self.h, which)
        if n == 0:
            return np.zeros(0, np.float64 if ctype is ctypes.c_double else np.int32)
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctype)), shape=(n,)).copy()

    def positions(self, ghosts=False):
        n = self.nlocal + (self.nghost if ghosts else 0)
        return self._arr(X, ctypes.c_double, 3 * n).reshape(n, 3)

    def velocities(self, ghosts=False):
        n = self.nlocal + (self.nghost if ghosts else 0)
        return self._arr(V, ctypes.c_double, 3 * n).reshape(n, 3)

    def forces(self):
        n = self.nlocal
        return self._arr(F, ctypes.c_double, 3 * n).reshape(n, 3)

    def types(self, ghosts=False):
        n = self.nlocal + (self.nghost if ghosts else 0)
        return self._arr(TYPE, ctypes.c_int, n)

    def numneighs(self):
        return self._arr(NUMNEIGHS, ctypes.c_int, self.nlocal)

    def neighbor_lists(self):
        """List of per-particle neighbour index arrays, in list order."""
        nn = self.numneighs()
        flat = self._arr(NEIGHBORLISTS, ctypes.c_int, self.nlocal * self.ncap).reshape(self.nlocal, self.ncap)
        return [flat[i, :nn[i]].copy() for i in range(self.nlocal)]

    def particle_cell(self):
        return self._arr(PARTICLE_CELL, ctypes.c_int, self.nlocal + self.nghost)

    def cells(self):
        nc = self.ncells
        sizes = self._arr(CELL_SIZES, ctypes.c_int, nc)
        parts = self._arr(CELL_PARTICLES, ctypes.c_int, nc * self.ccap).reshape(nc, self.ccap)
        return sizes, [parts[c, :sizes[c]].copy() for c in range(nc)]

    def send_map(self):
        n = self.nsend_all
        return self._arr(SEND_MAP, ctypes.c_int, n), self._arr(SEND_MULT, ctypes.c_int, 3 * n).reshape(n, 3)

    def nsend(self):
        return self._arr(NSEND, ctypes.c_int, 6)

    def stencil(self):
        return self._arr(STENCIL, ctypes.c_int, 27), self._arr(DIM_CELLS, ctypes.c_int, 3)


def fcc_system(n, rho=0.8442, temp=1.44, *, half=False, cap=None, ntypes=1, **kw):
    """The md.py set-up: lattice, adjust_thermo, stencil (generated main())."""
    L = fcc_box_length(n, rho)
    natoms = 4 * n ** 3
    cap = cap or max(4 * natoms + 4096, 8192)
    sim = OracleMD([(0.0, L)] * 3, cap, half=half, ntypes=ntypes, send_cap=max(200000, 2 * natoms), **kw)
    sim.fcc_lattice(n, rho, ntypes)
    sim.adjust_thermo(temp)
    sim.setup_cells()
    return sim
