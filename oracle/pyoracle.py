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
        for fn in ("omd_setup_cells", "omd_rebuild", "omd_force", "omd_sync", "omd_lj_older", "omd_euler"):
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
        L.omd_set_threads.argtypes = [ctypes.c_int]
        L.omd_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def set_threads(n):
    """OpenMP threads of the LJ oracle's per-particle loops (1 = serial)."""
    lib().omd_set_threads(int(n))


def threads():
    return lib().omd_threads()


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

    # the older API's modules (examples/lj_onetype.py:5-13)
    def lj_older(self):
        self.L.omd_lj_older(self.h)

    def euler(self):
        self.L.omd_euler(self.h)

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


# ---------------------------------------------------------------- DEM oracle

DEM_X, DEM_V, DEM_W, DEM_F, DEM_T, DEM_R, DEM_M, DEM_Q, DEM_ROT, DEM_UID, DEM_FLAGS, DEM_TYPE, \
    DEM_NCONT, DEM_CLIST, DEM_STICK, DEM_TSD, DEM_IVM, DEM_SHAPE, DEM_NORMAL, DEM_INVI, DEM_CUSED = range(21)


def _dem_lib():
    L = lib()
    if getattr(L, "_dem_ready", False):
        return L
    L.odem_create.restype = ctypes.c_void_p
    L.odem_create.argtypes = [ctypes.c_int] * 5 + [_D, _D, _I] + [ctypes.c_double] * 8 + [ctypes.c_int, _D, _D]
    L.odem_destroy.argtypes = [ctypes.c_void_p]
    L.odem_sc_grid.restype = ctypes.c_int
    L.odem_sc_grid.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 9 + [ctypes.c_int]
    L.odem_add_halfspace.restype = ctypes.c_int
    L.odem_add_halfspace.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _D,
                                     ctypes.c_int]
    for fn in ("odem_setup", "odem_setup_cells"):
        getattr(L, fn).argtypes = [ctypes.c_void_p]
    L.odem_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.odem_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.odem_ptr.restype = ctypes.c_void_p
    L.odem_ptr.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.odem_int.restype = ctypes.c_int
    L.odem_int.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.odem_add_sphere.restype = ctypes.c_int
    L.odem_add_sphere.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  _D, _D]
    L.odem_mt_next.restype = ctypes.c_uint32
    L.odem_mt_next.argtypes = [ctypes.c_void_p]
    L.odem_mt64_nth.restype = ctypes.c_uint64
    L.odem_mt64_nth.argtypes = [ctypes.c_uint64, ctypes.c_int]
    L.odem_set_stiffness.argtypes = [ctypes.c_void_p, ctypes.c_double]
    L.odem_jittered_grid.restype = ctypes.c_int
    L.odem_jittered_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_uint64]
    L._dem_ready = True
    return L


# examples/dem.py:113-140 (the reference's own DEM configuration)
DEM_REF = dict(diameter=0.0029, gravity=9.81, rho_f=1000.0, rho_p=2550.0, spacing=0.005, v0=1.0, dt=5e-5,
               mu=0.5, restitution=0.1, t_c=5e-4, poisson=0.22)


def dem_params(cfg=DEM_REF):
    kappa = 2.0 * (1.0 - cfg["poisson"]) / (2.0 - cfg["poisson"])
    return dict(kappa=kappa, ln_e=math.log(cfg["restitution"]), cell=1.01 * (cfg["diameter"] * 1.1))


class OracleDEM:
    """The CPU oracle's DEM system (spheres + half-spaces, contact history)."""

    def __init__(self, domain, periodic=(1, 1, 0), cap=4096, ccap=20, cellcap=64, send_cap=4096,
                 ncells_cap=200000, cfg=DEM_REF, ntypes=1, mu_s=None, mu_d=None, stiffness_norm=0.0):
        L = _dem_lib()
        p = dem_params(cfg)
        lo = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
        hi = (ctypes.c_double * 3)(*domain)
        per = (ctypes.c_int * 3)(*periodic)
        nt2 = ntypes * ntypes
        ms = (ctypes.c_double * nt2)(*(mu_s if mu_s is not None else [0.0] * nt2))
        md = (ctypes.c_double * nt2)(*(mu_d if mu_d is not None else [cfg["mu"]] * nt2))
        self.cfg, self.cell, self.ntypes = cfg, p["cell"], ntypes
        self.h = L.odem_create(cap, ccap, cellcap, send_cap, ncells_cap, lo, hi, per, p["cell"], cfg["dt"],
                               cfg["rho_p"], cfg["rho_f"], cfg["gravity"], p["kappa"], p["ln_e"], cfg["t_c"],
                               ntypes, ms, md)
        self.L = L
        self.cap, self.ccap = cap, ccap
        L.odem_set_stiffness(self.h, stiffness_norm)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.odem_destroy(self.h)
            self.h = None

    def sc_grid(self, domain, spacing=None, diameter=None, v0=None):
        c = self.cfg
        d = diameter if diameter is not None else c["diameter"]
        return self.L.odem_sc_grid(self.h, domain[0], domain[1], domain[2], spacing or c["spacing"], d, d * 0.9,
                                   d * 1.1, c["v0"] if v0 is None else v0, c["rho_p"], self.ntypes)

    def jittered_grid(self, n, origin, spacing, jitter, r_min, r_max, density=None, seed=7):
        o = (ctypes.c_double * 3)(*origin)
        sp = (ctypes.c_double * 3)(*spacing)
        return self.L.odem_jittered_grid(self.h, n[0], n[1], n[2], o, sp, jitter, r_min, r_max,
                                         self.cfg["rho_p"] if density is None else density, seed)

    def add_halfspace(self, uid, pos, normal, flags=13, type_=0, mass=1.0):
        p = (ctypes.c_double * 3)(*pos)
        n = (ctypes.c_double * 3)(*normal)
        return self.L.odem_add_halfspace(self.h, uid, type_, mass, p, n, flags)

    def add_sphere(self, uid, radius, pos, vel, density=None, type_=0):
        p = (ctypes.c_double * 3)(*pos)
        v = (ctypes.c_double * 3)(*vel)
        return self.L.odem_add_sphere(self.h, uid, type_, radius, self.cfg["rho_p"] if density is None else density,
                                      p, v)

    def add_reference_planes(self):
        """data/planes.input of the reference: z = 0 facing up, z = 0.2 facing down."""
        self.add_halfspace(100000, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
        self.add_halfspace(100001, (0.8, 0.015, 0.2), (0.0, 0.0, -1.0))

    def setup(self):
        self.L.odem_setup(self.h)
        self.L.odem_setup_cells(self.h)

    def step(self, ts):
        self.L.odem_step(self.h, ts)

    def run(self, a, b):
        self.L.odem_run(self.h, a, b)

    @property
    def nlocal(self):
        return self.L.odem_int(self.h, 0)

    @property
    def nghost(self):
        return self.L.odem_int(self.h, 1)

    @property
    def overflow(self):
        return self.L.odem_int(self.h, 3)

    def _arr(self, which, ctype, n, cols=None):
        ptr = self.L.odem_ptr(self.h, which)
        a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(n * (cols or 1),)).copy()
        return a.reshape(n, cols) if cols else a

    def get(self, name, ghosts=False):
        n = self.nlocal + (self.nghost if ghosts else 0)
        spec = {"position": (DEM_X, ctypes.c_double, 3), "linear_velocity": (DEM_V, ctypes.c_double, 3),
                "angular_velocity": (DEM_W, ctypes.c_double, 3), "force": (DEM_F, ctypes.c_double, 3),
                "torque": (DEM_T, ctypes.c_double, 3), "radius": (DEM_R, ctypes.c_double, None),
                "mass": (DEM_M, ctypes.c_double, None), "rotation_quat": (DEM_Q, ctypes.c_double, 4),
                "rotation_matrix": (DEM_ROT, ctypes.c_double, 9), "uid": (DEM_UID, ctypes.c_int, None),
                "flags": (DEM_FLAGS, ctypes.c_int, None), "type": (DEM_TYPE, ctypes.c_int, None),
                "shape": (DEM_SHAPE, ctypes.c_int, None), "normal": (DEM_NORMAL, ctypes.c_double, 3),
                "inv_inertia": (DEM_INVI, ctypes.c_double, 9), "num_contacts": (DEM_NCONT, ctypes.c_int, None)}
        which, ct, cols = spec[name]
        return self._arr(which, ct, n, cols)

    def contacts(self):
        """per local particle: list of (uid, is_sticking, tsd[3], impact) in slot order"""
        n, c = self.nlocal, self.ccap
        nc = self._arr(DEM_NCONT, ctypes.c_int, n)
        cl = self._arr(DEM_CLIST, ctypes.c_int, n, c)
        st = self._arr(DEM_STICK, ctypes.c_int, n, c)
        tsd = self._arr(DEM_TSD, ctypes.c_double, n * c, 3).reshape(n, c, 3)
        iv = self._arr(DEM_IVM, ctypes.c_double, n, c)
        return nc, cl, st, tsd, iv

    def mt_next(self):
        return self.L.odem_mt_next(self.h)


def dem_reference_system(domain=(0.04, 0.015, 0.04), **kw):
    """The reference's examples/dem.py setup at a given domain size."""
    o = OracleDEM(domain, **kw)
    o.sc_grid(domain)
    o.add_reference_planes()
    o.setup()
    return o


def mt64_nth(seed, n):
    """n-th output (n >= 1) of the oracle's MT19937-64 seeded with seed."""
    return _dem_lib().odem_mt64_nth(seed, n)
