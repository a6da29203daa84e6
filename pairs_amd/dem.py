"""DEM on the GPU: the host-side mirror of the reference's examples/dem.py
program, over the C ABI of include/pairs_hip_dem.h.

    s = DEMSystem(domain=(0.04, 0.015, 0.04))      # psim.set_domain / pbc / build_cell_lists
    s.sc_grid(...)                                 # psim.dem_sc_grid (runtime/dem_sc_grid.hpp)
    s.add_halfspace(...)                           # psim.read_particle_data(planes)
    s.setup()                                      # update_mass_and_inertia
    s.run(0, 201)                                  # the generated timestep loop

Parameters default to examples/dem.py:113-140.  Every call runs on the GPU;
without the built library this module raises on first use.
"""
import ctypes
import math

import numpy as np

from . import native

_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_DP, _IP = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)

OVERFLOW_PARTICLES, OVERFLOW_CONTACTS, OVERFLOW_SEND, OVERFLOW_CELLS = 1, 2, 3, 4
SPHERE, HALFSPACE = 0, 1
(POSITION, LINEAR_VELOCITY, ANGULAR_VELOCITY, FORCE, TORQUE, RADIUS, MASS, ROTATION_QUAT, ROTATION_MATRIX, UID,
 FLAGS, TYPE, SHAPE, NUM_CONTACTS) = range(14)
NLOCAL, NGHOST, NCELLS, CONTACTS, PARTICLE_CAPACITY, CONTACT_CAPACITY, SEND_CAPACITY, RESIZES = range(8)


class DEMConfig(ctypes.Structure):
    _fields_ = [
        ("particle_capacity", _I), ("contact_capacity", _I), ("send_capacity", _I), ("ncells_capacity", _I),
        ("box_lo", _D * 3), ("box_hi", _D * 3), ("periodic", _I * 3), ("cell_width", _D),
        ("dt", _D), ("density_particle", _D), ("density_fluid", _D), ("gravity", _D),
        ("kappa", _D), ("ln_restitution", _D), ("collision_time", _D), ("stiffness_norm", _D), ("ntypes", _I),
        ("friction_static", _D * 16), ("friction_dynamic", _D * 16), ("stream", _P),
    ]


native.register({
    "pairs_hip_dem_create": (_I, [ctypes.POINTER(_P), ctypes.POINTER(DEMConfig)]),
    "pairs_hip_dem_destroy": (_I, [_P]),
    "pairs_hip_dem_last_error": (ctypes.c_char_p, [_P]),
    "pairs_hip_dem_last_required": (_I, [_P, _I]),
    "pairs_hip_dem_sc_grid": (_I, [_P, _D, _D, _D, _D, _D, _D, _D, _D, _D, _I]),
    "pairs_hip_dem_jittered_grid": (_I, [_P, _I, _I, _I, _DP, _DP, _D, _D, _D, _D, ctypes.c_ulonglong]),
    "pairs_hip_dem_add_halfspace": (_I, [_P, _I, _I, _D, _DP, _DP, _I]),
    "pairs_hip_dem_add_spheres": (_I, [_P, _I, _IP, _IP, _DP, _DP, _DP, _DP, _IP]),
    "pairs_hip_dem_setup": (_I, [_P]),
    "pairs_hip_dem_step": (_I, [_P, _I]),
    "pairs_hip_dem_run": (_I, [_P, _I, _I]),
    "pairs_hip_dem_set_graphs": (_I, [_P, _I]),
    "pairs_hip_dem_set_check_every": (_I, [_P, _I]),
    "pairs_hip_dem_sync": (_I, [_P]),
    "pairs_hip_dem_count": (_I, [_P, _I]),
    "pairs_hip_dem_download": (_I, [_P, _I, _P, _I]),
    "pairs_hip_dem_contacts": (_I, [_P, _IP, _IP, _IP, _DP, _DP]),
    "pairs_hip_dem_set_timing": (_I, [_P, _I]),
    "pairs_hip_dem_contact_kernel_time": (_I, [_P, _DP, _IP]),
    "pairs_hip_dem_phase_times": (_I, [_P, _DP]),
})

# examples/dem.py:113-140
REFERENCE = dict(diameter=0.0029, gravity=9.81, density_fluid=1000.0, density_particle=2550.0, spacing=0.005,
                 initial_velocity=1.0, dt=5e-5, friction=0.5, restitution=0.1, collision_time=5e-4, poisson=0.22,
                 friction_static=0.0)


class DEMSystem:
    """One DEM domain on the GPU (the reference's PairsSimulation for dem.py)."""

    def __init__(self, domain, periodic=(True, True, False), particle_capacity=8192, contact_capacity=20,
                 send_capacity=8192, ncells_capacity=None, params=None, ntypes=1, kappa=None, ln_restitution=None,
                 cell_width=None):
        p = dict(REFERENCE)
        p.update(params or {})
        self.params = p
        if kappa is None:
            kappa = 2.0 * (1.0 - p["poisson"]) / (2.0 - p["poisson"])  # Thornton et al. (dem.py:132)
        width = cell_width if cell_width is not None else 1.01 * (p["diameter"] * 1.1)  # linkedCellWidth (dem.py:137)
        self.cell_width = width
        dims = [math.ceil((domain[d] + 2 * width) / width) + 1 for d in range(3)]
        ncells = dims[0] * dims[1] * dims[2] + 1
        cfg = DEMConfig()
        cfg.particle_capacity, cfg.contact_capacity = particle_capacity, contact_capacity
        cfg.send_capacity = send_capacity
        cfg.ncells_capacity = ncells_capacity or ncells + 16
        cfg.box_lo[:] = [0.0, 0.0, 0.0]
        cfg.box_hi[:] = list(domain)
        cfg.periodic[:] = [1 if x else 0 for x in periodic]
        cfg.cell_width = width
        cfg.dt, cfg.gravity = p["dt"], p["gravity"]
        cfg.density_particle, cfg.density_fluid = p["density_particle"], p["density_fluid"]
        lnr = math.log(p["restitution"]) if ln_restitution is None else ln_restitution
        cfg.kappa, cfg.ln_restitution, cfg.collision_time = kappa, lnr, p["collision_time"]
        cfg.stiffness_norm = p.get("stiffness_norm", 0.0) or 0.0
        cfg.ntypes = ntypes
        for k in range(ntypes * ntypes):
            cfg.friction_static[k] = p["friction_static"]
            cfg.friction_dynamic[k] = p["friction"]
        self.L = native.lib()
        self.ctx = _P()
        code = self.L.pairs_hip_dem_create(ctypes.byref(self.ctx), ctypes.byref(cfg))
        if code != 0:
            raise native.PairsHipError(f"pairs_hip_dem_create failed ({code})")
        self.domain, self.ntypes, self.ccap = tuple(domain), ntypes, contact_capacity

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            self.L.pairs_hip_dem_destroy(self.ctx)
            self.ctx = _P()

    def __del__(self):
        self.close()

    # entry points whose positive return is a capacity overflow (the others
    # return counts)
    _OVERFLOWING = ("setup", "step", "run")

    def _check(self, code, what):
        if code < 0:
            msg = self.L.pairs_hip_dem_last_error(self.ctx)
            raise native.PairsHipError(f"{what} failed ({code}): {msg.decode() if msg else ''}")
        if code > 0 and what in self._OVERFLOWING:
            need = self.L.pairs_hip_dem_last_required(self.ctx, code)
            raise native.PairsHipError(f"{what}: capacity {code} exceeded (need {need})")
        return code

    # ---- set-up (examples/dem.py:158-201)
    def sc_grid(self, domain=None, spacing=None, diameter=None, initial_velocity=None, density=None,
                min_diameter=None, max_diameter=None):
        p = self.params
        d = diameter if diameter is not None else p["diameter"]
        dom = domain or self.domain
        dmin = d * 0.9 if min_diameter is None else min_diameter
        dmax = d * 1.1 if max_diameter is None else max_diameter
        return self._check(self.L.pairs_hip_dem_sc_grid(
            self.ctx, dom[0], dom[1], dom[2], spacing or p["spacing"], d, dmin, dmax,
            p["initial_velocity"] if initial_velocity is None else initial_velocity,
            density or p["density_particle"], self.ntypes), "sc_grid")

    def jittered_grid(self, n, origin, spacing, jitter, r_min, r_max, density=None, seed=7):
        """Spheres at seeded random non-overlapping positions
        (pairs_hip_dem_jittered_grid): n = (nx, ny, nz) sites."""
        o = (_D * 3)(*origin)
        sp = (_D * 3)(*spacing)
        return self._check(self.L.pairs_hip_dem_jittered_grid(
            self.ctx, n[0], n[1], n[2], o, sp, jitter, r_min, r_max,
            density or self.params["density_particle"], seed), "jittered_grid")

    def add_halfspace(self, uid, position, normal, flags=13, type_=0, mass=1.0):
        pos = (_D * 3)(*position)
        nrm = (_D * 3)(*normal)
        return self._check(self.L.pairs_hip_dem_add_halfspace(self.ctx, uid, type_, mass, pos, nrm, flags),
                           "add_halfspace")

    def add_reference_planes(self):
        """data/planes.input of the reference: z = 0 facing up, z = 0.2 facing down."""
        self.add_halfspace(100000, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
        self.add_halfspace(100001, (0.8, 0.015, 0.2), (0.0, 0.0, -1.0))

    def add_spheres(self, uid, radius, position, velocity, density=None, type_=None, flags=None):
        n = len(uid)
        rho = density or self.params["density_particle"]
        r = np.ascontiguousarray(radius, np.float64)
        m = np.ascontiguousarray((4.0 / 3.0 * math.pi) * r * r * r * rho, np.float64)
        u = np.ascontiguousarray(uid, np.int32)
        t = np.ascontiguousarray(np.zeros(n) if type_ is None else type_, np.int32)
        f = np.ascontiguousarray(np.zeros(n) if flags is None else flags, np.int32)
        x = np.ascontiguousarray(position, np.float64).reshape(n, 3)
        v = np.ascontiguousarray(velocity, np.float64).reshape(n, 3)
        ip, dp = lambda a: a.ctypes.data_as(_IP), lambda a: a.ctypes.data_as(_DP)
        return self._check(self.L.pairs_hip_dem_add_spheres(self.ctx, n, ip(u), ip(t), dp(r), dp(m), dp(x), dp(v),
                                                            ip(f)), "add_spheres")

    def read_particle_data(self, filename, properties, shape):
        """psim.read_particle_data(filename, properties, shape) of examples/dem.py:
        rows of the CSV file (pairs_amd.io.read_particle_data) added as
        half-spaces or spheres; returns the number added."""
        from . import io
        d = io.read_particle_data(filename, properties, [(0.0, self.domain[k]) for k in range(3)])
        n = len(next(iter(d.values()))) if d else 0
        if shape == "halfspace":
            for k in range(n):
                self.add_halfspace(int(d["uid"][k]), d["position"][k], d["normal"][k],
                                   flags=int(d["flags"][k]) if "flags" in d else 13,
                                   type_=int(d["type"][k]) if "type" in d else 0,
                                   mass=float(d["mass"][k]) if "mass" in d else 1.0)
            return n
        if shape != "sphere":
            raise ValueError(f"unsupported shape {shape}")
        if n:
            # masses come from the density, as for the grid (the file's mass column is ignored)
            self.add_spheres(d["uid"], d["radius"], d["position"], d["linear_velocity"],
                             type_=d.get("type"), flags=d.get("flags"))
        return n

    def setup(self):
        self._check(self.L.pairs_hip_dem_setup(self.ctx), "setup")

    # ---- the timestep loop
    def step(self, ts):
        self._check(self.L.pairs_hip_dem_step(self.ctx, ts), "step")

    def run(self, ts_begin, ts_end):
        self._check(self.L.pairs_hip_dem_run(self.ctx, ts_begin, ts_end), "run")

    def set_graphs(self, enable):
        self._check(self.L.pairs_hip_dem_set_graphs(self.ctx, 1 if enable else 0), "set_graphs")

    def set_timing(self, every):
        self._check(self.L.pairs_hip_dem_set_timing(self.ctx, every), "set_timing")

    def contact_kernel_time(self):
        ms, n = _D(), _I()
        self._check(self.L.pairs_hip_dem_contact_kernel_time(self.ctx, ctypes.byref(ms), ctypes.byref(n)),
                    "contact_kernel_time")
        return ms.value, n.value

    def phase_times(self):
        """Average ms of the timed steps' phases: prepare, contact search,
        history slots, spring-dashpot, sums + history clean-up, euler."""
        ms = (_D * 6)()
        self._check(self.L.pairs_hip_dem_phase_times(self.ctx, ms), "phase_times")
        return tuple(ms)

    def sync(self):
        self._check(self.L.pairs_hip_dem_sync(self.ctx), "sync")

    # ---- state
    def set_check_every(self, steps):
        """Steps per checkpoint and overflow check of run() (resize protocol)."""
        self._check(self.L.pairs_hip_dem_set_check_every(self.ctx, steps), "set_check_every")

    def count(self, what):
        return self._check(self.L.pairs_hip_dem_count(self.ctx, what), "count")

    @property
    def capacities(self):
        """(particle, contact, send) capacities now, and the growths so far"""
        return (self.count(PARTICLE_CAPACITY), self.count(CONTACT_CAPACITY), self.count(SEND_CAPACITY),
                self.count(RESIZES))

    @property
    def nlocal(self):
        return self._check(self.L.pairs_hip_dem_count(self.ctx, NLOCAL), "count")

    @property
    def nghost(self):
        return self._check(self.L.pairs_hip_dem_count(self.ctx, NGHOST), "count")

    def get(self, name, ghosts=False):
        spec = {"position": (POSITION, np.float64, 3), "linear_velocity": (LINEAR_VELOCITY, np.float64, 3),
                "angular_velocity": (ANGULAR_VELOCITY, np.float64, 3), "force": (FORCE, np.float64, 3),
                "torque": (TORQUE, np.float64, 3), "radius": (RADIUS, np.float64, None),
                "mass": (MASS, np.float64, None), "rotation_quat": (ROTATION_QUAT, np.float64, 4),
                "rotation_matrix": (ROTATION_MATRIX, np.float64, 9), "uid": (UID, np.int32, None),
                "flags": (FLAGS, np.int32, None), "type": (TYPE, np.int32, None), "shape": (SHAPE, np.int32, None),
                "num_contacts": (NUM_CONTACTS, np.int32, None)}
        which, dt, cols = spec[name]
        n = self.nlocal + (self.nghost if ghosts else 0)
        out = np.zeros((n, cols) if cols else n, dt)
        self._check(self.L.pairs_hip_dem_download(self.ctx, which, out.ctypes.data_as(_P), n), "download")
        return out

    def contacts(self):
        """num_contacts [n], partner uid / sticking [n][C], tangential spring [n][C][3], impact velocity [n][C]"""
        n, c = self.nlocal, self.count(CONTACT_CAPACITY)
        nc = np.zeros(n, np.int32)
        uid = np.zeros((n, c), np.int32)
        st = np.zeros((n, c), np.int32)
        tsd = np.zeros((n, c, 3), np.float64)
        iv = np.zeros((n, c), np.float64)
        self._check(self.L.pairs_hip_dem_contacts(self.ctx, nc.ctypes.data_as(_IP), uid.ctypes.data_as(_IP),
                                                  st.ctypes.data_as(_IP), tsd.ctypes.data_as(_DP),
                                                  iv.ctypes.data_as(_DP)), "contacts")
        return nc, uid, st, tsd, iv


def reference_system(domain=(0.04, 0.015, 0.04), **kw):
    """examples/dem.py's setup on the given domain: grid, the two planes, masses."""
    s = DEMSystem(domain, **kw)
    s.sc_grid(domain)
    s.add_reference_planes()
    s.setup()
    return s
