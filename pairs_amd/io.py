"""Particle input (CSV) and output (legacy ASCII VTK), include/pairs_hip_io.h.

``read_particle_data`` reads a CSV file the way runtime/read_from_file.hpp:33-115
does: columns in the order of the property list, vectors as three columns,
rows kept when the position lies in the domain or the flags mark the particle
Infinite, Fixed or Global.  ``vtk_write`` writes the locals of an LJ or DEM
system as runtime/vtk.hpp:11-86 does.
"""
import ctypes

import numpy as np

from . import native as nat

_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_DP, _IP = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)

nat.register({
    "pairs_hip_read_particle_csv": (_I, [ctypes.c_char_p, _I, _IP, _I, _I, _DP, _DP, _DP, _I, _IP]),
    "pairs_hip_vtk_write_arrays": (_I, [ctypes.c_char_p, _I, _DP, _DP, _IP, _I, _I]),
    "pairs_hip_vtk_write": (_I, [_P, ctypes.c_char_p, _I, _I]),
    "pairs_hip_dem_vtk_write": (_I, [_P, ctypes.c_char_p, _I, _I]),
})

PROP_INT, PROP_REAL, PROP_VECTOR, PROP_QUATERNION, PROP_MATRIX = 0, 1, 3, 4, 9

# kinds of the properties the reference's examples read (examples/dem.py and
# the DSL's default property set)
KNOWN = {"uid": PROP_INT, "type": PROP_INT, "flags": PROP_INT, "mass": PROP_REAL, "radius": PROP_REAL,
         "position": PROP_VECTOR, "linear_velocity": PROP_VECTOR, "angular_velocity": PROP_VECTOR,
         "normal": PROP_VECTOR, "force": PROP_VECTOR, "torque": PROP_VECTOR, "rotation_quat": PROP_QUATERNION,
         "rotation_matrix": PROP_MATRIX, "inv_inertia": PROP_MATRIX}


def read_particle_data(filename, properties, domain, kinds=None, max_rows=1 << 20):
    """{property: array} of the rows kept (ints as int32, others float64;
    vectors [n, 3], quaternions [n, 4], matrices [n, 9])."""
    kinds = [(kinds or {}).get(p, KNOWN.get(p)) for p in properties]
    if any(k is None for k in kinds):
        raise ValueError(f"unknown property kind in {properties}")
    ncols = sum(1 if k in (PROP_INT, PROP_REAL) else k for k in kinds)
    kk = np.ascontiguousarray(kinds, np.int32)
    lo = np.ascontiguousarray([d[0] for d in domain], np.float64)
    hi = np.ascontiguousarray([d[1] for d in domain], np.float64)
    pos = properties.index("position") if "position" in properties else -1
    fl = properties.index("flags") if "flags" in properties else -1
    L = nat.lib()
    while True:
        out = np.zeros((max_rows, ncols), np.float64)
        need = ctypes.c_int()
        n = L.pairs_hip_read_particle_csv(str(filename).encode(), len(properties), kk.ctypes.data_as(_IP), pos, fl,
                                          lo.ctypes.data_as(_DP), hi.ctypes.data_as(_DP),
                                          out.ctypes.data_as(_DP), max_rows, ctypes.byref(need))
        if n == nat.ERR_STATE:
            max_rows = need.value
            continue
        if n < 0:
            raise nat.PairsHipError(f"read_particle_data({filename}) failed ({n})")
        break
    res, c = {}, 0
    for p, k in zip(properties, kinds):
        w = 1 if k in (PROP_INT, PROP_REAL) else k
        col = out[:n, c:c + w]
        res[p] = col[:, 0].astype(np.int32) if k == PROP_INT else (col[:, 0].copy() if w == 1 else col.copy())
        c += w
    return res


def vtk_write(system, prefix, timestep, frequency=0):
    """Write the locals of an LJ system (pairs_amd.md) or DEM system
    (pairs_amd.dem); returns True if a file was written."""
    L = nat.lib()
    if hasattr(system, "ctx"):
        code = L.pairs_hip_dem_vtk_write(system.ctx, str(prefix).encode(), timestep, frequency)
    else:
        code = L.pairs_hip_vtk_write(system.h, str(prefix).encode(), timestep, frequency)
    if code < 0:
        raise nat.PairsHipError(f"vtk_write({prefix}) failed ({code})")
    return code == 1


def vtk_write_arrays(prefix, position, mass, flags, timestep, frequency=0):
    x = np.ascontiguousarray(position, np.float64).reshape(-1, 3)
    m = np.ascontiguousarray(mass, np.float64)
    f = np.ascontiguousarray(flags, np.int32)
    code = nat.lib().pairs_hip_vtk_write_arrays(str(prefix).encode(), len(m), x.ctypes.data_as(_DP),
                                                m.ctypes.data_as(_DP), f.ctypes.data_as(_IP), timestep, frequency)
    if code < 0:
        raise nat.PairsHipError(f"vtk_write_arrays({prefix}) failed ({code})")
    return code == 1
