"""ctypes binding of libpairs_hip.so (the C ABI declared in include/pairs_hip.h).

The library is the only compute path: importing this module on a machine
without the built library raises, and every call that fails inside HIP raises
``PairsHipError`` with the library's own message.  There is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpairs_hip.so")

# return codes (include/pairs_hip.h)
OK = 0
OVERFLOW_PARTICLES, OVERFLOW_NEIGHBORS, OVERFLOW_SEND, OVERFLOW_CELLS = 1, 2, 3, 4
ERR_HIP, ERR_ARG, ERR_STATE = -1, -2, -3

# download selectors
POSITION, VELOCITY, FORCE, MASS, TYPE, FLAGS, UID, NUMNEIGHS, NEIGHBORLISTS, PARTICLE_CELL, \
    CELL_SIZES, CELL_PARTICLES, SEND_MAP, SEND_MULT, NSEND, ATOM_ENERGY = range(16)
NLOCAL, NGHOST, NCELLS, NSEND_ALL, NEIGHBOR_CAPACITY, PARTICLE_CAPACITY, SEND_CAPACITY, NEIGHBOR_TOTAL, TILE_CAPACITY, \
    REBUILDS_DEFERRED, REBUILD_ROLLBACKS = range(11)

FLAG_INFINITE, FLAG_GHOST, FLAG_FIXED, FLAG_GLOBAL = 1, 2, 4, 8


class PairsHipError(RuntimeError):
    pass


class MDConfig(ctypes.Structure):
    _fields_ = [
        ("particle_capacity", ctypes.c_int),
        ("neighbor_capacity", ctypes.c_int),
        ("send_capacity", ctypes.c_int),
        ("ncells_capacity", ctypes.c_int),
        ("box_lo", ctypes.c_double * 3),
        ("box_hi", ctypes.c_double * 3),
        ("periodic", ctypes.c_int * 3),
        ("cell_spacing", ctypes.c_double),
        ("cutoff", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("half", ctypes.c_int),
        ("ntypes", ctypes.c_int),
        ("epsilon", ctypes.POINTER(ctypes.c_double)),
        ("sigma6", ctypes.POINTER(ctypes.c_double)),
        ("reneighbor_every", ctypes.c_int),
        ("thermo_every", ctypes.c_int),
        ("stream", ctypes.c_void_p),
    ]


_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes)
_SIGNATURES = {
    "pairs_hip_create": (_I, [ctypes.POINTER(_P), ctypes.POINTER(MDConfig)]),
    "pairs_hip_destroy": (_I, [_P]),
    "pairs_hip_last_error": (ctypes.c_char_p, [_P]),
    "pairs_hip_last_required": (_I, [_P, _I]),
    "pairs_hip_reserve": (_I, [_P, _I, _I, _I]),
    "pairs_hip_copper_fcc_lattice": (_I, [_P, _I, _I, _I, _D, _I]),
    "pairs_hip_adjust_thermo": (_I, [_P, _D]),
    "pairs_hip_upload": (_I, [_P, _I, _DP, _DP, _DP, _IP, _IP, _IP]),
    "pairs_hip_download": (_I, [_P, _I, _P, _I]),
    "pairs_hip_count": (_I, [_P, _I]),
    "pairs_hip_setup_cells": (_I, [_P]),
    "pairs_hip_exchange": (_I, [_P]),
    "pairs_hip_borders": (_I, [_P]),
    "pairs_hip_sync_ghosts": (_I, [_P]),
    "pairs_hip_bin_cells": (_I, [_P]),
    "pairs_hip_build_neighbors": (_I, [_P]),
    "pairs_hip_lj": (_I, [_P]),
    "pairs_hip_vv_initial": (_I, [_P]),
    "pairs_hip_vv_final": (_I, [_P]),
    "pairs_hip_lj_older": (_I, [_P]),
    "pairs_hip_euler": (_I, [_P]),
    "pairs_hip_thermo": (_I, [_P, _DP]),
    "pairs_hip_energy": (_I, [_P, _DP, _DP]),
    "pairs_hip_energy_noalloc": (_I, [_P, _DP, _DP]),
    "pairs_hip_step": (_I, [_P, _I, _I]),
    "pairs_hip_run": (_I, [_P, _I, _I, _I, _DP, _I]),
    "pairs_hip_sync": (_I, [_P]),
    "pairs_hip_set_graphs": (_I, [_P, _I]),
    "pairs_hip_set_speculative_rebuilds": (_I, [_P, _I]),
    "pairs_hip_prepare_graphs": (_I, [_P]),
    "pairs_hip_set_force_timing": (_I, [_P, _I]),
    "pairs_hip_force_kernel_time": (_I, [_P, _DP, _IP]),
}


def exported_symbols():
    """Names the header declares; tests check that the library exports each."""
    return sorted(_SIGNATURES)


def lib():
    """Load libpairs_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PairsHipError(f"{LIB_PATH} is missing: build it with `python -m pairs_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def register(signatures):
    """Declare more entry points (other modules of the library)."""
    _SIGNATURES.update(signatures)
    if _lib is not None:
        for name, (res, args) in signatures.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args


def check(ctx, code, what):
    """Raise on a negative return; return positive overflow codes to the caller."""
    if code < 0:
        msg = lib().pairs_hip_last_error(ctx) if ctx else b""
        raise PairsHipError(f"{what} failed ({code}): {msg.decode() if msg else ''}")
    return code
