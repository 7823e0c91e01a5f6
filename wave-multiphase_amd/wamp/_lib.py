"""ctypes binding of libwamp.so (include/wamp.h).

The library is built in-tree by __graft_entry__.build() (or
`python -m wamp.build`).  There is no fallback: if the library is missing
or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC_DIR = os.path.join(os.path.dirname(PKG_DIR), "csrc")
LIB_PATH = os.environ.get("WAMP_LIB", os.path.join(CSRC_DIR, "libwamp.so"))


class WampError(RuntimeError):
    """A libwamp entry point returned a non-zero status."""


class WampGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("g", C.c_int32), ("ld", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("bc", C.c_int32 * 4)]


class WampEos(C.Structure):
    _fields_ = [("gamma1", C.c_double), ("pinf1", C.c_double),
                ("gamma2", C.c_double), ("pinf2", C.c_double)]


class WampScheme(C.Structure):
    _fields_ = [("base", C.c_int32), ("shear", C.c_int32), ("vf_thinc", C.c_int32),
                ("face_pp", C.c_int32), ("repair", C.c_int32), ("cfl", C.c_double)]


class WampDropletIC(C.Structure):
    _fields_ = [("xmin", C.c_double), ("ymin", C.c_double), ("shock_x", C.c_double),
                ("rho_post", C.c_double), ("p_post", C.c_double), ("u_post", C.c_double),
                ("rho_water", C.c_double), ("rho_air", C.c_double), ("p0", C.c_double),
                ("width", C.c_double), ("drops", C.c_void_p), ("n_drops", C.c_int32)]


class WampTripleIC(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("xmin", "ymin", "x_split", "y_split", "rho_left",
                                           "p_left", "rho_top", "p_top", "rho_bottom", "p_bottom")]


class WampGrid3(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("g", C.c_int32),
                ("ldz", C.c_int32), ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("bc", C.c_int32 * 6)]


PtrArr7 = C.c_void_p * 7
PtrArr6 = C.c_void_p * 6
PtrArr3 = C.c_void_p * 3
PtrArr28 = C.c_void_p * 28
PtrArr48 = C.c_void_p * 48

# (name, restype, argtypes) of every symbol include/wamp.h declares
_VP = C.c_void_p
SIGNATURES = [
    ("wamp_ctx_create", C.c_int, [C.POINTER(WampGrid), C.POINTER(WampEos), C.POINTER(WampScheme),
                                  C.c_int32, C.POINTER(_VP)]),
    ("wamp_ctx_destroy", C.c_int, [_VP]),
    ("wamp_last_error", C.c_char_p, []),
    ("wamp_ctx_scratch_bytes", C.c_int64, [_VP]),
    ("wamp_apply_bc", C.c_int, [_VP, PtrArr6, _VP, _VP]),
    ("wamp_cons_to_prim", C.c_int, [_VP, PtrArr6, PtrArr3, _VP]),
    ("wamp_enforce_admissible", C.c_int, [_VP, PtrArr6, _VP]),
    ("wamp_prepare", C.c_int, [_VP, PtrArr6, _VP]),
    ("wamp_compute_fluxes", C.c_int, [_VP, PtrArr6, PtrArr28, _VP]),
    ("wamp_limit_update", C.c_int, [_VP, PtrArr6, PtrArr6, PtrArr28, PtrArr6, C.c_double,
                                    C.c_double, _VP, _VP, _VP]),
    ("wamp_get_dt", C.c_int, [_VP, PtrArr6, _VP, _VP]),
    ("wamp_step", C.c_int, [_VP, PtrArr6, _VP, _VP, _VP]),
    ("wamp_step_noalloc", C.c_int, [C.POINTER(WampGrid), C.POINTER(WampEos), C.POINTER(WampScheme),
                                    C.c_int32, PtrArr6, _VP, _VP, _VP]),
    ("wamp_init_droplets", C.c_int, [C.POINTER(WampGrid), C.POINTER(WampEos),
                                     C.POINTER(WampDropletIC), PtrArr6, _VP]),
    ("wamp_init_triple_point", C.c_int, [C.POINTER(WampGrid), C.POINTER(WampEos),
                                         C.POINTER(WampTripleIC), PtrArr6, _VP]),
    ("wamp_batch_create", C.c_int, [C.POINTER(WampScheme), C.c_int32, C.c_int32, C.POINTER(_VP)]),
    ("wamp_batch_destroy", C.c_int, [_VP]),
    ("wamp_batch_set_case", C.c_int, [_VP, C.c_int32, C.POINTER(WampGrid), C.POINTER(WampEos),
                                      PtrArr6, C.c_double, C.c_double]),
    ("wamp_batch_clear_case", C.c_int, [_VP, C.c_int32]),
    ("wamp_batch_set_tuning", C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    ("wamp_batch_commit", C.c_int, [_VP, _VP]),
    ("wamp_batch_prepare", C.c_int, [_VP, C.POINTER(C.c_int32), C.c_int32, _VP]),
    ("wamp_batch_set_done", C.c_int, [_VP, C.c_int32, C.c_int32, _VP]),
    ("wamp_batch_graph", C.c_int, [_VP, C.c_int32]),
    ("wamp_batch_step", C.c_int, [_VP, _VP]),
    ("wamp_batch_status", C.c_int, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int32), _VP]),
    ("wamp_batch_work", C.c_int64, [_VP, C.c_int32]),
    ("wamp_timing_enable", C.c_int, [_VP, C.c_int32]),
    ("wamp_timing_read", C.c_int, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    ("wamp_selftest_arith", C.c_int, [_VP, _VP, _VP, C.c_int64, _VP]),
    ("wamp_selftest_elementary", C.c_int, [_VP, _VP, _VP, C.c_int64, _VP]),
    ("wamp_cpu_ctx_create", C.c_int, [C.POINTER(WampGrid), C.POINTER(WampEos),
                                      C.POINTER(WampScheme), C.c_int32, C.POINTER(_VP)]),
    ("wamp_cpu_ctx_destroy", C.c_int, [_VP]),
    ("wamp_cpu_threads", C.c_int32, [_VP]),
    ("wamp_cpu_prepare", C.c_int, [_VP, PtrArr6]),
    ("wamp_cpu_apply_bc", C.c_int, [_VP, PtrArr6, _VP]),
    ("wamp_cpu_cons_to_prim", C.c_int, [_VP, PtrArr6, PtrArr3]),
    ("wamp_cpu_compute_fluxes", C.c_int, [_VP, PtrArr6, PtrArr28]),
    ("wamp_cpu_get_dt", C.c_int, [_VP, PtrArr6, C.POINTER(C.c_double)]),
    ("wamp_cpu_step", C.c_int, [_VP, PtrArr6, C.c_double, C.POINTER(C.c_double)]),
    ("wamp3_ctx_create", C.c_int, [C.POINTER(WampGrid3), C.POINTER(WampEos), C.POINTER(WampScheme),
                                   C.c_int32, C.POINTER(_VP)]),
    ("wamp3_ctx_destroy", C.c_int, [_VP]),
    ("wamp3_ctx_scratch_bytes", C.c_int64, [_VP]),
    ("wamp3_prepare", C.c_int, [_VP, PtrArr7, _VP]),
    ("wamp3_get_dt", C.c_int, [_VP, PtrArr7, _VP, _VP]),
    ("wamp3_step", C.c_int, [_VP, PtrArr7, _VP, _VP, _VP]),
    ("wamp3_batch_create", C.c_int, [C.POINTER(WampScheme), C.c_int32, C.c_int32, C.POINTER(_VP)]),
    ("wamp3_batch_destroy", C.c_int, [_VP]),
    ("wamp3_batch_set_case", C.c_int, [_VP, C.c_int32, C.POINTER(WampGrid3), C.POINTER(WampEos),
                                       PtrArr7, C.c_double, C.c_double]),
    ("wamp3_batch_clear_case", C.c_int, [_VP, C.c_int32]),
    ("wamp3_batch_commit", C.c_int, [_VP, _VP]),
    ("wamp3_batch_prepare", C.c_int, [_VP, C.POINTER(C.c_int32), C.c_int32, _VP]),
    ("wamp3_batch_graph", C.c_int, [_VP, C.c_int32]),
    ("wamp3_batch_step", C.c_int, [_VP, _VP]),
    ("wamp3_batch_status", C.c_int, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int32), _VP]),
    ("wamp3_compute_fluxes", C.c_int, [_VP, PtrArr7, PtrArr48, _VP]),
    ("wamp3_limit_update", C.c_int, [_VP, PtrArr7, PtrArr7, PtrArr48, PtrArr7, C.c_double,
                                     C.c_double, _VP, _VP, _VP]),
]

_lib = None


def load(path: str | None = None):
    """Load libwamp.so once; raise if it is missing or lacks a declared symbol."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise WampError(f"libwamp.so not found at {p}: build it with "
                        f"`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(p)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)  # AttributeError if a declared symbol is missing
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().wamp_last_error()
        msg = msg.decode() if msg else ""
        raise WampError(f"{what} failed with status {rc}: {msg}")


def ptrs6(planes) -> PtrArr6:
    return PtrArr6(*[int(p) for p in planes])


def ptrs3(planes) -> PtrArr3:
    return PtrArr3(*[int(p) for p in planes])


def ptrs28(planes) -> PtrArr28:
    return PtrArr28(*[int(p) for p in planes])


def ptrs48(planes) -> PtrArr48:
    return PtrArr48(*[int(p) for p in planes])
