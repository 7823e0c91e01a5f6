"""ctypes wrapper of the CPU ORACLE (oracle/liborc.so) -- test infrastructure only.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker.  The product library never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
# the same source built with every double counted by operation kind (opcount.h)
COUNT_LIB_PATH = os.path.join(HERE, "liborc_count.so")
# the same source with the solver's own pow and tanh (shared_elementary.h)
WM_LIB_PATH = os.path.join(HERE, "liborc_wm.so")
OPCOUNT_KINDS = ("add", "mul", "div", "sqrt", "pow", "tanh", "fabs", "floor", "cmp")

# array ids, mirror of the enum in wamp_oracle.c
AR1, AR2, RU, RV, E, AL, UP, VP, PP, ALPRIM = range(10)
AR1O = 10
FX0, SRCX = 16, 22
GY0, SRCY = 23, 29
FXL0, SRCXL = 30, 36
GYL0, SRCYL = 37, 43
THETA, SIGX, SIGY = 44, 45, 46


class OrcEos(C.Structure):
    _fields_ = [("g1", C.c_double), ("p1", C.c_double), ("g2", C.c_double), ("p2", C.c_double)]


class OrcParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("g", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("eos", OrcEos),
                ("base", C.c_int), ("shear", C.c_int), ("vf_thinc", C.c_int),
                ("face_pp", C.c_int), ("repair", C.c_int), ("bc", C.c_int * 4),
                ("cfl", C.c_double)]


# orc_step's test hook: fn(user, stage) fills the face fluxes of the stage
FLUX_HOOK = C.CFUNCTYPE(None, C.c_void_p, C.c_int)


class S6(C.Structure):
    _fields_ = [("v%d" % k, C.c_double) for k in range(6)]

    def tolist(self):
        return [self.v0, self.v1, self.v2, self.v3, self.v4, self.v5]


def _path(counting: bool, shared_math: bool) -> str:
    if counting and shared_math:
        raise ValueError("the counting build uses the C library's pow and tanh")
    return COUNT_LIB_PATH if counting else WM_LIB_PATH if shared_math else LIB_PATH


def build(force: bool = False, counting: bool = False, shared_math: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -O2 -ffp-contract=off); with
    counting, the operation-counting build (g++, same flags, serial); with
    shared_math, the build whose pow and tanh are the solver's own."""
    path = _path(counting, shared_math)
    if force and os.path.exists(path):
        os.remove(path)
    subprocess.run(["make", "-s", "-C", HERE, os.path.basename(path)], check=True)
    return path


_libs = {}


def lib(counting: bool = False, shared_math: bool = False):
    path = _path(counting, shared_math)
    if path not in _libs:
        if not os.path.exists(path):
            build(counting=counting, shared_math=shared_math)
        L = C.CDLL(path)
        d, pd, ip = C.c_double, C.POINTER(C.c_double), C.c_void_p
        pe = C.POINTER(OrcEos)
        L.orc_create.restype = ip
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [ip]
        L.orc_array.restype = pd
        L.orc_array.argtypes = [ip, C.c_int]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        for n in ("orc_prepare", "orc_apply_bc", "orc_update_prims", "orc_enforce_admissible",
                  "orc_compute_fluxes", "orc_copy_old"):
            getattr(L, n).argtypes = [ip]
        L.orc_limit_fluxes.argtypes = [ip, d]
        L.orc_rk3_update.argtypes = [ip, d, d, d]
        L.orc_get_dt.restype = d
        L.orc_get_dt.argtypes = [ip]
        L.orc_step.argtypes = [ip, d]
        L.orc_run.argtypes = [ip, C.c_int, pd]
        L.orc_set_flux_hook.argtypes = [ip, FLUX_HOOK, ip]
        # scalar helpers
        L.orc_mix_gamma_pibig.argtypes = [pe, d, pd, pd]
        L.orc_mix_gamma_pinf.argtypes = [pe, d, pd, pd]
        for n, args in (("orc_pow", [d, d]), ("orc_tanh", [d]),
                        ("orc_sound_speed", [pe, d, d, d]),
                        ("orc_prim_to_energy", [pe, d, d, d, d, d, d]),
                        ("orc_clamp01_pp", [d]),
                        ("orc_pp_theta_lower", [d, d, d]),
                        ("orc_rho_ctilde2_cons", [pe, d, d, d, d, d, d]),
                        ("orc_theta_cons_stage", [pe, pd, pd, C.c_int]),
                        ("orc_cell_entropy", [pe, d, d, d, d]),
                        ("orc_cator2_val", [d] * 5),
                        ("orc_thinc_ul", [d] * 3), ("orc_thinc_ur", [d] * 3),
                        ("orc_mp5_ul", [d] * 5), ("orc_mp5_ur", [d] * 5),
                        ("orc_central6_avg", [d] * 6), ("orc_central4", [d] * 4),
                        ("orc_muscl_char_ul", [d] * 4), ("orc_muscl_char_ur", [d] * 4),
                        ("orc_hllc_x_src", [pe, pd, pd]), ("orc_hllc_y_src", [pe, pd, pd])):
            f = getattr(L, n)
            f.restype = d
            f.argtypes = args
        for n, args in (("orc_to_char_x", [pd, d, d, d, d, d, d]),
                        ("orc_to_char_y", [pd, d, d, d, d, d, d]),
                        ("orc_to_phys_x", [S6, d, d, d, d, d, d]),
                        ("orc_to_phys_y", [S6, d, d, d, d, d, d]),
                        ("orc_hllc_x", [pe, pd, pd]), ("orc_hllc_y", [pe, pd, pd]),
                        ("orc_pp_floor_pressure_state", [pe, S6]),
                        ("orc_pp_limit_reconstructed_state", [pe, S6, S6])):
            f = getattr(L, n)
            f.restype = S6
            f.argtypes = args
        if counting:
            L.orc_opcount_read.argtypes = [C.POINTER(C.c_uint64)]
            L.orc_opcount_categories.restype = C.c_int
            if L.orc_opcount_categories() != len(OPCOUNT_KINDS):
                raise RuntimeError("liborc_count.so counts a different set of operation kinds")
        _libs[path] = L
    return _libs[path]


def eos_struct(eos) -> OrcEos:
    t = eos.as_tuple() if hasattr(eos, "as_tuple") else tuple(eos)
    return OrcEos(*t)


def dvec(values) -> "C.Array":
    arr = (C.c_double * len(values))(*[float(v) for v in values])
    return arr


class Oracle:
    """One oracle solver context over the reference's 47 padded arrays."""

    def __init__(self, cfg, threads: int | None = None, counting: bool = False,
                 shared_math: bool = False):
        g = cfg.grid
        s = cfg.scheme
        p = OrcParams()
        p.nx, p.ny, p.g = g.nx, g.ny, g.g
        p.dx, p.dy = g.dx, g.dy
        p.eos = eos_struct(cfg.eos)
        p.base, p.shear, p.vf_thinc, p.face_pp = s.base, s.shear, s.vf_thinc, s.face_pp
        p.repair = 1 if s.repair else 0
        p.bc = (C.c_int * 4)(*g.bc)
        p.cfl = s.cfl
        self.params = p
        self.cfg = cfg
        self.shape = (g.nxg, g.nyg)
        self._L = lib(counting, shared_math)
        self.counting = counting
        if threads is not None:
            self._L.orc_set_threads(int(threads))
        self._ctx = self._L.orc_create(C.byref(p))
        if not self._ctx:
            raise MemoryError("orc_create failed")

    def close(self):
        if self._ctx:
            self._L.orc_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def array(self, which: int) -> np.ndarray:
        ptr = self._L.orc_array(self._ctx, which)
        return np.ctypeslib.as_array(ptr, shape=self.shape)

    def set_state(self, U: np.ndarray):
        for m in range(6):
            self.array(AR1 + m)[...] = U[m]

    def state(self) -> np.ndarray:
        return np.stack([self.array(AR1 + m).copy() for m in range(6)])

    def prims(self) -> np.ndarray:
        return np.stack([self.array(UP).copy(), self.array(VP).copy(), self.array(PP).copy()])

    def fluxes(self) -> np.ndarray:
        """(28, NXG, NYG): fx0..5, srcx, gy0..5, srcy, fxl0..5, srcxl, gyl0..5, srcyl."""
        return np.stack([self.array(FX0 + m).copy() for m in range(28)])

    def prepare(self):
        self._L.orc_prepare(self._ctx)

    def get_dt(self) -> float:
        return float(self._L.orc_get_dt(self._ctx))

    def step(self, dt: float):
        self._L.orc_step(self._ctx, float(dt))

    def run(self, nsteps: int) -> np.ndarray:
        dts = np.zeros(nsteps, dtype=np.float64)
        self._L.orc_run(self._ctx, int(nsteps), dts.ctypes.data_as(C.POINTER(C.c_double)))
        return dts

    def apply_bc(self):
        self._L.orc_apply_bc(self._ctx)

    def update_prims(self):
        self._L.orc_update_prims(self._ctx)

    def enforce_admissible(self):
        self._L.orc_enforce_admissible(self._ctx)

    def compute_fluxes(self):
        self._L.orc_compute_fluxes(self._ctx)

    def limit_fluxes(self, dt):
        self._L.orc_limit_fluxes(self._ctx, float(dt))

    def rk3_update(self, alpha, beta, dt):
        self._L.orc_rk3_update(self._ctx, float(alpha), float(beta), float(dt))

    def copy_old(self):
        self._L.orc_copy_old(self._ctx)

    def set_flux_hook(self, fn):
        """While set, step() calls fn(stage) for each RK stage instead of the
        flux kernels (fn writes the 28 face planes through array()); None
        restores the kernels."""
        self._hook = FLUX_HOOK(lambda _user, stage: fn(int(stage))) if fn else None
        self._L.orc_set_flux_hook(self._ctx, self._hook if fn else FLUX_HOOK(), None)

    def opcounts(self, reset: bool = False) -> dict:
        """Floating-point operations of each kind since the last reset (the
        counters are per library, not per context); counting builds only."""
        if not self.counting:
            raise RuntimeError("opcounts() needs Oracle(..., counting=True)")
        out = (C.c_uint64 * len(OPCOUNT_KINDS))()
        self._L.orc_opcount_read(out)
        if reset:
            self._L.orc_opcount_reset()
        return dict(zip(OPCOUNT_KINDS, (int(v) for v in out)))


def max_threads() -> int:
    return int(lib().orc_max_threads())


def set_threads(n: int):
    lib().orc_set_threads(int(n))
