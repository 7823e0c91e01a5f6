"""The CPU device: the reference's --cpu (SC:15-21, M10:14-20).

`CpuSolver` has the interface of `wamp.solver.Solver` with the state in host
memory (numpy) and libwamp.so's wamp_cpu_* entry points for every part of
the step: the FP64 step's own per-cell code compiled for the host, which
gives the FP64 GPU step's bits (subnormal values aside: below 2^-1022 the
kernels' shortened division may round differently).  It is chosen
explicitly (`--cpu` on the command line); nothing falls back to it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .config import CaseConfig
from .solver import advance


class CpuSolver:
    """One case on the host: conserved state (6 FP64 planes), a wamp_cpu context."""

    def __init__(self, cfg: CaseConfig, threads: int = 0, ld: int | None = None):
        if int(cfg.precision) != 0:
            raise _lib.WampError("the CPU device computes in FP64 only (precision 0)")
        self.L = _lib.load()
        self.cfg = cfg
        g = cfg.grid
        self.ld = int(ld) if ld else g.nyg
        if self.ld < g.nyg:
            raise ValueError("ld must be >= ny + 2g")
        self.U = np.zeros((6, g.nxg, self.ld), dtype=np.float64)
        self._u = _lib.ptrs6([self.U[m].ctypes.data for m in range(6)])
        grid = _lib.WampGrid(g.nx, g.ny, g.g, self.ld, g.dx, g.dy, (C.c_int32 * 4)(*g.bc))
        eos = _lib.WampEos(*cfg.eos.as_tuple())
        s = cfg.scheme
        sch = _lib.WampScheme(s.base, s.shear, s.vf_thinc, s.face_pp, 1 if s.repair else 0, s.cfl)
        ctx = C.c_void_p()
        _lib.check(self.L.wamp_cpu_ctx_create(C.byref(grid), C.byref(eos), C.byref(sch),
                                              int(threads), C.byref(ctx)), "wamp_cpu_ctx_create")
        self.ctx = ctx
        self.threads = int(self.L.wamp_cpu_threads(ctx))
        self.time = 0.0
        self.nstep = 0
        self._dt = C.c_double(0.0)

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            _lib.check(self.L.wamp_cpu_ctx_destroy(self.ctx), "wamp_cpu_ctx_destroy")
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state --------------------------------------------------------------
    def set_state(self, U):
        g = self.cfg.grid
        if tuple(U.shape[:2]) != (6, g.nxg) or U.shape[2] < g.nyg:
            raise ValueError(f"state shape {tuple(U.shape)} != {(6, g.nxg, g.nyg)}")
        self.U[:, :, : g.nyg] = np.asarray(U, dtype=np.float64)[:, :, : g.nyg]

    def state(self) -> np.ndarray:
        return self.U[:, :, : self.cfg.grid.nyg].copy()

    def prims(self) -> np.ndarray:
        """(3, NXG, NYG) u, v, p via cons_to_prim (SC:840-864) after a boundary fill."""
        g = self.cfg.grid
        P = np.zeros((3, g.nxg, self.ld), dtype=np.float64)
        p3 = _lib.ptrs3([P[m].ctypes.data for m in range(3)])
        self.apply_bc()
        _lib.check(self.L.wamp_cpu_cons_to_prim(self.ctx, self._u, p3), "wamp_cpu_cons_to_prim")
        return P[:, :, : g.nyg].copy()

    # -- the step's closures (SC:1894-1995) ----------------------------------
    def prepare(self):
        """Initial sequence SC:1998-2001 (repair + alpha clamp, then the ghost layers)."""
        _lib.check(self.L.wamp_cpu_prepare(self.ctx, self._u), "wamp_cpu_prepare")
        self.apply_bc()

    def apply_bc(self):
        _lib.check(self.L.wamp_cpu_apply_bc(self.ctx, self._u, None), "wamp_cpu_apply_bc")

    def get_dt(self) -> float:
        _lib.check(self.L.wamp_cpu_get_dt(self.ctx, self._u, C.byref(self._dt)), "wamp_cpu_get_dt")
        return self._dt.value

    def fluxes(self) -> np.ndarray:
        """The 28 face arrays of compute_fluxes + compute_low_fluxes (SC:1919-1947)."""
        g = self.cfg.grid
        F = np.zeros((28, g.nxg, self.ld), dtype=np.float64)
        _lib.check(self.L.wamp_cpu_compute_fluxes(self.ctx, self._u,
                                                  _lib.ptrs28([F[m].ctypes.data for m in range(28)])),
                   "wamp_cpu_compute_fluxes")
        return F[:, :, : g.nyg].copy()

    def step_dt(self, dt: float) -> float:
        """One SSP-RK3 step with time step dt; returns the next dt."""
        _lib.check(self.L.wamp_cpu_step(self.ctx, self._u, float(dt), C.byref(self._dt)),
                   "wamp_cpu_step")
        self.nstep += 1
        return self._dt.value

    def run(self, nsteps: int) -> np.ndarray:
        """n steps with dt from the state before each step (SC:2015, 2059); returns the dts."""
        dts = np.zeros(nsteps)
        dt = self.get_dt()
        for k in range(nsteps):
            dts[k] = dt
            dt = self.step_dt(dt)
        return dts

    def advance_to(self, t_end: float, save_times=(), on_save=None, log_every: int = 200,
                   log=print, max_steps: int | None = None):
        """Time loop of main() (SC:2010-2073): dt clipped to save times and t_end."""
        return advance(self, t_end, save_times, on_save, log_every, log, max_steps)
