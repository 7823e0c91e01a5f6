"""Error measures between two solutions of the same case (mixed vs FP64).

Reported quantities: relative L1 error of density, pressure and volume
fraction over the interior, and the shift of the interface centroid
(sum a1*x / sum a1, same for y) in cells.
"""
from __future__ import annotations

import numpy as np

from .config import CaseConfig


def _pressure(cfg: CaseConfig, U):
    e = cfg.eos
    al = U[5]
    rho = U[0] + U[1]
    u, v = U[2] / rho, U[3] / rho
    a2 = 1.0 - al
    gb = al / (e.gamma1 - 1.0) + a2 / (e.gamma2 - 1.0)
    pb = al * e.gamma1 * e.pinf1 / (e.gamma1 - 1.0) + a2 * e.gamma2 * e.pinf2 / (e.gamma2 - 1.0)
    return (U[4] - pb - 0.5 * rho * (u * u + v * v)) / gb


def rel_l1(a, b) -> float:
    den = float(np.abs(b).sum())
    return float(np.abs(a - b).sum()) / (den if den > 0 else 1.0)


def centroid(cfg: CaseConfig, U):
    g = cfg.grid
    si, sj = g.interior()
    a = U[5][si, sj]
    x = g.x_centres()[si][:, None]
    y = g.y_centres()[sj][None, :]
    m = a.sum()
    return float((a * x).sum() / m), float((a * y).sum() / m)


def compare(cfg: CaseConfig, U, Uref) -> dict:
    """Errors of state U against the reference state Uref, both (6, NXG, NYG) FP64."""
    g = cfg.grid
    si, sj = g.interior()
    rho, rho_r = (U[0] + U[1])[si, sj], (Uref[0] + Uref[1])[si, sj]
    p, p_r = _pressure(cfg, U)[si, sj], _pressure(cfg, Uref)[si, sj]
    cx, cy = centroid(cfg, U)
    rx, ry = centroid(cfg, Uref)
    return {
        "rho": rel_l1(rho, rho_r),
        "p": rel_l1(p, p_r),
        "alpha1": rel_l1(U[5][si, sj], Uref[5][si, sj]),
        "centroid_shift_cells": float(np.hypot((cx - rx) / g.dx, (cy - ry) / g.dy)),
        "finite": bool(np.isfinite(U[:, si, sj]).all()),
    }
