"""npz output and restart in the reference's format.

save_npz            SC:1712-1726 (keys time, x, y, rho, ar1, ar2, u, v, p, alpha1;
                    interior arrays of shape (NX, NY), axis 0 = x)
infer_restart_step  SC:1773-1776 (digits after the last '_' of the file stem)
load_restart        SC:1779-1821 (required keys, shape check, alpha clipped to
                    [1e-8, 1-1e-8], E rebuilt from p)

The reference rebuilds E with its hard-coded water/air constants
(SC:1811-1812), which equal its module EOS (SC:40-43); here E is rebuilt with
the run's EOS, both phase terms included.
"""
from __future__ import annotations

import os

import numpy as np

from .config import CaseConfig, Grid

REQUIRED = ["ar1", "ar2", "u", "v", "p", "alpha1"]


def output_name(prefix: str, step: int, ext: str = "npz") -> str:
    return f"{prefix}_{step:06d}.{ext}"


def save_npz(path: str, grid: Grid, U: np.ndarray, P: np.ndarray, t: float) -> str:
    """Write the interior of state U (6, NXG, NYG) and prims P (3, NXG, NYG) = (u, v, p)."""
    si, sj = grid.interior()
    ar1 = U[0][si, sj]
    ar2 = U[1][si, sj]
    np.savez_compressed(path,
                        time=np.float64(t),
                        x=grid.x_centres()[si], y=grid.y_centres()[sj],
                        rho=ar1 + ar2, ar1=ar1, ar2=ar2,
                        u=P[0][si, sj], v=P[1][si, sj], p=P[2][si, sj],
                        alpha1=U[5][si, sj])
    return path


def infer_restart_step(path: str) -> int:
    stem = os.path.splitext(os.path.basename(path))[0]
    tail = stem.split("_")[-1]
    return int(tail) if tail.isdigit() else 0


def load_restart(path: str, cfg: CaseConfig):
    """Return (U (6, NXG, NYG) float64 with zero ghosts, time)."""
    g, e = cfg.grid, cfg.eos
    with np.load(path, allow_pickle=False) as data:
        for key in REQUIRED:
            if key not in data:
                raise KeyError(f"Restart file {path} is missing '{key}'")
        f = {k: np.asarray(data[k], dtype=np.float64) for k in REQUIRED}
        t = float(data["time"]) if "time" in data else 0.0
    if f["ar1"].shape != (g.nx, g.ny):
        raise ValueError(f"Restart shape {f['ar1'].shape} does not match {(g.nx, g.ny)}")
    al = np.clip(f["alpha1"], 1.0e-8, 1.0 - 1.0e-8)
    rho = f["ar1"] + f["ar2"]
    ru, rv = rho * f["u"], rho * f["v"]
    a2 = 1.0 - al
    gb = al / (e.gamma1 - 1.0) + a2 / (e.gamma2 - 1.0)
    pb = al * e.gamma1 * e.pinf1 / (e.gamma1 - 1.0) + a2 * e.gamma2 * e.pinf2 / (e.gamma2 - 1.0)
    E = gb * f["p"] + pb + 0.5 * rho * (f["u"] * f["u"] + f["v"] * f["v"])
    U = np.zeros((6, g.nxg, g.nyg), dtype=np.float64)
    si, sj = g.interior()
    for m, a in enumerate((f["ar1"], f["ar2"], ru, rv, E, al)):
        U[m][si, sj] = a
    return U, t
