"""npz output and restart in the reference's format.

save_npz            SC:1712-1726 (keys time, x, y, rho, ar1, ar2, u, v, p, alpha1;
                    interior arrays of shape (NX, NY), axis 0 = x)
infer_restart_step  SC:1773-1776 (digits after the last '_' of the file stem)
load_restart        SC:1779-1821 (required keys, shape check, alpha clipped to
                    [1e-8, 1-1e-8], E rebuilt from p)
save_png            SC:1728-1769 (density-gradient and alpha1 panels; written
                    with numpy and zlib, no plotting library)

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


# ---------------------------------------------------------------------------
# Picture of a state: save_png SC:1728-1769 without matplotlib
# ---------------------------------------------------------------------------
def _png_bytes(rgb: np.ndarray) -> bytes:
    """8-bit RGB image (H, W, 3) as a PNG file (one IDAT, no filtering)."""
    import struct
    import zlib

    h, w, _ = rgb.shape
    raw = b"".join(b"\x00" + rgb[r].astype(np.uint8).tobytes() for r in range(h))

    def chunk(kind: bytes, data: bytes) -> bytes:
        return (struct.pack(">I", len(data)) + kind + data
                + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF))

    ihdr = struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(raw, 6))
            + chunk(b"IEND", b""))


def _blues(t: np.ndarray) -> np.ndarray:
    """t in [0, 1] -> RGB from near-white to dark blue (the reference's Blues map ends)."""
    lo = np.array([247.0, 251.0, 255.0])
    hi = np.array([8.0, 48.0, 107.0])
    t = np.clip(t, 0.0, 1.0)[..., None]
    return lo + (hi - lo) * t


def save_png(path: str, grid: Grid, U: np.ndarray, vmin: float | None = None,
             vmax: float | None = None) -> str:
    """Two panels, as SC:1745-1769: left the density-gradient magnitude
    (np.gradient central differences, SC:1735-1737) on a white-to-blue scale
    with the alpha1 = 0.5 contour in red; right alpha1 on a white-to-black
    scale.  Image rows are j, columns i (the transposed field, SC:1740); the
    gradient scale is [vmin, vmax] (default: 0 to the field's maximum)."""
    si, sj = grid.interior()
    rho = (U[0] + U[1])[si, sj]
    al = U[5][si, sj]
    xg = grid.x_centres()[si]
    yg = grid.y_centres()[sj]
    grad = np.sqrt(np.gradient(rho, xg, axis=0) ** 2 + np.gradient(rho, yg, axis=1) ** 2)
    lo = 0.0 if vmin is None else vmin
    hi = float(grad.max()) if vmax is None else vmax
    left = _blues((grad.T - lo) / (hi - lo if hi > lo else 1.0))
    # alpha1 = 0.5 contour: cells whose alpha1 - 0.5 changes sign towards a neighbour
    s = np.sign(al - 0.5)
    edge = np.zeros_like(al, dtype=bool)
    edge[:-1, :] |= s[:-1, :] != s[1:, :]
    edge[:, :-1] |= s[:, :-1] != s[:, 1:]
    left[edge.T] = (220.0, 20.0, 20.0)
    right = np.repeat((255.0 * (1.0 - np.clip(al.T, 0.0, 1.0)))[..., None], 3, axis=2)
    gap = np.full((al.shape[1], max(2, al.shape[0] // 64), 3), 255.0)
    img = np.concatenate([left, gap, right], axis=1)
    with open(path, "wb") as fh:
        fh.write(_png_bytes(np.rint(img)))
    return path
