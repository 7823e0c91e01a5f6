"""The state picture (io.save_png, SC:1728-1769): a valid PNG whose right panel
is alpha1 and whose left panel carries the alpha1 = 0.5 contour."""
import struct
import zlib

import numpy as np

from wamp import cases, io


def _decode(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype = hdr[:4]
    assert depth == 8 and ctype == 2
    raw = zlib.decompress(idat)
    rows = [raw[r * (1 + 3 * w):(r + 1) * (1 + 3 * w)] for r in range(h)]
    assert all(row[0] == 0 for row in rows)
    return np.array([np.frombuffer(row[1:], np.uint8).reshape(w, 3) for row in rows])


def test_png_panels(tmp_path):
    cfg, U = cases.shock_cylinder(64, 32)
    g = cfg.grid
    path = io.save_png(str(tmp_path / "mp5_000000.png"), g, U)
    img = _decode(path)
    gap = max(2, g.nx // 64)
    assert img.shape == (g.ny, 2 * g.nx + gap, 3)
    si, sj = g.interior()
    al = U[5][si, sj]
    right = img[:, g.nx + gap:, 0].astype(float)
    assert np.abs(right - np.rint(255.0 * (1.0 - al.T))).max() <= 0.5
    left = img[:, :g.nx]
    red = (left[..., 0] == 220) & (left[..., 1] == 20)
    # the contour ring of the r = 0.5 cylinder centred at (1.5, 1.0): about 2*pi*r/dx cells
    assert red.sum() > 2 * np.pi * 0.5 / g.dx
    rows, cols = np.nonzero(red)
    x, y = g.x_centres()[si][cols], g.y_centres()[sj][rows]
    assert np.abs(np.hypot(x - 1.5, y - 1.0) - 0.5).max() <= 2 * g.dx
    # the water-air interface has the largest density gradient: darkest blue on the rim
    blue = left[..., 2].astype(float)
    blue[red] = 255.0
    r0, c0 = np.unravel_index(np.argmin(blue), blue.shape)
    assert abs(np.hypot(g.x_centres()[si][c0] - 1.5, g.y_centres()[sj][r0] - 1.0) - 0.5) <= 2 * g.dx
    # with the scale set for the shock (jump 2.2 over 2 dx), the shock column is dark too
    img2 = _decode(io.save_png(str(tmp_path / "b.png"), g, U, vmin=0.0, vmax=2.2 / (2 * g.dx)))
    shock = np.argmin(np.abs(g.x_centres()[si] - 0.8))
    assert img2[g.ny // 8, shock - 1:shock + 1, 2].min() < 150
