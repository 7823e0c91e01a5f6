"""Host side of the 3D extension (wamp/threed.py) on the CPU: the embeddings
of a 2D state, the grid geometry and the shock-sphere initial state."""
import numpy as np

from wamp import cases, threed
from wamp.config import SCHEMES


def test_xy_embedding_replicates_the_slice_with_zero_w():
    cfg2, U2 = cases.shock_cylinder(24, 12, scheme=SCHEMES["m10"])
    cfg3, U3 = threed.from_2d_xy(cfg2, U2, nz=3)
    g3 = cfg3.grid
    assert (g3.nx, g3.ny, g3.nz) == (24, 12, 3) and U3.shape == (7, g3.nxg, g3.nyg, g3.nzg)
    assert g3.dx == cfg2.grid.dx and g3.dy == cfg2.grid.dy and g3.bc[:4] == tuple(cfg2.grid.bc)
    for k in range(g3.nzg):
        for m2, m3 in enumerate(threed.XY_MAP):
            assert np.array_equal(U3[m3][:, :, k], U2[m2])
    assert not U3[4].any()


def test_xz_embedding_maps_y_to_z_and_v_to_w():
    cfg2, U2 = cases.shock_cylinder(24, 12, scheme=SCHEMES["sc"])
    cfg3, U3 = threed.from_2d_xz(cfg2, U2, ny=2)
    g3, g2 = cfg3.grid, cfg2.grid
    assert (g3.nx, g3.ny, g3.nz) == (24, 2, 12)
    assert g3.dz == g2.dy and g3.bc[4:] == (g2.bc[2], g2.bc[3]) and g3.bc[:2] == (g2.bc[0], g2.bc[1])
    for j in range(g3.nyg):
        assert np.array_equal(U3[4][:, j, :], U2[3])  # 2D v momentum as w momentum
        assert np.array_equal(U3[5][:, j, :], U2[4])
    assert not U3[3].any()


def test_shock_sphere_state_is_admissible_and_y_z_symmetric():
    cfg, U = threed.shock_sphere(12)
    g = cfg.grid
    assert (g.nx, g.ny, g.nz) == (24, 12, 12) and abs(g.dx - g.dy) < 1e-15 and g.dy == g.dz
    si, sj, sk = g.interior()
    I = U[:, si, sj, sk]
    assert (I[0] > 0).all() and (I[1] > 0).all() and ((I[6] > 0) & (I[6] < 1)).all()
    # ghosts are zero (the kernels never read them), the interior is y <-> z symmetric
    assert not U[:, : g.g].any() and not U[:, :, :, -g.g:].any()
    T = np.swapaxes(U, 2, 3)
    for m, mt in enumerate([0, 1, 2, 4, 3, 5, 6]):
        assert np.array_equal(U[m], T[mt])
    # water sphere in the middle, post-shock gas on the left
    c = g.g + g.nx * 3 // 8, g.g + g.ny // 2, g.g + g.nz // 2
    assert U[6][c] > 0.5 and U[2][g.g, c[1], c[2]] > 0.0


def test_npz_round_trip(tmp_path):
    """save_npz3 then load_restart3 gives back the interior state (E rebuilt from
    p with the same EOS, to round-off) and the time; ghosts stay zero."""
    cfg, U = threed.shock_sphere(8)
    path = threed.save_npz3(str(tmp_path / "s_000010.npz"), cfg.grid, U, 0.125, cfg.eos)
    V, t = threed.load_restart3(path, cfg)
    assert t == 0.125 and V.shape == U.shape
    si, sj, sk = cfg.grid.interior()
    for m in range(7):
        a, b = V[m][si, sj, sk], U[m][si, sj, sk]
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max()), m
    assert not V[:, : cfg.grid.g].any()
    with np.load(path) as d:
        assert sorted(d.files) == sorted(["time", "x", "y", "z", "rho", "ar1", "ar2", "u", "v", "w",
                                          "p", "alpha1"])
        assert d["rho"].shape == (cfg.grid.nx, cfg.grid.ny, cfg.grid.nz)


def test_restart_rejects_a_wrong_shape(tmp_path):
    import pytest

    cfg, U = threed.shock_sphere(8)
    path = threed.save_npz3(str(tmp_path / "a.npz"), cfg.grid, U, 0.0, cfg.eos)
    other, _ = threed.shock_sphere(6)
    with pytest.raises(ValueError):
        threed.load_restart3(path, other)
