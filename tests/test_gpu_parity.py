"""HIP path vs the CPU oracle: face fluxes, one step, many steps (C-ABI through ctypes)."""
import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu

SCHEMES = ["sc", "m10", "mp5"]


def _setup(scheme, nx=256, ny=128, presteps=0):
    from oracle.oracle import Oracle
    from wamp import cases
    from wamp.config import SCHEMES as S

    cfg, U0 = cases.shock_cylinder(nx, ny, scheme=S[scheme])
    orc = Oracle(cfg)
    orc.set_state(U0)
    orc.prepare()
    if presteps:
        orc.run(presteps)
    return cfg, orc


def _interior_rel(cfg, got, ref):
    si, sj = cfg.grid.interior()
    return [rel_l1(got[m][si, sj], ref[m][si, sj]) for m in range(got.shape[0])]


@pytest.mark.parametrize("scheme", SCHEMES)
def test_face_fluxes_match_oracle(gpu_ok, scheme):
    """compute_fluxes + compute_low_fluxes (SC:1919-1947) on a shocked state."""
    from wamp.solver import Solver

    cfg, orc = _setup(scheme, presteps=60)
    U = orc.state()
    orc.apply_bc()
    orc.update_prims()
    orc.apply_bc()
    orc.compute_fluxes()
    ref = orc.fluxes()
    s = Solver(cfg)
    s.set_state(U)
    got = s.fluxes()
    g = cfg.grid
    G, NX, NY = g.g, g.nx, g.ny
    xs = (slice(None), slice(G - 1, G + NX), slice(G, G + NY))
    ys = (slice(None), slice(G, G + NX), slice(G - 1, G + NY))
    worst = 0.0
    for m in range(28):
        sl = xs if (m // 7) % 2 == 0 else ys
        a, b = got[m][sl[1:]], ref[m][sl[1:]]
        err = np.abs(a - b) / (np.abs(b) + 1e-300 + np.abs(b).max() * 1e-14)
        worst = max(worst, float(err.max()))
        assert np.isfinite(a).all(), m
    print(f"{scheme}: max face-flux relative error {worst:.3e}")
    assert worst < 1e-9


@pytest.mark.parametrize("scheme", SCHEMES)
def test_one_step_matches_oracle(gpu_ok, scheme):
    from wamp.solver import Solver

    cfg, orc = _setup(scheme, presteps=40)
    U = orc.state()
    dt = orc.get_dt()
    orc.step(dt)
    ref = orc.state()
    s = Solver(cfg)
    s.set_state(U)
    s.set_dt(dt)
    s.step_device()
    got = s.state()
    errs = _interior_rel(cfg, got, ref)
    print(scheme, ["%.2e" % e for e in errs])
    assert max(errs) <= 1e-12
    dt_next = float(s.dt_buf[0].item())
    assert abs(dt_next - orc.get_dt()) <= 1e-14 * abs(dt_next)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_200_steps_match_oracle(gpu_ok, scheme):
    """The 256x128 shock-cylinder case, 200 SSP-RK3 steps (shock meets the cylinder)."""
    from wamp.solver import Solver

    cfg, orc = _setup(scheme)
    U0 = orc.state()
    dts_ref = orc.run(200)
    ref = orc.state()
    s = Solver(cfg)
    s.set_state(U0)
    dts = s.run(200)
    got = s.state()
    errs = _interior_rel(cfg, got, ref)
    print(scheme, ["%.2e" % e for e in errs], "dt rel", float(np.abs(dts - dts_ref).max() / dts_ref.max()))
    assert max(errs) <= 1e-9


@pytest.mark.parametrize("nx, ny", [(1, 1), (1, 6), (6, 1), (3, 2), (37, 23), (130, 3)])
@pytest.mark.parametrize("scheme", SCHEMES)
def test_degenerate_and_ragged_grids_match_oracle(gpu_ok, scheme, nx, ny):
    """One cell, single rows and columns, and sizes that are no multiple of any
    tile or strip: 20 steps against the oracle (every ghost cell maps to an
    edge cell; most tiles are partly outside the grid)."""
    from wamp.solver import Solver

    cfg, orc = _setup(scheme, nx=nx, ny=ny)
    U0 = orc.state()
    dts_ref = orc.run(20)
    ref = orc.state()
    s = Solver(cfg)
    s.set_state(U0)
    dts = s.run(20)
    got = s.state()
    s.close()
    errs = _interior_rel(cfg, got, ref)
    assert np.isfinite(got).all()
    assert max(errs) <= 1e-9, errs
    assert np.abs(dts - dts_ref).max() <= 1e-12 * dts_ref.max()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("scheme", ["m10", "mp5"])
def test_limiter_kernels_agree_bitwise(gpu_ok, scheme, precision, monkeypatch):
    """The register-marching and the LDS-tile limiter/update kernels do the same
    arithmetic, on FP64 face planes and on the mixed storage."""
    import dataclasses

    from wamp.solver import Solver

    cfg, orc = _setup(scheme, nx=200, ny=120, presteps=30)
    cfg = dataclasses.replace(cfg, precision=precision)
    U = orc.state()
    dt = orc.get_dt()
    outs = {}
    segs = ("4", "8", "16", "20", "26", "28", "32", "64", "3", "37", "0")
    # the march that forms the first-order fluxes itself reads FP64 face storage only
    segs_lo = ("4", "8", "16", "26", "32", "48", "64", "3", "37", "0") if precision == 0 else ()
    for kind, seg in ((("0", "32"),) + tuple(("1", s) for s in segs)
                      + tuple(("2", s) for s in segs_lo)):
        monkeypatch.setenv("WAMP_LU_KIND", kind)
        monkeypatch.setenv("WAMP_LU_SEG", seg)
        s = Solver(cfg)
        s.set_state(U)
        s.set_dt(dt)
        s.step_device()
        outs[(kind, seg)] = (s.state(), float(s.dt_buf[0].item()))
        s.close()
    ref, dref = outs[("0", "32")]
    si, sj = cfg.grid.interior()
    for k, (st, dn) in outs.items():
        assert np.array_equal(st[:, si, sj], ref[:, si, sj]), k
        assert dn == dref, k


@pytest.mark.parametrize("nx, ny", [(1, 1), (6, 1), (3, 2), (37, 23), (130, 3), (9, 61)])
@pytest.mark.parametrize("precision", [0, 1])
def test_march_limiter_on_degenerate_grids(gpu_ok, nx, ny, precision, monkeypatch):
    """The marching limiter (chosen only above 2^18 cells) forced onto grids of
    one cell, one row, and widths that end mid-strip or span two strips: 10
    steps bitwise equal to the LDS-tile limiter, in FP64 and mixed storage."""
    import dataclasses

    from wamp.solver import Solver

    cfg, orc = _setup("m10", nx=nx, ny=ny)
    cfg = dataclasses.replace(cfg, precision=precision)
    U = orc.state()
    outs = {}
    for kind, seg in (("0", "0"), ("1", "4"), ("1", "26"), ("2", "4"), ("2", "26"), ("1", "0"),
                      ("2", "0")):
        monkeypatch.setenv("WAMP_LU_KIND", kind)
        monkeypatch.setenv("WAMP_LU_SEG", seg)
        s = Solver(cfg)
        s.set_state(U)
        s.run(10)
        outs[(kind, seg)] = s.state()
        s.close()
    si, sj = cfg.grid.interior()
    ref = outs[("0", "0")]
    assert np.isfinite(ref[:, si, sj]).all()
    for k, st in outs.items():
        assert np.array_equal(st[:, si, sj], ref[:, si, sj]), k


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (0, 0, 0, 0), (1, 0, 0, 1)])
@pytest.mark.parametrize("scheme", SCHEMES)
def test_boundary_variants_match_oracle(gpu_ok, scheme, bc):
    """Slip walls and zero-gradient copies on every side in every combination
    class (the reference's own setting is zero-gradient x, slip y): 40 steps of
    the 128x64 shock-cylinder case against the oracle."""
    import dataclasses

    from oracle.oracle import Oracle
    from wamp import cases
    from wamp.config import SCHEMES as S
    from wamp.solver import Solver

    cfg, U0 = cases.shock_cylinder(128, 64, scheme=S[scheme])
    cfg = dataclasses.replace(cfg, grid=dataclasses.replace(cfg.grid, bc=bc))
    orc = Oracle(cfg)
    orc.set_state(U0)
    orc.prepare()
    U = orc.state()
    orc.run(40)
    ref = orc.state()
    orc.close()
    s = Solver(cfg)
    s.set_state(U)
    s.run(40)
    got = s.state()
    s.close()
    errs = _interior_rel(cfg, got, ref)
    assert np.isfinite(got).all()
    assert max(errs) <= 1e-9, (bc, errs)


@pytest.mark.parametrize("lu_kind", ["0", "1", "2"])
@pytest.mark.parametrize("precision", [0, 1])
def test_padded_row_stride_gives_the_same_result(gpu_ok, precision, lu_kind, monkeypatch):
    """A state whose rows are padded (ld = ny + 2g + 13, not a multiple of
    anything) steps bitwise like the unpadded one, with either limiter kernel."""
    import dataclasses

    from wamp import cases
    from wamp.config import SCHEME_M10
    from wamp.solver import Solver

    monkeypatch.setenv("WAMP_LU_KIND", lu_kind)
    cfg, U0 = cases.shock_cylinder(200, 120, scheme=SCHEME_M10)
    cfg = dataclasses.replace(cfg, precision=precision)
    outs = []
    for ld in (None, cfg.grid.nyg + 13):
        s = Solver(cfg, ld=ld)
        s.set_state(U0)
        s.prepare()
        s.run(20)
        outs.append(s.state())
        s.close()
    si, sj = cfg.grid.interior()
    assert np.array_equal(outs[0][:, si, sj], outs[1][:, si, sj])


@pytest.mark.parametrize("eos", [(4.4, 6000.0, 2.0, 100.0), (1.67, 0.0, 1.4, 0.0), (6.12, 3430.0, 1.4, 0.0)])
@pytest.mark.parametrize("scheme", ["m10", "mp5"])
def test_eos_variants_match_oracle(gpu_ok, scheme, eos):
    """Other stiffened-gas pairs: two stiffened phases (both p_inf > 0), two ideal
    gases, a stiffer liquid.  20 steps of a 128x64 shock-droplet state against
    the oracle."""
    from oracle.oracle import Oracle
    from wamp import cases
    from wamp.config import EOS, SCHEMES as S, CaseConfig, Grid
    from wamp.solver import Solver

    e = EOS(*eos)
    grid = Grid(128, 64, 0.0, 4.0, 0.0, 2.0)
    cfg = CaseConfig(grid, e, S[scheme], name="eos_variant")
    U0 = cases.shock_droplets(grid, e, 2.4, 0.8, [(1.5, 1.0, 0.5, 0.2)])
    orc = Oracle(cfg)
    orc.set_state(U0)
    orc.prepare()
    U = orc.state()
    orc.run(20)
    ref = orc.state()
    orc.close()
    s = Solver(cfg)
    s.set_state(U)
    s.run(20)
    got = s.state()
    s.close()
    errs = _interior_rel(cfg, got, ref)
    assert np.isfinite(got).all()
    assert max(errs) <= 1e-9, (eos, errs)
