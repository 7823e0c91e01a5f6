"""Each closure-level entry point of the C-ABI against the oracle's closure of
the same name, and the emergency repair (SC:918-958) driven on the device.

  wamp_enforce_admissible  vs orc_enforce_admissible  (SC:1907-1911)
  wamp_cons_to_prim        vs orc_update_prims        (SC:1913-1917)
  wamp_apply_bc            vs orc_apply_bc            (SC:1899-1905)
  wamp_get_dt              vs orc_get_dt              (SC:1986-1995)
  wamp_prepare             vs orc_prepare             (SC:1998-2001)
  wamp_limit_update        vs orc_limit_fluxes + orc_rk3_update +
                              orc_enforce_admissible + orc_update_prims
                              (SC:1949-1984, 1907-1917)

The repair branches -- NaN partial density or momentum, a negative partial
density, energy below the p_inf floor, alpha outside [0, 1] -- are injected
into the device state (the cells of the oracle's own repair test) and driven
through the repair kernel, the initial sequence, both limiter/update kernels
(LDS tile and march) inside a full step, and both precisions.
"""
import dataclasses

import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu


def _p6(t):
    from wamp import _lib

    return _lib.ptrs6([t[m].data_ptr() for m in range(6)])


def _stream():
    import torch

    return torch.cuda.current_stream().cuda_stream


def _bad_cells(cfg, U):
    """Inject every repair branch: NaN a1*rho1, negative a2*rho2, E far below
    the floor, NaN rho*u, NaN E, NaN rho*v, alpha above 1 and below 0."""
    g = cfg.grid
    U = U.copy()
    i, j = g.g + 5, g.g + 7
    U[0, i, j] = np.nan
    U[1, i + 1, j] = -3.0
    U[4, i + 2, j] = -1e9
    U[2, i + 3, j] = np.nan
    U[4, i + 4, j + 1] = np.nan
    U[3, i + 5, j + 2] = np.nan
    U[5, i + 6, j] = 1.7
    U[5, i + 7, j + 3] = -0.4
    cells = [(i + q, j) for q in range(8)] + [(i + 4, j + 1), (i + 5, j + 2), (i + 7, j + 3)]
    return U, cells


def _case(scheme="sc", nx=48, ny=24):
    from wamp import cases
    from wamp.config import SCHEMES

    return cases.shock_cylinder(nx, ny, scheme=SCHEMES[scheme])


def _assert_close(got, ref, what, rtol=1e-15):
    assert got.shape == ref.shape
    nan_g, nan_r = np.isnan(got), np.isnan(ref)
    assert np.array_equal(nan_g, nan_r), (what, "NaN pattern differs")
    a, b = got[~nan_r], ref[~nan_r]
    err = np.abs(a - b) / (np.abs(b) + 1e-300)
    worst = float(err.max()) if err.size else 0.0
    assert worst <= rtol, (what, worst)
    return worst


def test_enforce_admissible_matches_oracle(gpu_ok):
    """Over the padded grid, ghost cells included, with every repair branch."""
    from oracle.oracle import Oracle
    from wamp import _lib
    from wamp.solver import Solver

    cfg, U0 = _case()
    U, cells = _bad_cells(cfg, U0)
    U[:, 1, 2] = U[:, cells[0][0], cells[0][1]]        # a bad ghost cell too
    orc = Oracle(cfg)
    orc.set_state(U)
    orc.enforce_admissible()
    ref = orc.state()
    s = Solver(cfg)
    s.set_state(U)
    _lib.check(s.L.wamp_enforce_admissible(s.ctx, s._u, _stream()), "wamp_enforce_admissible")
    got = s.state()
    s.close()
    assert np.isfinite(got).all()
    worst = _assert_close(got, ref, "enforce_admissible")
    for (i, j) in cells:
        assert got[0, i, j] >= 1e-10 and got[1, i, j] >= 1e-10
    print(f"enforce_admissible: max relative difference {worst:.2e}")


def test_cons_to_prim_matches_oracle(gpu_ok):
    """u, v, p over the padded grid and alpha clamped to [1e-8, 1-1e-8] in place."""
    import torch

    from oracle.oracle import Oracle
    from wamp import _lib
    from wamp.solver import Solver

    cfg, U0 = _case("m10")
    U = U0.copy()
    g = cfg.grid
    U[:, :g.g, :] = U[:, g.g:2 * g.g, :]        # ghosts hold states (no zero densities)
    U[:, -g.g:, :] = U[:, -2 * g.g:-g.g, :]
    U[:, :, :g.g] = U[:, :, g.g:2 * g.g]
    U[:, :, -g.g:] = U[:, :, -2 * g.g:-g.g]
    U[5, g.g + 3, g.g + 3] = 1.2
    U[5, g.g + 4, g.g + 3] = 1e-12
    U[5, 0, 0] = -0.5
    orc = Oracle(cfg)
    orc.set_state(U)
    orc.update_prims()
    ref_u, ref_p = orc.state(), orc.prims()
    s = Solver(cfg)
    s.set_state(U)
    P = torch.zeros((3, g.nxg, s.ld), dtype=torch.float64, device=s.device)
    _lib.check(s.L.wamp_cons_to_prim(s.ctx, s._u, _lib.ptrs3([P[m].data_ptr() for m in range(3)]),
                                     _stream()), "wamp_cons_to_prim")
    got_u, got_p = s.state(), P[:, :, :g.nyg].cpu().numpy()
    s.close()
    assert np.array_equal(got_u, ref_u)
    _assert_close(got_p, ref_p, "prims")
    assert got_u[5, g.g + 3, g.g + 3] == 1.0 - 1e-8 and got_u[5, 0, 0] == 1e-8


@pytest.mark.parametrize("bc", [(0, 0, 1, 1), (1, 1, 0, 0), (1, 0, 0, 1), (0, 0, 0, 0)])
def test_apply_bc_matches_oracle(gpu_ok, bc):
    """Every ghost cell of the state and of the primitive arrays."""
    import torch

    from oracle import oracle as O
    from wamp import _lib
    from wamp.solver import Solver

    cfg, _ = _case(nx=23, ny=17)
    cfg = dataclasses.replace(cfg, grid=dataclasses.replace(cfg.grid, bc=bc))
    g = cfg.grid
    rng = np.random.default_rng(7)
    U = rng.uniform(0.5, 2.0, (6, g.nxg, g.nyg))
    Pr = rng.uniform(-1.0, 1.0, (3, g.nxg, g.nyg))
    orc = O.Oracle(cfg)
    orc.set_state(U)
    for m in range(3):
        orc.array(O.UP + m)[...] = Pr[m]
    orc.apply_bc()
    ref_u, ref_p = orc.state(), orc.prims()
    s = Solver(cfg)
    s.set_state(U)
    P = torch.from_numpy(np.ascontiguousarray(Pr)).to(s.device)
    _lib.check(s.L.wamp_apply_bc(s.ctx, s._u, _lib.ptrs3([P[m].data_ptr() for m in range(3)]),
                                 _stream()), "wamp_apply_bc")
    got_u, got_p = s.state(), P.cpu().numpy()
    s.close()
    assert np.array_equal(got_u, ref_u)
    assert np.array_equal(got_p, ref_p)


@pytest.mark.parametrize("scheme", ["sc", "m10"])
def test_prepare_and_dt_match_oracle_with_repairs(gpu_ok, scheme):
    """The initial sequence on a state with bad cells, then get_dt."""
    from oracle.oracle import Oracle
    from wamp.solver import Solver

    cfg, U0 = _case(scheme)
    U, cells = _bad_cells(cfg, U0)
    orc = Oracle(cfg)
    orc.set_state(U)
    orc.prepare()
    ref = orc.state()
    dt_ref = orc.get_dt()
    s = Solver(cfg)
    s.set_state(U)
    s.prepare()
    got = s.state()
    dt = s.get_dt()
    s.close()
    si, sj = cfg.grid.interior()
    _assert_close(got[:, si, sj], ref[:, si, sj], "prepare")
    assert abs(dt - dt_ref) <= 1e-15 * dt_ref


def _force_repairs(orc, g):
    """Face values that drive every repair branch of the stage update: a huge
    partial-density and energy outflow (first- and high-order alike, so theta
    = 0 cannot save the cell), a NaN high-order momentum flux and a NaN
    first-order partial-density flux (theta's NaN branch, then NaN in the
    update), and a volume-fraction flux that pushes alpha out of [0, 1]."""
    from oracle import oracle as O

    G = g.g
    for base in (O.FX0, O.FXL0):
        orc.array(base)[G + 8, G + 5] = 1e6
    for base in (O.FX0 + 4, O.FXL0 + 4):
        orc.array(base)[G + 14, G + 9] = 1e9
    orc.array(O.FX0 + 2)[G + 20, G + 12] = np.nan
    orc.array(O.FXL0 + 1)[G + 26, G + 15] = np.nan
    for base in (O.GY0 + 5, O.GYL0 + 5):
        orc.array(base)[G + 30, G + 17] = 1e3


@pytest.mark.parametrize("lu_kind", ["0", "1"])
@pytest.mark.parametrize("forced", [False, True])
def test_limit_update_matches_oracle(gpu_ok, lu_kind, forced, monkeypatch):
    """wamp_limit_update (the three Wong stages, the RK combine of stage 2,
    repair and the alpha clamp) on the oracle's own face fluxes, against the
    oracle's limit_fluxes + rk3_update + enforce_admissible + update_prims,
    with either limiter/update kernel (LDS tile, march).  `forced` plants face
    values that make the update leave the admissible set in every way the
    repair handles (negative and NaN partial densities, NaN momentum, energy
    below the p_inf floor, alpha outside [0, 1])."""
    import ctypes as C

    import torch

    from oracle import oracle as O
    from wamp import _lib
    from wamp.solver import Solver

    monkeypatch.setenv("WAMP_LU_KIND", lu_kind)
    cfg, U0 = _case("sc")
    g = cfg.grid
    orc = O.Oracle(cfg)
    orc.set_state(U0)
    orc.prepare()
    orc.run(12)
    Un = orc.state()
    dt = orc.get_dt()
    orc.step(0.5 * dt)                      # a different state as the stage input
    U = orc.state()
    # oracle: stage 2 of a step from Un whose stage input is U
    orc.set_state(Un)
    orc.copy_old()
    orc.set_state(U)
    orc.apply_bc()
    orc.update_prims()
    orc.apply_bc()
    orc.compute_fluxes()
    if forced:
        _force_repairs(orc, g)
    F = orc.fluxes()
    orc.limit_fluxes(dt)
    orc.rk3_update(0.75, 0.25, dt)
    orc.enforce_admissible()
    orc.update_prims()
    ref = orc.state()
    orc.close()

    s = Solver(cfg)
    dev = s.device
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    Un_d, U_d = t(Un), t(U)
    Uo = torch.zeros_like(U_d)
    F_d = t(F)
    dt_d = torch.tensor([dt], dtype=torch.float64, device=dev)
    _lib.check(s.L.wamp_limit_update(s.ctx, _p6(Un_d), _p6(U_d),
                                     _lib.ptrs28([F_d[m].data_ptr() for m in range(28)]), _p6(Uo),
                                     0.75, 0.25, dt_d.data_ptr(), None, _stream()),
               "wamp_limit_update")
    got = Uo.cpu().numpy()
    s.close()
    si, sj = g.interior()
    a, b = got[:, si, sj], ref[:, si, sj]
    assert np.isfinite(a).all() and np.isfinite(b).all()
    errs = [rel_l1(a[m], b[m]) for m in range(6)]
    print(f"lu_kind {lu_kind}, forced {forced}: rel L1", ["%.1e" % e for e in errs])
    assert max(errs) <= 1e-13
    assert a[0].min() >= 1e-10 and a[1].min() >= 1e-10
    if forced:
        G = g.g
        assert got[0, G + 8, G + 5] == 1e-10                     # density floor
        assert got[2, G + 20, G + 12] == 0.0 and got[2, G + 21, G + 12] == 0.0   # NaN momentum
        assert got[1, G + 26, G + 15] == 1e-10                   # NaN partial density
        assert got[5, G + 30, G + 17] == 1e-8 and got[5, G + 30, G + 18] == 1.0 - 1e-8
        eos = O.eos_struct(cfg.eos)
        rc2 = O.lib().orc_rho_ctilde2_cons(C.byref(eos), *[float(x) for x in got[:, G + 14, G + 9]])
        assert rc2 == pytest.approx(1e-8, rel=1e-3)              # energy reset to the floor
        for (m, i, j) in ((0, 8, 5), (4, 14, 9), (1, 26, 15)):
            assert got[m, G + i, G + j] == ref[m, G + i, G + j]


def test_limiter_kernels_agree_under_heavy_limiting(gpu_ok, monkeypatch):
    """wamp_limit_update on face fluxes whose high-order values are scaled by
    seeded random factors in [-29, 31], so that the Wong stages limit about
    15 % of the cells (SC:1397-1548), against the oracle's limit_fluxes +
    rk3_update + enforce_admissible + update_prims, with the LDS-tile kernel
    and the march at every segment length, all bitwise equal to each other.
    The first output row i0 of a march segment depends on theta0 three rows
    above it (update <- theta2(i0-1) <- theta1(i0-2) <- theta0(i0-3)), whose
    lower face the march has to load before its first front."""
    import torch

    from oracle import oracle as O
    from wamp import _lib
    from wamp.solver import Solver

    cfg, U0 = _case("m10", nx=160, ny=120)
    g = cfg.grid
    orc = O.Oracle(cfg)
    orc.set_state(U0)
    orc.prepare()
    orc.run(12)
    Un = orc.state()
    dt = orc.get_dt()
    orc.step(0.5 * dt)
    U = orc.state()
    orc.set_state(Un)
    orc.copy_old()
    orc.set_state(U)
    orc.apply_bc()
    orc.update_prims()
    orc.apply_bc()
    orc.compute_fluxes()
    rng = np.random.default_rng(7)
    for base in (O.FX0, O.GY0):
        for m in range(7):
            a = orc.array(base + m)
            a *= 1.0 + 30.0 * rng.uniform(-1.0, 1.0, size=a.shape)
    F = orc.fluxes()
    orc.limit_fluxes(dt)
    limited = float((orc.array(O.THETA)[g.interior()] < 1.0).mean())
    orc.rk3_update(0.75, 0.25, dt)
    orc.enforce_admissible()
    orc.update_prims()
    ref = orc.state()
    orc.close()
    assert limited > 0.05

    si, sj = g.interior()
    outs = {}
    # rows per wave are run-time values: also 1, odd counts, the context's own
    # choice ("0") and one segment per strip (1000 rows > nx)
    variants = ([("0", "0")] + [("1", s) for s in ("4", "8", "16", "20", "26", "28", "32", "64",
                                                   "1", "5", "77", "0", "1000")]
                + [("2", s) for s in ("4", "8", "16", "26", "32", "48", "64",
                                      "1", "5", "151", "0", "1000")])
    for kind, seg in variants:
        monkeypatch.setenv("WAMP_LU_KIND", kind)
        monkeypatch.setenv("WAMP_LU_SEG", seg)
        s = Solver(cfg)
        dev = s.device
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        Un_d, U_d, F_d = t(Un), t(U), t(F)
        Uo = torch.zeros_like(U_d)
        dt_d = torch.tensor([dt], dtype=torch.float64, device=dev)
        _lib.check(s.L.wamp_limit_update(s.ctx, _p6(Un_d), _p6(U_d),
                                         _lib.ptrs28([F_d[m].data_ptr() for m in range(28)]),
                                         _p6(Uo), 0.75, 0.25, dt_d.data_ptr(), None, _stream()),
                   "wamp_limit_update")
        outs[(kind, seg)] = Uo.cpu().numpy()[:, si, sj]
        s.close()
    b = ref[:, si, sj]
    a = outs[("0", "0")]
    assert np.isfinite(a).all() and np.isfinite(b).all()
    errs = [rel_l1(a[m], b[m]) for m in range(6)]
    print(f"limited fraction {limited:.3f}; LDS-tile kernel vs oracle: rel L1",
          ["%.1e" % e for e in errs])
    assert max(errs) <= 1e-13
    for k, st in outs.items():
        assert np.array_equal(st, a), k


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("lu_kind", ["0", "1", "2"])
def test_repaired_state_steps_like_oracle(gpu_ok, lu_kind, precision, monkeypatch):
    """The bad cells repaired by the initial sequence (SC:1998-2001), then one
    step with either limiter/update kernel.  FP64: within 1e-12 of the oracle.
    Mixed precision: the repair (FP64 in both modes) equals the oracle's, and
    the step stays finite and admissible; its distance from the oracle is only
    reported, because the repaired cells are extreme states (a partial density
    floored to 1e-10 beside momentum of order one) whose FP32 face arithmetic
    differs from FP64 at the percent level in a small field like rho*v."""
    from oracle.oracle import Oracle
    from wamp.solver import Solver

    monkeypatch.setenv("WAMP_LU_KIND", lu_kind)
    cfg, U0 = _case("sc")
    U, cells = _bad_cells(cfg, U0)
    orc = Oracle(cfg)
    orc.set_state(U)
    orc.prepare()
    repaired = orc.state()
    dt = orc.get_dt()
    orc.step(dt)
    ref = orc.state()
    orc.close()
    s = Solver(dataclasses.replace(cfg, precision=precision))
    s.set_state(U)
    s.prepare()
    si, sj = cfg.grid.interior()
    assert np.array_equal(s.state()[:, si, sj], repaired[:, si, sj])
    s.set_dt(dt)
    s.step_device()
    got = s.state()
    s.close()
    a, b = got[:, si, sj], ref[:, si, sj]
    assert np.isfinite(b).all() and np.isfinite(a).all()
    errs = [rel_l1(a[m], b[m]) for m in range(6)]
    print(f"precision {precision}, lu_kind {lu_kind}: rel L1", ["%.1e" % e for e in errs])
    if precision == 0:
        assert max(errs) <= 1e-12
    assert a[0].min() >= 1e-10 and a[1].min() >= 1e-10
    assert a[5].min() >= 1e-8 and a[5].max() <= 1.0 - 1e-8
    for (i, j) in cells:
        assert got[0, i, j] >= 1e-10 and got[1, i, j] >= 1e-10
