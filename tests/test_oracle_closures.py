"""Pin the oracle's Ducros sensor, SSP-RK3 stage table, pressure floors and
flux-limiter blend against closed forms.

  * Ducros sensor (SC:875-916): for a linear velocity field u = a x - b y,
    v = b x + a y the dilatation factor is a^2/(a^2+b^2) exactly (1 for a pure
    dilatation, 0 for a pure rotation); the pressure factor
    |-p_-2+16p_-1-30p_0+16p_1-p_2| / (p_-2+16p_-1+30p_0+16p_1+p_2) is 0 for
    linear p and 24 k h^2 / (64 p_c + 40 k h^2) for p = p0 + k (x-x0)^2.
  * SSP-RK3 (SC:2028-2055, rk3_kernel SC:1553-1611): orc_step driven through
    its own three stages by fluxes whose residual is lambda*U multiplies U by
    1 + z + z^2/2 + z^3/6, z = -lambda dt (the Shu-Osher polynomial); the
    alpha source -alpha_prim div(u*) likewise with z = +kappa dt.  The stage
    table is the one inside orc_step, not one the test supplies.
  * Pressure floors (SC:180-187, SC:1335-1338, M10:1015-1021): a floored
    state lands on rho c~^2 = (p + pi)/(gamma - 1) = 1e-8.
  * Wong flux limiter (SC:1397-1548): theta lands the limited cell's
    forward-Euler partial density on the 1e-10 floor through its blended
    faces; a face shared by two limited cells takes the smaller theta.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from wamp import cases
from wamp.config import (BC_ZERO_GRADIENT, EOS, CaseConfig, Grid, SCHEME_M10, SCHEME_SC)

ZG = (BC_ZERO_GRADIENT,) * 4
AIR = EOS(1.4, 0.0, 1.4, 0.0)
L = O.lib()


# ------------------------------------------------------------ Ducros -------
def _ducros_oracle(nx, u, v, p):
    """Oracle sigma_x, sigma_y for primitive fields given on the padded grid."""
    grid = Grid(nx, nx, -0.5, 0.5, -0.5, 0.5, bc=ZG)
    cfg = CaseConfig(grid, AIR, SCHEME_SC)
    X = grid.x_centres()[:, None] * np.ones((1, grid.nyg))
    Y = grid.y_centres()[None, :] * np.ones((grid.nxg, 1))
    one = np.ones_like(X)
    U = cases.conserved_from_prims(AIR, 0.5 * one, one, one, u(X, Y), v(X, Y), p(X, Y))
    orc = O.Oracle(cfg)
    orc.set_state(U)
    orc.update_prims()          # no boundary fill: the fields hold everywhere
    orc.compute_fluxes()        # runs ducros_kernel (shear scheme)
    sx, sy = orc.array(O.SIGX).copy(), orc.array(O.SIGY).copy()
    P = orc.array(O.PP).copy()
    orc.close()
    return grid, sx, sy, P


def _inner(grid):
    # cells whose +-2 stencil lies inside the padded grid, away from its rim
    return slice(2, grid.nxg - 2), slice(2, grid.nyg - 2)


def test_ducros_dilatation_with_quadratic_pressure():
    a, p0, k, x0 = 3.0, 10.0, 5.0, 0.1
    grid, sx, sy, P = _ducros_oracle(
        24, lambda X, Y: a * X, lambda X, Y: a * Y, lambda X, Y: p0 + k * (X - x0) ** 2)
    si, sj = _inner(grid)
    h = grid.dx
    want = 24.0 * k * h * h / (64.0 * P[si, sj] + 40.0 * k * h * h)
    np.testing.assert_allclose(sx[si, sj], want, rtol=1e-9, atol=0)
    assert want.min() > 1e-4           # a real signal, not a vanishing one
    assert np.abs(sy[si, sj]).max() < 1e-13   # p does not vary along y


def test_ducros_pure_rotation_is_zero():
    b = 2.0
    grid, sx, sy, _ = _ducros_oracle(
        16, lambda X, Y: -b * Y, lambda X, Y: b * X, lambda X, Y: 3.0 + (X + 0.7) ** 2)
    si, sj = _inner(grid)
    assert np.all(sx[si, sj] == 0.0) and np.all(sy[si, sj] == 0.0)


def test_ducros_mixed_linear_field_and_linear_pressure():
    a, b, p0, k = 1.5, 2.5, 8.0, 4.0
    grid, sx, sy, P = _ducros_oracle(
        20, lambda X, Y: a * X - b * Y, lambda X, Y: b * X + a * Y,
        lambda X, Y: p0 + k * (Y + 0.2) ** 2)
    si, sj = _inner(grid)
    h = grid.dy
    ratio = 24.0 * k * h * h / (64.0 * P[si, sj] + 40.0 * k * h * h)
    np.testing.assert_allclose(sy[si, sj], ratio * a * a / (a * a + b * b), rtol=1e-9, atol=0)
    assert np.abs(sx[si, sj]).max() < 1e-13
    # linear pressure: the 4th-order difference vanishes
    grid, sx, sy, _ = _ducros_oracle(
        20, lambda X, Y: a * X, lambda X, Y: a * Y, lambda X, Y: 5.0 + 3.0 * X - 2.0 * Y)
    si, sj = _inner(grid)
    assert np.abs(sx[si, sj]).max() < 1e-13 and np.abs(sy[si, sj]).max() < 1e-13


def test_ducros_rim_is_one():
    grid, sx, sy, _ = _ducros_oracle(
        12, lambda X, Y: -Y, lambda X, Y: X, lambda X, Y: 1.0 + 0 * X)
    rim = np.ones_like(sx, dtype=bool)
    rim[2:-2, 2:-2] = False
    assert np.all(sx[rim] == 1.0) and np.all(sy[rim] == 1.0)


# ------------------------------------------------------------ SSP-RK3 ------
def _rk3_case(nx=8, ny=6, seed=0):
    grid = Grid(nx, ny, 0.0, 1.0, 0.0, 0.75, bc=ZG)
    cfg = CaseConfig(grid, AIR, SCHEME_M10)
    rng = np.random.default_rng(seed)
    shape = (grid.nxg, grid.nyg)
    a1 = rng.uniform(0.2, 0.45, shape)
    U = cases.conserved_from_prims(AIR, a1, rng.uniform(1.0, 2.0, shape),
                                   rng.uniform(0.5, 1.5, shape), rng.uniform(-0.5, 0.5, shape),
                                   rng.uniform(-0.5, 0.5, shape), rng.uniform(50.0, 80.0, shape))
    orc = O.Oracle(cfg)
    orc.set_state(U)
    orc.prepare()
    return grid, orc


def _shu_osher(z):
    return 1.0 + z + z * z / 2.0 + z ** 3 / 6.0


def test_ssp_rk3_stage_table_on_linear_residual():
    """Residual R(U) = lambda*U through orc_step's own stage sequence."""
    grid, orc = _rk3_case()
    lam, dt = 2.0, 0.15
    g, nx = grid.g, grid.nx
    calls = []

    def hook(stage):
        calls.append(stage)
        for m in range(6):
            u = orc.array(O.AR1 + m)
            fx = np.zeros(orc.shape)
            # fx[i] - fx[i-1] = dx * lam * U[i] over the interior x-faces
            fx[g - 1:g + nx, :] = np.concatenate(
                [np.zeros((1, orc.shape[1])), np.cumsum(grid.dx * lam * u[g:g + nx, :], axis=0)])
            for base in (O.FX0, O.FXL0):
                orc.array(base + m)[...] = fx
            for base in (O.GY0, O.GYL0):
                orc.array(base + m)[...] = 0.0
        for s in (O.SRCX, O.SRCXL, O.SRCY, O.SRCYL):
            orc.array(s)[...] = 0.0

    orc.set_flux_hook(hook)
    si, sj = grid.interior()
    U0 = orc.state()[:, si, sj]
    orc.step(dt)
    U1 = orc.state()[:, si, sj]
    orc.set_flux_hook(None)
    assert calls == [0, 1, 2]
    want = _shu_osher(-lam * dt) * U0
    np.testing.assert_allclose(U1, want, rtol=1e-13, atol=0)
    # a different stage table would give a different polynomial
    assert abs(_shu_osher(-lam * dt) - (1 - lam * dt) ** 3) > 1e-3


def test_ssp_rk3_alpha_source():
    """dalpha/dt = +alpha_prim * kappa from the contact-velocity source."""
    grid, orc = _rk3_case(seed=1)
    kappa, dt = 1.5, 0.2
    g, nx = grid.g, grid.nx

    def hook(stage):
        for m in range(6):
            for base in (O.FX0, O.FXL0, O.GY0, O.GYL0):
                orc.array(base + m)[...] = 0.0
        srcx = np.zeros(orc.shape)
        # (srcx[i] - srcx[i-1]) / dx = kappa
        srcx[g - 1:g + nx, :] = (kappa * grid.dx * np.arange(nx + 1))[:, None]
        for s in (O.SRCX, O.SRCXL):
            orc.array(s)[...] = srcx
        for s in (O.SRCY, O.SRCYL):
            orc.array(s)[...] = 0.0

    orc.set_flux_hook(hook)
    si, sj = grid.interior()
    U0 = orc.state()[:, si, sj]
    orc.step(dt)
    U1 = orc.state()[:, si, sj]
    np.testing.assert_allclose(U1[5], _shu_osher(kappa * dt) * U0[5], rtol=1e-13, atol=0)
    # zero residual: the stage combines reproduce U to round-off
    np.testing.assert_allclose(U1[:5], U0[:5], rtol=1e-15, atol=0)


# ------------------------------------------------------------ floors -------
@pytest.mark.parametrize("eos", [AIR, EOS(4.4, 6000.0, 4.4, 6000.0)])
def test_floored_face_state_lands_on_the_floor(eos):
    e = O.eos_struct(eos)
    pi = eos.pinf1
    gm1 = eos.gamma1 - 1.0
    tol = 4.0 * math.ulp(pi) / gm1 + 4.0 * math.ulp(1e-8)
    for p in (-pi - 1.0, -2.0 * pi - 5.0, float("nan"), -pi - 1e-3 * max(pi, 1.0)):
        s = L.orc_pp_floor_pressure_state(e, O.S6(1.0, 0.5, 0.3, -0.2, p, 0.4))
        assert abs((s.v4 + pi) / gm1 - 1e-8) <= tol, (p, s.v4)
        assert (s.v0, s.v1, s.v2, s.v3, s.v5) == (1.0, 0.5, 0.3, -0.2, 0.4)
    # admissible pressure: untouched; alpha is clamped to [1e-10, 1-1e-10]
    s = L.orc_pp_floor_pressure_state(e, O.S6(1.0, 0.5, 0.3, -0.2, 2.0, 1.3))
    assert s.v4 == 2.0 and s.v5 == 1.0 - 1e-10


@pytest.mark.parametrize("scheme", [SCHEME_SC, SCHEME_M10])
@pytest.mark.parametrize("eos", [AIR, EOS(4.4, 6000.0, 4.4, 6000.0)])
def test_flux_of_a_state_below_the_floor(scheme, eos):
    """A state at rest with p below the floor: the first-order flux (and the
    M10 high-order flux, whose face states are floored) carries the floored
    pressure as its normal momentum flux and nothing else."""
    grid = Grid(6, 5, 0.0, 1.0, 0.0, 1.0, bc=ZG)
    cfg = CaseConfig(grid, eos, scheme)
    one = np.ones((grid.nxg, grid.nyg))
    U = cases.conserved_from_prims(eos, 0.5 * one, one, one, 0 * one, 0 * one, 1.0 * one)
    orc = O.Oracle(cfg)
    orc.set_state(U)
    orc.prepare()
    orc.array(O.PP)[...] = -2.0 * eos.pinf1 - 1.0     # below the floor everywhere
    orc.compute_fluxes()
    F = orc.fluxes()
    orc.close()
    pi, gm1 = eos.pinf1, eos.gamma1 - 1.0
    tol = 8.0 * math.ulp(pi) / gm1 + 8.0 * math.ulp(1e-8)
    si, sj = grid.interior()
    xs = (slice(grid.g - 1, grid.g + grid.nx), sj)
    ys = (si, slice(grid.g - 1, grid.g + grid.ny))
    planes = [(14, xs, 2), (21, ys, 3)]              # first-order x / y
    if scheme.face_pp == 1:
        planes += [(0, xs, 2), (7, ys, 3)]          # M10 high-order x / y
    for base, sl, normal in planes:
        pstar = F[base + normal][sl]
        assert np.abs((pstar + pi) / gm1 - 1e-8).max() <= tol, base
        for m in range(6):
            if m != normal:
                assert np.abs(F[base + m][sl]).max() <= 1e-12 * max(pi, 1.0), (base, m)


# ------------------------------------------------------------ limiter ------
def test_limiter_lands_cells_on_the_floor_and_shares_the_min():
    """Two neighbouring cells A (i0) and B (i0+1) lose partial density 1
    through prescribed high-order faces; the first-order fluxes are zero.
    B's theta is the smaller, so the face they share takes B's theta, and B's
    forward-Euler density through its blended faces lands on 1e-10."""
    grid = Grid(8, 8, 0.0, 1.0, 0.0, 1.0, bc=ZG)
    cfg = CaseConfig(grid, AIR, SCHEME_SC)
    one = np.ones((grid.nxg, grid.nyg))
    U = cases.conserved_from_prims(AIR, 0.5 * one, 2.0 * one, 2.0 * one, 0 * one, 0 * one, 4.0 * one)
    orc = O.Oracle(cfg)
    orc.set_state(U)
    orc.prepare()
    orc.compute_fluxes()
    for m in range(28):
        orc.array(O.FX0 + m)[...] = 0.0
    dx, dy, dt = grid.dx, grid.dy, 0.01
    i0, j0 = grid.g + 3, grid.g + 4
    ar1 = orc.array(O.AR1)[i0, j0]                    # 1.0 in both cells
    qa, s, qb = 90.0 * dx, 40.0 * dx, 300.0 * dx      # fluxes (x-faces)
    fx0 = orc.array(O.FX0)
    fx0[i0 - 1, j0] = -qa      # A -> left neighbour
    fx0[i0, j0] = s            # A -> B
    fx0[i0 + 1, j0] = qb       # B -> right neighbour
    hA = ar1 - dt * (s + qa) / dx
    hB = ar1 - dt * (qb - s) / dx
    assert hA < 0 and hB < 0
    thA = (ar1 - 1e-10) / (ar1 - hA)
    thB = (ar1 - 1e-10) / (ar1 - hB)
    assert thB < thA < 1
    orc.limit_fluxes(dt)
    fx0 = orc.array(O.FX0)
    assert fx0[i0, j0] == pytest.approx(thB * s, rel=1e-14)         # shared: min
    assert fx0[i0 - 1, j0] == pytest.approx(-thA * qa, rel=1e-14)
    assert fx0[i0 + 1, j0] == pytest.approx(thB * qb, rel=1e-14)
    rB = (fx0[i0 + 1, j0] - fx0[i0, j0]) / dx
    assert abs(ar1 - dt * rB - 1e-10) <= 2e-15
    rA = (fx0[i0, j0] - fx0[i0 - 1, j0]) / dx
    assert ar1 - dt * rA > 1e-10
    # faces away from the two cells are untouched (first order = high order = 0)
    mask = np.ones(orc.shape, dtype=bool)
    mask[i0 - 1:i0 + 2, j0] = False
    assert np.all(fx0[mask] == 0.0)
    orc.close()
