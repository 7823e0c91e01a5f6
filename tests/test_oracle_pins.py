"""Pin the CPU oracle's scalar functions against mathematics, not against its own formulas.

The reference ships no golden vectors or tests, so every oracle function is
checked against what the mathematics fixes: closed forms at special cases,
polynomial exactness, symmetries, bounds and algebraic identities between
independently written functions (e.g. the left and right eigenvector
matrices).  Each check is chosen so that a plausible transcription mistake
(a dropped term, a wrong coefficient, sign or index) fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O

L = O.lib()
WATER = O.OrcEos(4.4, 6000.0, 1.4, 0.0)
DIM = O.OrcEos(6.12, 3.43e8, 1.4, 0.0)


def gpi(eos, a1):
    g, pi = C.c_double(), C.c_double()
    L.orc_mix_gamma_pinf(C.byref(eos), a1, C.byref(g), C.byref(pi))
    return g.value, pi.value


def gbpb(eos, a1):
    gb, pb = C.c_double(), C.c_double()
    L.orc_mix_gamma_pibig(C.byref(eos), a1, C.byref(gb), C.byref(pb))
    return gb.value, pb.value


# ---------------------------------------------------------------- EOS ------
@pytest.mark.parametrize("eos", [WATER, DIM])
def test_mixture_reduces_to_pure_phases(eos):
    g, pi = gpi(eos, 1.0)
    assert g == pytest.approx(eos.g1, rel=1e-14) and pi == pytest.approx(eos.p1, rel=1e-14)
    g, pi = gpi(eos, 0.0)
    assert g == pytest.approx(eos.g2, rel=1e-14) and pi == pytest.approx(eos.p2, abs=1e-9)


def test_mixture_energy_is_volume_fraction_weighted():
    # Stiffened gas: rho*e_k = (p + g_k pi_k)/(g_k - 1).  The five-equation
    # mixture energy at pressure equilibrium is the alpha-weighted sum.
    for a1 in (0.0, 0.2, 0.5, 0.93, 1.0):
        for p in (0.3, 1.0, 7.0):
            gb, pb = gbpb(WATER, a1)
            e_mix = gb * p + pb
            e1 = (p + 4.4 * 6000.0) / 3.4
            e2 = (p + 1.4 * 0.0) / 0.4
            assert e_mix == pytest.approx(a1 * e1 + (1 - a1) * e2, rel=1e-13)


def test_sound_speed_closed_form_pure_water():
    c = L.orc_sound_speed(C.byref(WATER), 1000.0, 1.0, 1.0)
    assert c == pytest.approx(math.sqrt(4.4 * (1.0 + 6000.0) / 1000.0), rel=1e-14)
    # floor at 1e-20 for non-physical input (SC:101-102)
    assert L.orc_sound_speed(C.byref(WATER), 1.0, -1e9, 0.5) == pytest.approx(1e-10)


def test_energy_pressure_round_trip():
    """prim_to_energy then cons_to_prim (via the oracle's update_prims) recover p."""
    from wamp.config import CaseConfig, Grid, SCHEME_SC, EOS

    cfg = CaseConfig(Grid(8, 8), EOS(4.4, 6000.0, 1.4, 0.0), SCHEME_SC)
    orc = O.Oracle(cfg)
    rng = np.random.default_rng(0)
    shape = orc.shape
    a1 = rng.uniform(0.01, 0.99, shape)
    rho1, rho2 = 1000.0, rng.uniform(0.5, 2.0, shape)
    u, v, p = rng.normal(size=shape), rng.normal(size=shape), rng.uniform(0.5, 5.0, shape)
    E = np.vectorize(lambda *a: L.orc_prim_to_energy(C.byref(WATER), *a))(
        a1 * rho1, (1 - a1) * rho2, u, v, p, a1)
    U = np.stack([a1 * rho1, (1 - a1) * rho2, (a1 * rho1 + (1 - a1) * rho2) * u,
                  (a1 * rho1 + (1 - a1) * rho2) * v, E, a1])
    orc.set_state(U)
    orc.update_prims()
    P = orc.prims()
    # cancellation against pb ~ 7765*a1 bounds the error by a few ulp of pb
    assert np.abs(P[2] - p).max() < 1e-11
    assert np.abs(P[0] - u).max() < 1e-13 and np.abs(P[1] - v).max() < 1e-13


def test_clamp01_pp():
    assert L.orc_clamp01_pp(float("nan")) == 1e-10
    assert L.orc_clamp01_pp(-3.0) == 1e-10
    assert L.orc_clamp01_pp(2.0) == 1.0 - 1e-10
    assert L.orc_clamp01_pp(0.3) == 0.3


# ------------------------------------------------------ projections ---------
def _rand_face(rng):
    ar1, ar2 = rng.uniform(0.1, 1000.0), rng.uniform(0.1, 3.0)
    return dict(c=rng.uniform(0.5, 6.0), rho=ar1 + ar2, ar1=ar1, ar2=ar2,
                u_avg=rng.normal(), v_avg=rng.normal())


@pytest.mark.parametrize("d", ["x", "y"])
def test_left_right_eigenvectors_are_inverse(d):
    """L.R = I (SC:462-478): to_phys(to_char(u)) = u and to_char(to_phys(w)) = w."""
    rng = np.random.default_rng(1)
    to_char = getattr(L, f"orc_to_char_{d}")
    to_phys = getattr(L, f"orc_to_phys_{d}")
    for _ in range(200):
        f = _rand_face(rng)
        args = (f["c"], f["rho"], f["ar1"], f["ar2"], f["u_avg"], f["v_avg"])
        u = rng.normal(size=6) * np.array([500, 1, 10, 10, 100, 0.3])
        w = to_char(O.dvec(u), *args)
        back = np.array(to_phys(w, *args).tolist())
        assert np.allclose(back, u, rtol=1e-12, atol=1e-12 * np.abs(u).max())
        w0 = rng.normal(size=6)
        s = to_phys(O.S6(*w0), *args)
        w1 = np.array(to_char(O.dvec(s.tolist()), *args).tolist())
        assert np.allclose(w1, w0, rtol=1e-11, atol=1e-11)


def test_right_eigenvectors_are_eigenvectors_of_the_flux_jacobian():
    """R's columns are eigenvectors of the x quasi-linear system in SC variables.

    For V = [m1, m2, ru, rv, p, a] the 1D x-system dV/dt + A dV/dx = 0 at
    (rho, u, v, c, Y1, Y2) has A with eigenvalues u-c, u (x4), u+c.  The
    columns of R (SC:472-478) must satisfy A r = lambda r.  A is built here
    from the conservation laws (finite differences of the physical flux in
    the SC variables), independently of the oracle's projection.
    """
    rng = np.random.default_rng(2)
    for _ in range(20):
        m1, m2 = rng.uniform(0.5, 900.0), rng.uniform(0.5, 2.0)
        rho = m1 + m2
        u, v, p = rng.normal(), rng.normal(), rng.uniform(0.5, 3.0)
        c = rng.uniform(0.5, 4.0)
        # A for V = [m1, m2, ru, rv, p, a]: m_k, ru, rv conservative with
        # fluxes m_k u, ru u + p, rv u;  p_t + u p_x + rho c^2 u_x = 0;  a_t + u a_x = 0
        A = np.zeros((6, 6))
        # d(u)/dV: u = ru/rho
        du = np.array([-u / rho, -u / rho, 1 / rho, 0, 0, 0])
        for k in range(2):  # m_k u
            A[k] = (np.eye(6)[k] * u) + (m1 if k == 0 else m2) * du
        A[2] = u * np.eye(6)[2] + rho * u * du + np.eye(6)[4]  # ru u + p
        A[3] = u * np.eye(6)[3] + rho * v * du                 # rv u
        A[4] = u * np.eye(6)[4] + rho * c * c * du             # non-conservative p
        A[5] = u * np.eye(6)[5]
        R = np.zeros((6, 6))
        for k in range(6):
            w = np.zeros(6)
            w[k] = 1.0
            R[:, k] = L.orc_to_phys_x(O.S6(*w), c, rho, m1, m2, u, v).tolist()
        lam = np.array([u - c, u, u, u, u, u + c])
        assert np.allclose(A @ R, R * lam, rtol=1e-10, atol=1e-10 * np.abs(A).max())


def test_y_projection_is_rotated_x_projection():
    """to_char_y(u; ua, va) = to_char_x(u rotated by +90 deg; va, -ua)."""
    rng = np.random.default_rng(3)
    for _ in range(100):
        f = _rand_face(rng)
        u = rng.normal(size=6)
        rot = u.copy()
        rot[2], rot[3] = u[3], -u[2]
        wy = L.orc_to_char_y(O.dvec(u), f["c"], f["rho"], f["ar1"], f["ar2"], f["u_avg"], f["v_avg"])
        wx = L.orc_to_char_x(O.dvec(rot), f["c"], f["rho"], f["ar1"], f["ar2"], f["v_avg"], -f["u_avg"])
        assert np.allclose(wy.tolist(), wx.tolist(), rtol=1e-14, atol=1e-14)


# --------------------------------------------------- reconstruction ---------
def cell_averages(poly, x0, n, h=1.0):
    """Exact cell averages of polynomial `poly` (np.poly1d) on cells [x0+kh, x0+(k+1)h]."""
    P = poly.integ()
    return [(P(x0 + (k + 1) * h) - P(x0 + k * h)) / h for k in range(n)]


def test_muscl_linear_data_and_extrema():
    for b in (0.7, -2.0):
        v = [3.0 + b * k for k in range(4)]
        assert L.orc_muscl_char_ul(*v) == pytest.approx(v[1] + b / 2, rel=1e-14)
        assert L.orc_muscl_char_ur(*v) == pytest.approx(v[2] - b / 2, rel=1e-14)
    # local maximum at v0: first order
    assert L.orc_muscl_char_ul(1.0, 2.0, 1.5, 0.0) == 2.0
    # TVD: never outside the neighbour range
    rng = np.random.default_rng(4)
    for _ in range(500):
        v = rng.normal(size=4)
        ul = L.orc_muscl_char_ul(*v)
        assert min(v[1], v[2]) - 1e-14 <= ul <= max(v[1], v[2]) + 1e-14


def test_central4_central6_polynomial_exactness():
    for deg in range(4):
        p = np.poly1d(np.random.default_rng(deg).normal(size=deg + 1))
        avg = cell_averages(p, -2.0, 4)  # cells -2..1, face at x=0
        assert L.orc_central4(*avg) == pytest.approx(p(0.0), rel=1e-12, abs=1e-12)
    for deg in range(6):
        p = np.poly1d(np.random.default_rng(10 + deg).normal(size=deg + 1))
        avg = cell_averages(p, -3.0, 6)  # cells -3..2
        assert L.orc_central6_avg(*avg) == pytest.approx(p(0.0), rel=1e-12, abs=1e-12)
    # degree 4 breaks central4 (it is 4th order, not 5th)
    p = np.poly1d([1.0, 0, 0, 0, 0])
    assert abs(L.orc_central4(*cell_averages(p, -2.0, 4)) - p(0.0)) > 1e-3


def test_mp5_is_fifth_order_on_smooth_monotone_data():
    # monotone quartic on the stencil: the MP limiter is inactive, the result
    # is the 5th-order interface value, exact for cell averages of degree <= 4
    p = np.poly1d([0.01, 0.02, 0.1, 1.0, 0.5])
    avg = cell_averages(p, -3.0, 6, h=1.0)  # cells -3..2 -> face between cells -1 and 0 at x=0
    ul = L.orc_mp5_ul(*avg[0:5])
    ur = L.orc_mp5_ur(*avg[1:6])
    assert ul == pytest.approx(p(0.0), rel=1e-12)
    assert ur == pytest.approx(p(0.0), rel=1e-12)


def test_mp5_mirror_symmetry_and_bounds():
    rng = np.random.default_rng(5)
    for _ in range(500):
        v = rng.normal(size=6) * rng.choice([1e-3, 1.0, 100.0])
        ur = L.orc_mp5_ur(*v[1:6])
        ul_mirror = L.orc_mp5_ul(*v[::-1][0:5])
        assert ur == pytest.approx(ul_mirror, rel=1e-13, abs=1e-13 * np.abs(v).max())
    # step data: bounded by the neighbours (no new extremum)
    for step in ([0, 0, 0, 1, 1, 1], [1, 1, 1, 0, 0, 0], [0, 0, 0.5, 1, 1, 1]):
        ul = L.orc_mp5_ul(*map(float, step[0:5]))
        assert -1e-14 <= ul <= 1.0 + 1e-14


def test_thinc_bounds_symmetry_and_centre_value():
    t1 = math.tanh(0.9)
    # centred step: face value qa + qd*tanh(beta/2)
    assert L.orc_thinc_ul(0.0, 0.5, 1.0) == pytest.approx(0.5 + 0.5 * t1, rel=1e-15)
    assert L.orc_thinc_ur(0.0, 0.5, 1.0) == pytest.approx(0.5 - 0.5 * t1, rel=1e-15)
    rng = np.random.default_rng(6)
    for _ in range(500):
        a, b = sorted(rng.normal(size=2))
        v0 = rng.uniform(a, b)
        ul = L.orc_thinc_ul(a, v0, b)
        assert a - 1e-14 <= ul <= b + 1e-14
        # thinc_ur(v0, v1, v2) == thinc_ul(v2, v1, v0)
        assert L.orc_thinc_ur(a, v0, b) == pytest.approx(L.orc_thinc_ul(b, v0, a), rel=1e-14, abs=1e-15)
    # non-monotone: identity
    assert L.orc_thinc_ul(1.0, 2.0, 1.0) == 2.0
    assert L.orc_thinc_ur(1.0, 2.0, 1.0) == 2.0


def test_entropy_indicator():
    # smooth (linear) entropy -> 1; constant -> 1; a jump -> below the 0.35 gate
    assert L.orc_cator2_val(-2.0, -1.0, 0.0, 1.0, 2.0) == pytest.approx(1.0, rel=1e-14)
    assert L.orc_cator2_val(5.0, 5.0, 5.0, 5.0, 5.0) == pytest.approx(1.0, rel=1e-14)
    assert L.orc_cator2_val(0.0, 0.0, 0.0, 1.0, 1.0) < 0.35
    # entropy p / rho^gamma for a pure phase
    s = L.orc_cell_entropy(C.byref(WATER), 0.0, 2.0, 3.0, 0.0)
    assert s == pytest.approx(3.0 / 2.0 ** 1.4, rel=1e-14)


# ------------------------------------------------------------- HLLC --------
def physical_flux_x(eos, P):
    ar1, ar2, u, v, p, a1 = P
    rho = ar1 + ar2
    E = L.orc_prim_to_energy(C.byref(eos), ar1, ar2, u, v, p, a1)
    return np.array([ar1 * u, ar2 * u, rho * u * u + p, rho * u * v, (E + p) * u, a1 * u])


def test_hllc_consistency_and_upwinding():
    rng = np.random.default_rng(7)
    for _ in range(200):
        a1 = rng.uniform(0.01, 0.99)
        P = [a1 * 1000.0, (1 - a1) * 1.2, rng.normal(), rng.normal(), rng.uniform(0.5, 5.0), a1]
        f = np.array(L.orc_hllc_x(C.byref(WATER), O.dvec(P), O.dvec(P)).tolist())
        ref = physical_flux_x(WATER, P)
        assert np.allclose(f, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
        us = L.orc_hllc_x_src(C.byref(WATER), O.dvec([P[0], P[1], P[2], P[4], P[5]]),
                              O.dvec([P[0], P[1], P[2], P[4], P[5]]))
        assert us == pytest.approx(P[2], rel=1e-10, abs=1e-12)
    # supersonic to the right: exactly the left flux
    PL = [1e-6, 1.0, 10.0, 0.3, 1.0, 1e-6]
    PR = [1e-6, 0.5, 9.0, -0.2, 0.4, 1e-6]
    f = np.array(L.orc_hllc_x(C.byref(WATER), O.dvec(PL), O.dvec(PR)).tolist())
    assert np.array_equal(f, physical_flux_x(WATER, PL))
    # y flux of a state equals the x flux of the rotated state
    Py = [2.0, 0.5, 0.3, -1.1, 1.7, 0.4]
    Pr = [2.0, 0.5, -1.1, -0.3, 1.7, 0.4]
    fy = np.array(L.orc_hllc_y(C.byref(WATER), O.dvec(Py), O.dvec(Py)).tolist())
    fx = np.array(L.orc_hllc_x(C.byref(WATER), O.dvec(Pr), O.dvec(Pr)).tolist())
    assert np.allclose(fy[[0, 1, 3, 4, 5]], fx[[0, 1, 2, 4, 5]], rtol=1e-12)
    assert fy[2] == pytest.approx(-fx[3], rel=1e-12)


def test_hllc_stationary_contact():
    # water | air at rest, equal pressure: no mass flux, momentum flux = p, u* = 0
    PL = [1000.0 * (1 - 1e-6), 1e-6, 0.0, 0.0, 1.0, 1 - 1e-6]
    PR = [1e-6 * 1000.0, 1.0, 0.0, 0.0, 1.0, 1e-6]
    f = np.array(L.orc_hllc_x(C.byref(WATER), O.dvec(PL), O.dvec(PR)).tolist())
    assert f[0] == 0.0 and f[1] == 0.0 and f[5] == 0.0 and f[4] == 0.0
    assert f[2] == pytest.approx(1.0, rel=1e-14)
    us = L.orc_hllc_x_src(C.byref(WATER), O.dvec([PL[0], PL[1], 0.0, 1.0, PL[5]]),
                          O.dvec([PR[0], PR[1], 0.0, 1.0, PR[5]]))
    assert us == 0.0


# ----------------------------------------------------------- Wong ----------
def test_theta_lower_lands_on_the_floor():
    for base, high, floor in ((1.0, -1.0, 1e-10), (5.0, -20.0, 1e-8), (0.3, 1e-12, 1e-10)):
        th = L.orc_pp_theta_lower(base, high, floor)
        assert 0.0 <= th < 1.0
        assert base + th * (high - base) == pytest.approx(floor, abs=1e-15 * abs(base))
    assert L.orc_pp_theta_lower(1.0, 0.5, 1e-10) == 1.0            # admissible: no limiting
    assert L.orc_pp_theta_lower(float("nan"), 0.5, 1e-10) == 0.0   # NaN: first order
    assert L.orc_pp_theta_lower(1e-10, -1.0, 1e-10) == 0.0         # base at the floor


def test_theta_cons_stage_selects_its_variables():
    l = O.dvec([1.0, 1.0, 0.0, 0.0, 10.0, 0.5])
    h = O.dvec([-1.0, 1.0, 0.0, 0.0, 10.0, 0.5])   # negative partial density only
    assert L.orc_theta_cons_stage(C.byref(WATER), l, h, 0) < 1.0
    assert L.orc_theta_cons_stage(C.byref(WATER), l, h, 1) == 1.0
    h2 = O.dvec([1.0, 1.0, 0.0, 0.0, 10.0, 1.5])   # alpha above 1 only
    assert L.orc_theta_cons_stage(C.byref(WATER), l, h2, 0) == 1.0
    assert L.orc_theta_cons_stage(C.byref(WATER), l, h2, 1) < 1.0
    # stage 2 limits on rho*c~^2 = E - kin - pi(alpha)
    h3 = O.dvec([1.0, 1.0, 0.0, 0.0, -1e4, 0.5])
    th = L.orc_theta_cons_stage(C.byref(WATER), l, h3, 2)
    assert 0.0 <= th < 1.0


def test_face_limiter_keeps_admissible_states_and_repairs_bad_ones():
    base = O.S6(2.0, 0.5, 1.0, 0.0, 1.0, 0.4)
    good = O.S6(2.1, 0.6, 1.1, 0.1, 1.1, 0.45)
    s = L.orc_pp_limit_reconstructed_state(C.byref(WATER), good, base)
    assert np.allclose(s.tolist(), good.tolist(), rtol=1e-14)
    bad = O.S6(-0.5, 0.6, 1.1, 0.1, 1.1, 1.3)
    s = L.orc_pp_limit_reconstructed_state(C.byref(WATER), bad, base)
    assert s.v0 >= 1e-11 and 1e-11 <= s.v5 <= 1 - 1e-11
