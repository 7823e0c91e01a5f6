"""Exact Riemann solver for the 1D Euler equations with stiffened-gas EOS per side.

Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", ch. 4,
with p replaced by p + pi_K on side K (the stiffened gas behaves as an ideal
gas in p + pi).  Each side may have its own (gamma, pi), which gives the
two-material solution the five-equation model converges to.
"""
from __future__ import annotations

import math


class Side:
    def __init__(self, rho, u, p, gamma, pinf):
        self.rho, self.u, self.p, self.g, self.pi = rho, u, p, gamma, pinf
        self.c = math.sqrt(gamma * (p + pinf) / rho)


def _f(p, K: Side):
    g, pb = K.g, K.p + K.pi
    P = p + K.pi
    if P > pb:  # shock
        A = 2.0 / ((g + 1.0) * K.rho)
        B = (g - 1.0) / (g + 1.0) * pb
        return (P - pb) * math.sqrt(A / (P + B))
    return 2.0 * K.c / (g - 1.0) * ((P / pb) ** ((g - 1.0) / (2.0 * g)) - 1.0)


def star_state(L: Side, R: Side):
    fun = lambda p: _f(p, L) + _f(p, R) + (R.u - L.u)
    lo = -min(L.pi, R.pi) + 1e-12 * max(1.0, L.p, R.p)
    hi = max(L.p, R.p, 1.0)
    while fun(hi) < 0.0:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if fun(mid) > 0.0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (L.u + R.u) + 0.5 * (_f(ps, R) - _f(ps, L))
    return ps, us


def sample(L: Side, R: Side, xi):
    """(rho, u, p, side) at similarity coordinate xi = x/t; side 0 = left material."""
    ps, us = star_state(L, R)
    if xi <= us:
        K, sgn, side = L, -1.0, 0
    else:
        K, sgn, side = R, 1.0, 1
    g, pb, Ps = K.g, K.p + K.pi, ps + K.pi
    if Ps > pb:  # shock
        r = Ps / pb
        gm = (g - 1.0) / (g + 1.0)
        rho_s = K.rho * (r + gm) / (gm * r + 1.0)
        S = K.u + sgn * K.c * math.sqrt((g + 1.0) / (2.0 * g) * r + (g - 1.0) / (2.0 * g))
        if (sgn < 0 and xi <= S) or (sgn > 0 and xi >= S):
            return K.rho, K.u, K.p, side
        return rho_s, us, ps, side
    # rarefaction
    rho_s = K.rho * (Ps / pb) ** (1.0 / g)
    c_s = K.c * (Ps / pb) ** ((g - 1.0) / (2.0 * g))
    head = K.u + sgn * K.c
    tail = us + sgn * c_s
    if (sgn < 0 and xi <= head) or (sgn > 0 and xi >= head):
        return K.rho, K.u, K.p, side
    if (sgn < 0 and xi >= tail) or (sgn > 0 and xi <= tail):
        return rho_s, us, ps, side
    # inside the fan (Toro 4.56 / 4.63 with p -> p + pi)
    u = 2.0 / (g + 1.0) * (-sgn * K.c + 0.5 * (g - 1.0) * K.u + xi)
    c = 2.0 / (g + 1.0) * (K.c - sgn * 0.5 * (g - 1.0) * (K.u - xi))
    rho = K.rho * (c / K.c) ** (2.0 / (g - 1.0))
    P = pb * (c / K.c) ** (2.0 * g / (g - 1.0))
    return rho, u, P - K.pi, side
