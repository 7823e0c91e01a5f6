"""The FP64 device math divides by shared or constant denominators as
q = RN(a*y), r = RN(a - b*q) (one FMA), RN(q + r*y) with y = RN(1/b)
(`quotient` in wave-multiphase_amd/csrc/wamp_device.h).  These tests check, with
exact rational arithmetic standing in for the FMAs, that this yields exactly
the IEEE quotient a/b -- the property the FP64 parity path relies on -- for
the EOS constants g-1 used on the device and for random operands."""
import random
import struct
from fractions import Fraction

import pytest


def rn(x: Fraction) -> float:
    return float(x)  # Fraction -> float rounds to nearest


def fma_quotient(a: float, y: float, b: float) -> float:
    q = a * y                                  # RN(a y)
    r = rn(Fraction(a) - Fraction(b) * Fraction(q))   # FMA: a - b q, rounded once
    return rn(Fraction(r) * Fraction(y) + Fraction(q))  # FMA: q + r y, rounded once


def rand_double(rng, emin=-40, emax=40):
    mant = rng.getrandbits(52)
    exp = rng.randint(emin, emax) + 1023
    sign = rng.getrandbits(1)
    return struct.unpack("<d", struct.pack("<Q", (sign << 63) | (exp << 52) | mant))[0]


@pytest.mark.parametrize("gamma", [4.4, 1.4, 1.5, 1.67, 6.12, 7.15])
def test_eos_constant_denominators(gamma):
    rng = random.Random(int(gamma * 100))
    b = gamma - 1.0
    y = 1.0 / b
    for _ in range(4000):
        a = rng.random() if rng.random() < 0.5 else rand_double(rng)
        assert fma_quotient(a, y, b) == a / b, (a, b)


def test_random_operands_and_all_ones_mantissas():
    rng = random.Random(7)
    for k in range(20000):
        a, b = rand_double(rng), rand_double(rng)
        if k % 4 == 0:  # denominators with (nearly) all-ones mantissas
            bits = struct.unpack("<Q", struct.pack("<d", b))[0] | ((1 << 52) - 1) ^ rng.getrandbits(3)
            b = struct.unpack("<d", struct.pack("<Q", bits))[0]
        y = 1.0 / b
        assert fma_quotient(a, y, b) == a / b, (a, b)


def test_zero_numerator():
    for b in (0.4, 3.4, 1.0e-10, 7.0e5):
        assert fma_quotient(0.0, 1.0 / b, b) == 0.0


def fma(x: float, y: float, z: float) -> float:
    return rn(Fraction(x) * Fraction(y) + Fraction(z))


def recip_nr(b: float, y0: float) -> float:
    """Two Newton steps from the hardware estimate y0 (recip_nr in wamp_device.h)."""
    y = y0
    for _ in range(2):
        e = fma(-b, y, 1.0)
        y = fma(y, e, y)
    return y


def estimate(rng, b: float, rel: float) -> float:
    """A reciprocal estimate with relative error up to `rel`, standing in for
    v_rcp_f64, whose exact output is not known here: the property holds for
    any estimate within 2^-20."""
    return rn(Fraction(1) / Fraction(b) * (1 + Fraction(rng.uniform(-rel, rel))))


def test_newton_reciprocal_then_correction_is_the_ieee_quotient():
    """ddiv / recip_nr: whatever the estimate within 2^-20, two Newton steps and
    the correction give exactly a / b -- the arithmetic of the compiler's IEEE
    divide sequence when no operand needs scaling -- including denominators
    with all-ones mantissas and the solver's divisors (densities, c^2,
    mixture gamma - 1, tanh(0.9))."""
    rng = random.Random(11)
    specials = [1.0e-10, 2.0e-10, 1.0e-20, 0.4, 3.4, 1000.0, 0.7162978701990245, 6000.0 * 4.4]
    for k in range(20000):
        a = rand_double(rng, -300, 300)
        if k % 5 == 0:
            b = specials[k // 5 % len(specials)]
        else:
            b = rand_double(rng, -300, 300)
        if k % 7 == 0:  # (nearly) all-ones mantissas
            bits = struct.unpack("<Q", struct.pack("<d", b))[0] | ((1 << 52) - 1) ^ rng.getrandbits(3)
            b = struct.unpack("<d", struct.pack("<Q", bits))[0]
        y = recip_nr(b, estimate(rng, b, 2.0 ** -20))
        assert fma_quotient(a, y, b) == a / b, (a, b)


def test_mixed_unit_division_is_within_one_ulp():
    """The mixed unit's ddiv: one Newton step and the correction, within one ulp."""
    import math

    rng = random.Random(12)
    for _ in range(20000):
        a, b = rand_double(rng, -200, 200), rand_double(rng, -200, 200)
        y0 = estimate(rng, b, 2.0 ** -26)
        e = fma(-b, y0, 1.0)
        y = fma(y0, e, y0)
        q = a * y
        r = fma(-b, q, a)
        got = fma(r, y, q)
        want = a / b
        assert abs(got - want) <= math.ulp(want), (a, b)
