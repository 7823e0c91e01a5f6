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
