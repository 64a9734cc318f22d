"""The exact path's short IEEE division (div_rn / div_by / recip_rn in
device_math.hpp), checked in exact rational arithmetic.

The device sequence is
    y0 = v_rcp_f32(d)                  (within 1 ulp of 1/d)
    y  = fma(fma(-d, y0, 1), y0, y0)   (= RN(1/d): the device premise)
    q0 = n * y
    q  = fma(fma(-d, q0, n), y, q0)
with every operation rounded once to FP32 (round to nearest, ties to even).
It claims the IEEE quotient RN(n / d) bit for bit for d > 0 normal with
d < 2^126, n = +0 or 2^-100 <= |n|, |n / d| < 2^96, given the premise that
the refined reciprocal y is the correctly rounded 1/d.  The premise is a
property of the hardware reciprocal: ausm_check_arithmetic verifies it on
the device for every FP32 divisor (tests/test_gpu_parity.py), and context
and batch creation refuse a device that fails it.  Here, with y = RN(1/d),
each operation is evaluated exactly and rounded to FP32 by hand, over
random operands spanning the claimed range and over quotients placed next
to a rounding midpoint (constructed to within ~2^-48 of it: the cases on
which the same sequence with a reciprocal one ulp off rounds the wrong
way, checked below too, so that the test can fail).  The end-to-end
bitwise parity tests of the exact path cover the operands the solver
actually divides.
"""
import math
import random
from fractions import Fraction

import pytest

P = 24  # FP32 significand bits


def rn32(x):
    """Round a rational to the nearest FP32 value (ties to even), with
    gradual underflow (the device keeps FP32 denormals)."""
    if x == 0:
        return Fraction(0)
    s = -1 if x < 0 else 1
    a = abs(x)
    e = math.floor(math.log2(a.numerator) - math.log2(a.denominator))
    # fix e so that 2^e <= a < 2^(e+1)
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    assert e < 128, e
    ulp = Fraction(2) ** (max(e, -126) - P + 1)
    m = a / ulp
    f = m.numerator // m.denominator
    rem = m - f
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and f % 2 == 1):
        f += 1
    return s * f * ulp


def ulp32(x):
    a = abs(x)
    e = math.floor(math.log2(a))
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    return Fraction(2) ** (e - P + 1)


def fma(a, b, c):
    return rn32(a * b + c)


def seq(n, d, y):
    """the device quotient from the reciprocal y of d"""
    q = rn32(n * y)
    return fma(fma(-d, q, n), y, q)


def recip_rn(d):
    """RN(1/d), what the refined hardware reciprocal returns"""
    return rn32(1 / d)


def off_by_one_ulp(d):
    """the two FP32 neighbours of RN(1/d)"""
    y = recip_rn(d)
    u = ulp32(y)
    return [y - u, y + u]


def rand32(rng, lo_exp, hi_exp):
    e = rng.randint(lo_exp, hi_exp - 1)
    m = rng.randint(2 ** (P - 1), 2 ** P - 1)
    return Fraction(m) * Fraction(2) ** (e - P + 1)


def hard_case(rng):
    """n, d with n / d within ~2^-48 (relative) of a rounding midpoint: with
    D odd, X = r0 / D mod 2^24 makes D X = N 2^24 + r0 for a small odd r0,
    so n / d = X / 2^24 - r0 / (2^24 D) sits just beside the midpoint X / 2^24
    (X odd, 25 bits) of two FP32 values."""
    while True:
        D = rng.randrange(2 ** 23 + 1, 2 ** 24, 2)
        r0 = rng.choice((1, -1, 3, -3, 5, -5))
        X = (r0 * pow(D, -1, 2 ** 24)) % 2 ** 24 + 2 ** 24
        N, rem = divmod(D * X - r0, 2 ** 24)
        if rem == 0 and 2 ** 23 <= N < 2 ** 24:
            break
    a = rng.randint(-60, 40)
    b = rng.randint(-60 - a, 40 - a)
    return Fraction(N) * Fraction(2) ** (a + b), Fraction(D) * Fraction(2) ** (b - 23)


def cases(seed, count):
    rng = random.Random(seed)
    out = []
    for _ in range(count):
        if rng.random() < 0.5:
            n, d = hard_case(rng)
        else:
            d = rand32(rng, -60, 60)
            n = rand32(rng, -100, 60)
        if rng.random() < 0.3:
            n = -n
        qe = n / d
        if qe != 0 and not (Fraction(2) ** -125 <= abs(qe) < Fraction(2) ** 96):
            continue
        out.append((n, d))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_short_division_is_the_ieee_quotient(seed):
    for n, d in cases(seed, 1500):
        assert seq(n, d, recip_rn(d)) == rn32(n / d), (n, d)


def test_zero_numerator_gives_plus_zero():
    for d in (Fraction(1), Fraction(3, 7), Fraction(2) ** -100, Fraction(2) ** 100):
        assert seq(Fraction(0), rn32(d), recip_rn(rn32(d))) == 0


def test_the_cases_need_the_correctly_rounded_reciprocal():
    """With a reciprocal one ulp off RN(1/d) some of the near-midpoint
    quotients round the wrong way after one correction, so the cases above
    can tell the premise from a plausible shortcut."""
    wrong = 0
    for n, d in cases(7, 600):
        want = rn32(n / d)
        if any(seq(n, d, y) != want for y in off_by_one_ulp(d)):
            wrong += 1
    assert wrong > 0


def seq_zs(n, d, y):
    """div_by_zs: the residual formed as d q - n and subtracted"""
    q = rn32(n * y)
    return fma(-fma(d, q, -n), y, q)


@pytest.mark.parametrize("seed", [4, 5])
def test_zero_sign_form_is_the_same_quotient(seed):
    """For nonzero numerators the residual is exact in both forms, so
    -(d q - n) = n - d q and div_by_zs is div_by bit for bit."""
    for n, d in cases(seed, 1500):
        y = recip_rn(d)
        assert seq_zs(n, d, y) == seq(n, d, y) == rn32(n / d), (n, d)


def _ieee_zero_steps(n, d, y, zs):
    """The two forms on a zero numerator in binary floating point, where
    every product and sum of the sequence is exact (so float64 gives the
    FP32 results, signs of zero included)."""
    q = n * y
    if zs:
        r = d * q + (-n)
        return (-r) * y + q
    r = (-d) * q + n
    return r * y + q


@pytest.mark.parametrize("d", [1.0, 3.0 / 7.0, 2.0 ** -100, 2.0 ** 25])
def test_zero_sign_form_keeps_the_sign_of_zero(d):
    """IEEE gives +-0 / d = +-0: div_by_zs keeps it; div_by's n - d q is +0
    for both zeros and turns -0 into +0 (why div_free uses div_by_zs)."""
    y = float(recip_rn(Fraction(d)))
    for n in (0.0, -0.0):
        out = _ieee_zero_steps(n, d, y, True)
        assert out == 0.0 and math.copysign(1.0, out) == math.copysign(1.0, n)
    assert math.copysign(1.0, _ieee_zero_steps(-0.0, d, y, False)) == 1.0
