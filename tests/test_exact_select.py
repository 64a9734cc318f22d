"""The exact path's select in place of the reference's mix(a, b, t) for the
step-weighted split functions (wsel in device_math.hpp, FP32 only).

The reference forms M4+-, P5+- (AUSM+-up, ausm.js:302-325) and SLAU's f_p+-
(ausm.js:361-373) as mix(subsonic, supersonic, step(1, |M|)) =
a (1 - t) + b t with t in {0, 1}.  For finite a and b that equals the select
t ? b : a except for the sign of a zero: a + (+0) = +0 when a = -0 (t = 0),
(+0) + b = +0 when b = -0 (t = 1).  At these sites the kept operand is never
-0, which is what this test checks by brute force in FP32 (numpy float32:
one IEEE rounding per operation, no contraction, as the exact path), bit for
bit including the sign of zero, on Mach numbers drawn uniformly and from the
neighbourhoods where a factor can vanish (M = +-1, 0, and sqrt(2) - 1, where
M2+ = (M + 1)^2 / 4 is closest to 1/2), with f_a both formed from the Mach
numbers as the flux forms it and drawn freely.  A mutation check (a site
whose kept operand can be -0) shows that the comparison sees the sign.
"""
import numpy as np
import pytest

f = np.float32


def wmix(a, b, t):
    return (a * (f(1) - t)) + (b * t)


def bits(x):
    return np.asarray(x, dtype=f).view(np.uint32)


def sites(ML, MR, fa):
    """(a, b, t) of the six step-weighted mixes, as the exact flux forms them."""
    Mp1L, Mp1R = ML + f(1), MR + f(1)
    Mm1L, Mm1R = ML - f(1), MR - f(1)
    M2pL, M2pR = (f(0.25) * Mp1L) * Mp1L, (f(0.25) * Mp1R) * Mp1R
    M2mL, M2mR = (-f(0.25) * Mm1L) * Mm1L, (-f(0.25) * Mm1R) * Mm1R
    supL = (np.abs(ML) >= f(1)).astype(f)
    supR = (np.abs(MR) >= f(1)).astype(f)
    alpha4 = f(0.75) * (-f(4) + (f(5) * fa) * fa)
    sgnL = np.sign(ML).astype(f)
    sgnR = np.sign(MR).astype(f)
    sqL = (f(0.5) * Mp1L) * Mp1L
    sqR = (f(0.5) * Mm1R) * Mm1R
    return [
        (M2pL * (f(1) + f(2) * M2mL), np.fmax(f(0), ML), supL),                     # M4+
        (M2mR * (f(1) - f(2) * M2pR), np.fmin(f(0), MR), supR),                     # M4-
        (M2pL * ((f(2) - ML) + (alpha4 * ML) * M2mL), (ML >= f(0)).astype(f), supL),  # P5+
        (M2mR * ((-f(2) - MR) - (alpha4 * MR) * M2pR), (-MR >= f(0)).astype(f), supR),  # P5-
        (sqL * (f(2) - f(1) * ML), f(1) + f(1) * sgnL, supL),                        # SLAU f_p+
        (sqR * (f(2) - -f(1) * MR), f(1) + -f(1) * sgnR, supR),                      # SLAU f_p-
    ]


def mach_samples(rng, n, kind):
    if kind == "uniform":
        return rng.uniform(-3, 3, n).astype(f)
    base = np.array([1, -1, 0, np.sqrt(2) - 1, 1 - np.sqrt(2)], dtype=f)
    k = rng.integers(-4096, 4097, n).astype(np.int32)
    m = (base[rng.integers(0, base.size, n)].view(np.int32) + k).view(f)
    m = m[np.isfinite(m)]
    return m


@pytest.mark.parametrize("kind", ["uniform", "near"])
@pytest.mark.parametrize("free_fa", [False, True])
def test_select_equals_mix_bitwise(kind, free_fa):
    rng = np.random.default_rng(7 + 2 * free_fa + (kind == "near"))
    n = 1_000_000
    with np.errstate(all="ignore"):
        ML = mach_samples(rng, n, kind)
        MR = mach_samples(rng, ML.size, kind)[: ML.size]
        ML = ML[: MR.size]
        if free_fa:
            fa = rng.uniform(0, 1, ML.size).astype(f)
        else:
            Mb = f(0.5) * (ML * ML + MR * MR)
            Minf = (rng.uniform(0, 1, ML.size) ** 4).astype(f)
            Mo2 = np.minimum(f(1), np.maximum(Mb, Minf))
            fa = Mo2 * (f(2) - Mo2)
        for k, (a, b, t) in enumerate(sites(ML, MR, fa)):
            sel = np.where(t != f(0), b, a).astype(f)
            assert np.array_equal(bits(wmix(a, b, t)), bits(sel)), k


def test_the_comparison_sees_the_sign_of_zero():
    """A site whose subsonic operand can be -0 (here -M2- (1 + 2 M2+) at
    M2+ = -1/2 forced) is told apart from the mix."""
    a = np.array([-0.0, 0.25], dtype=f)
    b = np.array([0.0, 0.0], dtype=f)
    t = np.array([0.0, 0.0], dtype=f)
    sel = np.where(t != f(0), b, a).astype(f)
    assert not np.array_equal(bits(wmix(a, b, t)), bits(sel))
