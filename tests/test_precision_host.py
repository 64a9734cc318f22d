"""Host-side logic of the precision study (no GPU): the critical pressure
coefficient, the shock locator and the seeded CST section."""
import math

import numpy as np

from ausm.precision import _cst_y, critical_cp, cst_coefficients, shock_positions
from oracle_lib import cst_boundary


def test_critical_cp_closed_forms():
    # sonic free stream: the critical pressure is the free-stream pressure
    assert abs(critical_cp(1.0)) < 1e-15
    # isentropic relation p*/p_inf = ((1 + (g-1)/2 M^2) / ((g+1)/2))^(g/(g-1))
    g, m = 1.4, 0.73
    pr = ((1 + 0.5 * (g - 1) * m * m) / (0.5 * (g + 1))) ** (g / (g - 1))
    assert math.isclose(critical_cp(m), (pr - 1) / (0.5 * g * m * m), rel_tol=1e-14)
    assert critical_cp(0.5) < critical_cp(0.8) < 0


def test_shock_locator_on_synthetic_cp():
    M = 512
    b = cst_boundary(M, [0.15] * 6, [-0.15] * 6)
    xm = 0.5 * (b[:, 0] + np.roll(b[:, 0], -1)) + 0.5
    cs = critical_cp(0.73)
    # supersonic plateau ending in a jump at x/c = 0.62 on the upper surface
    # only; the lower surface stays subsonic
    up = np.arange(M) < M // 2
    cp = np.where(up & (xm < 0.62), cs - 0.4, cs + 0.3)
    s = shock_positions(b, cp, 0.73)
    assert s["lower"] is None
    # the crossing is interpolated between the two segment midpoints that
    # straddle the jump
    below = xm[up & (xm < 0.62)].max()
    above = xm[up & (xm >= 0.62)].min()
    assert below <= s["upper"] <= above
    # a linear Cp ramp crosses the critical value where the line does
    cp2 = np.where(up, cs + (xm - 0.4), cs + 1.0)
    s2 = shock_positions(b, cp2, 0.73)
    assert math.isclose(s2["upper"], 0.4, abs_tol=1e-12)


def test_cst_section_is_seeded_thick_and_aft_cambered():
    au, al = cst_coefficients(0)
    au2, al2 = cst_coefficients(0)
    assert np.array_equal(au, au2) and np.array_equal(al, al2)
    assert not np.array_equal(au, cst_coefficients(1)[0])
    x = np.linspace(0, 1, 200001)
    yu, yl = _cst_y(au, x), _cst_y(al, x)
    assert math.isclose((yu - yl).max(), 0.11, rel_tol=1e-6)
    camber = 0.5 * (yu + yl)
    assert x[camber.argmax()] > 0.55  # aft loading
    assert au[0] ** 2 / 2 > 0.005     # round, blunt nose
