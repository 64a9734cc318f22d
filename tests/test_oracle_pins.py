"""Pins for the CPU oracle (ref32 / ref64).

The reference ships no tests, golden vectors or result fixtures for this
path (its only data files are airfoil coordinates), so every pin below comes
from what the mathematics fixes: flux consistency, supersonic closed forms,
free-stream preservation, discrete conservation, exact 90-degree rotation
equivariance, mirror symmetry, the RK3 amplification polynomial, the CFL
spectral radius on a Cartesian patch, the annulus closed form of the grid
solver, textbook NACA ordinates and ref32/ref64 agreement.  Each is chosen
so that a plausible restatement mistake (dropped term, wrong sign, wrong
index, transposed operand) fails it.
"""
import math

import numpy as np
import pytest

from oracle_lib import (AUSMUP, MINMOD, NONE, SLAU, SLAU2, VANALBADA, VANLEER, P_GAMMA, P_INP,
                        P_IS0, Oracle, Solver, circle_boundary, cst_boundary, make_params,
                        naca4_boundary, naca4_digits)

SCHEMES = [SLAU2, SLAU, AUSMUP]
LIMITERS = [VANALBADA, MINMOD, VANLEER, NONE]


def phys_flux(Q, n, gamma):
    """Exact Euler flux through a face of unit normal n (FP64)."""
    rho, mx, my, E = [float(x) for x in Q]
    u, v = mx / rho, my / rho
    p = (E - 0.5 * rho * (u * u + v * v)) * (gamma - 1)
    un = u * n[0] + v * n[1]
    H = (E + p) / rho
    return np.array([rho * un, rho * un * u + p * n[0], rho * un * v + p * n[1], rho * un * H])


def cons(rho, u, v, p, gamma=float(np.float32(1.4))):
    return np.array([rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u * u + v * v)])


def rand_state(rng, mach_max=3.0):
    rho = rng.uniform(0.3, 3.0)
    p = rng.uniform(0.2, 3.0)
    a = math.sqrt(1.4 * p / rho)
    ang = rng.uniform(0, 2 * math.pi)
    sp = rng.uniform(0, mach_max) * a
    return cons(rho, sp * math.cos(ang), sp * math.sin(ang), p)


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("limiter", LIMITERS)
def test_flux_consistency(scheme, limiter):
    """F(Q,Q,Q,Q) equals the physical flux: f+ + f- = 1, M4+ + M4- = M and
    P5+ + P5- = 1 make every scheme consistent, and uniform stencils leave
    the MUSCL reconstruction unchanged."""
    rng = np.random.default_rng(1234 + scheme * 10 + limiter)
    o = Oracle(64)
    par = make_params(0.7, 3.0, dtype=np.float64)
    for _ in range(200):
        Q = rand_state(rng)
        th = rng.uniform(0, 2 * math.pi)
        n = (math.cos(th), math.sin(th))
        F = o.face_flux(par, scheme, limiter, Q, Q, Q, Q, n)
        ref = phys_flux(Q, n, float(par[P_GAMMA]))
        scale = np.abs(ref).max() + 1.0
        assert np.abs(F - ref).max() / scale < 1e-12, (F, ref)


def test_flux_consistency_fp32():
    rng = np.random.default_rng(7)
    o = Oracle(32)
    par = make_params(0.7, 3.0)
    for scheme in SCHEMES:
        for _ in range(100):
            Q = rand_state(rng).astype(np.float32)
            n = (0.6, 0.8)
            F = o.face_flux(par, scheme, VANALBADA, Q, Q, Q, Q, n)
            ref = phys_flux(Q.astype(np.float64), n, float(np.float32(1.4)))
            assert np.abs(F - ref).max() / (np.abs(ref).max() + 1) < 2e-6


def _supersonic_pair(rng, sign=+1):
    """Left/right states both moving through the face faster than sound."""
    n = (1.0, 0.0)
    while True:
        QL = cons(rng.uniform(0.5, 2), sign * rng.uniform(2.5, 4.0), rng.uniform(-0.3, 0.3),
                  rng.uniform(0.3, 1.2))
        QR = cons(rng.uniform(0.5, 2), sign * rng.uniform(2.5, 4.0), rng.uniform(-0.3, 0.3),
                  rng.uniform(0.3, 1.2))
        return QL, QR, n


def test_ausmup_supersonic_is_upwind():
    """AUSM+-up with both interface Mach numbers beyond 1: M4- = P5- = 0 and
    (with K_p = 0) the flux is exactly the physical flux of the upwind side."""
    rng = np.random.default_rng(11)
    o = Oracle(64)
    par = make_params(2.0, 0.0, dtype=np.float64)
    for sign in (+1, -1):
        for _ in range(100):
            QL, QR, n = _supersonic_pair(rng, sign)
            F = o.face_flux(par, AUSMUP, NONE, QL, QL, QR, QR, n)
            up = QL if sign > 0 else QR
            ref = phys_flux(up, n, float(par[P_GAMMA]))
            assert np.allclose(F, ref, rtol=1e-12, atol=1e-12), (sign, F, ref)


@pytest.mark.parametrize("scheme", [SLAU, SLAU2])
def test_slau_supersonic_closed_form(scheme):
    """SLAU / SLAU2 with M_L, M_R >= 1: f+ = 1, f- = 0 so p_half = p_L; g = 0
    and chi = 0 so the mass flux is (rho_L u_L + rho_R u_R) rho_L/(rho_L+rho_R)."""
    rng = np.random.default_rng(13)
    o = Oracle(64)
    par = make_params(2.0, 0.0, dtype=np.float64)
    g = float(par[P_GAMMA])
    for _ in range(100):
        QL, QR, n = _supersonic_pair(rng, +1)
        F = o.face_flux(par, scheme, NONE, QL, QL, QR, QR, n)
        rL, uL, vL = QL[0], QL[1] / QL[0], QL[2] / QL[0]
        rR, uR = QR[0], QR[1] / QR[0]
        pL = (QL[3] - 0.5 * rL * (uL * uL + vL * vL)) * (g - 1)
        HL = (QL[3] + pL) / rL
        mdot = (rL * uL + rR * uR) * rL / (rL + rR)
        ref = np.array([mdot, mdot * uL + pL, mdot * vL, mdot * HL])
        assert np.allclose(F, ref, rtol=1e-12, atol=1e-12), (F, ref)


def test_muscl_limiters_reconstruction():
    """On linear data every limiter reproduces the exact face value (phi(1)=1
    for van Albada and van Leer, minmod of equal slopes); on a local extremum
    van Leer and minmod fall back to first order."""
    o = Oracle(64)
    par = make_params(0.5, 0.0, dtype=np.float64)
    n = (1.0, 0.0)
    # linear density ramp, same pressure/velocity: L/R face states both equal
    # the midpoint, so the flux equals the physical flux of the midpoint state
    for lim in (VANALBADA, VANLEER, MINMOD):
        rhos = [1.0, 1.1, 1.2, 1.3]
        Qs = [cons(r, 0.3, 0.0, 0.8) for r in rhos]
        F = o.face_flux(par, SLAU2, lim, *Qs, n)
        mid = cons(1.15, 0.3, 0.0, 0.8)
        assert np.allclose(F, phys_flux(mid, n, float(par[P_GAMMA])), rtol=1e-9, atol=1e-9), lim
    # extremum: rho 1.0 1.2 1.2 1.0 -> dL>0, dF=0, dR<0
    Qs = [cons(r, 0.3, 0.0, 0.8) for r in (1.0, 1.2, 1.2, 1.0)]
    for lim in (VANLEER, MINMOD):
        F = o.face_flux(par, SLAU2, lim, *Qs, n)
        F0 = o.face_flux(par, SLAU2, NONE, *Qs, n)
        assert np.allclose(F, F0, rtol=1e-12), lim


def _naca_grid(o, M, N, code=12, iters=300, radius=10.0, spacing=1, size=0.5, aoa_free=True):
    m, p, t = naca4_digits(code)
    body = naca4_boundary(M, m, p, t, size=size, x_offset=0.0, expo=1.4, spacing=spacing)
    outer = circle_boundary(M, radius)
    return o.grid(body, outer, N, iters)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_freestream_preservation(scheme):
    """Uniform free stream is an exact discrete steady state away from the
    wall: the quads close (sum of l*n = 0) and the far-field ghosts reproduce
    the free stream.  The reflected-velocity wall ghost and the far-field
    Riemann ghost (exact only up to the FP32 rounding of the free-stream
    energy) reach two cells further per RK stage, so rows 4 <= j < N-4 must
    be untouched after one step."""
    for prec, tol in ((64, 1e-12), (32, 2e-6)):
        o = Oracle(prec)
        v = _naca_grid(o, 64, 32, iters=100)
        par = make_params(0.8, 1.25, dtype=o.dt)
        s = Solver(prec, par, scheme, VANALBADA, v)
        s.step(1)
        fs = par[P_IS0:P_IS0 + 4].astype(np.float64)
        dev = np.abs(s.state[5:29].astype(np.float64) - fs).max() / np.abs(fs).max()
        assert dev < tol, (prec, dev)


def test_discrete_conservation():
    """Sum over cells of area*dQ/dt equals the net boundary flux: interior
    faces cancel pairwise, so a wrong residual index or sign shows here."""
    o = Oracle(64)
    M, N = 48, 24
    v = _naca_grid(o, M, N, code=2412, iters=100)
    par = make_params(0.6, 4.0, dtype=np.float64)
    rng = np.random.default_rng(3)
    st = o.init_state(par, M, N)
    st[1:N + 1] *= 1.0 + 0.05 * rng.standard_normal((N, M, 4))
    area, flen = o.finalize(v)
    res = o.residual(par, SLAU2, VANALBADA, v, st, area, flen)
    total = (res * area[:, :, None]).sum(axis=(0, 1))
    sb = o.boundary(par, v, st)
    fy = o.flux(par, SLAU2, VANALBADA, True, v, sb)
    # wall faces are the down faces of row 0, far-field faces the up faces of row N-1
    inflow = (fy[0] * flen[0, :, 3:4]).sum(axis=0)
    outflow = (fy[N] * flen[N - 1, :, 2:3]).sum(axis=0)
    expect = inflow - outflow
    assert np.allclose(total, expect, rtol=1e-10, atol=1e-10 * np.abs(expect).max()), (total,
                                                                                       expect)


def test_rotation_equivariance_exact():
    """Rotating grid, free stream and momentum by 90 degrees ((x,y)->(-y,x))
    is exact in IEEE arithmetic for every operation the scheme uses, so the
    residual rotates bit-for-bit in FP32."""
    o = Oracle(32)
    M, N = 40, 20
    v = _naca_grid(o, M, N, code=4412, iters=60)
    vr = np.stack([-v[..., 1], v[..., 0]], axis=-1).astype(np.float32)
    par = make_params(0.7, 5.0)
    parr = par.copy()
    parr[3], parr[4] = -par[4], par[3]
    parr[P_IS0 + 1], parr[P_IS0 + 2] = -par[P_IS0 + 2], par[P_IS0 + 1]
    rng = np.random.default_rng(5)
    st = o.init_state(par, M, N)
    st[1:N + 1] *= (1.0 + 0.03 * rng.standard_normal((N, M, 4))).astype(np.float32)
    str_ = st.copy()
    str_[..., 1], str_[..., 2] = -st[..., 2], st[..., 1]
    for scheme in SCHEMES:
        r = o.residual(par, scheme, VANALBADA, v, st)
        rr = o.residual(parr, scheme, VANALBADA, vr, str_)
        assert np.array_equal(rr[..., 0], r[..., 0])
        assert np.array_equal(rr[..., 3], r[..., 3])
        assert np.array_equal(rr[..., 1], -r[..., 2])
        assert np.array_equal(rr[..., 2], r[..., 1])


def test_symmetric_airfoil_zero_lift():
    """NACA 0012 at zero incidence on the mirror-symmetric grid stays
    mirror-symmetric (cell i <-> M-1-i) and carries no lift."""
    o = Oracle(64)
    M, N = 64, 32
    v = _naca_grid(o, M, N, code=12, iters=200, spacing=0)
    par = make_params(0.8, 0.0, dtype=np.float64)
    s = Solver(64, par, SLAU2, VANALBADA, v)
    s.step(40)
    st = s.state[1:N + 1]
    mir = st[:, ::-1]
    assert np.allclose(st[..., 0], mir[..., 0], rtol=1e-9, atol=1e-12)
    assert np.allclose(st[..., 2], -mir[..., 2], rtol=1e-9, atol=1e-9)
    f = o.forces(par, s.state, v, ref_len=1.0)
    assert abs(f[2]) < 1e-9, f


def test_rk3_amplification():
    """Applying the three stage formulas to dQ/dt = lambda*Q gives the
    third-order polynomial 1 + z + z^2/2 + z^3/6 (z = lambda*dt)."""
    o = Oracle(64)
    M, N = 4, 3
    rng = np.random.default_rng(9)
    Qn = rng.uniform(0.5, 2, (N + 3, M, 4))
    lam = -3.7
    dt = 0.11
    f = o.f("stage_pass")
    from oracle_lib import _ptr

    def res_of(Q):
        return np.ascontiguousarray(lam * Q[1:N + 1])

    s1 = np.zeros_like(Qn)
    f(1, M, N, dt, _ptr(res_of(Qn)), _ptr(Qn), None, _ptr(s1))
    s2 = np.zeros_like(Qn)
    f(2, M, N, dt, _ptr(res_of(s1)), _ptr(Qn), _ptr(s1), _ptr(s2))
    s3 = np.zeros_like(Qn)
    f(3, M, N, dt, _ptr(res_of(s2)), _ptr(Qn), _ptr(s2), _ptr(s3))
    z = lam * dt
    amp = 1 + z + z * z / 2 + z ** 3 / 6
    assert np.allclose(s3[1:N + 1], amp * Qn[1:N + 1], rtol=1e-14)


def test_cfl_cartesian_spectral_radius():
    """On rectangles dx*dy the wave speed is (|u|+a)(1/dx+1/dy)."""
    o = Oracle(64)
    M, N = 6, 5
    dx, dy = 0.3, 0.07
    xs = np.arange(M + 1) * dx
    ys = np.arange(N + 1) * dy
    # a Cartesian patch stored as (N+1) x M vertex rows; the finalize pass
    # wraps i, so build metrics from an explicit (M+1)-column patch instead
    area = np.full((N, M), dx * dy)
    flen = np.zeros((N, M, 4))
    flen[..., 0] = flen[..., 1] = dy
    flen[..., 2] = flen[..., 3] = dx
    par = make_params(0.5, 0.0, dtype=np.float64)
    st = o.init_state(par, M, N)
    g = float(par[P_GAMMA])
    rho, mx, my, E = st[1, 0]
    u = math.hypot(mx / rho, my / rho)
    c = math.sqrt(g * (E - 0.5 * rho * u * u) * (g - 1) / rho)
    assert abs(u - 0.5) < 1e-6 and abs(c - 1.0) < 1e-6
    ws = o.max_ws(par, st, area, flen)
    assert math.isclose(ws, (u + c) * (1 / dx + 1 / dy), rel_tol=1e-12)
    assert xs.size == M + 1 and ys.size == N + 1


def test_grid_annulus_closed_form():
    """Circle body in a concentric circular far field: the Winslow (P=Q=0)
    fixed point is r(eta) = r0 (r1/r0)^eta at uniform angle, to O(h^2)."""
    o = Oracle(64)
    M, N = 64, 16
    r0, r1 = 1.0, 3.0
    v = o.grid(circle_boundary(M, r0), circle_boundary(M, r1), N, 6000)
    r = np.hypot(v[..., 0], v[..., 1])
    eta = np.arange(N + 1) / N
    exact = r0 * (r1 / r0) ** eta
    assert np.abs(r.mean(axis=1) - exact).max() / r1 < 5e-3
    # linear interpolation alone (no smoothing) is off by much more
    v0 = o.grid(circle_boundary(M, r0), circle_boundary(M, r1), N, 0)
    r_lin = np.hypot(v0[..., 0], v0[..., 1]).mean(axis=1)
    assert np.abs(r_lin - exact).max() / r1 > 5e-2
    # angles stay uniform (rotational symmetry of the discrete problem)
    ang = np.unwrap(np.arctan2(v[N // 2, :, 1], v[N // 2, :, 0]))
    assert np.allclose(np.diff(ang), 2 * math.pi / M, atol=1e-9)


def test_grid_jacobian_guard():
    """A point whose local Jacobian is not negative is left where it is."""
    o = Oracle(64)
    M, N = 16, 6
    v = o.grid(circle_boundary(M, 1.0), circle_boundary(M, 2.0), N, 0)
    # mirror the grid (x -> -x) so the orientation flips: no point may move
    vm = v.copy()
    vm[..., 0] *= -1
    out = o.jacobi_sweep(vm)
    assert np.array_equal(out, vm)
    out2 = o.jacobi_sweep(v)
    assert not np.array_equal(out2, v)


def test_naca_textbook_ordinates():
    """NACA 4-digit closed forms: 12% thickness peaks near x=0.3c with
    y_t = 0.0600 (Abbott & von Doenhoff give 6.002% for the open-TE 0012),
    the closed trailing edge has zero thickness, and NACA 2412's mean line
    reaches 2% camber at 40% chord."""
    M = 400
    c = 1.0
    b = naca4_boundary(M, 0.0, 0.0, 0.12, size=c / 2, x_offset=0.0, expo=1.0, spacing=0)
    # spacing 0 with exponent 1: the upper x runs linearly; i = 140 -> x = 0.3
    i = int(round(M / 2 * (1 - 0.3)))
    assert math.isclose(b[i, 0], 0.3 - 0.5, abs_tol=1e-12)
    assert abs(b[i, 1] - 0.06002) < 1e-4
    assert abs(b[0, 1]) < 1e-12 and math.isclose(b[0, 0], 0.5, abs_tol=1e-12)
    assert math.isclose(b[M // 2, 0], -0.5, abs_tol=1e-12) and abs(b[M // 2, 1]) < 1e-12
    b2 = naca4_boundary(M, 0.02, 0.4, 0.12, size=0.5, expo=1.0, spacing=0)
    j = int(round(M / 2 * (1 - 0.4)))
    mid_y = 0.5 * (b2[j, 1] + b2[M - j, 1])
    assert math.isclose(mid_y, 0.02, abs_tol=1e-12)
    # counter-clockwise from the trailing edge: upper surface first
    assert b2[M // 4, 1] > 0 > b2[3 * M // 4, 1]


def test_cst_closed_forms():
    """Kulfan CST: equal coefficients reduce the Bernstein sum to 1 (partition
    of unity) so y = A sqrt(x)(1-x); degree-elevated coefficients give the
    same surface; the nose radius is A_0^2/2; mirrored coefficients mirror
    the section; the trailing-edge gap is te_gap."""
    M = 1024
    A = 0.17
    b = cst_boundary(M, [A] * 6, [-A] * 6, size=0.5, expo=1.4, spacing=0)
    x = b[:, 0] + 0.5
    up = np.arange(M) <= M // 2
    sgn = np.where(up, 1.0, -1.0)
    assert np.allclose(b[:, 1], sgn * A * np.sqrt(x) * (1 - x), rtol=0, atol=2e-16)
    # degree elevation n=5 -> 6: A'_i = i/6 A_{i-1} + (1 - i/6) A_i
    au = np.array([0.13, 0.13, 0.16, 0.14, 0.18, 0.20])
    al = np.array([-0.13, -0.14, -0.10, -0.03, 0.03, 0.06])

    def elevate(a):
        n1 = a.size
        return np.array([(i / n1) * (a[i - 1] if i > 0 else 0.0) +
                         (1 - i / n1) * (a[i] if i < n1 else 0.0) for i in range(n1 + 1)])

    b5 = cst_boundary(M, au, al, spacing=1)
    b6 = cst_boundary(M, elevate(au), elevate(al), spacing=1)
    assert np.allclose(b5, b6, rtol=0, atol=1e-15)
    # mirror: a_lower = -a_upper gives y_l(x) = -y_u(x) (indices i and M - i)
    bm = cst_boundary(M, au, -au, spacing=1)
    assert np.allclose(bm[1:M // 2, 1], -bm[M - 1:M // 2:-1, 1], rtol=0, atol=1e-16)
    # nose radius: y^2 / (2x) -> A_0^2 / 2 at the first stations off the nose
    bn = cst_boundary(4096, au, al, spacing=1)
    xn, yn = bn[2047, 0] + 0.5, bn[2047, 1]
    assert xn < 1e-5 and math.isclose(yn * yn / (2 * xn), au[0] ** 2 / 2, rel_tol=1e-4)
    # trailing-edge gap
    bt = cst_boundary(M, au, al, te_gap=0.004, spacing=1)
    assert math.isclose(bt[0, 1], 0.002, abs_tol=1e-16) and bt[0, 0] == 0.5


def test_bezier_closed_forms():
    """Bezier oracle (Bernstein form): endpoints interpolate the end control
    points; control points evenly spaced on a line give that line with
    x = t (linear precision); the quadratic (0,0), (1/2,h), (1,0) is the
    parabola y = 2h x(1-x); a degree-elevated polygon draws the same curve;
    a lower polygon mirrored from the upper mirrors the section."""
    from oracle_lib import bezier_boundary as o_bez

    M = 512
    up = np.arange(M) <= M // 2
    # linear precision on both sides: y = +-0.1 x, stations x = t
    n = 5
    line_u = np.array([[i / (n - 1), 0.1 * i / (n - 1)] for i in range(n)])
    line_l = line_u * [1.0, -1.0]
    b = o_bez(M, line_u, line_l, size=0.5, x_offset=0.0, expo=1.4, spacing=0)
    s = np.arange(M) / (M / 2)
    tc = np.where(up, (1 - np.minimum(s, 1)) ** 1.4, (np.maximum(s, 1) - 1) ** 1.4)
    assert np.allclose(b[:, 0] + 0.5, tc, rtol=0, atol=2e-16)
    assert np.allclose(b[:, 1], np.where(up, 0.1, -0.1) * tc, rtol=0, atol=2e-16)
    # endpoints: index 0 is the upper trailing edge, M/2 the leading edge
    cu = np.array([[0.0, 0.0], [0.0, 0.05], [0.3, 0.09], [0.7, 0.05], [1.0, 0.002]])
    cl = np.array([[0.0, 0.0], [0.0, -0.04], [0.4, -0.05], [1.0, -0.002]])
    bb = o_bez(M, cu, cl, size=0.5, spacing=1)
    assert np.allclose(bb[0], cu[-1] - [0.5, 0.0], rtol=0, atol=1e-16)
    assert np.allclose(bb[M // 2], cu[0] - [0.5, 0.0], rtol=0, atol=1e-16)
    # quadratic = parabola y = 2 h x (1 - x), x = t
    h = 0.12
    q = np.array([[0.0, 0.0], [0.5, h], [1.0, 0.0]])
    bq = o_bez(M, q, q * [1.0, -1.0], size=0.5, spacing=1)
    x = bq[:, 0] + 0.5
    assert np.allclose(bq[:, 1], np.where(up, 1.0, -1.0) * 2 * h * x * (1 - x), rtol=0, atol=1e-16)

    # degree elevation: P'_i = i/(n+1) P_(i-1) + (1 - i/(n+1)) P_i
    def elevate(p):
        n1 = p.shape[0]
        return np.array([(i / n1) * (p[i - 1] if i > 0 else 0.0) +
                         (1 - i / n1) * (p[i] if i < n1 else 0.0) for i in range(n1 + 1)])

    be = o_bez(M, elevate(cu), elevate(elevate(cl)), size=0.5, spacing=1)
    assert np.allclose(bb, be, rtol=0, atol=1e-15)
    # mirror
    bm = o_bez(M, cu, cu * [1.0, -1.0], size=0.5, spacing=1)
    # (stations i and M - i agree to an ulp of the cosine)
    assert np.allclose(bm[1:M // 2, 1], -bm[M - 1:M // 2:-1, 1], rtol=0, atol=4e-16)
    assert np.allclose(bm[1:M // 2, 0], bm[M - 1:M // 2:-1, 0], rtol=0, atol=4e-16)


def test_ref32_tracks_ref64():
    """On a smooth subsonic case the FP32 and FP64 oracles agree to FP32
    rounding growth over a few dozen steps."""
    o64 = Oracle(64)
    M, N = 48, 24
    v = _naca_grid(o64, M, N, code=12, iters=100)
    v32 = v.astype(np.float32)
    par = make_params(0.5, 2.0)
    s32 = Solver(32, par, SLAU2, VANALBADA, v32)
    s64 = Solver(64, par.astype(np.float64), SLAU2, VANALBADA, v32.astype(np.float64))
    s32.step(30)
    s64.step(30)
    d = np.abs(s32.state[1:N + 1] - s64.state[1:N + 1]).max(axis=(0, 1))
    scale = np.abs(s64.state[1:N + 1]).max(axis=(0, 1))
    assert (d / scale).max() < 1e-5, d / scale


def test_bitwise_repeatable():
    o = Oracle(32)
    v = _naca_grid(o, 32, 16, iters=50)
    par = make_params(1.5, 2.0)
    a = Solver(32, par, AUSMUP, VANALBADA, v)
    b = Solver(32, par, AUSMUP, VANALBADA, v)
    a.step(10)
    b.step(10)
    assert np.array_equal(a.state, b.state)


def test_dt_lag_cadence():
    """dt is refreshed from the state after every `refresh` steps (once per
    frame in the reference) and is constant in between."""
    o = Oracle(32)
    v = _naca_grid(o, 32, 16, iters=50)
    par = make_params(0.8, 0.0, maxdt=1.0)
    s = Solver(32, par, SLAU2, VANALBADA, v, refresh=3)
    ws0 = float(s.maxws[0])
    s.step(2)
    assert float(s.maxws[0]) == ws0
    s.step(1)
    assert float(s.maxws[0]) != ws0
    assert float(s.maxws[0]) == o.max_ws(par, s.state, s.area, s.flen)
    assert par[P_GAMMA] == np.float32(1.4) and par[P_INP] == np.float32(1 / 1.4)


def test_local_dt_reduces_to_global_when_maxdt_binds():
    """Local time stepping with maxdt below every cell's CFL step takes dt =
    maxdt everywhere, exactly the global step."""
    o = Oracle(64)
    v = _naca_grid(o, 48, 16, iters=60)
    par = make_params(0.7, 2.0, maxdt=1e-5, dtype=np.float64)
    area, flen = o.finalize(v)
    assert float(par[9]) < float(par[10]) / o.max_ws(par, o.init_state(par, 48, 16), area, flen)
    g = Solver(64, par, SLAU2, VANALBADA, v)
    loc = Solver(64, par, SLAU2, VANALBADA, v, local_dt=True)
    g.step(6)
    loc.step(6)
    assert np.array_equal(g.state, loc.state)


def test_local_dt_is_the_cells_own_cfl_step():
    """With a tiny CFL number one RK3 step is Q + dt_c L(Q) + O(dt^2): the
    ratio of the density change to the density residual recovers each
    cell's dt_c = cfl / ws_c (and the global mode's single min step)."""
    o = Oracle(64)
    M, N = 48, 16
    v = _naca_grid(o, M, N, iters=60)
    par = make_params(0.7, 2.0, cfl=1e-4, maxdt=1.0, dtype=np.float64)
    st = o.init_state(par, M, N)
    rng = np.random.default_rng(3)
    st[..., 0] *= 1 + 0.01 * rng.standard_normal(st.shape[:2])
    area, flen = o.finalize(v)
    _, ws = o.max_ws(par, st, area, flen, want_field=True)
    L0 = o.residual(par, SLAU2, VANALBADA, v, st)[..., 0]
    mask = np.abs(L0) > 1e-2 * np.abs(L0).max()
    for local, expect in ((True, 1e-4 / ws), (False, np.full_like(ws, 1e-4 / ws.max()))):
        s = Solver(64, par, SLAU2, VANALBADA, v, local_dt=local)
        s.set_state(st)
        s.step(1)
        ratio = (s.state[1:N + 1, :, 0] - st[1:N + 1, :, 0]) / L0
        assert np.allclose(ratio[mask], expect[mask], rtol=1e-3, atol=0), local
    assert ws.max() / ws.min() > 5  # the two modes differ on this grid


def test_local_dt_shares_the_steady_state():
    """Local time stepping converges to a fixed point of the discrete
    operator, which the global-dt scheme then leaves unchanged."""
    o = Oracle(64)
    M, N = 40, 12
    v = _naca_grid(o, M, N, iters=60)
    par = make_params(0.5, 1.0, cfl=0.5, maxdt=1.0, dtype=np.float64)
    loc = Solver(64, par, SLAU2, VANALBADA, v, local_dt=True)
    r0 = np.sqrt((o.residual(par, SLAU2, VANALBADA, v, loc.state)[..., 0] ** 2).sum())
    loc.step(6000)
    r1 = np.sqrt((o.residual(par, SLAU2, VANALBADA, v, loc.state)[..., 0] ** 2).sum())
    assert r1 < 1e-7 * r0, r1 / r0
    g = Solver(64, par, SLAU2, VANALBADA, v)
    g.set_state(loc.state)
    g.step(200)
    q0 = loc.state[1:N + 1]
    assert np.abs(g.state[1:N + 1] - q0).max() < 1e-9 * np.abs(q0).max()


def test_compensated_increment_form_is_the_same_scheme():
    """In FP64 the compensated increment form Q^n + (K1 + K2 + 4 K3)/6 and the
    reference's Shu-Osher stage form agree to round-off after many steps
    (a wrong stage weight would not)."""
    o = Oracle(64)
    M, N = 48, 16
    v = _naca_grid(o, M, N, iters=60)
    par = make_params(0.7, 2.0, cfl=0.5, maxdt=1.0, dtype=np.float64)
    st = o.init_state(par, M, N)
    rng = np.random.default_rng(9)
    st[..., 0] *= 1 + 0.02 * rng.standard_normal(st.shape[:2])
    for local in (False, True):
        a = Solver(64, par, SLAU2, VANALBADA, v, local_dt=local)
        b = Solver(64, par, SLAU2, VANALBADA, v, local_dt=local, comp=True)
        a.set_state(st)
        b.set_state(st)
        a.step(30)
        b.step(30)
        qa = a.state[1:N + 1]
        qb = b.state[1:N + 1] + b.lo[1:N + 1]
        assert np.abs(qa - qb).max() < 1e-12 * np.abs(qa).max(), local
        assert abs(a.last_dt - b.last_dt) <= 1e-12 * a.last_dt


def test_compensated_state_accumulates_sub_ulp_increments():
    """With increments far below half an ulp of Q (cfl = 1e-12), plain FP32
    accumulation cannot carry them: the state only moves by the round-off of
    the stage combinations (at most an ulp, uncorrelated with L).  The
    compensated pair accumulates them: after n steps hi is unchanged and lo
    equals n dt L(Q0) (all stages see Q0, and (1 + 1 + 4)/6 = 1)."""
    o32, o64 = Oracle(32), Oracle(64)
    M, N = 48, 16
    v = _naca_grid(o64, M, N, iters=60).astype(np.float32)
    par = make_params(0.7, 2.0, cfl=1e-12, maxdt=1.0)
    st = o32.init_state(par, M, N)
    rng = np.random.default_rng(5)
    st[..., 0] *= (1 + 0.02 * rng.standard_normal(st.shape[:2])).astype(np.float32)
    n = 50
    plain = Solver(32, par, SLAU2, VANALBADA, v, local_dt=True)
    comp = Solver(32, par, SLAU2, VANALBADA, v, local_dt=True, comp=True)
    plain.set_state(st)
    comp.set_state(st)
    plain.step(n)
    comp.step(n)
    q0 = st[1:N + 1].astype(np.float64)
    q0f = st[1:N + 1]
    ulp = np.spacing(np.abs(q0f)).astype(np.float64)
    assert (np.abs(plain.state[1:N + 1].astype(np.float64) - q0) <= ulp).all()
    assert np.array_equal(comp.state[1:N + 1], q0f)  # hi unchanged
    moved = comp.lo[1:N + 1].astype(np.float64)
    area, flen = o32.finalize(v)
    _, ws = o32.max_ws(par, st, area, flen, want_field=True)
    L = o32.residual(par, SLAU2, VANALBADA, v, st).astype(np.float64)
    expect = n * (float(par[10]) / ws.astype(np.float64))[..., None] * L
    mask = np.abs(expect) > 1e-3 * np.abs(expect).max()
    assert np.allclose(moved[mask], expect[mask], rtol=1e-3, atol=0)
    assert np.abs(moved).max() > 0


def test_openmp_build_is_bitwise_ref32():
    """The OpenMP -O3 build of ref32 (the CPU-baseline leg and the long-run
    checker) gives bit-for-bit the results of the plain build."""
    o = Oracle(64)
    M, N = 64, 24
    v = _naca_grid(o, M, N, iters=60).astype(np.float32)
    par = make_params(0.8, 1.25, cfl=0.5, maxdt=1.0)
    for local in (False, True):
        a = Solver(32, par, SLAU2, MINMOD, v, local_dt=local)
        b = Solver(32, par, SLAU2, MINMOD, v, local_dt=local, omp=True)
        a.step(25)
        b.step(25)
        assert np.array_equal(a.state, b.state) and a.last_dt == b.last_dt
