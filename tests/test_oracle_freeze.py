"""Limiter freezing on the oracle (a build extension with no counterpart in
the reference: the reference's MUSCL re-evaluates its limiter every stage).

With the reference's ratio r = (dL dF + eps) / (dF^2 + eps), psi = phi(r)/r
and w = dF^2 / (dF^2 + eps), the limited face value W_L + phi(r_L) dF / 2 is
exactly W_L + psi_L (w dL + (1 - w) dF) / 2 (and likewise on the right).
Freezing records psi and w of every face and reconstructs with them.

Pins, chosen so that a plausible mistake fails one of them:
  * the identity: at the recorded state the frozen residual is the limited
    one (FP64 to 1e-12 relative, FP32 to 1e-4), for van Albada and the
    clamped van Albada, on both face families (a dropped w, a swapped dL /
    dR or a mixed-up psi_L / psi_R breaks it by O(1));
  * the record's layout and values against the stencils of the
    boundary-applied state, formed here from the primitive variables
    (face order, component order, the clamped limiter's psi = 0 for r <= 0);
  * freezing fixes the operator: once the state has moved, frozen and
    limited residuals differ by more than rounding, and a new state ends
    freezing;
  * it converges a stalled solve: a 48 x 32 transonic section (CST, t/c
    0.11, Mach 0.73, AoA 2.5 deg, SLAU2 + van Albada, FP32 compensated
    state, local time step) holds a limit cycle above a 1e-3 drop of the
    density-residual norm over 8000 steps; frozen after 2000 steps it falls
    below 1e-5 within the same 8000.
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol
from oracle_lib import (SLAU2, AUSMUP, VANALBADA, VANALBADA_CLAMPED, Oracle, Solver,
                        make_params, make_params64)

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "ausm-webgpu_amd"))

EPS = 1e-6


def _grid(M, N, prec=64):
    body = ol.naca4_boundary(M, 0.02, 0.4, 0.12)
    v = Oracle(64).grid(body, ol.circle_boundary(M, 5.0), N, 400)
    return v.astype(np.float32) if prec == 32 else v


def _moved_solver(prec, scheme, limiter, M=40, N=20, steps=120):
    vtx = _grid(M, N, prec)
    if prec == 64:
        fs = (1.0, 1 / 1.4, 0.75 * math.cos(0.02), 0.75 * math.sin(0.02), 1.4)
        par = make_params64(fs, cfl=0.5, maxdt=1.0)
    else:
        par = make_params(0.75, math.degrees(0.02), cfl=0.5, maxdt=1.0)
    s = Solver(prec, par, scheme, limiter, vtx)
    s.step(steps)
    return s


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float((np.abs(a - b).max(axis=(0, 1)) / np.abs(b).max(axis=(0, 1))).max())


@pytest.mark.parametrize("limiter", [VANALBADA, VANALBADA_CLAMPED])
@pytest.mark.parametrize("scheme", [SLAU2, AUSMUP])
@pytest.mark.parametrize("prec,tol", [(64, 1e-12), (32, 1e-4)])
def test_frozen_residual_is_the_limited_one_at_the_record(prec, tol, scheme, limiter):
    s = _moved_solver(prec, scheme, limiter)
    r_lim = s.residual()
    s.freeze_limiter()
    r_frz = s.residual()
    assert s.psi is not None and np.isfinite(s.psi).all()
    assert _rel(r_frz, r_lim) < tol
    # both face families carry the identity: with one family's record
    # scrambled the residual moves by O(1) (so each family is exercised)
    M, N = s.M, s.N
    nv = (N + 1) * M * 12
    for sl in (slice(0, nv), slice(nv, None)):
        bad = s.psi.copy()
        bad[sl] = bad[sl][::-1]
        good = s.psi
        s.psi = bad
        r_bad = s.residual()
        s.psi = good
        assert _rel(r_bad, r_lim) > 1e-3


def _prims(st, gamma):
    rho = st[..., 0]
    u, v = st[..., 1] / rho, st[..., 2] / rho
    p = np.maximum(1e-7, (st[..., 3] - (0.5 * rho) * (u * u + v * v)) * (gamma - 1.0))
    return np.stack([rho, u, v, p], axis=-1)


@pytest.mark.parametrize("limiter", [VANALBADA, VANALBADA_CLAMPED])
def test_record_layout_and_values(limiter):
    s = _moved_solver(64, SLAU2, limiter)
    s.freeze_limiter()
    M, N = s.M, s.N
    W = _prims(s.o.boundary(s.par, s.vtx, s.s2), 1.4)  # rows -1 .. N+1 of the texture

    def row(y):  # texture row of cell row y (clamped at the body ghost)
        return W[max(y + 1, 0)]

    def expect(WL2, WL, WR, WR2):
        dL, dF, dR = WL - WL2, WR - WL, WR2 - WR
        rL = (dL * dF + EPS) / (dF * dF + EPS)
        rR = (dF * dR + EPS) / (dF * dF + EPS)

        def psi(r):
            out = (r + 1.0) / (r * r + 1.0)
            return np.where(r > 0, out, 0.0) if limiter == VANALBADA_CLAMPED else out
        return np.concatenate([psi(rL), psi(rR), dF * dF / (dF * dF + EPS)], axis=-1)

    rec = s.psi.reshape(-1, 12)
    vert = rec[:(N + 1) * M].reshape(N + 1, M, 12)
    horz = rec[(N + 1) * M:].reshape(N, M, 12)
    ev = np.stack([expect(row(y - 2), row(y - 1), row(y), row(y + 1)) for y in range(N + 1)])
    Wi = W[1:N + 1]
    eh = expect(np.roll(Wi, 2, axis=1), np.roll(Wi, 1, axis=1), Wi, np.roll(Wi, -1, axis=1))
    assert np.allclose(vert, ev, rtol=1e-12, atol=1e-14)
    assert np.allclose(horz, eh, rtol=1e-12, atol=1e-14)
    if limiter == VANALBADA_CLAMPED:
        assert (rec[:, :8] == 0).any() and (rec[:, :8] >= 0).all()
    else:
        assert (rec[:, :8] < 0).any()  # van Albada's psi is negative for r < -1


def test_freezing_fixes_the_operator_and_a_new_state_ends_it():
    s = _moved_solver(64, SLAU2, VANALBADA)
    s.freeze_limiter()
    s.step(200)
    r_frz = s.residual()
    psi = s.psi
    s.psi = None
    r_lim = s.residual()
    s.psi = psi
    assert _rel(r_frz, r_lim) > 1e-8
    assert np.array_equal(s.residual(), r_frz)
    s.set_state(s.state.copy())
    assert s.psi is None


def _cst_case(M=48, N=32):
    from ausm.precision import cst_coefficients
    au, al = cst_coefficients(0, 0.11)
    body = ol.cst_boundary(M, au, al, size=0.5, expo=1.4, spacing=0)
    vtx = Oracle(64).grid(body, ol.circle_boundary(M, 20.0), N, 3000).astype(np.float32)
    return Solver(32, make_params(0.73, 2.5, cfl=0.5, maxdt=1.0), SLAU2, VANALBADA, vtx,
                  local_dt=True, comp=True, omp=True)


def _drops(s, chunks, freeze_after=None):
    def norm():
        return math.sqrt((s.residual()[..., 0].astype(np.float64) ** 2).sum())
    r0 = norm()
    out = []
    for k in range(chunks):
        s.step(500)
        if freeze_after is not None and (k + 1) * 500 == freeze_after:
            s.freeze_limiter()
        out.append(norm() / r0)
    return out


def test_freezing_converges_a_stalled_muscl_solve():
    plain = _drops(_cst_case(), 16)
    frozen = _drops(_cst_case(), 16, freeze_after=2000)
    print("limited", " ".join(f"{d:.1e}" for d in plain[3::4]))
    print("frozen ", " ".join(f"{d:.1e}" for d in frozen[3::4]))
    assert min(plain[4:]) > 1e-3
    assert frozen[-1] < 1e-5
