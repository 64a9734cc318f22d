"""GPU parity of the all-FP64 solve against ref64 (the oracle in FP64).

The FP64 precision runs the exact stage arithmetic in double on FP64 state
and the FP64 grid, with the free stream unrounded.  Bar: bit-for-bit equal
to ref64 for residuals, states after any number of steps (across captured
graphs and dt-refresh cadences), forces and Cp; the residual norm within
relative 1e-12 (FP64 partial sums in another order); precision switches
carry the solution over exactly (FP32 -> FP64 widens, FP64 -> FP32 rounds).
"""
import ctypes
import math

import numpy as np
import pytest

import ausm
from oracle_lib import (AUSMUP, MINMOD, NONE, SLAU, SLAU2, VANALBADA, VANALBADA_CLAMPED, VANLEER,
                        Oracle, Solver,
                        circle_boundary, make_params64, naca4_boundary, naca4_digits)

pytestmark = pytest.mark.gpu

SCHEMES = [SLAU2, SLAU, AUSMUP]
LIMITERS = [VANALBADA, MINMOD, VANLEER, NONE, VANALBADA_CLAMPED]


def grid64(M, N, code=2412, iters=150, radius=10.0):
    m, p, t = naca4_digits(code)
    body = naca4_boundary(M, m, p, t, size=0.5, x_offset=0.0, expo=1.4, spacing=1)
    return Oracle(64).grid(body, circle_boundary(M, radius), N, iters)


def perturbed_state64(par, M, N, seed, amp=0.01):
    rng = np.random.default_rng(seed)
    g, rho0, p0 = par[0], par[5], par[2]
    u0, v0 = par[6] / rho0, par[7] / rho0
    jj, ii = np.meshgrid(np.arange(N + 3), np.arange(M), indexing="ij")
    ph = rng.uniform(0, 2 * math.pi, 4)
    k = rng.integers(1, 4, 4)
    w = [1 + amp * np.sin(2 * math.pi * k[c] * ii / M + 0.37 * jj + ph[c]) for c in range(4)]
    rho = rho0 * w[0]
    u = u0 * w[1] + 0.05 * amp * np.cos(0.5 * jj)
    v = v0 * w[2] + 0.05 * amp * np.sin(0.3 * jj + ii)
    p = p0 * w[3]
    return np.stack([rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)], axis=-1)


def new_fp64(vtx, mach, aoa, scheme, limiter, refresh=1, cfl=0.5, maxdt=1.0, tile=0):
    fs = ausm.freestream(mach, aoa)
    par = make_params64(fs, cfl=cfl, maxdt=maxdt)
    N1, M, _ = vtx.shape
    s = ausm.Solver(M, N1 - 1)
    s.set_vertices(vtx)
    s.set_freestream(*fs)
    s.set_scheme(scheme, limiter, 0.0, 0.75)
    s.set_cfl(cfl, maxdt, refresh)
    s.set_tile(tile)
    s.set_precision("fp64")
    return s, par


def max_rel(a, b):
    d = np.abs(a - b).reshape(-1, 4).max(axis=0)
    sc = np.abs(b).reshape(-1, 4).max(axis=0)
    return float((d / np.maximum(sc, 1e-300)).max())


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("limiter", LIMITERS)
def test_fp64_residual_bitwise(scheme, limiter):
    M, N = 96, 40
    vtx = grid64(M, N)
    s, par = new_fp64(vtx, 0.75, 2.0, scheme, limiter)
    st = perturbed_state64(par, M, N, seed=scheme * 7 + limiter)
    s.set_state(st)
    r_gpu = s.residual()
    assert r_gpu.dtype == np.float64
    r_ref = Oracle(64).residual(par, scheme, limiter, vtx, st)
    assert np.array_equal(r_gpu, r_ref), max_rel(r_gpu, r_ref)
    s.close()


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("mach,aoa", [(0.73, 2.5), (2.0, 0.0)])
def test_fp64_steps_bitwise(scheme, mach, aoa):
    M, N = 96, 40
    vtx = grid64(M, N, code=12)
    s, par = new_fp64(vtx, mach, aoa, scheme, VANALBADA)
    ref = Solver(64, par, scheme, VANALBADA, vtx)
    st = perturbed_state64(par, M, N, seed=11 + scheme)
    s.set_state(st)
    ref.set_state(st)
    for chunk in (1, 5, 40):  # 40 steps crosses one captured 32-step graph
        info = s.step(chunk)
        ref.step(chunk)
        got = s.state()
        assert np.array_equal(got[1:N + 1], ref.state[1:N + 1]), (chunk, max_rel(got, ref.state))
        assert info["dt"] == np.float32(ref.last_dt)
    s.close()


@pytest.mark.parametrize("tile", [1, 2, 4])
def test_fp64_tiles_and_cadence_bitwise(tile):
    """Every FP64 tile shape, with the dt refreshed every 7 steps."""
    M, N = 70, 24
    vtx = grid64(M, N, code=4412, iters=100)
    s, par = new_fp64(vtx, 0.6, 3.0, SLAU2, MINMOD, refresh=7, tile=tile)
    ref = Solver(64, par, SLAU2, MINMOD, vtx, refresh=7)
    st = perturbed_state64(par, M, N, seed=5)
    s.set_state(st)
    ref.set_state(st)
    for chunk in (3, 11, 70):
        s.step(chunk)
        ref.step(chunk)
        assert np.array_equal(s.state()[1:N + 1], ref.state[1:N + 1]), chunk
    s.close()


def test_fp64_ghost_rows_forces_norm():
    M, N = 96, 40
    vtx = grid64(M, N)
    s, par = new_fp64(vtx, 0.73, 2.5, SLAU2, VANALBADA)
    st = perturbed_state64(par, M, N, seed=17)
    s.set_state(st)
    o = Oracle(64)
    # interior as stored, ghost rows as the boundary pass forms them
    got, b = s.state(), o.boundary(par, vtx, st)
    assert np.array_equal(got[1:N + 1], st[1:N + 1])
    assert np.array_equal(got[0], b[0]) and np.array_equal(got[N + 1:], b[N + 1:])
    f = s.forces(ref_length=1.0, want_cp=True)
    out, cp = o.forces(par, st, vtx, 1.0, want_cp=True)
    assert [f["fx"], f["fy"], f["cl"], f["cd"]] == list(out)
    assert np.array_equal(f["cp"], cp)
    r = o.residual(par, SLAU2, VANALBADA, vtx, st)
    expect = math.sqrt(float((r[..., 0] ** 2).sum()))
    assert abs(s.residual_norm() - expect) <= 1e-12 * expect
    area, flen = o.finalize(vtx)
    ws, mx = s.wave_speeds()
    mref, wref = o.max_ws(par, st, area, flen, want_field=True)
    assert mx == np.float32(mref) and np.array_equal(ws, wref.astype(np.float32))
    s.close()


def test_fp64_on_the_gpu_grid():
    """FP64 solve on the grid generated on the GPU (ping-pong buffers), read
    back at full precision for the oracle."""
    M, N = 64, 24
    s = ausm.Solver(M, N)
    m, p, t = naca4_digits(2412)
    s.update_grid_boundaries(naca4_boundary(M, m, p, t, size=0.5, spacing=1),
                             circle_boundary(M, 8.0))
    s.prepare_grid(iters=37)  # odd: the result sits in the second buffer pair
    vtx = s.vertices(fp64=True)
    fs = ausm.freestream(0.8, 1.25)
    par = make_params64(fs, cfl=0.5, maxdt=1.0)
    s.set_freestream(*fs)
    s.set_scheme(SLAU, VANLEER, 0.0, 0.75)
    s.set_cfl(0.5, 1.0, 1)
    s.set_precision("fp64")
    s.prepare_state(freestream=True)
    ref = Solver(64, par, SLAU, VANLEER, vtx)
    s.step(33)
    ref.step(33)
    assert np.array_equal(s.state()[1:N + 1], ref.state[1:N + 1])
    s.close()


def test_precision_switch_carries_the_solution():
    M, N = 64, 24
    vtx = grid64(M, N)
    fs = ausm.freestream(0.8, 1.25)
    s = ausm.Solver(M, N)
    s.set_vertices(vtx.astype(np.float32))
    s.set_freestream(*fs)
    s.set_scheme(SLAU2, VANALBADA, 0.0, 0.75)
    s.set_cfl(0.5, 1.0, 1)
    s.prepare_state(freestream=True)
    s.step(10)
    q32 = s.state()
    s.set_precision("fp64")
    q64 = s.state()
    assert np.array_equal(q64[1:N + 1], q32[1:N + 1].astype(np.float64))
    s.step(5)
    q64b = s.state()
    s.set_precision("fp32")
    assert np.array_equal(s.state()[1:N + 1], q64b[1:N + 1].astype(np.float32))
    s.step(1)
    # the FP64 accessors refuse an FP32 context
    buf = np.zeros((N + 3, M, 4), dtype=np.float64)
    assert s._L.ausm_get_state64(s.h, buf.ctypes.data_as(ctypes.c_void_p)) == -4
    s.close()


def test_fp64_local_dt_bitwise():
    M, N = 96, 40
    vtx = grid64(M, N)
    s, par = new_fp64(vtx, 0.73, 2.5, SLAU2, VANALBADA)
    s.set_dt_mode("local")
    ref = Solver(64, par, SLAU2, VANALBADA, vtx, local_dt=True)
    st = perturbed_state64(par, M, N, seed=23)
    s.set_state(st)
    ref.set_state(st)
    for chunk in (1, 40):
        s.step(chunk)
        ref.step(chunk)
        assert np.array_equal(s.state()[1:N + 1], ref.state[1:N + 1]), chunk
    s.close()


def test_fp64_graphs_follow_a_new_grid():
    """A grid change on a context that has stepped through captured FP64
    graphs: the FP64 solve reads the grid solver's vertex pair, which an odd
    number of Jacobi sweeps leaves in the other ping-pong buffer, so graphs
    captured on the old grid must not be replayed.  Stepping after the
    regeneration is bit-for-bit a fresh context's run on the new grid."""
    M, N = 64, 24
    m, p, t = naca4_digits(2412)
    body = naca4_boundary(M, m, p, t, size=0.5, x_offset=0.0, expo=1.4, spacing=1)
    outer = circle_boundary(M, 10.0)
    fs = ausm.freestream(0.7, 2.0)

    def configured(iters):
        s = ausm.Solver(M, N)
        s.update_grid_boundaries(body, outer)
        s.prepare_grid(iters=iters)
        s.set_freestream(*fs)
        s.set_scheme(SLAU2, VANALBADA, 0.0, 0.75)
        s.set_cfl(0.5, 1.0, 1)
        s.set_precision("fp64")
        s.prepare_state(freestream=True)
        return s

    a = configured(40)   # even sweeps: the grid ends in the first pair
    a.step(64)           # two captured 32-step chunks
    a.prepare_grid(iters=41)  # odd: the second pair
    a.prepare_state(freestream=True)
    a.step(64)
    b = configured(41)
    b.step(64)
    assert np.array_equal(a.vertices(fp64=True), b.vertices(fp64=True))
    assert np.array_equal(a.state64(), b.state64())
    a.close()
    b.close()
