"""Grid shapes at the edges of what the stage kernel tiles: the smallest
grids the library accepts (4 x 4), strips narrower than and just past one
60-column wavefront strip, row counts below and across a workgroup's rows,
the largest ring (32768 x 4) and the tallest one (4 x 32768), and a large
grid (2048 x 1536).

Bar: exact math bit-for-bit equal to ref32 (the restated WGSL) for every
tile shape; fast math identical across tile shapes and within max-rel 1e-4
of ref32 per conserved component.  The grids are annuli (circle body,
circular far field, geometric radial spacing), built directly so that any
(M, N) is a valid O-grid.
"""
import numpy as np
import pytest

import ausm
from oracle_lib import AUSMUP, SLAU2, VANALBADA, Solver, make_params64
from test_gpu_fp64 import perturbed_state64
from test_gpu_parity import flow, max_rel, perturbed_state

pytestmark = pytest.mark.gpu

TILES = [1, 2, 4, 8]


def annulus(M, N, r0=0.5, r1=3.0, dtype=np.float32):
    """(N+1, M, 2) vertices: body row j = 0 counter-clockwise, far field at
    j = N."""
    th = 2.0 * np.pi * np.arange(M) / M
    r = r0 * (r1 / r0) ** (np.arange(N + 1) / N)
    x = r[:, None] * np.cos(th)[None, :]
    y = r[:, None] * np.sin(th)[None, :]
    return np.stack([x, y], axis=-1).astype(dtype)


def run(vtx, scheme, math_mode, tile, st, steps, fs, par):
    N1, M, _ = vtx.shape
    s = ausm.Solver(M, N1 - 1)
    s.set_vertices(vtx)
    s.set_freestream(*fs)
    s.set_scheme(scheme, VANALBADA, float(par[11]), float(par[12]))
    s.set_cfl(float(par[10]), float(par[9]), 1)
    s.set_math(math_mode)
    s.set_tile(tile)
    s.set_state(st)
    info = s.step(steps)
    q = s.state()
    s.close()
    return q, info


@pytest.mark.parametrize("M,N", [(4, 4), (5, 4), (59, 4), (60, 6), (61, 5), (64, 9), (121, 7),
                                 (125, 17)])
def test_small_and_ragged_shapes(M, N):
    vtx = annulus(M, N)
    fs, par = flow(0.6, 5.0)
    st = perturbed_state(par, M, N, seed=M * 31 + N)
    ref = Solver(32, par, SLAU2, VANALBADA, vtx)
    ref.set_state(st)
    ref.step(8)
    fast = []
    for tile in TILES:
        q, info = run(vtx, SLAU2, "exact", tile, st, 8, fs, par)
        assert np.array_equal(q[1:N + 1], ref.state[1:N + 1]), (M, N, tile, max_rel(q, ref.state))
        assert info["dt"] == ref.last_dt
        qf, _ = run(vtx, SLAU2, "fast", tile, st, 8, fs, par)
        fast.append(qf)
        assert max_rel(qf[1:N + 1], ref.state[1:N + 1]) < 1e-4, (M, N, tile)
    for qf in fast[1:]:
        assert np.array_equal(qf, fast[0])


@pytest.mark.parametrize("M,N", [(32768, 4), (4, 32768)])
def test_longest_ring_and_tallest_grid(M, N):
    vtx = annulus(M, N, r1=1.5 if N < 100 else 50.0)
    fs, par = flow(0.8, 2.0)
    st = perturbed_state(par, M, N, seed=5)
    ref = Solver(32, par, AUSMUP, VANALBADA, vtx, omp=True)
    ref.set_state(st)
    ref.step(3)
    for tile in (0, 1):
        q, info = run(vtx, AUSMUP, "exact", tile, st, 3, fs, par)
        assert np.array_equal(q[1:N + 1], ref.state[1:N + 1]), (M, N, tile, max_rel(q, ref.state))
        assert info["dt"] == ref.last_dt


def test_large_grid_exact_bitwise_and_fast_tiles():
    M, N = 2048, 1536
    vtx = annulus(M, N, r1=20.0)
    fs, par = flow(2.0, 0.0)
    st = perturbed_state(par, M, N, seed=17)
    ref = Solver(32, par, AUSMUP, VANALBADA, vtx, omp=True)
    ref.set_state(st)
    ref.step(4)
    q, info = run(vtx, AUSMUP, "exact", 0, st, 4, fs, par)
    assert np.array_equal(q[1:N + 1], ref.state[1:N + 1]), max_rel(q, ref.state)
    assert info["dt"] == ref.last_dt
    q2, _ = run(vtx, AUSMUP, "fast", 2, st, 4, fs, par)
    q4, _ = run(vtx, AUSMUP, "fast", 4, st, 4, fs, par)
    assert np.array_equal(q2, q4)
    assert max_rel(q2[1:N + 1], ref.state[1:N + 1]) < 1e-4


@pytest.mark.parametrize("M,N", [(4, 4), (61, 5), (125, 17)])
def test_shapes_fp64_and_compensated(M, N):
    """FP64 bit-for-bit ref64 and FP32C (exact math) bit-for-bit the oracle's
    compensated mode, hi and lo, for each tile shape of those kernels."""
    fs = ausm.freestream(0.6, 5.0)
    vtx64 = annulus(M, N, dtype=np.float64)
    par64 = make_params64(fs, cfl=0.5, maxdt=1.0)
    st64 = perturbed_state64(par64, M, N, seed=M + N)
    ref64 = Solver(64, par64, SLAU2, VANALBADA, vtx64)
    ref64.set_state(st64)
    ref64.step(6)
    vtx = annulus(M, N)
    _, par = flow(0.6, 5.0)
    st = perturbed_state(par, M, N, seed=M + N)
    refc = Solver(32, par, SLAU2, VANALBADA, vtx, comp=True)
    refc.set_state(st)
    refc.step(6)
    for tile in (1, 2, 4):
        s = ausm.Solver(M, N)
        s.set_vertices(vtx64)
        s.set_freestream(*fs)
        s.set_scheme(SLAU2, VANALBADA, 0.0, 0.75)
        s.set_cfl(0.5, 1.0, 1)
        s.set_tile(tile)
        s.set_precision("fp64")
        s.set_state64(st64)
        s.step(6)
        q = s.state64()
        assert np.array_equal(q[1:N + 1], ref64.state[1:N + 1]), (M, N, tile)
        s.close()
        s = ausm.Solver(M, N)
        s.set_vertices(vtx)
        s.set_freestream(*fs)
        s.set_scheme(SLAU2, VANALBADA, float(par[11]), float(par[12]))
        s.set_cfl(float(par[10]), float(par[9]), 1)
        s.set_math("exact")
        s.set_tile(tile)
        s.set_precision("fp32c")
        s.set_state(st)
        s.step(6)
        assert np.array_equal(s.state()[1:N + 1], refc.state[1:N + 1]), (M, N, tile)
        both = s.state64()[1:N + 1]
        expect = refc.state[1:N + 1].astype(np.float64) + refc.lo[1:N + 1].astype(np.float64)
        assert np.array_equal(both, expect), (M, N, tile)
        s.close()
