"""The plain-method configuration at full size on the GPU, against the oracle.

NACA 0012 (closed trailing edge, cosine spacing) O-grid of 256 x 128 cells,
circular far field at 20 chords, 2000 red-black SOR sweeps (omega 1.5) of
the FP64 elliptic grid solve; FP32 state from a uniform free stream at
Mach 0.8, AoA 1.25 deg, gamma 1.4; SLAU2 + MUSCL (minmod, and the reference's
van Albada), TVD-RK3, CFL 0.5, 1000 steps.

Bar: the exact-math GPU run is bit-for-bit the ref32 run of the same inputs
(the restated WGSL), and the FP32 results (Cl, Cd, the state) stay within
stated bounds of the all-FP64 run of the same 1000 steps on the same grid.
"""
import numpy as np
import pytest

import ausm
from oracle_lib import AUSMUP, MINMOD, SLAU2, VANALBADA, Solver, make_params

pytestmark = pytest.mark.gpu

M, N, STEPS = 256, 128, 1000


def setup_solver(limiter):
    s = ausm.Solver(M, N)
    s.update_grid_boundaries(ausm.naca4_boundary(M, code=12, size=0.5, spacing="cosine"),
                             ausm.circle_boundary(M, 20.0))
    s.prepare_grid(iters=2000, method="sor", omega=1.5)
    s.set_flow(0.8, 1.25)
    s.set_scheme(SLAU2, limiter, 0.0, 0.75)
    s.set_cfl(0.5, 1.0, 1)
    return s


@pytest.mark.parametrize("limiter", [MINMOD, VANALBADA])
def test_naca0012_256x128_1000_steps(limiter):
    s = setup_solver(limiter)
    vtx = s.vertices()
    s.prepare_state(freestream=True)
    info = s.step(STEPS)
    q = s.state()
    f32 = s.forces(ref_length=1.0)
    # the restated WGSL on the same FP32 grid and free stream
    ref = Solver(32, make_params(0.8, 1.25, cfl=0.5, maxdt=1.0), SLAU2, limiter, vtx, omp=True)
    ref.step(STEPS)
    assert np.array_equal(q[1:N + 1], ref.state[1:N + 1])
    assert info["dt"] == ref.last_dt
    # the same 1000 steps in FP64 (bit-for-bit ref64 on small grids)
    s.set_precision("fp64")
    s.prepare_state(freestream=True)
    s.step(STEPS)
    q64 = s.state()
    f64 = s.forces(ref_length=1.0)
    rel = np.abs(q[1:N + 1] - q64[1:N + 1]).max(axis=(0, 1)) / np.abs(q64[1:N + 1]).max(axis=(0, 1))
    print(f"limiter {limiter}: Cl32 {f32['cl']:.7f} Cl64 {f64['cl']:.7f} "
          f"dCl {f32['cl'] - f64['cl']:+.2e} dCd {f32['cd'] - f64['cd']:+.2e} "
          f"state max-rel {rel.max():.2e}")
    assert abs(f32["cl"] - f64["cl"]) < 1e-4
    assert abs(f32["cd"] - f64["cd"]) < 1e-4
    assert rel.max() < 1e-3
    s.close()


def test_literal_510x381_supersonic():
    """The reference README's literal grid size, 510 x 381 (M not a multiple
    of the 60-column wavefront strip, N odd, so partial strips and partial
    row blocks are both exercised at full size): NACA 2412 in the
    rectangular far field, 2000 Jacobi sweeps of the FP64 grid solve, Mach
    2.0, AUSM+-up + van Albada, CFL 0.5, 40 steps from the free stream.
    Exact math is bit-for-bit ref32; fast math stays within max-rel 1e-4
    per conserved component of it, the state bound of the other fast-math
    parity tests."""
    M, N, steps = 510, 381, 40
    s = ausm.Solver(M, N)
    s.update_grid_boundaries(ausm.naca4_boundary(M, code=2412, size=0.3, x_offset=-0.7, exponent=1.4),
                             ausm.rect_boundary(M, 1.5))
    s.prepare_grid(iters=2000)
    s.set_flow(2.0, 0.0)
    s.set_scheme(AUSMUP, VANALBADA, 0.0, 0.75)
    s.set_cfl(0.5, 1.0, 1)
    vtx = s.vertices()
    s.prepare_state(freestream=True)
    info = s.step(steps)
    q = s.state()
    ref = Solver(32, make_params(2.0, 0.0, cfl=0.5, maxdt=1.0), AUSMUP, VANALBADA, vtx, omp=True)
    ref.step(steps)
    assert np.array_equal(q[1:N + 1], ref.state[1:N + 1])
    assert info["dt"] == ref.last_dt
    s.set_math("fast")
    s.prepare_state(freestream=True)
    s.step(steps)
    qf = s.state().astype(np.float64)[1:N + 1]
    r = ref.state.astype(np.float64)[1:N + 1]
    rel = (np.abs(qf - r).max(axis=(0, 1)) / np.abs(r).max(axis=(0, 1))).max()
    print(f"510x381 fast vs ref32 after {steps} steps: max-rel {rel:.2e}")
    assert rel < 1e-4
    s.close()


def test_headline_fast_math_tracks_exact_over_2000_steps():
    """The single-grid headline configuration (512 x 384 NACA 2412, Mach 2,
    AUSM+-up + van Albada, CFL 0.5) run 2000 steps from the free stream with
    exact and with fast math: the bow-shock solution agrees in Cl and Cd to
    1e-4 and in the state to max-rel 1e-3 per conserved component (the
    bound of the FP32 vs FP64 comparison above, which is how far any FP32
    rounding choice moves this solution)."""
    M, N, steps = 512, 384, 2000
    s = ausm.Solver(M, N)
    s.update_grid_boundaries(ausm.naca4_boundary(M, code=2412, size=0.3, x_offset=-0.7, exponent=1.4),
                             ausm.rect_boundary(M, 1.5))
    s.prepare_grid(iters=2000)
    s.set_flow(2.0, 0.0)
    s.set_scheme(AUSMUP, VANALBADA, 0.0, 0.75)
    s.set_cfl(0.5, 1.0, 1)
    out = {}
    for mode in ("exact", "fast"):
        s.set_math(mode)
        s.prepare_state(freestream=True)
        info = s.step(steps)
        assert not info["nonfinite"]
        out[mode] = (s.state().astype(np.float64)[1:N + 1], s.forces(ref_length=1.0))
    (qe, fe), (qf, ff) = out["exact"], out["fast"]
    rel = (np.abs(qf - qe).max(axis=(0, 1)) / np.abs(qe).max(axis=(0, 1))).max()
    print(f"headline 2000 steps: Cl {fe['cl']:.6f} / {ff['cl']:.6f}, Cd {fe['cd']:.6f} / {ff['cd']:.6f}, "
          f"state max-rel {rel:.2e}")
    assert abs(ff["cl"] - fe["cl"]) < 1e-4
    assert abs(ff["cd"] - fe["cd"]) < 1e-4
    assert rel < 1e-3
    s.close()
