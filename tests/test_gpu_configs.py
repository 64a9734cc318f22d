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
from oracle_lib import MINMOD, SLAU2, VANALBADA, Solver, make_params

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
