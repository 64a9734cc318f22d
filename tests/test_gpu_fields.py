"""GPU parity of the derived display fields (ausm_derived_field) against the
oracle's restatement of the visualization pass: bitwise for the fields made
of division and square root (vorticity, density, temperature, pressure,
Mach, |v - V_inf|), within relative 1e-5 for those through exp / log / pow
(schlieren, entropy, total pressure: device and host libm differ by ulps);
in FP32 against ref32 and in FP64 against ref64 (rounded to FP32)."""
import numpy as np
import pytest

from oracle_lib import SLAU2, VANALBADA, Oracle
from test_gpu_fp64 import grid64, new_fp64, perturbed_state64
from test_gpu_parity import make_grid, new_pair, perturbed_state

pytestmark = pytest.mark.gpu

EXACT = [1, 2, 3, 4, 5, 6, 7]
LIBM = [0, 8, 9]


def check(got, fs, ref, fs_ref, mode):
    ref32 = np.asarray(ref, dtype=np.float64).astype(np.float32)
    if mode in EXACT:
        assert np.array_equal(got, ref32), (mode, np.abs(got - ref32).max())
        assert fs == np.float32(fs_ref)
    else:
        assert np.allclose(got, ref32, rtol=1e-5, atol=1e-6), (mode, np.abs(got - ref32).max())
        assert abs(fs - np.float32(fs_ref)) <= 1e-5 * max(1.0, abs(fs_ref))


@pytest.mark.parametrize("mode", EXACT + LIBM)
def test_fields_fp32(mode):
    M, N = 96, 40
    vtx, _, _ = make_grid(M, N, code=2412, iters=150)
    s, _, par = new_pair(vtx, 0.8, 3.0, SLAU2, VANALBADA)
    st = perturbed_state(par, M, N, seed=50 + mode, amp=0.05)
    s.set_state(st)
    got, fs = s.derived_field(mode)
    ref, fs_ref = Oracle(32).derived_field(par, s.state(), mode)
    check(got, fs, ref, fs_ref, mode)
    s.close()


@pytest.mark.parametrize("mode", EXACT + LIBM)
def test_fields_fp64(mode):
    M, N = 64, 24
    vtx = grid64(M, N)
    s, par = new_fp64(vtx, 1.5, -2.0, SLAU2, VANALBADA)
    st = perturbed_state64(par, M, N, seed=60 + mode, amp=0.05)
    s.set_state(st)
    got, fs = s.derived_field(mode)
    ref, fs_ref = Oracle(64).derived_field(par, s.state(), mode)
    check(got, fs, ref, fs_ref, mode)
    s.close()


def test_fields_of_the_free_stream():
    M, N = 64, 24
    vtx, _, _ = make_grid(M, N, code=12, iters=80)
    s, _, par = new_pair(vtx, 0.6, 2.0, SLAU2, VANALBADA)
    s.prepare_state(freestream=True)
    rho, fs = s.derived_field("density")
    assert np.all(rho == np.float32(par[1])) and fs == np.float32(par[1])
    w, _ = s.derived_field("vorticity")
    assert np.abs(w).max() < 1e-6
    sch, _ = s.derived_field("schlieren")
    assert np.allclose(sch, 1.0)
    mach, minf = s.derived_field("mach_inflow")
    assert np.allclose(mach, minf, rtol=1e-6) and abs(minf - 0.6) < 1e-6
    s.close()
