"""GPU parity of the compensated-state FP32 precision (AUSM_PREC_FP32C)
against the oracle's compensated stages (run_steps_comp).

Bar: exact math bit-for-bit equal to the oracle for hi and for lo (read back
as hi + lo in FP64) after any number of steps, with the global and the local
time step and every tile shape; fast math within the fast-math tolerance of
the per-step residual norm (relative 1e-5); precision switches carry the
solution over as documented (FP32 <-> FP32C keep hi, FP32C <-> FP64 through
hi + lo, which is exact in FP64).
"""
import numpy as np
import pytest

import ausm
from oracle_lib import AUSMUP, SLAU, SLAU2, VANALBADA, Oracle, Solver
from test_gpu_parity import (FAST_TOL_NORM, make_grid, max_rel, new_pair, perturbed_state,
                             rho_norm)

pytestmark = pytest.mark.gpu

SCHEMES = [SLAU2, SLAU, AUSMUP]


def comp_pair(vtx, mach, aoa, scheme, math_mode="exact", local=False, tile=0):
    s, _, par = new_pair(vtx, mach, aoa, scheme, VANALBADA, math_mode=math_mode, tile=tile)
    s.set_dt_mode("local" if local else "global")
    s.set_precision("fp32c")
    ref = Solver(32, par, scheme, VANALBADA, vtx, local_dt=local, comp=True)
    return s, ref, par


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("local", [False, True])
def test_comp_exact_bitwise(scheme, local):
    M, N = 96, 40
    vtx, _, _ = make_grid(M, N, code=2412, iters=150)
    s, ref, par = comp_pair(vtx, 0.73, 2.5, scheme, local=local)
    st = perturbed_state(par, M, N, seed=41 + scheme)
    s.set_state(st)
    ref.set_state(st)
    for chunk in (1, 5, 40):  # 40 steps crosses one captured 32-step graph
        info = s.step(chunk)
        ref.step(chunk)
        hi = s.state()
        assert np.array_equal(hi[1:N + 1], ref.state[1:N + 1]), (chunk, max_rel(hi, ref.state))
        both = s.state64()[1:N + 1]
        expect = ref.state[1:N + 1].astype(np.float64) + ref.lo[1:N + 1].astype(np.float64)
        assert np.array_equal(both, expect), chunk
        assert info["dt"] == ref.last_dt
    assert np.abs(ref.lo[1:N + 1]).max() > 0  # the low part is in use
    s.close()


@pytest.mark.parametrize("math_mode", ["exact", "fast"])
def test_comp_tiles_identical(math_mode):
    M, N = 130, 37
    vtx, _, _ = make_grid(M, N, code=4412, iters=100)
    out = []
    for tile in (1, 2, 4):
        s, _, par = comp_pair(vtx, 0.8, 1.25, SLAU2, math_mode=math_mode, local=True, tile=tile)
        s.set_state(perturbed_state(par, M, N, seed=2))
        s.step(45)
        out.append(s.state64())
        s.close()
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


def test_comp_fast_within_tolerance():
    M, N = 128, 48
    vtx, _, _ = make_grid(M, N, code=2412, iters=150)
    s, ref, par = comp_pair(vtx, 0.8, 1.25, SLAU2, math_mode="fast", local=True)
    st = perturbed_state(par, M, N, seed=4)
    s.set_state(st)
    ref.set_state(st)
    n_ref = rho_norm(Oracle(32).residual(par, SLAU2, VANALBADA, vtx, st))
    for k in range(30):
        info = s.step(1)
        ref.step(1)
        assert abs(info["res_l2"] - n_ref) <= FAST_TOL_NORM * n_ref, (k, info["res_l2"], n_ref)
        n_ref = rho_norm(Oracle(32).residual(par, SLAU2, VANALBADA, vtx, ref.state))
    assert max_rel(s.state()[1:N + 1], ref.state[1:N + 1]) < 1e-4
    s.close()


def test_comp_precision_switches():
    M, N = 64, 24
    vtx, _, _ = make_grid(M, N, code=12, iters=80)
    s, _, par = new_pair(vtx, 0.8, 1.25, SLAU2, VANALBADA)
    s.set_dt_mode("local")
    s.prepare_state(freestream=True)
    s.step(10)
    q32 = s.state()
    s.set_precision("fp32c")  # hi = the FP32 state, lo = 0
    assert np.array_equal(s.state()[1:N + 1], q32[1:N + 1])
    assert np.array_equal(s.state64()[1:N + 1], q32[1:N + 1].astype(np.float64))
    s.step(20)
    qc = s.state64()
    s.set_precision("fp64")  # hi + lo, exact in FP64
    assert np.array_equal(s.state()[1:N + 1], qc[1:N + 1])
    s.step(3)
    q64 = s.state()
    s.set_precision("fp32c")  # split: hi = fl(q), lo = fl(q - hi)
    hi = q64[1:N + 1].astype(np.float32)
    lo = (q64[1:N + 1] - hi.astype(np.float64)).astype(np.float32)
    assert np.array_equal(s.state()[1:N + 1], hi)
    assert np.array_equal(s.state64()[1:N + 1], hi.astype(np.float64) + lo.astype(np.float64))
    s.set_precision("fp32")
    assert np.array_equal(s.state()[1:N + 1], hi)
    # a state set in FP32C starts with lo = 0; set_state64 splits
    s.set_precision("fp32c")
    s.set_state(q32)
    assert np.array_equal(s.state64()[1:N + 1], q32[1:N + 1].astype(np.float64))
    full = q64.copy()
    s.set_state64(full)
    assert np.array_equal(s.state64()[1:N + 1], hi.astype(np.float64) + lo.astype(np.float64))
    s.close()
