"""The allocation-free step (ausm_step_noalloc) against ausm_step.

ausm_step_noalloc steps a caller-owned solution held in device memory, in
the reference's layouts, using only the library's fixed scratch array.  Two
calls in a row, the second on a smaller grid (so it reuses scratch the first
call left behind), must each give bit-for-bit the state, dt and residual
norm that ausm_step gives on a context set up with the same vertices, state
and settings -- in exact math (where both are also bit-for-bit the FP32
oracle) and in fast math, with global and local dt and a dt refresh
cadence.  A grid beyond the scratch array's size is refused before any GPU
work.
"""
import ctypes
import math

import numpy as np
import pytest

import ausm
from ausm import lib as L
from oracle_lib import Oracle, Solver, circle_boundary, make_params, naca4_boundary, naca4_digits


class Hip:
    """Device buffers for the caller's side of the call (the test is the
    caller: it allocates, the entry point does not)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        return p, a.shape

    def download(self, p, shape):
        out = np.zeros(shape, dtype=np.float32)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0
        return out

    def free(self, p):
        assert self.L.hipFree(p) == 0


def grid(M, N, code):
    m, p, t = naca4_digits(code)
    body = naca4_boundary(M, m, p, t, size=0.5, x_offset=0.0, expo=1.4, spacing=1)
    return Oracle(64).grid(body, circle_boundary(M, 10.0), N, 120).astype(np.float32)


def perturbed(par, M, N, seed):
    rng = np.random.default_rng(seed)
    st = Oracle(32).init_state(par, M, N).astype(np.float64)
    st[1:N + 1] *= 1.0 + 0.01 * rng.standard_normal((N, M, 4))
    return st.astype(np.float32)


def desc(M, N, fs, scheme, limiter, math_mode, dt_mode, refresh, cfl=0.5, maxdt=1.0):
    d = L.StepDesc()
    d.M, d.N = M, N
    d.rho, d.p, d.vx, d.vy, d.gamma = fs
    d.scheme, d.limiter = scheme, limiter
    d.K_p, d.K_u = 0.0, 0.75
    d.cfl, d.maxdt = cfl, maxdt
    d.steps_per_dt_refresh = refresh
    d.math = math_mode
    d.dt_mode = dt_mode
    d.rows_per_wave = 2
    return d


def context_run(vtx, st, fs, scheme, limiter, math_mode, dt_mode, refresh, steps):
    N1, M, _ = vtx.shape
    s = ausm.Solver(M, N1 - 1)
    s.set_vertices(vtx)
    s.set_freestream(*fs)
    s.set_scheme(scheme, limiter, 0.0, 0.75)
    s.set_cfl(0.5, 1.0, refresh)
    s.set_math(math_mode)
    s.set_dt_mode(dt_mode)
    s.set_tile(2)
    s.set_state(st)
    info = s.step(steps)
    q = s.state()
    s.close()
    return q, info


CASES = [  # (M, N, NACA, steps, scheme, limiter, dt mode, dt refresh cadence)
    (96, 40, 2412, 30, ausm.AUSMUP, ausm.VANALBADA, L.DT_GLOBAL, 1),
    (64, 24, 12, 17, ausm.SLAU2, ausm.MINMOD, L.DT_LOCAL, 3),
]


@pytest.mark.gpu
@pytest.mark.parametrize("math_mode", [L.MATH_EXACT, L.MATH_FAST])
def test_noalloc_step_matches_context_step_twice(math_mode):
    hip = Hip()
    lib = ausm.load()
    for k, (M, N, code, steps, scheme, limiter, dt_mode, refresh) in enumerate(CASES):
        print(f"[noalloc] call {k}: {M}x{N} {steps} steps math={math_mode}", flush=True)
        vtx = grid(M, N, code)
        mach, aoa = (2.0, 0.0) if k == 0 else (0.7, 2.5)
        fs = ausm.freestream(mach, aoa)
        st = perturbed(make_params(mach, aoa), M, N, seed=k)
        q_ref, info_ref = context_run(vtx, st, fs, scheme, limiter, math_mode, dt_mode, refresh,
                                      steps)
        dv, _ = hip.upload(vtx)
        dq, shape = hip.upload(st)
        d = desc(M, N, fs, scheme, limiter, math_mode, dt_mode, refresh)
        info = L.StepInfo()
        L.check(lib.ausm_step_noalloc(0, ctypes.byref(d), dv, dq, steps, None,
                                      ctypes.byref(info)))
        q = hip.download(dq, shape)
        hip.free(dv)
        hip.free(dq)
        assert np.array_equal(q, q_ref), (k, float(np.abs(q - q_ref).max()))
        assert info.dt == info_ref["dt"] and info.res_l2 == info_ref["res_l2"]
        assert info.max_wave_speed == info_ref["max_wave_speed"] and not info.nonfinite
        if math_mode == L.MATH_EXACT and dt_mode == L.DT_GLOBAL:
            par = make_params(mach, aoa, cfl=0.5, maxdt=1.0)
            ref = Solver(32, par, scheme, limiter, vtx, refresh=refresh)
            ref.set_state(st)
            ref.step(steps)
            assert np.array_equal(q[1:N + 1], ref.state[1:N + 1])


FIRST_CALL = r"""
import ctypes, os, sys
sys.path.insert(0, os.environ["AUSM_TEST_DIR"])
sys.path.insert(0, os.path.join(os.path.dirname(os.environ["AUSM_TEST_DIR"]), "ausm-webgpu_amd"))
import numpy as np
import ausm
from ausm import lib as L
from oracle_lib import Solver
from test_zz_no_alloc import Hip, desc, grid, perturbed, make_params
hip, lib = Hip(), ausm.load()
M, N, steps = 64, 24, 9
vtx = grid(M, N, 2412)
fs = ausm.freestream(2.0, 0.0)
st = perturbed(make_params(2.0, 0.0), M, N, seed=3)
dv, _ = hip.upload(vtx)
dq, shape = hip.upload(st)
# a fast-math call first (no check needed): the library's kernels and its
# scratch address are set up outside any capture
dqf, _ = hip.upload(st)
df = desc(M, N, fs, ausm.AUSMUP, ausm.VANALBADA, L.MATH_FAST, L.DT_GLOBAL, 1)
L.check(lib.ausm_step_noalloc(0, ctypes.byref(df), dv, dqf, steps, None, None))
assert hip.L.hipDeviceSynchronize() == 0
hip.free(dqf)
d = desc(M, N, fs, ausm.AUSMUP, ausm.VANALBADA, L.MATH_EXACT, L.DT_GLOBAL, 1)
# no context, batch or check has run in this process: under stream capture
# the first exact call is refused before it enqueues anything
hip.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
hip.L.hipStreamBeginCapture.argtypes = [ctypes.c_void_p, ctypes.c_int]
hip.L.hipStreamEndCapture.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
hip.L.hipGraphDestroy.argtypes = [ctypes.c_void_p]
hip.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
stream, graph = ctypes.c_void_p(), ctypes.c_void_p()
assert hip.L.hipStreamCreate(ctypes.byref(stream)) == 0
assert hip.L.hipStreamBeginCapture(stream, 2) == 0  # relaxed mode
rc = lib.ausm_step_noalloc(0, ctypes.byref(d), dv, dq, steps, stream, None)
msg = lib.ausm_last_error().decode()
assert hip.L.hipStreamEndCapture(stream, ctypes.byref(graph)) == 0
if graph.value:
    assert hip.L.hipGraphDestroy(graph) == 0
assert hip.L.hipStreamDestroy(stream) == 0
assert rc == -4 and "ausm_check_arithmetic" in msg, (rc, msg)
# outside capture the first exact call runs the check itself and steps
info = L.StepInfo()
L.check(lib.ausm_step_noalloc(0, ctypes.byref(d), dv, dq, steps, None, ctypes.byref(info)))
q = hip.download(dq, shape)
ref = Solver(32, make_params(2.0, 0.0, cfl=0.5, maxdt=1.0), ausm.AUSMUP, ausm.VANALBADA, vtx)
ref.set_state(st)
ref.step(steps)
assert np.array_equal(q[1:N + 1], ref.state[1:N + 1])
mism = ctypes.c_longlong(-1)
L.check(lib.ausm_check_arithmetic(0, ctypes.byref(mism)))
assert mism.value == 0
print("first-call check ok", flush=True)
"""


@pytest.mark.gpu
def test_noalloc_first_exact_call_runs_the_device_check():
    """In a process where no context or batch has run the arithmetic
    self-check, the first exact-math allocation-free call runs it (with its
    counter in the scratch array) and then steps bit-for-bit the oracle;
    under stream capture it is refused instead, before enqueuing anything.
    Needs a fresh process, so the case runs in a child."""
    import os
    import subprocess
    import sys
    env = dict(os.environ, AUSM_TEST_DIR=os.path.dirname(os.path.abspath(__file__)))
    print("[noalloc] first exact call in a child process", flush=True)
    r = subprocess.run([sys.executable, "-c", FIRST_CALL], env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "first-call check ok" in r.stdout


def test_noalloc_step_refuses_grids_beyond_its_scratch():
    """Checked before any GPU work: runs without a GPU."""
    lib = ausm.load()
    fs = ausm.freestream(0.8, 0.0)
    d = desc(4096, 2048, fs, ausm.SLAU2, ausm.VANALBADA, L.MATH_FAST, L.DT_GLOBAL, 1)
    fake = ctypes.c_void_p(256)
    rc = lib.ausm_step_noalloc(0, ctypes.byref(d), fake, fake, 1, None, None)
    assert rc == -1
    msg = lib.ausm_last_error().decode()
    assert "scratch" in msg and str(64 << 20) in msg, msg
    # the largest grid of the workloads, 1024 x 768, fits: the call gets as
    # far as the device check
    d = desc(1024, 768, fs, ausm.SLAU2, ausm.VANALBADA, L.MATH_FAST, L.DT_GLOBAL, 1)
    rc = lib.ausm_step_noalloc(-1, ctypes.byref(d), fake, fake, 1, None, None)
    assert rc == -1 and "device -1" in lib.ausm_last_error().decode()
    assert math.isfinite(fs[0])
