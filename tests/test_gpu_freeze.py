"""GPU parity of limiter freezing (ausm_freeze_limiter, the batch's
freeze_after) against the oracle's freezing (set_freeze).

Bar: exact math bit-for-bit equal to the oracle -- the state after frozen
steps (through captured graphs), the frozen residual, the FP64 solve against
ref64 and the compensated FP32 state (hi and lo) -- for both van Albada
limiters, on a grid with partial wavefront strips and row blocks; fast math
within the fast-math state bound of exact math; a frozen batched solve
bit-for-bit its one-by-one run; freezing converges a stalled MUSCL solve.
"""
import math

import numpy as np
import pytest

import ausm
from oracle_lib import (AUSMUP, MINMOD, SLAU2, VANALBADA, VANALBADA_CLAMPED, Oracle, Solver,
                        make_params64)
from test_gpu_parity import FAST_TOL_FIELD, make_grid, max_rel, new_pair, perturbed_state

pytestmark = pytest.mark.gpu

M, N = 190, 75  # 3 full 60-column strips and one of 10; odd N
LIMS = [VANALBADA, VANALBADA_CLAMPED]


def _grid():
    vtx, _, _ = make_grid(M, N, code=2412, iters=150)
    return vtx


@pytest.mark.parametrize("limiter", LIMS)
@pytest.mark.parametrize("scheme", [SLAU2, AUSMUP])
def test_frozen_steps_exact_bitwise(scheme, limiter):
    vtx = _grid()
    mach = 0.75 if scheme == SLAU2 else 1.6
    s, ref, par = new_pair(vtx, mach, 2.0, scheme, limiter)
    st = perturbed_state(par, M, N, seed=3 + scheme + limiter)
    s.set_state(st)
    ref.set_state(st)
    s.step(30)
    ref.step(30)
    assert np.array_equal(s.state()[1:N + 1], ref.state[1:N + 1])
    s.freeze_limiter()
    ref.freeze_limiter()
    assert s.limiter_frozen
    print(f"[freeze] scheme {scheme} limiter {limiter}: recorded after 30 steps", flush=True)
    for chunk in (1, 40):  # 40 steps run through captured graphs
        info = s.step(chunk)
        ref.step(chunk)
        got = s.state()
        assert np.array_equal(got[1:N + 1], ref.state[1:N + 1]), (chunk, max_rel(got, ref.state))
        assert info["dt"] == ref.last_dt
    assert np.array_equal(s.residual(), ref.residual())
    # the frozen operator is not the limited one any more
    limited = Oracle(32).residual(par, scheme, limiter, vtx, ref.state)
    assert not np.array_equal(s.residual(), limited)
    s.close()


def test_frozen_fp64_bitwise():
    vtx = _grid().astype(np.float64)
    fs = ausm.freestream(0.73, 2.5)
    par = make_params64(fs, cfl=0.5, maxdt=1.0)
    s = ausm.Solver(M, N)
    s.set_vertices(vtx.astype(np.float32))
    s.set_freestream(*fs)
    s.set_scheme(SLAU2, VANALBADA, 0.0, 0.75)
    s.set_cfl(0.5, 1.0, 1)
    s.set_precision("fp64")
    s.prepare_state(True)
    ref = Solver(64, par, SLAU2, VANALBADA, vtx)
    s.step(25)
    ref.step(25)
    s.freeze_limiter()
    ref.freeze_limiter()
    s.step(40)
    ref.step(40)
    assert np.array_equal(s.state()[1:N + 1], ref.state[1:N + 1])
    assert np.array_equal(s.residual(), ref.residual())
    s.close()


@pytest.mark.parametrize("local", [False, True])
def test_frozen_compensated_bitwise(local):
    vtx = _grid()
    s, _, par = new_pair(vtx, 0.8, 1.25, SLAU2, VANALBADA_CLAMPED)
    s.set_dt_mode("local" if local else "global")
    s.set_precision("fp32c")
    ref = Solver(32, par, SLAU2, VANALBADA_CLAMPED, vtx, local_dt=local, comp=True)
    st = perturbed_state(par, M, N, seed=9)
    s.set_state(st)
    ref.set_state(st)
    s.step(20)
    ref.step(20)
    s.freeze_limiter()
    ref.freeze_limiter()
    s.step(40)
    ref.step(40)
    assert np.array_equal(s.state()[1:N + 1], ref.state[1:N + 1])
    expect = ref.state[1:N + 1].astype(np.float64) + ref.lo[1:N + 1].astype(np.float64)
    assert np.array_equal(s.state64()[1:N + 1], expect)
    s.close()


def test_frozen_fast_within_the_state_bound():
    vtx = _grid()
    out = {}
    for mode in ("exact", "fast"):
        s, _, par = new_pair(vtx, 0.75, 2.0, SLAU2, VANALBADA, math_mode=mode)
        s.set_state(perturbed_state(par, M, N, seed=21))
        s.step(30)
        s.freeze_limiter()
        s.step(40)
        out[mode] = s.state()[1:N + 1]
        s.close()
    rel = max_rel(out["fast"], out["exact"])
    print(f"[freeze] fast vs exact after frozen steps: max-rel {rel:.2e}", flush=True)
    assert rel < FAST_TOL_FIELD


def test_freeze_errors_and_unfreeze():
    vtx = _grid()
    s, _, par = new_pair(vtx, 0.75, 2.0, SLAU2, MINMOD)
    with pytest.raises(ausm.AusmError):
        s.freeze_limiter()  # no state yet
    s.prepare_state(True)
    with pytest.raises(ausm.AusmError):
        s.freeze_limiter()  # not a van Albada limiter
    s.set_scheme(SLAU2, VANALBADA, 0.0, 0.75)
    s.freeze_limiter()
    assert s.limiter_frozen
    s.freeze_limiter(False)
    assert not s.limiter_frozen
    s.freeze_limiter()
    s.set_state(s.state())
    assert not s.limiter_frozen  # a new state ends it
    s.freeze_limiter()
    s.set_scheme(SLAU2, VANALBADA_CLAMPED, 0.0, 0.75)
    assert not s.limiter_frozen  # so does a new limiter
    s.close()


@pytest.mark.parametrize("math_mode,precision,dt_mode", [("fast", "fp32", "global"),
                                                         ("exact", "fp32c", "local"),
                                                         ("fast", "fp32c", "local")])
def test_batch_freeze_after_matches_single_grid(math_mode, precision, dt_mode):
    cases = [(M, N, 0.75, 2.0, SLAU2, VANALBADA), (128, 48, 1.4, 0.0, AUSMUP, VANALBADA_CLAMPED),
             (96, 40, 0.5, 3.0, SLAU2, VANALBADA)]
    grids = {(m, n): make_grid(m, n, code=2412, iters=120)[0] for m, n, *_ in cases}
    steps, freeze_after, check = 150, 60, 30
    with ausm.Batch(max_solves=4, math=math_mode, precision=precision) as b:
        ids = [b.submit(grids[(m, n)], ausm.freestream(mach, aoa), scheme=sc, limiter=li,
                        max_steps=steps, dt_mode=dt_mode,
                        freeze_after=freeze_after if k != 2 else 0)
               for k, (m, n, mach, aoa, sc, li) in enumerate(cases)]
        b.advance(steps, check_every=check)
        got = [b.state(sid) for sid in ids]
    for k, (m, n, mach, aoa, sc, li) in enumerate(cases):
        s = ausm.Solver(m, n)
        s.set_vertices(grids[(m, n)])
        s.set_flow(mach, aoa)
        s.set_scheme(sc, li, 0.0, 0.75)
        s.set_cfl(0.5, 1.0, 1)
        s.set_math(math_mode)
        s.set_dt_mode(dt_mode)
        s.set_precision(precision)
        s.prepare_state(True)
        if k != 2:
            s.step(freeze_after, info=False)
            s.freeze_limiter()
            s.step(steps - freeze_after, info=False)
        else:
            s.step(steps, info=False)
        assert np.array_equal(s.state()[1:n + 1], got[k][1:n + 1]), (k, max_rel(s.state(), got[k]))
        s.close()


def test_batch_rejects_freezing_other_limiters():
    vtx = make_grid(96, 40, code=12, iters=60)[0]
    with ausm.Batch(max_solves=2) as b:
        with pytest.raises(ausm.AusmError):
            b.submit(vtx, ausm.freestream(0.5, 0.0), limiter=MINMOD, freeze_after=10)


def test_freezing_converges_a_stalled_transonic_solve():
    """256 x 128 CST section (t/c 0.11), 20-chord circular far field, Mach
    0.73, AoA 2.5 deg, SLAU2 + the reference's van Albada, compensated FP32
    state, local time step, fast math: the limited solve holds a limit cycle
    above a 1e-3 drop of the density-residual norm (about 2.5e-2); frozen
    after 12000 steps it falls below a 1e-5 drop within 40000 steps (the
    FP32 residual then levels off near 6e-6, the rounding floor of its
    evaluation at this size), and Cl and Cd of the frozen steady state lie
    within 2e-3 of the limited solve's mean over its last 4000 steps."""
    from ausm.precision import cst_coefficients
    Mg, Ng = 256, 128
    au, al = cst_coefficients(0, 0.11)
    body = ausm.cst_boundary(Mg, au, al, size=0.5)
    s = ausm.Solver(Mg, Ng)
    s.update_grid_boundaries(body, ausm.circle_boundary(Mg, 20.0))
    s.prepare_grid(iters=3000)
    s.set_flow(0.73, 2.5)
    s.set_scheme(SLAU2, VANALBADA, 0.0, 0.75)
    s.set_cfl(0.5, 1.0, 1)
    s.set_math("fast")
    s.set_dt_mode("local")
    s.set_precision("fp32c")
    s.prepare_state(True)
    r0 = s.residual_norm()
    cls, cds, drops = [], [], []
    for k in range(24):  # 12000 limited steps
        info = s.step(500)
        drops.append(info["res_l2"] / r0)
        if k >= 16:
            f = s.forces()
            cls.append(f["cl"])
            cds.append(f["cd"])
    print("[freeze] limited drops " + " ".join(f"{d:.1e}" for d in drops[5::3]), flush=True)
    assert min(drops[8:]) > 1e-3
    s.freeze_limiter()
    steps, drop, traj = 0, 1.0, []
    while steps < 40000 and drop > 1e-5:
        drop = s.step(1000)["res_l2"] / r0
        steps += 1000
        if steps % 10000 == 0:
            traj.append(drop)
    print("[freeze] frozen drops per 10000 steps " + " ".join(f"{d:.1e}" for d in traj), flush=True)
    f = s.forces()
    print(f"[freeze] frozen: drop {drop:.2e} after {steps} steps; Cl {f['cl']:.5f} vs limited mean "
          f"{np.mean(cls):.5f}, Cd {f['cd']:.5f} vs {np.mean(cds):.5f}", flush=True)
    assert drop <= 1e-5
    assert abs(f["cl"] - np.mean(cls)) < 2e-3 and abs(f["cd"] - np.mean(cds)) < 2e-3
    s.close()


def test_served_freezing_equals_one_by_one():
    """The serving loop freezes a request's limiter at the same check in the
    batched service and in the one-by-one baseline: 6 requests of 128 x 64
    cells (drawn as the serving mix, MUSCL van Albada), frozen after 300
    steps, 1000-step cap -- state, steps, Cl and Cd agree exactly,
    whatever each request's outcome."""
    from ausm import serve
    reqs = serve.make_requests(6, seed=3, sizes=[(128, 64)], mach_range=(0.3, 3.0),
                               max_steps=1000, res_drop=1e-5)
    for r in reqs:
        r["limiter"] = VANALBADA
        r["freeze_after"] = 300
    serve.make_grids(reqs, iters=500)
    a = serve.run_batched(reqs, None, check_every=100, max_solves=8, dt_mode="local",
                          precision="fp32c")
    b = serve.run_sequential(reqs, None, check_every=100, dt_mode="local", precision="fp32c")
    for r in reqs:
        x, y = a["results"][r["id"]], b["results"][r["id"]]
        print(f"[freeze] served request {r['id']} Mach {r['mach']:.2f}: {x['state']} after "
              f"{x['steps']} steps, drop {x['res_l2'] / x['res_l2_0']:.1e}", flush=True)
        assert (x["state"], x["steps"], x.get("cl"), x.get("cd")) == \
            (y["state"], y["steps"], y.get("cl"), y.get("cd")), r["id"]
