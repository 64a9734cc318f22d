"""The batched, precision-study and serving workloads at their configured sizes.

C3  64 solves in one persistent launch per RK stage, sizes drawn (seed 2026)
    from {256x128, 384x192, 512x384, 768x512}, NACA m~U[0,.06], p~U[.2,.6],
    t~U[.08,.18], Mach~U[0.3,3], AoA~U[-4,8] deg, scheme uniform over
    SLAU2 / SLAU / AUSM+-up, 5000 steps, the reference's stepping (global CFL
    step, FP32): every batched solve is bit-for-bit its one-by-one run
    through the single-grid step, and one solve of each size class is
    bit-for-bit ref32 over its first 40 steps in exact math.
C4  the precision study: CST section (6+6 coefficients, t/c 0.11), 1024 x 768,
    Mach 0.73, AoA 2.5 deg, SLAU2, local time step.  First order: the FP64
    solve reaches the 1e-6 density-residual drop, and the compensated FP32
    fast-math solve lands within 1e-4 of it in Cl and Cd_p and within 1e-3
    chord in the shock position.  MUSCL van Albada (the reference's
    reconstruction, a limit cycle at a 3e-2 drop in every precision): the
    window means agree to the same bounds.
C5  24 requests (seed 5) with sizes 128x64 ... 1024x768, Mach log-U[0.1, 10],
    AoA U[-5, 10] deg, drawn as C3, each to a 1e-5 drop or a 20k-step cap,
    arriving as a Poisson stream at 4 req/s and served with continuous
    batching (local time step, compensated FP32 state, first order): every
    served result (state, steps, Cl, Cd) equals the one-by-one result, and
    every converged request's Cl and Cd are within 1e-4 of its FP64 re-solve
    to the same criterion.
"""
import numpy as np
import pytest

import ausm
from ausm import precision, serve
from oracle_lib import Solver as OSolver
from oracle_lib import make_params

pytestmark = pytest.mark.gpu


def test_c3_batched_64_uneven_solves_bitwise():
    sizes = [(256, 128), (384, 192), (512, 384), (768, 512)]
    steps = 5000
    reqs = serve.make_requests(64, seed=2026, sizes=sizes, max_steps=steps, res_drop=0.0)
    serve.make_grids(reqs, iters=2000)
    print(f"[c3] {len(reqs)} solves, sizes {sorted({(r['M'], r['N']) for r in reqs})}", flush=True)
    with ausm.Batch(max_solves=64, math="fast", precision="fp32") as b:
        ids = {r["id"]: b.submit(r["vertices"], ausm.freestream(r["mach"], r["aoa"]),
                                 scheme=r["scheme"], cfl=r["cfl"], max_steps=steps,
                                 dt_mode="global") for r in reqs}
        while b.advance(250, 250) > 0:
            pass
        polls = {k: b.poll(sid) for k, sid in ids.items()}
        states = {k: b.state(sid) for k, sid in ids.items()}
    assert all(p["state"] == "capped" and p["steps"] == steps for p in polls.values()), polls
    ctxs = {}
    for r in reqs:
        key = (r["M"], r["N"])
        if key not in ctxs:
            ctxs[key] = ausm.Solver(*key)
        s = ctxs[key]
        s.set_vertices(r["vertices"])
        s.set_flow(r["mach"], r["aoa"])
        s.set_scheme(r["scheme"], ausm.VANALBADA)
        s.set_cfl(r["cfl"], 1.0, 1)
        s.set_math("fast")
        s.prepare_state(True)
        s.step(steps, info=False)
        assert np.array_equal(s.state(), states[r["id"]]), r["id"]
    for s in ctxs.values():
        s.close()

    # one solve per size class, exact math, against the FP32 oracle
    firsts = {}
    for r in reqs:
        firsts.setdefault((r["M"], r["N"]), r)
    with ausm.Batch(max_solves=8, math="exact", precision="fp32") as b:
        ids = {k: b.submit(r["vertices"], ausm.freestream(r["mach"], r["aoa"]), scheme=r["scheme"],
                           cfl=r["cfl"], max_steps=40, dt_mode="global")
               for k, r in firsts.items()}
        b.advance(40, 40)
        got = {k: b.state(sid) for k, sid in ids.items()}
    for k, r in firsts.items():
        ref = OSolver(32, make_params(r["mach"], r["aoa"], cfl=r["cfl"], maxdt=1.0), r["scheme"],
                      ausm.VANALBADA, r["vertices"], omp=True)
        ref.step(40)
        N = k[1]
        assert np.array_equal(got[k][1:N + 1], ref.state[1:N + 1]), k


def _study(limiter):
    return precision.run_study(M=1024, N=768, runs=(("fp64", "exact", limiter),
                                                    ("fp32c", "fast", limiter)),
                               drop=1e-6, max_steps=150000, check_every=5000, window=0.3)


def _check_study(res, converge):
    fp64, fp32c = res["runs"]
    e = fp32c["error_vs_fp64"]
    print(f"[c4] {fp64['limiter']}: FP64 {fp64['status']} drop {fp64['drop']:.2e} after "
          f"{fp64['steps']} steps, Cl {fp64['cl_mean']:.6f}; FP32C fast drop {fp32c['drop']:.2e}, "
          f"dCl {e['cl_mean']:+.2e} dCd_p {e['cd_p_mean']:+.2e} dx_shock {e['shock_mean_cp_upper']}",
          flush=True)
    if converge:
        assert fp64["status"] == "converged" and fp64["drop"] <= 1e-6
    assert abs(e["cl_mean"]) < 1e-4 and abs(e["cd_p_mean"]) < 1e-4
    assert e["shock_mean_cp_upper"] is not None and abs(e["shock_mean_cp_upper"]) < 1e-3


def test_c4_precision_study_first_order():
    _check_study(_study("none"), converge=True)


def test_c4_precision_study_van_albada():
    _check_study(_study("vanalbada"), converge=False)


def test_c5_serving_stream_equals_one_by_one_and_fp64():
    sizes = [(128, 64), (256, 128), (384, 192), (512, 384), (768, 512), (1024, 768)]
    reqs = serve.make_requests(24, seed=5, sizes=sizes, mach_range=(0.1, 10.0), log_mach=True,
                               aoa_range=(-5.0, 10.0), max_steps=20000, res_drop=1e-5)
    for r in reqs:
        r["limiter"] = ausm.NOLIMITER
    serve.make_grids(reqs, iters=2000)
    arr = serve.poisson_arrivals(len(reqs), 4.0, seed=11)
    served = serve.run_batched(reqs, arr, check_every=100, max_solves=64, dt_mode="local",
                               precision="fp32c")
    seq = serve.run_sequential(reqs, None, check_every=100, dt_mode="local", precision="fp32c")
    print(f"[c5] served {served['converged']} converged, {served['capped']} capped, p50 "
          f"{served['p50_latency_s']:.2f} s, p99 {served['p99_latency_s']:.2f} s", flush=True)
    assert served["converged"] > 0 and served["failed"] == 0 and seq["failed"] == 0
    for r in reqs:
        a, b = served["results"][r["id"]], seq["results"][r["id"]]
        assert (a["state"], a["steps"], a["cl"], a["cd"]) == (b["state"], b["steps"], b["cl"],
                                                               b["cd"]), r["id"]
    # FP64 re-solves of every converged request, stopped at the same
    # residual-drop criterion
    acc = serve.accuracy_vs_fp64(reqs, served["results"], n_check=len(reqs), deeper=1.0)
    assert len(acc["checked"]) == served["converged"]
    for c in acc["checked"]:
        print(f"[c5] request {c['id']} {c['M']}x{c['N']} Mach {c['mach']:.2f}: {c['steps']} "
              f"steps, FP64 {c['fp64_steps_at_criterion']}: dCl {c['cl_err_same_criterion']:+.2e} "
              f"dCd {c['cd_err_same_criterion']:+.2e}", flush=True)
        assert abs(c["cl_err_same_criterion"]) < 1e-4 and abs(c["cd_err_same_criterion"]) < 1e-4
