"""Continuous-batching service for airfoil solves, and its sequential baseline.

Requests (airfoil, Mach, AoA, scheme, grid size) arrive over time; the
service keeps every admitted solve resident in HBM, admits new ones as table
slots and arena memory free up, advances all running solves together and
retires each one at its residual-drop criterion or step cap.  The baseline
runs the same requests one at a time through the plain single-grid step.

Grid generation (FP64 elliptic smoothing on the GPU) happens before the
timed loops and is not part of a solve's latency.

Steady solves step by default with local time stepping (each cell its own
CFL step) and the compensated FP32 state: with the reference's single
global step the far field of a fine O-grid barely moves, and plain FP32
accumulation stagnates at a residual drop of ~1e-5 with Cl off by ~1e-4
or more.  Both are per-run options (dt_mode, precision).

A request may carry "freeze_after" (steps): its van Albada limiter is
frozen (Solver.freeze_limiter) at the first convergence check at or after
that many steps, in the batched service, the one-by-one baseline and the
FP64 re-solve alike, so that a MUSCL solve held in a limit cycle by its
limiter can still converge.
"""
import math
import time

import numpy as np

from .batch import Batch
from .solver import Solver, circle_boundary, freestream, naca4_boundary

SCHEME_NAMES = ["SLAU2", "SLAU", "AUSM+-up"]


def make_requests(n, seed, sizes, mach_range=(0.3, 3.0), log_mach=False, aoa_range=(-4.0, 8.0),
                  max_steps=5000, res_drop=0.0, cfl=0.5):
    """Requests drawn as in the batched configuration: NACA 4-digit sections
    with m ~ U[0, 0.06], p ~ U[0.2, 0.6], t ~ U[0.08, 0.18], Mach and AoA
    uniform (or log-uniform Mach), scheme uniform over SLAU2 / SLAU /
    AUSM+-up, grid size drawn from `sizes`."""
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        M, N = sizes[rng.integers(len(sizes))]
        lo, hi = mach_range
        mach = math.exp(rng.uniform(math.log(lo), math.log(hi))) if log_mach else rng.uniform(lo, hi)
        reqs.append(dict(id=k, M=int(M), N=int(N), m=rng.uniform(0, 0.06), p=rng.uniform(0.2, 0.6),
                         t=rng.uniform(0.08, 0.18), mach=float(mach),
                         aoa=float(rng.uniform(*aoa_range)), scheme=int(rng.integers(3)),
                         max_steps=max_steps, res_drop=res_drop, cfl=cfl))
    return reqs


def make_grids(reqs, iters=1000, radius_chords=20.0, device=0):
    """FP32 vertices of each request's O-grid (NACA body of unit chord,
    circular far field at radius_chords), by the library's FP64 grid solve."""
    ctxs = {}
    for r in reqs:
        key = (r["M"], r["N"])
        if key not in ctxs:
            ctxs[key] = Solver(r["M"], r["N"], device=device)
        s = ctxs[key]
        body = naca4_boundary(r["M"], m=r["m"], p=r["p"], t=r["t"], size=0.5, spacing="cosine")
        s.update_grid_boundaries(body, circle_boundary(r["M"], radius_chords))
        s.prepare_grid(iters=iters)
        r["vertices"] = s.vertices()
    for s in ctxs.values():
        s.close()
    return reqs


def percentile(xs, q):
    if not xs:
        return None
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


def run_batched(reqs, arrivals=None, check_every=100, max_solves=64, hbm_budget_bytes=0,
                math="fast", device=0, release=True, dt_mode="local", precision="fp32c"):
    """Serve the requests with continuous batching.  arrivals[k] is request
    k's arrival time in seconds from the start (None: all at t = 0).
    Returns per-request results and aggregate throughput / latency."""
    n = len(reqs)
    arrivals = np.zeros(n) if arrivals is None else np.asarray(arrivals, dtype=np.float64)
    order = np.argsort(arrivals, kind="stable")
    b = Batch(max_solves=max_solves, hbm_budget_bytes=hbm_budget_bytes, device=device, math=math,
              precision=precision)
    sids = {}
    done = {}
    nxt = 0
    t0 = time.perf_counter()
    while len(done) < n:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[order[nxt]] <= now:
            r = reqs[order[nxt]]
            sids[r["id"]] = b.submit(r["vertices"], freestream(r["mach"], r["aoa"]),
                                     scheme=r["scheme"], limiter=r.get("limiter", 0),
                                     cfl=r["cfl"], max_steps=r["max_steps"],
                                     res_drop=r["res_drop"], ref_length=1.0, dt_mode=dt_mode,
                                     freeze_after=r.get("freeze_after", 0))
            nxt += 1
        active = [k for k in sids if k not in done]
        if not active:
            time.sleep(max(0.0, min(0.01, arrivals[order[nxt]] - now)))
            continue
        b.advance(check_every, check_every)
        tnow = time.perf_counter() - t0
        for k in active:
            st = b.poll(sids[k])
            if st["state"] in ("converged", "capped", "failed"):
                st["latency_s"] = tnow - arrivals[k]
                st["done_s"] = tnow
                done[k] = st
                if release:
                    b.release(sids[k])
    makespan = max(d["done_s"] for d in done.values()) - float(arrivals.min())
    b.close()
    return summarize(reqs, done, makespan)


def run_sequential(reqs, arrivals=None, check_every=100, math="fast", device=0, dt_mode="local",
                   precision="fp32c"):
    """The same requests one by one through the plain single-grid step, in
    arrival order, each run to its criterion (checked every check_every
    steps, as in the batched service)."""
    n = len(reqs)
    arrivals = np.zeros(n) if arrivals is None else np.asarray(arrivals, dtype=np.float64)
    order = np.argsort(arrivals, kind="stable")
    ctxs = {}
    done = {}
    t0 = time.perf_counter()
    for k in order:
        r = reqs[k]
        now = time.perf_counter() - t0
        if arrivals[k] > now:
            time.sleep(arrivals[k] - now)
        key = (r["M"], r["N"])
        if key not in ctxs:
            ctxs[key] = Solver(r["M"], r["N"], device=device)
        s = ctxs[key]
        s.set_vertices(r["vertices"])
        s.set_flow(r["mach"], r["aoa"])
        s.set_scheme(r["scheme"], r.get("limiter", 0))
        s.set_cfl(r["cfl"], 1.0, 1)
        s.set_math(math)
        s.set_dt_mode(dt_mode)
        s.set_precision(precision)
        s.prepare_state(True)
        res0 = s.residual_norm()
        steps, state, res = 0, "running", res0
        fa, frozen = r.get("freeze_after", 0), False
        while state == "running":
            try:
                info = s.step(check_every)
            except Exception:
                state = "failed"
                break
            steps += check_every
            res = info["res_l2"]
            if not math_isfinite(res):
                state = "failed"
            elif r["res_drop"] > 0 and res <= r["res_drop"] * res0:
                state = "converged"
            elif steps >= r["max_steps"]:
                state = "capped"
            elif fa > 0 and not frozen and steps >= fa:
                s.freeze_limiter()
                frozen = True
        st = {"state": state, "steps": steps, "res_l2": res, "res_l2_0": res0}
        if state != "failed":
            f = s.forces(1.0)
            st["cl"], st["cd"] = f["cl"], f["cd"]
        tnow = time.perf_counter() - t0
        st["latency_s"] = tnow - arrivals[k]
        st["done_s"] = tnow
        done[r["id"]] = st
    for s in ctxs.values():
        s.close()
    makespan = max(d["done_s"] for d in done.values()) - float(arrivals.min())
    return summarize(reqs, done, makespan)


def accuracy_vs_fp64(reqs, done, n_check=4, check_every=100, dt_mode="local", device=0,
                     deeper=1e-3, step_factor=3):
    """Cl / Cd error of served results against the all-FP64 solve of the same
    request (exact arithmetic in double, bit-for-bit the FP64 oracle on
    small grids), for the first n_check converged requests, at two points:
    where the FP64 run meets the same residual-drop criterion (the error
    the arithmetic adds) and after it has gone `deeper` times further down
    (or step_factor times the cap): the converged answer, so that error also
    holds what the criterion leaves unconverged."""
    checked = []
    for r in reqs:
        st = done[r["id"]]
        if st["state"] != "converged" or len(checked) >= n_check:
            continue
        s = Solver(r["M"], r["N"], device=device)
        s.set_vertices(r["vertices"])
        s.set_flow(r["mach"], r["aoa"])
        s.set_scheme(r["scheme"], r.get("limiter", 0))
        s.set_cfl(r["cfl"], 1.0, 1)
        s.set_dt_mode(dt_mode)
        s.set_precision("fp64")
        s.prepare_state(True)
        res0 = s.residual_norm()
        steps, res = 0, res0
        at_criterion = None
        fa, frozen = r.get("freeze_after", 0), False
        while steps < step_factor * r["max_steps"] and res > deeper * r["res_drop"] * res0:
            res = s.step(check_every)["res_l2"]
            steps += check_every
            met_now = at_criterion is None and res <= r["res_drop"] * res0
            if met_now:
                f = s.forces(1.0)
                at_criterion = (steps, f["cl"], f["cd"])
            # frozen at the check the served solve freezes at (one where it
            # does not retire), or later on the way to the deeper answer
            if fa > 0 and not frozen and steps >= fa and not met_now:
                s.freeze_limiter()
                frozen = True
        f = s.forces(1.0)
        s.close()
        c = {"id": r["id"], "M": r["M"], "N": r["N"], "mach": r["mach"], "steps": st["steps"],
             "fp64_steps": steps, "fp64_drop": res / res0,
             "cl_err": st["cl"] - f["cl"], "cd_err": st["cd"] - f["cd"]}
        if at_criterion is not None:
            c["fp64_steps_at_criterion"] = at_criterion[0]
            c["cl_err_same_criterion"] = st["cl"] - at_criterion[1]
            c["cd_err_same_criterion"] = st["cd"] - at_criterion[2]
        checked.append(c)
    mx = lambda key: max((abs(c[key]) for c in checked if key in c), default=None)
    return {"checked": checked,
            "max_abs_cl_err_same_criterion": mx("cl_err_same_criterion"),
            "max_abs_cd_err_same_criterion": mx("cd_err_same_criterion"),
            "max_abs_cl_err": mx("cl_err"), "max_abs_cd_err": mx("cd_err")}


def math_isfinite(x):
    return x == x and abs(x) != float("inf")


def summarize(reqs, done, makespan):
    lat = [done[r["id"]]["latency_s"] for r in reqs]
    conv = [r for r in reqs if done[r["id"]]["state"] == "converged"]
    finished = [r for r in reqs if done[r["id"]]["state"] in ("converged", "capped")]
    cell_steps = sum(r["M"] * r["N"] * done[r["id"]]["steps"] for r in reqs)
    return {
        "n": len(reqs),
        "converged": len(conv),
        "capped": sum(1 for r in reqs if done[r["id"]]["state"] == "capped"),
        "failed": sum(1 for r in reqs if done[r["id"]]["state"] == "failed"),
        "makespan_s": makespan,
        "solves_per_s": len(finished) / makespan,
        "converged_solves_per_s": len(conv) / makespan,
        "cell_steps_per_s": cell_steps / makespan,
        "p50_latency_s": percentile(lat, 50),
        "p99_latency_s": percentile(lat, 99),
        "results": done,
    }


def poisson_arrivals(n, rate, seed):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n)) - 0.0
