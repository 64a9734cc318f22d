"""Precision study: how far FP32 flux and stage arithmetic moves the steady
transonic answer away from the same scheme run entirely in FP64.

Workload: an aft-cambered transonic section from a seeded Kulfan CST
parameterisation (6 upper + 6 lower coefficients, t/c = 0.11, round nose),
on a 1024 x 768 O-grid (elliptic FP64 grid solve on the GPU), Mach 0.73,
AoA 2.5 deg, SLAU2 + MUSCL (van Albada), TVD-RK3 from a uniform free stream
until the density-residual L2 norm has dropped by `drop` (or a step cap).
The reference's single global CFL step is set by the smallest nose cells
and leaves the far field frozen (a 3e-2 drop after 4e5 steps on this grid),
so the study steps with local time stepping by default (each cell its own
CFL step: the same discrete operator and steady state).  Every run uses the same FP64 grid
(the FP32 runs see it rounded to FP32); the grid solve, the residual norms
and the Cp / Cl / Cd_p integration are FP64 in every run.

Runs: FP32 exact (the reference's f32 arithmetic, bit-for-bit the WGSL
semantics), FP32 fast (the production arithmetic) and FP64 (exact
arithmetic in double; bit-for-bit the FP64 oracle, checked on small grids
by the GPU tests).  Reported per FP32 run: Cl, Cd_p and upper/lower shock
positions, with their differences from the FP64 run.

    python -m ausm.precision [--M 1024 --N 768 --drop 1e-6 --out result.json]
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import lib as L
from .solver import Solver, circle_boundary, cst_boundary, freestream

# aft-cambered baseline (flat upper surface, aft-loaded lower surface); the
# seed perturbs it and the thickness is then scaled to the target t/c
CST_UPPER = (0.13, 0.13, 0.16, 0.14, 0.18, 0.20)
CST_LOWER = (-0.13, -0.14, -0.10, -0.03, 0.03, 0.06)


def _cst_y(a, x):
    n = len(a) - 1
    s = sum(a[i] * math.comb(n, i) * x ** i * (1 - x) ** (n - i) for i in range(n + 1))
    return np.sqrt(x) * (1 - x) * s


def cst_coefficients(seed=0, thickness=0.11, jitter=0.005):
    """Seeded (a_upper, a_lower) of the study section: the baseline plus
    N(0, jitter) per coefficient, scaled so that max(y_u - y_l) = thickness
    (chord 1).  The nose radius is a_upper[0]^2 / 2."""
    rng = np.random.default_rng(seed)
    au = np.asarray(CST_UPPER) + rng.normal(0.0, jitter, 6)
    al = np.asarray(CST_LOWER) + rng.normal(0.0, jitter, 6)
    x = np.linspace(0.0, 1.0, 20001)
    t = float((_cst_y(au, x) - _cst_y(al, x)).max())
    k = thickness / t
    return au * k, al * k


def critical_cp(mach, gamma=1.4):
    """Pressure coefficient at which the local Mach number is 1."""
    g = gamma
    return (2.0 / (g * mach * mach)) * (
        ((2.0 + (g - 1.0) * mach * mach) / (g + 1.0)) ** (g / (g - 1.0)) - 1.0)


def shock_positions(body_xy, cp, mach, gamma=1.4, chord=1.0, x_le=-0.5):
    """Chordwise shock position x/c on each surface: the last place, going
    from the nose to the trailing edge, where the segment Cp rises through
    the critical Cp (supersonic -> subsonic), linearly interpolated between
    segment midpoints; None where the surface stays subsonic.  Segment i
    joins body points i and i+1 (points CCW from the trailing edge, upper
    surface first)."""
    M = body_xy.shape[0]
    xm = 0.5 * (body_xy[:, 0] + np.roll(body_xy[:, 0], -1))
    xc = (xm - x_le) / chord
    cs = critical_cp(mach, gamma)
    out = {}
    for name, idx in (("upper", np.arange(M // 2 - 1, -1, -1)),   # nose -> TE
                      ("lower", np.arange(M // 2, M - 1))):
        c, x = cp[idx], xc[idx]
        pos = None
        for k in range(len(idx) - 1):
            if c[k] < cs <= c[k + 1]:
                f = (cs - c[k]) / (c[k + 1] - c[k])
                pos = float(x[k] + f * (x[k + 1] - x[k]))
        out[name] = pos
    return out


def make_grid(M, N, au, al, radius=20.0, iters=10000, device=0):
    """FP64 O-grid around the CST body (power-law spacing, exponent 1.4, as
    the reference's generator), circular far field at `radius` chords."""
    s = Solver(M, N, device=device)
    s.update_grid_boundaries(cst_boundary(M, au, al, size=0.5, spacing="power"),
                             circle_boundary(M, radius))
    s.prepare_grid(iters=iters)
    v = s.vertices(fp64=True)
    s.close()
    return v


def solve(vtx, mach, aoa, prec="fp32", math_mode="exact", cfl=0.5, drop=1e-6,
          max_steps=100_000, check_every=2000, scheme=L.SLAU2, limiter=L.VANALBADA,
          device=0, log=None, time_limit=None, dt_mode="local", window=0.5):
    """One steady solve from a uniform free stream, to `drop` or max_steps.
    Cl, Cd_p (chord 1) and the segment Cp are sampled at every check of the
    last `window` fraction of the steps: a solution that settles into a
    limit cycle instead of converging is reported by its window mean and
    peak-to-peak amplitude (a converged one has amplitude ~ 0)."""
    N1, M, _ = vtx.shape
    s = Solver(M, N1 - 1, device=device)
    s.set_vertices(vtx)  # float64: exact for the FP64 run, rounded for FP32
    fs = freestream(mach, aoa)
    s.set_freestream(*fs)
    s.set_scheme(scheme, limiter, 0.0, 0.75)
    s.set_cfl(cfl, 1.0, 1)
    s.set_math(math_mode)
    s.set_dt_mode(dt_mode)
    s.set_precision(prec)
    s.prepare_state(freestream=True)
    t0 = time.perf_counter()
    info = s.step(1)
    r0 = info["res_l2"]
    f = s.forces(ref_length=1.0)
    hist = [(1, r0, f["cl"], f["cd"])]
    steps = 1
    status = "cap"
    best = r0
    w_start = int((1.0 - window) * max_steps)
    samples = []
    while steps < max_steps:
        n = min(check_every, max_steps - steps)
        info = s.step(n)
        steps += n
        r = info["res_l2"]
        f = s.forces(ref_length=1.0, want_cp=steps >= w_start)
        hist.append((steps, r, f["cl"], f["cd"]))
        best = min(best, r)
        if log:
            log(f"  {prec}/{math_mode}/lim{limiter} step {steps} res {r:.3e} drop {r / r0:.3e}")
        if info["nonfinite"] or not math.isfinite(r):
            status = "diverged"
            break
        if r <= drop * r0:
            status = "converged"
            break
        if steps >= w_start:
            samples.append((f["cl"], f["cd"], f["cp"]))
        if time_limit is not None and time.perf_counter() - t0 > time_limit:
            status = "time limit"
            break
    s.sync()
    wall = time.perf_counter() - t0
    f = s.forces(ref_length=1.0, want_cp=True)
    if status == "converged" or not samples:
        samples = [(f["cl"], f["cd"], f["cp"])]
    cl = np.array([x[0] for x in samples])
    cd = np.array([x[1] for x in samples])
    cp_mean = np.mean([x[2] for x in samples], axis=0)
    body = vtx[0]
    res = {
        "precision": prec, "math": math_mode, "limiter": limiter, "status": status,
        "steps": steps, "res_l2_0": r0, "res_l2": hist[-1][1], "drop": hist[-1][1] / r0,
        "best_drop": best / r0,
        "cl": float(f["cl"]), "cd_p": float(f["cd"]),
        "cl_mean": float(cl.mean()), "cl_amp": float(cl.max() - cl.min()),
        "cd_p_mean": float(cd.mean()), "cd_p_amp": float(cd.max() - cd.min()),
        "n_samples": len(samples),
        "shock": shock_positions(body, f["cp"], mach),
        "shock_mean_cp": shock_positions(body, cp_mean, mach),
        "seconds": wall, "cell_steps_per_s": M * (N1 - 1) * steps / wall,
        "history": hist,
    }
    res["_cp"] = cp_mean
    s.close()
    return res


LIMITER_NAMES = {"vanalbada": L.VANALBADA, "minmod": L.MINMOD, "vanleer": L.VANLEER,
                 "none": L.NOLIMITER, "vanalbada_clamped": L.VANALBADA_CLAMPED}
DEFAULT_RUNS = tuple((p, m, lim) for lim in ("vanalbada", "none")
                     for p, m in (("fp64", "exact"), ("fp32", "exact"), ("fp32", "fast"),
                                  ("fp32c", "exact"), ("fp32c", "fast")))


def run_study(M=1024, N=768, mach=0.73, aoa=2.5, seed=0, thickness=0.11, radius=20.0,
              grid_iters=10000, cfl=0.5, drop=1e-6, max_steps=100_000, check_every=2000,
              runs=DEFAULT_RUNS, device=0, log=None, time_limit=None, dt_mode="local",
              window=0.5):
    """All runs share one FP64 grid.  Each FP32 run is compared with the FP64
    run of the same limiter (MUSCL van Albada = the reference's scheme;
    "none" = first order, which converges to the arithmetic's floor)."""
    au, al = cst_coefficients(seed, thickness)
    vtx = make_grid(M, N, au, al, radius, grid_iters, device)
    out = {"workload": f"CST seed {seed} t/c {thickness} O-grid {M}x{N}, far field {radius} chords",
           "mach": mach, "aoa": aoa, "scheme": "SLAU2", "cfl": cfl, "dt_mode": dt_mode,
           "drop_target": drop, "max_steps": max_steps, "window": window,
           "a_upper": list(map(float, au)), "a_lower": list(map(float, al)),
           "cp_critical": critical_cp(mach), "runs": []}
    refs = {}
    for prec, mm, lim in runs:
        limiter = LIMITER_NAMES[lim]
        r = solve(vtx, mach, aoa, prec, mm, cfl, drop, max_steps, check_every, limiter=limiter,
                  device=device, log=log, time_limit=time_limit, dt_mode=dt_mode, window=window)
        r["limiter"] = lim
        if prec == "fp64" and lim not in refs:
            refs[lim] = r
        out["runs"].append(r)
    for r in out["runs"]:
        ref = refs.get(r["limiter"])
        if ref is None or r is ref:
            continue
        e = {"cl": r["cl"] - ref["cl"], "cd_p": r["cd_p"] - ref["cd_p"],
             "cl_mean": r["cl_mean"] - ref["cl_mean"], "cd_p_mean": r["cd_p_mean"] - ref["cd_p_mean"],
             "cp_mean_max_abs": float(np.abs(r["_cp"] - ref["_cp"]).max())}
        for key in ("shock", "shock_mean_cp"):
            for side in ("upper", "lower"):
                a, b = r[key][side], ref[key][side]
                e[key + "_" + side] = None if a is None or b is None else a - b
        r["error_vs_fp64"] = e
    for r in out["runs"]:
        del r["_cp"]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--M", type=int, default=1024)
    ap.add_argument("--N", type=int, default=768)
    ap.add_argument("--mach", type=float, default=0.73)
    ap.add_argument("--aoa", type=float, default=2.5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--radius", type=float, default=20.0)
    ap.add_argument("--grid-iters", type=int, default=10000)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--drop", type=float, default=1e-6)
    ap.add_argument("--max-steps", type=int, default=100_000)
    ap.add_argument("--check-every", type=int, default=2000)
    ap.add_argument("--window", type=float, default=0.5,
                    help="final fraction of the steps over which forces are averaged")
    ap.add_argument("--time-limit", type=float, default=None,
                    help="seconds per run before it is stopped (reported as such)")
    ap.add_argument("--runs", default=",".join(":".join(r) for r in DEFAULT_RUNS),
                    help="precision:math:limiter, comma separated")
    ap.add_argument("--dt-mode", default="local", choices=["local", "global"],
                    help="local time stepping (steady-state accelerator) or the global CFL step")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    runs = tuple(tuple(r.split(":")) for r in a.runs.split(","))
    if any(len(r) != 3 or r[2] not in LIMITER_NAMES for r in runs):
        ap.error("--runs entries are precision:math:limiter")

    def log(msg):
        print(msg, flush=True)

    res = run_study(a.M, a.N, a.mach, a.aoa, a.seed, 0.11, a.radius, a.grid_iters, a.cfl, a.drop,
                    a.max_steps, a.check_every, runs, log=log, time_limit=a.time_limit,
                    dt_mode=a.dt_mode, window=a.window)
    txt = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt)
    summary = {k: v for k, v in res.items() if k != "runs"}
    summary["runs"] = [{k: v for k, v in r.items() if k != "history"} for r in res["runs"]]
    print(json.dumps(summary, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
