#!/usr/bin/env python3
"""Benchmark of the explicit O-grid solver step on one MI355X.

Workload (the single-instance configuration the headline metric is quoted
on): NACA 2412 O-grid of 512 x 384 cells at the reference's default geometry
(chord 0.6 at x = -0.7 inside the +-1.5 x +-1 rectangular far field,
power-law 7/5 body spacing, 10^4 FP64 Jacobi grid sweeps on the GPU), FP32
state, Mach 2.0, AoA 0, AUSM+-up with MUSCL (van Albada), TVD-RK3, CFL 0.5,
uniform free-stream start.  A "step" is one TVD-RK3 time step of the whole
grid (three fused stage kernels, dt refreshed every step).

Prints one JSON line: cell-steps/s over K timed steps (inputs resident in
HBM) in exact math -- the arithmetic that is bit-for-bit the FP32 oracle,
the restated reference WGSL -- the stage kernel's roofline position (its
average launch duration from HIP events over the timed region, three
launches per step), and the CPU baseline (the FP32 oracle in the
reference's unfused 15-pass structure, OpenMP over rows) on a bounded
sample of the same workload.  Beside them: the fast-math figure (hardware
reciprocal / square root, explicit fused multiply-adds), which is not held
to the per-step field-wise residual bound of 1e-5 against the oracle, with
its measured residual error at the final state next to the FP32 rounding
floor (the oracle's own residual error against FP64 at that state).  Two
more objects ride on the same line:
  batched  64 solves with sizes drawn from {256x128, 384x192, 512x384,
           768x512}, random NACA 4-digit sections, Mach U[0.3,3], AoA
           U[-4,8] deg, scheme uniform over SLAU2/SLAU/AUSM+-up, 5000 steps
           each: persistent batched kernel vs the same solves one by one;
  serving  Poisson arrivals of requests with sizes 128x64 ... 1024x768, Mach
           log-U[0.1,10], AoA U[-5,10] deg, each run to a residual drop or a
           step cap, served with continuous batching; converged solves/s and
           p50/p99 time-to-converged against the one-by-one baseline.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ausm-webgpu_amd"))

ALG_BYTES_PER_CELL_STEP = 152.0  # 8 x 16 B state streams + 3 x 8 B vertex
HBM_PEAK_GBPS = 8000.0
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9  # FP32 lane-ops/s, single issue, no packed math
METRIC = "cell-steps/s & % HBM roofline; batched converged solves/s + p50/p99 latency"
# steps timed for the supplementary single-grid figures (510x381, SLAU2):
# whatever --steps is, long enough that launch and graph latencies amortise
SUPPLEMENTARY_STEPS = 2000
TILE_ROWS = 2


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class HipEvents:
    """HIP events on an explicit stream (the solver's own stream)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.L.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def event(self):
        e = ctypes.c_void_p()
        assert self.L.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.L.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, e0, e1):
        assert self.L.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert self.L.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        return ms.value

    def device_sync(self):
        assert self.L.hipDeviceSynchronize() == 0


def build_workload(ausm, M, N, grid_iters):
    body = ausm.naca4_boundary(M, code=2412, size=0.3, x_offset=-0.7, exponent=1.4)
    outer = ausm.rect_boundary(M, 1.5)
    s = ausm.Solver(M, N, device=0)
    s.update_grid_boundaries(body, outer)
    t0 = time.perf_counter()
    s.prepare_grid(iters=grid_iters)
    t_grid = time.perf_counter() - t0
    s.set_flow(2.0, 0.0)
    s.set_scheme(ausm.AUSMUP, ausm.VANALBADA, 0.0, 0.75)
    s.set_cfl(0.5, 1.0, 1)
    # two rows per wavefront tile: in a sweep over 1 / 2 / 4 rows at this
    # size, 2 is the fastest (19.9 us per step against 21.0 and 22.6; 8 rows
    # were slower still in an earlier sweep); fixed so that timing, trace and
    # counter runs all measure the same kernel
    s.set_tile(TILE_ROWS)
    return s, t_grid


def time_steps(s, ev, steps, warmup):
    s.prepare_state(freestream=True)
    # graph set-up (capture, instantiation and upload of the chunks the
    # warm-up and the timed call launch) is done first, so that the warm-up
    # steps run straight into the timed ones and the timed region holds the
    # steps' kernels only
    s.prepare_steps(warmup)
    s.prepare_steps(steps)
    s.sync()
    ev.device_sync()
    e0, e1 = ev.event(), ev.event()
    stream = s.stream
    s.step(warmup, info=False)
    t0 = time.perf_counter()
    ev.record(e0, stream)
    s.step(steps, info=False)
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1)
    s.sync()
    wall = time.perf_counter() - t0
    info = s.step(1)
    return ms, wall, info


def cpu_baseline(vtx, M, N, target_s=15.0, q_gpu=None, r_gpu=None):
    """The FP32 oracle, unfused 15 passes per step, OpenMP over rows, on a
    bounded number of steps of the same 512 x 384 workload.  With the GPU's
    state at the end of its timed run and the exact-math residual of that
    state (q_gpu, r_gpu), the oracle also evaluates the residual of that
    state: the measured step's residual must be bit-for-bit the restated
    WGSL's."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import AUSMUP, VANALBADA, Oracle, make_params

    libp = os.path.join(ROOT, "oracle", "libbaseline.so")
    if not os.path.exists(libp):
        import subprocess

        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    B = ctypes.CDLL(libp)
    vp, ci = ctypes.c_void_p, ctypes.c_int
    B.base32_run_steps.argtypes = [vp, ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, ci, ci, vp, vp, vp]
    B.base32_run_steps.restype = ctypes.c_float
    B.base32_grid_finalize.argtypes = [ci, ci, vp, vp, vp]
    B.base32_init_state.argtypes = [vp, ci, ci, ci, vp]
    B.base32_max_wave_speed.argtypes = [vp, ci, ci, vp, vp, vp, vp]
    B.base32_max_wave_speed.restype = ctypes.c_float
    B.base32_num_threads.restype = ci
    threads = B.base32_num_threads()
    P = lambda a: a.ctypes.data_as(vp)
    par = make_params(2.0, 0.0, cfl=0.5, maxdt=1.0)
    v = np.ascontiguousarray(vtx, dtype=np.float32)
    area = np.zeros((N, M), np.float32)
    flen = np.zeros((N, M, 4), np.float32)
    B.base32_grid_finalize(M, N, P(v), P(area), P(flen))
    st = np.zeros((N + 3, M, 4), np.float32)
    B.base32_init_state(P(par), M, N, 1, P(st))
    s0, s1, s2 = st.copy(), st.copy(), st.copy()
    cnt = ctypes.c_longlong(0)
    mws = np.array([B.base32_max_wave_speed(P(par), M, N, P(s2), P(flen), P(area), None)],
                   np.float32)

    def run(n):
        t0 = time.perf_counter()
        B.base32_run_steps(P(par), 2, 0, M, N, P(v), P(flen), P(area), P(s0), P(s1), P(s2), n, 1,
                           ctypes.byref(cnt), P(mws), None)
        return time.perf_counter() - t0

    run(1)  # warm-up
    t1 = run(2) / 2
    n = max(2, min(2000, int(target_s / max(t1, 1e-6))))
    dt = run(n)
    # the same on one thread, a shorter bounded sample
    B.base32_set_threads.argtypes = [ci]
    B.base32_set_threads(1)
    t1s = run(1)
    n1 = max(1, min(200, int(target_s / 3 / max(t1s, 1e-6))))
    dt1 = run(n1)
    B.base32_set_threads(threads)
    agreement = None
    if q_gpu is not None and r_gpu is not None:
        r_ref = Oracle(32).residual(par, AUSMUP, VANALBADA, v, q_gpu)
        agreement = {"state": "GPU state after the timed steps",
                     "exact_math_residual_bitwise_equal_to_ref32": bool(np.array_equal(r_gpu, r_ref)),
                     "max_abs_diff": float(np.abs(r_gpu.astype(np.float64) - r_ref).max())}
    return {"value": M * N * n / dt, "unit": "cell-steps/s", "cores": threads,
            "kind": "port", "host": host_cpu(),
            "residual_agreement_vs_ref32": agreement,
            "sample": f"{n} steps of the 512x384 workload (ref32 oracle, 15 unfused passes per "
                      f"step, OpenMP over rows, {threads} threads, {dt:.1f} s)",
            "steps_per_s": n / dt,
            "single_thread": {"value": M * N * n1 / dt1, "unit": "cell-steps/s", "cores": 1,
                              "sample": f"{n1} steps, {dt1:.1f} s"}}


def grid_solve_record(M, N, sweeps, seconds):
    """The FP64 elliptic grid solve that built the workload's grid, timed on
    the host around the whole call (upload, interpolation, the Jacobi sweeps
    and the FP32 vertex and metric planes): sweeps per second and the sweeps'
    algorithmic traffic, 32 B per vertex per sweep (both FP64 planes read and
    written)."""
    verts = (N + 1) * M
    gbps = 32.0 * verts * sweeps / seconds / 1e9 if seconds > 0 else None
    return {"vertices": verts, "sweeps": sweeps, "seconds": seconds,
            "us_per_sweep": seconds / max(sweeps, 1) * 1e6,
            "achieved_GBps": gbps, "frac_hbm": None if gbps is None else gbps / HBM_PEAK_GBPS,
            "kernel": "ausm::grid_block_kernel<false> (four FP64 Jacobi sweeps per launch "
                      "through LDS)",
            "source": "host clock around ausm_grid_from_boundary"}


def host_cpu():
    """The host's CPU model and the logical CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def max_rel(a, ref):
    """max |a - ref| / max |ref| per conserved component (N x M x 4 fields)."""
    import numpy as np

    a = a.astype(np.float64).reshape(-1, 4)
    ref = ref.astype(np.float64).reshape(-1, 4)
    d = np.abs(a - ref).max(axis=0) / np.maximum(np.abs(ref).max(axis=0), 1e-300)
    return [float(x) for x in d]


def rho_norm_rel(a, ref):
    import numpy as np

    na = math.sqrt(float((a[..., 0].astype(np.float64) ** 2).sum()))
    nr = math.sqrt(float((ref[..., 0].astype(np.float64) ** 2).sum()))
    return abs(na - nr) / nr


def residual_parity(ausm, s, q):
    """Residuals of one state (the final state of the timed run) in exact
    math, fast math and FP64 arithmetic on the same FP32 vertices (the FP64
    solve is bit-for-bit the FP64 oracle): how far fast math is from the
    reference arithmetic, and how far the reference arithmetic itself is
    from the FP64 residual (its rounding floor)."""
    s.set_math("exact")
    r_exact = s.residual()
    s.set_math("fast")
    r_fast = s.residual()
    s.set_math("exact")
    M, N = s.M, s.N
    s64 = ausm.Solver(M, N, device=0)
    s64.set_vertices(s.vertices())
    s64.set_flow(2.0, 0.0)
    s64.set_scheme(ausm.AUSMUP, ausm.VANALBADA, 0.0, 0.75)
    s64.set_precision("fp64")
    s64.set_state(q)
    r64 = s64.residual()
    s64.close()
    return r_exact, {
        "state": "final state of the timed exact-math run",
        "fast_vs_exact_max_rel": max_rel(r_fast, r_exact),
        "fast_vs_exact_rho_norm_rel": rho_norm_rel(r_fast, r_exact),
        "exact_vs_fp64_max_rel": max_rel(r_exact, r64),
        "fast_vs_fp64_max_rel": max_rel(r_fast, r64),
        "bound": "per-step residual max-rel 1e-5 against the oracle: exact math meets it "
                 "bit-for-bit; fast math is reported, not held to it (the oracle's own FP32 "
                 "residual is exact_vs_fp64 away from the FP64 residual)"}


def calibration_copies(ev):
    """Three 512 MiB device-to-device copies: the kernel of known traffic
    the FETCH_SIZE / WRITE_SIZE counters are scaled against."""
    H = ev.L
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    n = 512 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(a), n) == 0 and H.hipMalloc(ctypes.byref(b), n) == 0
    assert H.hipMemset(a, 1, n) == 0
    for _ in range(3):
        assert H.hipMemcpy(b, a, n, 3) == 0  # device to device
    ev.device_sync()
    H.hipFree(a)
    H.hipFree(b)
    return n


def probe(ausm, ev, M, N, args):
    n = calibration_copies(ev)
    s, _ = build_workload(ausm, M, N, args.grid_iters)
    s.set_math(args.math)
    s.prepare_state(freestream=True)
    s.step(args.steps, info=False)
    s.sync()
    log(f"[bench] probe: 3 x {n} B copies, {args.steps} steps")
    s.close()


def probe_batched(ausm, ev, args):
    """Counter-collection run of the batched engine: the calibration copies,
    then the C3 mix (grids from --grid-iters sweeps) advanced --batch-steps
    steps in fast math, no timing and no JSON."""
    from ausm import serve

    n = calibration_copies(ev)
    reqs = serve.make_requests(args.batch_solves, seed=2026, sizes=BATCH_SIZES,
                               max_steps=args.batch_steps, res_drop=0.0)
    serve.make_grids(reqs, iters=args.grid_iters)
    with ausm.Batch(max_solves=len(reqs), math=args.math, precision="fp32") as b:
        for r in reqs:
            b.submit(r["vertices"], ausm.freestream(r["mach"], r["aoa"]), scheme=r["scheme"],
                     cfl=r["cfl"], max_steps=args.batch_steps, dt_mode="global")
        b.advance(args.batch_steps, args.batch_steps)
    log(f"[bench] batched probe: 3 x {n} B copies, {len(reqs)} solves x {args.batch_steps} steps")


BATCH_SIZES = [(256, 128), (384, 192), (512, 384), (768, 512)]


def batched_leg(serve, n, steps, grid_iters):
    sizes = BATCH_SIZES
    reqs = serve.make_requests(n, seed=2026, sizes=sizes, max_steps=steps, res_drop=0.0)
    serve.make_grids(reqs, iters=grid_iters)
    log(f"[bench] batched: {n} solves x {steps} steps")
    # the reference's stepping: global CFL step, plain FP32 state
    bat = serve.run_batched(reqs, None, check_every=250, max_solves=n, dt_mode="global",
                            precision="fp32")
    log(f"[bench] batched {bat['solves_per_s']:.2f} solves/s, {bat['cell_steps_per_s']:.3e} cell-steps/s")
    seq = serve.run_sequential(reqs, None, check_every=250, dt_mode="global", precision="fp32")
    log(f"[bench] one by one {seq['solves_per_s']:.2f} solves/s, {seq['cell_steps_per_s']:.3e} cell-steps/s")
    same = all(bat["results"][r["id"]]["state"] == seq["results"][r["id"]]["state"] and
               (bat["results"][r["id"]]["state"] == "failed" or
                bat["results"][r["id"]]["res_l2"] == seq["results"][r["id"]]["res_l2"])
               for r in reqs)
    strip = lambda d: {k: v for k, v in d.items() if k != "results"}
    # the whole batched run on the 152-B accounting (persistent stage
    # kernels plus the per-step dt kernel, convergence checks and admission)
    ach = ALG_BYTES_PER_CELL_STEP * bat["cell_steps_per_s"] / 1e9
    return {"workload": f"{n} solves, sizes from {sizes}, NACA m~U[0,.06] p~U[.2,.6] "
                        f"t~U[.08,.18], Mach~U[0.3,3], AoA~U[-4,8], scheme~U{{SLAU2,SLAU,AUSM+-up}}, "
                        f"{steps} steps each",
            "batched": strip(bat), "one_by_one": strip(seq),
            "speedup_solves_per_s": bat["solves_per_s"] / seq["solves_per_s"],
            "results_identical": same,
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBPS,
                         "kernel": "ausm::batch_stage_kernel (persistent fused RK stage, fast math)",
                         "source": "152 B per cell-step x the batched run's cell-steps/s (whole "
                                   "run, makespan timed on the host); per-launch durations in "
                                   "profiles/batch_stage_kernel_c3.md"}}


SERVE_VARIANTS = (("MUSCL van Albada (the reference's reconstruction)", 0),
                  ("first order (MUSCL off)", 3))


def serving_leg(serve, n, cap, drop, rates, grid_iters, dt_mode, precision, n_check):
    """The same request stream served twice: with the reference's MUSCL van
    Albada reconstruction (which settles into limit cycles at shocks and
    extrema, so requests end at the step cap) and first order (which
    converges): converged solves/s, p50/p99 latency, batched vs one by one,
    and the error of converged results against the FP64 solve."""
    sizes = [(128, 64), (256, 128), (384, 192), (512, 384), (768, 512), (1024, 768)]
    reqs = serve.make_requests(n, seed=5, sizes=sizes, mach_range=(0.1, 10.0), log_mach=True,
                               aoa_range=(-5.0, 10.0), max_steps=cap, res_drop=drop)
    serve.make_grids(reqs, iters=grid_iters)
    strip = lambda d: {k: v for k, v in d.items() if k != "results"}
    out = {"workload": f"{n} requests, sizes from {sizes}, Mach~logU[0.1,10], AoA~U[-5,10], "
                       f"residual drop {drop:g} or {cap}-step cap, checks every 100 steps",
           "dt_mode": dt_mode, "precision": precision, "variants": []}
    for label, limiter in SERVE_VARIANTS:
        for r in reqs:
            r["limiter"] = limiter
        var = {"reconstruction": label, "rates": []}
        seq = serve.run_sequential(reqs, None, check_every=100, dt_mode=dt_mode,
                                   precision=precision)
        var["one_by_one_closed_loop"] = strip(seq)
        log(f"[bench] serving ({label}) one by one: {seq['solves_per_s']:.2f} solves/s, "
            f"{seq['converged']} converged")
        sat = None
        for rate in rates:
            arr = None if rate == 0 else serve.poisson_arrivals(n, rate, seed=11)
            r = serve.run_batched(reqs, arr, check_every=100, max_solves=64, dt_mode=dt_mode,
                                  precision=precision)
            rec = dict(strip(r), arrival_rate_per_s=("all at t=0" if rate == 0 else rate))
            if rate:
                sq = serve.run_sequential(reqs, arr, check_every=100, dt_mode=dt_mode,
                                          precision=precision)
                rec["one_by_one_p50_s"] = sq["p50_latency_s"]
                rec["one_by_one_p99_s"] = sq["p99_latency_s"]
            else:
                sat = r
            var["rates"].append(rec)
            log(f"[bench] serving ({label}) rate {rate}: {r['solves_per_s']:.2f} solves/s, "
                f"{r['converged_solves_per_s']:.2f} converged/s, p50 {r['p50_latency_s']:.3f} s, "
                f"p99 {r['p99_latency_s']:.3f} s")
        if sat is not None:
            var["speedup_solves_per_s"] = sat["solves_per_s"] / seq["solves_per_s"]
            var["speedup_converged_solves_per_s"] = (
                sat["converged_solves_per_s"] / seq["converged_solves_per_s"]
                if seq["converged_solves_per_s"] > 0 else None)
            # served results equal the one-by-one results (same kernels, same
            # check points)
            a, b = sat["results"], seq["results"]
            var["results_identical_to_one_by_one"] = all(
                (a[k]["state"], a[k]["steps"], a[k].get("cl")) ==
                (b[k]["state"], b[k]["steps"], b[k].get("cl")) for k in a)
        if n_check > 0 and seq["converged"] > 0:
            acc = serve.accuracy_vs_fp64(reqs, seq["results"], n_check=n_check, dt_mode=dt_mode)
            var["accuracy_vs_fp64"] = acc
            log(f"[bench] serving ({label}) accuracy vs FP64 over {len(acc['checked'])} converged "
                f"requests: max |dCl| {acc['max_abs_cl_err_same_criterion']} at the same "
                f"criterion, {acc['max_abs_cl_err']} against the converged FP64 answer")
        out["variants"].append(var)
    # the engine where single solves leave the GPU idle: 64 small requests
    # (128 x 64 cells, first order, all converge)
    small = serve.make_requests(64, seed=7, sizes=[(128, 64)], mach_range=(0.1, 10.0),
                                log_mach=True, aoa_range=(-5.0, 10.0), max_steps=cap, res_drop=drop)
    for r in small:
        r["limiter"] = 3
    serve.make_grids(small, iters=grid_iters)
    sb = serve.run_batched(small, None, check_every=100, max_solves=64, dt_mode=dt_mode,
                           precision=precision)
    ss = serve.run_sequential(small, None, check_every=100, dt_mode=dt_mode, precision=precision)
    out["small_grid_mix"] = {
        "workload": "64 requests of 128x64 cells, first order, drawn as above",
        "batched": strip(sb), "one_by_one": strip(ss),
        "speedup_converged_solves_per_s": (sb["converged_solves_per_s"] / ss["converged_solves_per_s"]
                                           if ss["converged_solves_per_s"] > 0 else None)}
    log(f"[bench] serving small-grid mix: {sb['converged_solves_per_s']:.1f} vs "
        f"{ss['converged_solves_per_s']:.1f} converged solves/s one by one, p99 "
        f"{sb['p99_latency_s']:.2f} s vs {ss['p99_latency_s']:.2f} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--grid-iters", type=int, default=10000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--math", default="exact", choices=["fast", "exact"],
                    help="arithmetic of the headline (exact: bit-for-bit the FP32 oracle)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-literal", action="store_true",
                    help="skip the 510x381 figure")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--batch-solves", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-cap", type=int, default=20000)
    ap.add_argument("--serve-drop", type=float, default=1e-5)
    ap.add_argument("--serve-rates", default="0,4")
    ap.add_argument("--serve-dt", default="local", choices=["local", "global"])
    ap.add_argument("--serve-precision", default="fp32c", choices=["fp32c", "fp32"])
    ap.add_argument("--serve-check", type=int, default=3,
                    help="converged requests re-solved in FP64 for the accuracy figure")
    ap.add_argument("--probe", action="store_true",
                    help="counter-collection run: three 512 MiB device copies (the calibration "
                         "kernel for FETCH_SIZE / WRITE_SIZE) then --steps steps of the workload, "
                         "no timing and no JSON")
    ap.add_argument("--probe-batched", action="store_true",
                    help="counter-collection run of the batched engine: the calibration copies, "
                         "then the batched mix (grids from --grid-iters sweeps) advanced "
                         "--batch-steps steps, no timing and no JSON")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: the solver step has no multi-GPU path in this "
              "build; batched solves shard by solve, one process per GPU.",
              file=sys.stderr)
        sys.exit(2)

    import ausm

    M, N = 512, 384
    ev = HipEvents()
    if args.probe:
        probe(ausm, ev, M, N, args)
        return
    if args.probe_batched:
        probe_batched(ausm, ev, args)
        return
    log(f"[bench] building {M}x{N} NACA 2412 grid ({args.grid_iters} FP64 Jacobi sweeps)")
    s, t_grid = build_workload(ausm, M, N, args.grid_iters)
    log(f"[bench] grid in {t_grid:.2f} s")

    other = "fast" if args.math == "exact" else "exact"
    # The timed runs of both arithmetics use contexts of their own whose step
    # graphs are captured here, before any timing: a capture (or any other
    # host work) right before a timed region leaves the GPU idle, and after
    # about 10 ms of idle it runs the next ~10 ms of steps about 6 % slower
    # (lower clocks), which a short timed region would measure.
    so, _ = build_workload(ausm, M, N, args.grid_iters)
    s.set_math(args.math)
    so.set_math(other)
    for ctx in (s, so):
        ctx.prepare_state(freestream=True)
        ctx.prepare_steps(args.warmup)
        ctx.prepare_steps(args.steps)

    # Supplementary single-grid figures first: they also bring the GPU to its
    # sustained clocks, so that a short timed region (--steps 20 --warmup 5)
    # measures a warm GPU.
    # The same workload at the reference README's literal 510 x 381 (the
    # O-grid the reference's default grid settings give), for comparability
    literal = None
    if not args.no_literal:
        ml, nl = 510, 381
        sl, _ = build_workload(ausm, ml, nl, args.grid_iters)
        sl.set_math(args.math)
        nst = SUPPLEMENTARY_STEPS
        ms_l, _, _ = time_steps(sl, ev, nst, min(args.warmup, 100))
        literal = {"M": ml, "N": nl, "value": ml * nl * nst / (ms_l / 1e3), "unit": "cell-steps/s",
                   "ms_per_step": ms_l / nst, "steps": nst}
        log(f"[bench] {ml}x{nl}: {literal['value']:.3e} cell-steps/s")
        sl.close()

    # the same workload with SLAU2 (the reference's default scheme), both
    # arithmetics, on a context of its own; the timed runs follow at once
    s2, _ = build_workload(ausm, M, N, args.grid_iters)
    s2.set_scheme(ausm.SLAU2, ausm.VANALBADA, 0.0, 0.75)
    n_sl = SUPPLEMENTARY_STEPS
    slau2 = {}
    for mm in (other, args.math):
        s2.set_math(mm)
        ms_sl, _, _ = time_steps(s2, ev, n_sl, min(args.warmup, 100))
        slau2[mm] = M * N * n_sl / (ms_sl / 1e3)
        log(f"[bench] SLAU2 {mm} math: {slau2[mm]:.3e} cell-steps/s")

    log(f"[bench] timing {args.steps} steps ({args.math} math) after {args.warmup} warm-up")
    ms, wall, info = time_steps(s, ev, args.steps, args.warmup)
    # the other arithmetic beside it (same steps, same warm-up), straight after
    ms_o, _, _ = time_steps(so, ev, args.steps, args.warmup)
    # supplementary: each stage alone, launched back to back on warm caches
    # (and a warm GPU: straight after the timed runs)
    ms3 = (ctypes.c_float * 3)()
    ausm.lib.check(s._L.ausm_profile_stages(s.h, 200, ms3))
    stage_ms = [ms3[0], ms3[1], ms3[2]]
    s2.close()
    so.close()
    rate = M * N * args.steps / (ms / 1e3)
    step_ms = ms / args.steps
    log(f"[bench] {rate:.3e} cell-steps/s, {step_ms * 1e3:.2f} us/step, wall {wall:.3f} s, "
        f"res_l2 {info['res_l2']:.3e}")
    rate_o = M * N * args.steps / (ms_o / 1e3)
    launch_o = ms_o / 1e3 / (3 * args.steps)
    log(f"[bench] {other} math: {rate_o:.3e} cell-steps/s, {ms_o / args.steps * 1e3:.2f} us/step")
    q_end = s.state()
    r_end, parity = residual_parity(ausm, s, q_end)
    log(f"[bench] residual at the final state: fast vs exact max-rel "
        f"{max(parity['fast_vs_exact_max_rel']):.2e}, exact vs FP64 "
        f"{max(parity['exact_vs_fp64_max_rel']):.2e}")

    # the stage kernel's average launch duration over the timed region: the
    # HIP events above bracket exactly 3 x steps stage launches on the
    # solver's stream (a step is nothing but its three stage kernels)
    avg_launch_s = ms / 1e3 / (3 * args.steps)
    per_launch_bytes = ALG_BYTES_PER_CELL_STEP / 3 * M * N
    s.set_math(args.math)

    log(f"[bench] stage kernel {avg_launch_s * 1e6:.2f} us per launch in the timed region; "
        f"isolated stages {[round(x * 1e3, 2) for x in stage_ms]} us")


    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline")
        cpu = cpu_baseline(s.vertices(), M, N, q_gpu=q_end, r_gpu=r_end)
        if cpu["residual_agreement_vs_ref32"]:
            a = cpu["residual_agreement_vs_ref32"]
            log(f"[bench] exact-math residual bitwise ref32: "
                f"{a['exact_math_residual_bitwise_equal_to_ref32']}")
        log(f"[bench] CPU: {cpu['value']:.3e} cell-steps/s on {cpu['cores']} threads")

    from ausm import serve

    batched = None
    if not args.no_batched:
        batched = batched_leg(serve, args.batch_solves, args.batch_steps, 2000)
    serving = None
    if not args.no_serving:
        rates = [float(x) for x in args.serve_rates.split(",") if x != ""]
        serving = serving_leg(serve, args.serve_requests, args.serve_cap, args.serve_drop, rates,
                              2000, args.serve_dt, args.serve_precision, args.serve_check)

    # HBM traffic and VALU instructions per launch of each arithmetic's
    # stage kernel, from the committed rocprofv3 counter runs
    counters = {}
    tf = os.path.join(ROOT, "profiles", "stage_kernel_counters.json")
    if os.path.exists(tf):
        with open(tf) as f:
            counters = json.load(f).get("per_math", {})

    def roofline(math_mode, launch_s):
        c = counters.get(math_mode, {})
        valu = c.get("valu_insts_per_launch")
        ach = per_launch_bytes / launch_s / 1e9
        return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBPS, "traffic": c.get("traffic_bytes_per_launch"),
                "traffic_source": "profiles/stage_kernel_counters.json (rocprofv3 --pmc "
                                  "FETCH_SIZE x2.0 + WRITE_SIZE x1.0, bytes per launch)",
                "valu": (None if valu is None else
                         valu * 64 / (launch_s * VALU_PEAK_LANE_OPS)),
                "valu_source": "SQ_INSTS_VALU per launch x 64 lanes / (launch time x 39.3e12 "
                               "FP32 lane-ops/s: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, one "
                               "wave64 instruction per 4 cycles per SIMD, no packed math; two "
                               "plain FP32 instructions of two wavefronts can issue together, "
                               "at most twice that)",
                "valu_active_share": c.get("valu_active_share_of_wave_life"),
                "valu_coissue_share": c.get("valu_coissue_share"),
                "valu_active_source": "SQ_ACTIVE_INST_VALU / SQ_WAVE_CYCLES (share of a "
                                      "wavefront's life spent issuing VALU) and "
                                      "SQ_ACTIVE_INST_VALU2 / SQ_ACTIVE_INST_VALU (co-issued), "
                                      "profiles/stage_kernel_counters.json",
                "kernel": f"ausm::stage_kernel (fused RK stage, {math_mode} math)",
                "launch_us": launch_s * 1e6,
                "launch_source": "HIP events over the timed region / (3 x steps)"}

    whole_gbps = ALG_BYTES_PER_CELL_STEP * rate / 1e9
    out = {
        "metric": METRIC,
        "value": rate,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "NACA 2412 O-grid 512x384 (reference default geometry), Mach 2.0, "
                               "AoA 0, AUSM+-up + MUSCL van Albada, TVD-RK3, CFL 0.5, "
                               "free-stream start",
                   "M": M, "N": N, "math": args.math, "dt_refresh_every": 1,
                   "tile_rows_per_wavefront": TILE_ROWS,
                   "parallelism": "single GPU"},
        "pct_hbm_roofline_whole_step": 100.0 * whole_gbps / HBM_PEAK_GBPS,
        "steps_per_s": args.steps / (ms / 1e3),
        "slau2_cell_steps_per_s": slau2[args.math],
        "grid_solve": grid_solve_record(M, N, args.grid_iters, t_grid),
        "roofline": dict(roofline(args.math, avg_launch_s),
                         isolated_stage_us=[x * 1e3 for x in stage_ms]),
        "parity": parity,
        f"{other}_math": {"value": rate_o, "unit": "cell-steps/s",
                          "ms_per_step": ms_o / args.steps,
                          "slau2_cell_steps_per_s": slau2[other],
                          "roofline": roofline(other, launch_o),
                          "note": ("hardware reciprocal / square root and explicit fused "
                                   "multiply-adds; residual error at the final state in "
                                   "'parity' (not held to the 1e-5 field-wise bound)"
                                   if other == "fast" else
                                   "bit-for-bit the FP32 oracle")},
        "cpu_baseline": cpu,
        "literal_510x381": literal,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)
    s.close()


if __name__ == "__main__":
    main()
