"""GPU tests of batched solves (ausm_batch_*): every solve in a batch must
step exactly like the single-grid path (bitwise, same kernels and dt rule),
be bitwise the same alone and in any mix, converge / cap / fail cleanly
without disturbing its neighbours, and be admitted as memory frees up."""
import numpy as np
import pytest

import ausm
from oracle_lib import Oracle, circle_boundary, naca4_boundary, naca4_digits

pytestmark = pytest.mark.gpu


def grid(M, N, code=12, iters=80, radius=8.0):
    m, p, t = naca4_digits(code)
    body = naca4_boundary(M, m, p, t, size=0.5, expo=1.4, spacing=1)
    return Oracle(64).grid(body, circle_boundary(M, radius), N, iters).astype(np.float32)


CASES = [
    dict(M=96, N=32, code=12, mach=0.8, aoa=1.25, scheme=ausm.SLAU2),
    dict(M=150, N=41, code=2412, mach=2.0, aoa=0.0, scheme=ausm.AUSMUP),
    dict(M=64, N=20, code=4412, mach=0.5, aoa=4.0, scheme=ausm.SLAU),
    dict(M=200, N=64, code=2412, mach=1.2, aoa=-2.0, scheme=ausm.SLAU2),
]


def single_run(c, vtx, steps, math, tile=0, dt_mode="global", precision="fp32"):
    s = ausm.Solver(c["M"], c["N"])
    s.set_tile(tile)
    s.set_dt_mode(dt_mode)
    s.set_precision(precision)
    s.set_vertices(vtx)
    s.set_flow(c["mach"], c["aoa"])
    s.set_scheme(c["scheme"], c.get("limiter", ausm.VANALBADA), 0.0, 0.75)
    s.set_cfl(0.5, 1.0, 1)
    s.set_math(math)
    s.prepare_state(True)
    s.step(steps, info=False)
    out = s.state()
    s.close()
    return out


@pytest.mark.parametrize("math,sched,tile", [("exact", (0, 0), 0), ("fast", (2, 2), 4),
                                             ("fast", (1, 4), 1)])
def test_batch_compensated_matches_single_grid_bitwise(math, sched, tile):
    """Compensated-state batches (local time step) equal single-grid FP32C
    runs; hi is compared (the batch returns hi as the state)."""
    steps = 60
    grids = [grid(c["M"], c["N"], c["code"]) for c in CASES]
    with ausm.Batch(max_solves=8, math=math, precision="fp32c") as b:
        b.set_schedule(*sched)
        ids = [b.submit(g, ausm.freestream(c["mach"], c["aoa"]), scheme=c["scheme"],
                        max_steps=steps, dt_mode="local") for c, g in zip(CASES, grids)]
        b.advance(steps, check_every=20)
        for c, g, sid in zip(CASES, grids, ids):
            got = b.state(sid)
            ref = single_run(c, g, steps, math, tile, "local", "fp32c")
            N = c["N"]
            assert np.array_equal(got[1:N + 1], ref[1:N + 1]), c
        # the precision is fixed while a solve is queued or running
        b.submit(grids[0], ausm.freestream(0.5, 0.0), max_steps=10)
        with pytest.raises(ausm.AusmError):
            b.set_precision("fp32")


@pytest.mark.parametrize("math,sched,tile,dt_mode", [
    ("exact", (0, 0), 0, "global"), ("fast", (0, 0), 0, "global"), ("fast", (4, 2), 1, "global"),
    ("fast", (1, 1), 2, "global"), ("exact", (0, 0), 0, "local"), ("fast", (2, 4), 4, "local")])
def test_batch_matches_single_grid_bitwise(math, sched, tile, dt_mode):
    """A batch with any schedule equals single-grid runs with any tile shape,
    with the global and the local time step."""
    steps = 60
    grids = [grid(c["M"], c["N"], c["code"]) for c in CASES]
    with ausm.Batch(max_solves=8, math=math) as b:
        b.set_schedule(*sched)
        ids = [b.submit(g, ausm.freestream(c["mach"], c["aoa"]), scheme=c["scheme"],
                        max_steps=steps, dt_mode=dt_mode) for c, g in zip(CASES, grids)]
        b.advance(steps, check_every=20)
        for c, g, sid in zip(CASES, grids, ids):
            st = b.poll(sid)
            assert st["state"] == "capped" and st["steps"] == steps
            got = b.state(sid)
            ref = single_run(c, g, steps, math, tile, dt_mode)
            N = c["N"]
            assert np.array_equal(got[1:N + 1], ref[1:N + 1]), c


@pytest.mark.parametrize("sched", [(1, 1), (2, 1), (4, 2), (0, 16)])
def test_batch_uneven_runs_bitwise(sched):
    """More items than workgroups, so that the persistent workgroups take
    runs of unequal length (one row per wavefront and one workgroup per CU:
    about 520 items on 256 workgroups; more rows or workgroups: runs of one
    or two, or workgroups left without items; 16 workgroups per CU with the
    rows tuned: a workgroup count outside the tuner's own candidates).
    Every solve still equals its single-grid run bitwise."""
    steps = 30
    cases = CASES + [dict(M=480, N=200, code=2412, mach=0.7, aoa=2.0, scheme=ausm.AUSMUP)]
    grids = [grid(c["M"], c["N"], c["code"]) for c in cases]
    with ausm.Batch(max_solves=8, math="fast") as b:
        b.set_schedule(*sched)
        ids = [b.submit(g, ausm.freestream(c["mach"], c["aoa"]), scheme=c["scheme"],
                        max_steps=steps) for c, g in zip(cases, grids)]
        b.advance(steps, check_every=10)
        for c, g, sid in zip(cases, grids, ids):
            got = b.state(sid)
            ref = single_run(c, g, steps, "fast", 2)
            N = c["N"]
            assert np.array_equal(got[1:N + 1], ref[1:N + 1]), (sched, c)


def annulus(M, N, r0=0.5, r1=3.0):
    th = 2.0 * np.pi * np.arange(M) / M
    r = r0 * (r1 / r0) ** (np.arange(N + 1) / N)
    return np.stack([r[:, None] * np.cos(th)[None, :], r[:, None] * np.sin(th)[None, :]],
                    axis=-1).astype(np.float32)


@pytest.mark.parametrize("math", ["exact", "fast"])
def test_batch_limiters_and_edge_shapes(math):
    """Every limiter (van Albada and first order take code specialised at
    compile time, minmod and van Leer the run-time path) and grids down to
    4 x 4 in one batch: each solve equals its single-grid run."""
    steps = 40
    cases = [
        (dict(M=4, N=4, mach=0.6, aoa=5.0, scheme=ausm.SLAU2, limiter=ausm.NOLIMITER), annulus(4, 4)),
        (dict(M=61, N=5, mach=0.7, aoa=3.0, scheme=ausm.AUSMUP, limiter=ausm.MINMOD), annulus(61, 5)),
        (dict(M=125, N=17, mach=0.5, aoa=-2.0, scheme=ausm.SLAU, limiter=ausm.VANLEER),
         annulus(125, 17)),
        (dict(M=96, N=32, mach=0.8, aoa=1.25, scheme=ausm.SLAU2, limiter=ausm.VANALBADA),
         grid(96, 32, 12)),
        (dict(M=150, N=41, mach=2.0, aoa=0.0, scheme=ausm.AUSMUP, limiter=ausm.NOLIMITER),
         grid(150, 41, 2412)),
    ]
    with ausm.Batch(max_solves=8, math=math) as b:
        ids = [b.submit(g, ausm.freestream(c["mach"], c["aoa"]), scheme=c["scheme"],
                        limiter=c["limiter"], max_steps=steps) for c, g in cases]
        b.advance(steps, check_every=20)
        for (c, g), sid in zip(cases, ids):
            assert b.poll(sid)["steps"] == steps
            got = b.state(sid)
            ref = single_run(c, g, steps, math)
            N = c["N"]
            assert np.array_equal(got[1:N + 1], ref[1:N + 1]), c


def test_alone_and_in_mix_identical():
    c = CASES[1]
    g = grid(c["M"], c["N"], c["code"])
    fs = ausm.freestream(c["mach"], c["aoa"])
    with ausm.Batch(max_solves=4) as b:
        a = b.submit(g, fs, scheme=c["scheme"], max_steps=80)
        b.advance(80, check_every=40)
        alone, alone_st = b.state(a), b.poll(a)
    with ausm.Batch(max_solves=8) as b:
        others = [b.submit(grid(x["M"], x["N"], x["code"]), ausm.freestream(x["mach"], x["aoa"]),
                           scheme=x["scheme"], max_steps=200) for x in CASES]
        a = b.submit(g, fs, scheme=c["scheme"], max_steps=80)
        b.advance(80, check_every=40)
        mixed, mixed_st = b.state(a), b.poll(a)
        assert all(b.poll(o)["state"] == "running" for o in others)
    assert np.array_equal(alone, mixed)
    assert alone_st["res_l2"] == mixed_st["res_l2"]


def test_convergence_and_forces():
    c = CASES[0]
    g = grid(c["M"], c["N"], c["code"])
    with ausm.Batch(max_solves=4) as b:
        sid = b.submit(g, ausm.freestream(0.5, 2.0), scheme=ausm.SLAU2, max_steps=20000,
                       res_drop=0.3, ref_length=1.0)
        pend = b.advance(20000, check_every=100)
        st = b.poll(sid)
        assert st["state"] == "converged" and pend == 0
        assert st["res_l2"] <= 0.3 * st["res_l2_0"]
        # forces at retirement equal the single-grid forces of the same state
        s = ausm.Solver(c["M"], c["N"])
        s.set_vertices(g)
        s.set_flow(0.5, 2.0)
        s.set_state(b.state(sid))
        f = s.forces(1.0)
        assert f["cl"] == st["cl"] and f["cd"] == st["cd"]
        s.close()


def test_failed_solve_is_isolated():
    """An unstable solve (CFL 20) blows up and is retired as failed; its
    neighbour is bitwise unaffected."""
    c = CASES[2]
    g = grid(c["M"], c["N"], c["code"])
    fs = ausm.freestream(c["mach"], c["aoa"])
    with ausm.Batch(max_solves=4) as b:
        good = b.submit(g, fs, scheme=c["scheme"], max_steps=400)
        bad = b.submit(g, fs, scheme=c["scheme"], cfl=20.0, max_steps=400)
        b.advance(400, check_every=20)
        assert b.poll(bad)["state"] == "failed"
        assert b.poll(good)["state"] == "capped"
        got = b.state(good)
    ref = single_run(dict(c, scheme=c["scheme"]), g, 400, "fast")
    assert np.array_equal(got[1:c["N"] + 1], ref[1:c["N"] + 1])


def test_admission_by_memory_budget():
    """A budget that holds one solve runs three submissions one after another."""
    c = CASES[3]
    g = grid(c["M"], c["N"], c["code"])
    fs = ausm.freestream(c["mach"], c["aoa"])
    per = 3 * 4 * 4 * 256 * 64 + 2 * 4 * 256 * 65 + 4 * 64 * 1024  # a little more than one solve
    with ausm.Batch(max_solves=8, hbm_budget_bytes=per) as b:
        ids = [b.submit(g, fs, scheme=c["scheme"], max_steps=40) for _ in range(3)]
        b.advance(40, check_every=40)
        states = [b.poll(i)["state"] for i in ids]
        assert states[0] == "capped" and states[1:] == ["queued", "queued"]
        b.release(ids[0])
        b.advance(40, check_every=40)
        assert b.poll(ids[1])["state"] == "capped"
        b.release(ids[1])
        b.advance(40, check_every=40)
        assert b.poll(ids[2])["state"] == "capped"
    with pytest.raises(ausm.AusmError):
        with ausm.Batch(max_solves=2, hbm_budget_bytes=1 << 16) as b:
            b.submit(g, fs)


def test_bad_descriptor_rejected():
    with ausm.Batch(max_solves=2) as b:
        g = grid(64, 20)
        with pytest.raises(ausm.AusmError):
            b.submit(g, (1.0, 1 / 1.4, 0.5, 0.0, 0.9))  # gamma <= 1
        with pytest.raises(ausm.AusmError):
            b.poll(12345)


def test_service_matches_sequential_solves():
    """Poisson arrivals into a 4-slot continuous-batching service: every
    request ends in the same state, step count, residual and Cl as when run
    alone through the single-grid step."""
    from ausm import serve

    reqs = serve.make_requests(6, seed=3, sizes=[(64, 20), (96, 32)], max_steps=400,
                               res_drop=0.5)
    serve.make_grids(reqs, iters=100, radius_chords=8.0)
    arr = serve.poisson_arrivals(6, rate=50.0, seed=1)
    a = serve.run_batched(reqs, arr, check_every=50, max_solves=4)
    s = serve.run_sequential(reqs, arr, check_every=50)
    for r in reqs:
        x, y = a["results"][r["id"]], s["results"][r["id"]]
        assert x["state"] == y["state"], (r, x, y)
        if x["state"] != "failed":
            assert x["steps"] == y["steps"] and x["res_l2"] == y["res_l2"]
            assert x["cl"] == y["cl"] and x["cd"] == y["cd"]
    assert a["p99_latency_s"] >= a["p50_latency_s"] > 0
