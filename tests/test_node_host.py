"""The Node.js host (N-API addon over the C ABI + the JS mirror of the
reference's main.js loop).  CPU: the addon loads in node and its geometry
matches the library's.  GPU: a frame loop driven from JavaScript (inflow
ramp, per-frame dt, scheme switch, async steps) ends bit-for-bit where the
same sequence through the Python binding ends."""
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "ausm-webgpu_amd", "node", "ausm_napi.node")

needs_node = pytest.mark.skipif(NODE is None, reason="node is not installed")


@needs_node
def test_addon_loads_and_geometry_matches():
    import ausm

    assert os.path.exists(ADDON), "build the addon first (__graft_entry__.build())"
    js = ("const a=require(process.argv[1]);"
          "console.log(JSON.stringify({v:a.version(),"
          "b:Array.from(a.boundaryNaca4(0.02,0.4,0.12,0.3,-0.7,1.4,0,64)),"
          "r:Array.from(a.boundaryRect(1.5,64))}))")
    out = json.loads(subprocess.check_output([NODE, "-e", js, ADDON], text=True))
    assert out["v"] == 1
    assert np.array_equal(np.array(out["b"]).reshape(64, 2),
                          ausm.naca4_boundary(64, code=2412, size=0.3, x_offset=-0.7))
    assert np.array_equal(np.array(out["r"]).reshape(64, 2), ausm.rect_boundary(64, 1.5))


def _oracle_par(act, xy, cfl=1.5, maxdt=1e-4):
    """The uniforms of one frame as the reference forms them on the host
    (src/main.js:779-784, util.js:76-99): inflow act * (cos, sin) of the AoA,
    inState with rhoE = p/(gamma-1) + act^2 rho/2, all in double, stored as
    f32."""
    from oracle_lib import P_COUNT, P_CFL, P_GAMMA, P_INP, P_INRHO, P_IS0, P_KP, P_KU, P_MAXDT

    g, rho, p = 1.4, 1.0, 1.0 / 1.4
    par = np.zeros(P_COUNT, dtype=np.float64)
    par[P_GAMMA], par[P_INRHO], par[P_INP] = g, rho, p
    vx, vy = act * xy[0], act * xy[1]
    par[3], par[4] = vx, vy
    par[P_IS0:P_IS0 + 4] = [rho, vx * rho, vy * rho, p / (g - 1.0) + 0.5 * (act * act) * rho]
    par[P_MAXDT], par[P_CFL], par[P_KP], par[P_KU] = maxdt, cfl, 0.0, 0.75
    return par.astype(np.float32)


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize("pause", [False, True])
def test_node_frame_loop_matches_python_host_and_oracle(tmp_path, pause):
    """The reference's interactive loop (src/main.js:714-792) driven from
    JavaScript: air at rest, the inflow ramped once per stepping frame, one
    dt per frame (refreshed from the frame's last state), a scheme switch
    mid-run, optionally two paused frames (no steps, no ramp) and an async
    last frame.  The final state is bit-for-bit the same sequence through
    the Python binding and bit-for-bit the FP32 oracle stepped frame by
    frame with the reference's per-frame uniforms."""
    import ausm
    from oracle_lib import AUSMUP, SLAU2 as O_SLAU2, VANALBADA, Solver as OSolver

    M, N, frames, spf = 96, 32, 8, 7
    out = tmp_path / "state.bin"
    args = [NODE, os.path.join(ROOT, "tests", "node_frames.js"), str(M), str(N), str(frames),
            str(spf), str(out)] + (["pause"] if pause else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout.strip().splitlines()[-1])
    q_js = np.fromfile(out, dtype=np.float32).reshape(N + 3, M, 4)
    paused = {2, 3} if pause else set()
    assert [i is None for i in js["infos"]] == [f in paused for f in range(frames)]

    # the same sequence through the Python binding
    s = ausm.Solver(M, N)
    s.update_grid_boundaries(ausm.naca4_boundary(M, code=12, size=0.3, x_offset=-0.7),
                             ausm.rect_boundary(M, 1.5))
    s.prepare_grid(iters=300)
    aoa = math.radians(2)
    xy = (math.cos(aoa), math.sin(aoa))
    act, scheme = 0.0, ausm.SLAU2

    def upload():
        s.set_freestream(1.0, 1.0 / 1.4, act * xy[0], act * xy[1], 1.4)
        s.set_scheme(scheme, ausm.VANALBADA, 0.0, 0.75)
        s.set_cfl(1.5, 1e-4, spf)
        s.set_math(ausm.MATH_EXACT)

    # and through the oracle: the reference's three-texture driver with dt
    # refreshed every spf steps (once per frame), uniforms per frame
    ref = OSolver(32, _oracle_par(0.0, xy), O_SLAU2, VANALBADA, s.vertices(), refresh=spf,
                  freestream=False)
    upload()
    s.prepare_state(freestream=False)
    for f in range(frames):
        if f == frames // 2:
            scheme = ausm.AUSMUP
            ref.scheme = AUSMUP
        if f in paused:
            continue
        upload()
        s.step(spf, info=False)
        ref.par[...] = _oracle_par(act, xy)
        ref.step(spf)
        act += (0.8 - act) / (128 / spf * 50)
    assert np.array_equal(s.state(), q_js)
    assert np.array_equal(q_js[1:N + 1], ref.state[1:N + 1])
    f = s.forces(0.6)
    assert f["cl"] == js["forces"]["cl"] and f["cd"] == js["forces"]["cd"]
    assert js["actualInflowVel"] == act
    s.close()


@needs_node
@pytest.mark.gpu
def test_node_rejects_arrays_of_the_wrong_length():
    """Typed arrays are checked against the context's own grid size before
    the library reads or writes them: a short state, vertex or boundary
    array, or a read-back asked for another size, is a RangeError."""
    js = r"""
const m = require(process.argv[1]);
const M = 32, N = 12;
const ctx = m.native.createContext(M, N, 0);
const cases = {
  state: () => m.native.setState(ctx, new Float32Array(4 * M * (N + 3) - 1)),
  vertices: () => m.native.setVertices(ctx, new Float32Array(2 * M * N)),
  body: () => m.native.gridFromBoundary(ctx, new Float64Array(2 * M - 2), new Float64Array(2 * M), 10, 0, 1.0),
  outer: () => m.native.gridFromBoundary(ctx, new Float64Array(2 * M), new Float64Array(2), 10, 0, 1.0),
  getState: () => m.native.getState(ctx, M + 1, N),
  residual: () => m.native.residual(ctx, M, N + 1),
  forces: () => m.native.forces(ctx, 1.0, M - 1),
  field: () => m.native.derivedField(ctx, 0, M, N - 1),
};
const out = {};
for (const [k, f] of Object.entries(cases)) {
  try { f(); out[k] = 'no error'; } catch (e) { out[k] = e.constructor.name; }
}
// the right sizes still work
m.native.gridFromBoundary(ctx, m.naca4Boundary(M, 12, 0.3, -0.7), m.rectangularOuterBoundary(M, 1.5), 10, 0, 1.0);
m.native.initState(ctx, 1);
out.ok = m.native.getState(ctx, M, N).length === 4 * M * (N + 3);
console.log(JSON.stringify(out));
"""
    r = subprocess.run([NODE, "-e", js, os.path.join(ROOT, "ausm-webgpu_amd", "node", "ausm.js")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout.strip().splitlines()[-1])
    assert o.pop("ok") is True
    assert all(v == "RangeError" for v in o.values()), o


@needs_node
def test_node_bodies_and_controls_match_library():
    """The JS object menu (src/grid.js:258-263) and the new generator
    bindings produce the library's points; pause and the adaptive
    steps-per-frame controller (main.js:720-731, util.js:192-201) behave as
    the reference's frame loop does, without a GPU (no context needed)."""
    import ausm

    dat = os.path.join(ROOT, "tests", "golden", "sc20012.dat")
    js = r"""
const m = require(process.argv[1]);
const fs = require('fs');
const M = 128;
const text = fs.readFileSync(process.argv[2], 'utf8');
const out = {
  polygon: Array.from(m.objectBoundary('polygon', M)),
  circle: Array.from(m.objectBoundary('circle', M)),
  sears: Array.from(m.objectBoundary('sears-haack', M)),
  haack: Array.from(m.objectBoundary('haack-projectile', M)),
  dat: Array.from(m.objectBoundary('airfoil-dat', M, text)),
  cst: Array.from(m.cstBoundary(M, [0.15, 0.15], [-0.1, -0.1])),
  bez: Array.from(m.bezierBoundary(M, [[0, 0], [0, 0.05], [0.5, 0.08], [1, 0]], [[0, 0], [0, -0.03], [1, 0]])),
};
// controller arithmetic without a context: a solver-like object
const S = Object.create(m.AusmSolver.prototype);
Object.assign(S, {autoStepsPerFrame: true, targetFrameTime: 16, _lastTotal: 5, _lastFrameStart: 0,
                  stepsPerFrame: 70, maxdt: 1e-4, _oldDt: null});
S._adapt(20);            // interval 20 ms: overshoot 4 ms, compute 5 ms < 14 ms
const a1 = S.stepsPerFrame;
S.togglePause();
const paused = !S.running;
S._adapt(40);            // paused: unchanged
const a2 = S.stepsPerFrame;
S.togglePause();
out.ctrl = [a1, paused, a2, S.running, S.maxdt];
console.log(JSON.stringify(out));
"""
    r = subprocess.run([NODE, "-e", js, os.path.join(ROOT, "ausm-webgpu_amd", "node", "ausm.js"), dat],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout.strip().splitlines()[-1])
    M = 128
    get = lambda k: np.array(o[k]).reshape(M, 2)
    assert np.array_equal(get("polygon"), ausm.polygon_boundary(M, 0.1, 4, 0.1, 0.0, -0.7))
    assert np.array_equal(get("circle"), ausm.polygon_boundary(M, 0.1, 400, 1.0, 0.0, -0.7))
    assert np.array_equal(get("sears"), ausm.sears_haack_boundary(M, 0.0002, 0.4, -0.7))
    assert np.array_equal(get("haack"), ausm.haack_projectile_boundary(M))
    with open(dat) as f:
        assert np.array_equal(get("dat"), ausm.dat_boundary(M, f.read(), 0.3, -0.7)[0])
    assert np.array_equal(get("cst"), ausm.cst_boundary(M, [0.15, 0.15], [-0.1, -0.1]))
    assert np.array_equal(get("bez"), ausm.bezier_boundary(
        M, [[0, 0], [0, 0.05], [0.5, 0.08], [1, 0]], [[0, 0], [0, -0.03], [1, 0]]))
    # 70 + 1 (under budget) - round(4 * 0.05) = 71; paused: unchanged; resumed
    assert o["ctrl"] == [71, True, 71, True, 1e-4]
