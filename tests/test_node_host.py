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


@needs_node
@pytest.mark.gpu
def test_node_frame_loop_matches_python_host(tmp_path):
    import ausm

    M, N, frames, spf = 96, 32, 6, 7
    out = tmp_path / "state.bin"
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node_frames.js"), str(M), str(N),
                        str(frames), str(spf), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout.strip().splitlines()[-1])
    q_js = np.fromfile(out, dtype=np.float32).reshape(N + 3, M, 4)

    # the same sequence through the Python binding (src/main.js:735-784)
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

    upload()
    s.prepare_state(freestream=False)
    for f in range(frames):
        if f == frames // 2:
            scheme = ausm.AUSMUP
        upload()
        s.step(spf, info=False)
        act += (0.8 - act) / (128 / spf * 50)
    assert np.array_equal(s.state(), q_js)
    f = s.forces(0.6)
    assert f["cl"] == js["forces"]["cl"] and f["cd"] == js["forces"]["cd"]
    assert js["actualInflowVel"] == act
    s.close()


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
