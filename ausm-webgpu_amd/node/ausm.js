'use strict';
// Node.js host for the MI355X solver step, keeping the reference's host
// surface (huj31415/ausm-webgpu src/main.js, src/grid.js, src/util.js):
// updateGridBoundaries -> prepareGrid -> prepareState -> frame loop, with
// updateSolver to swap the flux scheme.  The compute runs in libausm through
// the Node-API addon (ausm_napi.node); nothing here does arithmetic on the
// solution.
const path = require('path');

const native = require(path.join(__dirname, 'ausm_napi.node'));

// flux scheme table, in the reference order (src/util.js:114-118)
const solvers = Object.freeze({ 'SLAU2': 0, 'SLAU': 1, 'AUSM+-up': 2 });
const limiters = Object.freeze({ vanAlbada: 0, minmod: 1, vanLeer: 2, none: 3, vanAlbadaClamped: 4 });
const INIT_REST = 0;
const INIT_FREESTREAM = 1;

const now = () => Number(process.hrtime.bigint()) / 1e6;  // ms

function toF64Pairs(coords) {
  if (coords instanceof Float64Array) return coords;
  const out = new Float64Array(coords.length * 2);
  coords.forEach((p, i) => { out[2 * i] = p[0]; out[2 * i + 1] = p[1]; });
  return out;
}

// generateNACA4Boundary (src/grid.js:38-52) for all M points at once
function naca4Boundary(M, naca4, size = 1.0, xOffset = 0, pointDistExponent = 7 / 5, spacing = 0) {
  const t = naca4 % 100 / 100;
  const m = Math.floor(naca4 / 1000) / 100;
  const p = Math.floor(naca4 / 100) % 10 / 10;
  return native.boundaryNaca4(m, p, t, size, xOffset, pointDistExponent, spacing, M);
}
const circularOuterBoundary = (M, radius = 1) => native.boundaryCircle(radius, M);
const rectangularOuterBoundary = (M, lengthRatio = 1.5) => native.boundaryRect(lengthRatio, M);
// generatePolygonBoundary, generateSearsHaackBoundary, generateHaackProjectileBoundary
// (src/grid.js:53-231) and sampleAirfoil (src/airfoilDatParser.js:276-293)
const polygonBoundary = (M, R = 0.2, N = 4, F = 0.1, a = 0, xOffset = -0.7) =>
  native.boundaryPolygon(R, N, F, a, xOffset, M);
const searsHaackBoundary = (M, V = 0.2, L = 10, xOffset = -0.7) =>
  native.boundarySearsHaack(V, L, xOffset, M);
const haackProjectileBoundary = (M, F = 6, R = 0.03, L = 0.5, Rb = 0.5, Lb = 0.2, C = 0, xOffset = -0.7) =>
  native.boundaryHaackProjectile(F, R, L, Rb, Lb, C, xOffset, M);
const datBoundary = (M, datText, scale = 1.0, xOffset = 0.0) => native.boundaryDat(datText, scale, xOffset, M);
const cstBoundary = (M, aUpper, aLower, teGap = 0, size = 0.5, xOffset = 0, exponent = 7 / 5, spacing = 0) =>
  native.boundaryCst(Float64Array.from(aUpper), Float64Array.from(aLower), teGap, size, xOffset, exponent,
                     spacing, M);

// upper / lower: control polygons [[x, y], ...] from the shared leading edge
const bezierBoundary = (M, upper, lower, size = 0.5, xOffset = 0, exponent = 7 / 5, spacing = 0) =>
  native.boundaryBezier(Float64Array.from(upper.flat()), Float64Array.from(lower.flat()), size, xOffset,
                        exponent, spacing, M);

// the object menu of the reference with its settings (src/grid.js:258-263)
function objectBoundary(type, M, datText) {
  switch (type) {
    case 'polygon': return polygonBoundary(M, 0.1, 4, 0.1, 0, -0.7);
    case 'circle': return polygonBoundary(M, 0.1, 400, 1, 0, -0.7);
    case 'naca-4': return naca4Boundary(M, 4412, 0.2, -0.7);
    case 'airfoil-dat': return datBoundary(M, datText, 0.3, -0.7);
    case 'sears-haack': return searsHaackBoundary(M, 0.0002, 0.4, -0.7);
    case 'haack-projectile': return haackProjectileBoundary(M);
    default: throw new Error(`unknown object type ${type}`);
  }
}

// display modes of the visualization pass (src/shaders/ausm.js:586-710)
const fields = Object.freeze({
  schlieren: 0, vorticity: 1, density: 2, temperature: 3, pressure: 4, mach: 5, machInflow: 6,
  velocityRel: 7, entropy: 8, totalPressure: 9,
});
const precisions = Object.freeze({ fp32: 0, fp64: 1, fp32c: 2 });

class AusmSolver {
  // simulationDomain = [M, N] (src/grid.js:12)
  constructor({ M, N, device = 0 }) {
    this.M = M;
    this.N = N;
    this.ctx = native.createContext(M, N, device);
    // physics and solver defaults (src/util.js:76-99, src/main.js:874-887)
    this.gamma = 1.4;
    this.inPressure = 1.0 / this.gamma;
    this.inRho = 1.0;
    this.maxdt = 1e-4;
    this.cflFactor = 1.5;
    this.K_p = 0;
    this.K_u = 0.75;
    this.muscl = true;
    this.limiter = 'vanAlbada';
    this.inflowVel = 3.8;
    this.actualInflowVel = 0;
    this.velRampUpStrength = 128;
    this.xyAoA = [1, 0];
    this.stepsPerFrame = 70;
    this.solver = 'SLAU2';
    this.mathMode = 0;
    // adaptive steps per frame (main.js:720-731)
    this.autoStepsPerFrame = false;
    this.targetFrameTime = 1000 / 60;
    this._lastFrameStart = 0;
    this._lastTotal = 0;
    this._oldDt = null;
    this._body = null;
    this._outer = null;
  }

  setAoA(deg) {
    const r = deg * Math.PI / 180;
    this.xyAoA = [Math.cos(r), Math.sin(r)];
  }

  // updateGridBoundaries (src/grid.js:267-296): body row and far-field row
  updateGridBoundaries(objCoords, boundCoords) {
    this._body = toF64Pairs(objCoords);
    this._outer = toF64Pairs(boundCoords);
  }

  // prepareGrid (src/main.js:658-707): interpolation, elliptic smoothing and
  // metrics on the GPU (FP64); the reference runs 1e4 Jacobi sweeps
  prepareGrid(iterations = 1e4) {
    native.gridFromBoundary(this.ctx, this._body, this._outer, iterations, 0, 1.0);
  }

  _uploadUniforms() {
    const v = this.actualInflowVel;
    native.setFreestream(this.ctx, this.inRho, this.inPressure, v * this.xyAoA[0], v * this.xyAoA[1],
                         this.gamma);
    native.setScheme(this.ctx, solvers[this.solver], this.muscl ? limiters[this.limiter] : limiters.none,
                     this.K_p, this.K_u);
    native.setCfl(this.ctx, this.cflFactor, this.maxdt, this.stepsPerFrame);
    native.setMath(this.ctx, this.mathMode);
  }

  // prepareState (src/main.js:631-655): air at rest, ramp restarted
  prepareState(mode = INIT_REST) {
    this._uploadUniforms();
    native.initState(this.ctx, mode);
    this.actualInflowVel = 0;
  }

  // updateSolver (src/main.js:629): swap the flux scheme, keep the state
  updateSolver(name) {
    if (!(name in solvers)) throw new Error(`unknown solver ${name}`);
    this.solver = name;
    native.setScheme(this.ctx, solvers[name], this.muscl ? limiters[this.limiter] : limiters.none,
                     this.K_p, this.K_u);
  }

  // body row from the reference's object menu (src/grid.js:258-263)
  setObject(type, datText) { this._body = objectBoundary(type, this.M, datText); }

  // play / pause and restart (src/util.js:192-201): pausing sets maxdt to 0,
  // which render() treats as "do not step" (main.js:719)
  togglePause() {
    if (this._oldDt !== null) {
      this.maxdt = this._oldDt;
      this._oldDt = null;
    } else {
      this._oldDt = this.maxdt;
      this.maxdt = 0;
    }
  }
  get running() { return this.maxdt > 0; }
  restart() { this.prepareState(); }

  // adaptive stepsPerFrame (main.js:720-731): grow by one while the last
  // frame's compute time stays 2 ms under the target, shrink by 5 % of the
  // frame-interval overshoot (at most 100 ms)
  _adapt(start) {
    if (this.autoStepsPerFrame && this._lastTotal > 0 && this.running) {
      const timeDifference = (start - this._lastFrameStart) - this.targetFrameTime;
      if (this._lastTotal < this.targetFrameTime - 2) this.stepsPerFrame += 1;
      this.stepsPerFrame = Math.max(1, Math.round(this.stepsPerFrame - Math.min(100, timeDifference) * 0.05));
    }
    this._lastFrameStart = start;
  }

  // One frame of render() (src/main.js:735-784): upload the uniforms, take
  // stepsPerFrame steps with one dt for the frame, then ramp the inflow for
  // the next frame.
  frame() {
    const start = now();
    this._adapt(start);
    this._uploadUniforms();
    let info = null;
    if (this.running) info = native.step(this.ctx, this.stepsPerFrame);
    this._lastTotal = now() - start;
    if (this.running) this._ramp();  // only frames that step ramp (main.js:742-784)
    return info;
  }

  async frameAsync() {
    const start = now();
    this._adapt(start);
    this._uploadUniforms();
    let info = null;
    if (this.running) info = await native.stepAsync(this.ctx, this.stepsPerFrame);
    this._lastTotal = now() - start;
    if (this.running) this._ramp();
    return info;
  }

  setPrecision(name) { native.setPrecision(this.ctx, precisions[name]); }
  setDtMode(local) { native.setDtMode(this.ctx, local ? 1 : 0); }
  field(kind) { return native.derivedField(this.ctx, typeof kind === 'string' ? fields[kind] : kind, this.M, this.N); }

  _ramp() {
    this.actualInflowVel += (this.inflowVel - this.actualInflowVel) /
                            (this.velRampUpStrength / this.stepsPerFrame * 50);
  }

  step(n) { return native.step(this.ctx, n); }
  stepAsync(n) { return native.stepAsync(this.ctx, n); }
  state() { return native.getState(this.ctx, this.M, this.N); }
  setState(q) { native.setState(this.ctx, q); }
  vertices() { return native.getVertices(this.ctx, this.M, this.N); }
  setVertices(v) { native.setVertices(this.ctx, v); }
  residual() { return native.residual(this.ctx, this.M, this.N); }
  forces(refLength = 1.0) { return native.forces(this.ctx, refLength, this.M); }
}

module.exports = {
  AusmSolver, solvers, limiters, fields, precisions, naca4Boundary, circularOuterBoundary,
  rectangularOuterBoundary, polygonBoundary, searsHaackBoundary, haackProjectileBoundary, datBoundary,
  cstBoundary, bezierBoundary, objectBoundary, INIT_REST, INIT_FREESTREAM, native,
};
