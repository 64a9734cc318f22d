'use strict';
// Drives the Node.js host (ausm-webgpu_amd/node/ausm.js) through the
// reference's start-up and frame loop: NACA 0012 in the default rectangular
// far field, grid solve, air at rest, inflow ramped once per frame, a scheme
// switch mid-run (optionally a pause of two frames) and the last frame taken
// with stepAsync.  Writes the final
// state (Float32, reference layout) to argv[6] and prints a JSON summary.
const fs = require('fs');
const path = require('path');
const ausm = require(path.join(__dirname, '..', 'ausm-webgpu_amd', 'node', 'ausm.js'));

async function main() {
  const [M, N, frames, spf] = process.argv.slice(2, 6).map(Number);
  const out = process.argv[6];
  // optional: pause before frame 2 and resume before frame 4 (two frames
  // that neither step nor ramp the inflow, src/main.js:719-784)
  const pause = process.argv[7] === 'pause';
  const s = new ausm.AusmSolver({ M, N });
  s.updateGridBoundaries(ausm.naca4Boundary(M, 12, 0.3, -0.7), ausm.rectangularOuterBoundary(M, 1.5));
  s.prepareGrid(300);
  s.stepsPerFrame = spf;
  s.inflowVel = 0.8;
  s.setAoA(2);
  s.prepareState(ausm.INIT_REST);
  const infos = [];
  for (let f = 0; f < frames - 1; ++f) {
    if (pause && (f === 2 || f === 4)) s.togglePause();
    if (f === Math.floor(frames / 2)) s.updateSolver('AUSM+-up');
    infos.push(s.frame());
  }
  infos.push(await s.frameAsync());
  fs.writeFileSync(out, Buffer.from(s.state().buffer));
  console.log(JSON.stringify({ infos, forces: (({ fx, fy, cl, cd }) => ({ fx, fy, cl, cd }))(s.forces(0.6)),
                               actualInflowVel: s.actualInflowVel }));
}
main().catch((e) => { console.error(e); process.exit(1); });
