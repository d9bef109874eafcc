'use strict';
// GPU check of the JS Sim surface on the small preset: the reference's
// verify harness (main.js:836-942: scan monotonic, scan total, CPU density
// on 200 particles, finite values, box bounds) plus the interactive calls
// (pour, box-wall move, ray impulse, body drag).  Prints one JSON line.
const path = require('path');
const root = path.join(__dirname, '..', '..', 'particles4all_amd', 'node');
const { Sim } = require(path.join(root, 'sim.js'));
const { loadPreset } = require(path.join(root, 'settingsini.js'));
const { advanceFrame, makeUi, pourStep } = require(path.join(root, 'driver.js'));

const out = {};
const { params } = loadPreset('small', 'particles=12000');
const sim = new Sim(0);
sim.reset(params);
out.n0 = sim.n;
out.cap = sim.cap;
out.nBodies = sim.nBodies;
out.bodyShapes = sim.bodies.map(b => b.shape);
out.sub = sim.step(1 / 60);
// the autotuner leaves the particles where they are and stepping goes on
{
  const before = Array.from(sim.read('pos'));
  out.tuned = sim.tune();
  const after = Array.from(sim.read('pos'));
  out.tuneKeepsState = before.length === after.length && before.every((v, i) => Object.is(v, after[i]));
  out.subAfterTune = sim.step(1 / 60);
}
// the density buffer of a finished step belongs to the last solver
// iteration's input; run predict, grid and one lambda so that pred and
// density refer to the same positions
const dt = 1 / 60 / params.substeps;
sim.runStage('predict', dt);
sim.runStage('grid', dt);
sim.runStage('lambda', dt);

// verify harness (main.js:836-942)
const n = sim.n, nCells = sim.nCells, h = sim.h, h2 = h * h;
const cs = sim.read('cellStart');
let bad = 0, prev = 0;
for (let i = 0; i <= nCells; i++) { if (cs[i] < prev) bad++; prev = cs[i]; }
out.scanMonotonic = bad === 0;
out.scanTotal = cs[nCells];
const pred = sim.read('pred'), dens = sim.read('density'), lam = sim.read('lambda');
const coef = 315 / (64 * Math.PI * Math.pow(h, 9));
const gd = sim.gridDim;
const cellIdx = (x, y, z) => {
  const c = [Math.min(gd[0] - 1, Math.max(0, Math.floor(x / h))),
             Math.min(gd[1] - 1, Math.max(0, Math.floor(y / h))),
             Math.min(gd[2] - 1, Math.max(0, Math.floor(z / h)))];
  return (c[2] * gd[1] + c[1]) * gd[0] + c[0];
};
// the grid is the one built this substep from the pre-solve positions; the
// CPU check uses the same cells (cellStart) and the final positions, and
// adds the wall samples the GPU density includes
const bpos = sim.read('bpos'), bcs = sim.read('bcellStart');
let worst = 0;
for (let s = 0; s < 200; s++) {
  const i = Math.floor(s * n / 200);
  const px = pred[i * 4], py = pred[i * 4 + 1], pz = pred[i * 4 + 2];
  let rho = sim.scene.mass * coef * h2 * h2 * h2;
  const c = [Math.min(gd[0] - 1, Math.max(0, Math.floor(px / h))),
             Math.min(gd[1] - 1, Math.max(0, Math.floor(py / h))),
             Math.min(gd[2] - 1, Math.max(0, Math.floor(pz / h)))];
  for (let dz = -1; dz <= 1; dz++) for (let dy = -1; dy <= 1; dy++) {
    const z = c[2] + dz, y = c[1] + dy;
    if (z < 0 || z >= gd[2] || y < 0 || y >= gd[1]) continue;
    const x0 = Math.max(c[0] - 1, 0), x1 = Math.min(c[0] + 1, gd[0] - 1);
    const rb = (z * gd[1] + y) * gd[0];
    for (let j = cs[rb + x0]; j < cs[rb + x1 + 1]; j++) {
      if (j === i) continue;
      const r2 = (px - pred[j * 4]) ** 2 + (py - pred[j * 4 + 1]) ** 2 + (pz - pred[j * 4 + 2]) ** 2;
      if (r2 >= h2) continue;
      const t = h2 - r2;
      rho += sim.scene.mass * coef * t * t * t;
    }
    for (let k = bcs[rb + x0]; k < bcs[rb + x1 + 1]; k++) {
      const r2 = (px - bpos[k * 4]) ** 2 + (py - bpos[k * 4 + 1]) ** 2 + (pz - bpos[k * 4 + 2]) ** 2;
      if (r2 >= h2 || r2 < 1e-12) continue;
      const t = h2 - r2;
      rho += bpos[k * 4 + 3] * coef * t * t * t;
    }
  }
  void cellIdx;
  worst = Math.max(worst, Math.abs(dens[i] - rho) / Math.max(1, rho));
}
out.worstDensityRelError = worst;
let nan = 0, lmax = -1e30;
for (let i = 0; i < n; i++) {
  if (!Number.isFinite(lam[i]) || !Number.isFinite(dens[i])) nan++;
  lmax = Math.max(lmax, lam[i]);
}
out.nonFinite = nan;
out.lambdaMax = lmax;
for (let it = 0; it < params.iterations; it++) {
  if (it > 0) sim.runStage('lambda', dt);
  sim.runStage('delta', dt);
  if (sim.nBodies) sim.runStage('bodySolve', dt);
}
sim.runStage('finish', dt);
const pos = sim.read('pos');
let outside = 0;
for (let i = 0; i < n; i++)
  for (let k = 0; k < 3; k++)
    if (pos[i * 4 + k] < 0.5 * params.spacing - 1e-6 ||
        pos[i * 4 + k] > params.box[k] - 0.5 * params.spacing + 1e-6) outside++;
out.outsideBox = outside;

// pour, box wall, impulse, drag through the JS surface
const ui = makeUi(sim, { pouring: true, pourSpeed: 4, pourWidth: 0.13, pourTilt: 25 });
let poured = 0;
for (let f = 0; f < 10; f++) poured += advanceFrame(sim, ui, 1 / 60);
out.poured = poured;
out.nAfterPour = sim.n;
ui.boxScaleX = 0.8;
for (let f = 0; f < 5; f++) advanceFrame(sim, Object.assign(ui, { pouring: false }), 1 / 60);
out.boxX = sim.params.box[0];
out.gen = sim.gen;
const v0 = sim.read('vel');
sim.applyRayImpulse([0.3, 0.3, -1], [0, 0, 1], [0, 5, 0], 0.2, 10);
const v1 = sim.read('vel');
let changed = 0;
for (let i = 0; i < sim.n; i++) if (v1[i * 4 + 1] !== v0[i * 4 + 1]) changed++;
out.impulsed = changed;
sim.holdBody(1, [0.7, 0.8, 0.5], 10, 4);
for (let f = 0; f < 3; f++) sim.step(1 / 60);
sim.releaseBody();
out.bodyCentre = sim.bodyPose[0].centre;
out.stats = sim.stats;
sim.destroy();
// neighbour masks through the JS surface: the default skin and skin 0 (full
// walks) give the same bits; the mask state reads back
{
  const a = new Sim(0), b = new Sim(0);
  a.reset(Object.assign({}, params, { neighbourMinParticles: 0 }));
  b.reset(Object.assign({}, params, { neighbourSkin: 0 }));
  for (let f = 0; f < 4; f++) { a.step(1 / 60); b.step(1 / 60); }
  const pa = a.read('pos'), pb = b.read('pos');
  out.masksSameBits = pa.length === pb.length && pa.every((v, i) => Object.is(v, pb[i]));
  out.neighbourStateLen = a.read('neighbourState').length;
  a.destroy();
  b.destroy();
}
console.log(JSON.stringify(out));
