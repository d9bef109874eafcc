'use strict';
// Headless counterparts of the Sim callers in the reference's app shell
// (behaviour of src/main.js:89-160 and 974-1000): a pour hose feeding
// Sim.appendFluid, a box wall moving toward its target, and the per-frame
// loop that steps the simulation around them.  State lives on a `ui`
// object with the reference's field names (pouring, pourLeft, pourSpeed,
// pourWidth, pourHeight, pourTilt, pourExtruded, pourNextLayer, pourSeed,
// timeScale, paused, boxBaseX, boxScaleX).

// The jet is emitted as square layers of particles one spacing apart
// across, LAYER_PITCH spacings apart along the jet.
const LAYER_PITCH = 1.25;
const MAX_LAYERS_PER_CALL = 32;
const MAX_SLICES_PER_FRAME = 64;

// Particles across the nozzle: its width in spacings, rounded, at least 1.
function pourLatticeAcross(width, spacing) {
  return Math.max(1, Math.round(width / Math.max(spacing, 1e-4)));
}

// Particles per second the hose delivers: a layer of k x k every
// LAYER_PITCH spacings of jet.
function pourRatePerSecond(ui, spacing) {
  const k = pourLatticeAcross(ui.pourWidth, spacing);
  return Math.floor(k * k * ui.pourSpeed / Math.max(spacing * LAYER_PITCH, 1e-4));
}

// Placement jitter of the hose, +-0.001 spacing: a 31-bit linear
// congruential generator (multiplier 1103515245, increment 12345, taken in
// double precision as JavaScript evaluates it), bits 8-23 as the fraction.
function pourJitter(ui, spacing) {
  ui.pourSeed = (ui.pourSeed * 1103515245 + 12345) & 0x7fffffff;
  const frac = ((ui.pourSeed >>> 8) & 0xffff) / 65535;
  return (frac - 0.5) * 0.002 * spacing;
}

function makeUi(sim, opts) {
  const budget = sim.pourBudget || 0;
  const ui = {
    pouring: false, pourLeft: budget, pourBudget: budget,
    pourSpeed: 3.0, pourWidth: 0.24, pourHeight: 0.85, pourTilt: 0, pourExtruded: 0,
    pourNextLayer: 0, pourSeed: 22222, timeScale: 1, paused: false,
    boxBaseX: sim.params.box[0], boxScaleX: 1,
  };
  return Object.assign(ui, opts || {});
}

// Layers whose place on the jet the extruded length has reached since the
// last emission (layer m sits m * pitch behind the jet's front).
function dueLayers(ui, pitch) {
  let due = 0;
  while ((ui.pourNextLayer + due) * pitch <= ui.pourExtruded) due++;
  return due;
}

// Centre and vertical velocity of layer m: a ballistic jet leaving the
// nozzle (one spacing inside the wall at x = h + d, at pourHeight of the
// box, mid-depth) at pourSpeed, tilted down by pourTilt degrees.
function layerState(sim, ui, m, pitch) {
  const p = sim.params;
  const tilt = ui.pourTilt * Math.PI / 180;
  const vx = ui.pourSpeed * Math.cos(tilt), vy0 = -ui.pourSpeed * Math.sin(tilt);
  const t = (ui.pourExtruded - m * pitch) / Math.max(ui.pourSpeed, 1e-6);
  return {
    x: sim.h + p.spacing + vx * t,
    y: ui.pourHeight * p.box[1] + vy0 * t - 0.5 * p.gravity * t * t,
    z: 0.5 * p.box[2],
    vx,
    vy: vy0 - p.gravity * t,
  };
}

// Advances the hose by simDt of simulated time and appends the layers that
// came due (at most the remaining budget).  Returns the particles added.
function pourStep(sim, ui, simDt) {
  if (!ui.pouring || ui.pourLeft <= 0 || simDt <= 0) return 0;
  const d = sim.params.spacing;
  const box = sim.params.box;
  const pitch = d * LAYER_PITCH;
  ui.pourExtruded += ui.pourSpeed * simDt;
  const due = Math.min(dueLayers(ui, pitch), MAX_LAYERS_PER_CALL);
  if (due === 0) return 0;
  const k = pourLatticeAcross(ui.pourWidth, d);
  const half = 0.5 * (k - 1) * d;
  const pos = [], vel = [];
  let count = 0;
  emit:
  for (let l = 0; l < due; l++) {
    if (count >= ui.pourLeft) break;
    const c = layerState(sim, ui, ui.pourNextLayer + l, pitch);
    // the jet has left the box: this layer and the later ones are dropped
    if (c.x >= box[0] - d || c.y <= d) break;
    for (let a = 0; a < k; a++) {
      for (let b = 0; b < k; b++) {
        if (count >= ui.pourLeft) break emit;
        const y = c.y - half + a * d + pourJitter(ui, d);
        if (y <= d || y >= box[1] - d) continue;
        const x = c.x + pourJitter(ui, d);
        const z = c.z - half + b * d + pourJitter(ui, d);
        pos.push(x, y, z);
        vel.push(c.vx, c.vy, 0);
        count++;
      }
    }
  }
  ui.pourNextLayer += due;
  if (count === 0) return 0;
  const added = sim.appendFluid(pos, vel);
  ui.pourLeft = Math.max(0, ui.pourLeft - added);
  if (ui.pourLeft === 0) ui.pouring = false;
  return added;
}

// Moves the box's x wall toward boxBaseX * boxScaleX by at most 2 dt.
// Returns whether it moved.
function boxWallStep(sim, ui, dt) {
  const box = sim.params.box;
  const gap = ui.boxBaseX * ui.boxScaleX - box[0];
  const move = Math.sign(gap) * Math.min(Math.abs(gap), 2 * dt);
  if (Math.abs(move) <= 1e-6) return false;
  sim.resizeBox([box[0] + move, box[1], box[2]]);
  return true;
}

// Lengths of the slices a frame of `total` seconds is stepped in while the
// hose runs: at most one substep of a 1/60 s frame each, at most 64.
function frameSlices(total, substeps) {
  const most = 1 / 60 / Math.max(1, substeps);
  const out = [];
  let left = total;
  while (left > 1e-6 && out.length < MAX_SLICES_PER_FRAME) {
    const s = Math.min(left, most);
    out.push(s);
    left -= s;
  }
  return out;
}

// One rendered frame of simulation work: the wall, then the frame stepped
// whole, or in slices with the hose fed after each while it pours.
// Returns the particles poured.
function advanceFrame(sim, ui, dt) {
  if (ui.paused) return 0;
  boxWallStep(sim, ui, dt);
  const total = dt * ui.timeScale;
  if (!(ui.pouring && ui.pourLeft > 0)) {
    sim.step(total);
    return 0;
  }
  let poured = 0;
  for (const s of frameSlices(total, sim.params.substeps)) {
    sim.step(s);
    poured += pourStep(sim, ui, sim.lastAdvanced || 0);
  }
  return poured;
}

module.exports = { LAYER_PITCH, pourLatticeAcross, pourRatePerSecond, pourJitter, pourStep,
                   boxWallStep, advanceFrame, frameSlices, makeUi };
