'use strict';
// The callers of Sim that live in the reference's app shell, restated for a
// headless host: the pour hose (src/main.js:89-160), the box-wall driver
// (main.js:974-985) and the per-frame stepping loop (main.js:986-1000).
// `ui` carries the same fields the reference keeps on its ui object.

const POUR_SLACK = 1.25;

function pourLatticeAcross(width, spacing) {
  return Math.max(1, Math.round(width / Math.max(spacing, 1e-4)));
}

function pourRatePerSecond(ui, spacing) {
  const k = pourLatticeAcross(ui.pourWidth, spacing);
  return Math.floor(k * k * ui.pourSpeed / Math.max(spacing * POUR_SLACK, 1e-4));
}

// The hose's LCG (main.js:100-104); state lives on `ui.pourSeed`.
function pourJitter(ui, step) {
  ui.pourSeed = (ui.pourSeed * 1103515245 + 12345) & 0x7fffffff;
  return ((ui.pourSeed >>> 8 & 0xffff) / 65535 - 0.5) * 0.002 * step;
}

function makeUi(sim, opts) {
  opts = opts || {};
  return Object.assign({
    pouring: false, pourLeft: sim.pourBudget || 0, pourBudget: sim.pourBudget || 0,
    pourSpeed: 3.0, pourWidth: 0.24, pourHeight: 0.85, pourTilt: 0, pourExtruded: 0,
    pourNextLayer: 0, pourSeed: 22222, timeScale: 1, paused: false,
    boxBaseX: sim.params.box[0], boxScaleX: 1,
  }, opts);
}

// pourStep(sim, ui, simDt) (main.js:106-160)
function pourStep(sim, ui, simDt) {
  if (!ui.pouring || ui.pourLeft <= 0 || simDt <= 0) return 0;
  const p = sim.params;
  const d = p.spacing;
  const box = p.box;
  const step = d;
  const layerStep = d * POUR_SLACK;
  ui.pourExtruded += ui.pourSpeed * simDt;
  let layers = 0;
  for (let m = ui.pourNextLayer; m * layerStep <= ui.pourExtruded; ++m) ++layers;
  if (layers <= 0) return 0;
  layers = Math.min(layers, 32);
  const k = pourLatticeAcross(ui.pourWidth, d);
  const span = (k - 1) * step;
  const nozzle = [sim.h + d, ui.pourHeight * box[1], 0.5 * box[2]];
  const tilt = ui.pourTilt * Math.PI / 180;
  const jet = [ui.pourSpeed * Math.cos(tilt), -ui.pourSpeed * Math.sin(tilt), 0];
  const pos = [], vel = [];
  const g = p.gravity;
  for (let l = 0; l < layers && pos.length / 3 < ui.pourLeft; ++l) {
    const m = ui.pourNextLayer + l;
    const behind = ui.pourExtruded - m * layerStep;
    const t = behind / Math.max(ui.pourSpeed, 1e-6);
    const cx = nozzle[0] + jet[0] * t;
    const cy = nozzle[1] + jet[1] * t - 0.5 * g * t * t;
    const vy = jet[1] - g * t;
    if (cx >= box[0] - d || cy <= d) break;
    for (let iy = 0; iy < k && pos.length / 3 < ui.pourLeft; ++iy)
      for (let iz = 0; iz < k && pos.length / 3 < ui.pourLeft; ++iz) {
        const y = cy - 0.5 * span + iy * step + pourJitter(ui, step);
        if (y <= d || y >= box[1] - d) continue;
        pos.push(cx + pourJitter(ui, step), y,
                 nozzle[2] - 0.5 * span + iz * step + pourJitter(ui, step));
        vel.push(jet[0], vy, 0);
      }
  }
  ui.pourNextLayer += layers;
  if (!pos.length) return 0;
  const added = sim.appendFluid(pos, vel);
  ui.pourLeft = Math.max(0, ui.pourLeft - added);
  if (ui.pourLeft === 0) ui.pouring = false;
  return added;
}

// Box-wall driver (main.js:974-985): move wall x toward the target at
// most 2*dt per frame.
function boxWallStep(sim, ui, dt) {
  const want = ui.boxBaseX * ui.boxScaleX;
  const gap = want - sim.params.box[0];
  const step = Math.sign(gap) * Math.min(Math.abs(gap), 2 * dt);
  if (Math.abs(step) > 1e-6) {
    sim.resizeBox([sim.params.box[0] + step, sim.params.box[1], sim.params.box[2]]);
    return true;
  }
  return false;
}

// One rendered frame's simulation work (main.js:974-1000).
function advanceFrame(sim, ui, dt) {
  if (ui.paused) return 0;
  boxWallStep(sim, ui, dt);
  let poured = 0;
  if (ui.pouring && ui.pourLeft > 0) {
    const sub = 1 / 60 / Math.max(1, sim.params.substeps);
    let want = dt * ui.timeScale;
    for (let guard = 0; want > 1e-6 && guard < 64; ++guard) {
      const slice = Math.min(want, sub);
      sim.step(slice);
      poured += pourStep(sim, ui, sim.lastAdvanced || 0);
      want -= slice;
    }
  } else {
    sim.step(dt * ui.timeScale);
  }
  return poured;
}

module.exports = { POUR_SLACK, pourLatticeAcross, pourRatePerSecond, pourJitter, pourStep,
                   boxWallStep, advanceFrame, makeUi };
