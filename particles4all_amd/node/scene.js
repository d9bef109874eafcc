'use strict';
// Scene construction for the JS host: the reference's src/scene.js surface
// (PRESETS, defaultParams, blockCapacity, boundaryFor, buildScene) with the
// O(n * 27 cells) work done natively by libp4a (p4a_scene_build,
// p4a_boundary_build), which restates the same algorithm in FP64.

const native = require('./native');

// PRESETS (scene.js:3-9); the ini presets ship in ../presets.
const PRESETS = {
  small: { ini: 'presets/small.ini' },
  medium: { ini: 'presets/medium.ini' },
  large: { ini: 'presets/large.ini' },
  extreme: { particles: 1000000, box: [3.0, 3.0, 3.0], substeps: 4, iterations: 4,
             bodies: ['sphere:0.5', 'torus:0.8', 'box:1.4'] },
};

// defaultParams(preset) (scene.js:11-39)
function defaultParams(preset) {
  if (preset === undefined) preset = 'small';
  const p = PRESETS[preset] || {};
  const out = {
    preset,
    spacing: 0.02,
    restDensity: 1000,
    gravity: 9.81,
    substeps: 4,
    iterations: 4,
    cfmEpsilonRel: 0.01,
    sCorrK: 0.1,
    sCorrDq: 0.3,
    xsphC: 0.05,
    omega: 1.0,
    sorAverage: false,
    surfaceTensionK: 0,
    targetParticleCount: 60000,
    box: [1.5, 1.5, 1.0],
    bodies: ['sphere:0.5', 'torus:0.8', 'box:1.4'],
    bodySize: 0,
  };
  if (p.substeps !== undefined) out.substeps = p.substeps;
  if (p.iterations !== undefined) out.iterations = p.iterations;
  if (p.particles !== undefined) out.targetParticleCount = p.particles;
  if (p.box !== undefined) out.box = p.box.slice();
  if (p.bodies !== undefined) out.bodies = p.bodies.slice();
  return Object.assign(out, p.params || {});
}

// blockCapacity (scene.js:66-72)
function blockCapacity(p, d) {
  const bx = p.box[0], by = p.box[1], bz = p.box[2];
  const margin = d;
  return Math.max(1, Math.floor(bx * 0.35 / d)) *
         Math.max(1, Math.floor((bz - 2 * margin) / d)) *
         Math.max(1, Math.floor((by - 2 * margin) / d));
}

// boundaryFor (scene.js:74) -> {pts, psi, count}
function boundaryFor(p, d, h) {
  return native.boundaryBuild(p.box, d, h, p.restDensity);
}

const SHAPES = { sphere: 0, torus: 1 };

// 'shape:density:startY' strings (scene.js:167-171)
function parseBodies(list) {
  return (list || []).map(s => {
    const bits = String(s).split(':');
    const shape = bits[0];
    const density = Number(bits[1] !== undefined ? bits[1] : 0.5) || 0.5;
    const startY = Number(bits[2] !== undefined ? bits[2] : 0.75) || 0.75;
    return { name: shape, shape: SHAPES[shape] !== undefined ? SHAPES[shape] : 2,
             density, startY };
  });
}

// buildScene(p) (scene.js:159-315): returns the reference's scene object
// {n, nFluid, nBody, pos, vel, rest, phase, boundary, bodies, mass, h, d,
// denomRest} plus the native handle that Sim.reset uploads from.
function buildScene(p) {
  const specs = parseBodies(p.bodies);
  const sc = native.sceneBuild(p, { kind: 0, bodies: specs });
  const bx = p.box[0], by = p.box[1], bz = p.box[2];
  const sz = p.bodySize > 0 ? p.bodySize : 0.12 * Math.min(bx, by, bz);
  // bodies that sampled no particle are dropped (scene.js:183); b.spec
  // names the spec each kept body came from
  for (const b of sc.bodies) {
    b.shape = specs[b.spec].name;
    b.size = sz;
  }
  if (p.verbose) console.info(`boundary ${sc.boundary.count} samples`);
  return sc;
}

// A jittered lattice scene (synthetic workloads): spec = {lattice: [nx,ny,nz],
// origin, jitter, seed, calibrateUnjittered, bodies: [{shape, size,
// density, centre}]} with shapes 3 = lattice cube, 4 = lattice sphere.
function buildLatticeScene(p, spec) {
  return native.sceneBuild(p, Object.assign({ kind: 1 }, spec));
}

module.exports = { PRESETS, defaultParams, blockCapacity, boundaryFor, buildScene,
                   buildLatticeScene, parseBodies };
