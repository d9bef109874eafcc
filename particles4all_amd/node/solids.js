'use strict';
// The pose packing of the reference's Solids (src/solids.js): the static
// per-body data of build (107-145: half extents and shape id, colour) and
// pack (147-161, bodyPackWGSL solid_wgsl.js:3-26) from the live body poses.
// The body meshes and the draw call stay with the renderer, which reads
// `packed` ({ptr, bytes}: 6 float4 per body) in place.
const native = require('./native');
const { BUF } = require('./sim');

// solids.js:3-25
function bodyColor(shape) {
  if (shape === 'box') return [0.55, 0.72, 0.95];
  if (shape === 'sphere') return [0.60, 0.88, 0.66];
  if (shape === 'torus') return [0.96, 0.60, 0.62];
  return [1.0, 0.55, 0.10];
}

function shapeId(shape) {
  return shape === 'sphere' ? 1 : shape === 'torus' ? 2 : 0;
}

function halfExtent(b) {
  const s = b.size;
  if (b.shape === 'sphere') return [s, s, s];
  if (b.shape === 'torus') return [s, 0.4 * s, 0.0];
  return [s * 0.9, s * 0.7, s * 0.9];
}

class Solids {
  constructor() {
    this.count = 0;
    this.gen = 0;
    this.packed = null;
    this.stat = null;
  }

  // build(sim): static data of the scene's bodies
  build(sim) {
    const bodies = sim.scene.bodies || [];
    this.count = bodies.length;
    this.gen++;
    if (!this.count) { this.packed = null; return; }
    const stat = new Float32Array(this.count * 8);
    bodies.forEach((b, i) => {
      const h = halfExtent(b), c = bodyColor(b.shape);
      stat.set([h[0], h[1], h[2], shapeId(b.shape)], i * 8);
      stat.set([c[0], c[1], c[2], 0], i * 8 + 4);
    });
    this.stat = stat;
    native.bodyPack(sim.handle, stat, this.count);
    this.packed = native.devicePtr(sim.handle, BUF.bodyPacked);
  }

  // pack(enc, sim): the packed poses after a step
  pack(sim) {
    if (!this.count) return;
    native.bodyPack(sim.handle, null, this.count);
  }

  // the packed poses as a Float32Array (6 float4 per body)
  read(sim) {
    return native.read(sim.handle, BUF.bodyPacked, 'f32', this.count * 24);
  }
}

module.exports = { Solids, bodyColor, shapeId, halfExtent };
