'use strict';
// The compute part of the reference's SurfaceMesh (src/mesh.js): the
// anisotropy pass (buildAnisotropy, mesh.js:207-261), the density field
// (configure 96-128, generateField 263-271), its blur (smoothField 273-291,
// buildNormalField 293-313) and the marching-tetrahedra surface (generate
// 315-331, pollTriangles 333-345).  Draw calls stay with the renderer: the
// packed vertices and the draw-indirect arguments are device buffers
// (`verts`, `indirect`: {ptr, bytes} descriptors) it can read in place.
const native = require('./native');
const { BUF } = require('./sim');

class SurfaceMesh {
  constructor() {
    this.maxTriangles = 3000000;     // mesh.js:20-21
    this.normalSmoothPasses = 0;
    this.lastTriangles = 0;
    // mesh.js:81-90 defaults
    this.anisoRatio = 4.0;
    this.anisoKs = 1.0;
    this.anisoStretch = 2.0;
    this.anisoLonely = 1.0;
    this.anisoMinNeighbours = 25;
    this.anisoLambda = 0.9;
    this.anisoRadiusScale = 2.0;
    this.anisoLimitToField = true;
    this.voxel = 0;
    this.vertDim = [0, 0, 0];
    this.nVerts = 0;
    this.origin = [0, 0, 0];
    this.lastFp64Solves = 0;
  }

  // configure(boxMin, boxMax, resolution, maxTriangles): the voxel-vertex
  // lattice and the triangle budget
  configure(sim, boxMin, boxMax, resolution, maxTriangles = this.maxTriangles) {
    native.fieldConfigure(sim.handle, boxMin, boxMax, resolution, maxTriangles);
    const f = native.fieldInfo(sim.handle);
    this.voxel = f.voxel;
    this.vertDim = f.vertDim;
    this.nVerts = f.nVerts;
    this.origin = f.origin;
    this.boxMin = boxMin;
    this.vertBudget = f.maxTriangles;
    this.verts = native.devicePtr(sim.handle, BUF.meshVerts);
    this.indirect = native.devicePtr(sim.handle, BUF.meshIndirect);
  }

  // buildAnisotropy(enc, sim): smoothed centres and G; minExtent follows
  // anisoLimitToField (mesh.js:238)
  buildAnisotropy(sim) {
    if (sim.n <= 0) return 0;
    this.lastFp64Solves = native.aniso(sim.handle, {
      radiusScale: this.anisoRadiusScale, lambda: this.anisoLambda, kr: this.anisoRatio,
      kn: this.anisoLonely, ks: this.anisoKs, stretch: this.anisoStretch,
      minExtent: this.anisoLimitToField && this.voxel > 0 ? 1.25 * this.voxel / sim.h : 0,
      neps: this.anisoMinNeighbours,
    });
    return this.lastFp64Solves;
  }

  // generateField(enc, sim, iso): the isotropic density field of
  // meshDensityWGSL; generateField(sim, {anisotropic: true}) splats the
  // last buildAnisotropy instead.  Bodies are skipped when the scene has
  // any (mesh.js:145).
  generateField(sim, opts = {}) {
    native.fieldGenerate(sim.handle, opts.anisotropic ? 1 : 0,
                         opts.skipBodies === undefined ? -1 : (opts.skipBodies ? 1 : 0));
  }

  // smoothField(enc, sim, iso, passes): blur the field in place
  smoothField(sim, iso, passes) {
    if (passes > 0) native.fieldSmooth(sim.handle, passes);
  }

  // buildNormalField(enc, sim, iso, passes): the blurred copy the march
  // differentiates for normals (none when passes <= 0)
  buildNormalField(sim, iso, passes) {
    this.normalFieldReady = passes > 0;
    native.meshNormalField(sim.handle, passes);
  }

  // generate(enc, sim, iso, smoothPasses): field, blur, normal field and
  // the marching-tetrahedra surface into `verts`; opts.anisotropic splats
  // the last buildAnisotropy instead of the isotropic field
  generate(sim, iso, smoothPasses = 0, opts = {}) {
    this.generateField(sim, opts);
    this.smoothField(sim, iso, smoothPasses);
    this.buildNormalField(sim, iso, this.normalSmoothPasses);
    native.meshMarch(sim.handle, iso);
  }

  // pollTriangles(): the triangle counter of the last generate (waits for
  // it; the reference reads it back a few frames late)
  pollTriangles(sim) {
    const t = native.meshTriangles(sim.handle);
    this.lastTriangles = t.total;
    return t;
  }

  // packed vertices of the written triangles (2 words each, 3 per
  // triangle) as a Uint32Array
  vertices(sim) {
    const t = native.meshTriangles(sim.handle);
    if (t.emitted <= 0) return new Uint32Array(0);
    return native.read(sim.handle, BUF.meshVerts, 'u32', t.emitted * 6);
  }

  // the field as a Float32Array, x fastest
  field(sim) { return native.read(sim.handle, BUF.field, 'f32', this.nVerts); }
  smoothPos(sim) { return native.read(sim.handle, BUF.smoothPos, 'f32', sim.n * 4); }
  aniso(sim) { return native.read(sim.handle, BUF.anisoG, 'f32', sim.n * 8); }
}

module.exports = { SurfaceMesh };
