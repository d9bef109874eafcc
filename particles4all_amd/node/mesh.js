'use strict';
// The simulation-side part of the reference's SurfaceMesh (src/mesh.js):
// the anisotropy pass (buildAnisotropy, mesh.js:207-261) and the density
// field (configure 96-128, generateField 263-271).  The marching-tetrahedra
// mesher and every draw call stay with the renderer.
const native = require('./native');
const { BUF } = require('./sim');

class SurfaceMesh {
  constructor() {
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

  // configure(boxMin, boxMax, resolution): the voxel-vertex lattice
  configure(sim, boxMin, boxMax, resolution) {
    native.fieldConfigure(sim.handle, boxMin, boxMax, resolution);
    const f = native.fieldInfo(sim.handle);
    this.voxel = f.voxel;
    this.vertDim = f.vertDim;
    this.nVerts = f.nVerts;
    this.origin = f.origin;
    this.boxMin = boxMin;
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

  // the field as a Float32Array, x fastest
  field(sim) { return native.read(sim.handle, BUF.field, 'f32', this.nVerts); }
  smoothPos(sim) { return native.read(sim.handle, BUF.smoothPos, 'f32', sim.n * 4); }
  aniso(sim) { return native.read(sim.handle, BUF.anisoG, 'f32', sim.n * 8); }
}

module.exports = { SurfaceMesh };
