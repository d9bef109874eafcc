'use strict';
// GPU check of the JS SurfaceMesh surface (node/mesh.js): configure, the
// anisotropy pass, the isotropic field against a brute-force double sum
// over all particles at sampled vertices, the anisotropic field's total
// against the isotropic one (both integrate each particle's volume), and
// generate's marching-tetrahedra surface (counter, vertex buffer, unpacking).
// Prints one JSON line.
const path = require('path');
const root = path.join(__dirname, '..', '..', 'particles4all_amd', 'node');
const { Sim } = require(path.join(root, 'sim.js'));
const { SurfaceMesh } = require(path.join(root, 'mesh.js'));
const { loadPreset } = require(path.join(root, 'settingsini.js'));
const { Solids, halfExtent, shapeId } = require(path.join(root, 'solids.js'));

const out = {};
const { params } = loadPreset('small', 'particles=12000');
const sim = new Sim(0);
sim.reset(params);
for (let f = 0; f < 10; f++) sim.step(1 / 60);
// a grid built from the live positions (predict with dt = 0 keeps them), so
// the gather sees every particle within h
sim.runStage('predict', 0);
sim.runStage('grid', 0);
const mesh = new SurfaceMesh();
// four voxels per h, so the lattice sums approximate the fields' integrals
const res = Math.ceil(4 * Math.max(...sim.params.box) / sim.h);
mesh.configure(sim, [0, 0, 0], sim.params.box, res);
out.resolution = res;
out.vertDim = mesh.vertDim;
out.nVerts = mesh.nVerts;
out.fp64Solves = mesh.buildAnisotropy(sim);
mesh.generateField(sim);
const iso = mesh.field(sim);
const pos = sim.read('pos'), dens = sim.read('density'), body = sim.read('body');
const n = sim.n, h = sim.h, h2 = h * h;
const coef = 315 / (64 * Math.PI * Math.pow(h, 9));
const mass = sim.scene.mass, rho0 = sim.params.restDensity;
const skip = sim.nBodyParts > 0;
let worst = 0, checked = 0, seed = 7;
const vd = mesh.vertDim;
for (let s = 0; s < 4000 && checked < 60; s++) {
  seed = (seed * 1103515245 + 12345) & 0x7fffffff;
  const idx = seed % mesh.nVerts;
  if (!(iso[idx] > 1e-3)) continue;
  const x = idx % vd[0], y = Math.floor(idx / vd[0]) % vd[1], z = Math.floor(idx / (vd[0] * vd[1]));
  const p = [mesh.origin[0] + mesh.voxel * x, mesh.origin[1] + mesh.voxel * y,
             mesh.origin[2] + mesh.voxel * z];
  let acc = 0;
  for (let j = 0; j < n; j++) {
    if (skip && body[4 * j] !== 0) continue;
    const dx = p[0] - pos[4 * j], dy = p[1] - pos[4 * j + 1], dz = p[2] - pos[4 * j + 2];
    const r2 = dx * dx + dy * dy + dz * dz;
    if (r2 >= h2) continue;
    const t = h2 - r2;
    acc += mass / Math.max(dens[j], 0.5 * rho0) * coef * t * t * t;
  }
  worst = Math.max(worst, Math.abs(iso[idx] - acc) / acc);
  checked++;
}
out.checked = checked;
out.isoWorstRelError = worst;
let sumIso = 0, nonFinite = 0;
for (let i = 0; i < iso.length; i++) { sumIso += iso[i]; if (!Number.isFinite(iso[i])) nonFinite++; }
mesh.generateField(sim, { anisotropic: true });
const an = mesh.field(sim);
let sumAn = 0, maxAn = 0;
for (let i = 0; i < an.length; i++) {
  sumAn += an[i]; maxAn = Math.max(maxAn, an[i]); if (!Number.isFinite(an[i])) nonFinite++;
}
const v3 = Math.pow(mesh.voxel, 3);
out.isoIntegral = sumIso * v3;
out.anisoIntegral = sumAn * v3;
out.anisoMax = maxAn;
out.nonFinite = nonFinite;
// generate(sim, iso, smoothPasses): field, blur, normal field, march
mesh.normalSmoothPasses = 2;
mesh.generate(sim, 0.4, 1);
const tri = mesh.pollTriangles(sim);
out.triangles = tri.total;
out.emitted = tri.emitted;
out.lastTriangles = mesh.lastTriangles;
out.vertsBytes = mesh.verts.bytes;
out.vertsPtrNonZero = mesh.verts.ptr !== 0n;
const vw = mesh.vertices(sim);
out.vertWords = vw.length;
// unpack as the mesh vertex shader does (mesh_wgsl.js:271-283)
let outside = 0, badNormal = 0;
for (let k = 0; k < vw.length; k += 2) {
  const q = [vw[k] & 0xffff, vw[k] >>> 16, vw[k + 1] & 0xffff];
  for (let a = 0; a < 3; a++) {
    const w = mesh.origin[a] + q[a] / 65535 * (vd[a] - 1) * mesh.voxel;
    if (w < -mesh.voxel - 1e-6 || w > sim.params.box[a] + 2 * mesh.voxel + 1e-6) outside++;
  }
  const oct = vw[k + 1] >>> 16;
  const e = [(oct & 0xff) / 255 * 2 - 1, (oct >>> 8) / 255 * 2 - 1];
  if (Math.abs(e[0]) > 1 || Math.abs(e[1]) > 1) badNormal++;
}
out.vertsOutside = outside;
out.badNormal = badNormal;
// renderer exports: the per-slot phase on the device, and the packed poses
out.liveBodyBytes = sim.liveBody().bytes;
out.liveBodyPtrNonZero = sim.liveBody().ptr !== 0n;
const solids = new Solids();
solids.build(sim);
sim.step(1 / 60);
solids.pack(sim);
const pk = solids.read(sim);
const centre = sim.read('bodyCentre'), rot = sim.read('bodyRot');
let packErr = 0;
sim.scene.bodies.forEach((b, i) => {
  const o = i * 24, hx = halfExtent(b);
  for (let a = 0; a < 3; a++) {
    packErr += Math.abs(pk[o + a] - centre[i * 4 + a]);
    packErr += Math.abs(pk[o + 4 + a] - Math.fround(hx[a]));
    for (let r = 0; r < 3; r++) packErr += Math.abs(pk[o + 8 + 4 * a + r] - rot[r * 4 + a + i * 12]);
  }
  packErr += Math.abs(pk[o + 3] - shapeId(b.shape));
});
out.solidsCount = solids.count;
out.packErr = packErr;
out.packedBytes = solids.packed.bytes;
// snapshot: a second Sim loaded from it continues bit for bit
const snap = sim.snapshot();
const sim2 = new Sim(0);
sim2.loadSnapshot(snap, sim.params, sim.scene);
for (let f = 0; f < 3; f++) { sim.step(1 / 60); sim2.step(1 / 60); }
const pa = sim.read('pos'), pb = sim2.read('pos');
let snapDiff = pa.length === pb.length ? 0 : -1;
for (let k = 0; k < pa.length && snapDiff >= 0; k++) if (pa[k] !== pb[k]) snapDiff++;
out.snapshotBytes = snap.length;
out.snapshotDiff = snapDiff;
out.snapshotBodyPose = JSON.stringify(sim.bodyPose) === JSON.stringify(sim2.bodyPose);
console.log(JSON.stringify(out));
