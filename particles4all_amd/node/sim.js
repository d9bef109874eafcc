'use strict';
// class Sim - the reference's simulation object (src/sim.js:9-814) over the
// MI355X library.  Same constructor, methods and fields; the GPU buffers it
// owns are device allocations of libp4a, exposed as {ptr (BigInt), bytes}
// descriptors instead of WebGPU buffers.

const native = require('./native');
const { buildScene } = require('./scene');

// buffer ids of include/p4a.h (enum p4a_buf)
const BUF = {
  pos: 0, vel: 1, pred: 2, body: 3, rest: 4, lambda: 5, density: 6, slot: 7, corr: 8,
  normal: 9, cellStart: 10, bpos: 11, bcellStart: 12, bodyCentre: 13, bodyRot: 14,
  pid: 15, key: 16, order: 17, bpsi: 18, smoothPos: 19, anisoG: 20, ofs: 21, cellOfSlot: 22,
  drift: 23, field: 24, fieldNormal: 25, meshVerts: 26, meshIndirect: 27, bodyPacked: 28,
  neighbourState: 29,
};
const STAGE = { predict: 0, grid: 1, bodyPre: 2, lambda: 3, delta: 4, bodySolve: 5,
                finish: 6, stats: 7 };

class Sim {
  // new Sim(device) (sim.js:10-18); `device` is a HIP device ordinal.
  constructor(device) {
    this.dev = typeof device === 'number' ? device : 0;
    this.handle = native.create(this.dev);
    this.buf = {};
    this.stats = { avgRho: 1, maxRho: 1, maxSpeed: 0, frames: 0 };
    this.simTime = 0;
    this.gen = 0;
    this.parity = 0;
    this.timer = null;
    this.bodyPhases = false;
    this.bodyPose = [];
  }

  // reset(params) (sim.js:66-221)
  reset(params) {
    this.params = params;
    const sc = buildScene(params);
    this.resetWithScene(params, sc);
  }

  // reset from an already built scene (buildScene or buildLatticeScene)
  resetWithScene(params, sc) {
    this.params = params;
    this.scene = sc;
    native.reset(this.handle, params, sc);
    this._refresh();
    this.bodies = sc.bodies;
    this.bodyPose = sc.bodies.map(b => ({
      centre: [b.centre[0], b.centre[1], b.centre[2]],
      rot: [1, 0, 0, 0, 1, 0, 0, 0, 1],
    }));
    this.heldBody = -1;
  }

  // The whole simulation state as a Uint8Array (p4a_snapshot_save): a sim
  // loaded from it continues bit for bit.  params and the scene's body
  // list are JS-side and are not part of it; keep them beside it.
  snapshot() { return native.snapshotSave(this.handle); }

  // Reset from snapshot bytes.  `params` and `scene` (the JS objects the
  // saved sim was reset with) replace this.params / this.scene when given;
  // step() needs params.
  loadSnapshot(bytes, params, scene) {
    native.snapshotLoad(this.handle, bytes);
    if (params) this.params = params;
    if (scene) {
      this.scene = scene;
      this.bodies = scene.bodies;
    }
    this._refresh();
    this.bodyPose = [];
    if (this.nBodies) {
      const pose = native.bodyPose(this.handle, this.nBodies);
      for (let i = 0; i < this.nBodies; i++)
        this.bodyPose.push({ centre: Array.from(pose.centre.slice(i * 3, i * 3 + 3)),
                             rot: Array.from(pose.rot.slice(i * 9, i * 9 + 9)) });
    }
    this.heldBody = -1;
  }

  _refresh() {
    const i = native.info(this.handle);
    this.n = i.n;
    this.cap = i.cap;
    this.h = i.h;
    this.gridDim = i.gridDim;
    this.nCells = i.nCells;
    this.nBoundary = i.nBoundary;
    this.nBodies = i.nBodies;
    this.nBodyParts = i.nBodyParts;
    this.pourBudget = i.pourBudget;
    this.parity = i.parity;
    this.predParity = i.parity;
    this.gen = i.gen;
    this.simTime = i.simTime;
    this.lastAdvanced = i.lastAdvanced;
    this.lastSubsteps = i.lastSubsteps;
    this.timeBank = i.timeBank;
    this.bodyScale = i.bodyScale;
    this.covScale = i.covScale;
    const buf = {};
    for (const k of Object.keys(BUF)) {
      if (k === 'rest') continue;  // synthesised on read
      // buf.body: asking for its device pointer keeps the per-slot phase on
      // the device, current after every step (what the SSFR splat binds)
      try { buf[k] = native.devicePtr(this.handle, BUF[k]); } catch (e) { /* empty buffer */ }
    }
    this.buf = buf;
    return i;
  }

  // uploadParams (sim.js:371-413): live parameters are re-read each step
  uploadParams() { native.setParams(this.handle, this.params); }

  // step(frameDt) (sim.js:415-656).  stats and bodyPose are read back at
  // the end of the step (the reference fills them a frame or more later).
  // Autotuner (no counterpart in the reference): times the neighbour
  // kernels' traversal shapes on the live scene and keeps the fastest for
  // later steps; returns the shapes chosen.  The particle state is unchanged.
  tune() {
    return native.tune(this.handle);
  }

  step(frameDt) {
    native.setParams(this.handle, this.params);
    const sub = native.step(this.handle, frameDt);
    this._refresh();
    if (sub >= 1) {
      const s = native.stats(this.handle);
      this.stats.avgRho = s.avgRho;
      this.stats.maxRho = s.maxRho;
      this.stats.maxSpeed = s.maxSpeed;
      this.stats.ke = s.ke;
      this.stats.frames += 1;
      if (this.nBodies) {
        const pose = native.bodyPose(this.handle, this.nBodies);
        for (let i = 0; i < this.nBodies; i++) {
          const p = this.bodyPose[i];
          for (let k = 0; k < 3; k++) p.centre[k] = pose.centre[i * 3 + k];
          for (let k = 0; k < 9; k++) p.rot[k] = pose.rot[i * 9 + k];
        }
      }
    }
    return sub;
  }

  // applyRayImpulse(origin, dir, impulse, radius, speedLimit) (sim.js:658-676)
  applyRayImpulse(origin, dir, impulse, radius, speedLimit) {
    native.applyRayImpulse(this.handle, origin, dir, impulse, radius, speedLimit);
  }

  // holdBody / releaseBody (sim.js:678-685)
  holdBody(body1Based, target, rate, limit) {
    this.heldBody = (body1Based >= 1 && body1Based <= this.nBodies) ? body1Based : -1;
    this.heldTarget = target;
    this.heldRate = rate;
    this.heldLimit = limit;
    native.holdBody(this.handle, body1Based, target, rate, limit);
  }

  releaseBody() {
    this.heldBody = -1;
    native.releaseBody(this.handle);
  }

  // resizeBox(newBox) (sim.js:687-772)
  resizeBox(newBox) {
    const box = [newBox[0], newBox[1], newBox[2]];
    native.resizeBox(this.handle, box);
    for (let a = 0; a < 3; a++) this.params.box[a] = box[a];
    this._refresh();
  }

  // appendFluid(pos, vel) -> number appended (sim.js:774-809)
  appendFluid(pos, vel) {
    const k = native.appendFluid(this.handle, pos, vel);
    this._refresh();
    return k;
  }

  livePos() { return this.buf.pos; }
  liveVel() { return this.buf.vel; }
  liveBody() { return this.buf.body; }

  // readback of a buffer as a typed array (the reference's __readBuf,
  // main.js:810-820); particle arrays are in the live order.
  read(name) {
    const n = this.n, nc = this.nCells;
    const shape = {
      pos: ['f32', n * 4], vel: ['f32', n * 4], pred: ['f32', n * 4], body: ['u32', n * 4],
      rest: ['f32', n * 4], lambda: ['f32', n], density: ['f32', n], slot: ['u32', n],
      corr: ['f32', n * 4], normal: ['f32', n * 4], cellStart: ['u32', nc + 2],
      bpos: ['f32', this.nBoundary * 4], bcellStart: ['u32', nc + 2],
      bodyCentre: ['f32', this.nBodies * 4], bodyRot: ['f32', this.nBodies * 12],
      pid: ['u32', n], key: ['u32', n], order: ['u32', n], bpsi: ['f32', this.nBoundary],
      // FP16-offset records as raw 32-bit words: float x, then half y | half z
      ofs: ['u32', n * 2], cellOfSlot: ['u32', n],
      // neighbour masks: travel bound (float bits), then the queue counts
      neighbourState: ['u32', 1025],
    }[name];
    if (!shape) throw new Error(`unknown buffer ${name}`);
    return native.read(this.handle, BUF[name], shape[0], shape[1]);
  }

  // one kernel group of a substep (parity harness)
  runStage(name, dt) {
    native.runStage(this.handle, STAGE[name], dt);
    this._refresh();
  }

  destroy() {
    if (this.handle) native.destroy(this.handle);
    this.handle = null;
  }
}

module.exports = { Sim, BUF, STAGE };
