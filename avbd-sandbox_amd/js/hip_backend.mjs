// HipBackend — the MI355X backend for a JavaScript host, with the method
// surface of the reference's GpuBackend (src/physics/gpu/backend.js:243-2226)
// so a host can switch backends at the same call sites (boot.js:2145-2156
// doStep, 1774-1813 backend switch). Compute runs in libavbd_hip.so through
// the Node-API addon avbd.node; nothing here falls back to the CPU.
import { createRequire } from 'module';
import { dirname, join } from 'path';
import { fileURLToPath } from 'url';

const here = dirname(fileURLToPath(import.meta.url));
const require = createRequire(import.meta.url);
const addon = require(join(here, 'avbd.node'));

export const PRECISION = { exact: 0, f64: 1, mixed: 2 };
const PHASE_NAMES = ['broad', 'narrow', 'topology', 'colour', 'solve', 'total'];

export class HipBackend {
  constructor({ precision = 'mixed', team = 0, device = 0, manifoldCap = 0, pairCap = 0 } = {}) {
    this.precision = precision;
    this.team = team;
    this.device = device;
    this.manifoldCap = manifoldCap;
    this.pairCap = pairCap;
    this.ctx = null;
    this.ready = false;
    this.adapterInfo = null;
    this.gpuErrors = [];
    this.topologyDirty = true;
    this.bodyIndex = new Map();
    this.jointIndex = new Map();
    this.springIndex = new Map();
    this.forceRealloc = false;
    this.numBodies = 0;
    this.lastProfile = null;
    this._profiling = false;
    this._queued = 0;
  }

  // --- lifecycle (backend.js:337-469)
  // initDevice + initPipelines in one (backend.js:337-469); HIP code objects
  // are loaded with the library, so there is no pipeline build to wait for.
  async initDevice() {
    if (!this.ctx) {
      this.ctx = addon.create({
        device: this.device, precision: PRECISION[this.precision], team: this.team,
        manifoldCap: this.manifoldCap, pairCap: this.pairCap,
      });
      this.adapterInfo = 'AMD Instinct MI355X (gfx950, HIP)';
    }
    return this;
  }

  async initPipelines(onProgress) {
    await this.initDevice();
    if (onProgress) onProgress(1);
    this.ready = true;
    return this;
  }

  async init() {
    return this.initPipelines();
  }

  // --- topology (backend.js:662-1301, 1898-1915)
  // A topology edit keeps state where it can (backend.js:662-1301): bodies
  // found by identity in the last pack keep their device state, and, when the
  // survivors kept their order and the joints are exactly the surviving ones,
  // their contacts (warm start), joints (fracture included) and colours too
  // (avbd_repack). Otherwise the scene is packed afresh from the survivors'
  // device state, as the reference does when it cannot keep contact state.
  pack(solver) {
    const bodies = solver.bodies;
    const n = bodies.length;
    const prev = this.bodyIndex;
    const prevJoints = this.packedJoints || [];
    const A = {
      shape: new Int32Array(n), size: new Float64Array(n * 3), density: new Float64Array(n),
      friction: new Float64Array(n), pos: new Float64Array(n * 3), quat: new Float64Array(n * 4),
      vel: new Float64Array(n * 3), angvel: new Float64Array(n * 3),
    };
    this.bodyIndex = new Map();
    bodies.forEach((b, i) => {
      this.bodyIndex.set(b, i);
      A.shape[i] = b.shape === 'sphere' ? 1 : 0;
      A.size.set(b.size, i * 3);
      A.density[i] = b.density;
      A.friction[i] = b.friction;
      A.pos.set(b.positionLin, i * 3);
      A.quat.set(b.positionAng, i * 4);
      A.vel.set(b.velocityLin, i * 3);
      A.angvel.set(b.velocityAng, i * 3);
    });
    // joints and springs (spring.js: stiffness, rest, tearStrain), in the
    // force list's creation order
    const isSpring = (f) => 'stiffness' in f && 'rest' in f && !('stiffnessLin' in f);
    const joints = solver.forces.filter((f) => f.rA && f.rB && ('stiffnessLin' in f || isSpring(f)));
    const m = joints.length;
    let oldIndex = null;
    if (prev && this.numBodies > 0 && !this.forceRealloc) {
      oldIndex = new Int32Array(n);
      let survivors = 0;
      let last = -1;
      let ordered = true;
      bodies.forEach((b, i) => {
        const o = prev.get(b);
        oldIndex[i] = o === undefined ? -1 : o;
        if (o !== undefined) {
          survivors++;
          if (o <= last) ordered = false;
          last = o;
        }
      });
      const alive = (b) => b === null || this.bodyIndex.has(b);
      const kept = prevJoints.filter((j) => alive(j.bodyA || null) && alive(j.bodyB));
      const sameJoints = kept.length === joints.length && kept.every((j, k) => j === joints[k]);
      if (survivors === 0) {
        oldIndex = null;
      } else if (!ordered || !sameJoints) {
        // fresh pack: carry the survivors' device state into the solver first
        const N0 = this.numBodies;
        const d = {
          pos: new Float64Array(N0 * 3), quat: new Float64Array(N0 * 4), vel: new Float64Array(N0 * 3),
          angvel: new Float64Array(N0 * 3),
        };
        addon.readState(this.ctx, d.pos, d.quat, d.vel, d.angvel, null);
        bodies.forEach((b, i) => {
          const o = oldIndex[i];
          if (o < 0) return;
          A.pos.set(d.pos.subarray(o * 3, o * 3 + 3), i * 3);
          A.quat.set(d.quat.subarray(o * 4, o * 4 + 4), i * 4);
          A.vel.set(d.vel.subarray(o * 3, o * 3 + 3), i * 3);
          A.angvel.set(d.angvel.subarray(o * 3, o * 3 + 3), i * 3);
        });
        oldIndex = null;
      }
    }
    if (m) {
      A.jointA = new Int32Array(m);
      A.jointB = new Int32Array(m);
      A.jointRA = new Float64Array(m * 3);
      A.jointRB = new Float64Array(m * 3);
      A.jointStiffLin = new Float64Array(m);
      A.jointStiffAng = new Float64Array(m);
      A.jointFracture = new Float64Array(m);
      A.jointKind = new Int32Array(m);
      joints.forEach((j, k) => {
        A.jointA[k] = j.bodyA ? this.bodyIndex.get(j.bodyA) : -1;
        A.jointB[k] = this.bodyIndex.get(j.bodyB);
        A.jointRA.set(j.rA, k * 3);
        A.jointRB.set(j.rB, k * 3);
        if (isSpring(j)) {
          A.jointKind[k] = 1;
          A.jointStiffLin[k] = j.stiffness;
          A.jointStiffAng[k] = j.rest;
          A.jointFracture[k] = j.tearStrain === undefined ? Infinity : j.tearStrain;
        } else {
          A.jointStiffLin[k] = j.stiffnessLin;
          A.jointStiffAng[k] = j.stiffnessAng;
          A.jointFracture[k] = j.fracture;
        }
      });
    }
    addon.setParams(this.ctx, {
      dt: solver.dt, gravity: solver.gravity, iterations: solver.iterations, betaLin: solver.betaLin,
      betaAng: solver.betaAng, alpha: solver.alpha, gamma: solver.gamma, penaltyMin: solver.penaltyMin,
      penaltyMax: solver.penaltyMax, collisionMargin: solver.collisionMargin,
      stickThreshold: solver.stickThreshold, postStabilize: solver.postStabilize ? 1 : 0,
      rotatedInertia: solver.rotatedInertia === false ? 0 : 1, cachedContactJacobians: 1,
      colorOrder: solver.colorOrder === 'priority' || solver.colorOrder === 1 ? 1 : 0,
    });
    if (oldIndex) addon.pack(this.ctx, A, oldIndex);
    else addon.pack(this.ctx, A);
    this.lastPackPreservedContacts = oldIndex !== null;
    this.forceRealloc = false;
    this.packedJoints = joints;
    this.jointIndex = new Map();
    this.springIndex = new Map();
    joints.forEach((j, k) => (isSpring(j) ? this.springIndex : this.jointIndex).set(j, k));
    this.numBodies = n;
    this.topologyDirty = false;
  }

  prepareTopologyChange() {
    this.topologyDirty = true;
  }

  // A synchronous one-shot flush: true when it packed (backend.js:1911-1915).
  flushTopology(solver) {
    if (!this.topologyDirty) return false;
    this.pack(solver);
    return true;
  }

  /** Push a drag joint's world anchor without a repack (backend.js:1969-1977). */
  updateJointAnchor(joint) {
    const idx = this.jointIndex.get(joint);
    if (idx === undefined) return false;
    addon.setJointAnchor(this.ctx, idx, Float64Array.from(joint.rA));
    return true;
  }

  /**
   * Where a renderer finds constraint break state (backend.js:1939-1966): here
   * a host snapshot of the packed joints and springs, 16 numbers each, the
   * broken flag at offset 12, indexed through jointIndex / springIndex.
   */
  getConstraintGpuState() {
    const n = (this.packedJoints || []).length;
    if (!this.ctx || n === 0) return null;
    const records = new Float64Array(n * 16);
    addon.readJoints(this.ctx, records);
    return {
      records, stride: 16, jointBrokenOffset: 12, springBrokenOffset: 12,
      jointIndex: this.jointIndex, springIndex: this.springIndex,
    };
  }

  getSpringGpuState() {
    const s = this.getConstraintGpuState();
    return s && { records: s.records, springStride: 16, brokenOffset: 12, springIndex: s.springIndex };
  }

  // --- stepping (backend.js:1393-1620); returns the host encode time in ms
  // Never throws (backend.js:366-381, 1683-1739): a failed step is recorded
  // in gpuErrors; a capacity overflow also marks the topology dirty, so that
  // the next step's state-keeping repack grows the arena it overflowed.
  step(solver) {
    const t0 = Date.now();
    try {
      this.flushTopology(solver);
      addon.step(this.ctx, 1);
      this._queued++;
      if (this._profiling) this._readProfile();
    } catch (e) {
      this.gpuErrors.push(String(e && e.message ? e.message : e));
      if (e && String(e.code) === '-3') this.topologyDirty = true;  // AVBD_E_CAPACITY
    }
    return Date.now() - t0;
  }

  get queuedSteps() {
    return this._queued;
  }

  async idle() {
    await addon.sync(this.ctx);
    this._queued = 0;
  }

  // --- state (backend.js:2154-2192, 1984-2039)
  async readState() {
    await this.idle();
    const n = this.numBodies;
    const data = {
      pos: new Float64Array(n * 3), quat: new Float64Array(n * 4), vel: new Float64Array(n * 3),
      angvel: new Float64Array(n * 3), prevVel: new Float64Array(n * 3),
    };
    addon.readState(this.ctx, data.pos, data.quat, data.vel, data.angvel, data.prevVel);
    return { data, numBodies: n, counters: addon.readCounters(this.ctx) };
  }

  // Device state back into the scene graph (backend.js:1984-2039): bodies by
  // their packed index, including the step-start pose (initialLin /
  // initialAng), and the packed joints' penalties, multipliers and fracture,
  // and the springs' penalty, tearing and peak strain.
  async syncToCPU(solver) {
    const { data } = await this.readState();
    const n = this.numBodies;
    const p0 = new Float64Array(n * 3);
    const q0 = new Float64Array(n * 4);
    addon.readStepStart(this.ctx, p0, q0);
    const setv = (obj, key, src) => {
      if (obj[key] && obj[key].set) obj[key].set(src);
      else obj[key] = Float64Array.from(src);
    };
    solver.bodies.forEach((b, i) => {
      b.positionLin.set(data.pos.subarray(i * 3, i * 3 + 3));
      b.positionAng.set(data.quat.subarray(i * 4, i * 4 + 4));
      b.velocityLin.set(data.vel.subarray(i * 3, i * 3 + 3));
      b.velocityAng.set(data.angvel.subarray(i * 3, i * 3 + 3));
      b.prevVelocityLin.set(data.prevVel.subarray(i * 3, i * 3 + 3));
      setv(b, 'initialLin', p0.subarray(i * 3, i * 3 + 3));
      setv(b, 'initialAng', q0.subarray(i * 4, i * 4 + 4));
    });
    const joints = this.packedJoints || [];
    if (joints.length) {
      const st = new Float64Array(joints.length * 16);
      addon.readJoints(this.ctx, st);
      joints.forEach((j, k) => {
        const o = st.subarray(k * 16, k * 16 + 16);
        if ('stiffnessLin' in j) {
          setv(j, 'penaltyLin', o.subarray(0, 3));
          setv(j, 'penaltyAng', o.subarray(3, 6));
          setv(j, 'lambdaLin', o.subarray(6, 9));
          setv(j, 'lambdaAng', o.subarray(9, 12));
          if (o[12] > 0.5) j.broken = true;
        } else {
          j.penalty = o[13];
          j.broken = o[12] > 0.5;
          j.peakStrain = o[15];
        }
      });
    }
  }

  // --- poses for drawing (renderer.js:1327-1383; boot.js:2223-2231): per
  // body {p, 0, q} as f32 at a fraction alpha of the last step
  async exportPoses(alpha = 1, out = null) {
    await this.idle();
    const poses = out || new Float32Array(this.numBodies * 8);
    addon.exportPoses(this.ctx, alpha, poses);
    return poses;
  }

  // --- verification (backend.js:2053-2151)
  async readColoring() {
    await this.idle();
    const colorOf = new Int32Array(this.numBodies);
    const numColors = addon.readColoring(this.ctx, colorOf);
    const count = addon.readPairs(this.ctx, null);
    const contactPairs = new Int32Array(count * 2);
    addon.readPairs(this.ctx, contactPairs);
    return { colorOf, numColors, contactPairs };
  }

  // Verification hook (backend.js:1915-1931): install a colouring given as
  // body indices grouped by colour (entries, offsets[c]..offsets[c+1]) for
  // the following steps; clearColoring() returns to the greedy colouring.
  uploadColoring(entries, offsets, counts) {
    const colorOf = new Int32Array(this.numBodies).fill(-1);
    for (let c = 0; c + 1 < offsets.length; c++)
      for (let k = offsets[c]; k < offsets[c] + (counts ? counts[c] : offsets[c + 1] - offsets[c]); k++)
        colorOf[entries[k]] = c;
    addon.uploadColoring(this.ctx, colorOf);
  }

  clearColoring() {
    addon.uploadColoring(this.ctx, null);
  }

  // The colouring's claim order (the device counterpart of Solver.bodyOrder,
  // solver.js:225-233): rank[i] is body i's position; null for the default.
  setBodyOrder(rank) {
    addon.setBodyOrder(this.ctx, rank ? Int32Array.from(rank) : null);
  }

  // --- metrics and profiling (backend.js:309-324, 1836-1891)
  get lastStats() {
    return this.ctx ? addon.readCounters(this.ctx) : null;
  }

  setProfiling(on) {
    addon.setProfiling(this.ctx, !!on);
    this._profiling = !!on;
    return true;
  }

  _readProfile() {
    const ms = addon.readProfile(this.ctx);
    this.lastProfile = Object.fromEntries(PHASE_NAMES.map((k, i) => [k, ms[i]]));
  }
}

export const abiVersion = () => addon.abiVersion();
