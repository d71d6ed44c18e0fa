// HipBackend — the MI355X backend for a JavaScript host, with the method
// surface of the reference's GpuBackend (src/physics/gpu/backend.js:243-2226)
// so a host can switch backends at the same call sites (boot.js:2145-2156
// doStep, 1774-1813 backend switch). Compute runs in libavbd_hip.so through
// the Node-API addon avbd.node; nothing here falls back to the CPU.
import { createRequire } from 'module';
import { isSpring, num, packArrays } from './pack_arrays.mjs';
import { dirname, join } from 'path';
import { fileURLToPath } from 'url';

const here = dirname(fileURLToPath(import.meta.url));
const require = createRequire(import.meta.url);
const addon = require(join(here, 'avbd.node'));

export const PRECISION = { exact: 0, f64: 1, mixed: 2 };
// avbd_config.graphs
export const GRAPHS = { solve: 0, step: 1, none: -1 };
const PHASE_NAMES = ['broad', 'narrow', 'topology', 'colour', 'solve', 'total'];
export const MAX_COLORS = 256;

// The body arena readState returns: the sections of the reference GPU's
// BODY_F32 buffer (layout.js BODY_SECTIONS), 4 floats per body each except
// bStatic (16), one section after the other, so `data[layout.body.bPos + i * 4
// + k]` addresses a body's position as GpuBackend's readState does.
const BODY_SECTIONS = [
  ['bPos', 4], ['bQuat', 4], ['bPosB', 4], ['bQuatB', 4], ['bVel', 4], ['bVelA', 4], ['bPrevV', 4],
  ['bInitP', 4], ['bInitQ', 4], ['bInerP', 4], ['bInerQ', 4], ['bStatic', 16],
];
// bStatic element indices (layout.js BS_*)
export const BS = {
  MASS: 0, MOMENT: 1, SIZE: 4, STATIC_FRICTION: 7, DYNAMIC_FRICTION: 8, RESTITUTION: 9, RADIUS: 10, FLAGS: 11,
  CONTACT_OFFSET: 12, REST_OFFSET: 13,
};

export { packArrays };

export function bodyLayout(n) {
  const body = {};
  let o = 0;
  for (const [name, per] of BODY_SECTIONS) {
    body[name] = o;
    o += per * n;
  }
  return { body, bodyF32Len: o };
}

export class HipBackend {
  // graphs: 'solve' replays each step's solve loop as a graph (the step reads
  // three counts on the host); 'step' replays whole steps with no host read
  // inside them (the reference's step never waits either, backend.js:1606-1611),
  // reporting a capacity overflow on a later step; 'none' launches directly.
  constructor({ precision = 'mixed', team = 0, device = 0, manifoldCap = 0, pairCap = 0, graphs = 'solve' } = {}) {
    if (!(graphs in GRAPHS)) throw new Error(`graphs must be one of ${Object.keys(GRAPHS).join(', ')}`);
    this.graphs = graphs;
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
    // Colouring controls of GpuBackend (backend.js:256, 270, 299): rounds of
    // refinement per step and colours per step; null = none (the exact greedy
    // colouring, this backend's default; the reference ships 5 rounds and 16
    // colours and tolerates the conflicts they leave).
    this.colorRounds = null;
    this.encColors = null;
    this._limits = '0,0';
    this._override = null;
    this._uploaded = null;
  }

  // A truthy value makes the following steps use the colouring given to
  // uploadColoring (backend.js:299, 1508-1516); null returns to the greedy one.
  get overrideColoring() {
    return this._override;
  }

  set overrideColoring(v) {
    this._override = v || null;
    if (!this.ctx || this.numBodies === 0) return;  // applied by the next uploadColoring
    if (this._override && this._uploaded) addon.uploadColoring(this.ctx, this._uploaded);
    else addon.uploadColoring(this.ctx, null);
  }

  // --- lifecycle (backend.js:337-469)
  // initDevice + initPipelines in one (backend.js:337-469); HIP code objects
  // are loaded with the library, so there is no pipeline build to wait for.
  async initDevice() {
    if (!this.ctx) {
      this.ctx = addon.create({
        device: this.device, precision: PRECISION[this.precision], team: this.team,
        manifoldCap: this.manifoldCap, pairCap: this.pairCap, graphs: GRAPHS[this.graphs],
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
    const { arrays: A, index, joints, params } = packArrays(solver);
    const bodies = solver.bodies;
    const n = bodies.length;
    const prev = this.bodyIndex;
    const prevJoints = this.packedJoints || [];
    this.bodyIndex = index;
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
    addon.setParams(this.ctx, params);
    this._contactModel = params.contactModel;
    this._offsets = { contact: params.contactOffset, rest: params.restOffset };
    if (oldIndex) addon.pack(this.ctx, A, oldIndex);
    else addon.pack(this.ctx, A);
    this._packed = A;
    this._uploaded = null;
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
      const limits = `${this.colorRounds || 0},${this.encColors || 0}`;
      if (limits !== this._limits) {
        addon.setColoringLimits(this.ctx, this.colorRounds || 0, this.encColors || 0);
        this._limits = limits;
      }
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
  // Body state in f64, straight from the device.
  async readBodies() {
    await this.idle();
    const n = this.numBodies;
    const d = {
      pos: new Float64Array(n * 3), quat: new Float64Array(n * 4), vel: new Float64Array(n * 3),
      angvel: new Float64Array(n * 3), prevVel: new Float64Array(n * 3),
    };
    addon.readState(this.ctx, d.pos, d.quat, d.vel, d.angvel, d.prevVel);
    return d;
  }

  // GpuBackend.readState's result (backend.js:2154-2192): `data` is the body
  // arena in the reference's BODY_F32 layout (f32, addressed through
  // `layout.body`), with `counters` {pairs, slots, adj, overflow, contacts} and
  // `dispatch`; `f64` carries the same state at full precision.
  async readState() {
    const f64 = await this.readBodies();
    const n = this.numBodies;
    const L = bodyLayout(n);
    const data = new Float32Array(L.bodyF32Len);
    const put = (sec, src, w) => {
      for (let i = 0; i < n; i++) for (let k = 0; k < w; k++) data[L.body[sec] + i * 4 + k] = src[i * w + k];
    };
    put('bPos', f64.pos, 3); put('bQuat', f64.quat, 4);
    put('bPosB', f64.pos, 3); put('bQuatB', f64.quat, 4);
    put('bVel', f64.vel, 3); put('bVelA', f64.angvel, 3); put('bPrevV', f64.prevVel, 3);
    const p0 = new Float64Array(n * 3), q0 = new Float64Array(n * 4);
    addon.readStepStart(this.ctx, p0, q0);
    put('bInitP', p0, 3); put('bInitQ', q0, 4);
    const py = new Float64Array(n * 3), qy = new Float64Array(n * 4);
    addon.readInertialTarget(this.ctx, py, qy);
    put('bInerP', py, 3); put('bInerQ', qy, 4);
    const A = this._packed;
    if (A) {
      for (let i = 0; i < n; i++) {
        const o = L.body.bStatic + i * 16;
        const sx = A.size[i * 3], sy = A.size[i * 3 + 1], sz = A.size[i * 3 + 2];
        const sphere = A.shape[i] === 1;
        let mass, m0, m1, m2, radius;
        if (sphere) {  // rigid.js:105-133
          const r = sx * 0.5;
          mass = (4 / 3) * Math.PI * r * r * r * A.density[i];
          m0 = m1 = m2 = 0.4 * mass * r * r;
          radius = r;
        } else {  // rigid.js:29-98
          mass = sx * sy * sz * A.density[i];
          m0 = ((sy * sy + sz * sz) / 12) * mass; m1 = ((sx * sx + sz * sz) / 12) * mass;
          m2 = ((sx * sx + sy * sy) / 12) * mass;
          radius = Math.sqrt(sx * sx / 4 + sy * sy / 4 + sz * sz / 4);
        }
        data[o + BS.MASS] = mass;
        data[o + BS.MOMENT] = m0; data[o + BS.MOMENT + 1] = m1; data[o + BS.MOMENT + 2] = m2;
        data[o + BS.SIZE] = sx; data[o + BS.SIZE + 1] = sy; data[o + BS.SIZE + 2] = sz;
        // the contact material as the GPU law resolves it (backend.js:156-185)
        const co = Math.max(0, Number.isFinite(A.contactOffset[i]) ? A.contactOffset[i] : this._offsets.contact);
        const ro = Math.min(Number.isFinite(A.restOffset[i]) ? A.restOffset[i] : this._offsets.rest, co);
        data[o + BS.STATIC_FRICTION] = A.staticFriction[i];
        data[o + BS.DYNAMIC_FRICTION] = A.dynamicFriction[i];
        data[o + BS.RESTITUTION] = A.restitution[i];
        data[o + BS.CONTACT_OFFSET] = co;
        data[o + BS.REST_OFFSET] = ro;
        data[o + BS.RADIUS] = radius;
        data[o + BS.FLAGS] = sphere ? 1 : 0;
      }
    }
    const c = addon.readCounters(this.ctx);
    const joints = this.packedJoints || [];
    const springs = joints.filter((j) => this.springIndex.has(j)).length;
    return {
      data, layout: L, numBodies: n, f64,
      counters: {
        pairs: c.pairs, slots: c.manifolds, adj: 2 * (c.manifolds + c.joints), overflow: c.overflow,
        contacts: c.contacts,
      },
      dispatch: {
        numBodies: n, numJoints: joints.length - springs, numSprings: springs,
        encColors: this.encColors || c.colors,
      },
    };
  }

  // Device state back into the scene graph (backend.js:1984-2039): bodies by
  // their packed index, including the step-start pose (initialLin /
  // initialAng), and the packed joints' penalties, multipliers and fracture,
  // and the springs' penalty, tearing and peak strain.
  async syncToCPU(solver) {
    const data = await this.readBodies();
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
  // GpuBackend.readColoring's result: the colouring the last step used
  // (colorOf, -1 for static bodies; counts and offsets per colour; entries,
  // the bodies grouped by colour, by index within a colour) and the contact
  // slots it coloured around: contactPairs ([A, B] per manifold), slots,
  // manifolds (distinct pairs) and slotInfo (per manifold: contact count,
  // friction, basis, and per contact the body-frame anchors, C0, penalty,
  // lambda, stick flag and feature key). numColors and broadPhasePairs (the
  // last broad phase's pairs, A > B) are this backend's additions.
  async readColoring() {
    await this.idle();
    const n = this.numBodies;
    const colorOf = new Int32Array(n);
    const numColors = addon.readColoring(this.ctx, colorOf);
    const counts = new Uint32Array(MAX_COLORS);
    for (let i = 0; i < n; i++) if (colorOf[i] >= 0) counts[colorOf[i]]++;
    const offsets = new Uint32Array(MAX_COLORS + 1);
    for (let c = 0; c < MAX_COLORS; c++) offsets[c + 1] = offsets[c] + counts[c];
    const entries = new Uint32Array(n);
    const cursor = offsets.slice(0, MAX_COLORS);
    for (let i = 0; i < n; i++) if (colorOf[i] >= 0) entries[cursor[colorOf[i]]++] = i;
    const count = addon.readPairs(this.ctx, null);
    const broadPhasePairs = new Int32Array(count * 2);
    addon.readPairs(this.ctx, broadPhasePairs);
    const M = addon.readManifolds(this.ctx);
    const contactModel = this._contactModel;
    const m = M.head.length / 4;
    const nj = M.forces.length / 4 - m;
    const contactPairs = [];
    const slotInfo = [];
    const pairSet = new Set();
    let k = 0;
    for (let t = 0; t < m; t++) {
      const a = M.head[t * 4], b = M.head[t * 4 + 1], nc = M.head[t * 4 + 2];
      contactPairs.push([a, b]);
      pairSet.add(`${Math.min(a, b)}:${Math.max(a, b)}`);
      const bm = M.basisMu.subarray(t * 10, t * 10 + 10);
      const contacts = [];
      let sticking = 0, maxPenalty = 0;
      for (let s = 0; s < nc; s++, k++) {
        const r = M.contacts.subarray(k * 16, k * 16 + 16);
        const stick = r[15] > 0.5;
        sticking += stick ? 1 : 0;
        maxPenalty = Math.max(maxPenalty, r[9], r[10], r[11]);
        contacts.push({
          rA: [r[0], r[1], r[2]], rB: [r[3], r[4], r[5]], constraint: [r[6], r[7], r[8]],
          penalty: [r[9], r[10], r[11]], lambda: [r[12], r[13], r[14]], sticking: stick,
          feature: M.features[(nj + t) * 8 + s], restitutionBias: M.bias[k],
        });
      }
      slotInfo.push({
        numContacts: nc, chainLen: 1, staticFriction: M.material[t * 3],
        dynamicFriction: contactModel ? M.material[t * 3 + 1] : M.material[t * 3],
        basis: { normal: [bm[0], bm[1], bm[2]], tangent1: [bm[3], bm[4], bm[5]], tangent2: [bm[6], bm[7], bm[8]] },
        contacts, stickingContacts: sticking, maxPenalty,
      });
    }
    return {
      colorOf, counts, offsets, entries, contactPairs, slots: m, manifolds: pairSet.size, slotInfo,
      numColors, broadPhasePairs,
    };
  }

  // Verification hook (backend.js:1915-1931): install a colouring given as
  // body indices grouped by colour (entries, offsets[c]..offsets[c+1]) for
  // the following steps; clearColoring() returns to the greedy colouring.
  // The colouring is used while overrideColoring is truthy (backend.js:299).
  uploadColoring(entries, offsets, counts) {
    const colorOf = new Int32Array(this.numBodies).fill(-1);
    for (let c = 0; c + 1 < offsets.length; c++)
      for (let k = offsets[c]; k < offsets[c] + (counts ? counts[c] : offsets[c + 1] - offsets[c]); k++)
        colorOf[entries[k]] = c;
    this._uploaded = colorOf;
    if (this._override === null) this._override = true;
    addon.uploadColoring(this.ctx, colorOf);
  }

  clearColoring() {
    this.overrideColoring = null;
  }

  // The colouring's claim order (the device counterpart of Solver.bodyOrder,
  // solver.js:225-233): rank[i] is body i's position; null for the default.
  setBodyOrder(rank) {
    addon.setBodyOrder(this.ctx, rank ? Int32Array.from(rank) : null);
  }

  // --- launch shapes (no counterpart in the reference): one step that times
  // the solve's team (lanes per body) and workgroup shapes on the step's own
  // data and keeps the fastest; trials maps "team/block" to ms.
  autotune() {
    const r = addon.autotune(this.ctx);
    const trials = {};
    [1, 4, 8, 16, 32, 64].forEach((t, i) => [64, 128, 256, 512].forEach((b, j) => {
      if (r.trials[i * 4 + j] > 0) trials[`${t}/${b}`] = r.trials[i * 4 + j];
    }));
    return { team: r.team, block: r.block, trials };
  }

  setLaunchShape(team, block) {
    addon.setLaunchShape(this.ctx, team, block);
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
