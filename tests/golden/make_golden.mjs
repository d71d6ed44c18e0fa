// Writes golden vectors for the f64 oracle by running the reference CPU engine.
//
// Run in the development container only (the reference is not present on the
// GPU machine), from the repository root:
//
//   node --experimental-loader ./tests/golden/nullish_loader.mjs \
//        tests/golden/make_golden.mjs /root/reference tests/golden
//
// Each scene is defined here as plain data (bodies, joints, parameters),
// built into the reference `Solver`, stepped, and dumped as JSON:
//   * per step: SHA-256 of the state (13 f64 per body: p, q, v, w, with -0
//     folded to +0), SHA-256 of the force list ([kind, a, b, nc, features...]
//     as u32), and under the colour schedule the colouring used;
//   * exact states at a few sample steps.
// The oracle must reproduce every hash bit for bit.
import { createHash } from 'crypto';
import { writeFileSync } from 'fs';
import { join } from 'path';
import { pathToFileURL } from 'url';

const REF = process.argv[2] || '/root/reference';
const OUT = process.argv[3] || 'tests/golden';
const imp = (p) => import(pathToFileURL(join(REF, p)).href);
let Solver, Rigid, Joint, Spring, Manifold, Coloring, colorOrder;

// Seeded LCG (Numerical Recipes constants), uniform in [0, 1).
function lcg(seed) {
  let s = seed >>> 0;
  return () => {
    s = (s * 1664525 + 1013904223) >>> 0;
    return s / 4294967296;
  };
}

const box = (size, density, friction, pos, extra = {}) => ({ shape: 'box', size, density, friction, pos, ...extra });
const sphere = (diameter, density, friction, pos, extra = {}) =>
  ({ shape: 'sphere', size: [diameter, diameter, diameter], density, friction, pos, ...extra });
const ground = (friction = 0.5) => box([400, 400, 4], 0, friction, [0, 0, -2]);

function axisAngle(ax, ay, az, angle) {
  const len = Math.hypot(ax, ay, az);
  const h = angle * 0.5;
  const s = Math.sin(h);
  return [(ax / len) * s, (ay / len) * s, (az / len) * s, Math.cos(h)];
}

// ---------------------------------------------------------------- scenes
function sceneStack10() {
  const r = lcg(1);
  const bodies = [ground(0.5)];
  for (let i = 0; i < 10; i++) {
    const jx = (r() * 2 - 1) * 0.01;
    const jy = (r() * 2 - 1) * 0.01;
    bodies.push(box([1, 1, 1], 1000, 0.5, [jx, jy, 0.501 + i * 1.001]));
  }
  return { bodies, joints: [], params: { gravity: -9.81, iterations: 10 }, steps: 300 };
}

function sceneDrop() {
  return {
    bodies: [box([100, 100, 2], 0, 0.6, [0, 0, -1]),
      box([1, 1, 1], 1, 0.5, [0.13, -0.07, 1.2], { vel: [1.5, 0.4, -6] })],
    joints: [], params: { gravity: -10, iterations: 8 }, steps: 60,
  };
}

function sceneSpheres() {
  return {
    bodies: [box([100, 100, 2], 0, 0.6, [0, 0, -1]),
      sphere(1, 1, 0.55, [0.17, -0.11, 0.48], { vel: [1.2, 0.4, -1], angvel: [0.7, -1.1, 0] }),
      sphere(1, 0, 0.5, [2.2, 0, 0.5]),
      sphere(1, 1, 0.5, [1.25, 0.02, 0.5], { vel: [3.5, 0.6, 0], angvel: [0, 0, 1.3] })],
    joints: [], params: { gravity: -10, iterations: 8 }, steps: 90,
  };
}

function sceneChain() {
  const bodies = [box([100, 100, 2], 0, 0.5, [0, 0, -30])];
  const joints = [];
  for (let i = 0; i < 8; i++) {
    bodies.push(box([0.8, 0.4, 0.4], i === 0 ? 0 : 1, 0.5, [i * 1.4, 0, 10]));
    if (i > 0) joints.push({ a: i, b: i + 1, rA: [0.7, 0, 0], rB: [-0.7, 0, 0], kLin: Infinity, kAng: Infinity, fracture: Infinity });
  }
  // a world-anchored joint and a breakable one
  bodies.push(box([0.5, 0.5, 0.5], 2, 0.5, [0, 3, 10]));
  joints.push({ a: -1, b: 9, rA: [0, 3, 11], rB: [0, 0, 0.25], kLin: Infinity, kAng: 0, fracture: Infinity });
  bodies.push(box([0.5, 0.5, 0.5], 40, 0.5, [0, 3, 8.5]));
  joints.push({ a: 9, b: 10, rA: [0, 0, -0.25], rB: [0, 0, 0.25], kLin: Infinity, kAng: 1e4, fracture: 50 });
  return { bodies, joints, params: { gravity: -10, iterations: 8 }, steps: 90 };
}

function scenePendulum() {
  // Chain of hard ball-socket links ending in a heavy mass (about 26,000:1).
  const N = 12, link = 0.5, anchor = [0, 0, 12];
  const angle = 40 * 0.01745329251994329577;
  const q = axisAngle(0, 1, 0, angle);
  const along = (s) => [anchor[0] - Math.sin(angle) * s, anchor[1], anchor[2] - Math.cos(angle) * s];
  const bodies = [ground(0.5)];
  const joints = [];
  for (let i = 0; i < N; i++) {
    bodies.push(box([link, link * 0.5, link], i === 0 ? 0 : 1, 0.4, along(i * link), { quat: q }));
    if (i > 0) joints.push({ a: i, b: i + 1, rA: [0, 0, -link / 2], rB: [0, 0, link / 2], kLin: Infinity, kAng: 0, fracture: Infinity });
  }
  bodies.push(box([3, 3, 3], 60, 0.5, along((N - 1) * link + link / 2 + 1.5), { quat: q }));
  joints.push({ a: N, b: N + 1, rA: [0, 0, -link / 2], rB: [0, 0, 1.5], kLin: Infinity, kAng: 0, fracture: Infinity });
  return { bodies, joints, params: { gravity: -9.81, iterations: 20 }, steps: 120 };
}

function sceneRamp() {
  const r = lcg(5);
  const angle = 30 * 0.01745329251994329577;
  const q = axisAngle(0, 1, 0, angle);
  const bodies = [ground(0.5), box([30, 12, 1], 0, 1.0, [0, 0, 6], { quat: q })];
  const c = Math.cos(angle), s = Math.sin(angle);
  // ramp frame: tangent (c, 0, -s) points down-slope along x, normal (s, 0, c)
  for (let i = 0; i < 12; i++) {
    const u = -8 + (i % 4) * 2.0, w = -4 + Math.floor(i / 4) * 2.5;
    const h = 0.5 + 0.25 + 0.001;
    const pos = [c * u + s * h, w, 6 - s * u + c * h];
    const v0 = r() * 2;
    bodies.push(box([0.5, 0.5, 0.5], 500, 0.1 + r() * 0.8, pos, { quat: q, vel: [c * v0, 0, -s * v0] }));
  }
  return { bodies, joints: [], params: { gravity: -9.81, iterations: 10 }, steps: 120 };
}

function sceneRain() {
  const r = lcg(11);
  const bodies = [ground(0.6)];
  bodies.push(box([0.5, 12, 6], 0, 0.5, [-6, 0, 3]));
  bodies.push(box([0.5, 12, 6], 0, 0.5, [6, 0, 3]));
  bodies.push(box([12, 0.5, 6], 0, 0.5, [0, -6, 3]));
  bodies.push(box([12, 0.5, 6], 0, 0.5, [0, 6, 3]));
  for (let i = 0; i < 48; i++) {
    const pos = [(r() * 2 - 1) * 4, (r() * 2 - 1) * 4, 1 + r() * 10];
    const density = Math.exp(Math.log(1) + r() * Math.log(1000));
    const friction = 0.2 + r() * 0.7;
    const q = axisAngle(r() - 0.5, r() - 0.5, r() - 0.5 + 1e-3, r() * 6.28);
    if (r() < 0.7) {
      bodies.push(box([0.2 + r() * 0.8, 0.2 + r() * 0.8, 0.2 + r() * 0.8], density, friction, pos, { quat: q }));
    } else {
      bodies.push(sphere(2 * (0.1 + r() * 0.4), density, friction, pos, { quat: q }));
    }
  }
  return { bodies, joints: [], params: { gravity: -9.81, iterations: 12 }, steps: 150 };
}

function scenePyramid(layers, broadphase, steps) {
  const bodies = [ground(0.6)];
  for (let layer = 0; layer < layers; layer++) {
    const w = layers - layer;
    const off = (w - 1) / 2;
    for (let x = 0; x < w; x++)
      for (let y = 0; y < w; y++) bodies.push(box([1, 1, 1], 1, 0.6, [x - off, y - off, layer + 0.5]));
  }
  bodies.push(sphere(1, 10, 0.6, [0, 0, layers + 3], { vel: [0, 0, -20] }));
  return { bodies, joints: [], params: { gravity: -9.81, iterations: 20, broadphase }, steps };
}

function sceneSprings() {
  // The parity harness's two spring scenes side by side (gen_scenes.mjs:213-237),
  // a spring that tears, and a spring of current rest length between two
  // touching boxes (the spring keeps them from making a contact).
  const bodies = [box([100, 100, 1], 0, 0.5, [0, 0, 0])];
  const joints = [];
  const spring = (a, b, rA, rB, stiffness, rest, extra = {}) =>
    joints.push({ type: 'spring', a, b, rA, rB, stiffness, rest, tearStrain: Infinity, ...extra });
  bodies.push(box([1, 1, 1], 0, 0.5, [0, 0, 14]));
  bodies.push(box([2, 2, 2], 1, 0.5, [0, 0, 8]));
  spring(1, 2, [0, 0, 0], [0, 0, 0], 100, 4);
  let prev = -1;
  for (let i = 0; i < 8; i++) {
    bodies.push(box([1, 0.75, 0.75], i === 0 || i === 7 ? 0 : 1, 0.5, [(i - 3.5) * 3.0, 10, 12]));
    const cur = bodies.length - 1;
    if (prev >= 0) spring(prev, cur, [0.5, 0, 0], [-0.5, 0, 0], i % 2 === 0 ? 10 : 10000, 3);
    prev = cur;
  }
  bodies.push(box([1, 1, 1], 0, 0.5, [0, -10, 14]));
  bodies.push(box([1.5, 1.5, 1.5], 5, 0.5, [0, -10, 9]));
  spring(bodies.length - 2, bodies.length - 1, [0, 0, -0.5], [0, 0, 0.75], 2000, 3.75, { tearStrain: 0.03 });
  bodies.push(box([1, 1, 1], 1, 0.5, [8, 0, 0.5]));
  bodies.push(box([1, 1, 1], 1, 0.5, [9.002, 0, 0.5]));
  spring(bodies.length - 2, bodies.length - 1, [0.5, 0, 0], [-0.5, 0, 0], 500, -1);
  return { bodies, joints, params: { gravity: -10, iterations: 10 }, steps: 120 };
}

function scenePostStab() {
  // post-stabilisation (solver.js:115-129, 411-417): a 3-cube stack dropped
  // with a gap, a sphere landing on a box, and a 4-link jointed pendulum
  const bodies = [ground(0.6)];
  for (let i = 0; i < 3; i++) bodies.push(box([1, 1, 1], 1, 0.6, [0.02 * i, 0, 0.55 + 1.05 * i]));
  bodies.push(box([1.5, 1.5, 0.5], 2, 0.5, [4, 0, 0.3]));
  bodies.push(sphere(0.8, 1, 0.5, [4.1, 0, 1.6], { vel: [0, 0, -2] }));
  const joints = [];
  bodies.push(box([0.4, 0.4, 0.4], 0, 0.5, [-4, 0, 6]));
  let prev = bodies.length - 1;
  for (let i = 0; i < 4; i++) {
    bodies.push(box([1, 0.3, 0.3], 1, 0.5, [-4 + 0.7 + i * 1.0, 0, 6]));
    const cur = bodies.length - 1;
    joints.push({ a: prev, b: cur, rA: i === 0 ? [0.2, 0, 0] : [0.5, 0, 0], rB: [-0.5, 0, 0],
                  kLin: Infinity, kAng: 0, fracture: Infinity });
    prev = cur;
  }
  return { bodies, joints, params: { gravity: -10, iterations: 8, postStabilize: true }, steps: 90 };
}

function sceneBodyInertia() {
  // rotatedInertia = false (solver.js:130-140): the body-frame inertia in the
  // mass matrix; tumbling boxes of unequal moments landing on the ground
  const bodies = [ground(0.5)];
  bodies.push(box([2, 0.5, 0.25], 1, 0.5, [0, 0, 2], { quat: axisAngle(1, 1, 0, 0.7), angvel: [3, -2, 5] }));
  bodies.push(box([0.3, 1.2, 0.6], 2, 0.4, [2.5, 0, 1.5], { quat: axisAngle(0, 1, 1, 1.1), angvel: [-4, 1, 2] }));
  bodies.push(box([1, 1, 1], 1, 0.6, [-2.5, 0, 0.5]));
  bodies.push(box([0.6, 0.6, 1.8], 1, 0.6, [-2.5, 0.1, 2.0], { quat: axisAngle(1, 0, 0, 0.4), angvel: [0, 2, 0] }));
  return { bodies, joints: [], params: { gravity: -10, iterations: 10, rotatedInertia: false }, steps: 90 };
}

function sceneEdits() {
  // Solver.removeBodies (solver.js:261-300) and bodies created mid-run: a
  // 4-layer pyramid and a 3-link rope; after step 20 three pyramid cubes and a
  // rope link go (their contacts and joints with them), after step 30 a cube
  // and a sphere are dropped in
  const bodies = [ground(0.6)];
  for (let l = 0; l < 4; l++) {
    const n = 4 - l;
    for (let i = 0; i < n; i++) {
      for (let j = 0; j < n; j++) bodies.push(box([1, 1, 1], 1, 0.6, [i - (n - 1) / 2, j - (n - 1) / 2, 0.5 + l]));
    }
  }
  const joints = [];
  bodies.push(box([0.4, 0.4, 0.4], 0, 0.5, [6, 0, 5]));
  let prev = bodies.length - 1;
  for (let i = 0; i < 3; i++) {
    bodies.push(box([0.8, 0.3, 0.3], 1, 0.5, [6.6 + i * 0.9, 0, 5]));
    const cur = bodies.length - 1;
    joints.push({ a: prev, b: cur, rA: i === 0 ? [0.2, 0, 0] : [0.4, 0, 0], rB: [-0.4, 0, 0],
                  kLin: Infinity, kAng: 0, fracture: Infinity });
    prev = cur;
  }
  const edits = [
    { step: 20, remove: [3, 17, 28, bodies.length - 2] },
    { step: 30, add: [box([1, 1, 1], 1, 0.6, [0.3, 0.2, 6]), sphere(0.6, 2, 0.5, [-1.5, 0, 3], { vel: [2, 0, 0] })] },
  ];
  return { bodies, joints, params: { gravity: -10, iterations: 10 }, steps: 60, edits };
}

// ---------------------------------------------------------------- runner
function makeBody(solver, d) {
  const vel = d.vel || [0, 0, 0];
  const b = d.shape === 'sphere'
    ? Rigid.sphere(solver, d.size[0], d.density, d.friction, d.pos, vel)
    : new Rigid(solver, d.size, d.density, d.friction, d.pos, vel);
  if (d.quat) { b.positionAng.set(d.quat); b.initialAng.set(d.quat); }
  if (d.angvel) b.velocityAng.set(d.angvel);
  return b;
}

// A topology edit after a step: bodies removed by their index at that moment
// (Solver.removeBodies), then bodies appended (new Rigid on the running solver).
function applyEdit(solver, edit) {
  if (edit.remove) {
    const doomed = new Set(edit.remove.map((i) => solver.bodies[i]));
    solver.removeBodies((b) => doomed.has(b));
  }
  for (const d of edit.add || []) makeBody(solver, d);
}

function build(def) {
  const solver = new Solver();
  solver.dt = 1 / 60;
  solver.gravity = def.params.gravity;
  solver.iterations = def.params.iterations;
  if (def.params.postStabilize) solver.postStabilize = true;
  if (def.params.rotatedInertia === false) solver.rotatedInertia = false;
  solver.broadphase = def.params.broadphase || 'bruteforce';
  const bodies = def.bodies.map((d) => {
    const vel = d.vel || [0, 0, 0];
    const b = d.shape === 'sphere'
      ? Rigid.sphere(solver, d.size[0], d.density, d.friction, d.pos, vel)
      : new Rigid(solver, d.size, d.density, d.friction, d.pos, vel);
    if (d.quat) { b.positionAng.set(d.quat); b.initialAng.set(d.quat); }
    if (d.angvel) b.velocityAng.set(d.angvel);
    return b;
  });
  for (const j of def.joints) {
    if (j.type === 'spring') {
      const sp = new Spring(solver, bodies[j.a], bodies[j.b], j.rA, j.rB, j.stiffness, j.rest);
      sp.tearStrain = j.tearStrain;
      continue;
    }
    new Joint(solver, j.a < 0 ? null : bodies[j.a], bodies[j.b], j.rA, j.rB, j.kLin, j.kAng, j.fracture);
  }
  return solver;
}

// Greedy colouring visiting bodies in the reference GPU's claim order
// (layout.js colorOrder: lowbias32 priority, ties to the lower index), the
// fixed point its Jacobi rounds converge to; bodies grouped by colour in index
// order, as Coloring.flatOrder() groups them.
function priorityColoring(solver) {
  const n = solver.bodies.length;
  const index = new Map();
  solver.bodies.forEach((b, i) => index.set(b, i));
  const adj = Array.from({ length: n }, () => []);
  for (const f of solver.forces) {
    const a = f.bodyA ? index.get(f.bodyA) : -1;
    const b = f.bodyB ? index.get(f.bodyB) : -1;
    if (a >= 0 && b >= 0) { adj[a].push(b); adj[b].push(a); }
  }
  const colorOf = new Array(n).fill(-1);
  let numColors = 0;
  for (const i of colorOrder(n)) {
    if (solver.bodies[i].mass <= 0) continue;
    const used = new Set();
    for (const j of adj[i]) if (colorOf[j] >= 0) used.add(colorOf[j]);
    let c = 0;
    while (used.has(c)) c++;
    colorOf[i] = c;
    numColors = Math.max(numColors, c + 1);
  }
  const order = [];
  for (let c = 0; c < numColors; c++) for (let i = 0; i < n; i++) if (colorOf[i] === c) order.push(i);
  return { colorOf, numColors, order };
}

function installColorSchedule(solver, sink, priority) {
  const coloring = new Coloring();
  let stepNo = 0, builtFor = -1, order = null;
  const step = solver.step.bind(solver);
  solver.step = () => { stepNo++; step(); };
  Object.defineProperty(solver, 'bodyOrder', {
    configurable: true,
    get() {
      if (builtFor !== stepNo) {
        if (priority) {
          const pc = priorityColoring(solver);
          order = pc.order;
          sink.colorOf = pc.colorOf;
          sink.numColors = pc.numColors;
        } else {
          const index = new Map();
          solver.bodies.forEach((b, i) => index.set(b, i));
          coloring.build(solver.bodies, solver.forces, index);
          order = Array.from(coloring.flatOrder());
          sink.colorOf = Array.from(coloring.colorOf.subarray(0, solver.bodies.length));
          sink.numColors = coloring.numColors;
        }
        builtFor = stepNo;
      }
      return order;
    },
    set() {},
  });
}

const sha = (typed) => createHash('sha256').update(Buffer.from(typed.buffer, typed.byteOffset, typed.byteLength)).digest('hex');

function stateOf(solver) {
  const out = new Float64Array(solver.bodies.length * 13);
  solver.bodies.forEach((b, i) => {
    const v = [...b.positionLin, ...b.positionAng, ...b.velocityLin, ...b.velocityAng];
    for (let k = 0; k < 13; k++) out[i * 13 + k] = v[k] === 0 ? 0 : v[k];
  });
  return out;
}

function forceListOf(solver) {
  const index = new Map();
  solver.bodies.forEach((b, i) => index.set(b, i));
  const words = [];
  for (const f of solver.forces) {
    const isM = f instanceof Manifold;
    words.push(isM ? 0 : 1, f.bodyA ? index.get(f.bodyA) : 0xffffffff, index.get(f.bodyB));
    const nc = isM ? f.numContacts : 0;
    words.push(nc);
    for (let i = 0; i < nc; i++) words.push(f.contacts[i].feature >>> 0);
  }
  return Uint32Array.from(words);
}

function run(name, def, schedule) {
  if (ONLY && !ONLY.has(name)) return;
  const solver = build(def);
  const sink = {};
  if (schedule === 'color' || schedule === 'prio') installColorSchedule(solver, sink, schedule === 'prio');
  const sampleAt = new Set([1, 2, 3, 5, 10, 20, 50, 100, 200, def.steps]);
  const steps = [];
  const samples = {};
  for (let s = 1; s <= def.steps; s++) {
    solver.step();
    const st = stateOf(solver);
    const fl = forceListOf(solver);
    const rec = { state: sha(st), forces: sha(fl), nforces: solver.forces.length };
    if (schedule !== 'reverse') { rec.colorOf = sink.colorOf; rec.numColors = sink.numColors; }
    if (s <= 3) rec.forceList = Array.from(fl);
    steps.push(rec);
    if (sampleAt.has(s)) samples[s] = Array.from(st);
    for (const e of def.edits || []) if (e.step === s) applyEdit(solver, e);
  }
  const enc = (x) => (x === Infinity ? 'inf' : x);
  const scene = {
    bodies: def.bodies,
    joints: def.joints.map((j) => (j.type === 'spring'
      ? { ...j, stiffness: enc(j.stiffness), tearStrain: enc(j.tearStrain) }
      : { ...j, kLin: enc(j.kLin), kAng: enc(j.kAng), fracture: enc(j.fracture) })),
  };
  const doc = { name, schedule, params: { dt: 1 / 60, ...def.params }, steps: def.steps, scene, perStep: steps, samples };
  if (def.edits) doc.edits = def.edits;
  writeFileSync(join(OUT, `${name}.json`), JSON.stringify(doc));
  console.log(`${name}: ${def.steps} steps, ${def.bodies.length} bodies, final forces ${solver.forces.length}`);
}

// argv[4], when given: a comma-separated list of the fixtures to write
const ONLY = process.argv[4] ? new Set(process.argv[4].split(',')) : null;

async function main() {
  ({ Solver } = await imp('src/physics/solver.js'));
  ({ Rigid } = await imp('src/physics/rigid.js'));
  ({ Joint } = await imp('src/physics/joint.js'));
  ({ Spring } = await imp('src/physics/spring.js'));
  ({ Manifold } = await imp('src/physics/manifold.js'));
  ({ Coloring } = await imp('src/physics/coloring.js'));
  ({ colorOrder } = await imp('src/physics/gpu/layout.js'));
run('stack10_reverse', sceneStack10(), 'reverse');
run('stack10_color', sceneStack10(), 'color');
run('drop_color', sceneDrop(), 'color');
run('spheres_color', sceneSpheres(), 'color');
run('spheres_reverse', sceneSpheres(), 'reverse');
run('chain_color', sceneChain(), 'color');
run('pendulum_color', scenePendulum(), 'color');
run('ramp_color', sceneRamp(), 'color');
run('rain_color', sceneRain(), 'color');
run('rain_reverse', sceneRain(), 'reverse');
run('pyramid5_color', scenePyramid(5, 'bruteforce', 60), 'color');
run('pyramid8_hash_reverse', scenePyramid(8, 'auto', 20), 'reverse');
run('pyramid5_prio', scenePyramid(5, 'bruteforce', 60), 'prio');
run('rain_prio', sceneRain(), 'prio');
run('pendulum_prio', scenePendulum(), 'prio');
run('chain_prio', sceneChain(), 'prio');
run('springs_color', sceneSprings(), 'color');
run('springs_reverse', sceneSprings(), 'reverse');
run('poststab_color', scenePostStab(), 'color');
run('bodyinertia_color', sceneBodyInertia(), 'color');
run('edits_color', sceneEdits(), 'color');
}

main().catch((e) => { console.error(e); process.exit(1); });
