// Minimal scene graph with the reference's Solver / Rigid / Joint shapes
// (src/physics/solver.js:59-236, rigid.js:29-133, joint.js:78-113), so that a
// host written against the reference can hand its solver to HipBackend.pack()
// and .step() unchanged. Only the fields the backend reads are modelled; the
// dynamics run on the GPU.

export class Solver {
  constructor() {
    this.bodies = [];
    this.forces = [];
    this.defaultParams();
  }

  defaultParams() {
    this.dt = 1.0 / 60.0;
    this.gravity = -10.0;
    this.iterations = 10;
    this.betaLin = 10000.0;
    this.betaAng = 100.0;
    this.alpha = 0.95;
    this.postStabilize = false;
    this.gamma = 0.999;
    this.penaltyMin = 1.0;
    this.penaltyMax = 1.0e10;
    this.collisionMargin = 0.01;
    this.stickThreshold = 0.00001;
    this.rotatedInertia = true;
    this.cachedContactJacobians = true;
    // the GPU backend's contact law (solver.js:142-162); contactModel 'gpu'
    // selects it on this backend ('cpu', the CPU engine's, by default)
    this.contactModel = 'cpu';
    this.contactOffset = 0.005;
    this.restOffset = -0.005;
    this.restitutionThreshold = 1.0;
    this.contactPersistence = 0.002;
  }

  /**
   * Remove every body `predicate` accepts and every force touching one,
   * keeping the order of the surviving bodies and forces (solver.js:261-300).
   * A backend told of the change (prepareTopologyChange) keeps the
   * survivors' state on its next pack.
   */
  removeBodies(predicate) {
    const doomed = new Set(this.bodies.filter((b) => predicate(b)));
    if (doomed.size === 0) return 0;
    const retained = (f) => !doomed.has(f.bodyA) && !doomed.has(f.bodyB);
    this.forces = this.forces.filter(retained);
    this.bodies = this.bodies.filter((b) => {
      if (doomed.has(b)) {
        b.forces.length = 0;
        return false;
      }
      b.forces = b.forces.filter(retained);
      return true;
    });
    return doomed.size;
  }
}

export class Rigid {
  constructor(solver, size, density, friction, position, velocity = [0, 0, 0]) {
    this.solver = solver;
    this.shape = 'box';
    this.forces = [];
    this.size = Float64Array.from(size);
    this.density = density;
    this.friction = friction;
    // contact material of the GPU law (rigid.js:56-70): null offsets inherit
    // the solver's
    this.staticFriction = friction;
    this.dynamicFriction = friction;
    this.restitution = 0;
    this.contactOffset = null;
    this.restOffset = null;
    this.positionLin = Float64Array.from(position);
    this.positionAng = Float64Array.from([0, 0, 0, 1]);
    this.velocityLin = Float64Array.from(velocity);
    this.velocityAng = new Float64Array(3);
    this.prevVelocityLin = Float64Array.from(velocity);
    this.mass = size[0] * size[1] * size[2] * density;
    solver.bodies.push(this);
  }

  static sphere(solver, diameter, density, friction, position, velocity = [0, 0, 0]) {
    if (!Number.isFinite(diameter) || diameter <= 0) {
      throw new RangeError('sphere diameter must be a finite positive number');
    }
    const b = new Rigid(solver, [diameter, diameter, diameter], density, friction, position, velocity);
    b.shape = 'sphere';
    const r = diameter * 0.5;
    b.mass = (4 / 3) * Math.PI * r * r * r * density;
    return b;
  }

  get isStatic() {
    return this.mass <= 0;
  }

  /**
   * Rigid.setMaterial (rigid.js:144-214): the whole candidate material is
   * validated before any field changes; `friction` sets both coefficients.
   */
  setMaterial({ friction, staticFriction, dynamicFriction, restitution, contactOffset, restOffset } = {}) {
    const nonNegative = (name, v) => {
      if (!Number.isFinite(v) || v < 0) throw new RangeError(`${name} must be a finite non-negative number`);
      return v;
    };
    let f = this.friction, sf = this.staticFriction, df = this.dynamicFriction, e = this.restitution;
    let co = this.contactOffset, ro = this.restOffset;
    if (friction !== undefined) f = sf = df = nonNegative('friction', friction);
    if (staticFriction !== undefined) sf = nonNegative('staticFriction', staticFriction);
    if (dynamicFriction !== undefined) df = nonNegative('dynamicFriction', dynamicFriction);
    if (restitution !== undefined) {
      if (!Number.isFinite(restitution) || restitution < 0 || restitution > 1) {
        throw new RangeError('restitution must be a finite number in [0, 1]');
      }
      e = restitution;
    }
    if (contactOffset !== undefined) co = contactOffset === null ? null : nonNegative('contactOffset', contactOffset);
    if (restOffset !== undefined) {
      if (restOffset !== null && !Number.isFinite(restOffset)) {
        throw new RangeError('restOffset must be null or a finite number');
      }
      ro = restOffset;
    }
    const resolvedContact = co === null ? this.solver.contactOffset : co;
    const resolvedRest = ro === null ? this.solver.restOffset : ro;
    if (df > sf) throw new RangeError('dynamicFriction must not exceed staticFriction');
    if (resolvedRest > resolvedContact) throw new RangeError('restOffset must not exceed contactOffset');
    this.friction = f;
    this.staticFriction = sf;
    this.dynamicFriction = df;
    this.restitution = e;
    this.contactOffset = co;
    this.restOffset = ro;
    return this;
  }
}

export class Joint {
  constructor(solver, bodyA, bodyB, rA, rB, stiffnessLin = Infinity, stiffnessAng = 0, fracture = Infinity) {
    this.solver = solver;
    this.bodyA = bodyA;
    this.bodyB = bodyB;
    this.rA = Float64Array.from(rA);
    this.rB = Float64Array.from(rB);
    this.stiffnessLin = stiffnessLin;
    this.stiffnessAng = stiffnessAng;
    this.fracture = fracture;
    this.broken = false;
    solver.forces.push(this);
    if (bodyA) bodyA.forces.push(this);
    bodyB.forces.push(this);
  }
}

/** Distance spring (spring.js:63-105): rest < 0 takes the anchors' current
 * distance; tearStrain (tensile strain at which it tears) starts at Infinity. */
export class Spring {
  constructor(solver, bodyA, bodyB, rA, rB, stiffness, rest = -1) {
    this.solver = solver;
    this.bodyA = bodyA;
    this.bodyB = bodyB;
    this.rA = Float64Array.from(rA);
    this.rB = Float64Array.from(rB);
    this.stiffness = stiffness;
    if (rest < 0) {
      const pA = transformPoint(bodyA.positionLin, bodyA.positionAng, this.rA);
      const pB = transformPoint(bodyB.positionLin, bodyB.positionAng, this.rB);
      const d = [pA[0] - pB[0], pA[1] - pB[1], pA[2] - pB[2]];
      this.rest = Math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    } else {
      this.rest = rest;
    }
    this.tearStrain = Infinity;
    this.broken = false;
    solver.forces.push(this);
    bodyA.forces.push(this);
    bodyB.forces.push(this);
  }
}

// p + q v q* (maths.js:604-627)
function transformPoint(p, q, v) {
  const [ux, uy, uz, w] = q;
  const [vx, vy, vz] = v;
  const tx = (uy * vz - uz * vy) * 2;
  const ty = (uz * vx - ux * vz) * 2;
  const tz = (ux * vy - uy * vx) * 2;
  return [vx + tx * w + (uy * tz - uz * ty) + p[0], vy + ty * w + (uz * tx - ux * tz) + p[1],
    vz + tz * w + (ux * ty - uy * tx) + p[2]];
}

/** Build a scene graph from the JSON scene description the Python side uses. */
export function fromDescription(desc, params = {}) {
  const solver = new Solver();
  Object.assign(solver, params);
  const num = (x) => (x === 'inf' ? Infinity : x === '-inf' ? -Infinity : x);
  const bodies = desc.bodies.map((d) => {
    const b = d.shape === 'sphere'
      ? Rigid.sphere(solver, d.size[0], d.density, d.friction, d.pos, d.vel || [0, 0, 0])
      : new Rigid(solver, d.size, d.density, d.friction, d.pos, d.vel || [0, 0, 0]);
    if (d.quat) b.positionAng.set(d.quat);
    if (d.angvel) b.velocityAng.set(d.angvel);
    // the GPU contact law's material, where the description gives one
    const mat = {};
    for (const k of ['staticFriction', 'dynamicFriction', 'restitution', 'contactOffset', 'restOffset']) {
      if (d[k] !== undefined && d[k] !== null) mat[k] = d[k];
    }
    if (Object.keys(mat).length) b.setMaterial(mat);
    return b;
  });
  for (const j of desc.joints || []) {
    if (j.type === 'spring') {
      const sp = new Spring(solver, bodies[j.a], bodies[j.b], j.rA, j.rB, num(j.stiffness), j.rest === undefined ? -1 : j.rest);
      sp.tearStrain = j.tearStrain === undefined ? Infinity : num(j.tearStrain);
      continue;
    }
    new Joint(solver, j.a < 0 ? null : bodies[j.a], bodies[j.b], j.rA, j.rB, num(j.kLin), num(j.kAng),
      num(j.fracture));
  }
  return solver;
}
