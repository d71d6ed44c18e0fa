// The arrays and parameters HipBackend.pack hands the library, read from a
// Solver's scene graph. No device call and no addon: a host, a test or a
// fixture script can run it on any Solver (the reference's own included).
export const num = (v, d) => (typeof v === 'number' ? v : d);
// A joint-like force is a spring when it carries spring.js's fields
// (stiffness, rest) and no joint stiffness.
export const isSpring = (f) => 'stiffness' in f && 'rest' in f && !('stiffnessLin' in f);

/**
 * What pack hands avbd_pack and avbd_set_params, read from a Solver's scene
 * graph (GpuBackend.pack's arenas, backend.js:662-1301): body arrays named as
 * avbd_scene's fields (the GPU law's material from Rigid.setMaterial's fields,
 * rigid.js:144-214, a null offset as NaN: the solver's), the joints and
 * springs of forces[] in creation order, the parameters (solver.js:92-236),
 * and each body's index. Pure: no device call.
 */
export function packArrays(solver) {
  const bodies = solver.bodies;
  const n = bodies.length;
  const A = {
    shape: new Int32Array(n), size: new Float64Array(n * 3), density: new Float64Array(n),
    friction: new Float64Array(n), pos: new Float64Array(n * 3), quat: new Float64Array(n * 4),
    vel: new Float64Array(n * 3), angvel: new Float64Array(n * 3),
    staticFriction: new Float64Array(n), dynamicFriction: new Float64Array(n), restitution: new Float64Array(n),
    contactOffset: new Float64Array(n), restOffset: new Float64Array(n),
  };
  const index = new Map();
  bodies.forEach((b, i) => {
    index.set(b, i);
    A.shape[i] = b.shape === 'sphere' ? 1 : 0;
    A.size.set(b.size, i * 3);
    A.density[i] = b.density;
    A.friction[i] = b.friction;
    A.pos.set(b.positionLin, i * 3);
    A.quat.set(b.positionAng, i * 4);
    A.vel.set(b.velocityLin, i * 3);
    A.angvel.set(b.velocityAng, i * 3);
    A.staticFriction[i] = num(b.staticFriction, b.friction);
    A.dynamicFriction[i] = num(b.dynamicFriction, b.friction);
    A.restitution[i] = num(b.restitution, 0);
    A.contactOffset[i] = num(b.contactOffset, NaN);
    A.restOffset[i] = num(b.restOffset, NaN);
  });
  const joints = solver.forces.filter((f) => f.rA && f.rB && ('stiffnessLin' in f || isSpring(f)));
  const m = joints.length;
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
      A.jointA[k] = j.bodyA ? index.get(j.bodyA) : -1;
      A.jointB[k] = index.get(j.bodyB);
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
  const params = {
    dt: solver.dt, gravity: solver.gravity, iterations: solver.iterations, betaLin: solver.betaLin,
    betaAng: solver.betaAng, alpha: solver.alpha, gamma: solver.gamma, penaltyMin: solver.penaltyMin,
    penaltyMax: solver.penaltyMax, collisionMargin: solver.collisionMargin,
    stickThreshold: solver.stickThreshold, postStabilize: solver.postStabilize ? 1 : 0,
    rotatedInertia: solver.rotatedInertia === false ? 0 : 1, cachedContactJacobians: 1,
    colorOrder: solver.colorOrder === 'priority' || solver.colorOrder === 1 ? 1 : 0,
    // the GPU backend's contact law when the solver asks for it (this
    // backend's addition: contactModel 'gpu'), with the solver's offsets
    contactModel: solver.contactModel === 'gpu' || solver.contactModel === 1 ? 1 : 0,
    contactOffset: num(solver.contactOffset, 0.005), restOffset: num(solver.restOffset, -0.005),
    restitutionThreshold: num(solver.restitutionThreshold, 1.0),
    contactPersistence: num(solver.contactPersistence, 0.002),
  };
  return { arrays: A, index, joints, params };
}

