// Scenes built through the Solver / Rigid / Joint / Spring constructors and
// setMaterial, with whichever classes are passed in: the reference's own
// (tests/golden/make_pack_vectors.mjs) or this package's scene graph
// (tests/js/check_pack_arrays.mjs). The same calls must give the same arrays.
export function buildScenes({ Solver, Rigid, Joint, Spring }) {
  const scenes = {};

  {
    // boxes and spheres, orientations and velocities, materials
    const s = new Solver();
    s.defaultParams();
    s.iterations = 7;
    s.gravity = -9.81;
    s.postStabilize = true;
    new Rigid(s, [20, 20, 1], 0, 0.7, [0, 0, -0.5]);
    const a = new Rigid(s, [1, 2, 0.5], 2.5, 0.4, [0.5, -1, 1.2], [1, 0, -2]);
    a.positionAng.set([0.1, 0.2, 0.3, 0.9273618495495703]);
    a.velocityAng.set([0.5, -1, 2]);
    a.setMaterial({ staticFriction: 0.9, dynamicFriction: 0.3, restitution: 0.6 });
    const b = Rigid.sphere(s, 0.8, 3, 0.5, [2, 2, 3], [0, -1, 0]);
    b.setMaterial({ contactOffset: 0.02, restOffset: 0.001 });
    const c = new Rigid(s, [0.5, 0.5, 0.5], 1, 0.6, [-2, 1, 0.25]);
    c.setMaterial({ friction: 0.25, restitution: 1 });
    scenes.shapes = s;
  }

  {
    // joints (world-anchored, body-body, breakable, angular) and springs
    // (given rest, current length, tearing) in creation order
    const s = new Solver();
    s.defaultParams();
    s.rotatedInertia = false;
    const anchor = new Rigid(s, [0.5, 0.5, 0.5], 0, 0.5, [0, 0, 5]);
    const links = [];
    for (let i = 0; i < 4; i++) links.push(new Rigid(s, [1, 0.3, 0.3], 1, 0.5, [0.75 + i, 0, 5]));
    new Joint(s, anchor, links[0], [0.25, 0, 0], [-0.5, 0, 0]);
    new Joint(s, links[0], links[1], [0.5, 0, 0], [-0.5, 0, 0], Infinity, 1000, 500);
    const spring = new Spring(s, links[1], links[2], [0.5, 0, 0], [-0.5, 0, 0], 800, 0.1);
    spring.tearStrain = 0.5;
    new Joint(s, null, links[3], [4, 0, 6], [0.5, 0, 0], 5e5, 0);
    new Spring(s, links[2], links[3], [0, 0, 0.15], [0, 0, -0.15], 50);
    scenes.joints = s;
  }
  return scenes;
}
