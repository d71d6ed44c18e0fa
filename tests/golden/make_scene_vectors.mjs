// Writes tests/golden/scenes/reference_loader.json: what the reference's own
// `.scene` reader (test/parity/scene_loader.mjs) builds from each fixture under
// tests/golden/scenes/, as data (per body: size, density, friction, mass,
// position, quaternion, velocity, angular velocity; per joint and spring:
// bodies, anchors and parameters). Run in the development container only (the
// reference is not present on the GPU machine), from the repository root:
//
//   node --experimental-loader ./tests/golden/nullish_loader.mjs \
//        tests/golden/make_scene_vectors.mjs /root/reference tests/golden
import { readdirSync, writeFileSync } from 'fs';
import { join } from 'path';
import { pathToFileURL } from 'url';

const REF = process.argv[2] || '/root/reference';
const OUT = process.argv[3] || 'tests/golden';
const imp = (p) => import(pathToFileURL(join(REF, p)).href);

async function main() {
  const { Solver } = await imp('src/physics/solver.js');
  const { loadScene } = await imp('test/parity/scene_loader.mjs');
  const { Joint } = await imp('src/physics/joint.js');
  const { Spring } = await imp('src/physics/spring.js');
  const dir = join(OUT, 'scenes');
  const enc = (x) => (x === Infinity ? 'inf' : x === -Infinity ? '-inf' : x);
  const doc = {};
  for (const f of readdirSync(dir).filter((n) => n.endsWith('.scene')).sort()) {
    const solver = new Solver();
    const bodies = loadScene(solver, join(dir, f));
    const index = new Map(bodies.map((b, i) => [b, i]));
    doc[f.replace('.scene', '')] = {
      bodies: bodies.map((b) => [...b.size, b.density, b.friction, b.mass, ...b.positionLin, ...b.positionAng,
        ...b.velocityLin, ...b.velocityAng]),
      forces: solver.forces.map((c) => (c instanceof Spring
        ? ['spring', index.get(c.bodyA), index.get(c.bodyB), ...c.rA, ...c.rB, c.stiffness, c.rest]
        : ['joint', c.bodyA ? index.get(c.bodyA) : -1, index.get(c.bodyB), ...c.rA, ...c.rB,
          enc(c.stiffnessLin), enc(c.stiffnessAng), enc(c.fracture)])),
    };
  }
  writeFileSync(join(dir, 'reference_loader.json'), JSON.stringify(doc));
  console.log(`${Object.keys(doc).length} scenes`);
}

main().catch((e) => { console.error(e); process.exit(1); });
