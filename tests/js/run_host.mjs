// Drives the JavaScript host end to end: builds the scene graph from a JSON
// scene description, packs it into HipBackend, steps, and prints the final
// state as JSON on stdout. Usage:
//   node tests/js/run_host.mjs <scene.json> <steps> <precision>
import { readFileSync } from 'fs';
import { HipBackend } from '../../avbd-sandbox_amd/js/hip_backend.mjs';
import { fromDescription } from '../../avbd-sandbox_amd/js/scene_graph.mjs';

async function main() {
  const [path, steps, precision] = process.argv.slice(2);
  const doc = JSON.parse(readFileSync(path, 'utf8'));
  const solver = fromDescription(doc.scene, doc.params);
  const backend = new HipBackend({ precision: precision || 'exact' });
  await backend.init();
  backend.setProfiling(true);
  for (let s = 0; s < Number(steps); s++) backend.step(solver);
  await backend.idle();
  await backend.syncToCPU(solver);
  const col = await backend.readColoring();
  const state = solver.bodies.map((b) => [...b.positionLin, ...b.positionAng, ...b.velocityLin, ...b.velocityAng]);
  process.stdout.write(JSON.stringify({
    state, numColors: col.numColors, colorOf: Array.from(col.colorOf), pairs: col.broadPhasePairs.length / 2,
    materials: col.slotInfo.map((sl) => [sl.staticFriction, sl.dynamicFriction]),
    bias: col.slotInfo.flatMap((sl) => sl.contacts.map((c) => c.restitutionBias)),
    stats: backend.lastStats, profile: backend.lastProfile, adapter: backend.adapterInfo,
  }));
}

main().catch((e) => { console.error(e); process.exit(1); });
