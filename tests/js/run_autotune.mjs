// The launch-shape surface of the JavaScript host: packs a scene, steps,
// autotunes, then fixes a shape and steps on; prints what it saw as JSON.
// Usage: node tests/js/run_autotune.mjs <scene.json>
import { readFileSync } from 'fs';
import { HipBackend } from '../../avbd-sandbox_amd/js/hip_backend.mjs';
import { fromDescription } from '../../avbd-sandbox_amd/js/scene_graph.mjs';

async function main() {
  const doc = JSON.parse(readFileSync(process.argv[2], 'utf8'));
  const solver = fromDescription(doc.scene, doc.params);
  const backend = new HipBackend({ precision: 'mixed' });
  await backend.init();
  for (let s = 0; s < 2; s++) backend.step(solver);
  await backend.idle();
  const tuned = backend.autotune();
  backend.setLaunchShape(8, 128);
  for (let s = 0; s < 2; s++) backend.step(solver);
  await backend.idle();
  let rejected = null;
  try {
    backend.setLaunchShape(8, 96);
  } catch (e) {
    rejected = String(e.message || e);
  }
  await backend.syncToCPU(solver);
  const finite = solver.bodies.every((b) => [...b.positionLin, ...b.positionAng].every(Number.isFinite));
  process.stdout.write(JSON.stringify({ tuned, rejected, finite, steps: backend.lastStats.steps }));
}

main().catch((e) => { console.error(e); process.exit(1); });
