// Drives the JavaScript `.scene` reader and trajectory dump. Usage:
//   node tests/js/run_scene.mjs format <numbers.json>   one %.17g string per line
//   node tests/js/run_scene.mjs dump0 <file.scene>      the frame of step 0
//   node tests/js/run_scene.mjs run <file.scene> <steps> <precision>
//                                                        steps on the GPU, a frame per step
//   node tests/js/run_scene.mjs repack <file.scene> <steps> <precision>
//                                                        steps, erases and appends, steps (JSON)
//   node tests/js/run_scene.mjs poses <file.scene> <steps> <precision>
//                                                        the poses drawn at alpha 0.5 (JSON)
import { readFileSync } from 'fs';
import { dumpState, formatG17, loadScene } from '../../avbd-sandbox_amd/js/scene_io.mjs';
import { Rigid, Solver } from '../../avbd-sandbox_amd/js/scene_graph.mjs';

async function main() {
  const [mode, path, steps, precision] = process.argv.slice(2);
  if (mode === 'format') {
    const vals = JSON.parse(readFileSync(path, 'utf8')).map((x) =>
      (x === 'inf' ? Infinity : x === '-inf' ? -Infinity : x === '-0' ? -0 : x));
    process.stdout.write(vals.map(formatG17).join('\n') + '\n');
    return;
  }
  const solver = new Solver();
  const bodies = loadScene(solver, path);
  if (mode === 'dump0') {
    process.stdout.write(dumpState(bodies, 0));
    return;
  }
  const { HipBackend } = await import('../../avbd-sandbox_amd/js/hip_backend.mjs');
  const backend = new HipBackend({ precision: precision || 'exact' });
  await backend.init();
  if (mode === 'poses') {
    // steps, then the poses at alpha = 0.5 and the two states they blend
    const pose = (bs) => bs.map((b) => [...b.positionLin, ...b.positionAng]);
    for (let s = 1; s < Number(steps); s++) backend.step(solver);
    await backend.syncToCPU(solver);
    const s0 = pose(bodies);
    backend.step(solver);
    await backend.syncToCPU(solver);
    const poses = await backend.exportPoses(0.5);
    process.stdout.write(JSON.stringify({ s0, s1: pose(bodies), poses: Array.from(poses) }));
    return;
  }
  if (mode === 'repack') {
    // steps, erases body 8 and appends a cube (a topology edit that keeps
    // state), steps again; prints the final frame and whether the pack kept
    // the contact state
    for (let s = 0; s < Number(steps); s++) backend.step(solver);
    solver.removeBodies((b) => b === bodies[8]);
    new Rigid(solver, [1, 1, 1], 1, 0.5, [3, 0, -15]);
    backend.prepareTopologyChange();
    for (let s = 0; s < Number(steps); s++) backend.step(solver);
    await backend.syncToCPU(solver);
    process.stdout.write(JSON.stringify({ kept: backend.lastPackPreservedContacts,
      frame: dumpState(solver.bodies, 2 * Number(steps)) }));
    return;
  }
  let out = '';
  for (let s = 1; s <= Number(steps); s++) {
    backend.step(solver);
    await backend.syncToCPU(solver);
    out += dumpState(bodies, s);
  }
  process.stdout.write(out);
}

main().catch((e) => { console.error(e); process.exit(1); });
