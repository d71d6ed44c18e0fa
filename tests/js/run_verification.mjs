// The verification surface of the JS host against a scene: packs it, steps,
// and prints as JSON what GpuBackend's verification readbacks return
// (readState: the body arena through its layout, counters, dispatch;
// readColoring: colorOf, counts, offsets, entries, contactPairs, slots,
// manifolds, slotInfo), then the same after steps under the colouring
// controls: encColors (a colour cap), colorRounds (a round cap) and an
// uploaded colouring with overrideColoring (the gputest.js pattern).
//   node tests/js/run_verification.mjs <scene.json> <steps>
import { readFileSync } from 'fs';
import { HipBackend } from '../../avbd-sandbox_amd/js/hip_backend.mjs';
import { fromDescription } from '../../avbd-sandbox_amd/js/scene_graph.mjs';

async function main() {
  const [path, steps] = process.argv.slice(2);
  const doc = JSON.parse(readFileSync(path, 'utf8'));
  const solver = fromDescription(doc.scene, doc.params);
  const backend = new HipBackend({ precision: 'exact' });
  await backend.init();
  for (let s = 0; s < Number(steps); s++) backend.step(solver);
  const st = await backend.readState();
  const n = st.numBodies;
  const pos = [], quat = [];
  for (let i = 0; i < n; i++) {
    pos.push([0, 1, 2].map((k) => st.data[st.layout.body.bPos + i * 4 + k]));
    quat.push([0, 1, 2, 3].map((k) => st.data[st.layout.body.bQuat + i * 4 + k]));
  }
  const mass = Array.from({ length: n }, (_, i) => st.data[st.layout.body.bStatic + i * 16]);
  const col = await backend.readColoring();
  const out = {
    pos, quat, mass, counters: st.counters, dispatch: st.dispatch, numBodies: n,
    colorOf: Array.from(col.colorOf), counts: Array.from(col.counts), offsets: Array.from(col.offsets),
    entries: Array.from(col.entries), contactPairs: col.contactPairs, slots: col.slots, manifolds: col.manifolds,
    slotInfo: col.slotInfo, numColors: col.numColors,
  };
  // a colour cap of 2: colours above take the last one (conflicts tolerated)
  backend.encColors = 2;
  backend.step(solver);
  const capped = await backend.readColoring();
  out.capped = { maxColor: Math.max(...capped.colorOf), overflow: (await backend.readState()).counters.overflow };
  backend.encColors = null;
  // the caller's colouring, as gputest.js installs it: one colour per body in
  // index order (a proper colouring of any graph)
  const entries = new Uint32Array(n), offsets = new Uint32Array(n + 1), counts = new Uint32Array(n);
  let k = 0;
  for (let i = 0; i < n; i++) {
    offsets[k] = k;
    if (solver.bodies[i].mass > 0) { entries[k] = i; counts[k] = 1; k++; }
  }
  for (let c = k; c <= n; c++) offsets[c] = k;
  backend.overrideColoring = true;
  backend.uploadColoring(entries.subarray(0, k), offsets.subarray(0, Math.min(k + 1, 257)), counts.subarray(0, k));
  backend.step(solver);
  const forced = await backend.readColoring();
  out.forced = { colorOf: Array.from(forced.colorOf) };
  backend.overrideColoring = null;
  backend.step(solver);
  const greedy = await backend.readColoring();
  out.released = { colorOf: Array.from(greedy.colorOf) };
  process.stdout.write(JSON.stringify(out));
}

main().catch((e) => { console.error(e); process.exit(1); });
