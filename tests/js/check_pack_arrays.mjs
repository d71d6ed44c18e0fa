// Builds tests/js/pack_scenes.mjs with this package's scene graph and prints,
// as JSON, the differences between what packArrays reads from it and what it
// read from the reference's own classes (tests/golden/pack_arrays.json).
import { readFileSync } from 'fs';
import { dirname, join } from 'path';
import { fileURLToPath } from 'url';
import { packArrays } from '../../avbd-sandbox_amd/js/pack_arrays.mjs';
import { Joint, Rigid, Solver, Spring } from '../../avbd-sandbox_amd/js/scene_graph.mjs';
import { buildScenes } from './pack_scenes.mjs';
import { encodeArrays } from './pack_encode.mjs';

const here = dirname(fileURLToPath(import.meta.url));
const want = JSON.parse(readFileSync(join(here, '..', 'golden', 'pack_arrays.json'), 'utf8'));
const got = {};
for (const [name, solver] of Object.entries(buildScenes({ Solver, Rigid, Joint, Spring }))) {
  got[name] = encodeArrays(packArrays(solver));
}
const diffs = [];
for (const name of Object.keys(want)) {
  for (const part of ['arrays', 'params']) {
    const w = want[name][part], g = (got[name] || {})[part] || {};
    for (const k of new Set([...Object.keys(w), ...Object.keys(g)])) {
      if (JSON.stringify(w[k]) !== JSON.stringify(g[k])) diffs.push({ name, part, key: k, want: w[k], got: g[k] });
    }
  }
}
process.stdout.write(JSON.stringify({ scenes: Object.keys(want), diffs }));
