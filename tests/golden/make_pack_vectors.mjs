// Writes tests/golden/pack_arrays.json: what HipBackend.pack's extraction
// (avbd-sandbox_amd/js/pack_arrays.mjs, packArrays) reads from scenes built
// with the reference's own Solver, Rigid, Joint and Spring
// (src/physics/{solver,rigid,joint,spring}.js) through tests/js/pack_scenes.mjs.
// tests/js/check_pack_arrays.mjs builds the same scenes with this package's
// scene graph and must read the same arrays.
//
// Run in the development container only (the reference is not present on the
// GPU machine), from the repository root:
//
//   node --experimental-loader ./tests/golden/nullish_loader.mjs \
//        tests/golden/make_pack_vectors.mjs /root/reference tests/golden
import { writeFileSync } from 'fs';
import { join } from 'path';
import { pathToFileURL } from 'url';
import { packArrays } from '../../avbd-sandbox_amd/js/pack_arrays.mjs';
import { buildScenes } from '../js/pack_scenes.mjs';
import { encodeArrays } from '../js/pack_encode.mjs';

const REF = process.argv[2] || '/root/reference';
const OUT = process.argv[3] || 'tests/golden';
const imp = (p) => import(pathToFileURL(join(REF, p)).href);

async function main() {
  const { Solver } = await imp('src/physics/solver.js');
  const { Rigid } = await imp('src/physics/rigid.js');
  const { Joint } = await imp('src/physics/joint.js');
  const { Spring } = await imp('src/physics/spring.js');
  const scenes = buildScenes({ Solver, Rigid, Joint, Spring });
  const doc = {};
  for (const [name, solver] of Object.entries(scenes)) doc[name] = encodeArrays(packArrays(solver));
  writeFileSync(join(OUT, 'pack_arrays.json'), JSON.stringify(doc, null, 1));
  console.log(`pack_arrays.json: ${Object.keys(doc).join(', ')}`);
}

main().catch((e) => { console.error(e); process.exit(1); });
