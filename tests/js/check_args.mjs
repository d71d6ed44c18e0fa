// Feeds the addon's entry points arrays of the wrong element type or length
// and prints, per case, the error each call threw, as JSON on stdout. With no
// context (the first argument null) a call whose arrays are well formed
// reaches the context check; a malformed one must be rejected before it, so
// the library is never handed an array shorter than what it reads or
// writes. With `--ctx <precision>` the same calls run on a real context (a
// GPU host), where the well-formed ones must succeed.
//   node tests/js/check_args.mjs [--ctx exact]
import { createRequire } from 'module';
import { dirname, join } from 'path';
import { fileURLToPath } from 'url';

const here = dirname(fileURLToPath(import.meta.url));
const require = createRequire(import.meta.url);
const addon = require(join(here, '..', '..', 'avbd-sandbox_amd', 'js', 'avbd.node'));

const n = 2;
const good = () => ({
  shape: new Int32Array(n), size: new Float64Array(3 * n).fill(1), density: Float64Array.from([0, 1]),
  friction: new Float64Array(n).fill(0.5), pos: Float64Array.from([0, 0, -0.5, 0, 0, 1]),
  quat: Float64Array.from([0, 0, 0, 1, 0, 0, 0, 1]), vel: new Float64Array(3 * n), angvel: new Float64Array(3 * n),
});

function attempt(name, fn) {
  try {
    fn();
    return { name, ok: true };
  } catch (e) {
    return { name, ok: false, type: e.constructor.name, message: String(e.message) };
  }
}

async function main() {
  const useCtx = process.argv[2] === '--ctx';
  const ctx = useCtx ? addon.create({ precision: { exact: 0, f64: 1, mixed: 2 }[process.argv[3] || 'exact'] }) : null;
  const out = [];
  if (useCtx) addon.setParams(ctx, {});
  out.push(attempt('pack good', () => addon.pack(ctx, good())));
  out.push(attempt('pack size short', () => addon.pack(ctx, { ...good(), size: new Float64Array(3 * n - 1) })));
  out.push(attempt('pack pos float32', () => addon.pack(ctx, { ...good(), pos: new Float32Array(3 * n) })));
  out.push(attempt('pack quat long', () => addon.pack(ctx, { ...good(), quat: new Float64Array(4 * n + 4) })));
  out.push(attempt('pack shape float64', () => addon.pack(ctx, { ...good(), shape: new Float64Array(n) })));
  out.push(attempt('pack density missing', () => addon.pack(ctx, { ...good(), density: undefined })));
  out.push(attempt('pack jointB short', () => addon.pack(ctx, {
    ...good(), jointA: Int32Array.from([-1]), jointB: new Int32Array(0), jointRA: new Float64Array(3),
    jointRB: new Float64Array(3), jointStiffLin: new Float64Array(1), jointStiffAng: new Float64Array(1),
    jointFracture: new Float64Array(1),
  })));
  out.push(attempt('pack oldIndex short', () => addon.pack(ctx, good(), new Int32Array(n - 1))));
  out.push(attempt('readState good', () => addon.readState(ctx, new Float64Array(3 * n), new Float64Array(4 * n),
    new Float64Array(3 * n), new Float64Array(3 * n), new Float64Array(3 * n))));
  out.push(attempt('readState quat short', () => addon.readState(ctx, new Float64Array(3 * n), new Float64Array(4 * n - 1))));
  out.push(attempt('readState vel float32', () => addon.readState(ctx, new Float64Array(3 * n), null,
    new Float32Array(3 * n))));
  out.push(attempt('readState pos ragged', () => addon.readState(ctx, new Float64Array(3 * n + 1))));
  out.push(attempt('writeState prevVel short', () => addon.writeState(ctx, new Float64Array(3 * n), null, null, null,
    new Float64Array(3))));
  out.push(attempt('readStepStart quat int32', () => addon.readStepStart(ctx, new Float64Array(3 * n),
    new Int32Array(4 * n))));
  out.push(attempt('exportPoses float64', () => addon.exportPoses(ctx, 0.5, new Float64Array(8 * n))));
  out.push(attempt('exportPoses ragged', () => addon.exportPoses(ctx, 0.5, new Float32Array(8 * n - 3))));
  out.push(attempt('readJoints ragged', () => addon.readJoints(ctx, new Float64Array(15))));
  out.push(attempt('setJointAnchor short', () => addon.setJointAnchor(ctx, 0, new Float64Array(2))));
  out.push(attempt('readColoring uint32', () => addon.readColoring(ctx, new Uint32Array(n))));
  out.push(attempt('uploadColoring float64', () => addon.uploadColoring(ctx, new Float64Array(n))));
  out.push(attempt('setBodyOrder float32', () => addon.setBodyOrder(ctx, new Float32Array(n))));
  out.push(attempt('readPairs ragged', () => addon.readPairs(ctx, new Int32Array(3))));
  if (useCtx) {
    // a sync in flight: every other call on the context throws until it settles
    const p = addon.sync(ctx);
    out.push(attempt('step during sync', () => addon.step(ctx, 1)));
    await p;
    out.push(attempt('step after sync', () => addon.step(ctx, 1)));
    await addon.sync(ctx);
  }
  process.stdout.write(JSON.stringify(out));
}

main().catch((e) => { console.error(e); process.exit(1); });
