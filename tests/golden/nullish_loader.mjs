// ES-module loader hook for Node 12: rewrites `a ?? b` to the equivalent
// conditional at load time, so the reference engine's modules (which use `??`
// at four sites, all of the form `identifier ?? member.expression`) parse under
// Node 12. Nothing is written to disk; used only by make_golden.mjs.
export async function transformSource(source, context, defaultTransformSource) {
  const text = typeof source === 'string' ? source : Buffer.from(source).toString('utf8');
  const lowered = text.replace(
    /([A-Za-z_$][\w$.]*)\s*\?\?\s*([A-Za-z_$][\w$.]*)/g,
    '((($1) !== null && ($1) !== undefined) ? ($1) : ($2))'
  );
  return defaultTransformSource(lowered, context, defaultTransformSource);
}
