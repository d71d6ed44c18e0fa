// packArrays' result as plain JSON: typed arrays as lists, non-finite numbers
// as the strings 'nan', 'inf' and '-inf'.
const enc = (x) => (Number.isNaN(x) ? 'nan' : x === Infinity ? 'inf' : x === -Infinity ? '-inf' : x);

export function encodeArrays({ arrays, params }) {
  const out = { arrays: {}, params: {} };
  for (const [k, v] of Object.entries(arrays)) out.arrays[k] = Array.from(v, enc);
  for (const [k, v] of Object.entries(params)) out.params[k] = enc(v);
  return out;
}
