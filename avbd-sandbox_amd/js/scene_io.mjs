// The `.scene` interchange format and the %.17g trajectory dump, for hosts
// that hold the scene graph in JavaScript.
//
// loadScene(solver, path) builds Rigid and Joint objects from a `.scene` file
// record for record, in creation order (the reader of
// test/parity/scene_loader.mjs:20-84 and reference_main.cpp:45-124; format
// gen_scenes.mjs:8-14). dumpState(bodies, step) returns the lines reference_main.cpp
// prints after each step (lines 126-140), numbers formatted exactly as C's
// printf %.17g would, so dumps written here and by a C program compare as
// text.
import { readFileSync } from 'node:fs';
import { Joint, Rigid, Spring } from './scene_graph.mjs';

function value(tok) {
  if (tok === 'inf') return Infinity;
  if (tok === '-inf') return -Infinity;
  const v = Number(tok);
  if (Number.isNaN(v) && tok !== 'nan') throw new Error(`not a number: ${tok}`);
  return v;
}

export function parseScene(solver, text) {
  const bodies = [];
  text.split('\n').forEach((raw, i) => {
    const tok = raw.split('#')[0].split(/[ \t\r]+/).filter((t) => t.length > 0);
    if (tok.length === 0) return;
    const line = i + 1;
    if (tok[0] === 'body') {
      if (tok.length < 19) throw new Error(`line ${line}: body record needs 18 numbers`);
      const v = tok.slice(1, 19).map(value);
      const b = new Rigid(solver, v.slice(0, 3), v[3], v[4], v.slice(5, 8), v.slice(8, 11));
      b.positionAng.set(v.slice(11, 15));
      b.initialAng = Float64Array.from(b.positionAng);
      b.velocityAng.set(v.slice(15, 18));
      bodies.push(b);
    } else if (tok[0] === 'joint') {
      if (tok.length < 12) throw new Error(`line ${line}: joint record needs 11 fields`);
      const ia = Math.trunc(value(tok[1]));
      const ib = Math.trunc(value(tok[2]));
      if (!(ib >= 0 && ib < bodies.length) || ia >= bodies.length) {
        throw new Error(`line ${line}: joint names a body not yet defined`);
      }
      const v = tok.slice(3, 12).map(value);
      new Joint(solver, ia < 0 ? null : bodies[ia], bodies[ib], v.slice(0, 3), v.slice(3, 6), v[6], v[7], v[8]);
    } else if (tok[0] === 'spring') {
      if (tok.length < 11) throw new Error(`line ${line}: spring record needs 10 fields`);
      const ia = Math.trunc(value(tok[1]));
      const ib = Math.trunc(value(tok[2]));
      if (!(ia >= 0 && ia < bodies.length) || !(ib >= 0 && ib < bodies.length)) {
        throw new Error(`line ${line}: spring names a body not yet defined`);
      }
      const v = tok.slice(3, 11).map(value);
      new Spring(solver, bodies[ia], bodies[ib], v.slice(0, 3), v.slice(3, 6), v[6], v[7]);
    } else {
      throw new Error(`line ${line}: unknown record: ${tok[0]}`);
    }
  });
  return bodies;
}

export function loadScene(solver, path) {
  return parseScene(solver, readFileSync(path, 'utf8'));
}

const P = 17;

function stripZeros(s) {
  if (!s.includes('.')) return s;
  return s.replace(/0+$/, '').replace(/\.$/, '');
}

/**
 * C's printf("%.17g", v). Number's own formatters round ties away from zero
 * and keep trailing zeros; %g rounds the exact binary value half to even and
 * drops them. So the exact decimal value is formed with BigInt (the double is
 * m * 2^e), rounded to 17 significant digits, and laid out by the %g rule:
 * fixed notation when the decimal exponent X satisfies -4 <= X < 17, else
 * d.ddde+XX with at least two exponent digits.
 */
export function formatG17(v) {
  if (Number.isNaN(v)) return 'nan';
  if (v === Infinity) return 'inf';
  if (v === -Infinity) return '-inf';
  const neg = v < 0 || Object.is(v, -0);
  if (v === 0) return neg ? '-0' : '0';
  const dv = new DataView(new ArrayBuffer(8));
  dv.setFloat64(0, Math.abs(v));
  const hi = dv.getUint32(0);
  const lo = dv.getUint32(4);
  const be = (hi >>> 20) & 0x7ff;
  let mant = (BigInt(hi & 0xfffff) << 32n) | BigInt(lo);
  let exp = -1074;
  if (be !== 0) {
    mant |= 1n << 52n;
    exp = be - 1075;
  }
  // |v| = digits * 10^-k exactly
  let digits;
  let k = 0;
  if (exp >= 0) digits = (mant << BigInt(exp)).toString();
  else {
    digits = (mant * 5n ** BigInt(-exp)).toString();
    k = -exp;
  }
  let X = digits.length - 1 - k;
  let s;
  if (digits.length > P) {
    let q = BigInt(digits.slice(0, P));
    const first = digits.charCodeAt(P) - 48;
    const beyond = /[1-9]/.test(digits.slice(P + 1));
    if (first > 5 || (first === 5 && (beyond || (q & 1n) === 1n))) q += 1n;
    s = q.toString();
    if (s.length > P) {
      s = s.slice(0, P);
      X += 1;
    }
  } else {
    s = digits.padEnd(P, '0');
  }
  let out;
  if (X >= -4 && X < P) {
    out = X >= 0 ? `${s.slice(0, X + 1)}.${s.slice(X + 1)}` : `0.${'0'.repeat(-X - 1)}${s}`;
    out = stripZeros(out);
  } else {
    out = `${stripZeros(`${s[0]}.${s.slice(1)}`)}e${X < 0 ? '-' : '+'}${String(Math.abs(X)).padStart(2, '0')}`;
  }
  return neg ? `-${out}` : out;
}

/** One frame of the trajectory dump: "step index p[3] q[4] v[3] w[3]" per body. */
export function dumpState(bodies, step) {
  let out = '';
  bodies.forEach((b, i) => {
    const f = [...b.positionLin, ...b.positionAng, ...b.velocityLin, ...b.velocityAng];
    out += `${step} ${i} ${f.map(formatG17).join(' ')}\n`;
  });
  return out;
}
