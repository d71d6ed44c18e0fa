// avbd_oracle.cpp — CPU f64 restatement of the reference AVBD step.
//
// TEST INFRASTRUCTURE (the oracle): only tests, the smoke check and the CPU
// baseline leg of bench.py use it. Build with -O2 -ffp-contract=off and no
// fast-math: every expression keeps the reference's operation order, so the
// results are the reference's bit for bit (pinned by tests/golden).
//
// Citations are file:line in the reference repository.
#include "avbd_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace orc {
namespace {

// ---------------------------------------------------------------------------
// Scalar helpers with JavaScript Math.min / Math.max semantics (NaN
// propagates; min(-0, +0) = -0, max(-0, +0) = +0).
// ---------------------------------------------------------------------------
inline double jmin(double a, double b) {
  if (a != a || b != b) return std::numeric_limits<double>::quiet_NaN();
  if (a == b) return std::signbit(a) ? a : b;
  return a < b ? a : b;
}
inline double jmax(double a, double b) {
  if (a != a || b != b) return std::numeric_limits<double>::quiet_NaN();
  if (a == b) return std::signbit(a) ? b : a;
  return a > b ? a : b;
}
// maths.js:34-36 — sign(0) = 0
inline double sgn(double x) { return x < 0 ? -1.0 : x > 0 ? 1.0 : 0.0; }
// maths.js:38-40
inline double clampd(double x, double a, double b) { return jmax(a, jmin(b, x)); }

// --- vec3 (maths.js:48-191) ---
inline void v_set(double *o, double x, double y, double z) { o[0] = x; o[1] = y; o[2] = z; }
inline void v_copy(double *o, const double *a) { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; }
inline void v_add(double *o, const double *a, const double *b) {
  o[0] = a[0] + b[0]; o[1] = a[1] + b[1]; o[2] = a[2] + b[2];
}
inline void v_sub(double *o, const double *a, const double *b) {
  o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}
inline void v_scale(double *o, const double *a, double s) {
  o[0] = a[0] * s; o[1] = a[1] * s; o[2] = a[2] * s;
}
inline void v_div(double *o, const double *a, double s) {
  o[0] = a[0] / s; o[1] = a[1] / s; o[2] = a[2] / s;
}
inline void v_addScaled(double *o, const double *a, double s) {
  o[0] += a[0] * s; o[1] += a[1] * s; o[2] += a[2] * s;
}
inline void v_neg(double *o, const double *a) { o[0] = -a[0]; o[1] = -a[1]; o[2] = -a[2]; }
inline double v_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double v_lenSq(const double *a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
inline void v_cross(double *o, const double *a, const double *b) {
  const double ax = a[0], ay = a[1], az = a[2], bx = b[0], by = b[1], bz = b[2];
  o[0] = ay * bz - az * by;
  o[1] = az * bx - ax * bz;
  o[2] = ax * by - ay * bx;
}
inline void v_abs(double *o, const double *a) {
  o[0] = std::fabs(a[0]); o[1] = std::fabs(a[1]); o[2] = std::fabs(a[2]);
}
inline void v_minScalar(double *o, const double *a, double s) {
  o[0] = jmin(a[0], s); o[1] = jmin(a[1], s); o[2] = jmin(a[2], s);
}
inline void v_clampScalar(double *o, const double *a, double lo, double hi) {
  o[0] = clampd(a[0], lo, hi); o[1] = clampd(a[1], lo, hi); o[2] = clampd(a[2], lo, hi);
}

// --- mat3, row-major (maths.js:197-446) ---
inline void m_zero(double *o) { for (int i = 0; i < 9; i++) o[i] = 0.0; }
inline void m_copy(double *o, const double *a) { for (int i = 0; i < 9; i++) o[i] = a[i]; }
inline void m_diag(double *o, double a, double b, double c) {
  o[0] = a; o[1] = 0; o[2] = 0; o[3] = 0; o[4] = b; o[5] = 0; o[6] = 0; o[7] = 0; o[8] = c;
}
inline void m_addInto(double *o, const double *a) { for (int i = 0; i < 9; i++) o[i] += a[i]; }
inline void m_scale(double *o, const double *a, double s) { for (int i = 0; i < 9; i++) o[i] = a[i] * s; }
inline void m_div(double *o, const double *a, double s) { for (int i = 0; i < 9; i++) o[i] = a[i] / s; }
inline void m_neg(double *o, const double *a) { for (int i = 0; i < 9; i++) o[i] = -a[i]; }
inline void m_mulVec(double *o, const double *a, const double *v) {
  const double x = v[0], y = v[1], z = v[2];
  o[0] = a[0] * x + a[1] * y + a[2] * z;
  o[1] = a[3] * x + a[4] * y + a[5] * z;
  o[2] = a[6] * x + a[7] * y + a[8] * z;
}
inline void m_mul(double *o, const double *a, const double *b) {
  double r[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      r[i * 3 + j] = a[i * 3 + 0] * b[0 * 3 + j] + a[i * 3 + 1] * b[1 * 3 + j] + a[i * 3 + 2] * b[2 * 3 + j];
  for (int i = 0; i < 9; i++) o[i] = r[i];
}
inline void m_mulDiagRight(double *o, const double *a, double d0, double d1, double d2) {
  o[0] = a[0] * d0; o[1] = a[1] * d1; o[2] = a[2] * d2;
  o[3] = a[3] * d0; o[4] = a[4] * d1; o[5] = a[5] * d2;
  o[6] = a[6] * d0; o[7] = a[7] * d1; o[8] = a[8] * d2;
}
inline void m_transpose(double *o, const double *a) {
  double r[9] = {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]};
  for (int i = 0; i < 9; i++) o[i] = r[i];
}
inline void m_skew(double *o, const double *r) {
  o[0] = 0; o[1] = -r[2]; o[2] = r[1];
  o[3] = r[2]; o[4] = 0; o[5] = -r[0];
  o[6] = -r[1]; o[7] = r[0]; o[8] = 0;
}
// maths.js:373-381 — Section 3.5 column-norm lumping
inline void m_diagonalize(double *o, const double *m) {
  const double c0 = std::sqrt(m[0] * m[0] + m[3] * m[3] + m[6] * m[6]);
  const double c1 = std::sqrt(m[1] * m[1] + m[4] * m[4] + m[7] * m[7]);
  const double c2 = std::sqrt(m[2] * m[2] + m[5] * m[5] + m[8] * m[8]);
  m_diag(o, c0, c1, c2);
}
// maths.js:389-414 — contact basis, row 0 = normal
inline void m_orthonormal(double *o, const double *n) {
  const double nx = n[0], ny = n[1], nz = n[2];
  double tx, ty, tz;
  if (std::fabs(nx) > std::fabs(nz)) { tx = -ny; ty = nx; tz = 0; }
  else { tx = 0; ty = -nz; tz = ny; }
  const double len = std::sqrt(tx * tx + ty * ty + tz * tz);
  tx /= len; ty /= len; tz /= len;
  const double sx = ny * tz - nz * ty;
  const double sy = nz * tx - nx * tz;
  const double sz = nx * ty - ny * tx;
  o[0] = nx; o[1] = ny; o[2] = nz;
  o[3] = tx; o[4] = ty; o[5] = tz;
  o[6] = sx; o[7] = sy; o[8] = sz;
}
// maths.js:417-436 — yields R^T under row-wise multiplication
inline void m_fromQuat(double *o, const double *q) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  const double xx = x * x, yy = y * y, zz = z * z;
  const double xy = x * y, xz = x * z, yz = y * z;
  const double wx = w * x, wy = w * y, wz = w * z;
  o[0] = 1 - 2 * (yy + zz); o[1] = 2 * (xy + wz); o[2] = 2 * (xz - wy);
  o[3] = 2 * (xy - wz); o[4] = 1 - 2 * (xx + zz); o[5] = 2 * (yz + wx);
  o[6] = 2 * (xz + wy); o[7] = 2 * (yz - wx); o[8] = 1 - 2 * (xx + yy);
}

// --- quat, xyzw (maths.js:452-627) ---
inline void q_copy(double *o, const double *a) { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; o[3] = a[3]; }
inline void q_conj(double *o, const double *q) { o[0] = -q[0]; o[1] = -q[1]; o[2] = -q[2]; o[3] = q[3]; }
// maths.js:541-555 — Eq. 20, vector part of 2 a b^-1
inline void q_subtract(double *o, const double *a, const double *b) {
  const double lsq = b[0] * b[0] + b[1] * b[1] + b[2] * b[2] + b[3] * b[3];
  const double ix = -b[0] / lsq, iy = -b[1] / lsq, iz = -b[2] / lsq, iw = b[3] / lsq;
  const double ax = a[0], ay = a[1], az = a[2], aw = a[3];
  o[0] = (aw * ix + ax * iw + ay * iz - az * iy) * 2;
  o[1] = (aw * iy - ax * iz + ay * iw + az * ix) * 2;
  o[2] = (aw * iz + ax * iy - ay * ix + az * iw) * 2;
}
// maths.js:568-592 — Eq. 21, including the 0*q terms of the Hamilton product
inline void q_integrate(double *o, const double *q, const double *dw) {
  const double qx = q[0], qy = q[1], qz = q[2], qw = q[3];
  const double wx = dw[0], wy = dw[1], wz = dw[2];
  const double px = 0 * qx + wx * qw + wy * qz - wz * qy;
  const double py = 0 * qy - wx * qz + wy * qw + wz * qx;
  const double pz = 0 * qz + wx * qy - wy * qx + wz * qw;
  const double pw = 0 * qw - wx * qx - wy * qy - wz * qz;
  const double rx = qx + px * 0.5, ry = qy + py * 0.5, rz = qz + pz * 0.5, rw = qw + pw * 0.5;
  const double len = std::sqrt(rx * rx + ry * ry + rz * rz + rw * rw);
  o[0] = rx / len; o[1] = ry / len; o[2] = rz / len; o[3] = rw / len;
}
// maths.js:599-613
inline void q_rotate(double *o, const double *q, const double *v) {
  const double ux = q[0], uy = q[1], uz = q[2], w = q[3];
  const double vx = v[0], vy = v[1], vz = v[2];
  const double tx = (uy * vz - uz * vy) * 2;
  const double ty = (uz * vx - ux * vz) * 2;
  const double tz = (ux * vy - uy * vx) * 2;
  o[0] = vx + tx * w + (uy * tz - uz * ty);
  o[1] = vy + ty * w + (uz * tx - ux * tz);
  o[2] = vz + tz * w + (ux * ty - uy * tx);
}
// maths.js:621-627
inline void transformPoint(double *o, const double *p, const double *q, const double *v) {
  q_rotate(o, q, v);
  o[0] += p[0]; o[1] += p[1]; o[2] += p[2];
}

// rigid.js:228-247 — world inertia R I R^T (identity-rotated for spheres)
void worldMoment(double *o, const Body &b, bool rotated) {
  if (!rotated || b.shape == SPHERE) { m_diag(o, b.moment[0], b.moment[1], b.moment[2]); return; }
  double R[9], T[9];
  m_fromQuat(R, b.q);
  m_transpose(R, R);
  m_diag(T, b.moment[0], b.moment[1], b.moment[2]);
  m_mul(o, R, T);
  m_transpose(R, R);
  m_mul(o, o, R);
}

// ===========================================================================
// Narrow phase (collide.js)
// ===========================================================================
constexpr int MAX_POLY = 16;                      // collide.js:23
constexpr double SAT_EPS = 1.0e-6;                // collide.js:24
constexpr double SPHERE_EPS_SQ = 1.0e-20;         // collide.js:28
constexpr double PLANE_EPS = 1.0e-5;              // collide.js:29
constexpr double MERGE_DIST_SQ = 1.0e-6;          // collide.js:30
enum { AXIS_FACE_A = 0, AXIS_FACE_B = 1, AXIS_EDGE = 2 };

struct OBB { double c[3], h[3], ax[3][3]; };
struct SatAxis { int type, ia, ib; double sep; double n[3]; bool valid; };

// collide.js:107-112
void makeOBB(OBB &o, const Body &b) {
  static const double ux[3] = {1, 0, 0}, uy[3] = {0, 1, 0}, uz[3] = {0, 0, 1};
  v_copy(o.c, b.p);
  v_scale(o.h, b.size, 0.5);
  q_rotate(o.ax[0], b.q, ux);
  q_rotate(o.ax[1], b.q, uy);
  q_rotate(o.ax[2], b.q, uz);
}

// collide.js:119-129
void supportPoint(double *o, const OBB &b, const double *dir) {
  const double sx = v_dot(dir, b.ax[0]) >= 0 ? 1 : -1;
  const double sy = v_dot(dir, b.ax[1]) >= 0 ? 1 : -1;
  const double sz = v_dot(dir, b.ax[2]) >= 0 ? 1 : -1;
  v_copy(o, b.c);
  v_addScaled(o, b.ax[0], b.h[0] * sx);
  v_addScaled(o, b.ax[1], b.h[1] * sy);
  v_addScaled(o, b.ax[2], b.h[2] * sz);
}

// collide.js:141-158 — in-plane axes and half extents of face `ai`
void faceAxes(const OBB &b, int ai, double *u, double *v, double &eu, double &ev) {
  if (ai == 0) { v_copy(u, b.ax[1]); v_copy(v, b.ax[2]); eu = b.h[1]; ev = b.h[2]; }
  else if (ai == 1) { v_copy(u, b.ax[0]); v_copy(v, b.ax[2]); eu = b.h[0]; ev = b.h[2]; }
  else { v_copy(u, b.ax[0]); v_copy(v, b.ax[1]); eu = b.h[0]; ev = b.h[1]; }
}

// collide.js:202-247 — Sutherland-Hodgman against n.x <= off
int clipPlane(const double *in, int n, const double *pn, double off, double *out) {
  if (n <= 0) return 0;
  int cnt = 0;
  double ax = in[(n - 1) * 3], ay = in[(n - 1) * 3 + 1], az = in[(n - 1) * 3 + 2];
  double da = pn[0] * ax + pn[1] * ay + pn[2] * az - off;
  for (int i = 0; i < n; i++) {
    const double bx = in[i * 3], by = in[i * 3 + 1], bz = in[i * 3 + 2];
    const double db = pn[0] * bx + pn[1] * by + pn[2] * bz - off;
    const bool aIn = da <= PLANE_EPS, bIn = db <= PLANE_EPS;
    if (aIn != bIn) {
      double t = 0;
      const double den = da - db;
      if (std::fabs(den) > SAT_EPS) t = clampd(da / den, 0, 1);
      if (cnt < MAX_POLY) {
        out[cnt * 3] = ax + (bx - ax) * t;
        out[cnt * 3 + 1] = ay + (by - ay) * t;
        out[cnt * 3 + 2] = az + (bz - az) * t;
        cnt++;
      }
    }
    if (bIn && cnt < MAX_POLY) {
      out[cnt * 3] = bx; out[cnt * 3 + 1] = by; out[cnt * 3 + 2] = bz;
      cnt++;
    }
    ax = bx; ay = by; az = bz; da = db;
  }
  return cnt;
}

struct ContactSink {
  const Body &A, &B;
  Contact *out;
  int count = 0;
  double mid[MAX_CONTACTS * 3];
  ContactSink(const Body &a, const Body &b, Contact *o) : A(a), B(b), out(o) {}
  // collide.js:254-285 — dedup by midpoint, anchors in body frames
  bool add(const double *xA, const double *xB, uint32_t key) {
    const double mx = (xA[0] + xB[0]) * 0.5, my = (xA[1] + xB[1]) * 0.5, mz = (xA[2] + xB[2]) * 0.5;
    for (int i = 0; i < count; i++) {
      const double dx = mx - mid[i * 3], dy = my - mid[i * 3 + 1], dz = mz - mid[i * 3 + 2];
      if (dx * dx + dy * dy + dz * dz < MERGE_DIST_SQ) return false;
    }
    if (count >= MAX_CONTACTS) return false;
    Contact &c = out[count];
    c.feature = key;
    double inv[4], t[3];
    q_conj(inv, A.q); v_sub(t, xA, A.p); q_rotate(c.rA, inv, t);
    q_conj(inv, B.q); v_sub(t, xB, B.p); q_rotate(c.rB, inv, t);
    mid[count * 3] = mx; mid[count * 3 + 1] = my; mid[count * 3 + 2] = mz;
    count++;
    return true;
  }
};

// collide.js:291-324
bool testAxis(const OBB &A, const OBB &B, const double *delta, const double *axis, int type,
              int ia, int ib, SatAxis &t, double skin) {
  const double lenSq = v_lenSq(axis);
  if (lenSq < SAT_EPS) return true;
  const double invLen = 1.0 / std::sqrt(lenSq);
  double n[3];
  v_scale(n, axis, invLen);
  if (v_dot(n, delta) < 0) v_neg(n, n);
  const double dist = std::fabs(v_dot(delta, n));
  const double rA = A.h[0] * std::fabs(v_dot(n, A.ax[0])) + A.h[1] * std::fabs(v_dot(n, A.ax[1])) +
                    A.h[2] * std::fabs(v_dot(n, A.ax[2]));
  const double rB = B.h[0] * std::fabs(v_dot(n, B.ax[0])) + B.h[1] * std::fabs(v_dot(n, B.ax[1])) +
                    B.h[2] * std::fabs(v_dot(n, B.ax[2]));
  const double sep = dist - (rA + rB);
  // skin: the GPU contact model's speculative contact distance (0 otherwise)
  if (sep > skin) return false;
  if (!t.valid || sep > t.sep) {
    t.valid = true; t.type = type; t.ia = ia; t.ib = ib; t.sep = sep; v_copy(t.n, n);
  }
  return true;
}

// collide.js:327-343
void supportEdge(const OBB &b, int ai, const double *dir, double *e0, double *e1) {
  const int a1 = (ai + 1) % 3, a2 = (ai + 2) % 3;
  const double s1 = v_dot(dir, b.ax[a1]) >= 0 ? 1 : -1;
  const double s2 = v_dot(dir, b.ax[a2]) >= 0 ? 1 : -1;
  double t[3];
  v_copy(t, b.c);
  v_addScaled(t, b.ax[a1], b.h[a1] * s1);
  v_addScaled(t, b.ax[a2], b.h[a2] * s2);
  v_copy(e0, t); v_addScaled(e0, b.ax[ai], -b.h[ai]);
  v_copy(e1, t); v_addScaled(e1, b.ax[ai], b.h[ai]);
}

// collide.js:345-393 (Ericson's closest points between segments)
void closestSeg(const double *p0, const double *p1, const double *q0, const double *q1, double *c0, double *c1) {
  const double d1x = p1[0] - p0[0], d1y = p1[1] - p0[1], d1z = p1[2] - p0[2];
  const double d2x = q1[0] - q0[0], d2y = q1[1] - q0[1], d2z = q1[2] - q0[2];
  const double rx = p0[0] - q0[0], ry = p0[1] - q0[1], rz = p0[2] - q0[2];
  const double a = d1x * d1x + d1y * d1y + d1z * d1z;
  const double e = d2x * d2x + d2y * d2y + d2z * d2z;
  const double f = d2x * rx + d2y * ry + d2z * rz;
  double s = 0, t = 0;
  if (a <= SAT_EPS && e <= SAT_EPS) { v_copy(c0, p0); v_copy(c1, q0); return; }
  if (a <= SAT_EPS) {
    t = clampd(f / e, 0, 1);
  } else {
    const double c = d1x * rx + d1y * ry + d1z * rz;
    if (e <= SAT_EPS) {
      s = clampd(-c / a, 0, 1);
    } else {
      const double bb = d1x * d2x + d1y * d2y + d1z * d2z;
      const double den = a * e - bb * bb;
      if (std::fabs(den) > SAT_EPS) s = clampd((bb * f - c * e) / den, 0, 1);
      t = (bb * s + f) / e;
      if (t < 0) { t = 0; s = clampd(-c / a, 0, 1); }
      else if (t > 1) { t = 1; s = clampd((bb - c) / a, 0, 1); }
    }
  }
  c0[0] = p0[0] + d1x * s; c0[1] = p0[1] + d1y * s; c0[2] = p0[2] + d1z * s;
  c1[0] = q0[0] + d2x * t; c1[1] = q0[1] + d2y * t; c1[2] = q0[2] + d2z * t;
}

// collide.js:403-472
int faceManifold(const OBB &oA, const OBB &oB, ContactSink &sink, bool refIsA, int refAxis, const double *nAB,
                 double skin) {
  const OBB &ref = refIsA ? oA : oB;
  const OBB &inc = refIsA ? oB : oA;
  double on[3];
  if (refIsA) v_copy(on, nAB); else v_neg(on, nAB);
  // reference face frame (collide.js:161-169)
  double fn[3], fc[3], fu[3], fv[3], eu, ev;
  const double s = v_dot(on, ref.ax[refAxis]) >= 0 ? 1 : -1;
  v_scale(fn, ref.ax[refAxis], s);
  v_copy(fc, ref.c);
  v_addScaled(fc, fn, ref.h[refAxis]);
  faceAxes(ref, refAxis, fu, fv, eu, ev);
  // incident face (collide.js:172-200)
  int incAxis = 0;
  double best = -std::numeric_limits<double>::infinity();
  for (int i = 0; i < 3; i++) {
    const double d = std::fabs(v_dot(inc.ax[i], fn));
    if (d > best) { best = d; incAxis = i; }
  }
  double c0[MAX_POLY * 3], c1[MAX_POLY * 3];
  {
    const double si = v_dot(inc.ax[incAxis], fn) > 0 ? -1 : 1;
    double nrm[3], ctr[3], u[3], v[3], iu, iv;
    v_scale(nrm, inc.ax[incAxis], si);
    v_copy(ctr, inc.c);
    v_addScaled(ctr, nrm, inc.h[incAxis]);
    faceAxes(inc, incAxis, u, v, iu, iv);
    for (int i = 0; i < 4; i++) {
      const double su = (i == 0 || i == 3) ? 1 : -1;
      const double sv = (i == 0 || i == 1) ? 1 : -1;
      for (int k = 0; k < 3; k++) c0[i * 3 + k] = ctr[k] + u[k] * (iu * su) + v[k] * (iv * sv);
    }
  }
  int count = 4;
  double pn[3];
  v_copy(pn, fu);
  count = clipPlane(c0, count, pn, v_dot(pn, fc) + eu, c1);
  if (!count) return 0;
  v_neg(pn, fu);
  count = clipPlane(c1, count, pn, v_dot(pn, fc) + eu, c0);
  if (!count) return 0;
  v_copy(pn, fv);
  count = clipPlane(c0, count, pn, v_dot(pn, fc) + ev, c1);
  if (!count) return 0;
  v_neg(pn, fv);
  count = clipPlane(c1, count, pn, v_dot(pn, fc) + ev, c0);
  if (!count) return 0;

  sink.count = 0;
  uint32_t prefix = uint32_t(refIsA ? AXIS_FACE_A : AXIS_FACE_B) << 24;
  prefix |= uint32_t(refAxis & 0xff) << 16;
  prefix |= uint32_t(incAxis & 0xff) << 8;
  for (int i = 0; i < count && sink.count < MAX_CONTACTS; i++) {
    double p[3], t[3];
    v_copy(p, c0 + i * 3);
    v_sub(t, p, fc);
    const double dist = v_dot(t, fn);
    if (dist > skin + PLANE_EPS) continue;
    v_copy(t, p);
    v_addScaled(t, fn, -dist);
    if (refIsA) sink.add(t, p, prefix | uint32_t(i & 0xff));
    else sink.add(p, t, prefix | uint32_t(i & 0xff));
  }
  if (!sink.count) {
    double xA[3], xB[3], nn[3];
    supportPoint(xA, oA, nAB);
    v_neg(nn, nAB);
    supportPoint(xB, oB, nn);
    sink.add(xA, xB, prefix);
  }
  return sink.count;
}

// collide.js:475-493
int edgeContact(const OBB &oA, const OBB &oB, ContactSink &sink, int axA, int axB, const double *nAB) {
  double nn[3], a0[3], a1[3], b0[3], b1[3], xA[3], xB[3];
  v_neg(nn, nAB);
  supportEdge(oA, axA, nAB, a0, a1);
  supportEdge(oB, axB, nn, b0, b1);
  closestSeg(a0, a1, b0, b1, xA, xB);
  sink.count = 0;
  const uint32_t key = (uint32_t(AXIS_EDGE) << 24) | (uint32_t(axA & 0xff) << 8) | uint32_t(axB & 0xff);
  sink.add(xA, xB, key);
  if (!sink.count) {
    supportPoint(xA, oA, nAB);
    supportPoint(xB, oB, nn);
    sink.add(xA, xB, key);
  }
  return sink.count;
}

// collide.js:496-520
int sphereSphere(const Body &A, const Body &B, ContactSink &sink, double *basis, double skin) {
  double d[3], n[3], t[3], xA[3], xB[3];
  v_sub(d, B.p, A.p);
  const double dsq = v_lenSq(d);
  const double rs = A.radius + B.radius + skin;
  if (dsq > rs * rs) return 0;
  if (dsq > SPHERE_EPS_SQ) v_scale(n, d, 1 / std::sqrt(dsq));
  else v_set(n, 1, 0, 0);
  v_neg(t, n);
  m_orthonormal(basis, t);
  v_copy(xA, A.p); v_addScaled(xA, n, A.radius);
  v_copy(xB, B.p); v_addScaled(xB, n, -B.radius);
  sink.count = 0;
  sink.add(xA, xB, 0x31000001u);
  return sink.count;
}

// collide.js:526-600
int sphereBox(const Body &A, const Body &B, ContactSink &sink, double *basis, bool sphereIsA, double skin) {
  const Body &sph = sphereIsA ? A : B;
  const Body &bx = sphereIsA ? B : A;
  OBB o;
  makeOBB(o, bx);
  double d[3], cp[3], sv[3], n[3], xA[3], xB[3], t[3];
  v_sub(d, sph.p, o.c);
  const double lx = v_dot(d, o.ax[0]), ly = v_dot(d, o.ax[1]), lz = v_dot(d, o.ax[2]);
  const double cx = clampd(lx, -o.h[0], o.h[0]);
  const double cy = clampd(ly, -o.h[1], o.h[1]);
  const double cz = clampd(lz, -o.h[2], o.h[2]);
  v_copy(cp, o.c);
  v_addScaled(cp, o.ax[0], cx);
  v_addScaled(cp, o.ax[1], cy);
  v_addScaled(cp, o.ax[2], cz);
  v_sub(sv, sph.p, cp);
  const double dsq = v_lenSq(sv);
  int fa = 0;
  if (dsq > SPHERE_EPS_SQ) {
    const double dist = std::sqrt(dsq);
    if (dist > sph.radius + skin) return 0;
    v_scale(n, sv, 1 / dist);
  } else {
    const double gx = o.h[0] - std::fabs(lx), gy = o.h[1] - std::fabs(ly), gz = o.h[2] - std::fabs(lz);
    fa = gx < gy ? (gx < gz ? 0 : 2) : (gy < gz ? 1 : 2);
    const double local = fa == 0 ? lx : fa == 1 ? ly : lz;
    const double s = local >= 0 ? 1 : -1;
    v_scale(n, o.ax[fa], s);
    const double gap = fa == 0 ? gx : fa == 1 ? gy : gz;
    v_copy(cp, sph.p);
    v_addScaled(cp, n, gap);
  }
  v_copy(sv, sph.p);
  v_addScaled(sv, n, -sph.radius);
  if (sphereIsA) {
    v_copy(xA, sv); v_copy(xB, cp);
    m_orthonormal(basis, n);
  } else {
    v_copy(xA, cp); v_copy(xB, sv);
    v_neg(t, n);
    m_orthonormal(basis, t);
  }
  sink.count = 0;
  int sX = lx < -o.h[0] ? 1 : lx > o.h[0] ? 2 : 0;
  int sY = ly < -o.h[1] ? 1 : ly > o.h[1] ? 2 : 0;
  int sZ = lz < -o.h[2] ? 1 : lz > o.h[2] ? 2 : 0;
  if ((sX | sY | sZ) == 0) {
    const double local = fa == 0 ? lx : fa == 1 ? ly : lz;
    const int side = local >= 0 ? 2 : 1;
    if (fa == 0) sX = side; else if (fa == 1) sY = side; else sZ = side;
  }
  const uint32_t region = uint32_t(sX | (sY << 2) | (sZ << 4));
  const uint32_t feature = 0x32000000u | (region << 2) | (sphereIsA ? 1u : 2u);
  sink.add(xA, xB, feature);
  return sink.count;
}

}  // namespace

// collide.js:612-685
int collideInto(const Body &A, const Body &B, ContactSink &sink, double basis[9], double skin) {
  const bool sA = A.shape == SPHERE, sB = B.shape == SPHERE;
  if (sA && sB) return sphereSphere(A, B, sink, basis, skin);
  if (sA || sB) return sphereBox(A, B, sink, basis, sA, skin);

  OBB oA, oB;
  makeOBB(oA, A);
  makeOBB(oB, B);
  double delta[3];
  v_sub(delta, oB.c, oA.c);
  SatAxis face{0, -1, -1, -std::numeric_limits<double>::infinity(), {0, 0, 0}, false};
  SatAxis edge = face;
  for (int i = 0; i < 3; i++)
    if (!testAxis(oA, oB, delta, oA.ax[i], AXIS_FACE_A, i, -1, face, skin)) return 0;
  for (int i = 0; i < 3; i++)
    if (!testAxis(oA, oB, delta, oB.ax[i], AXIS_FACE_B, -1, i, face, skin)) return 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double ax[3];
      v_cross(ax, oA.ax[i], oB.ax[j]);
      if (!testAxis(oA, oB, delta, ax, AXIS_EDGE, i, j, edge, skin)) return 0;
    }
  if (!face.valid) return 0;
  SatAxis best = face;
  if (edge.valid) {
    const double relTol = 0.95, absTol = 0.01;
    if (relTol * edge.sep > face.sep + absTol) best = edge;
  }
  double t[3];
  v_neg(t, best.n);
  m_orthonormal(basis, t);
  if (best.type == AXIS_EDGE) return edgeContact(oA, oB, sink, best.ia, best.ib, best.n);
  if (best.type == AXIS_FACE_A) return faceManifold(oA, oB, sink, true, best.ia, best.n, skin);
  return faceManifold(oA, oB, sink, false, best.ib, best.n, skin);
}

int collide(const Body &A, const Body &B, Contact *contacts, double basis[9], double skin, double *mids) {
  ContactSink sink(A, B, contacts);
  const int n = collideInto(A, B, sink, basis, skin);
  if (mids)
    for (int i = 0; i < 3 * n; i++) mids[i] = sink.mid[i];
  return n;
}

// maths.js:660-725 — unrolled LDL^T on the lower triangle
void solve6(const double aLin[9], const double aAng[9], const double aCross[9],
            const double bLin[3], const double bAng[3], double xLin[3], double xAng[3]) {
  const double A11 = aLin[0];
  const double A21 = aLin[3], A22 = aLin[4];
  const double A31 = aLin[6], A32 = aLin[7], A33 = aLin[8];
  const double A41 = aCross[0], A42 = aCross[1], A43 = aCross[2], A44 = aAng[0];
  const double A51 = aCross[3], A52 = aCross[4], A53 = aCross[5], A54 = aAng[3], A55 = aAng[4];
  const double A61 = aCross[6], A62 = aCross[7], A63 = aCross[8], A64 = aAng[6], A65 = aAng[7], A66 = aAng[8];
  const double L21 = A21 / A11, L31 = A31 / A11, L41 = A41 / A11, L51 = A51 / A11, L61 = A61 / A11;
  const double D1 = A11;
  const double D2 = A22 - L21 * L21 * D1;
  const double L32 = (A32 - L21 * L31 * D1) / D2;
  const double L42 = (A42 - L21 * L41 * D1) / D2;
  const double L52 = (A52 - L21 * L51 * D1) / D2;
  const double L62 = (A62 - L21 * L61 * D1) / D2;
  const double D3 = A33 - (L31 * L31 * D1 + L32 * L32 * D2);
  const double L43 = (A43 - L31 * L41 * D1 - L32 * L42 * D2) / D3;
  const double L53 = (A53 - L31 * L51 * D1 - L32 * L52 * D2) / D3;
  const double L63 = (A63 - L31 * L61 * D1 - L32 * L62 * D2) / D3;
  const double D4 = A44 - (L41 * L41 * D1 + L42 * L42 * D2 + L43 * L43 * D3);
  const double L54 = (A54 - L41 * L51 * D1 - L42 * L52 * D2 - L43 * L53 * D3) / D4;
  const double L64 = (A64 - L41 * L61 * D1 - L42 * L62 * D2 - L43 * L63 * D3) / D4;
  const double D5 = A55 - (L51 * L51 * D1 + L52 * L52 * D2 + L53 * L53 * D3 + L54 * L54 * D4);
  const double L65 = (A65 - L51 * L61 * D1 - L52 * L62 * D2 - L53 * L63 * D3 - L54 * L64 * D4) / D5;
  const double D6 = A66 - (L61 * L61 * D1 + L62 * L62 * D2 + L63 * L63 * D3 + L64 * L64 * D4 + L65 * L65 * D5);
  const double y1 = bLin[0];
  const double y2 = bLin[1] - L21 * y1;
  const double y3 = bLin[2] - L31 * y1 - L32 * y2;
  const double y4 = bAng[0] - L41 * y1 - L42 * y2 - L43 * y3;
  const double y5 = bAng[1] - L51 * y1 - L52 * y2 - L53 * y3 - L54 * y4;
  const double y6 = bAng[2] - L61 * y1 - L62 * y2 - L63 * y3 - L64 * y4 - L65 * y5;
  const double z1 = y1 / D1, z2 = y2 / D2, z3 = y3 / D3, z4 = y4 / D4, z5 = y5 / D5, z6 = y6 / D6;
  xAng[2] = z6;
  xAng[1] = z5 - L65 * xAng[2];
  xAng[0] = z4 - L54 * xAng[1] - L64 * xAng[2];
  xLin[2] = z3 - L43 * xAng[0] - L53 * xAng[1] - L63 * xAng[2];
  xLin[1] = z2 - L32 * xLin[2] - L42 * xAng[0] - L52 * xAng[1] - L62 * xAng[2];
  xLin[0] = z1 - L21 * xLin[1] - L31 * xLin[2] - L41 * xAng[0] - L51 * xAng[1] - L61 * xAng[2];
}

namespace {

// ===========================================================================
// Forces
// ===========================================================================

// manifold.js:73-98 — basis transport of a row-basis vector (sphere contacts)
void transportComponents(double *o, const double *val, const double *ob, const double *nb) {
  const double x = val[0], y = val[1], z = val[2];
  for (int r = 0; r < 3; r++) {
    const int n = r * 3;
    o[r] = x * (nb[n] * ob[0] + nb[n + 1] * ob[1] + nb[n + 2] * ob[2]) +
           y * (nb[n] * ob[3] + nb[n + 1] * ob[4] + nb[n + 2] * ob[5]) +
           z * (nb[n] * ob[6] + nb[n + 1] * ob[7] + nb[n + 2] * ob[8]);
  }
}
// manifold.js:143-162
void transportPenalty(double *o, const double *val, const double *ob, const double *nb) {
  const double x = val[0], y = val[1], z = val[2];
  for (int r = 0; r < 3; r++) {
    const int n = r * 3;
    const double dx = nb[n] * ob[0] + nb[n + 1] * ob[1] + nb[n + 2] * ob[2];
    const double dy = nb[n] * ob[3] + nb[n + 1] * ob[4] + nb[n + 2] * ob[5];
    const double dz = nb[n] * ob[6] + nb[n + 1] * ob[7] + nb[n + 2] * ob[8];
    o[r] = x * dx * dx + y * dy * dy + z * dz * dz;
  }
}

// manifold.js:104-120 — dC/dtheta = r x (row of the linear block)
void angularBlocks(const double *jALin, const double *jBLin, const double *rAw, const double *rBw,
                   double *jAAng, double *jBAng) {
  for (int r = 0; r < 3; r++) {
    double row[3], acc[3];
    v_copy(row, jALin + r * 3);
    v_cross(acc, rAw, row);
    jAAng[r * 3] = acc[0]; jAAng[r * 3 + 1] = acc[1]; jAAng[r * 3 + 2] = acc[2];
    v_copy(row, jBLin + r * 3);
    v_cross(acc, rBw, row);
    jBAng[r * 3] = acc[0]; jBAng[r * 3 + 1] = acc[1]; jBAng[r * 3 + 2] = acc[2];
  }
}

// A body's contact and rest offsets in the GPU contact model: its own, else
// the solver's, with the contact offset at least 0 and the rest offset at
// most the contact offset (bodyContactSignature, backend.js:176-185).
double bodyContactOffset(const Params &P, const Body &b) {
  return jmax(0.0, std::isnan(b.contactOffset) ? P.contactOffset : b.contactOffset);
}
double bodyRestOffset(const Params &P, const Body &b) {
  return jmin(std::isnan(b.restOffset) ? P.restOffset : b.restOffset, bodyContactOffset(P, b));
}

// manifold.js:214-318
bool manifoldInitialize(World &W, Force &f) {
  const Params &P = W.prm;
  const Body &A = W.bodies[f.a], &B = W.bodies[f.b];
  f.jacValid = false;
  const bool hasSphere = A.shape == SPHERE || B.shape == SPHERE;
  const bool hadContacts = f.nc > 0;
  double oldBasis[9] = {0};
  const bool gpu = P.contactModel == 1;
  if ((hasSphere || gpu) && hadContacts) m_copy(oldBasis, f.basis);
  // GPU contact model (shaders.js:1105-1149): material of the pair, the
  // contact skin and the rest distance are sums of the bodies' offsets
  double restDistance = 0.0, restitution = 0.0;
  if (gpu) {
    f.mu = std::sqrt(A.staticFriction * B.staticFriction);
    f.muD = std::sqrt(A.dynamicFriction * B.dynamicFriction);
    f.contactDistance = bodyContactOffset(P, A) + bodyContactOffset(P, B);
    restDistance = bodyRestOffset(P, A) + bodyRestOffset(P, B);
    restitution = jmax(A.restitution, B.restitution);
  } else {
    f.mu = std::sqrt(A.friction * B.friction);
    f.muD = f.mu;
    f.contactDistance = 0.0;
  }

  Contact nw[MAX_CONTACTS];
  bool found[MAX_CONTACTS];
  for (int i = 0; i < MAX_CONTACTS; i++) {
    nw[i].reset();
    nw[i].stick = gpu;  // GPU model: a new contact starts in the static regime (shaders.js:1207-1210)
    found[i] = false;
  }
  double mids[3 * MAX_CONTACTS];
  const int nn = collide(A, B, nw, f.basis, f.contactDistance, mids);
  const bool compatible = !hasSphere || !hadContacts ||
                          (oldBasis[0] * f.basis[0] + oldBasis[1] * f.basis[1] + oldBasis[2] * f.basis[2]) > 0.85;
  if (gpu) {
    // shaders.js:1174-1319: an old contact matches a new one by feature key
    // or, failing that, geometrically: the old anchors, placed at the current
    // poses, lie within the persistence distance of the new contact (normal
    // gap, tangential offset and midpoint); the nearest such one wins, an
    // exact feature match ends the search (for a pair with a sphere at once),
    // and each old contact is claimed once. k and lambda are carried into the
    // new basis for every pair.
    const bool basisCompatible =
        hadContacts && (oldBasis[0] * f.basis[0] + oldBasis[1] * f.basis[1] + oldBasis[2] * f.basis[2]) > 0.85;
    const double minExtent = 2.0 * jmin(jmin(A.size[0] * 0.5, jmin(A.size[1] * 0.5, A.size[2] * 0.5)),
                                        jmin(B.size[0] * 0.5, jmin(B.size[1] * 0.5, B.size[2] * 0.5)));
    const double persistence = jmax(P.contactPersistence, jmax(f.contactDistance, minExtent * 0.02));
    const double persistenceSq = persistence * persistence;
    uint32_t oldUsed = 0;
    for (int i = 0; i < nn; i++) {
      int match = -1;
      bool near = false, exact = false;
      double best = persistenceSq;
      for (int j = 0; basisCompatible && j < f.nc; j++) {
        if (oldUsed & (1u << j)) continue;
        const bool same = f.c[j].feature == nw[i].feature;
        if (hasSphere && same) { match = j; exact = true; break; }
        double oxA[3], oxB[3], od[3], tan[3], md[3];
        transformPoint(oxA, A.p, A.q, f.c[j].rA);
        transformPoint(oxB, B.p, B.q, f.c[j].rB);
        v_sub(od, oxA, oxB);
        const double sep = v_dot(f.basis, od);
        for (int k = 0; k < 3; k++) tan[k] = od[k] - f.basis[k] * sep;
        if (sep > f.contactDistance + persistence) continue;
        if (v_lenSq(tan) > persistenceSq) continue;
        for (int k = 0; k < 3; k++) md[k] = mids[i * 3 + k] - (oxA[k] + oxB[k]) * 0.5;
        const double mdSq = v_lenSq(md);
        if (mdSq > persistenceSq) continue;
        if (same) { match = j; exact = true; break; }
        if (!near || mdSq < best) { near = true; match = j; best = mdSq; }
      }
      if (!near && !exact) continue;
      if (!exact) W.geometricMatches++;
      oldUsed |= 1u << match;
      found[i] = true;
      double rA[3], rB[3];
      v_copy(rA, nw[i].rA);
      v_copy(rB, nw[i].rB);
      const uint32_t feature = nw[i].feature;  // the new contact's key is kept
      nw[i] = f.c[match];
      nw[i].feature = feature;
      transportComponents(nw[i].lam, nw[i].lam, oldBasis, f.basis);
      transportPenalty(nw[i].k, nw[i].k, oldBasis, f.basis);
      if (!f.c[match].stick || hasSphere) {
        v_copy(nw[i].rA, rA);
        v_copy(nw[i].rB, rB);
      }
    }
  }
  for (int i = 0; i < nn && !gpu; i++) {
    for (int j = 0; j < f.nc; j++) {
      if (compatible && nw[i].feature == f.c[j].feature) {
        found[i] = true;
        double rA[3], rB[3];
        v_copy(rA, nw[i].rA);
        v_copy(rB, nw[i].rB);
        nw[i] = f.c[j];
        if (hasSphere) {
          transportComponents(nw[i].lam, nw[i].lam, oldBasis, f.basis);
          transportPenalty(nw[i].k, nw[i].k, oldBasis, f.basis);
        }
        if (!f.c[j].stick || hasSphere) {
          v_copy(nw[i].rA, rA);
          v_copy(nw[i].rB, rB);
        }
        break;
      }
    }
  }
  f.nc = nn;
  for (int i = 0; i < nn; i++) f.c[i] = nw[i];
  for (int i = 0; i < f.nc; i++) {
    Contact &c = f.c[i];
    double xA[3], xB[3], d[3];
    transformPoint(xA, A.p, A.q, c.rA);
    transformPoint(xB, B.p, B.q, c.rB);
    v_sub(d, xA, xB);
    if (gpu) {
      // shaders.js:1322-1350: C0 against the rest distance; restitution is a
      // one-step target for a new impact (or one a speculative-only contact
      // kept armed), when the closing motion crosses the rest surface this
      // step faster than the threshold
      const bool priorOutsideRest = c.C0[0] > 0.0, priorHadBias = c.bias > 0.0;
      m_mulVec(c.C0, f.basis, d);
      c.C0[0] = c.C0[0] - restDistance;
      double bias = 0.0;
      const bool armed = !found[i] || (priorOutsideRest && !priorHadBias);
      if (armed && restitution > 0.0) {
        double ra[3], rb[3], va[3], vb[3], rel[3];
        q_rotate(ra, A.q, c.rA);
        q_rotate(rb, B.q, c.rB);
        v_cross(va, A.w, ra);
        v_add(va, A.v, va);
        v_cross(vb, B.w, rb);
        v_add(vb, B.v, vb);
        v_sub(rel, va, vb);
        const double vn = v_dot(f.basis, rel);
        const double closing = jmax(0.0, -vn * P.dt);
        const double toRest = jmax(0.0, c.C0[0]);
        if (vn < -P.restitutionThreshold && closing > toRest) bias = restitution * (closing - toRest);
      }
      c.bias = bias;
    } else {
      m_mulVec(c.C0, f.basis, d);
      c.C0[0] += P.collisionMargin;
    }
    if (!P.postStabilize) {
      v_scale(c.lam, c.lam, P.alpha);
      v_scale(c.lam, c.lam, P.gamma);
    }
    v_scale(c.k, c.k, P.gamma);
    v_clampScalar(c.k, c.k, P.penaltyMin, P.penaltyMax);
  }
  return f.nc > 0;
}

struct PassState {   // manifold.js:325-350 (_beginPass)
  double dqALin[3], dqAAng[3], dqBLin[3], dqBAng[3];
  double jALin[9], jBLin[9];
};

void beginPass(World &W, Force &f, PassState &s) {
  const Body &A = W.bodies[f.a], &B = W.bodies[f.b];
  v_sub(s.dqALin, A.p, A.p0);
  q_subtract(s.dqAAng, A.q, A.q0);
  v_sub(s.dqBLin, B.p, B.p0);
  q_subtract(s.dqBAng, B.q, B.q0);
  m_copy(s.jALin, f.basis);
  m_neg(s.jBLin, f.basis);
  if (W.prm.cachedContactJacobians && !f.jacValid) {
    for (int i = 0; i < f.nc; i++) {
      double rAw[3], rBw[3];
      q_rotate(rAw, A.q0, f.c[i].rA);
      q_rotate(rBw, B.q0, f.c[i].rB);
      angularBlocks(s.jALin, s.jBLin, rAw, rBw, f.jac + i * 18, f.jac + i * 18 + 9);
    }
    f.jacValid = true;
  }
}

struct EvalOut { double jAAng[9], jBAng[9], C[3], F[3]; bool within; };

// manifold.js:358-409
void evaluate(World &W, Force &f, const PassState &s, int i, double alpha, EvalOut &e) {
  const Contact &c = f.c[i];
  if (W.prm.cachedContactJacobians) {
    for (int k = 0; k < 9; k++) { e.jAAng[k] = f.jac[i * 18 + k]; e.jBAng[k] = f.jac[i * 18 + 9 + k]; }
  } else {
    const Body &A = W.bodies[f.a], &B = W.bodies[f.b];
    double rAw[3], rBw[3];
    q_rotate(rAw, A.q, c.rA);
    q_rotate(rBw, B.q, c.rB);
    angularBlocks(s.jALin, s.jBLin, rAw, rBw, e.jAAng, e.jBAng);
  }
  double K[9], acc[3];
  m_diag(K, c.k[0], c.k[1], c.k[2]);
  v_scale(e.C, c.C0, 1 - alpha);
  const bool gpu = W.prm.contactModel == 1;
  if (gpu) {
    // shaders.js:1805-1819: a positive speculative gap is kept whole, and the
    // restitution target is subtracted outside the projection pass
    if (f.contactDistance > 0.0 && c.C0[0] > 0.0) e.C[0] = c.C0[0];
    if (!(W.prm.postStabilize && alpha < 0.5)) e.C[0] = e.C[0] - c.bias;
  }
  m_mulVec(acc, s.jALin, s.dqALin); v_add(e.C, e.C, acc);
  m_mulVec(acc, s.jBLin, s.dqBLin); v_add(e.C, e.C, acc);
  m_mulVec(acc, e.jAAng, s.dqAAng); v_add(e.C, e.C, acc);
  m_mulVec(acc, e.jBAng, s.dqBAng); v_add(e.C, e.C, acc);
  m_mulVec(e.F, K, e.C);
  v_add(e.F, e.F, c.lam);
  e.F[0] = jmin(e.F[0], 0);
  if (gpu) {
    // shaders.js:1824-1844: the static cone while the contact sticks; a
    // force beyond it is clamped to the dynamic cone at once
    const double fs = std::sqrt(e.F[1] * e.F[1] + e.F[2] * e.F[2]);
    const double active = std::fabs(e.F[0]) * (c.stick ? f.mu : f.muD);
    e.within = fs <= active;
    if (!e.within && fs > 0.0) {
      const double dyn = std::fabs(e.F[0]) * f.muD;
      e.F[1] = e.F[1] * dyn / fs;
      e.F[2] = e.F[2] * dyn / fs;
    }
    return;
  }
  const double bounds = std::fabs(e.F[0]) * f.mu;
  const double fs = std::sqrt(e.F[1] * e.F[1] + e.F[2] * e.F[2]);
  if (fs > bounds && fs > 0) {
    e.F[1] *= bounds / fs;
    e.F[2] *= bounds / fs;
  }
  e.within = fs <= bounds;
}

// manifold.js:411-445
void manifoldPrimal(World &W, Force &f, int body, double alpha, double *lhsLin, double *lhsAng,
                    double *lhsCross, double *rhsLin, double *rhsAng) {
  const bool isA = body == f.a;
  PassState s;
  beginPass(W, f, s);
  for (int i = 0; i < f.nc; i++) {
    const Contact &c = f.c[i];
    EvalOut e;
    evaluate(W, f, s, i, alpha, e);
    const double *jLin = isA ? s.jALin : s.jBLin;
    const double *jAng = isA ? e.jAAng : e.jBAng;
    double jLinT[9], jAngT[9], jAngTk[9], m0[9], acc[3];
    m_transpose(jLinT, jLin);
    m_transpose(jAngT, jAng);
    m_mulDiagRight(jAngTk, jAngT, c.k[0], c.k[1], c.k[2]);
    m_mulDiagRight(m0, jLinT, c.k[0], c.k[1], c.k[2]);
    m_mul(m0, m0, jLin);
    m_addInto(lhsLin, m0);
    m_mul(m0, jAngTk, jAng);
    m_addInto(lhsAng, m0);
    m_mul(m0, jAngTk, jLin);
    m_addInto(lhsCross, m0);
    m_mulVec(acc, jLinT, e.F); v_add(rhsLin, rhsLin, acc);
    m_mulVec(acc, jAngT, e.F); v_add(rhsAng, rhsAng, acc);
  }
}

// manifold.js:447-491
void manifoldDual(World &W, Force &f, double alpha) {
  const Params &P = W.prm;
  PassState s;
  beginPass(W, f, s);
  const bool hasSphere = W.bodies[f.a].shape == SPHERE || W.bodies[f.b].shape == SPHERE;
  for (int i = 0; i < f.nc; i++) {
    Contact &c = f.c[i];
    EvalOut e;
    evaluate(W, f, s, i, alpha, e);
    v_copy(c.lam, e.F);
    v_copy(c.C, e.C);
    if (e.F[0] < 0) c.k[0] = jmin(c.k[0] + P.betaLin * std::fabs(e.C[0]), P.penaltyMax);
    if (P.contactModel == 1) {
      // shaders.js:2248-2258: the tangential ramp inside the cone, and the
      // contact sticks exactly when the force is inside it
      if (e.within) {
        c.k[1] = jmin(c.k[1] + P.betaLin * std::fabs(e.C[1]), P.penaltyMax);
        c.k[2] = jmin(c.k[2] + P.betaLin * std::fabs(e.C[2]), P.penaltyMax);
      }
      c.stick = e.within;
    } else if (e.within) {
      c.k[1] = jmin(c.k[1] + P.betaLin * std::fabs(e.C[1]), P.penaltyMax);
      c.k[2] = jmin(c.k[2] + P.betaLin * std::fabs(e.C[2]), P.penaltyMax);
      c.stick = std::sqrt(e.C[1] * e.C[1] + e.C[2] * e.C[2]) < P.stickThreshold;
    } else if (hasSphere) {
      c.stick = false;
    }
  }
}

const double IDENTITY_Q[4] = {0, 0, 0, 1};

// joint.js:115-128
void jointEvalLinear(const World &W, const Force &f, double *o) {
  double pA[3], pB[3];
  if (f.a >= 0) transformPoint(pA, W.bodies[f.a].p, W.bodies[f.a].q, f.jrA);
  else v_copy(pA, f.jrA);
  transformPoint(pB, W.bodies[f.b].p, W.bodies[f.b].q, f.jrB);
  v_sub(o, pA, pB);
}
void jointEvalAngular(const World &W, const Force &f, double *o) {
  q_subtract(o, f.a >= 0 ? W.bodies[f.a].q : IDENTITY_Q, W.bodies[f.b].q);
  v_scale(o, o, f.torqueArm);
}

// joint.js:131-160
bool jointInitialize(World &W, Force &f) {
  const Params &P = W.prm;
  jointEvalLinear(W, f, f.C0Lin);
  jointEvalAngular(W, f, f.C0Ang);
  if (!P.postStabilize) {
    v_scale(f.lamLin, f.lamLin, P.alpha);
    v_scale(f.lamLin, f.lamLin, P.gamma);
    v_scale(f.lamAng, f.lamAng, P.alpha);
    v_scale(f.lamAng, f.lamAng, P.gamma);
  }
  v_scale(f.kLin, f.kLin, P.gamma);
  v_clampScalar(f.kLin, f.kLin, P.penaltyMin, P.penaltyMax);
  v_scale(f.kAng, f.kAng, P.gamma);
  v_clampScalar(f.kAng, f.kAng, P.penaltyMin, P.penaltyMax);
  v_minScalar(f.kLin, f.kLin, f.stiffLin);
  v_minScalar(f.kAng, f.kAng, f.stiffAng);
  return !f.broken;
}

// joint.js:57-63 — second derivative of component k of a rotated offset
void geomStiffness(double *o, int k, const double *v) {
  m_diag(o, -v[k], -v[k], -v[k]);
  o[0 * 3 + k] += v[0];
  o[1 * 3 + k] += v[1];
  o[2 * 3 + k] += v[2];
}

// joint.js:162-253
void jointPrimal(World &W, Force &f, int body, double alpha, double *lhsLin, double *lhsAng,
                 double *lhsCross, double *rhsLin, double *rhsAng) {
  const bool isA = body == f.a;
  const double INF = std::numeric_limits<double>::infinity();
  if (v_lenSq(f.kLin) > 0) {
    double K[9], C[3], F[3], jLin[9], jAng[9], r[3], rW[3];
    double jLinT[9], jAngT[9], jAngTk[9], m0[9], H[9], G[9], acc[3];
    m_diag(K, f.kLin[0], f.kLin[1], f.kLin[2]);
    jointEvalLinear(W, f, C);
    if (f.stiffLin == INF) v_addScaled(C, f.C0Lin, -alpha);
    m_mulVec(F, K, C);
    v_add(F, F, f.lamLin);
    if (isA) {
      m_diag(jLin, 1, 1, 1);
      q_rotate(rW, W.bodies[f.a].q, f.jrA);
      v_neg(r, rW);
      m_skew(jAng, r);
      v_copy(r, rW);
    } else {
      m_diag(jLin, -1, -1, -1);
      q_rotate(rW, W.bodies[f.b].q, f.jrB);
      m_skew(jAng, rW);
      v_neg(r, rW);
    }
    const double k0 = f.kLin[0], k1 = f.kLin[1], k2 = f.kLin[2];
    m_transpose(jLinT, jLin);
    m_transpose(jAngT, jAng);
    m_mulDiagRight(jAngTk, jAngT, k0, k1, k2);
    m_mulDiagRight(m0, jLinT, k0, k1, k2);
    m_mul(m0, m0, jLin);
    m_addInto(lhsLin, m0);
    m_mul(m0, jAngTk, jAng);
    m_addInto(lhsAng, m0);
    m_mul(m0, jAngTk, jLin);
    m_addInto(lhsCross, m0);
    m_zero(H);
    for (int k = 0; k < 3; k++) {
      geomStiffness(G, k, r);
      m_scale(G, G, F[k]);
      m_addInto(H, G);
    }
    m_diagonalize(G, H);
    m_addInto(lhsAng, G);
    m_mulVec(acc, jLinT, F); v_add(rhsLin, rhsLin, acc);
    m_mulVec(acc, jAngT, F); v_add(rhsAng, rhsAng, acc);
  }
  if (v_lenSq(f.kAng) > 0) {
    double K[9], C[3], F[3], jAng[9], jAngT[9], m0[9], acc[3];
    m_diag(K, f.kAng[0], f.kAng[1], f.kAng[2]);
    jointEvalAngular(W, f, C);
    if (f.stiffAng == INF) v_addScaled(C, f.C0Ang, -alpha);
    m_mulVec(F, K, C);
    v_add(F, F, f.lamAng);
    const double s = isA ? f.torqueArm : -f.torqueArm;
    m_diag(jAng, s, s, s);
    m_transpose(jAngT, jAng);
    m_mul(m0, jAngT, K);
    m_mul(m0, m0, jAng);
    m_addInto(lhsAng, m0);
    m_mulVec(acc, jAngT, F); v_add(rhsAng, rhsAng, acc);
  }
}

// joint.js:255-317
void jointDual(World &W, Force &f, double alpha) {
  const Params &P = W.prm;
  const double INF = std::numeric_limits<double>::infinity();
  if (v_lenSq(f.kLin) > 0) {
    double K[9], C[3], F[3], acc[3];
    m_diag(K, f.kLin[0], f.kLin[1], f.kLin[2]);
    jointEvalLinear(W, f, C);
    if (f.stiffLin == INF) {
      v_addScaled(C, f.C0Lin, -alpha);
      m_mulVec(F, K, C);
      v_add(f.lamLin, F, f.lamLin);
    }
    v_copy(f.CLin, C);
    v_abs(acc, C);
    v_scale(acc, acc, P.betaLin);
    v_add(f.kLin, f.kLin, acc);
    v_minScalar(f.kLin, f.kLin, jmin(f.stiffLin, P.penaltyMax));
  }
  if (v_lenSq(f.kAng) > 0) {
    double K[9], C[3], F[3], acc[3];
    m_diag(K, f.kAng[0], f.kAng[1], f.kAng[2]);
    jointEvalAngular(W, f, C);
    if (f.stiffAng == INF) {
      v_addScaled(C, f.C0Ang, -alpha);
      m_mulVec(F, K, C);
      v_add(f.lamAng, F, f.lamAng);
    }
    v_copy(f.CAng, C);
    v_abs(acc, C);
    v_scale(acc, acc, P.betaAng);
    v_add(f.kAng, f.kAng, acc);
    v_minScalar(f.kAng, f.kAng, jmin(f.stiffAng, P.penaltyMax));
  }
  if (v_lenSq(f.lamAng) > f.fracture * f.fracture) {
    for (int k = 0; k < 3; k++) f.kLin[k] = f.kAng[k] = f.lamLin[k] = f.lamAng[k] = 0.0;
    f.broken = true;
  }
}

// ------------------------------------------------------------------ springs
// spring.js:107-118 (_evaluate): d = pA - pB and n = d / |d|; NaN when the
// anchors (nearly) coincide.
double springEvaluate(const World &W, const Force &f, double *d, double *n) {
  const Body &A = W.bodies[f.a], &B = W.bodies[f.b];
  double pA[3], pB[3];
  transformPoint(pA, A.p, A.q, f.jrA);
  transformPoint(pB, B.p, B.q, f.jrB);
  v_sub(d, pA, pB);
  const double dLen = std::sqrt(v_lenSq(d));
  if (dLen <= 1.0e-6) return std::numeric_limits<double>::quiet_NaN();
  v_div(n, d, dLen);
  return dLen - f.rest;
}

// spring.js:121-131 (_updateTearing)
bool springTear(Force &f, double C) {
  if (f.rest <= 1.0e-6) return false;
  const double strain = jmax(0.0, C / f.rest);
  f.peakStrain = jmax(f.peakStrain, strain);
  if (strain <= f.tearStrain) return false;
  f.broken = true;
  f.penalty = 0;
  f.springC = 0;
  return true;
}

// spring.js:133-166 with the stiffness ramp off (solver.js:187, the default):
// the penalty is the material stiffness. A torn spring stays in the force
// list as an inactive sentinel.
bool springInitialize(World &W, Force &f) {
  if (f.broken) {
    f.penalty = 0;
    f.springC = 0;
    return true;
  }
  if (std::isfinite(f.tearStrain)) {
    double d[3], n[3];
    const double C = springEvaluate(W, f, d, n);
    if (!std::isnan(C)) {
      f.springC = C;
      if (springTear(f, C)) return true;
    }
  }
  f.penalty = f.stiffness;
  return f.stiffness > 0;
}

inline void m_outer(double *o, const double *a, const double *b) {
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) o[r * 3 + c] = a[r] * b[c];
}

// spring.js:168-212 (updatePrimal) and 214-269 (the geometric stiffness of
// Equation 17, on by default: solver.js:190-193)
void springPrimal(World &W, Force &f, int body, double *lhsLin, double *lhsAng, double *lhsCross, double *rhsLin,
                  double *rhsAng) {
  if (f.broken) return;
  double d[3], n[3];
  const double C = springEvaluate(W, f, d, n);
  if (std::isnan(C)) return;
  const bool isA = body == f.a;
  const double k = f.penalty;
  const double fo = k * C;
  double rW[3], jl[3], ja[3], m0[9];
  if (isA) {
    q_rotate(rW, W.bodies[f.a].q, f.jrA);
    v_copy(jl, n);
    v_cross(ja, rW, n);
  } else {
    q_rotate(rW, W.bodies[f.b].q, f.jrB);
    v_neg(jl, n);
    v_cross(ja, rW, n);
    v_neg(ja, ja);
  }
  m_outer(m0, jl, jl);
  m_scale(m0, m0, k);
  m_addInto(lhsLin, m0);
  m_outer(m0, ja, ja);
  m_scale(m0, m0, k);
  m_addInto(lhsAng, m0);
  m_outer(m0, ja, jl);
  m_scale(m0, m0, k);
  m_addInto(lhsCross, m0);
  v_addScaled(rhsLin, jl, fo);
  v_addScaled(rhsAng, ja, fo);

  const double dLen = std::sqrt(v_lenSq(d));
  const double s = isA ? 1.0 : -1.0;
  double P[9], S[9], Gll[9], Gla[9], Gal[9], Gaa[9];
  m_outer(P, n, n);
  m_neg(P, P);
  P[0] += 1;
  P[4] += 1;
  P[8] += 1;
  m_div(P, P, dLen);
  m_skew(S, rW);
  m_copy(Gll, P);
  m_mul(Gla, P, S);
  m_neg(Gla, Gla);
  m_mul(Gal, S, P);
  m_mul(Gaa, S, P);
  m_mul(Gaa, Gaa, S);
  m_neg(Gaa, Gaa);
  const double ndotr = v_dot(n, rW);
  m_outer(m0, rW, n);
  m0[0] -= ndotr;
  m0[4] -= ndotr;
  m0[8] -= ndotr;
  m_scale(m0, m0, s);
  m_addInto(Gaa, m0);
  const double af = std::fabs(fo);
  for (int c = 0; c < 3; c++) {
    const double a0 = Gll[c], a1 = Gll[3 + c], a2 = Gll[6 + c];
    const double a3 = Gal[c], a4 = Gal[3 + c], a5 = Gal[6 + c];
    lhsLin[c * 3 + c] += std::sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3 + a4 * a4 + a5 * a5) * af;
  }
  for (int c = 0; c < 3; c++) {
    const double a0 = Gla[c], a1 = Gla[3 + c], a2 = Gla[6 + c];
    const double a3 = Gaa[c], a4 = Gaa[3 + c], a5 = Gaa[6 + c];
    lhsAng[c * 3 + c] += std::sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3 + a4 * a4 + a5 * a5) * af;
  }
  f.springC = C;
}

// spring.js:271-288 (updateDual): the constraint value and tearing only,
// with the ramp off
void springDual(World &W, Force &f) {
  if (f.broken) return;
  double d[3], n[3];
  const double C = springEvaluate(W, f, d, n);
  if (std::isnan(C)) return;
  f.springC = C;
  if (std::isfinite(f.tearStrain)) springTear(f, C);
}

// rigid.js:250-259
bool constrainedTo(const World &W, int a, int b) {
  const std::vector<int> &fl = W.bodies[a].forces;
  for (int i = int(fl.size()) - 1; i >= 0; i--) {
    const Force &f = W.pool[fl[i]];
    if (f.kind != MANIFOLD && f.broken) continue;
    if ((f.a == a && f.b == b) || (f.a == b && f.b == a)) return true;
  }
  return false;
}

bool jointConnected(const World &W, int a, int b) {
  const std::vector<int> &fl = W.bodies[a].forces;
  for (int i = int(fl.size()) - 1; i >= 0; i--) {
    const Force &f = W.pool[fl[i]];
    if (f.kind == MANIFOLD || f.broken) continue;
    if ((f.a == a && f.b == b) || (f.a == b && f.b == a)) return true;
  }
  return false;
}

void removeFromBody(std::vector<int> &fl, int id) {
  auto it = std::find(fl.begin(), fl.end(), id);
  if (it != fl.end()) fl.erase(it);
}

// Broad-phase candidates: every (i, j), j < i, whose bounding spheres may
// overlap, enumerated i = n-1..1 and j descending (the order of
// broadphase.js:236-243). A uniform grid over the bodies' centres keeps the
// cost near-linear; bodies far larger than the rest are tested against all
// others. The set visited is exactly the brute-force set.
double bodyContactOffset(const Params &P, const Body &b);

template <class Visit>
void forEachPair(const World &W, Visit visit, bool brute) {
  const int n = int(W.bodies.size());
  if (n < 2) return;
  if (brute) {
    for (int i = n - 1; i >= 1; i--)
      for (int j = i - 1; j >= 0; j--) visit(i, j);
    return;
  }
  double rsum = 0;
  for (int i = 0; i < n; i++) rsum += W.bodies[i].radius;
  const double large = std::max(rsum / n * 8.0, 1e-6);
  double rmax = 0;
  std::vector<char> isLarge(n, 0);
  std::vector<int> larges;
  for (int i = 0; i < n; i++) {
    if (W.bodies[i].radius > large) { isLarge[i] = 1; larges.push_back(i); }
    else rmax = std::max(rmax, W.bodies[i].radius);
  }
  // the GPU contact model grows every pair's reach by the two contact offsets
  double reach = rmax;
  if (W.prm.contactModel == 1) {
    double off = 0.0;
    for (int i = 0; i < n; i++) off = std::max(off, bodyContactOffset(W.prm, W.bodies[i]));
    reach = rmax + off;
  }
  const double h = std::max(2.0 * reach * (1.0 + 1e-9), 1e-6);
  struct Key { int64_t cell; int idx; };
  std::vector<Key> keys;
  keys.reserve(n);
  auto cellOf = [&](int i, int64_t &cx, int64_t &cy, int64_t &cz) {
    cx = int64_t(std::floor(W.bodies[i].p[0] / h));
    cy = int64_t(std::floor(W.bodies[i].p[1] / h));
    cz = int64_t(std::floor(W.bodies[i].p[2] / h));
  };
  auto pack = [](int64_t x, int64_t y, int64_t z) {
    return ((x + (1 << 20)) << 42) | ((y + (1 << 20)) << 21) | (z + (1 << 20));
  };
  for (int i = 0; i < n; i++) {
    if (isLarge[i]) continue;
    int64_t x, y, z;
    cellOf(i, x, y, z);
    keys.push_back({pack(x, y, z), i});
  }
  std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
    return a.cell < b.cell || (a.cell == b.cell && a.idx < b.idx);
  });
  std::vector<int> cand;
  for (int i = n - 1; i >= 1; i--) {
    if (isLarge[i]) {
      for (int j = i - 1; j >= 0; j--) visit(i, j);
      continue;
    }
    cand.clear();
    for (int j : larges) if (j < i) cand.push_back(j);
    int64_t x, y, z;
    cellOf(i, x, y, z);
    for (int64_t dx = -1; dx <= 1; dx++)
      for (int64_t dy = -1; dy <= 1; dy++)
        for (int64_t dz = -1; dz <= 1; dz++) {
          const int64_t key = pack(x + dx, y + dy, z + dz);
          auto lo = std::lower_bound(keys.begin(), keys.end(), key,
                                     [](const Key &k, int64_t v) { return k.cell < v; });
          for (auto it = lo; it != keys.end() && it->cell == key && it->idx < i; ++it) cand.push_back(it->idx);
        }
    std::sort(cand.begin(), cand.end(), std::greater<int>());
    for (int j : cand) visit(i, j);
  }
}

int newForce(World &W, int kind, int a, int b) {
  Force f;
  f.kind = kind;
  f.a = a;
  f.b = b;
  f.seq = W.nextSeq++;
  // a dropped contact's slot is taken again (the reference's manifolds are
  // garbage-collected objects; ids are only handles, never ordered)
  int id;
  if (!W.freeForces.empty()) {
    id = W.freeForces.back();
    W.freeForces.pop_back();
    W.pool[size_t(id)] = f;
  } else {
    W.pool.push_back(f);
    id = int(W.pool.size()) - 1;
  }
  W.forces.push_back(id);
  if (a >= 0) W.bodies[a].forces.push_back(id);
  if (b >= 0) W.bodies[b].forces.push_back(id);
  return id;
}

}  // namespace

// rigid.js:29-98
int World::addBox(const double size[3], double density, double friction, const double pos[3], const double vel[3]) {
  Body b;
  b.shape = BOX;
  for (int k = 0; k < 3; k++) {
    b.size[k] = size[k];
    b.p[k] = b.p0[k] = pos[k];
    b.v[k] = b.vprev[k] = vel[k];
    b.w[k] = 0; b.py[k] = 0;
  }
  b.q[0] = b.q[1] = b.q[2] = 0; b.q[3] = 1;
  q_copy(b.q0, b.q);
  b.qy[0] = b.qy[1] = b.qy[2] = 0; b.qy[3] = 1;
  b.density = density;
  b.friction = friction;
  b.staticFriction = b.dynamicFriction = friction;  // rigid.js:56-58
  b.mass = size[0] * size[1] * size[2] * density;
  b.moment[0] = ((size[1] * size[1] + size[2] * size[2]) / 12.0) * b.mass;
  b.moment[1] = ((size[0] * size[0] + size[2] * size[2]) / 12.0) * b.mass;
  b.moment[2] = ((size[0] * size[0] + size[1] * size[1]) / 12.0) * b.mass;
  const double hx = size[0] * 0.5, hy = size[1] * 0.5, hz = size[2] * 0.5;
  b.radius = std::sqrt(hx * hx + hy * hy + hz * hz);
  bodies.push_back(b);
  return int(bodies.size()) - 1;
}

// rigid.js:105-133
int World::addSphere(double diameter, double density, double friction, const double pos[3], const double vel[3]) {
  const double size[3] = {diameter, diameter, diameter};
  const int id = addBox(size, density, friction, pos, vel);
  Body &b = bodies[id];
  const double r = diameter * 0.5;
  b.shape = SPHERE;
  b.radius = r;
  b.mass = (4.0 / 3.0) * M_PI * r * r * r * density;
  const double inertia = (2.0 / 5.0) * b.mass * r * r;
  b.moment[0] = b.moment[1] = b.moment[2] = inertia;
  return id;
}

// joint.js:78-113
int World::addJoint(int a, int b, const double rA[3], const double rB[3], double stiffLin, double stiffAng,
                    double fracture) {
  const int id = newForce(*this, JOINT, a, b);
  Force &f = pool[id];
  v_copy(f.jrA, rA);
  v_copy(f.jrB, rB);
  f.stiffLin = stiffLin;
  f.stiffAng = stiffAng;
  f.fracture = fracture;
  const double *sA = a >= 0 ? bodies[a].size : nullptr;
  const double sx = (sA ? sA[0] : 0) + bodies[b].size[0];
  const double sy = (sA ? sA[1] : 0) + bodies[b].size[1];
  const double sz = (sA ? sA[2] : 0) + bodies[b].size[2];
  f.torqueArm = sx * sx + sy * sy + sz * sz;
  return id;
}

int World::addSpring(int a, int b, const double rA[3], const double rB[3], double stiffness, double rest,
                     double tearStrain) {
  const int id = newForce(*this, SPRING, a, b);
  Force &f = pool[id];
  v_copy(f.jrA, rA);
  v_copy(f.jrB, rB);
  f.stiffness = stiffness;
  f.tearStrain = tearStrain;
  if (rest < 0) {
    // spring.js:78-84: the anchors' current distance
    double pA[3], pB[3], d[3];
    transformPoint(pA, bodies[a].p, bodies[a].q, f.jrA);
    transformPoint(pB, bodies[b].p, bodies[b].q, f.jrB);
    v_sub(d, pA, pB);
    f.rest = std::sqrt(v_lenSq(d));
  } else {
    f.rest = rest;
  }
  return id;
}

// layout.js:50-57 — lowbias32 integer hash
uint32_t colorPriority(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// coloring.js:43-118 — greedy, statics uncoloured
void World::buildColoring(Coloring &out) const {
  const int n = int(bodies.size());
  out.colorOf.assign(n, -1);
  std::vector<int> degree(n, 0);
  for (int id : forces) {
    const Force &f = pool[id];
    if (f.a >= 0 && f.b >= 0) { degree[f.a]++; degree[f.b]++; }
  }
  std::vector<int> start(n + 1, 0);
  for (int i = 0; i < n; i++) start[i + 1] = start[i] + degree[i];
  std::vector<int> adj(start[n]);
  std::vector<int> cur(start.begin(), start.end() - 1);
  for (int id : forces) {
    const Force &f = pool[id];
    if (f.a >= 0 && f.b >= 0) { adj[cur[f.a]++] = f.b; adj[cur[f.b]++] = f.a; }
  }
  // Visiting order: body index (coloring.js:75-92), or the reference GPU's
  // claim order, lowbias32 priority with ties to the lower index
  // (layout.js:50-66, shaders.js:1571-1585); greedy in that order is the
  // fixed point its Jacobi rounds converge to.
  std::vector<int> visit(n);
  for (int i = 0; i < n; i++) visit[i] = i;
  if (colorOrder == 1)
    std::stable_sort(visit.begin(), visit.end(), [](int a, int b) {
      const uint32_t pa = colorPriority(uint32_t(a)), pb = colorPriority(uint32_t(b));
      return pa < pb || (pa == pb && a < b);
    });
  int numColors = 0, stamp = 0;
  std::vector<int> stampOf;
  for (int v = 0; v < n; v++) {
    const int i = visit[v];
    if (bodies[i].mass <= 0) continue;
    stamp++;
    for (int k = start[i]; k < start[i + 1]; k++) {
      const int c = out.colorOf[adj[k]];
      if (c >= 0) {
        if (int(stampOf.size()) <= c) stampOf.resize(c + 1, 0);
        stampOf[c] = stamp;
      }
    }
    int c = 0;
    while (c < numColors && c < int(stampOf.size()) && stampOf[c] == stamp) c++;
    if (c == numColors) numColors++;
    out.colorOf[i] = c;
  }
  std::vector<int> counts(numColors, 0);
  for (int i = 0; i < n; i++) if (out.colorOf[i] >= 0) counts[out.colorOf[i]]++;
  out.offsets.assign(numColors + 1, 0);
  for (int c = 0; c < numColors; c++) out.offsets[c + 1] = out.offsets[c] + counts[c];
  out.entries.assign(out.offsets[numColors], 0);
  std::vector<int> fill(out.offsets.begin(), out.offsets.end() - 1);
  for (int i = 0; i < n; i++) {
    const int c = out.colorOf[i];
    if (c >= 0) out.entries[fill[c]++] = i;
  }
  out.numColors = numColors;
}

// solver.js:307-493 — Algorithm 1
// Solver.removeBodies (solver.js:261-300): drops every flagged body and every
// force touching one (a world-anchored joint touches only its body B), keeps
// the order of the surviving bodies, of the force list and of each body's
// force list, and renumbers the bodies the surviving forces name. Returns the
// number of bodies removed.
int World::removeBodies(const std::vector<char> &doomed) {
  const int n = int(bodies.size());
  std::vector<int> remap(n, -1);
  int kept = 0;
  for (int i = 0; i < n; i++)
    if (!doomed[i]) remap[i] = kept++;
  if (kept == n) return 0;
  auto retained = [&](const Force &f) { return (f.a < 0 || !doomed[f.a]) && !doomed[f.b]; };
  int write = 0;
  for (size_t k = 0; k < forces.size(); k++) {
    if (retained(pool[forces[k]])) forces[write++] = forces[k];
    else if (pool[forces[k]].kind == MANIFOLD) freeForces.push_back(forces[k]);
  }
  forces.resize(write);
  std::vector<Body> survivors;
  survivors.reserve(kept);
  for (int i = 0; i < n; i++) {
    if (doomed[i]) continue;
    Body b = bodies[i];
    int fw = 0;
    for (size_t k = 0; k < b.forces.size(); k++)
      if (retained(pool[b.forces[k]])) b.forces[fw++] = b.forces[k];
    b.forces.resize(fw);
    survivors.push_back(b);
  }
  bodies.swap(survivors);
  for (int id : forces) {
    Force &f = pool[id];
    if (f.a >= 0) f.a = remap[f.a];
    f.b = remap[f.b];
  }
  bodyOrder.clear();
  lastPairs.clear();
  return n - kept;
}

void World::step() {
  const Params &P = prm;
  const double dt = P.dt;
  const int n = int(bodies.size());

  // 1. broad phase (solver.js:316-336)
  if (recordPairs) lastPairs.clear();
  // The recorded pair list is every candidate the predicate accepts with at
  // least one dynamic body and no live joint between the two, whether or not a
  // manifold already exists; a manifold is created for those without one.
  forEachPair(*this, [&](int i, int j) {
    const Body &A = bodies[i], &B = bodies[j];
    const double dx = A.p[0] - B.p[0], dy = A.p[1] - B.p[1], dz = A.p[2] - B.p[2];
    // GPU contact model: bounding spheres grown by the pair's contact skin
    // (shaders.js:440-441)
    const double r = prm.contactModel == 1
                         ? A.radius + B.radius + (bodyContactOffset(prm, A) + bodyContactOffset(prm, B))
                         : A.radius + B.radius;
    if (dx * dx + dy * dy + dz * dz <= r * r) {
      if (recordPairs && (A.mass > 0 || B.mass > 0) && !jointConnected(*this, i, j)) {
        lastPairs.push_back(i);
        lastPairs.push_back(j);
      }
      if (!constrainedTo(*this, i, j) && (A.mass > 0 || B.mass > 0)) newForce(*this, MANIFOLD, i, j);
    }
  }, broadphase == 1);

  // 2. force initialisation and warm start (solver.js:344-364)
  int write = 0;
  for (size_t k = 0; k < forces.size(); k++) {
    const int id = forces[k];
    Force &f = pool[id];
    const bool keep = f.kind == MANIFOLD ? manifoldInitialize(*this, f)
                      : f.kind == JOINT  ? jointInitialize(*this, f)
                                         : springInitialize(*this, f);
    if (keep) {
      forces[write++] = id;
    } else {
      if (f.a >= 0) removeFromBody(bodies[f.a].forces, id);
      if (f.b >= 0) removeFromBody(bodies[f.b].forces, id);
      // only contact slots are taken again: joints and springs keep theirs,
      // whose order numbers them (orc_set_joint_anchor)
      if (f.kind == MANIFOLD) freeForces.push_back(id);
    }
  }
  forces.resize(write);

  if (schedule == SCHED_COLOR) buildColoring(coloring);

  // 3. body initialisation (solver.js:367-401)
  const double g[3] = {0, 0, P.gravity};
  for (int i = 0; i < n; i++) {
    Body &b = bodies[i];
    v_copy(b.py, b.p);
    v_addScaled(b.py, b.v, dt);
    if (b.mass > 0) v_addScaled(b.py, g, dt * dt);
    double wdt[3] = {b.w[0] * dt, b.w[1] * dt, b.w[2] * dt};
    q_integrate(b.qy, b.q, wdt);
    const double accelZ = (b.v[2] - b.vprev[2]) / dt;
    const double accelExt = accelZ * sgn(P.gravity);
    double wgt = clampd(accelExt / std::fabs(P.gravity), 0.0, 1.0);
    if (!std::isfinite(wgt)) wgt = 0.0;
    v_copy(b.p0, b.p);
    q_copy(b.q0, b.q);
    if (b.mass > 0) {
      v_addScaled(b.p, b.v, dt);
      v_addScaled(b.p, g, wgt * dt * dt);
      double wdt2[3] = {b.w[0] * dt, b.w[1] * dt, b.w[2] * dt};
      q_integrate(b.q, b.q, wdt2);
    }
  }

  // 4. iterations (solver.js:404-489)
  const double dt2 = dt * dt;
  const int total = P.iterations + (P.postStabilize ? 1 : 0);
  const std::vector<int> *order = nullptr;
  if (schedule == SCHED_ORDER) order = &bodyOrder;
  else if (schedule == SCHED_COLOR) order = &coloring.entries;
  for (int it = 0; it < total; it++) {
    const double alpha = P.postStabilize ? (it < P.iterations ? 1.0 : 0.0) : P.alpha;
    const int sweep = order ? int(order->size()) : n;
    for (int s = 0; s < sweep; s++) {
      const int bi = order ? (*order)[s] : n - 1 - s;
      Body &b = bodies[bi];
      if (b.mass <= 0) continue;
      double MLin[9], MAng[9], lhsLin[9], lhsAng[9], lhsCross[9], rhsLin[3], rhsAng[3], tmp[3], dq[3];
      m_diag(MLin, b.mass, b.mass, b.mass);
      worldMoment(MAng, b, P.rotatedInertia);
      m_div(lhsLin, MLin, dt2);
      m_div(lhsAng, MAng, dt2);
      m_zero(lhsCross);
      v_sub(tmp, b.p, b.py);
      m_mulVec(rhsLin, lhsLin, tmp);
      q_subtract(dq, b.q, b.qy);
      m_mulVec(rhsAng, lhsAng, dq);
      for (int k = int(b.forces.size()) - 1; k >= 0; k--) {
        Force &f = pool[b.forces[k]];
        if (f.kind == MANIFOLD) manifoldPrimal(*this, f, bi, alpha, lhsLin, lhsAng, lhsCross, rhsLin, rhsAng);
        else if (f.kind == JOINT) jointPrimal(*this, f, bi, alpha, lhsLin, lhsAng, lhsCross, rhsLin, rhsAng);
        else springPrimal(*this, f, bi, lhsLin, lhsAng, lhsCross, rhsLin, rhsAng);
      }
      v_neg(rhsLin, rhsLin);
      v_neg(rhsAng, rhsAng);
      double dxL[3], dxA[3];
      solve6(lhsLin, lhsAng, lhsCross, rhsLin, rhsAng, dxL, dxA);
      v_add(b.p, b.p, dxL);
      q_integrate(b.q, b.q, dxA);
    }
    if (it < P.iterations) {
      for (int k = int(forces.size()) - 1; k >= 0; k--) {
        Force &f = pool[forces[k]];
        if (f.kind == MANIFOLD) manifoldDual(*this, f, alpha);
        else if (f.kind == JOINT) jointDual(*this, f, alpha);
        else springDual(*this, f);
      }
    }
    if (it == P.iterations - 1) {
      for (int i = 0; i < n; i++) {
        Body &b = bodies[i];
        v_copy(b.vprev, b.v);
        if (b.mass > 0) {
          v_sub(b.v, b.p, b.p0);
          v_div(b.v, b.v, dt);
          q_subtract(b.w, b.q, b.q0);
          v_div(b.w, b.w, dt);
        }
      }
    }
  }
}

}  // namespace orc
