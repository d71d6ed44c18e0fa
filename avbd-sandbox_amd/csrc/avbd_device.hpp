// avbd_device.hpp — device data layout and the per-thread building blocks of
// the AVBD step: narrow phase (collide.js), warm-start merge
// (manifold.js:214-318), contact evaluation (manifold.js:358-409) and the
// joint rows (joint.js). The operation order follows the reference's f64
// engine; instantiated with double it reproduces it exactly.
#pragma once
#include <cstdint>

#include "avbd_math.hpp"

namespace avbd {

constexpr int MAXC = 8;            // contacts per manifold (collide.js:22)

// Static per-body data, read-only after pack.
struct BodyProp {
  double mass, m0, m1, m2;         // mass and body-frame principal moments
  double hx, hy, hz, radius;       // half extents (size * 0.5), bounding radius
  double sx, sy, sz, friction;     // full size, friction coefficient
  int32_t shape, scene, large, pad;
};

// Contact material of the GPU contact model, per body, resolved when it is
// uploaded (bodyContactSignature, backend.js:156-185): static and dynamic
// friction, restitution, the contact offset (at least 0) and the rest offset
// (at most the contact offset), each its own or else the solver's.
struct BodyMat {
  double staticFriction, dynamicFriction, restitution, contactOffset, restOffset, pad;
};

// Body state rows, double, 8 per body: X = [p0 p1 p2 q0 q1 q2 q3 -].
constexpr int XS = 8;
// Velocities, 12 per body: v[3] w[3] vprev[3] -.
constexpr int VS = 12;

// Joint, always double (joint.js:78-129).
// kind AVBD_JOINT_SPRING: a distance spring (spring.js:63-105) in the same
// record: rA/rB its anchors, stiffLin the stiffness, stiffAng the rest
// length, fracture the tearing strain, kLin[0] the penalty, CLin[0] the last
// constraint value, lamLin[0] the peak strain; broken 1 = torn (an inactive
// entry that stays in the force list), 2 = left the force list.
struct JointDev {
  int32_t a, b, broken, kind;
  double rA[3], rB[3];
  double stiffLin, stiffAng, fracture, torqueArm;
  double C0Lin[3], C0Ang[3], kLin[3], kAng[3], lamLin[3], lamAng[3], CLin[3], CAng[3];
};

// One manifold buffer (two are kept: last step's and this step's).
// Per manifold: head = {a, b, nc, flags}, basis[12] = row-major basis, mu.
// Per contact slot s of manifold m (index m * MAXC + s):
//   hot[16]: C0[3] k[3] lam[3] rAw[3] rBw[3] -  (rAw/rBw: anchors rotated
//            by the x_t orientations, the cached Jacobian of manifold.js:338-349)
//   anchor[6]: rA[3] rB[3] body-frame anchors (double)
//   feat, stick, Cd[4] (last constraint value, diagnostics)
template <class R> struct ManBuf {
  int4 *head;
  R *basis;
  R *hot;
  double *anchor;
  uint32_t *feat;
  int32_t *stick;
  R *Cd;
};
constexpr int HOT = 16;
constexpr int BAS = 12;
// Offsets in a hot contact record. k and lambda lead so that the dual update
// writes them as aligned vectors.
constexpr int HK = 0, HL = 3, HC0 = 6, HRA = 9, HRB = 12;
// GPU contact model: hot[HB] the contact's restitution target; basis[BMU]
// the static friction coefficient (the CPU engine's only one), basis[BMUD]
// the dynamic one, basis[BCD] the pair's contact distance (skin).
constexpr int HB = 15, BMU = 9, BMUD = 10, BCD = 11;

// The contact law an evaluation follows: the CPU engine's (manifold.js:
// 358-491), the reference GPU backend's (shaders.js:1777-1850, 2229-2264),
// or the GPU backend's in a post-stabilisation projection pass (alpha below
// 0.5), which leaves the restitution target out.
constexpr int LAW_CPU = 0, LAW_GPU = 1, LAW_GPU_PROJ = 2;
__host__ __device__ inline int contactLaw(int model, int postStab, double alpha) {
  return model ? (postStab && alpha < 0.5 ? LAW_GPU_PROJ : LAW_GPU) : LAW_CPU;
}

// Aligned 16-byte loads of N consecutive numbers (N a multiple of 4 floats or
// 2 doubles; the records are 16-byte aligned).
template <int N> __device__ __forceinline__ void vload(float *o, const float *p) {
  const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
  for (int k = 0; k < N / 4; k++) {
    const float4 v = q[k];
    o[4 * k] = v.x; o[4 * k + 1] = v.y; o[4 * k + 2] = v.z; o[4 * k + 3] = v.w;
  }
}
template <int N> __device__ __forceinline__ void vload(double *o, const double *p) {
  const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
  for (int k = 0; k < N / 2; k++) {
    const double2 v = q[k];
    o[2 * k] = v.x; o[2 * k + 1] = v.y;
  }
}
// Store the six numbers k[3], lambda[3] at the head of a hot record.
__device__ __forceinline__ void vstoreKL(float *p, const float *kl) {
  reinterpret_cast<float4 *>(p)[0] = make_float4(kl[0], kl[1], kl[2], kl[3]);
  reinterpret_cast<float2 *>(p)[2] = make_float2(kl[4], kl[5]);
}
__device__ __forceinline__ void vstoreKL(double *p, const double *kl) {
  double2 *q = reinterpret_cast<double2 *>(p);
  q[0] = make_double2(kl[0], kl[1]);
  q[1] = make_double2(kl[2], kl[3]);
  q[2] = make_double2(kl[4], kl[5]);
}
constexpr uint32_t MF_SPHERE = 1u;  // head.w flag: a sphere is involved

// ---------------------------------------------------------------------------
// Narrow phase, double (collide.js).
// ---------------------------------------------------------------------------
constexpr int MAXPOLY = 16;
constexpr double SAT_EPS = 1.0e-6;
constexpr double SPHERE_EPS_SQ = 1.0e-20;
constexpr double PLANE_EPS = 1.0e-5;
constexpr double MERGE_DIST_SQ = 1.0e-6;

struct Obb { double c[3], h[3], ax[3][3]; };

struct BodyView { const double *p, *q; double hx, hy, hz, radius; int shape; };

__device__ inline void makeObb(Obb &o, const BodyView &b) {
  const double ux[3] = {1, 0, 0}, uy[3] = {0, 1, 0}, uz[3] = {0, 0, 1};
  o.c[0] = b.p[0]; o.c[1] = b.p[1]; o.c[2] = b.p[2];
  o.h[0] = b.hx; o.h[1] = b.hy; o.h[2] = b.hz;
  qrot(o.ax[0], b.q, ux);
  qrot(o.ax[1], b.q, uy);
  qrot(o.ax[2], b.q, uz);
}

__device__ inline void supportPt(double *o, const Obb &b, const double *dir) {
  const double sx = dot3(dir, b.ax[0]) >= 0 ? 1.0 : -1.0;
  const double sy = dot3(dir, b.ax[1]) >= 0 ? 1.0 : -1.0;
  const double sz = dot3(dir, b.ax[2]) >= 0 ? 1.0 : -1.0;
  for (int k = 0; k < 3; k++) o[k] = b.c[k];
  for (int k = 0; k < 3; k++) o[k] += b.ax[0][k] * (b.h[0] * sx);
  for (int k = 0; k < 3; k++) o[k] += b.ax[1][k] * (b.h[1] * sy);
  for (int k = 0; k < 3; k++) o[k] += b.ax[2][k] * (b.h[2] * sz);
}

// A value the optimiser cannot trace back to the memory it was read from: a
// select between two such values stays a select of registers. (A select
// between two loads of one structure is otherwise turned into a load at a
// computed offset, which keeps the whole structure in private memory.)
__device__ __forceinline__ double opaque(double x) {
  asm("" : "+v"(x));
  return x;
}

// Axis i and half extent i of a box by selects, not by indexing, so that a
// box stays in registers.
__device__ __forceinline__ void axOf(const Obb &b, int i, double *o) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double a0 = opaque(b.ax[0][k]), a1 = opaque(b.ax[1][k]), a2 = opaque(b.ax[2][k]);
    o[k] = i == 0 ? a0 : (i == 1 ? a1 : a2);
  }
}
__device__ __forceinline__ double hOf(const Obb &b, int i) {
  const double h0 = opaque(b.h[0]), h1 = opaque(b.h[1]), h2 = opaque(b.h[2]);
  return i == 0 ? h0 : (i == 1 ? h1 : h2);
}

__device__ inline void faceAx(const Obb &b, int ai, double *u, double *v, double &eu, double &ev) {
  const int iu = ai == 0 ? 1 : 0, iv = ai == 2 ? 1 : 2;
  axOf(b, iu, u);
  axOf(b, iv, v);
  eu = hOf(b, iu);
  ev = hOf(b, iv);
}

__device__ __forceinline__ void pickObb(Obb &o, bool first, const Obb &a, const Obb &b) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    o.c[k] = first ? opaque(a.c[k]) : opaque(b.c[k]);
    o.h[k] = first ? opaque(a.h[k]) : opaque(b.h[k]);
#pragma unroll
    for (int j = 0; j < 3; j++) o.ax[j][k] = first ? opaque(a.ax[j][k]) : opaque(b.ax[j][k]);
  }
}

// Sutherland-Hodgman against the half-space n.x <= off (collide.js:202-247).
// Vertex i coordinate k lives at buf[(3 i + k) * STRIDE]. CAP vertices fit;
// with SIGNAL set, needing more returns -1 (the caller redoes the manifold
// with MAXPOLY room), otherwise extra vertices are dropped as the reference
// drops them beyond MAXPOLY.
template <int STRIDE, int CAP, bool SIGNAL>
__device__ inline int clipPolyT(const double *in, int n, const double *pn, double off, double *out) {
  if (n <= 0) return 0;
  int cnt = 0;
  double ax = in[((n - 1) * 3) * STRIDE], ay = in[((n - 1) * 3 + 1) * STRIDE], az = in[((n - 1) * 3 + 2) * STRIDE];
  double da = pn[0] * ax + pn[1] * ay + pn[2] * az - off;
  for (int i = 0; i < n; i++) {
    const double bx = in[(i * 3) * STRIDE], by = in[(i * 3 + 1) * STRIDE], bz = in[(i * 3 + 2) * STRIDE];
    const double db = pn[0] * bx + pn[1] * by + pn[2] * bz - off;
    const bool aIn = da <= PLANE_EPS, bIn = db <= PLANE_EPS;
    if (aIn != bIn) {
      double t = 0;
      const double den = da - db;
      if (xabs(den) > SAT_EPS) t = clampj(da / den, 0.0, 1.0);
      if (cnt < CAP) {
        out[(cnt * 3) * STRIDE] = ax + (bx - ax) * t;
        out[(cnt * 3 + 1) * STRIDE] = ay + (by - ay) * t;
        out[(cnt * 3 + 2) * STRIDE] = az + (bz - az) * t;
        cnt++;
      } else if (SIGNAL) {
        return -1;
      }
    }
    if (bIn) {
      if (cnt < CAP) {
        out[(cnt * 3) * STRIDE] = bx; out[(cnt * 3 + 1) * STRIDE] = by; out[(cnt * 3 + 2) * STRIDE] = bz;
        cnt++;
      } else if (SIGNAL) {
        return -1;
      }
    }
    ax = bx; ay = by; az = bz; da = db;
  }
  return cnt;
}

__device__ inline int clipPoly(const double *in, int n, const double *pn, double off, double *out) {
  return clipPolyT<1, MAXPOLY, false>(in, n, pn, off, out);
}

// A double that may share memory with other types (the merge midpoints below
// live for a while in space that later holds a slot's R-typed hot record).
typedef double __attribute__((may_alias)) AliasDouble;

// Where the contacts of one manifold go as they are generated: the bodies'
// pose pointers, the manifold's output slots (body-frame anchors rA, rB: 6
// doubles per contact; feature keys) and, for the merge test, the contact
// midpoints (contact i at mid + i * midStride: the slots' hot-record space,
// free until the records are written). Nothing is kept in private memory.
struct Sink {
  const double *pA, *qA, *pB, *qB;
  double *anchor;
  uint32_t *feat;
  AliasDouble *mid;
  int midStride;
  int count;
  // the contact skin of the GPU contact model (0 otherwise): pairs within it
  // produce speculative contacts (shaders.js:821-1055)
  double skin;
  // collide.js:254-285
  __device__ bool add(const double *xA, const double *xB, uint32_t key) {
    const double mx = (xA[0] + xB[0]) * 0.5, my = (xA[1] + xB[1]) * 0.5, mz = (xA[2] + xB[2]) * 0.5;
    for (int i = 0; i < count; i++) {
      const AliasDouble *m = mid + i * midStride;
      const double dx = mx - m[0], dy = my - m[1], dz = mz - m[2];
      if (dx * dx + dy * dy + dz * dz < MERGE_DIST_SQ) return false;
    }
    if (count >= MAXC) return false;
    feat[count] = key;
    double inv[4], t[3], r[3];
    double *an = anchor + count * 6;
    inv[0] = -qA[0]; inv[1] = -qA[1]; inv[2] = -qA[2]; inv[3] = qA[3];
    t[0] = xA[0] - pA[0]; t[1] = xA[1] - pA[1]; t[2] = xA[2] - pA[2];
    qrot(r, inv, t);
    an[0] = r[0]; an[1] = r[1]; an[2] = r[2];
    inv[0] = -qB[0]; inv[1] = -qB[1]; inv[2] = -qB[2]; inv[3] = qB[3];
    t[0] = xB[0] - pB[0]; t[1] = xB[1] - pB[1]; t[2] = xB[2] - pB[2];
    qrot(r, inv, t);
    an[3] = r[0]; an[4] = r[1]; an[5] = r[2];
    AliasDouble *m = mid + count * midStride;
    m[0] = mx; m[1] = my; m[2] = mz;
    count++;
    return true;
  }
};

struct Sat { int type, ia, ib; double sep; double n[3]; bool valid; };

// collide.js:291-324
__device__ inline bool satAxis(const Obb &A, const Obb &B, const double *delta, const double *axis, int type,
                               int ia, int ib, Sat &t, double skin) {
  const double lenSq = dot3(axis, axis);
  if (lenSq < SAT_EPS) return true;
  const double invLen = 1.0 / xsqrt(lenSq);
  double n[3] = {axis[0] * invLen, axis[1] * invLen, axis[2] * invLen};
  if (dot3(n, delta) < 0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
  const double dist = xabs(dot3(delta, n));
  const double rA = A.h[0] * xabs(dot3(n, A.ax[0])) + A.h[1] * xabs(dot3(n, A.ax[1])) + A.h[2] * xabs(dot3(n, A.ax[2]));
  const double rB = B.h[0] * xabs(dot3(n, B.ax[0])) + B.h[1] * xabs(dot3(n, B.ax[1])) + B.h[2] * xabs(dot3(n, B.ax[2]));
  const double sep = dist - (rA + rB);
  if (sep > skin) return false;
  if (!t.valid || sep > t.sep) {
    t.valid = true; t.type = type; t.ia = ia; t.ib = ib; t.sep = sep;
    t.n[0] = n[0]; t.n[1] = n[1]; t.n[2] = n[2];
  }
  return true;
}

// collide.js:327-343
__device__ inline void supportEdge(const Obb &b, int ai, const double *dir, double *e0, double *e1) {
  const int a1 = (ai + 1) % 3, a2 = (ai + 2) % 3;
  double x0[3], x1[3], x2[3];
  axOf(b, ai, x0);
  axOf(b, a1, x1);
  axOf(b, a2, x2);
  const double h0 = hOf(b, ai), h1 = hOf(b, a1), h2 = hOf(b, a2);
  const double s1 = dot3(dir, x1) >= 0 ? 1.0 : -1.0;
  const double s2 = dot3(dir, x2) >= 0 ? 1.0 : -1.0;
  double t[3] = {b.c[0], b.c[1], b.c[2]};
  for (int k = 0; k < 3; k++) t[k] += x1[k] * (h1 * s1);
  for (int k = 0; k < 3; k++) t[k] += x2[k] * (h2 * s2);
  for (int k = 0; k < 3; k++) { e0[k] = t[k]; e0[k] += x0[k] * -h0; }
  for (int k = 0; k < 3; k++) { e1[k] = t[k]; e1[k] += x0[k] * h0; }
}

// collide.js:345-393
__device__ inline void closestSeg(const double *p0, const double *p1, const double *q0, const double *q1, double *c0,
                                  double *c1) {
  const double d1x = p1[0] - p0[0], d1y = p1[1] - p0[1], d1z = p1[2] - p0[2];
  const double d2x = q1[0] - q0[0], d2y = q1[1] - q0[1], d2z = q1[2] - q0[2];
  const double rx = p0[0] - q0[0], ry = p0[1] - q0[1], rz = p0[2] - q0[2];
  const double a = d1x * d1x + d1y * d1y + d1z * d1z;
  const double e = d2x * d2x + d2y * d2y + d2z * d2z;
  const double f = d2x * rx + d2y * ry + d2z * rz;
  double s = 0, t = 0;
  if (a <= SAT_EPS && e <= SAT_EPS) {
    for (int k = 0; k < 3; k++) { c0[k] = p0[k]; c1[k] = q0[k]; }
    return;
  }
  if (a <= SAT_EPS) {
    t = clampj(f / e, 0.0, 1.0);
  } else {
    const double c = d1x * rx + d1y * ry + d1z * rz;
    if (e <= SAT_EPS) {
      s = clampj(-c / a, 0.0, 1.0);
    } else {
      const double bb = d1x * d2x + d1y * d2y + d1z * d2z;
      const double den = a * e - bb * bb;
      if (xabs(den) > SAT_EPS) s = clampj((bb * f - c * e) / den, 0.0, 1.0);
      t = (bb * s + f) / e;
      if (t < 0) { t = 0; s = clampj(-c / a, 0.0, 1.0); }
      else if (t > 1) { t = 1; s = clampj((bb - c) / a, 0.0, 1.0); }
    }
  }
  c0[0] = p0[0] + d1x * s; c0[1] = p0[1] + d1y * s; c0[2] = p0[2] + d1z * s;
  c1[0] = q0[0] + d2x * t; c1[1] = q0[1] + d2y * t; c1[2] = q0[2] + d2z * t;
}

// collide.js:403-472. The clip polygons c0, c1 hold vertex i coordinate k at
// [(3 i + k) * STRIDE]; returns -1 if CAP vertices do not suffice (SIGNAL).
template <int STRIDE, int CAP, bool SIGNAL>
__device__ inline int faceManifoldT(const Obb &oA, const Obb &oB, Sink &sink, bool refIsA, int refAxis,
                                    const double *nAB, double *c0, double *c1) {
  Obb ref, inc;
  pickObb(ref, refIsA, oA, oB);
  pickObb(inc, !refIsA, oA, oB);
  double on[3];
  for (int k = 0; k < 3; k++) on[k] = refIsA ? nAB[k] : -nAB[k];
  double fn[3], fc[3], fu[3], fv[3], eu, ev, rax[3];
  axOf(ref, refAxis, rax);
  const double s = dot3(on, rax) >= 0 ? 1.0 : -1.0;
  for (int k = 0; k < 3; k++) fn[k] = rax[k] * s;
  const double rh = hOf(ref, refAxis);
  for (int k = 0; k < 3; k++) { fc[k] = ref.c[k]; fc[k] += fn[k] * rh; }
  faceAx(ref, refAxis, fu, fv, eu, ev);
  int incAxis = 0;
  double best = -__builtin_inf();
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double d = xabs(dot3(inc.ax[i], fn));
    if (d > best) { best = d; incAxis = i; }
  }
  {
    double iax[3];
    axOf(inc, incAxis, iax);
    const double si = dot3(iax, fn) > 0 ? -1.0 : 1.0;
    double nrm[3], ctr[3], u[3], v[3], iu, iv;
    for (int k = 0; k < 3; k++) nrm[k] = iax[k] * si;
    const double ih = hOf(inc, incAxis);
    for (int k = 0; k < 3; k++) { ctr[k] = inc.c[k]; ctr[k] += nrm[k] * ih; }
    faceAx(inc, incAxis, u, v, iu, iv);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const double su = (i == 0 || i == 3) ? 1.0 : -1.0;
      const double sv = (i == 0 || i == 1) ? 1.0 : -1.0;
      for (int k = 0; k < 3; k++) c0[(i * 3 + k) * STRIDE] = ctr[k] + u[k] * (iu * su) + v[k] * (iv * sv);
    }
  }
  int count = 4;
  double pn[3];
  for (int k = 0; k < 3; k++) pn[k] = fu[k];
  count = clipPolyT<STRIDE, CAP, SIGNAL>(c0, count, pn, dot3(pn, fc) + eu, c1);
  if (count <= 0) return count;
  for (int k = 0; k < 3; k++) pn[k] = -fu[k];
  count = clipPolyT<STRIDE, CAP, SIGNAL>(c1, count, pn, dot3(pn, fc) + eu, c0);
  if (count <= 0) return count;
  for (int k = 0; k < 3; k++) pn[k] = fv[k];
  count = clipPolyT<STRIDE, CAP, SIGNAL>(c0, count, pn, dot3(pn, fc) + ev, c1);
  if (count <= 0) return count;
  for (int k = 0; k < 3; k++) pn[k] = -fv[k];
  count = clipPolyT<STRIDE, CAP, SIGNAL>(c1, count, pn, dot3(pn, fc) + ev, c0);
  if (count <= 0) return count;

  sink.count = 0;
  uint32_t prefix = uint32_t(refIsA ? 0 : 1) << 24;
  prefix |= uint32_t(refAxis & 0xff) << 16;
  prefix |= uint32_t(incAxis & 0xff) << 8;
  for (int i = 0; i < count && sink.count < MAXC; i++) {
    double p[3], t[3];
    for (int k = 0; k < 3; k++) p[k] = c0[(i * 3 + k) * STRIDE];
    for (int k = 0; k < 3; k++) t[k] = p[k] - fc[k];
    const double dist = dot3(t, fn);
    if (dist > sink.skin + PLANE_EPS) continue;
    for (int k = 0; k < 3; k++) { t[k] = p[k]; t[k] += fn[k] * -dist; }
    if (refIsA) sink.add(t, p, prefix | uint32_t(i & 0xff));
    else sink.add(p, t, prefix | uint32_t(i & 0xff));
  }
  if (!sink.count) {
    double xA[3], xB[3], nn[3] = {-nAB[0], -nAB[1], -nAB[2]};
    supportPt(xA, oA, nAB);
    supportPt(xB, oB, nn);
    sink.add(xA, xB, prefix);
  }
  return sink.count;
}

// Face manifold with the clip polygons in the caller's LDS column (POLY_LDS
// vertices, lane-interleaved); a manifold that needs more is redone with
// MAXPOLY room in private memory.
constexpr int POLY_LDS = 8, LDS_LANES = 64;
__device__ inline int faceManifold(const Obb &oA, const Obb &oB, Sink &sink, bool refIsA, int refAxis,
                                   const double *nAB, double *lds) {
  if (lds) {
    const int r = faceManifoldT<LDS_LANES, POLY_LDS, true>(oA, oB, sink, refIsA, refAxis, nAB, lds,
                                                           lds + POLY_LDS * 3 * LDS_LANES);
    if (r >= 0) return r;
  }
  double c0[MAXPOLY * 3], c1[MAXPOLY * 3];
  return faceManifoldT<1, MAXPOLY, false>(oA, oB, sink, refIsA, refAxis, nAB, c0, c1);
}

// collide.js:475-493
__device__ inline int edgeContact(const Obb &oA, const Obb &oB, Sink &sink, int axA, int axB, const double *nAB) {
  double nn[3] = {-nAB[0], -nAB[1], -nAB[2]}, a0[3], a1[3], b0[3], b1[3], xA[3], xB[3];
  supportEdge(oA, axA, nAB, a0, a1);
  supportEdge(oB, axB, nn, b0, b1);
  closestSeg(a0, a1, b0, b1, xA, xB);
  sink.count = 0;
  const uint32_t key = (2u << 24) | (uint32_t(axA & 0xff) << 8) | uint32_t(axB & 0xff);
  sink.add(xA, xB, key);
  if (!sink.count) {
    supportPt(xA, oA, nAB);
    supportPt(xB, oB, nn);
    sink.add(xA, xB, key);
  }
  return sink.count;
}

// collide.js:496-520
__device__ inline int sphereSphere(const BodyView &A, const BodyView &B, Sink &sink, double *basis) {
  double d[3] = {B.p[0] - A.p[0], B.p[1] - A.p[1], B.p[2] - A.p[2]}, n[3], t[3], xA[3], xB[3];
  const double dsq = dot3(d, d);
  const double rs = A.radius + B.radius + sink.skin;
  if (dsq > rs * rs) return 0;
  if (dsq > SPHERE_EPS_SQ) {
    const double inv = 1 / xsqrt(dsq);
    n[0] = d[0] * inv; n[1] = d[1] * inv; n[2] = d[2] * inv;
  } else {
    n[0] = 1; n[1] = 0; n[2] = 0;
  }
  t[0] = -n[0]; t[1] = -n[1]; t[2] = -n[2];
  orthonormal(basis, t);
  for (int k = 0; k < 3; k++) { xA[k] = A.p[k]; xA[k] += n[k] * A.radius; }
  for (int k = 0; k < 3; k++) { xB[k] = B.p[k]; xB[k] += n[k] * -B.radius; }
  sink.count = 0;
  sink.add(xA, xB, 0x31000001u);
  return sink.count;
}

// collide.js:526-600
__device__ inline int sphereBox(const BodyView &A, const BodyView &B, Sink &sink, double *basis, bool sphereIsA) {
  // the two roles by field selects (a select of references would keep both
  // views in private memory)
  const BodyView sph{sphereIsA ? A.p : B.p, sphereIsA ? A.q : B.q, sphereIsA ? A.hx : B.hx,
                     sphereIsA ? A.hy : B.hy, sphereIsA ? A.hz : B.hz, sphereIsA ? A.radius : B.radius,
                     sphereIsA ? A.shape : B.shape};
  const BodyView bx{sphereIsA ? B.p : A.p, sphereIsA ? B.q : A.q, sphereIsA ? B.hx : A.hx,
                    sphereIsA ? B.hy : A.hy, sphereIsA ? B.hz : A.hz, sphereIsA ? B.radius : A.radius,
                    sphereIsA ? B.shape : A.shape};
  Obb o;
  makeObb(o, bx);
  double d[3], cp[3], sv[3], n[3], xA[3], xB[3];
  for (int k = 0; k < 3; k++) d[k] = sph.p[k] - o.c[k];
  const double lx = dot3(d, o.ax[0]), ly = dot3(d, o.ax[1]), lz = dot3(d, o.ax[2]);
  const double cx = clampj(lx, -o.h[0], o.h[0]);
  const double cy = clampj(ly, -o.h[1], o.h[1]);
  const double cz = clampj(lz, -o.h[2], o.h[2]);
  for (int k = 0; k < 3; k++) cp[k] = o.c[k];
  for (int k = 0; k < 3; k++) cp[k] += o.ax[0][k] * cx;
  for (int k = 0; k < 3; k++) cp[k] += o.ax[1][k] * cy;
  for (int k = 0; k < 3; k++) cp[k] += o.ax[2][k] * cz;
  for (int k = 0; k < 3; k++) sv[k] = sph.p[k] - cp[k];
  const double dsq = dot3(sv, sv);
  int fa = 0;
  if (dsq > SPHERE_EPS_SQ) {
    const double dist = xsqrt(dsq);
    if (dist > sph.radius + sink.skin) return 0;
    const double inv = 1 / dist;
    for (int k = 0; k < 3; k++) n[k] = sv[k] * inv;
  } else {
    const double gx = o.h[0] - xabs(lx), gy = o.h[1] - xabs(ly), gz = o.h[2] - xabs(lz);
    fa = gx < gy ? (gx < gz ? 0 : 2) : (gy < gz ? 1 : 2);
    const double local = fa == 0 ? lx : fa == 1 ? ly : lz;
    const double s = local >= 0 ? 1.0 : -1.0;
    double af[3];
    axOf(o, fa, af);
    for (int k = 0; k < 3; k++) n[k] = af[k] * s;
    const double gap = fa == 0 ? gx : fa == 1 ? gy : gz;
    for (int k = 0; k < 3; k++) { cp[k] = sph.p[k]; cp[k] += n[k] * gap; }
  }
  for (int k = 0; k < 3; k++) { sv[k] = sph.p[k]; sv[k] += n[k] * -sph.radius; }
  if (sphereIsA) {
    for (int k = 0; k < 3; k++) { xA[k] = sv[k]; xB[k] = cp[k]; }
    orthonormal(basis, n);
  } else {
    double t[3] = {-n[0], -n[1], -n[2]};
    for (int k = 0; k < 3; k++) { xA[k] = cp[k]; xB[k] = sv[k]; }
    orthonormal(basis, t);
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
  sink.add(xA, xB, 0x32000000u | (region << 2) | (sphereIsA ? 1u : 2u));
  return sink.count;
}

// collide.js:612-685. Writes the basis only where the reference does.
// skin: the GPU contact model's contact distance, 0 for the CPU engine's
// (the same comparisons then, bit for bit).
__device__ inline int collide(const BodyView &A, const BodyView &B, double *anchor, uint32_t *feat, AliasDouble *mid,
                              int midStride, double *basis, double *lds = nullptr, double skin = 0.0) {
  Sink sink{A.p, A.q, B.p, B.q, anchor, feat, mid, midStride, 0, skin};
  const bool sA = A.shape == 1, sB = B.shape == 1;
  if (sA && sB) return sphereSphere(A, B, sink, basis);
  if (sA || sB) return sphereBox(A, B, sink, basis, sA);
  Obb oA, oB;
  makeObb(oA, A);
  makeObb(oB, B);
  double delta[3] = {oB.c[0] - oA.c[0], oB.c[1] - oA.c[1], oB.c[2] - oA.c[2]};
  Sat face{0, -1, -1, -__builtin_inf(), {0, 0, 0}, false};
  Sat edge = face;
#pragma unroll
  for (int i = 0; i < 3; i++)
    if (!satAxis(oA, oB, delta, oA.ax[i], 0, i, -1, face, skin)) return 0;
#pragma unroll
  for (int i = 0; i < 3; i++)
    if (!satAxis(oA, oB, delta, oB.ax[i], 1, -1, i, face, skin)) return 0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double ax[3];
      cross3(ax, oA.ax[i], oB.ax[j]);
      if (!satAxis(oA, oB, delta, ax, 2, i, j, edge, skin)) return 0;
    }
  if (!face.valid) return 0;
  Sat best = face;
  if (edge.valid) {
    if (0.95 * edge.sep > face.sep + 0.01) best = edge;
  }
  double t[3] = {-best.n[0], -best.n[1], -best.n[2]};
  orthonormal(basis, t);
  if (best.type == 2) return edgeContact(oA, oB, sink, best.ia, best.ib, best.n);
  // one face routine for both reference sides (one inlined copy)
  const bool refIsA = best.type == 0;
  return faceManifold(oA, oB, sink, refIsA, refIsA ? best.ia : best.ib, best.n, lds);
}

// ---------------------------------------------------------------------------
// Contact evaluation (manifold.js:358-409), in the arithmetic type R.
// jLinA = basis, jLinB = -basis; jAngA row r = rAw x basis_r, jAngB row r =
// rBw x (-basis_r) (manifold.js:104-120).
// ---------------------------------------------------------------------------
template <class R> struct Eval {
  R jA[9], jB[9];   // angular blocks of A and B
  R C[3], F[3];
  bool within;
};

// Constraint value C = (1 - alpha) C0 + J dq of a contact from the cached
// Jacobian at x_t (manifold.js:325-410), and the angular Jacobian blocks.
template <class R, bool EXACT>
__device__ __forceinline__ void evalC(const R *bas, const R *negb, const R *hot, const R *dqA, const R *dqB, R alpha,
                                      Eval<R> &e, int law = LAW_CPU) {
  // angular blocks from the cached world anchors
#pragma unroll
  for (int r = 0; r < 3; r++) {
    R acc[3];
    cross3(acc, hot + HRA, bas + r * 3);
    e.jA[r * 3] = acc[0]; e.jA[r * 3 + 1] = acc[1]; e.jA[r * 3 + 2] = acc[2];
    cross3(acc, hot + HRB, negb + r * 3);
    e.jB[r * 3] = acc[0]; e.jB[r * 3 + 1] = acc[1]; e.jB[r * 3 + 2] = acc[2];
  }
  const R *C0 = hot + HC0;
  R acc[3];
  const R oma = R(1) - alpha;
  e.C[0] = C0[0] * oma; e.C[1] = C0[1] * oma; e.C[2] = C0[2] * oma;
  if (law != LAW_CPU) {
    // a positive speculative gap is kept whole; the restitution target is
    // subtracted outside a projection pass (shaders.js:1805-1819)
    if (bas[BCD] > R(0) && C0[0] > R(0)) e.C[0] = C0[0];
    if (law == LAW_GPU) e.C[0] = e.C[0] - hot[HB];
  }
  mulv(acc, bas, dqA);
  e.C[0] = e.C[0] + acc[0]; e.C[1] = e.C[1] + acc[1]; e.C[2] = e.C[2] + acc[2];
  mulv(acc, negb, dqB);
  e.C[0] = e.C[0] + acc[0]; e.C[1] = e.C[1] + acc[1]; e.C[2] = e.C[2] + acc[2];
  mulv(acc, e.jA, dqA + 3);
  e.C[0] = e.C[0] + acc[0]; e.C[1] = e.C[1] + acc[1]; e.C[2] = e.C[2] + acc[2];
  mulv(acc, e.jB, dqB + 3);
  e.C[0] = e.C[0] + acc[0]; e.C[1] = e.C[1] + acc[1]; e.C[2] = e.C[2] + acc[2];
}

// Force F = clamp(K C + lambda) into the friction cone (manifold.js:411-430).
// Under the GPU law (shaders.js:1824-1844) the cone is the static one while
// the contact sticks, the dynamic one otherwise, and a force outside it is
// scaled onto the dynamic cone.
template <class R, bool EXACT>
__device__ __forceinline__ void evalF(const R *k, const R *lam, R mu, Eval<R> &e, int law = LAW_CPU, R muD = R(0),
                                      int stick = 0) {
  if (EXACT) {
    // K C with the diagonal K written out as the reference's full product
    e.F[0] = k[0] * e.C[0] + R(0) * e.C[1] + R(0) * e.C[2];
    e.F[1] = R(0) * e.C[0] + k[1] * e.C[1] + R(0) * e.C[2];
    e.F[2] = R(0) * e.C[0] + R(0) * e.C[1] + k[2] * e.C[2];
  } else {
    e.F[0] = k[0] * e.C[0]; e.F[1] = k[1] * e.C[1]; e.F[2] = k[2] * e.C[2];
  }
  e.F[0] = e.F[0] + lam[0]; e.F[1] = e.F[1] + lam[1]; e.F[2] = e.F[2] + lam[2];
  e.F[0] = jmin(e.F[0], R(0));
  if (law != LAW_CPU) {
    const R fs = xsqrt(e.F[1] * e.F[1] + e.F[2] * e.F[2]);
    const R active = xabs(e.F[0]) * (stick ? mu : muD);
    e.within = fs <= active;
    if (!e.within && fs > R(0)) {
      const R dyn = xabs(e.F[0]) * muD;
      e.F[1] = e.F[1] * dyn / fs;
      e.F[2] = e.F[2] * dyn / fs;
    }
    return;
  }
  const R bounds = xabs(e.F[0]) * mu;
  const R fs = xsqrt(e.F[1] * e.F[1] + e.F[2] * e.F[2]);
  if (fs > bounds && fs > R(0)) {
    e.F[1] *= bounds / fs;
    e.F[2] *= bounds / fs;
  }
  e.within = fs <= bounds;
}

template <class R, bool EXACT>
__device__ __forceinline__ void evalContact(const R *bas, const R *negb, R mu, const R *hot, const R *dqA,
                                            const R *dqB, R alpha, Eval<R> &e, int law = LAW_CPU, int stick = 0) {
  evalC<R, EXACT>(bas, negb, hot, dqA, dqB, alpha, e, law);
  evalF<R, EXACT>(hot + HK, hot + HL, mu, e, law, bas[BMUD], stick);
}

// Manifold.updateDual for one contact (manifold.js:447-491) from its
// evaluation e with the current k: the new k and lambda into kl[6] (k then
// lambda), and the stick flag to store, or -1 to leave it as it is.
// Under the GPU law the contact sticks exactly when its force is inside the
// cone (shaders.js:2248-2258).
template <class R>
__device__ __forceinline__ int dualUpdate(const R *k, const Eval<R> &e, R beta, R pmax, R stickT, bool hasSphere,
                                          R *kl, int law = LAW_CPU) {
  kl[0] = k[0]; kl[1] = k[1]; kl[2] = k[2];
  kl[3] = e.F[0]; kl[4] = e.F[1]; kl[5] = e.F[2];
  if (e.F[0] < R(0)) kl[0] = jmin(k[0] + beta * xabs(e.C[0]), pmax);
  if (law != LAW_CPU) {
    if (e.within) {
      kl[1] = jmin(k[1] + beta * xabs(e.C[1]), pmax);
      kl[2] = jmin(k[2] + beta * xabs(e.C[2]), pmax);
    }
    return e.within ? 1 : 0;
  }
  int stick = -1;
  if (e.within) {
    kl[1] = jmin(k[1] + beta * xabs(e.C[1]), pmax);
    kl[2] = jmin(k[2] + beta * xabs(e.C[2]), pmax);
    stick = xsqrt(e.C[1] * e.C[1] + e.C[2] * e.C[2]) < stickT ? 1 : 0;
  } else if (hasSphere) {
    stick = 0;
  }
  return stick;
}

// ---------------------------------------------------------------- fast f64
// Reciprocal to about 1 ulp: the hardware estimate refined by two
// Newton-Raphson steps, r <- r + r (1 - d r).
__device__ __forceinline__ double frcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  return __builtin_fma(r, e, r);
}

// LDL^T of the 6x6 system (solve6's factorisation) with one reciprocal per
// pivot instead of a division per entry, kept for callers that solve twice
// with one matrix (the adjoint: the update, then its transpose solve).
struct Ldlt6 {
  double L21, L31, L41, L51, L61, L32, L42, L52, L62, L43, L53, L63, L54, L64, L65;
  double i1, i2, i3, i4, i5, i6;
};

__device__ __forceinline__ void ldlt6(const double *aLin, const double *aAng, const double *aCross, Ldlt6 &f) {
  // multiply-adds contracted: the throughput precisions and the adjoint use
  // this factorisation, never the exact mode (solve6 keeps its roundings)
#pragma clang fp contract(fast)
  const double A11 = aLin[0];
  const double A21 = aLin[3], A22 = aLin[4];
  const double A31 = aLin[6], A32 = aLin[7], A33 = aLin[8];
  const double A41 = aCross[0], A42 = aCross[1], A43 = aCross[2], A44 = aAng[0];
  const double A51 = aCross[3], A52 = aCross[4], A53 = aCross[5], A54 = aAng[3], A55 = aAng[4];
  const double A61 = aCross[6], A62 = aCross[7], A63 = aCross[8], A64 = aAng[6], A65 = aAng[7], A66 = aAng[8];
  const double D1 = A11, i1 = frcp(D1);
  const double L21 = A21 * i1, L31 = A31 * i1, L41 = A41 * i1, L51 = A51 * i1, L61 = A61 * i1;
  const double D2 = A22 - L21 * L21 * D1, i2 = frcp(D2);
  const double L32 = (A32 - L21 * L31 * D1) * i2;
  const double L42 = (A42 - L21 * L41 * D1) * i2;
  const double L52 = (A52 - L21 * L51 * D1) * i2;
  const double L62 = (A62 - L21 * L61 * D1) * i2;
  const double D3 = A33 - (L31 * L31 * D1 + L32 * L32 * D2), i3 = frcp(D3);
  const double L43 = (A43 - L31 * L41 * D1 - L32 * L42 * D2) * i3;
  const double L53 = (A53 - L31 * L51 * D1 - L32 * L52 * D2) * i3;
  const double L63 = (A63 - L31 * L61 * D1 - L32 * L62 * D2) * i3;
  const double D4 = A44 - (L41 * L41 * D1 + L42 * L42 * D2 + L43 * L43 * D3), i4 = frcp(D4);
  const double L54 = (A54 - L41 * L51 * D1 - L42 * L52 * D2 - L43 * L53 * D3) * i4;
  const double L64 = (A64 - L41 * L61 * D1 - L42 * L62 * D2 - L43 * L63 * D3) * i4;
  const double D5 = A55 - (L51 * L51 * D1 + L52 * L52 * D2 + L53 * L53 * D3 + L54 * L54 * D4), i5 = frcp(D5);
  const double L65 = (A65 - L51 * L61 * D1 - L52 * L62 * D2 - L53 * L63 * D3 - L54 * L64 * D4) * i5;
  const double D6 = A66 - (L61 * L61 * D1 + L62 * L62 * D2 + L63 * L63 * D3 + L64 * L64 * D4 + L65 * L65 * D5);
  f = Ldlt6{L21, L31, L41, L51, L61, L32, L42, L52, L62, L43, L53, L63, L54, L64, L65,
            i1, i2, i3, i4, i5, frcp(D6)};
}

__device__ __forceinline__ void ldlt6Solve(const Ldlt6 &f, const double *bLin, const double *bAng, double *xLin,
                                           double *xAng) {
#pragma clang fp contract(fast)
  const double y1 = bLin[0];
  const double y2 = bLin[1] - f.L21 * y1;
  const double y3 = bLin[2] - f.L31 * y1 - f.L32 * y2;
  const double y4 = bAng[0] - f.L41 * y1 - f.L42 * y2 - f.L43 * y3;
  const double y5 = bAng[1] - f.L51 * y1 - f.L52 * y2 - f.L53 * y3 - f.L54 * y4;
  const double y6 = bAng[2] - f.L61 * y1 - f.L62 * y2 - f.L63 * y3 - f.L64 * y4 - f.L65 * y5;
  xAng[2] = y6 * f.i6;
  xAng[1] = y5 * f.i5 - f.L65 * xAng[2];
  xAng[0] = y4 * f.i4 - f.L54 * xAng[1] - f.L64 * xAng[2];
  xLin[2] = y3 * f.i3 - f.L43 * xAng[0] - f.L53 * xAng[1] - f.L63 * xAng[2];
  xLin[1] = y2 * f.i2 - f.L32 * xLin[2] - f.L42 * xAng[0] - f.L52 * xAng[1] - f.L62 * xAng[2];
  xLin[0] = y1 * f.i1 - f.L21 * xLin[1] - f.L31 * xLin[2] - f.L41 * xAng[0] - f.L51 * xAng[1] - f.L61 * xAng[2];
}

// solve6 through that factorisation (the f32/f64 team modes; the exact mode
// keeps the reference's divisions).
__device__ __forceinline__ void solve6f(const double *aLin, const double *aAng, const double *aCross,
                                        const double *bLin, const double *bAng, double *xLin, double *xAng) {
  Ldlt6 f;
  ldlt6(aLin, aAng, aCross, f);
  ldlt6Solve(f, bLin, bAng, xLin, xAng);
}

// Eq. 21 quaternion update with the normalisation as one reciprocal.
__device__ __forceinline__ void qintf(double *o, const double *q, const double *dw) {
#pragma clang fp contract(fast)
  const double qx = q[0], qy = q[1], qz = q[2], qw = q[3];
  const double wx = dw[0], wy = dw[1], wz = dw[2];
  const double rx = qx + (wx * qw + wy * qz - wz * qy) * 0.5;
  const double ry = qy + (-wx * qz + wy * qw + wz * qx) * 0.5;
  const double rz = qz + (wx * qy - wy * qx + wz * qw) * 0.5;
  const double rw = qw + (-wx * qx - wy * qy - wz * qz) * 0.5;
  const double il = frcp(xsqrt(rx * rx + ry * ry + rz * rz + rw * rw));
  o[0] = rx * il; o[1] = ry * il; o[2] = rz * il; o[3] = rw * il;
}

// ------------------------------------------------------------ team reductions
// A 64-bit lane move under a DPP pattern (quad permutes and row mirrors stay
// inside a row of 16 lanes, so every source lane is one of the team's).
template <int CTRL> __device__ __forceinline__ double dppMove(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(unsigned(b)), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(unsigned(b >> 32)), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)unsigned(hi) << 32) | unsigned(lo)));
}

// Sum over the TEAM lanes of an aligned team; every lane gets the total. The
// pairs are those of an xor butterfly (lane ^ 1, ^ 2, ^ 4, ^ 8, ...): after
// the quad steps all lanes of a quad hold the quad's sum, so mirroring within
// 8 and then 16 lanes pairs each group with the other half.
template <int TEAM> __device__ __forceinline__ double teamSum(double v) {
  if constexpr (TEAM >= 2) v += dppMove<0xB1>(v);   // quad_perm [1,0,3,2]
  if constexpr (TEAM >= 4) v += dppMove<0x4E>(v);   // quad_perm [2,3,0,1]
  if constexpr (TEAM >= 8) v += dppMove<0x141>(v);  // row_half_mirror
  if constexpr (TEAM >= 16) v += dppMove<0x140>(v); // row_mirror
  if constexpr (TEAM >= 32) v += __shfl_xor(v, 16, TEAM);
  if constexpr (TEAM >= 64) v += __shfl_xor(v, 32, TEAM);
  return v;
}

// The same butterfly in f32: each step is one add with a DPP source operand
// (the compiler folds the lane move into the add), a third of the f64 form's
// instructions.
template <int CTRL> __device__ __forceinline__ float dppMoveF(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int TEAM> __device__ __forceinline__ float teamSumF(float v) {
  if constexpr (TEAM >= 2) v += dppMoveF<0xB1>(v);
  if constexpr (TEAM >= 4) v += dppMoveF<0x4E>(v);
  if constexpr (TEAM >= 8) v += dppMoveF<0x141>(v);
  if constexpr (TEAM >= 16) v += dppMoveF<0x140>(v);
  if constexpr (TEAM >= 32) v += __shfl_xor(v, 16, TEAM);
  if constexpr (TEAM >= 64) v += __shfl_xor(v, 32, TEAM);
  return v;
}

// Accumulated 6x6 system: lower triangles of the Lin and Ang blocks, the full
// Cross block, and the right-hand side (the 27 numbers solve6 reads).
// lin: 0:(0,0) 1:(1,0) 2:(1,1) 3:(2,0) 4:(2,1) 5:(2,2)  (row, col)
template <class A> struct Sys {
  A lin[6], ang[6], cross[9], rl[3], ra[3];
  __device__ void zero() {
    for (int i = 0; i < 6; i++) lin[i] = ang[i] = A(0);
    for (int i = 0; i < 9; i++) cross[i] = A(0);
    for (int i = 0; i < 3; i++) rl[i] = ra[i] = A(0);
  }
};
__device__ constexpr int LT_R[6] = {0, 1, 1, 2, 2, 2};
__device__ constexpr int LT_C[6] = {0, 0, 1, 0, 1, 2};

// One contact's contribution to the side `isA` (manifold.js:411-445):
//   Lin += Jl^T K Jl, Ang += Ja^T K Ja, Cross += Ja^T K Jl, rhs += J^T F.
// Entry (r, c) of Jx^T K Jy is sum_j (Jx[j][r] k_j) Jy[j][c] in the reference's
// association.
template <class R, class A>
__device__ __forceinline__ void stampContact(Sys<A> &s, const R *jl, const R *ja, const R *k, const R *F) {
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const int r = LT_R[i], c = LT_C[i];
    s.lin[i] += A((jl[r] * k[0]) * jl[c] + (jl[3 + r] * k[1]) * jl[3 + c] + (jl[6 + r] * k[2]) * jl[6 + c]);
    s.ang[i] += A((ja[r] * k[0]) * ja[c] + (ja[3 + r] * k[1]) * ja[3 + c] + (ja[6 + r] * k[2]) * ja[6 + c]);
  }
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      s.cross[r * 3 + c] +=
          A((ja[r] * k[0]) * jl[c] + (ja[3 + r] * k[1]) * jl[3 + c] + (ja[6 + r] * k[2]) * jl[6 + c]);
#pragma unroll
  for (int r = 0; r < 3; r++) {
    s.rl[r] += A(jl[r] * F[0] + jl[3 + r] * F[1] + jl[6 + r] * F[2]);
    s.ra[r] += A(ja[r] * F[0] + ja[3 + r] * F[1] + ja[6 + r] * F[2]);
  }
}

// ----------------------------------------------- fused contact arithmetic
// The team modes' contact evaluation and stamp (the exact mode keeps the
// reference's expression order above): the same quantities with fused
// multiply-adds, the basis signs folded into the products, and the cone
// scale from a reciprocal square root.
__device__ __forceinline__ float fmaR(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaR(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float rsqrtR(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrtR(double x) { return 1.0 / __builtin_sqrt(x); }

template <class R> struct FEval {
  R v[3];  // dA - dB + wA x rAw + rBw x wB: C_r = (1 - alpha) C0_r + B_r . v
  R C[3], F[3];
  bool within;
};

template <class R>
__device__ __forceinline__ void fcross(R *o, const R *a, const R *b) {
  o[0] = fmaR(a[1], b[2], -(a[2] * b[1]));
  o[1] = fmaR(a[2], b[0], -(a[0] * b[2]));
  o[2] = fmaR(a[0], b[1], -(a[1] * b[0]));
}

// C = (1 - alpha) C0 + J (x - x_t) with the cached angular rows folded into
// one vector: (rAw x B_r) . wA = B_r . (wA x rAw) and (B_r x rBw) . wB =
// B_r . (rBw x wB), so every row is B_r . v.
template <class R>
__device__ __forceinline__ void fevalC(const R *B, const R *hot, const R *dqA, const R *dqB, R alpha, FEval<R> &e,
                                       int law = LAW_CPU) {
  R wa[3], wb[3];
  fcross(wa, dqA + 3, hot + HRA);
  fcross(wb, hot + HRB, dqB + 3);
#pragma unroll
  for (int k = 0; k < 3; k++) e.v[k] = (dqA[k] - dqB[k]) + (wa[k] + wb[k]);
  fevalCAlpha(B, hot, alpha, e, law);
}

// C from the stored v under another alpha.
template <class R>
__device__ __forceinline__ void fevalCAlpha(const R *B, const R *hot, R alpha, FEval<R> &e, int law = LAW_CPU) {
  const R oma = R(1) - alpha;
  R base[3] = {oma * hot[HC0], oma * hot[HC0 + 1], oma * hot[HC0 + 2]};
  if (law != LAW_CPU) {
    if (B[BCD] > R(0) && hot[HC0] > R(0)) base[0] = hot[HC0];
    if (law == LAW_GPU) base[0] = base[0] - hot[HB];
  }
#pragma unroll
  for (int r = 0; r < 3; r++)
    e.C[r] = fmaR(B[r * 3], e.v[0], fmaR(B[r * 3 + 1], e.v[1], fmaR(B[r * 3 + 2], e.v[2], base[r])));
}

template <class R>
__device__ __forceinline__ void fevalF(const R *k, const R *lam, R mu, FEval<R> &e, int law = LAW_CPU,
                                       R muD = R(0), int stick = 0) {
  const R z0 = fmaR(k[0], e.C[0], lam[0]);
  const R z1 = fmaR(k[1], e.C[1], lam[1]);
  const R z2 = fmaR(k[2], e.C[2], lam[2]);
  const R f0 = z0 < R(0) ? z0 : R(0);
  if (law != LAW_CPU) {
    // the GPU law: static cone while sticking, a force outside it scaled
    // onto the dynamic cone
    const R active = -f0 * (stick ? mu : muD), dyn = -f0 * muD;
    const R fs2 = fmaR(z1, z1, z2 * z2);
    e.F[0] = f0;
    e.within = fs2 <= active * active;
    const R tiny = sizeof(R) == 4 ? R(1.1754944e-38f) : R(2.2250738585072014e-308);
    const R scale = e.within ? R(1) : (dyn > R(0) ? dyn * rsqrtR(fs2 > tiny ? fs2 : tiny) : R(0));
    e.F[1] = z1 * scale;
    e.F[2] = z2 * scale;
    return;
  }
  const R bounds = -f0 * mu;
  const R fs2 = fmaR(z1, z1, z2 * z2);
  e.F[0] = f0;
  e.within = fs2 <= bounds * bounds;
  // outside the cone: scale by bounds / |z_t| (zero when the normal force is
  // zero); the square is kept above the smallest normal number so that the
  // hardware reciprocal square root never sees a flushed denormal
  const R tiny = sizeof(R) == 4 ? R(1.1754944e-38f) : R(2.2250738585072014e-308);
  const R scale = e.within ? R(1) : (bounds > R(0) ? bounds * rsqrtR(fs2 > tiny ? fs2 : tiny) : R(0));
  e.F[1] = z1 * scale;
  e.F[2] = z2 * scale;
}

// Adds one contact to the system of one of its bodies: linear rows s B_r
// (s = +1 for body A, -1 for body B) and angular rows s (rW x B_r) with rW
// that body's world anchor. With L = sum_r k_r B_r B_r^T and g = B^T F the
// blocks are
//   Lin += L,  Cross += [rW]x L,  Ang += [rW]x L [rW]x^T,
//   rl += s g,  ra += s (rW x g),
// the same sums as J^T K J and J^T F, with fewer products.
template <class R, class A>
__device__ __forceinline__ void fstamp(Sys<A> &sy, const R *B, R s, const R *rW, const R *k, const R *F) {
  R kB[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int i = 0; i < 3; i++) kB[r * 3 + i] = k[r] * B[r * 3 + i];
  R L[9];  // full symmetric
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int i = LT_R[q], j = LT_C[q];
    const R l = fmaR(kB[i], B[j], fmaR(kB[3 + i], B[3 + j], kB[6 + i] * B[6 + j]));
    L[i * 3 + j] = l;
    L[j * 3 + i] = l;
    sy.lin[q] = sy.lin[q] + A(l);
  }
  // M = [rW]x L: column j is rW x L_j
  R M[9];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    M[0 * 3 + j] = fmaR(rW[1], L[2 * 3 + j], -(rW[2] * L[1 * 3 + j]));
    M[1 * 3 + j] = fmaR(rW[2], L[0 * 3 + j], -(rW[0] * L[2 * 3 + j]));
    M[2 * 3 + j] = fmaR(rW[0], L[1 * 3 + j], -(rW[1] * L[0 * 3 + j]));
  }
#pragma unroll
  for (int q = 0; q < 9; q++) sy.cross[q] = sy.cross[q] + A(M[q]);
  // Ang_ij = (rW x M_row_i)_j
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int i = LT_R[q], j = LT_C[q];
    const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    sy.ang[q] = sy.ang[q] + A(fmaR(rW[j1], M[i * 3 + j2], -(rW[j2] * M[i * 3 + j1])));
  }
  R g[3], rg[3];
#pragma unroll
  for (int i = 0; i < 3; i++) g[i] = fmaR(F[0], B[i], fmaR(F[1], B[3 + i], F[2] * B[6 + i]));
  fcross(rg, rW, g);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    sy.rl[i] = fmaR(s, g[i], sy.rl[i]);
    sy.ra[i] = fmaR(s, rg[i], sy.ra[i]);
  }
}

// The dual update of a contact (dualUpdate) for the fused evaluation.
template <class R>
__device__ __forceinline__ int fdualUpdate(const R *k, const FEval<R> &e, R beta, R pmax, R stickT, bool hasSphere,
                                           R *kl, int law = LAW_CPU) {
  kl[0] = k[0]; kl[1] = k[1]; kl[2] = k[2];
  kl[3] = e.F[0]; kl[4] = e.F[1]; kl[5] = e.F[2];
  if (e.F[0] < R(0)) kl[0] = jmin(fmaR(beta, xabs(e.C[0]), k[0]), pmax);
  if (law != LAW_CPU) {
    if (e.within) {
      kl[1] = jmin(fmaR(beta, xabs(e.C[1]), k[1]), pmax);
      kl[2] = jmin(fmaR(beta, xabs(e.C[2]), k[2]), pmax);
    }
    return e.within ? 1 : 0;
  }
  int stick = -1;
  if (e.within) {
    kl[1] = jmin(fmaR(beta, xabs(e.C[1]), k[1]), pmax);
    kl[2] = jmin(fmaR(beta, xabs(e.C[2]), k[2]), pmax);
    stick = fmaR(e.C[1], e.C[1], e.C[2] * e.C[2]) < stickT * stickT ? 1 : 0;
  } else if (hasSphere) {
    stick = 0;
  }
  return stick;
}

}  // namespace avbd
