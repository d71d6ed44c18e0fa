// avbd_adjoint.hpp — reverse-mode derivatives of the AVBD step (f64).
//
// The forward step is recorded per step (poses after every sweep, contact
// k/lambda after every dual update, the step's contacts and colouring) and
// reversed operation by operation: velocity update, then per iteration the
// dual update and the colour sweeps in reverse order, then body
// initialisation and force initialisation (solver.js:366-489,
// manifold.js:214-491). The derivative is that of the computed discrete
// trajectory (a fixed number of iterations, not the converged solution).
//
// Contact geometry (basis, feature keys, local anchors) is piecewise constant
// in the poses and is not differentiated, with one exception that matters for
// sliding: a contact whose anchors were re-derived this step has its points on
// the contact normal through each other, so its tangential C0 is zero for every
// pose and only the normal row of C0 carries a derivative. Anchors kept from a
// sticking contact are constants and C0 is differentiated in full.
//
// Layouts: pose adjoints use the pose layout (p[3], q[4], one pad); contact
// adjoints use the hot-record order (k[3], lambda[3]).
#pragma once

#include "avbd_kernels.hpp"

namespace avbd {

// ------------------------------------------------- VJPs of the pose operators
// qsub (Eq. 20): o = 2 vec(a (x) b^-1), b^-1 = conj(b) / |b|^2.
__device__ inline void qsubVjp(const double *a, const double *b, const double *ob, double *ab, double *bb) {
  const double lsq = b[0] * b[0] + b[1] * b[1] + b[2] * b[2] + b[3] * b[3];
  const double c[4] = {-b[0], -b[1], -b[2], b[3]};
  const double il = frcp(lsq);  // the adjoint needs no division-exact rounding
  const double ix = c[0] * il, iy = c[1] * il, iz = c[2] * il, iw = c[3] * il;
  const double o0 = 2 * ob[0], o1 = 2 * ob[1], o2 = 2 * ob[2];
  ab[0] += o0 * iw - o1 * iz + o2 * iy;
  ab[1] += o0 * iz + o1 * iw - o2 * ix;
  ab[2] += -o0 * iy + o1 * ix + o2 * iw;
  ab[3] += o0 * ix + o1 * iy + o2 * iz;
  const double ax = a[0], ay = a[1], az = a[2], aw = a[3];
  const double ib[4] = {o0 * aw + o1 * az - o2 * ay, -o0 * az + o1 * aw + o2 * ax, o0 * ay - o1 * ax + o2 * aw,
                        o0 * ax + o1 * ay + o2 * az};
  const double s[4] = {-1, -1, -1, 1};
  const double ci = c[0] * ib[0] + c[1] * ib[1] + c[2] * ib[2] + c[3] * ib[3];
  const double il2 = il * il;
  for (int j = 0; j < 4; j++) bb[j] += s[j] * ib[j] * il - 2 * b[j] * ci * il2;
}

// qint (Eq. 21): o = r / |r|, r = q + 1/2 (w, 0) (x) q.
__device__ inline void qintVjp(const double *q, const double *w, const double *ob, double *qb, double *wb) {
  const double qx = q[0], qy = q[1], qz = q[2], qw = q[3];
  const double wx = w[0], wy = w[1], wz = w[2];
  const double r[4] = {qx + 0.5 * (wx * qw + wy * qz - wz * qy), qy + 0.5 * (-wx * qz + wy * qw + wz * qx),
                       qz + 0.5 * (wx * qy - wy * qx + wz * qw), qw + 0.5 * (-wx * qx - wy * qy - wz * qz)};
  const double len = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
  const double il = frcp(len);
  const double o[4] = {r[0] * il, r[1] * il, r[2] * il, r[3] * il};
  const double od = o[0] * ob[0] + o[1] * ob[1] + o[2] * ob[2] + o[3] * ob[3];
  double rb[4];
  for (int k = 0; k < 4; k++) rb[k] = (ob[k] - o[k] * od) * il;
  wb[0] += 0.5 * (rb[0] * qw - rb[1] * qz + rb[2] * qy - rb[3] * qx);
  wb[1] += 0.5 * (rb[0] * qz + rb[1] * qw - rb[2] * qx - rb[3] * qy);
  wb[2] += 0.5 * (-rb[0] * qy + rb[1] * qx + rb[2] * qw - rb[3] * qz);
  qb[0] += rb[0] + 0.5 * (rb[1] * wz - rb[2] * wy - rb[3] * wx);
  qb[1] += rb[1] + 0.5 * (-rb[0] * wz + rb[2] * wx - rb[3] * wy);
  qb[2] += rb[2] + 0.5 * (rb[0] * wy - rb[1] * wx - rb[3] * wz);
  qb[3] += rb[3] + 0.5 * (rb[0] * wx + rb[1] * wy + rb[2] * wz);
}

// qrot: o = v + w t + u x t, t = 2 u x v; the adjoint of q only (v fixed).
__device__ inline void qrotVjp(const double *q, const double *v, const double *ob, double *qb) {
  const double u[3] = {q[0], q[1], q[2]}, w = q[3];
  double t[3], tb[3], ub[3];
  cross3(t, u, v);
  for (int k = 0; k < 3; k++) t[k] *= 2;
  // o = v + w t + u x t
  qb[3] += ob[0] * t[0] + ob[1] * t[1] + ob[2] * t[2];
  cross3(ub, t, ob);  // d(u x t) / du
  cross3(tb, ob, u);  // d(u x t) / dt
  for (int k = 0; k < 3; k++) tb[k] += w * ob[k];
  // t = 2 u x v
  double e[3];
  cross3(e, v, tb);
  for (int k = 0; k < 3; k++) ub[k] += 2 * e[k];
  for (int k = 0; k < 3; k++) qb[k] += ub[k];
}

// worldMoment: M = Rt D R with R = fromQuat(q); Mb (3x3) -> q adjoint.
__device__ inline void worldMomentVjp(const double *q, double m0, double m1, double m2, const double *Mb,
                                      double *qb) {
  double R[9];
  fromQuat(R, q);
  const double D[3] = {m0, m1, m2};
  double Rb[9];
  // Rb = D R (Mb + Mb^T)
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double acc = 0;
      for (int k = 0; k < 3; k++) acc += R[i * 3 + k] * (Mb[k * 3 + j] + Mb[j * 3 + k]);
      Rb[i * 3 + j] = D[i] * acc;
    }
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  qb[0] += 2 * y * Rb[1] + 2 * z * Rb[2] + 2 * y * Rb[3] - 4 * x * Rb[4] + 2 * w * Rb[5] + 2 * z * Rb[6] -
           2 * w * Rb[7] - 4 * x * Rb[8];
  qb[1] += -4 * y * Rb[0] + 2 * x * Rb[1] - 2 * w * Rb[2] + 2 * x * Rb[3] + 2 * z * Rb[5] + 2 * w * Rb[6] +
           2 * z * Rb[7] - 4 * y * Rb[8];
  qb[2] += -4 * z * Rb[0] + 2 * w * Rb[1] + 2 * x * Rb[2] - 2 * w * Rb[3] - 4 * z * Rb[4] + 2 * y * Rb[5] +
           2 * x * Rb[6] + 2 * y * Rb[7];
  qb[3] += 2 * z * Rb[1] - 2 * y * Rb[2] - 2 * z * Rb[3] + 2 * x * Rb[5] + 2 * y * Rb[6] - 2 * x * Rb[7];
}

// ------------------------------------------------------------ contact pieces
// A contact re-evaluated from its record: C (cached Jacobian at x_t) and the
// cone-projected force.
struct CEval {
  double jA[9], jB[9], dqA[6], dqB[6], C[3];
  double z[3], F[3], bounds, fs;
  bool clamped, within;
};

__device__ inline void bodyDelta(const double *x, const double *x0, double *dq) {
  for (int k = 0; k < 3; k++) dq[k] = x[k] - x0[k];
  qsub(dq + 3, x + 3, x0 + 3);
}

// With the body deltas given (dqA, dqB: 6 each).
__device__ inline void cEvalDq(const double *B, double mu, const double *hot, const double *k, const double *lam,
                               const double *dqA, const double *dqB, double alpha, CEval &e) {
  for (int j = 0; j < 6; j++) { e.dqA[j] = dqA[j]; e.dqB[j] = dqB[j]; }
  for (int r = 0; r < 3; r++) {
    const double nb[3] = {-B[r * 3], -B[r * 3 + 1], -B[r * 3 + 2]};
    cross3(e.jA + r * 3, hot + HRA, B + r * 3);
    cross3(e.jB + r * 3, hot + HRB, nb);
  }
  for (int r = 0; r < 3; r++) {
    double c = hot[HC0 + r] * (1 - alpha);
    for (int j = 0; j < 3; j++)
      c += B[r * 3 + j] * e.dqA[j] - B[r * 3 + j] * e.dqB[j] + e.jA[r * 3 + j] * e.dqA[3 + j] +
           e.jB[r * 3 + j] * e.dqB[3 + j];
    e.C[r] = c;
  }
  for (int r = 0; r < 3; r++) e.z[r] = k[r] * e.C[r] + lam[r];
  e.F[0] = jmin(e.z[0], 0.0);
  e.bounds = xabs(e.F[0]) * mu;
  e.fs = sqrt(e.z[1] * e.z[1] + e.z[2] * e.z[2]);
  e.clamped = e.fs > e.bounds && e.fs > 0;
  const double sc = e.bounds * frcp(e.fs);  // used only where clamped (fs > 0)
  e.F[1] = e.clamped ? e.z[1] * sc : e.z[1];
  e.F[2] = e.clamped ? e.z[2] * sc : e.z[2];
  e.within = e.fs <= e.bounds;
}

// The deltas of a contact's bodies A and B from those of body i and of the
// other body o (element selects: no pointer into either array is chosen).
__device__ inline void orderDeltas(bool isA, const double *dqI, const double *dqO, double *dqA, double *dqB) {
  for (int j = 0; j < 6; j++) {
    dqA[j] = isA ? dqI[j] : dqO[j];
    dqB[j] = isA ? dqO[j] : dqI[j];
  }
}

// Cone projection adjoint: Fb -> zb, mub (added).
__device__ inline void coneVjp(const CEval &e, double mu, const double *Fb, double *zb, double &mub) {
  double bb = 0;
  if (e.clamped) {
    const double ifs = frcp(e.fs);
    const double s = e.bounds * ifs, s3 = s * ifs * ifs;
    const double fz = Fb[1] * e.z[1] + Fb[2] * e.z[2];
    zb[1] = s * Fb[1] - s3 * e.z[1] * fz;
    zb[2] = s * Fb[2] - s3 * e.z[2] * fz;
    bb = fz * ifs;
  } else {
    zb[1] = Fb[1];
    zb[2] = Fb[2];
  }
  // bounds = |F0| mu with F0 = min(z0, 0) <= 0
  double F0b = Fb[0];
  if (e.F[0] < 0) {
    F0b += -mu * bb;
    mub += bb * -e.F[0];
  }
  zb[0] = e.z[0] < 0 ? F0b : 0.0;
}

// C adjoint: Cb -> dqA, dqB, C0 and the world anchors rAw, rBw (all added).
__device__ inline void cVjp(const double *B, const CEval &e, const double *Cb, double alpha, double *dqAb,
                            double *dqBb, double *C0b, double *rAb, double *rBb) {
  for (int r = 0; r < 3; r++) C0b[r] += (1 - alpha) * Cb[r];
  for (int j = 0; j < 3; j++) {
    double l = 0, a = 0, b = 0;
    for (int r = 0; r < 3; r++) {
      l += B[r * 3 + j] * Cb[r];
      a += e.jA[r * 3 + j] * Cb[r];
      b += e.jB[r * 3 + j] * Cb[r];
    }
    dqAb[j] += l;
    dqBb[j] -= l;
    dqAb[3 + j] += a;
    dqBb[3 + j] += b;
  }
  for (int r = 0; r < 3; r++) {
    const double ja[3] = {Cb[r] * e.dqA[3], Cb[r] * e.dqA[4], Cb[r] * e.dqA[5]};
    const double jb[3] = {Cb[r] * e.dqB[3], Cb[r] * e.dqB[4], Cb[r] * e.dqB[5]};
    const double nb[3] = {-B[r * 3], -B[r * 3 + 1], -B[r * 3 + 2]};
    double o[3];
    cross3(o, B + r * 3, ja);  // jA_r = rAw x B_r
    for (int k = 0; k < 3; k++) rAb[k] += o[k];
    cross3(o, nb, jb);  // jB_r = rBw x (-B_r)
    for (int k = 0; k < 3; k++) rBb[k] += o[k];
  }
}

// Adjoint of a body-delta dq = (x - x0, qsub(q, q0)) into the pose adjoint xb
// and the step-start adjoint x0b, atomically (other bodies share them).
__device__ inline void deltaVjpAtomic(const double *x, const double *x0, const double *dqb, double *xb, double *x0b) {
  double qb[4] = {0, 0, 0, 0}, q0b[4] = {0, 0, 0, 0};
  qsubVjp(x + 3, x0 + 3, dqb + 3, qb, q0b);
  for (int k = 0; k < 3; k++) {
    atomicAdd(xb + k, dqb[k]);
    atomicAdd(x0b + k, -dqb[k]);
  }
  for (int k = 0; k < 4; k++) {
    atomicAdd(xb + 3 + k, qb[k]);
    atomicAdd(x0b + 3 + k, q0b[k]);
  }
}

// One recorded step, as the reverse kernels read it.
struct AdjStep {
  int nLive, numColors;
  const int *live;     // live manifolds
  const int4 *head;    // bodies, contact count, flags
  const double *basis; // per manifold: basis rows, mu
  const double *anchor;
  const int *src;      // per contact: matched old slot | kept-anchor bit, or -1
  const double *hot;   // snapshots l = 0..I of the hot records
  const double *X;     // snapshots l = 0..I of the poses (X^0 after init)
  const double *X0, *Y, *V;
  const int *color;
  const int *cstart;
  const int4 *cwork;
  const uint2 *items;
  size_t hotStride, xStride;  // doubles per snapshot
};

struct AdjBufs {
  double *Xb, *X0b, *Yb, *Vb, *VbS;  // poses, step-start poses, inertial targets, end / start velocities
  double *LK, *LKprev;               // contact k/lambda adjoints (this step's slots, the previous step's)
  double *C0b, *RWb, *MUb;           // per contact C0 and world anchors; per manifold mu
  double *muGrad;                    // per body
};

constexpr int SRC_KEPT = 0x40000000;

// ---------------------------------------------------------- reverse kernels
// BDF1 velocities (solver.js:476-488) at the step's end.
__device__ inline void adjVelocity(int i, const BodyProp *props, const AdjStep &T, const AdjBufs &A, int iters,
                                   double dt) {
  double *vb = A.Vb + size_t(i) * VS, *vs = A.VbS + size_t(i) * VS;
  for (int k = 0; k < 12; k++) vs[k] = 0;
  if (!(props[i].mass > 0)) {
    // velocities unchanged, previous velocity <- velocity
    for (int k = 0; k < 6; k++) vs[k] = vb[k];
    for (int k = 0; k < 3; k++) vs[k] += vb[6 + k];
    return;
  }
  const double *x = T.X + T.xStride * iters + size_t(i) * XS, *x0 = T.X0 + size_t(i) * XS;
  double *xb = A.Xb + size_t(i) * XS, *x0b = A.X0b + size_t(i) * XS;
  for (int k = 0; k < 3; k++) {
    xb[k] += vb[k] / dt;
    x0b[k] -= vb[k] / dt;
  }
  const double wb[3] = {vb[3] / dt, vb[4] / dt, vb[5] / dt};
  double qb[4] = {0, 0, 0, 0}, q0b[4] = {0, 0, 0, 0};
  qsubVjp(x + 3, x0 + 3, wb, qb, q0b);
  for (int k = 0; k < 4; k++) {
    xb[3 + k] += qb[k];
    x0b[3 + k] += q0b[k];
  }
  for (int k = 0; k < 3; k++) vs[k] = vb[6 + k];  // prevVelocity <- the step-start velocity
}

// The dual update of iteration l (manifold.js:447-491) for one contact slot si
// at the poses after the iteration's sweep: the contact adjoints LK turn from
// those of k, lambda after the update into those before it; the body deltas'
// adjoints dqAb, dqBb and mub are added to.
__device__ inline void adjDualContact(const double *B, size_t si, const double *hot, const AdjBufs &A,
                                      const double *dqA, const double *dqB, const StepParams &P, double *dqAb,
                                      double *dqBb, double &mub) {
  // the slot's adjoints are read first and written last: no load waits on a
  // store to memory it might alias
  double *lk = A.LK + si * 6, *rwb = A.RWb + si * 6, *c0b = A.C0b + si * 3;
  double lkv[6], rwv[6], c0v[3];
  for (int k = 0; k < 6; k++) { lkv[k] = lk[k]; rwv[k] = rwb[k]; }
  for (int k = 0; k < 3; k++) c0v[k] = c0b[k];
  CEval e;
  cEvalDq(B, B[9], hot, hot + HK, hot + HL, dqA, dqB, P.alpha, e);
  double kb[3] = {lkv[0], lkv[1], lkv[2]}, Cb[3] = {0, 0, 0};
  const double Fb[3] = {lkv[3], lkv[4], lkv[5]};
  // k' = min(k + beta |C|, kmax) where the update applies (normal: pushing;
  // tangents: inside the cone)
  for (int r = 0; r < 3; r++) {
    const bool applies = r == 0 ? e.F[0] < 0 : e.within;
    if (!applies) continue;
    const double raw = hot[HK + r] + P.betaLin * xabs(e.C[r]);
    if (raw < P.pmax) {
      Cb[r] += kb[r] * P.betaLin * sgn(e.C[r]);
    } else {
      kb[r] = 0;
    }
  }
  double zb[3];
  coneVjp(e, B[9], Fb, zb, mub);
  for (int r = 0; r < 3; r++) {
    Cb[r] += hot[HK + r] * zb[r];
    kb[r] += zb[r] * e.C[r];
  }
  cVjp(B, e, Cb, P.alpha, dqAb, dqBb, c0v, rwv, rwv + 3);
  lk[0] = kb[0]; lk[1] = kb[1]; lk[2] = kb[2];
  lk[3] = zb[0]; lk[4] = zb[1]; lk[5] = zb[2];
  for (int k = 0; k < 6; k++) rwb[k] = rwv[k];
  for (int k = 0; k < 3; k++) c0b[k] = c0v[k];
}

// The primal update of one colour in iteration l (solver.js:419-458), a team
// of TEAM lanes per body: the pose adjoint of each body turns from that after
// its update into that before it; the neighbours, the contacts and the step's
// start and inertial target receive their shares.
//
// The team also applies, first, the dual adjoint of iteration l for the
// contacts of which its body is the reverse sweep's first visitor (the other
// body has a lower colour or is static). The forward's dual of iteration l
// follows the whole sweep, so its adjoint must precede the primal adjoints of
// both bodies of iteration l, and follow those of iteration l + 1: the first
// visitor's team is such a place. Each contact is handled by the lane that
// holds its item, in both parts, so that its LK slot is read and written in
// program order; the body's own share stays in registers.
template <int TEAM>
__device__ inline void adjPrimalTeam(int c, int team, int lane, const AdjStep &T, const AdjBufs &A,
                                     const BodyProp *props, int l, const StepParams &P) {
  const int c0 = T.cstart[c];
  const double dt2 = P.dt * P.dt, idt2 = 1.0 / dt2;
  const double *Xbef = T.X + T.xStride * l, *Xaft = T.X + T.xStride * (l + 1);
  const double *hotL = T.hot + T.hotStride * l;
  {
    const int4 w4 = T.cwork[c0 + team];
    const int i = w4.x, b0 = w4.y, b1 = w4.z;
    const BodyProp &bp = props[i];
    const double *x = Xbef + size_t(i) * XS, *x0 = T.X0 + size_t(i) * XS, *y = T.Y + size_t(i) * XS;
    // pose of a neighbour as body i saw it: updated already in this sweep if
    // its colour comes first
    auto poseOf = [&](int j) -> const double * {
      const int cj = T.color[j];
      return (cj >= 0 && cj < c ? Xaft : Xbef) + size_t(j) * XS;
    };
    // 0. the dual update's adjoint at the poses after the sweep
    double dqDual[6] = {0, 0, 0, 0, 0, 0};
    {
      const double *xa = Xaft + size_t(i) * XS;
      for (int t = b0 + lane; t < b1; t += TEAM) {
        const uint2 iw = T.items[t];
        const int o = int(iw.y & ITEM_BODY);
        if (T.color[o] > c) continue;  // the other body's team came first
        const int m = int(iw.x >> 4);
        const size_t si = size_t(m) * MAXC + ((iw.x >> 1) & 7u);
        const bool isA = iw.x & 1u;
        const bool oDyn = props[o].mass > 0;
        double xo[7], x0o[7];
        for (int k = 0; k < 7; k++) { xo[k] = Xaft[size_t(o) * XS + k]; x0o[k] = T.X0[size_t(o) * XS + k]; }
        double dqI[6], dqO[6], dqA[6], dqB[6];
        bodyDelta(xa, x0, dqI);
        bodyDelta(xo, x0o, dqO);
        orderDeltas(isA, dqI, dqO, dqA, dqB);
        double dqAb[6] = {0, 0, 0, 0, 0, 0}, dqBb[6] = {0, 0, 0, 0, 0, 0}, mub = 0;
        adjDualContact(T.basis + size_t(m) * BAS, si, hotL + si * HOT, A, dqA, dqB, P, dqAb, dqBb, mub);
        double dqo[6];
        for (int k = 0; k < 6; k++) {
          dqDual[k] += isA ? dqAb[k] : dqBb[k];
          dqo[k] = isA ? dqBb[k] : dqAb[k];
        }
        atomicAdd(A.MUb + m, mub);
        if (oDyn) deltaVjpAtomic(xo, x0o, dqo, A.Xb + size_t(o) * XS, A.X0b + size_t(o) * XS);
      }
      if constexpr (TEAM > 1)
        for (int k = 0; k < 6; k++) dqDual[k] = teamSum<TEAM>(dqDual[k]);
    }
    // 1. the system H, g as the forward built it
    Sys<double> s;
    s.zero();
    for (int t = b0 + lane; t < b1; t += TEAM) {
      const uint2 iw = T.items[t];
      const int m = int(iw.x >> 4);
      const size_t si = size_t(m) * MAXC + ((iw.x >> 1) & 7u);
      const bool isA = iw.x & 1u;
      const int o = int(iw.y & ITEM_BODY);
      const double *B = T.basis + size_t(m) * BAS;
      const double *hot = hotL + si * HOT;
      double dqI[6], dqO[6], dqA[6], dqB[6];
      bodyDelta(x, x0, dqI);
      bodyDelta(poseOf(o), T.X0 + size_t(o) * XS, dqO);
      orderDeltas(isA, dqI, dqO, dqA, dqB);
      CEval e;
      cEvalDq(B, B[9], hot, hot + HK, hot + HL, dqA, dqB, P.alpha, e);
      double jl[9], ja[9];
      for (int k = 0; k < 9; k++) {
        jl[k] = isA ? B[k] : -B[k];
        ja[k] = isA ? e.jA[k] : e.jB[k];
      }
      stampContact<double, double>(s, jl, ja, hot + HK, e.F);
    }
    if constexpr (TEAM > 1) {
      for (int k = 0; k < 6; k++) { s.lin[k] = teamSum<TEAM>(s.lin[k]); s.ang[k] = teamSum<TEAM>(s.ang[k]); }
      for (int k = 0; k < 9; k++) s.cross[k] = teamSum<TEAM>(s.cross[k]);
      for (int k = 0; k < 3; k++) { s.rl[k] = teamSum<TEAM>(s.rl[k]); s.ra[k] = teamSum<TEAM>(s.ra[k]); }
    }
    double M[9], ei[3];
    worldMoment(M, x + 3, bp.m0, bp.m1, bp.m2, !P.rotated || bp.shape == 1);
    const double ml = bp.mass * idt2;
    qsub(ei, x + 3, y + 3);
    s.lin[0] += ml; s.lin[2] += ml; s.lin[5] += ml;
    s.ang[0] += M[0] * idt2; s.ang[1] += M[3] * idt2; s.ang[2] += M[4] * idt2;
    s.ang[3] += M[6] * idt2; s.ang[4] += M[7] * idt2; s.ang[5] += M[8] * idt2;
    for (int k = 0; k < 3; k++) {
      s.rl[k] += ml * (x[k] - y[k]);
      s.ra[k] += (M[k * 3] * ei[0] + M[k * 3 + 1] * ei[1] + M[k * 3 + 2] * ei[2]) * idt2;
    }
    const double aLin[9] = {s.lin[0], 0, 0, s.lin[1], s.lin[2], 0, s.lin[3], s.lin[4], s.lin[5]};
    const double aAng[9] = {s.ang[0], 0, 0, s.ang[1], s.ang[2], 0, s.ang[3], s.ang[4], s.ang[5]};
    double Dl[3], Da[3];
    Ldlt6 H;  // factored once for the update and its transpose solve
    ldlt6(aLin, aAng, s.cross, H);
    {
      const double bl[3] = {-s.rl[0], -s.rl[1], -s.rl[2]}, ba[3] = {-s.ra[0], -s.ra[1], -s.ra[2]};
      ldlt6Solve(H, bl, ba, Dl, Da);
    }
    const double Dv[6] = {Dl[0], Dl[1], Dl[2], Da[0], Da[1], Da[2]};
    // 2. the update x' = (p + D_lin, qint(q, D_ang)) and the solve
    const double *xbv = A.Xb + size_t(i) * XS;
    double pb1[3] = {xbv[0], xbv[1], xbv[2]}, qb1[4] = {xbv[3], xbv[4], xbv[5], xbv[6]};
    double q0Dual[4] = {0, 0, 0, 0};
    {
      // the body's own share of its dual updates, as deltaVjpAtomic would add it
      double qbD[4] = {0, 0, 0, 0};
      qsubVjp(Xaft + size_t(i) * XS + 3, x0 + 3, dqDual + 3, qbD, q0Dual);
      for (int k = 0; k < 3; k++) pb1[k] += dqDual[k];
      for (int k = 0; k < 4; k++) qb1[k] += qbD[k];
    }
    double qb[4] = {0, 0, 0, 0}, Dab[3] = {0, 0, 0};
    qintVjp(x + 3, Da, qb1, qb, Dab);
    double wl[3], wa[3];
    {
      const double bl[3] = {pb1[0], pb1[1], pb1[2]};
      ldlt6Solve(H, bl, Dab, wl, wa);
    }
    const double w[6] = {wl[0], wl[1], wl[2], wa[0], wa[1], wa[2]};  // H^-1 Db; gb = -w, Hb = -w D^T
    // 3. the contacts' shares (each item's inputs read first, its outputs
    // written last)
    double dqSelf[6] = {0, 0, 0, 0, 0, 0};
    for (int t = b0 + lane; t < b1; t += TEAM) {
      const uint2 iw = T.items[t];
      const int m = int(iw.x >> 4);
      const size_t si = size_t(m) * MAXC + ((iw.x >> 1) & 7u);
      const bool isA = iw.x & 1u;
      const int o = int(iw.y & ITEM_BODY);
      const bool oDyn = props[o].mass > 0;
      const double *B = T.basis + size_t(m) * BAS;
      const double *hot = hotL + si * HOT;
      const double *xoP = poseOf(o), *x0oP = T.X0 + size_t(o) * XS;
      double xo[7], x0o[7];
      for (int k = 0; k < 7; k++) { xo[k] = xoP[k]; x0o[k] = x0oP[k]; }
      double *lk = A.LK + si * 6, *rwb = A.RWb + si * 6, *c0b = A.C0b + si * 3;
      double lkv[6], rwv[6], c0v[3];
      for (int k = 0; k < 6; k++) { lkv[k] = lk[k]; rwv[k] = rwb[k]; }
      for (int k = 0; k < 3; k++) c0v[k] = c0b[k];
      double dqI[6], dqO[6], dqA[6], dqB[6];
      bodyDelta(x, x0, dqI);
      bodyDelta(xo, x0o, dqO);
      orderDeltas(isA, dqI, dqO, dqA, dqB);
      CEval e;
      cEvalDq(B, B[9], hot, hot + HK, hot + HL, dqA, dqB, P.alpha, e);
      double Fb[3], Cb[3], zb[3], mub = 0;
      for (int r = 0; r < 3; r++) {
        double jl[3], ja[3];
        for (int k = 0; k < 3; k++) {
          jl[k] = isA ? B[r * 3 + k] : -B[r * 3 + k];
          ja[k] = isA ? e.jA[r * 3 + k] : e.jB[r * 3 + k];
        }
        const double Jw = jl[0] * w[0] + jl[1] * w[1] + jl[2] * w[2] + ja[0] * w[3] + ja[1] * w[4] + ja[2] * w[5];
        const double JD = jl[0] * Dv[0] + jl[1] * Dv[1] + jl[2] * Dv[2] + ja[0] * Dv[3] + ja[1] * Dv[4] +
                          ja[2] * Dv[5];
        Fb[r] = -Jw;
        lkv[r] += -Jw * JD;  // k from H = J^T K J
        // angular row adjoint: F_r gb + k_r J_r (Hb + Hb^T)
        double jab[3];
        for (int k = 0; k < 3; k++) jab[k] = -e.F[r] * w[3 + k] - hot[HK + r] * (Jw * Dv[3 + k] + JD * w[3 + k]);
        double o3[3];
        cross3(o3, jl, jab);  // Ja_r = rW x Jl_r
        for (int k = 0; k < 3; k++) {
          rwv[k] += isA ? o3[k] : 0.0;
          rwv[3 + k] += isA ? 0.0 : o3[k];
        }
      }
      coneVjp(e, B[9], Fb, zb, mub);
      for (int r = 0; r < 3; r++) {
        Cb[r] = hot[HK + r] * zb[r];
        lkv[r] += zb[r] * e.C[r];
        lkv[3 + r] += zb[r];
      }
      double dqAb[6] = {0, 0, 0, 0, 0, 0}, dqBb[6] = {0, 0, 0, 0, 0, 0};
      cVjp(B, e, Cb, P.alpha, dqAb, dqBb, c0v, rwv, rwv + 3);
      double dqo[6];
      for (int k = 0; k < 6; k++) {
        dqSelf[k] += isA ? dqAb[k] : dqBb[k];
        dqo[k] = isA ? dqBb[k] : dqAb[k];
      }
      for (int k = 0; k < 6; k++) { lk[k] = lkv[k]; rwb[k] = rwv[k]; }
      for (int k = 0; k < 3; k++) c0b[k] = c0v[k];
      atomicAdd(A.MUb + m, mub);
      if (oDyn) deltaVjpAtomic(xo, x0o, dqo, A.Xb + size_t(o) * XS, A.X0b + size_t(o) * XS);
    }
    if constexpr (TEAM > 1)
      for (int k = 0; k < 6; k++) dqSelf[k] = teamSum<TEAM>(dqSelf[k]);
    if (lane != 0) return;
    // 4. the body's own terms: inertia, its delta, the update (inputs read
    // first, outputs written last)
    double *x0b = A.X0b + size_t(i) * XS, *yb = A.Yb + size_t(i) * XS, *xb = A.Xb + size_t(i) * XS;
    double x0v[7], yv[7], xq[4], x0q[4], yq[4];
    for (int k = 0; k < 7; k++) { x0v[k] = x0b[k]; yv[k] = yb[k]; }
    for (int k = 0; k < 4; k++) { xq[k] = x[3 + k]; x0q[k] = x0[3 + k]; yq[k] = y[3 + k]; }
    double pb[3];
    for (int k = 0; k < 3; k++) {
      pb[k] = pb1[k] - ml * w[k] + dqSelf[k];  // gb_lin = -w_lin
      x0v[k] -= dqSelf[k] + dqDual[k];
      yv[k] += ml * w[k];
    }
    for (int k = 0; k < 4; k++) x0v[3 + k] += q0Dual[k];
    {
      double q0b[4] = {0, 0, 0, 0};
      qsubVjp(xq, x0q, dqSelf + 3, qb, q0b);
      for (int k = 0; k < 4; k++) x0v[3 + k] += q0b[k];
    }
    // g_ang = M e / dt^2, H_ang = M / dt^2
    double eb[3], Mb[9];
    for (int k = 0; k < 3; k++) eb[k] = -(M[k] * w[3] + M[3 + k] * w[4] + M[6 + k] * w[5]) * idt2;
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < 3; k++) Mb[r * 3 + k] = (-w[3 + r] * ei[k] - w[3 + r] * Dv[3 + k]) * idt2;
    {
      double yqb[4] = {0, 0, 0, 0};
      qsubVjp(xq, yq, eb, qb, yqb);
      for (int k = 0; k < 4; k++) yv[3 + k] += yqb[k];
    }
    if (P.rotated && bp.shape != 1) worldMomentVjp(xq, bp.m0, bp.m1, bp.m2, Mb, qb);
    for (int k = 0; k < 7; k++) { x0b[k] = x0v[k]; yb[k] = yv[k]; }
    for (int k = 0; k < 3; k++) xb[k] = pb[k];
    for (int k = 0; k < 4; k++) xb[3 + k] = qb[k];
  }
}

// Force initialisation (manifold.js:214-318): C0 from the step-start poses,
// the world anchors, and the warm start from the previous step's contacts.
__device__ inline void adjNarrow(int t, const AdjStep &T, const AdjBufs &A, const BodyProp *props,
                                 const StepParams &P) {
  const int m = T.live[t];
  const int4 h = T.head[m];
  const double *B = T.basis + size_t(m) * BAS;
  const double *x0A = T.X0 + size_t(h.x) * XS, *x0B = T.X0 + size_t(h.y) * XS;
  double qAb[4] = {0, 0, 0, 0}, qBb[4] = {0, 0, 0, 0}, pAb[3] = {0, 0, 0};
  for (int s = 0; s < h.z; s++) {
    const size_t si = size_t(m) * MAXC + s;
    const int src = T.src[si];
    const double *c0b = A.C0b + si * 3, *rwb = A.RWb + si * 6;
    const double *an = T.anchor + si * 6;
    // C0 = B (xA - xB) + margin, x = p + qrot(q, r)
    const bool kept = src >= 0 && (src & SRC_KEPT);
    const double cb[3] = {c0b[0], kept ? c0b[1] : 0.0, kept ? c0b[2] : 0.0};
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = B[k] * cb[0] + B[3 + k] * cb[1] + B[6 + k] * cb[2];
    const double nd[3] = {-d[0], -d[1], -d[2]};
    for (int k = 0; k < 3; k++) pAb[k] += d[k];
    qrotVjp(x0A + 3, an, d, qAb);
    qrotVjp(x0B + 3, an + 3, nd, qBb);
    qrotVjp(x0A + 3, an, rwb, qAb);
    qrotVjp(x0B + 3, an + 3, rwb + 3, qBb);
    // warm start: lambda0 = (lambda alpha) gamma, k0 = clamp(k gamma)
    if (src >= 0) {
      const int os = src & ~SRC_KEPT;
      const double *hot0 = T.hot + si * HOT;
      const double *lk = A.LK + si * 6;
      double *pl = A.LKprev + size_t(os) * 6;
      for (int r = 0; r < 3; r++) {
        const double k0 = hot0[HK + r];
        if (k0 > P.pmin && k0 < P.pmax) atomicAdd(pl + r, lk[r] * P.gamma);
        if (!P.postStab) atomicAdd(pl + 3 + r, lk[3 + r] * P.alpha * P.gamma);
      }
    }
  }
  double *xAb = A.X0b + size_t(h.x) * XS, *xBb = A.X0b + size_t(h.y) * XS;
  for (int k = 0; k < 3; k++) {
    atomicAdd(xAb + k, pAb[k]);
    atomicAdd(xBb + k, -pAb[k]);
  }
  for (int k = 0; k < 4; k++) {
    atomicAdd(xAb + 3 + k, qAb[k]);
    atomicAdd(xBb + 3 + k, qBb[k]);
  }
  // mu = sqrt(muA muB)
  const double mu = B[9], mA = props[h.x].friction, mB = props[h.y].friction;
  if (mu > 0) {
    atomicAdd(A.muGrad + h.x, A.MUb[m] * mB / (2 * mu));
    atomicAdd(A.muGrad + h.y, A.MUb[m] * mA / (2 * mu));
  }
}

// Body initialisation (solver.js:366-401): the adjoint of the step-start
// state, written as the pose and velocity adjoints of the previous step's end.
__device__ inline void adjInit(int i, const BodyProp *props, const AdjStep &T, const AdjBufs &A,
                               const StepParams &P) {
  const double dt = P.dt;
  const double *x0 = T.X0 + size_t(i) * XS, *v = T.V + size_t(i) * VS;
  double *xb = A.Xb + size_t(i) * XS, *x0b = A.X0b + size_t(i) * XS, *yb = A.Yb + size_t(i) * XS;
  double *vs = A.VbS + size_t(i) * VS;
  double p0b[3], q0b[4], vb[3] = {vs[0], vs[1], vs[2]}, wb[3] = {vs[3], vs[4], vs[5]};
  double prevb[3] = {vs[6], vs[7], vs[8]};
  for (int k = 0; k < 3; k++) p0b[k] = x0b[k] + xb[k] + yb[k];
  for (int k = 0; k < 4; k++) q0b[k] = x0b[3 + k];
  // y = x + v dt for every body; x^0 = x + v dt + ... for dynamic bodies only
  // (a static body never moves: its pose adjoint is a sink)
  const bool dyn = props[i].mass > 0;
  for (int k = 0; k < 3; k++) vb[k] += ((dyn ? xb[k] : 0.0) + yb[k]) * dt;
  const double wdt[3] = {v[3] * dt, v[4] * dt, v[5] * dt};
  double wdtb[3] = {0, 0, 0};
  if (dyn) {
    // x^0 = p + v dt + g wgt dt^2, wgt = clamp(a_z sgn(g) / |g|, 0, 1)
    const double accelExt = (v[2] - v[8]) / dt * sgn(P.gravity);
    const double raw = accelExt / xabs(P.gravity);
    if (raw > 0 && raw < 1) {
      const double wgtb = xb[2] * P.gravity * dt * dt;
      const double d = wgtb * sgn(P.gravity) / (dt * xabs(P.gravity));
      vb[2] += d;
      prevb[2] -= d;
    }
    qintVjp(x0 + 3, wdt, xb + 3, q0b, wdtb);
  } else {
    for (int k = 0; k < 4; k++) q0b[k] += xb[3 + k];
  }
  qintVjp(x0 + 3, wdt, yb + 3, q0b, wdtb);
  for (int k = 0; k < 3; k++) wb[k] += wdtb[k] * dt;
  for (int k = 0; k < 3; k++) xb[k] = p0b[k];
  for (int k = 0; k < 4; k++) xb[3 + k] = q0b[k];
  double *vo = A.Vb + size_t(i) * VS;
  for (int k = 0; k < 3; k++) { vo[k] = vb[k]; vo[3 + k] = wb[k]; vo[6 + k] = prevb[k]; }
  for (int k = 0; k < 8; k++) { x0b[k] = 0; yb[k] = 0; }
}

// ------------------------------------------------------------- the kernels
__global__ void k_adj_velocity(int n, const BodyProp *props, AdjStep T, AdjBufs A, int iters, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) adjVelocity(i, props, T, A, iters, dt);
}
template <int TEAM>
__global__ void __launch_bounds__(256) k_adj_primal(int c, AdjStep T, AdjBufs A, const BodyProp *props, int l,
                                                    StepParams P) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = gt % TEAM, teams = gridDim.x * blockDim.x / TEAM;
  const int cnt = T.cstart[c + 1] - T.cstart[c];
  for (int team = gt / TEAM; team < cnt; team += teams) adjPrimalTeam<TEAM>(c, team, lane, T, A, props, l, P);
}
__global__ void k_adj_narrow(AdjStep T, AdjBufs A, const BodyProp *props, StepParams P) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < T.nLive) adjNarrow(t, T, A, props, P);
}
__global__ void k_adj_init(int n, const BodyProp *props, AdjStep T, AdjBufs A, StepParams P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) adjInit(i, props, T, A, P);
}

// The whole reverse of one step in a single workgroup, for small scenes where
// the per-kernel launches would dominate: the phases of the per-kernel path
// (one per colour and iteration, the dual adjoints folded in), separated by
// workgroup barriers (one CU, so global writes of one phase are visible to
// the next), then the per-step contact adjoints cleared. One wave per SIMD: the phases are
// chains of dependent f64 work with little parallelism, and a wave of its own
// per SIMD has the whole register file (accumulation registers included)
// where two would share it and spill to memory.
constexpr int ADJ_SMALL_THREADS = 256;
template <int TEAM>
__global__ void __launch_bounds__(ADJ_SMALL_THREADS) k_adj_step_small(int n, int nTot, const BodyProp *props,
                                                                      AdjStep T, AdjBufs A, int iters,
                                                                      StepParams P) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = tid; i < n; i += nt) adjVelocity(i, props, T, A, iters, P.dt);
  __syncthreads();
  const int lane = tid % TEAM, teams = nt / TEAM;
  // each iteration's dual adjoint rides with its first colour in reverse
  for (int l = iters - 1; l >= 0; l--) {
    for (int c = T.numColors - 1; c >= 0; c--) {
      const int cnt = T.cstart[c + 1] - T.cstart[c];
      for (int team = tid / TEAM; team < cnt; team += teams)
        adjPrimalTeam<TEAM>(c, team, lane, T, A, props, l, P);
      __syncthreads();
    }
  }
  for (int t = tid; t < T.nLive; t += nt) adjNarrow(t, T, A, props, P);
  __syncthreads();
  for (int i = tid; i < n; i += nt) adjInit(i, props, T, A, P);
  const size_t slots = size_t(nTot) * MAXC;
  for (size_t k = tid; k < slots * 6; k += nt) { A.LK[k] = 0; A.RWb[k] = 0; }
  for (size_t k = tid; k < slots * 3; k += nt) A.C0b[k] = 0;
  for (int k = tid; k < nTot; k += nt) A.MUb[k] = 0;
}

// Two device copies in one launch (a recording snapshot of poses and hot
// records).
__global__ void k_snapshot(const double *a, size_t na, double *da, const double *b, size_t nb, double *db) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < na; k += stride) da[k] = a[k];
  for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nb; k += stride) db[k] = b[k];
}

// The recorded work items, as many as the step has (read on the device).
__global__ void k_copy_items(const int *count, const uint2 *src, uint2 *dst, int cap) {
  const int m = min(*count, cap);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) dst[k] = src[k];
}

}  // namespace avbd
