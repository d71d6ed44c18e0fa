// avbd_math.hpp — device math for the AVBD step (gfx950).
//
// Vectors are plain arrays; quaternions are xyzw with the scalar last; 3x3
// matrices are row-major. The operation order of every routine follows the
// reference's f64 engine (src/math/maths.js) so that, evaluated in double with
// contraction off, results are identical to the reference's; evaluated in
// float the same code is the FP32 variant of the step.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace avbd {

#define AVBD_HD __host__ __device__ __forceinline__

// Math.min / Math.max semantics: NaN propagates, min(-0, +0) = -0.
template <class T> AVBD_HD T jmin(T a, T b) {
  if (a != a || b != b) return a + b;
  if (a == b) return __builtin_signbit(a) ? a : b;
  return a < b ? a : b;
}
template <class T> AVBD_HD T jmax(T a, T b) {
  if (a != a || b != b) return a + b;
  if (a == b) return __builtin_signbit(a) ? b : a;
  return a > b ? a : b;
}
template <class T> AVBD_HD T clampj(T x, T lo, T hi) { return jmax(lo, jmin(hi, x)); }  // maths.js:38-40
template <class T> AVBD_HD T sgn(T x) { return x < T(0) ? T(-1) : x > T(0) ? T(1) : T(0); }  // maths.js:34-36
AVBD_HD double xsqrt(double x) { return __builtin_sqrt(x); }
AVBD_HD float xsqrt(float x) { return __builtin_sqrtf(x); }
AVBD_HD double xabs(double x) { return __builtin_fabs(x); }
AVBD_HD float xabs(float x) { return __builtin_fabsf(x); }

template <class T> AVBD_HD T dot3(const T *a, const T *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <class T> AVBD_HD void cross3(T *o, const T *a, const T *b) {
  const T ax = a[0], ay = a[1], az = a[2], bx = b[0], by = b[1], bz = b[2];
  o[0] = ay * bz - az * by;
  o[1] = az * bx - ax * bz;
  o[2] = ax * by - ay * bx;
}

// Row-basis with row 0 = n (maths.js:389-414).
template <class T> AVBD_HD void orthonormal(T *o, const T *n) {
  const T nx = n[0], ny = n[1], nz = n[2];
  T tx, ty, tz;
  if (xabs(nx) > xabs(nz)) { tx = -ny; ty = nx; tz = 0; }
  else { tx = 0; ty = -nz; tz = ny; }
  const T len = xsqrt(tx * tx + ty * ty + tz * tz);
  tx /= len; ty /= len; tz /= len;
  const T sx = ny * tz - nz * ty, sy = nz * tx - nx * tz, sz = nx * ty - ny * tx;
  o[0] = nx; o[1] = ny; o[2] = nz;
  o[3] = tx; o[4] = ty; o[5] = tz;
  o[6] = sx; o[7] = sy; o[8] = sz;
}

// Eq. 20: vector part of 2 a b^-1 (maths.js:541-555).
template <class T> AVBD_HD void qsub(T *o, const T *a, const T *b) {
  const T lsq = b[0] * b[0] + b[1] * b[1] + b[2] * b[2] + b[3] * b[3];
  const T ix = -b[0] / lsq, iy = -b[1] / lsq, iz = -b[2] / lsq, iw = b[3] / lsq;
  const T ax = a[0], ay = a[1], az = a[2], aw = a[3];
  o[0] = (aw * ix + ax * iw + ay * iz - az * iy) * T(2);
  o[1] = (aw * iy - ax * iz + ay * iw + az * ix) * T(2);
  o[2] = (aw * iz + ax * iy - ay * ix + az * iw) * T(2);
}

// Eq. 21: normalize(q + 1/2 (0, dw) q), keeping the 0*q terms (maths.js:568-592).
template <class T> AVBD_HD void qint(T *o, const T *q, const T *dw) {
  const T qx = q[0], qy = q[1], qz = q[2], qw = q[3];
  const T wx = dw[0], wy = dw[1], wz = dw[2];
  const T px = T(0) * qx + wx * qw + wy * qz - wz * qy;
  const T py = T(0) * qy - wx * qz + wy * qw + wz * qx;
  const T pz = T(0) * qz + wx * qy - wy * qx + wz * qw;
  const T pw = T(0) * qw - wx * qx - wy * qy - wz * qz;
  const T rx = qx + px * T(0.5), ry = qy + py * T(0.5), rz = qz + pz * T(0.5), rw = qw + pw * T(0.5);
  const T len = xsqrt(rx * rx + ry * ry + rz * rz + rw * rw);
  o[0] = rx / len; o[1] = ry / len; o[2] = rz / len; o[3] = rw / len;
}

// Rotate v by q, t = 2 (u x v) form (maths.js:599-613).
template <class T> AVBD_HD void qrot(T *o, const T *q, const T *v) {
  const T ux = q[0], uy = q[1], uz = q[2], w = q[3];
  const T vx = v[0], vy = v[1], vz = v[2];
  const T tx = (uy * vz - uz * vy) * T(2);
  const T ty = (uz * vx - ux * vz) * T(2);
  const T tz = (ux * vy - uy * vx) * T(2);
  o[0] = vx + tx * w + (uy * tz - uz * ty);
  o[1] = vy + ty * w + (uz * tx - ux * tz);
  o[2] = vz + tz * w + (ux * ty - uy * tx);
}

// rotate(q, v) + p (maths.js:621-627).
template <class T> AVBD_HD void xform(T *o, const T *p, const T *q, const T *v) {
  qrot(o, q, v);
  o[0] += p[0]; o[1] += p[1]; o[2] += p[2];
}

// Row-wise 3x3 * vector.
template <class T> AVBD_HD void mulv(T *o, const T *a, const T *v) {
  const T x = v[0], y = v[1], z = v[2];
  o[0] = a[0] * x + a[1] * y + a[2] * z;
  o[1] = a[3] * x + a[4] * y + a[5] * z;
  o[2] = a[6] * x + a[7] * y + a[8] * z;
}

template <class T> AVBD_HD void mul33(T *o, const T *a, const T *b) {
  T r[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) r[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
#pragma unroll
  for (int i = 0; i < 9; i++) o[i] = r[i];
}

// R^T under row-wise multiplication (maths.js:417-436).
template <class T> AVBD_HD void fromQuat(T *o, const T *q) {
  const T x = q[0], y = q[1], z = q[2], w = q[3];
  const T xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  const T wx = w * x, wy = w * y, wz = w * z;
  o[0] = T(1) - T(2) * (yy + zz); o[1] = T(2) * (xy + wz); o[2] = T(2) * (xz - wy);
  o[3] = T(2) * (xy - wz); o[4] = T(1) - T(2) * (xx + zz); o[5] = T(2) * (yz + wx);
  o[6] = T(2) * (xz + wy); o[7] = T(2) * (yz - wx); o[8] = T(1) - T(2) * (xx + yy);
}

// World inertia R I R^T (rigid.js:228-247), the full products of the reference.
template <class T> AVBD_HD void worldMoment(T *o, const T *q, T m0, T m1, T m2, bool isotropic) {
  if (isotropic) {
    o[0] = m0; o[1] = 0; o[2] = 0; o[3] = 0; o[4] = m1; o[5] = 0; o[6] = 0; o[7] = 0; o[8] = m2;
    return;
  }
  T R[9], Rt[9], D[9] = {m0, 0, 0, 0, m1, 0, 0, 0, m2};
  fromQuat(R, q);
  Rt[0] = R[0]; Rt[1] = R[3]; Rt[2] = R[6]; Rt[3] = R[1]; Rt[4] = R[4]; Rt[5] = R[7];
  Rt[6] = R[2]; Rt[7] = R[5]; Rt[8] = R[8];
  mul33(o, Rt, D);
  mul33(o, o, R);
}

// Unrolled LDL^T of the 6x6 system held as lower-triangle blocks
// (maths.js:660-725). Always double.
AVBD_HD void solve6(const double *aLin, const double *aAng, const double *aCross, const double *bLin,
                    const double *bAng, double *xLin, double *xAng) {
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

}  // namespace avbd
