// Small FP64/FP32 matrix, vector and quaternion helpers for the rigid-body
// kernels (forward chain rule, rigid update and its adjoint).
#pragma once
#include "dfr_kernels.h"

namespace dfr {

// Small fixed-size row-major matrices for device code.
template <typename T, int R, int C> struct DM {
  T a[R * C];
  __device__ T& operator()(int r, int c) { return a[r * C + c]; }
  __device__ const T& operator()(int r, int c) const { return a[r * C + c]; }
  __device__ static DM zero() { DM m; for (int i = 0; i < R * C; ++i) m.a[i] = T(0); return m; }
  __device__ static DM eye() { DM m = zero(); for (int i = 0; i < (R < C ? R : C); ++i) m(i, i) = T(1); return m; }
  __device__ DM operator+(const DM& o) const { DM m; for (int i = 0; i < R * C; ++i) m.a[i] = a[i] + o.a[i]; return m; }
  __device__ DM operator-(const DM& o) const { DM m; for (int i = 0; i < R * C; ++i) m.a[i] = a[i] - o.a[i]; return m; }
  __device__ DM operator*(T s) const { DM m; for (int i = 0; i < R * C; ++i) m.a[i] = a[i] * s; return m; }
  __device__ DM<T, C, R> tr() const { DM<T, C, R> m; for (int r = 0; r < R; ++r) for (int c = 0; c < C; ++c) m(c, r) = (*this)(r, c); return m; }
};
template <typename T, int R, int K, int C>
__device__ inline DM<T, R, C> mm(const DM<T, R, K>& A, const DM<T, K, C>& B) {
  DM<T, R, C> m;
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < C; ++c) {
      T s = T(0);
      for (int k = 0; k < K; ++k) s += A(r, k) * B(k, c);
      m(r, c) = s;
    }
  return m;
}
template <typename T> using DM3 = DM<T, 3, 3>;
template <typename T> struct DV3 { T x, y, z; };
template <typename T> __device__ inline DV3<T> mv(const DM3<T>& A, DV3<T> v) {
  return {A(0, 0) * v.x + A(0, 1) * v.y + A(0, 2) * v.z, A(1, 0) * v.x + A(1, 1) * v.y + A(1, 2) * v.z,
          A(2, 0) * v.x + A(2, 1) * v.y + A(2, 2) * v.z};
}
template <typename T> __device__ inline DV3<T> cross(DV3<T> a, DV3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> __device__ inline DM3<T> skew(DV3<T> v) {
  DM3<T> m = DM3<T>::zero();
  m(0, 1) = -v.z; m(0, 2) = v.y; m(1, 0) = v.z; m(1, 2) = -v.x; m(2, 0) = -v.y; m(2, 1) = v.x;
  return m;
}
template <typename T> __device__ inline DM3<T> outer(DV3<T> a, DV3<T> b) {
  DM3<T> m;
  const T av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
  for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) m(r, c) = av[r] * bv[c];
  return m;
}
// 3x3 inverse by the adjugate (the cofactor scheme used for fixed 3x3 inverses).
template <typename T> __device__ inline DM3<T> inverse3(const DM3<T>& m) {
  auto cof = [&](int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m(i1, j1) * m(i2, j2) - m(i1, j2) * m(i2, j1);
  };
  const T c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
  const T det = c00 * m(0, 0) + c10 * m(1, 0) + c20 * m(2, 0);
  const T inv = T(1) / det;
  DM3<T> r;
  r(0, 0) = c00 * inv; r(0, 1) = c10 * inv; r(0, 2) = c20 * inv;
  r(1, 0) = cof(0, 1) * inv; r(1, 1) = cof(1, 1) * inv; r(1, 2) = cof(2, 1) * inv;
  r(2, 0) = cof(0, 2) * inv; r(2, 1) = cof(1, 2) * inv; r(2, 2) = cof(2, 2) * inv;
  return r;
}
// d(R(q) p)/dq and d(R(q)^T p)/dq, columns (w,x,y,z): GradientUtils.cpp:9-34.
template <typename T> __device__ inline DM<T, 3, 4> grad_Rqp(const T q[4], DV3<T> p, T sgn) {
  const DV3<T> qv = {q[1], q[2], q[3]};
  const DV3<T> pc = cross(p, qv);
  const DV3<T> t1 = {T(2) * (q[0] * p.x - sgn * pc.x), T(2) * (q[0] * p.y - sgn * pc.y), T(2) * (q[0] * p.z - sgn * pc.z)};
  const T d = qv.x * p.x + qv.y * p.y + qv.z * p.z;
  DM3<T> t2 = DM3<T>::eye() * d + outer(qv, p) - outer(p, qv) - skew(p) * (sgn * q[0]);
  DM<T, 3, 4> r;
  const T tv[3] = {t1.x, t1.y, t1.z};
  for (int i = 0; i < 3; ++i) { r(i, 0) = tv[i]; for (int j = 0; j < 3; ++j) r(i, 1 + j) = T(2) * t2(i, j); }
  return r;
}
__device__ inline void quat_mul(const double p[4], const double q[4], double o[4]) {  // (w,x,y,z)
  o[0] = p[0] * q[0] - p[1] * q[1] - p[2] * q[2] - p[3] * q[3];
  o[1] = p[0] * q[1] + p[1] * q[0] + p[2] * q[3] - p[3] * q[2];
  o[2] = p[0] * q[2] + p[2] * q[0] + p[3] * q[1] - p[1] * q[3];
  o[3] = p[0] * q[3] + p[3] * q[0] + p[1] * q[2] - p[2] * q[1];
}
__device__ inline DM3<double> quat_to_R(const double q[4]) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  const double tx = 2 * x, ty = 2 * y, tz = 2 * z, twx = tx * w, twy = ty * w, twz = tz * w;
  const double txx = tx * x, txy = ty * x, txz = tz * x, tyy = ty * y, tyz = tz * y, tzz = tz * z;
  DM3<double> r;
  r(0, 0) = 1 - (tyy + tzz); r(0, 1) = txy - twz; r(0, 2) = txz + twy;
  r(1, 0) = txy + twz; r(1, 1) = 1 - (txx + tzz); r(1, 2) = tyz - twx;
  r(2, 0) = txz - twy; r(2, 1) = tyz + twx; r(2, 2) = 1 - (txx + tyy);
  return r;
}
__device__ inline DV3<double> quat_rotate(const double q[4], DV3<double> v) {  // Eigen _transformVector form
  const DV3<double> qv = {q[1], q[2], q[3]};
  DV3<double> uv = cross(qv, v);
  uv = {uv.x + uv.x, uv.y + uv.y, uv.z + uv.z};
  const DV3<double> c = cross(qv, uv);
  return {v.x + q[0] * uv.x + c.x, v.y + q[0] * uv.y + c.y, v.z + q[0] * uv.z + c.z};
}

using M3d = DM3<double>;
using M34d = DM<double, 3, 4>;
using M43d = DM<double, 4, 3>;
using M4d = DM<double, 4, 4>;
using V3d = DV3<double>;

__device__ inline M3d ld3(const double* p) { M3d m; for (int k = 0; k < 9; ++k) m.a[k] = p[k]; return m; }
__device__ inline M43d ld43(const double* p) { M43d m; for (int k = 0; k < 12; ++k) m.a[k] = p[k]; return m; }
__device__ inline M34d ld34(const double* p) { M34d m; for (int k = 0; k < 12; ++k) m.a[k] = p[k]; return m; }
__device__ inline void st3(double* p, const M3d& m) { for (int k = 0; k < 9; ++k) p[k] = m.a[k]; }
__device__ inline void st43(double* p, const M43d& m) { for (int k = 0; k < 12; ++k) p[k] = m.a[k]; }

}  // namespace dfr
