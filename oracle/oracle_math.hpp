// Small fixed-size linear algebra for the CPU oracle (test infrastructure only).
//
// This file is part of the ORACLE: a CPU restatement of DiffFR's differentiable
// DFSPH step used by the tests, by __graft_entry__.smoke() and by bench.py's
// cpu_baseline leg as the checker.  Nothing in the product (difffr_amd/) links or
// calls it.
//
// The operations mirror the Eigen 3.4 expressions the reference evaluates
// (Quaternion * Vector3 via _transformVector, toRotationMatrix, 3x3 inverse by
// cofactors), so that FP32 restatements round the way the reference does.
#pragma once
#include <cmath>
#include <cstring>

namespace orc {

// Sum of K terms in the order Eigen's unrolled fixed-size reductions use:
// split at K/2 and add the halves (3 terms: a0 + (a1 + a2); 4: (a0+a1)+(a2+a3)).
template <typename T, int K> struct EigenSum {
  static T run(const T* v) {
    if (K == 1) return v[0];
    return EigenSum<T, K / 2>::run(v) + EigenSum<T, K - K / 2>::run(v + K / 2);
  }
};
template <typename T> struct EigenSum<T, 1> { static T run(const T* v) { return v[0]; } };
template <typename T> struct EigenSum<T, 0> { static T run(const T*) { return T(0); } };

template <typename T> struct V3 {
  T x = 0, y = 0, z = 0;
  V3() = default;
  V3(T a, T b, T c) : x(a), y(b), z(c) {}
  template <typename U> explicit V3(const V3<U>& o) : x(T(o.x)), y(T(o.y)), z(T(o.z)) {}
  T& operator[](int i) { return (&x)[i]; }
  const T& operator[](int i) const { return (&x)[i]; }
  V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  V3 operator-() const { return {-x, -y, -z}; }
  V3 operator*(T s) const { return {x * s, y * s, z * s}; }
  V3 operator/(T s) const { return {x / s, y / s, z / s}; }
  V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
  V3& operator-=(const V3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
  V3& operator*=(T s) { x *= s; y *= s; z *= s; return *this; }
  T dot(const V3& o) const { return x * o.x + (y * o.y + z * o.z); }
  T squaredNorm() const { return x * x + (y * y + z * z); }
  T norm() const { return std::sqrt(squaredNorm()); }
  V3 cross(const V3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
  static V3 Zero() { return {0, 0, 0}; }
};
template <typename T> inline V3<T> operator*(T s, const V3<T>& v) { return v * s; }

// Row-major R x C matrix.
template <typename T, int R, int C> struct Mat {
  T a[R * C];
  Mat() { for (int i = 0; i < R * C; ++i) a[i] = 0; }
  template <typename U> explicit Mat(const Mat<U, R, C>& o) { for (int i = 0; i < R * C; ++i) a[i] = T(o.a[i]); }
  T& operator()(int r, int c) { return a[r * C + c]; }
  const T& operator()(int r, int c) const { return a[r * C + c]; }
  static Mat Zero() { return Mat(); }
  static Mat Identity() { Mat m; for (int i = 0; i < (R < C ? R : C); ++i) m(i, i) = 1; return m; }
  Mat operator+(const Mat& o) const { Mat m; for (int i = 0; i < R * C; ++i) m.a[i] = a[i] + o.a[i]; return m; }
  Mat operator-(const Mat& o) const { Mat m; for (int i = 0; i < R * C; ++i) m.a[i] = a[i] - o.a[i]; return m; }
  Mat operator-() const { Mat m; for (int i = 0; i < R * C; ++i) m.a[i] = -a[i]; return m; }
  Mat operator*(T s) const { Mat m; for (int i = 0; i < R * C; ++i) m.a[i] = a[i] * s; return m; }
  Mat operator/(T s) const { Mat m; for (int i = 0; i < R * C; ++i) m.a[i] = a[i] / s; return m; }
  Mat& operator+=(const Mat& o) { for (int i = 0; i < R * C; ++i) a[i] += o.a[i]; return *this; }
  Mat<T, C, R> transpose() const { Mat<T, C, R> m; for (int r = 0; r < R; ++r) for (int c = 0; c < C; ++c) m(c, r) = (*this)(r, c); return m; }
  void setZero() { for (int i = 0; i < R * C; ++i) a[i] = 0; }
  void setIdentity() { *this = Identity(); }
};
template <typename T, int R, int C> inline Mat<T, R, C> operator*(T s, const Mat<T, R, C>& m) { return m * s; }
template <typename T, int R, int K, int C>
inline Mat<T, R, C> operator*(const Mat<T, R, K>& A, const Mat<T, K, C>& B) {
  Mat<T, R, C> m;
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < C; ++c) {
      T t[K];
      for (int k = 0; k < K; ++k) t[k] = A(r, k) * B(k, c);
      m(r, c) = EigenSum<T, K>::run(t);
    }
  return m;
}
template <typename T> inline V3<T> operator*(const Mat<T, 3, 3>& A, const V3<T>& v) {
  return {A(0, 0) * v.x + (A(0, 1) * v.y + A(0, 2) * v.z),
          A(1, 0) * v.x + (A(1, 1) * v.y + A(1, 2) * v.z),
          A(2, 0) * v.x + (A(2, 1) * v.y + A(2, 2) * v.z)};
}
template <typename T> using M3 = Mat<T, 3, 3>;
template <typename T> using M34 = Mat<T, 3, 4>;
template <typename T> using M43 = Mat<T, 4, 3>;
template <typename T> using M4 = Mat<T, 4, 4>;

// Outer product a b^T.
template <typename T> inline M3<T> outer(const V3<T>& a, const V3<T>& b) {
  M3<T> m;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m(r, c) = a[r] * b[c];
  return m;
}

// 3x3 inverse via the adjugate, the cofactor scheme Eigen uses for fixed 3x3
// matrices (Eigen/src/LU/InverseImpl.h, compute_inverse_size3_helper).
template <typename T> inline M3<T> inverse(const M3<T>& m) {
  auto cof = [&](int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m(i1, j1) * m(i2, j2) - m(i1, j2) * m(i2, j1);
  };
  const T c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
  const T det = c00 * m(0, 0) + c10 * m(1, 0) + c20 * m(2, 0);
  const T inv = T(1) / det;
  M3<T> r;
  r(0, 0) = c00 * inv; r(0, 1) = c10 * inv; r(0, 2) = c20 * inv;
  r(1, 0) = cof(0, 1) * inv; r(1, 1) = cof(1, 1) * inv; r(1, 2) = cof(2, 1) * inv;
  r(2, 0) = cof(0, 2) * inv; r(2, 1) = cof(1, 2) * inv; r(2, 2) = cof(2, 2) * inv;
  return r;
}

// Skew-symmetric cross-product matrix, SPlisHSPlasH/GradientUtils.cpp:4-6.
template <typename T> inline M3<T> skew(const V3<T>& v) {
  M3<T> m;
  m(0, 1) = -v.z; m(0, 2) = v.y;
  m(1, 0) = v.z;  m(1, 2) = -v.x;
  m(2, 0) = -v.y; m(2, 1) = v.x;
  return m;
}

// Hamilton quaternion stored as (w, x, y, z).
template <typename T> struct Quat {
  T w = 1, x = 0, y = 0, z = 0;
  Quat() = default;
  Quat(T w_, T x_, T y_, T z_) : w(w_), x(x_), y(y_), z(z_) {}
  template <typename U> explicit Quat(const Quat<U>& o) : w(T(o.w)), x(T(o.x)), y(T(o.y)), z(T(o.z)) {}
  V3<T> vec() const { return {x, y, z}; }
  Quat operator*(const Quat& o) const {
    return {w * o.w - x * o.x - y * o.y - z * o.z,
            w * o.x + x * o.w + y * o.z - z * o.y,
            w * o.y + y * o.w + z * o.x - x * o.z,
            w * o.z + z * o.w + x * o.y - y * o.x};
  }
  // Eigen normalises coeffs() = (x, y, z, w): (x^2 + y^2) + (z^2 + w^2)
  T norm() const { return std::sqrt((x * x + y * y) + (z * z + w * w)); }
  void normalize() { const T n = norm(); if (n > T(0)) { x /= n; y /= n; z /= n; w /= n; } }
  Quat normalized() const { Quat q = *this; q.normalize(); return q; }
  // Eigen QuaternionBase::_transformVector.
  V3<T> rotate(const V3<T>& v) const {
    V3<T> uv = vec().cross(v);
    uv += uv;
    return v + uv * w + vec().cross(uv);
  }
  // Eigen QuaternionBase::toRotationMatrix.
  M3<T> toRotationMatrix() const {
    M3<T> r;
    const T tx = T(2) * x, ty = T(2) * y, tz = T(2) * z;
    const T twx = tx * w, twy = ty * w, twz = tz * w;
    const T txx = tx * x, txy = ty * x, txz = tz * x;
    const T tyy = ty * y, tyz = tz * y, tzz = tz * z;
    r(0, 0) = T(1) - (tyy + tzz); r(0, 1) = txy - twz; r(0, 2) = txz + twy;
    r(1, 0) = txy + twz; r(1, 1) = T(1) - (txx + tzz); r(1, 2) = tyz - twx;
    r(2, 0) = txz - twy; r(2, 1) = tyz + twx; r(2, 2) = T(1) - (txx + tyy);
    return r;
  }
};

// d(R(q) p)/dq, columns (w, x, y, z): SPlisHSPlasH/GradientUtils.cpp:9-20.
template <typename T> inline M34<T> grad_Rqp_to_q(const Quat<T>& q, const V3<T>& p) {
  const V3<T> qv = q.vec();
  const V3<T> t1 = (p * q.w - p.cross(qv)) * T(2);
  M3<T> t2 = (M3<T>::Identity() * qv.dot(p) + outer(qv, p) - outer(p, qv) - skew(p) * q.w) * T(2);
  M34<T> r;
  for (int i = 0; i < 3; ++i) {
    r(i, 0) = t1[i];
    for (int j = 0; j < 3; ++j) r(i, 1 + j) = t2(i, j);
  }
  return r;
}

// d(R(q)^T p)/dq, columns (w, x, y, z): SPlisHSPlasH/GradientUtils.cpp:23-34.
template <typename T> inline M34<T> grad_RqTp_to_q(const Quat<T>& q, const V3<T>& p) {
  const V3<T> qv = q.vec();
  const V3<T> t1 = (p * q.w + p.cross(qv)) * T(2);
  M3<T> t2 = (M3<T>::Identity() * qv.dot(p) + outer(qv, p) - outer(p, qv) + skew(p) * q.w) * T(2);
  M34<T> r;
  for (int i = 0; i < 3; ++i) {
    r(i, 0) = t1[i];
    for (int j = 0; j < 3; ++j) r(i, 1 + j) = t2(i, j);
  }
  return r;
}

// d(p (x) q)/dq for fixed p, (w,x,y,z) ordering: SPlisHSPlasH/GradientUtils.cpp:36-43.
template <typename T> inline M4<T> grad_p_q_product_to_q(const Quat<T>& p) {
  M4<T> m;
  const T v[16] = {p.w, -p.x, -p.y, -p.z,
                   p.x,  p.w, -p.z,  p.y,
                   p.y,  p.z,  p.w, -p.x,
                   p.z, -p.y,  p.x,  p.w};
  for (int i = 0; i < 16; ++i) m.a[i] = v[i];
  return m;
}

// d((0,omega) (x) q)/d omega: SPlisHSPlasH/GradientUtils.cpp:45-53.
template <typename T> inline M43<T> grad_omega_q_product_to_omega(const Quat<T>& q) {
  M43<T> m;
  const T v[12] = {-q.x, -q.y, -q.z,
                    q.w,  q.z, -q.y,
                   -q.z,  q.w,  q.x,
                    q.y, -q.x,  q.w};
  for (int i = 0; i < 12; ++i) m.a[i] = v[i];
  return m;
}

}  // namespace orc
