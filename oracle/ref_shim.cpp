// C entry points over the reference's own cubic, cohesion and adhesion kernels and gradient helpers
// (test infrastructure).  Compiled by oracle/Makefile.ref together with the
// reference sources SPlisHSPlasH/SPHKernels.cpp and SPlisHSPlasH/GradientUtils.cpp
// taken in place from the reference tree (FP32 build, Real = float); the result
// lands in oracle/_ref/ and is only used to check the oracle.
#include "SPlisHSPlasH/SPHKernels.h"
#include "SPlisHSPlasH/GradientUtils.h"

extern "C" {

void ref_cubic_eval(double radius, const double* r, int n, double* W, double* gW, double* ggW) {
  SPH::CubicKernel::setRadius(static_cast<Real>(radius));
  for (int i = 0; i < n; ++i) {
    const Vector3r v(static_cast<Real>(r[3 * i]), static_cast<Real>(r[3 * i + 1]), static_cast<Real>(r[3 * i + 2]));
    W[i] = SPH::CubicKernel::W(v);
    const Vector3r g = SPH::CubicKernel::gradW(v);
    for (int k = 0; k < 3; ++k) gW[3 * i + k] = g[k];
    const Matrix3r H = SPH::CubicKernel::gradGradW(v);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) ggW[9 * i + 3 * a + b] = H(a, b);
  }
}

// Akinci et al. 2013 cohesion and adhesion kernels (radius = support radius).
void ref_cohesion_adhesion_eval(double radius, const double* r, int n, double* C, double* A) {
  SPH::CohesionKernel::setRadius(static_cast<Real>(radius));
  SPH::AdhesionKernel::setRadius(static_cast<Real>(radius));
  for (int i = 0; i < n; ++i) {
    const Vector3r v(static_cast<Real>(r[3 * i]), static_cast<Real>(r[3 * i + 1]), static_cast<Real>(r[3 * i + 2]));
    C[i] = SPH::CohesionKernel::W(v);
    A[i] = SPH::AdhesionKernel::W(v);
  }
}

double ref_cubic_w_zero(double radius) {
  SPH::CubicKernel::setRadius(static_cast<Real>(radius));
  return SPH::CubicKernel::W_zero();
}

// q = (w, x, y, z)
void ref_grad_utils(const double* q, const double* p, double* Rqp, double* RqTp, double* PQ, double* WQ,
                    double* skew) {
  const Quaternionr qq(static_cast<Real>(q[0]), static_cast<Real>(q[1]), static_cast<Real>(q[2]),
                       static_cast<Real>(q[3]));
  const Vector3r pp(static_cast<Real>(p[0]), static_cast<Real>(p[1]), static_cast<Real>(p[2]));
  const Matrix34r a = get_grad_Rqp_to_q(qq, pp);
  const Matrix34r b = get_grad_RqTp_to_q(qq, pp);
  const Matrix4r c = get_grad_p_q_product_to_q(qq);
  const Matrix43r d = get_grad_omega_q_product_to_omega(qq);
  const Matrix3r s = skewMatrix(pp);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) { Rqp[4 * i + j] = a(i, j); RqTp[4 * i + j] = b(i, j); }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) PQ[4 * i + j] = c(i, j);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) WQ[3 * i + j] = d(i, j);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) skew[3 * i + j] = s(i, j);
}
}
