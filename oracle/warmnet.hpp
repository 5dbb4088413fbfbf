// The learned warm-start network for the CPU oracle (test infrastructure only).
//
// 6 -> 32 -> 32 -> 32 -> 2 MLP with SiLU.  Parameters: flat float[2402] =
// W1[32][6] b1[32] W2[32][32] b2[32] W3[32][32] b3[32] W4[2][32] b4[2];
// norm[16] = feature mean[6], feature std[6], target mean[2], target std[2].
// Two arithmetic modes:
//   exact:  double everywhere (finite-difference pinning);
//   bf16:   float with weights, layer inputs and backward operands rounded to
//           bfloat16 (nearest even) where the GPU's matrix cores take them.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

inline float bf16_round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  std::memcpy(&x, &u, 4);
  return x;
}

struct WarmNet {
  static constexpr int NP = 2402;
  static constexpr int O[4] = {32, 32, 32, 2}, I[4] = {6, 32, 32, 32};
  static constexpr int OW[4] = {0, 224, 1280, 2336}, OB[4] = {192, 1248, 2304, 2400};
  std::vector<double> P;
  double mu_in[6], sd_in[6], mu_out[2], sd_out[2];
  bool bf16 = false;
  bool enabled = false;

  void set(const float* params, const float* norm, int use_bf16) {
    P.assign(params, params + NP);
    for (int k = 0; k < 6; ++k) { mu_in[k] = norm[k]; sd_in[k] = norm[6 + k]; }
    for (int k = 0; k < 2; ++k) { mu_out[k] = norm[12 + k]; sd_out[k] = norm[14 + k]; }
    bf16 = use_bf16 != 0;
    enabled = true;
  }

  double r(double x) const { return bf16 ? double(bf16_round(float(x))) : x; }
  double w(int L, int o, int i) const { return r(P[OW[L] + o * I[L] + i]); }
  static double silu(double z) { return z / (1.0 + std::exp(-z)); }
  static double silu_d(double z) {
    const double s = 1.0 / (1.0 + std::exp(-z));
    return s * (1.0 + z * (1.0 - s));
  }
  // float accumulation in bf16 mode, double otherwise
  double acc_add(double a, double b) const { return bf16 ? double(float(a) + float(b)) : a + b; }

  struct Cache { double h[4][64]; double z[4][64]; };

  // Predictions (kappa_v, kappa) in raw units.
  void forward(const double f[6], double y[2], Cache* cache = nullptr) const {
    Cache loc;
    Cache& C = cache ? *cache : loc;
    for (int k = 0; k < 6; ++k) {
      const double x = bf16 ? double((float(f[k]) - float(mu_in[k])) * (1.f / float(sd_in[k]))) : (f[k] - mu_in[k]) / sd_in[k];
      C.h[0][k] = r(x);
    }
    for (int L = 0; L < 4; ++L)
      for (int o = 0; o < O[L]; ++o) {
        double a = 0;
        for (int i = 0; i < I[L]; ++i) a = acc_add(a, w(L, o, i) * C.h[L][i]);
        const double z = acc_add(a, P[OB[L] + o]);
        C.z[L][o] = z;
        if (L < 3) C.h[L + 1][o] = r(silu(z));
        else y[o] = z;
      }
    for (int o = 0; o < 2; ++o) y[o] = y[o] * sd_out[o] + mu_out[o];
  }

  // dL/d(raw features) for dL/d(raw predictions) = ybar.
  void input_grad(const double f[6], const double ybar[2], double fbar[6]) const {
    Cache C;
    double y[2];
    forward(f, y, &C);
    double dz[64] = {0}, dh[64];
    for (int o = 0; o < 2; ++o) dz[o] = ybar[o] * sd_out[o];
    for (int L = 3; L >= 0; --L) {
      for (int i = 0; i < I[L]; ++i) {
        double a = 0;
        for (int o = 0; o < O[L]; ++o) a = acc_add(a, w(L, o, i) * r(dz[o]));
        dh[i] = a;
      }
      if (L == 0) break;
      for (int i = 0; i < I[L]; ++i) dz[i] = dh[i] * silu_d(C.z[L - 1][i]);
    }
    for (int k = 0; k < 6; ++k) fbar[k] = dh[k] / sd_in[k];
  }
};

}  // namespace orc
