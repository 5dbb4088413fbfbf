// Device-side types and math shared by the gfx950 kernels of libdfr_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DFR_BLOCK 256
#define DFR_MAX_BODIES 16

// ---------------------------------------------------------------------------
// Cell grid: cells of edge >= h, keyed in 4x4x4 bricks (bricks row-major,
// cells inside a brick in Morton order) so that consecutive keys are close in
// space and the key space stays within ~1.2x the cell count.
// ---------------------------------------------------------------------------
struct GridDesc {
  float lo[3];
  float inv_cs;
  int n[3];      // cells per axis (multiples of 4)
  int bn[3];     // bricks per axis
  uint32_t nkeys;
};

__host__ __device__ inline uint32_t morton2x3(int x, int y, int z) {
  // two bits per axis interleaved: bit0 x0, bit1 y0, bit2 z0, bit3 x1, bit4 y1, bit5 z1
  return (uint32_t)((x & 1) | ((y & 1) << 1) | ((z & 1) << 2) | ((x & 2) << 2) | ((y & 2) << 3) | ((z & 2) << 4));
}

__host__ __device__ inline uint32_t cell_key(const GridDesc& g, int cx, int cy, int cz) {
  const uint32_t brick = ((uint32_t)(cz >> 2) * (uint32_t)g.bn[1] + (uint32_t)(cy >> 2)) * (uint32_t)g.bn[0] + (uint32_t)(cx >> 2);
  return brick * 64u + morton2x3(cx & 3, cy & 3, cz & 3);
}

__device__ inline void cell_coords(const GridDesc& g, float x, float y, float z, int& cx, int& cy, int& cz) {
  cx = (int)floorf((x - g.lo[0]) * g.inv_cs);
  cy = (int)floorf((y - g.lo[1]) * g.inv_cs);
  cz = (int)floorf((z - g.lo[2]) * g.inv_cs);
  cx = min(max(cx, 0), g.n[0] - 1);
  cy = min(max(cy, 0), g.n[1] - 1);
  cz = min(max(cz, 0), g.n[2] - 1);
}

// Squared distance with each product and sum rounded separately (no fused
// multiply-add), so the neighbour predicate matches the CPU oracle bit for bit.
__device__ inline float mul_unfused(float a, float b) {
  // an opaque product: the compiler cannot contract it with the following add
  float r;
  asm("v_mul_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ inline float dist2_exact(float ax, float ay, float az, float bx, float by, float bz) {
  const float dx = ax - bx, dy = ay - by, dz = az - bz;
  return (mul_unfused(dx, dx) + mul_unfused(dy, dy)) + mul_unfused(dz, dz);
}

// ---------------------------------------------------------------------------
// Cubic spline kernel, SPlisHSPlasH/SPHKernels.h:16-176 (analytic form).
// ---------------------------------------------------------------------------
struct KernelConst {
  float h, k, l, w0;
  float inv_h;   // 1/h
  float l_ih;    // l/h
  float l_ih2;   // l/h^2
};

__device__ inline float cubic_W(const KernelConst& K, float rl) {
  const float q = rl / K.h;
  float res = 0.f;
  if (q <= 1.f) {
    if (q <= 0.5f) {
      const float q2 = q * q, q3 = q2 * q;
      res = K.k * (6.f * q3 - 6.f * q2 + 1.f);
    } else {
      const float f = 1.f - q;
      res = K.k * (2.f * (f * f * f));
    }
  }
  return res;
}

// Returns the scalar c with gradW(r) = c * r (zero outside the support or for
// |r| <= 1e-5), i.e. c = l*q*(3q-2)/(|r| h) or -l*(1-q)^2/(|r| h).
__device__ inline float cubic_gradW_scale(const KernelConst& K, float rl) {
  const float q = rl / K.h;
  if (rl > 1.0e-5f && q <= 1.f) {
    const float inv = 1.f / (rl * K.h);
    if (q <= 0.5f) return K.l * q * (3.f * q - 2.f) * inv;
    const float f = 1.f - q;
    return K.l * (-f * f) * inv;
  }
  return 0.f;
}

// Division-free forms used by the gather passes.  With q = |r|/h:
//   gradW(r) = c r,  c = l (3q - 2) / h^2          for q <= 1/2
//                    c = -l (1 - q)^2 / (h |r|)     for 1/2 < q <= 1
// (the q / (|r| h) factor of SPHKernels.h:81-101 is 1/h^2), and 1/|r| comes from
// one hardware reciprocal square root.  Results differ from the reference's
// sqrt/divide form by a few ulp.
struct PairKernel {
  float W;    // W(|r|)
  float c;    // gradW(r) = c * r
};
// Both evaluations are branch-free (selects only): a wavefront evaluates the
// inner and outer polynomial for every pair instead of diverging on q <= 1/2.
// on = false gives W = 0 and c = 0 (a slot that contributes nothing).
__device__ inline PairKernel cubic_pair(const KernelConst& K, float r2, bool on = true) {
  const float irl = __builtin_amdgcn_rsqf(r2);
  const float rl = r2 * irl;
  const float q = rl * K.inv_h;
  const float f = 1.f - q;
  const bool inner = q <= 0.5f;
  const float Win = K.k * (6.f * q * q * (q - 1.f) + 1.f);
  const float Wout = 2.f * K.k * (f * f * f);
  const float cin = K.l_ih2 * (3.f * q - 2.f);
  const float cout = -K.l_ih * (f * f) * irl;
  const bool far = !(q <= 1.f);
  PairKernel o;
  // |r| <= 1e-5: gradient is zero, W = W(0)
  o.W = on ? ((r2 > 1.0e-10f) ? (far ? 0.f : (inner ? Win : Wout)) : K.k) : 0.f;
  o.c = (on && r2 > 1.0e-10f && !far) ? (inner ? cin : cout) : 0.f;
  return o;
}
// on = false gives 0 (a pair that contributes nothing), folded into the
// support test's select.
__device__ inline float cubic_grad_c(const KernelConst& K, float r2, bool on = true) {
  const float irl = __builtin_amdgcn_rsqf(r2);
  const float q = r2 * irl * K.inv_h;
  const float f = 1.f - q;
  const float cin = fmaf(q, 3.f * K.l_ih2, -2.f * K.l_ih2);
  const float cout = -K.l_ih * (f * f) * irl;
  const float cc = (q <= 0.5f) ? cin : cout;
  return (on && r2 > 1.0e-10f && q <= 1.f) ? cc : 0.f;
}

// gradGradW(r) = a * I + b * r r^T (SPHKernels.h:122-149 rewritten with
// gradq = r/(|r| h), gradGradq = (I - r r^T/|r|^2)/(|r| h)).
__device__ inline void cubic_gradGradW_ab(const KernelConst& K, float rl, float& a, float& b) {
  const float q = rl / K.h;
  a = 0.f; b = 0.f;
  if (rl > 1.0e-5f && q <= 1.f) {
    const float inv = 1.f / (rl * K.h);
    float c1, c2;  // coefficients of gradGradq and gradq gradq^T
    if (q <= 0.5f) { c1 = K.l * q * (3.f * q - 2.f); c2 = K.l * (6.f * q - 2.f); }
    else { const float f = 1.f - q; c1 = K.l * (-f * f); c2 = K.l * 2.f * f; }
    a = c1 * inv;
    // -c1*inv/|r|^2 from gradGradq plus c2 * inv^2 from gradq gradq^T
    b = -c1 * inv / (rl * rl) + c2 * inv * inv;
  }
}

// ---------------------------------------------------------------------------
// Rigid body state on the device (FP64), one per boundary body.
// ---------------------------------------------------------------------------
struct RigidDev {
  double x[3], q[4], v[3], w[3];     // q = (w, x, y, z)
  double v0[3], w0[3];               // init_v_rb / init_omega_rb
  double mass, invMass;
  double I0[9], I[9], Iinv[9];
  double F[3], T[3];                 // force/torque accumulated in the current step
  double Flast[3], Tlast[3];         // after the step: what the rigid update used
  // forward-mode Jacobians: Jxv Jxw Jvv Jvw Jwv Jww (3x3 row-major), Jqv Jqw (4x3)
  double J[78];
  // per-step pair-Jacobian sums: Fv Fx Fw Tw Tx Tv (3x3), Fq Tq (3x4)
  double net[78];
  double netlast[78];                // net of the last completed step
  // RigidBodyGradientManager path (R == RR entries): its Jacobians, same layout
  // as J, and the per-step sums it was last handed (RigidBodyGradientManager.h)
  double Jm[78];
  double mnet[78];
  int dynamic, animated, grad_active, pad;
};

// Step constants passed by value.
struct StepConst {
  KernelConst K;
  float dt, rho0, V_f, m_f;
  float g[3];
  float eps;               // 1e-5
  int nf, nb;              // fluid / boundary counts
  int maxn;                // ELL stride multiplier (slots per fluid particle)
  int maxbf;               // slots per dynamic boundary particle
  int nbodies;
};

// Wave-level sum of a double (64-lane wavefront).
__device__ inline double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Logical block index remapped so that each of the 8 XCDs of an MI355X (which
// receive workgroups round-robin) processes one contiguous range of the grid:
// neighbouring particles/bricks then share an L2.
__device__ inline uint32_t xcd_block() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t per = nb / 8u, rem = nb % 8u;
  const uint32_t x = b % 8u, k = b / 8u;
  return x * per + min(x, rem) + k;
}
__device__ inline uint32_t gid() { return xcd_block() * blockDim.x + threadIdx.x; }

// Exclusive scan of one value per thread over an NT-thread block.
// lds needs NT/64 + 1 entries.
template <int NT = DFR_BLOCK>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) lds[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int w = 0; w < NT / 64; ++w) { const uint32_t t = lds[w]; lds[w] = acc; acc += t; }
    lds[NT / 64] = acc;
  }
  __syncthreads();
  const uint32_t res = inc - v + lds[wid];
  total = lds[NT / 64];
  __syncthreads();
  return res;
}

// Block-level FP64 sum of one value per thread, stored to *out (fixed order).
template <int NT = DFR_BLOCK>
__device__ inline void block_sum_double(double v, double* out) {
  __shared__ double part[NT / 64];
  v = wave_sum_d(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; ++w) t += part[w];
    *out = t;
  }
}
