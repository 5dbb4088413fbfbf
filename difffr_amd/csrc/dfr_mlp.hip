// Learned kappa warm start: a 6 -> 32 -> 32 -> 32 -> 2 MLP (SiLU) that
// predicts, per fluid particle, the stiffness the divergence and pressure
// solves converge to (kappa_v, kappa after the step, in the stored units of
// TimeStepDiffDFSPH: kappa_v * dt, kappa * dt^2), from six features known
// before the solves:
//   rho/rho0 - 1, dt * (velocity divergence), neighbour count,
//   boundary density sum_b V_b W_ib, kappa and kappa_v of the previous step.
// The predictions replace the halved previous-step stiffness of
// warmstartDivergenceSolve / warmstartPressureSolve (TimeStepDiffDFSPH.cpp:964-1059,
// :1698-1788); the Jacobi solves still iterate to the same tolerance.
//
// All layers run on bf16 MFMA with FP32 accumulation
// (v_mfma_f32_16x16x32_bf16, gfx950).  One wavefront evaluates 64 particles:
// activations live in LDS as bf16, particle-major (row p: the layer's units,
// padded) for the B operand of W * H and, when training, unit-major for the
// operands of dW = dZ * H^T.  Weights are kept in FP32 (Adam master copy) and
// repacked after every update into bf16 A-operand fragments (W and W^T), so an
// MFMA operand is one 16-byte load per lane.
//
// MFMA 16x16x32 operand layout (lane l, j = 0..7 / 0..3):
//   A[m][k]: m = l % 16, k = 8 (l / 16) + j;  B[k][n]: k = 8 (l / 16) + j, n = l % 16;
//   D[m][n]: m = 4 (l / 16) + j, n = l % 16.
#include "dfr_kernels.h"

namespace dfr {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4;

// layer l: W_l is (MLP_O[l] x MLP_I[l]), row-major, followed by its bias
constexpr int MLP_H = 32;           // hidden units per layer (a multiple of 32: whole K chunks)
constexpr int MLP_HT = MLP_H / 16;  // D tiles of a hidden layer
constexpr int MLP_KC = MLP_H / 32;  // K chunks of a hidden layer's output
__device__ constexpr int MLP_O[4] = {MLP_H, MLP_H, MLP_H, 2};
__device__ constexpr int MLP_I[4] = {6, MLP_H, MLP_H, MLP_H};
__device__ constexpr int MLP_OW[4] = {0, 224, 1280, 2336};    // weight offsets in the flat parameter vector
__device__ constexpr int MLP_OB[4] = {192, 1248, 2304, 2400}; // bias offsets
// fragment blocks (64 lanes x bf16x8 each): forward (A = W) and transposed (A = W^T)
__device__ constexpr int MLP_FW[4] = {0, 2, 4, 6};
__device__ constexpr int MLP_FT[4] = {13, 11, 9, 7};
constexpr int MLP_NFRAG = 14;
// A second set of the same 44 fragments, K-permuted for register chaining
// (fragment f of the chain set is at MLP_NFRAG + f).  The D tile of a layer
// gives lane group g the rows 4g..4g+3; the two D tiles 2c, 2c+1 therefore
// form lane group g's eight B elements of K chunk c if element j of that group
// stands for unit chain_unit(c, g, j) -- and the A fragments use the same
// unit for that K slot, so the product is unchanged.
constexpr int MLP_NFRAG_ALL = 2 * MLP_NFRAG;
__host__ __device__ constexpr int chain_unit(int kc, int g, int j) {
  return 32 * kc + (j < 4 ? 4 * g + j : 12 + 4 * g + j);
}

constexpr int LD0 = 40;   // particle-major stride of the padded input (32 + 8)
constexpr int LDH = 72;   // particle-major / unit-major stride of 64-wide layers (64 + 8)

__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// FP32 master parameters -> bf16 A fragments.  Fragment (mt, kc) of a layer,
// lane l, element j holds A[16 mt + l % 16][32 kc + 8 (l / 16) + j] with
// A = W (forward) or A = W^T (backward to the layer input), zero outside.
// ---------------------------------------------------------------------------
__global__ void k_mlp_pack(const float* __restrict__ P, bf16x8* __restrict__ F) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= MLP_NFRAG_ALL * 64) return;
  const int l = g & 63;
  const bool chain = (g >> 6) >= MLP_NFRAG;
  const int f = (g >> 6) - (chain ? MLP_NFRAG : 0);  // fragment index within its set
  int layer = -1, tr = 0, local = 0;
  for (int L = 0; L < 4; ++L) {
    const int nf = ceil_div(MLP_O[L], 16) * ceil_div(MLP_I[L], 32);
    const int nt = ceil_div(MLP_I[L], 16) * ceil_div(MLP_O[L], 32);
    if (f >= MLP_FW[L] && f < MLP_FW[L] + nf) { layer = L; tr = 0; local = f - MLP_FW[L]; }
    if (f >= MLP_FT[L] && f < MLP_FT[L] + nt) { layer = L; tr = 1; local = f - MLP_FT[L]; }
  }
  bf16x8 v;
  for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.f;
  if (layer >= 0) {
    const int O = MLP_O[layer], I = MLP_I[layer];
    const int kt = tr ? ceil_div(O, 32) : ceil_div(I, 32);
    const int mt = local / kt, kc = local % kt;
    const int m = 16 * mt + (l & 15);
    for (int j = 0; j < 8; ++j) {
      const int k = chain ? chain_unit(kc, l >> 4, j) : 32 * kc + 8 * (l >> 4) + j;
      const int o = tr ? k : m, i = tr ? m : k;
      const float w = (o < O && i < I) ? P[MLP_OW[layer] + o * I + i] : 0.f;
      v[j] = (__bf16)w;
    }
  }
  F[g] = v;
}

// ---------------------------------------------------------------------------
// Feature sources.  Raw features per particle:
//   f0 = rho/rho0 - 1, f1 = dt * delta, f2 = neighbour count,
//   f3 = sum_b V_b W_ib, f4 = kappa (previous step), f5 = kappa_v (previous step)
// ---------------------------------------------------------------------------
__device__ inline void features_from_fields(const MlpFields& s, uint32_t i, float f[6]) {
  const float2 a = s.aux[i];
  const uint32_t c = s.cnt[i];
  f[0] = s.rho[i] / s.rho0 - 1.f;
  f[1] = s.dt * a.x;
  f[2] = (float)((c & 0xffffu) + (c >> 16));
  f[3] = a.y;
  f[4] = s.kappa[i];
  f[5] = s.kappav[i];
}

enum MlpMode { MLP_INFER = 0, MLP_TRAIN = 1, MLP_INPUT_GRAD = 2 };

struct MlpArgs {
  const bf16x8* frag;
  const float* P;             // FP32 parameters (biases are read from here)
  MlpNorm norm;
  int n;                      // particles (or batch entries)
  // feature source: raw array (feat != nullptr; optional index list) or step fields
  const float* feat;          // [n_total][6]
  const uint32_t* index;      // batch entries -> rows of feat/targ (training)
  MlpFields fields;
  // outputs / inputs per mode
  float* out;                 // INFER: [n][2] predictions (raw units)
  const float* targ;          // TRAIN: [n_total][2] targets (raw units)
  float loss_scale;           // TRAIN: 1 / batch
  float* grad;                // TRAIN: [MLP_NP] accumulated dL/dP
  double* loss;               // TRAIN: sum of squared normalised errors
  const float* ybar;          // INPUT_GRAD: [n][2] dL/d(prediction)
  float* fbar;                // INPUT_GRAD: [n][6] dL/d(raw features)
  // INFER on the step's fields, out == nullptr: the warm starts written directly
  const float* dmask;         // d of the divergence warm start (nullptr: kappa_v untouched)
  float inv_dt;
  float* kappav;
  float* kpred;
  // INPUT_GRAD, fbar == nullptr: the feature adjoints applied directly
  LearnedAdjSink sink;
};

// LDS layout (bytes) per mode.  Layer outputs alternate between two
// particle-major buffers (layer 0 -> A, 1 -> B, 2 -> A); the padded input
// lives in B until layer 1 overwrites it.
struct MlpLds {
  static constexpr int PA = 0;                        // bf16 [64][LDH]
  static constexpr int PB = PA + 64 * LDH * 2;        // bf16 [64][LDH]; input bf16 [64][LD0] at the start
  static constexpr int YB = PB + 64 * LDH * 2;        // f32 [64][2]
  static constexpr int INFER_BYTES = YB + 64 * 2 * 4;
  static constexpr int Z = INFER_BYTES;               // f32 [3][16 tiles][64 lanes][4]
  static constexpr int DZ = Z + 3 * 64 * 64 * 4;      // bf16 [64][LDH] particle-major dz
  static constexpr int IG_BYTES = DZ + 64 * LDH * 2;
  static constexpr int DZN = IG_BYTES;                // bf16 [64][LDH] unit-major dz
  static constexpr int HN0 = DZN + 64 * LDH * 2;      // bf16 [16][LDH] unit-major h0
  static constexpr int HN1 = HN0 + 16 * LDH * 2;      // bf16 [64][LDH] x3 unit-major h1..h3
  static constexpr int DB = HN1 + 3 * 64 * LDH * 2;   // f32 [64] bias-gradient partials
  static constexpr int TRAIN_BYTES = DB + 64 * 4;
};

// SiLU with the hardware exp2 and reciprocal (v_exp_f32, v_rcp_f32, ~1 ulp):
// sigma(z) = 1 / (1 + 2^(-z log2 e)).
__device__ inline float sigmoid_fast(float z) {
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.44269504088896341f * z));
}
__device__ inline float silu(float z) { return z * sigmoid_fast(z); }
__device__ inline float silu_d(float z) {
  const float s = sigmoid_fast(z);
  return s * (1.f + z * (1.f - s));
}

template <int MODE>
__global__ void __launch_bounds__(64) k_mlp(MlpArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __bf16* h0 = reinterpret_cast<__bf16*>(smem + MlpLds::PB);
  __bf16* pa = reinterpret_cast<__bf16*>(smem + MlpLds::PA);
  __bf16* pb = reinterpret_cast<__bf16*>(smem + MlpLds::PB);
  float* yb = reinterpret_cast<float*>(smem + MlpLds::YB);
  float* zb = reinterpret_cast<float*>(smem + MlpLds::Z);
  __bf16* dz = reinterpret_cast<__bf16*>(smem + MlpLds::DZ);
  __bf16* dzn = reinterpret_cast<__bf16*>(smem + MlpLds::DZN);
  __bf16* hn0 = reinterpret_cast<__bf16*>(smem + MlpLds::HN0);
  __bf16* hnl = reinterpret_cast<__bf16*>(smem + MlpLds::HN1);  // unit-major layer k output at hnl + k * 64 * LDH
  float* dbs = reinterpret_cast<float*>(smem + MlpLds::DB);
  constexpr int HS = 64 * LDH;
  const int l = threadIdx.x;
  const int lr = l & 15, lg = l >> 4;
  const int p0 = blockIdx.x * 64;
  const int p = p0 + l;
  const bool valid = p < a.n;
  const MlpNorm& N = a.norm;

  // ---- input layer: normalised features, bf16, padded to 32 (and 16 unit-major rows)
  uint32_t row = 0;
  float fr[6];
  if (valid) {
    if (a.feat) {
      row = a.index ? a.index[p] : (uint32_t)p;
      for (int k = 0; k < 6; ++k) fr[k] = a.feat[(size_t)row * 6 + k];
    } else {
      row = (uint32_t)p;
      features_from_fields(a.fields, row, fr);
    }
  } else {
    for (int k = 0; k < 6; ++k) fr[k] = N.mu_in[k];
  }
  for (int k = 0; k < 32; ++k) {
    const float x = (k < 6) ? (fr[k] - N.mu_in[k]) * N.is_in[k] : 0.f;
    h0[l * LD0 + k] = (__bf16)x;
    if (MODE == MLP_TRAIN && k < 16) hn0[k * LDH + l] = (__bf16)x;
  }
  __syncthreads();

  // ---- forward: layers 0..2 (SiLU), layer 3 (linear)
#pragma unroll
  for (int L = 0; L < 4; ++L) {
    const __bf16* in = (L == 0) ? h0 : ((L & 1) ? pa : pb);
    __bf16* outp = (L & 1) ? pb : pa;
    const int ldi = (L == 0) ? LD0 : LDH;
    const int mtiles = ceil_div(MLP_O[L], 16), kt = ceil_div(MLP_I[L], 32);
    const bf16x8* Fw = a.frag + MLP_FW[L] * 64;
    const float* bias = a.P + MLP_OB[L];
    // the layer's weight fragments, loaded together once per wavefront
    bf16x8 Aw[4][2];
#pragma unroll
    for (int mt = 0; mt < mtiles; ++mt)
#pragma unroll
      for (int kc = 0; kc < kt; ++kc) Aw[mt][kc] = Fw[(mt * kt + kc) * 64 + l];
    float bv[4][4];  // the lane's biases: unit 16 mt + 4 lg + j
#pragma unroll
    for (int mt = 0; mt < mtiles; ++mt)
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[mt][j] = (16 * mt + 4 * lg + j < MLP_O[L]) ? bias[16 * mt + 4 * lg + j] : 0.f;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      bf16x8 Bf[2];  // this particle tile's B fragments, shared by every M tile
#pragma unroll
      for (int kc = 0; kc < kt; ++kc)
        Bf[kc] = *reinterpret_cast<const bf16x8*>(in + (16 * nt + lr) * ldi + 32 * kc + 8 * lg);
#pragma unroll
      for (int mt = 0; mt < mtiles; ++mt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < kt; ++kc) acc = mfma(Aw[mt][kc], Bf[kc], acc);
        const int pp = 16 * nt + lr;
        if (L < 3) {
          bf16x4 hv;
          for (int j = 0; j < 4; ++j) {
            const int m = 16 * mt + 4 * lg + j;
            const float z = acc[j] + bv[mt][j];
            if (MODE != MLP_INFER) zb[((L * 16 + nt * 4 + mt) * 64 + l) * 4 + j] = z;
            const __bf16 h = (__bf16)silu(z);
            hv[j] = h;
            if (MODE == MLP_TRAIN) hnl[L * HS + m * LDH + pp] = h;
          }
          *reinterpret_cast<bf16x4*>(outp + pp * LDH + 16 * mt + 4 * lg) = hv;
        } else if (lg == 0) {
          yb[pp * 2 + 0] = acc[0] + bv[0][0];
          yb[pp * 2 + 1] = acc[1] + bv[0][1];
        }
      }
    }
    __syncthreads();
  }
  const float y0 = yb[l * 2 + 0], y1 = yb[l * 2 + 1];  // normalised predictions of particle p

  if (MODE == MLP_INFER) {
    if (valid) {
      a.out[(size_t)p * 2 + 0] = y0 * N.sd_out[0] + N.mu_out[0];
      a.out[(size_t)p * 2 + 1] = y1 * N.sd_out[1] + N.mu_out[1];
    }
    return;
  }

  // ---- output adjoint (normalised units)
  float g0 = 0.f, g1 = 0.f;
  if (MODE == MLP_TRAIN) {
    double e = 0.0;
    if (valid) {
      const float t0 = (a.targ[(size_t)row * 2 + 0] - N.mu_out[0]) / N.sd_out[0];
      const float t1 = (a.targ[(size_t)row * 2 + 1] - N.mu_out[1]) / N.sd_out[1];
      const float e0 = y0 - t0, e1 = y1 - t1;
      g0 = 2.f * e0 * a.loss_scale;
      g1 = 2.f * e1 * a.loss_scale;
      e = (double)e0 * e0 + (double)e1 * e1;
    }
    e = wave_sum_d(e);
    if (l == 0) atomicAdd(a.loss, e);
  } else if (valid) {
    g0 = a.ybar[(size_t)p * 2 + 0] * N.sd_out[0];
    g1 = a.ybar[(size_t)p * 2 + 1] * N.sd_out[1];
  }
  for (int k = 0; k < 32; ++k) {
    const float g = (k == 0) ? g0 : (k == 1) ? g1 : 0.f;
    dz[l * LDH + k] = (__bf16)g;
    if (MODE == MLP_TRAIN && k < 16) dzn[k * LDH + l] = (__bf16)g;
  }
  if (MODE == MLP_TRAIN) {
    const double b0 = wave_sum_d((double)g0), b1 = wave_sum_d((double)g1);
    if (l == 0) { atomicAdd(&a.grad[MLP_OB[3] + 0], (float)b0); atomicAdd(&a.grad[MLP_OB[3] + 1], (float)b1); }
  }
  __syncthreads();

  // ---- backward through layers 3..1 (and to the input for INPUT_GRAD)
#pragma unroll
  for (int L = 3; L >= 0; --L) {
    const int O = MLP_O[L], I = MLP_I[L];
    if (MODE == MLP_TRAIN) {
      // dW_L = dZ_L (O x 64 particles) * H_{L-1}^T (64 particles x I)
      const __bf16* hprev = (L == 0) ? hn0 : hnl + (L - 1) * HS;
      const int mtiles = ceil_div(O, 16), ntiles = ceil_div(I, 16);
      for (int mt = 0; mt < mtiles; ++mt)
        for (int nt = 0; nt < ntiles; ++nt) {
          f32x4 acc = {0.f, 0.f, 0.f, 0.f};
          for (int kc = 0; kc < 2; ++kc) {
            const bf16x8 A = *reinterpret_cast<const bf16x8*>(dzn + (16 * mt + lr) * LDH + 32 * kc + 8 * lg);
            const bf16x8 B = *reinterpret_cast<const bf16x8*>(hprev + (16 * nt + lr) * LDH + 32 * kc + 8 * lg);
            acc = mfma(A, B, acc);
          }
          const int i = 16 * nt + lr;
          for (int j = 0; j < 4; ++j) {
            const int o = 16 * mt + 4 * lg + j;
            if (o < O && i < I) atomicAdd(&a.grad[MLP_OW[L] + o * I + i], acc[j]);
          }
        }
    }
    if (L == 0 && MODE == MLP_TRAIN) break;
    // dH_{L-1} = W_L^T (I x O) * dZ_L (O x 64 particles)
    const int mtiles = ceil_div(I, 16), kt = ceil_div(O, 32);
    const bf16x8* Ft = a.frag + MLP_FT[L] * 64;
    bf16x8 At[4][2];
#pragma unroll
    for (int mt = 0; mt < mtiles; ++mt)
#pragma unroll
      for (int kc = 0; kc < kt; ++kc) At[mt][kc] = Ft[(mt * kt + kc) * 64 + l];
    f32x4 dh[4][4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int mt = 0; mt < mtiles; ++mt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < kt; ++kc) {
          const bf16x8 A = At[mt][kc];
          const bf16x8 B = *reinterpret_cast<const bf16x8*>(dz + (16 * nt + lr) * LDH + 32 * kc + 8 * lg);
          acc = mfma(A, B, acc);
        }
        dh[nt][mt] = acc;
      }
    __syncthreads();  // every lane has read dz
    if (L == 0) {
      // input features: x = (f - mu) * is  ->  dL/df = dL/dx * is; lanes of group 0/1 hold units 0..7
      for (int nt = 0; nt < 4; ++nt)
        for (int j = 0; j < 4; ++j) {
          const int k = 4 * lg + j;
          const int pp = 16 * nt + lr;
          if (k < 6 && p0 + pp < a.n) a.fbar[(size_t)(p0 + pp) * 6 + k] = dh[nt][0][j] * N.is_in[k];
        }
      break;
    }
    // dZ_{L-1} = dH_{L-1} * silu'(Z_{L-1})
    if (MODE == MLP_TRAIN && l < 64) dbs[l] = 0.f;
    __syncthreads();
    for (int nt = 0; nt < 4; ++nt)
      for (int mt = 0; mt < MLP_HT; ++mt) {
        bf16x4 dv;
        const int pp = 16 * nt + lr;
        for (int j = 0; j < 4; ++j) {
          const int m = 16 * mt + 4 * lg + j;
          const float z = zb[(((L - 1) * 16 + nt * 4 + mt) * 64 + l) * 4 + j];
          const float g = (p0 + pp < a.n) ? dh[nt][mt][j] * silu_d(z) : 0.f;
          dv[j] = (__bf16)g;
          if (MODE == MLP_TRAIN) {
            dzn[m * LDH + pp] = (__bf16)g;
            atomicAdd(&dbs[m], g);
          }
        }
        *reinterpret_cast<bf16x4*>(dz + pp * LDH + 16 * mt + 4 * lg) = dv;
      }
    __syncthreads();
    if (MODE == MLP_TRAIN && l < MLP_O[L - 1]) atomicAdd(&a.grad[MLP_OB[L - 1] + l], dbs[l]);
  }
}

// ---------------------------------------------------------------------------
// Register-chained network: inference and the input gradient dL/d(features)
// for dL/d(predictions) (used by the exact adjoint through the learned warm
// start).  A workgroup stages the K-permuted fragments (22 forward, + 22
// transposed for the gradient) and the biases in LDS once and its wavefronts
// then loop over groups of T 16-particle tiles (grid-stride): each fragment
// read from LDS (ds_read_b128) feeds the MFMAs of all T tiles, which gives
// the MFMA pipe T independent accumulators per fragment.  A tile is carried
// through every layer in registers: the D tiles of a layer, rounded to bf16,
// are the next layer's B operand as they stand.  The gradient keeps silu'(z)
// of the three hidden layers (D layout, FP32) from the tile's forward sweep,
// so nothing is recomputed.
// ---------------------------------------------------------------------------
// the gradient's workgroup is twice as large, so eight wavefronts share one
// staged copy of its 44 fragments
template <bool IG> constexpr int chain_block() { return IG ? 512 : 256; }
// 16-particle tiles per group (measured on the 1M tank: two or four tiles per
// fragment read ran slower than one, the extra accumulators costing occupancy)
template <bool IG> constexpr int chain_tiles() { return 1; }
constexpr int CHAIN_BIAS = 3 * 64 + 16;        // biases of layers 0..3 (output layer zero-padded to a tile)
template <bool IG> struct ChainLds {
  static constexpr int NF = IG ? MLP_NFRAG : MLP_FT[3];  // chain fragments staged (forward ones first)
  static constexpr int FRAG = 0;
  static constexpr int BIAS = NF * 64 * 16;
  static constexpr int BYTES = BIAS + CHAIN_BIAS * 4;
};
__device__ constexpr int CHAIN_OB[4] = {0, 64, 128, 192};  // bias offsets in the staged copy

__device__ inline bf16x8 pack2(f32x4 lo, f32x4 hi) {
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 4; ++j) { v[j] = (__bf16)lo[j]; v[4 + j] = (__bf16)hi[j]; }
  return v;
}

template <bool IG>
__global__ void __launch_bounds__(chain_block<IG>()) k_mlp_chain(MlpArgs a) {
  constexpr int CB = chain_block<IG>();
  constexpr int T = chain_tiles<IG>();
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  using L = ChainLds<IG>;
  bf16x8* F = reinterpret_cast<bf16x8*>(smem + L::FRAG);
  float* bias = reinterpret_cast<float*>(smem + L::BIAS);
  for (int k = threadIdx.x; k < L::NF * 64; k += CB) F[k] = a.frag[MLP_NFRAG * 64 + k];
  for (int k = threadIdx.x; k < CHAIN_BIAS; k += CB) {
    const int layer = k < 192 ? k / 64 : 3;
    const int u = k - CHAIN_OB[layer];
    bias[k] = u < MLP_O[layer] ? a.P[MLP_OB[layer] + u] : 0.f;
  }
  __syncthreads();
  const int l = threadIdx.x & 63;
  const int lr = l & 15, lg = l >> 4;
  const MlpNorm& N = a.norm;
  const int groups = ceil_div(a.n, 16 * T);
  const int stride = gridDim.x * (CB / 64);
#pragma unroll 1
  for (int grp = blockIdx.x * (CB / 64) + (threadIdx.x >> 6); grp < groups; grp += stride) {
    // Fragments and biases are re-read from LDS for every group: the empty asm
    // hides the loop invariance of `o`, so they are not hoisted into registers
    // (the index stays an LDS address: ds_read_b128 with the fragment offset).
    int o = l;
    asm volatile("" : "+v"(o));
    auto frag = [&](int f) { return F[f * 64 + o]; };
    auto lbias = [&](int layer, int mt) {
      return *reinterpret_cast<const f32x4*>(bias + (o - l) + CHAIN_OB[layer] + 16 * mt + 4 * lg);
    };
    int p[T];
    bool valid[T];
    bf16x8 B[T][2];  // K chunks of the next product (the input uses one)
#pragma unroll
    for (int t = 0; t < T; ++t) {
      p[t] = grp * 16 * T + 16 * t + lr;
      valid[t] = p[t] < a.n;
      // ---- input: lane group 0 holds features 0..3, group 1 features 4, 5 (chain_unit(0, g, j))
      float x[6];
      if (valid[t]) {
        if (a.feat) {
          for (int k = 0; k < 6; ++k) x[k] = a.feat[(size_t)p[t] * 6 + k];
        } else {
          features_from_fields(a.fields, (uint32_t)p[t], x);
        }
        for (int k = 0; k < 6; ++k) x[k] = (x[k] - N.mu_in[k]) * N.is_in[k];
      } else {
        for (int k = 0; k < 6; ++k) x[k] = 0.f;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float v = (j < 4 && lg == 0) ? x[j] : (j < 2 && lg == 1) ? x[4 + j] : 0.f;
        B[t][0][j] = (__bf16)v;
        B[t][1][j] = (__bf16)0.f;
      }
    }
    // ---- hidden layers 0..2: z = W h + b, h' = silu(z)
    f32x4 sp[T][3][MLP_HT];  // silu'(z) per hidden layer (IG only)
#pragma unroll
    for (int layer = 0; layer < 3; ++layer) {
      const int kt = layer == 0 ? 1 : MLP_KC;
      f32x4 h[T][MLP_HT];
#pragma unroll
      for (int mt = 0; mt < MLP_HT; ++mt) {
        const f32x4 b0 = lbias(layer, mt);
#pragma unroll
        for (int t = 0; t < T; ++t) h[t][mt] = b0;
#pragma unroll
        for (int kc = 0; kc < 2; ++kc)
          if (kc < kt) {
            const bf16x8 A = frag(MLP_FW[layer] + mt * kt + kc);
#pragma unroll
            for (int t = 0; t < T; ++t) h[t][mt] = mfma(A, B[t][kc], h[t][mt]);
          }
      }
#pragma unroll
      for (int t = 0; t < T; ++t) {
#pragma unroll
        for (int mt = 0; mt < MLP_HT; ++mt)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float z = h[t][mt][j];
            const float sg = sigmoid_fast(z);
            const float hz = z * sg;
            h[t][mt][j] = hz;
            if (IG) sp[t][layer][mt][j] = fmaf(hz, 1.f - sg, sg);  // silu'(z) = s + z s (1 - s)
          }
#pragma unroll
        for (int c = 0; c < MLP_KC; ++c) B[t][c] = pack2(h[t][2 * c], h[t][2 * c + 1]);
      }
    }
    // ---- output layer (2 units, rows 0 and 1 of lane group 0)
    f32x4 y[T];
    {
      const f32x4 b3 = lbias(3, 0);
#pragma unroll
      for (int t = 0; t < T; ++t) y[t] = b3;
#pragma unroll
      for (int c = 0; c < MLP_KC; ++c) {
        const bf16x8 A = frag(MLP_FW[3] + c);
#pragma unroll
        for (int t = 0; t < T; ++t) y[t] = mfma(A, B[t][c], y[t]);
      }
    }
    if (!IG) {
#pragma unroll
      for (int t = 0; t < T; ++t) {
        if (!valid[t] || lg != 0) continue;
        const float yv = y[t][0] * N.sd_out[0] + N.mu_out[0], yp = y[t][1] * N.sd_out[1] + N.mu_out[1];
        if (a.out) {
          a.out[(size_t)p[t] * 2 + 0] = yv;
          a.out[(size_t)p[t] * 2 + 1] = yp;
        } else {
          // the learned warm starts: kappa_v = y_v / dt where d > 0 (warmstartDivergenceSolve's
          // mask; no divergence warm start: kappa_v untouched), kappa prediction kept for the
          // pressure warm start (applied where rho* > 1)
          if (a.dmask) a.kappav[p[t]] = (a.dmask[p[t]] > 0.f) ? yv * a.inv_dt : 0.f;
          a.kpred[p[t]] = yp;
        }
      }
      continue;
    }
    // ---- backward: dZ_3 = ybar * sd_out (normalised units), then dH = W^T dZ, dZ = dH * silu'(z)
#pragma unroll
    for (int t = 0; t < T; ++t) {
      float g0 = 0.f, g1 = 0.f;
      if (valid[t] && lg == 0) {
        g0 = a.ybar[(size_t)p[t] * 2 + 0] * N.sd_out[0];
        g1 = a.ybar[(size_t)p[t] * 2 + 1] * N.sd_out[1];
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        B[t][0][j] = (__bf16)(j == 0 ? g0 : j == 1 ? g1 : 0.f);
        B[t][1][j] = (__bf16)0.f;
      }
    }
#pragma unroll
    for (int layer = 3; layer >= 1; --layer) {
      const int kt = layer == 3 ? 1 : MLP_KC;
      f32x4 dz[T][MLP_HT];
#pragma unroll
      for (int mt = 0; mt < MLP_HT; ++mt) {
#pragma unroll
        for (int t = 0; t < T; ++t) dz[t][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < 2; ++kc)
          if (kc < kt) {
            const bf16x8 A = frag(MLP_FT[layer] + mt * kt + kc);
#pragma unroll
            for (int t = 0; t < T; ++t) dz[t][mt] = mfma(A, B[t][kc], dz[t][mt]);
          }
      }
#pragma unroll
      for (int t = 0; t < T; ++t) {
#pragma unroll
        for (int mt = 0; mt < MLP_HT; ++mt)
#pragma unroll
          for (int j = 0; j < 4; ++j) dz[t][mt][j] *= sp[t][layer - 1][mt][j];
#pragma unroll
        for (int c = 0; c < MLP_KC; ++c) B[t][c] = pack2(dz[t][2 * c], dz[t][2 * c + 1]);
      }
    }
    // dH_0 = W_1^T dZ_1 (features 0..3 in lane group 0, 4..5 in group 1): dL/df = dL/dx * is
#pragma unroll
    for (int t = 0; t < T; ++t) {
      f32x4 din = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < MLP_KC; ++c) din = mfma(frag(MLP_FT[0] + c), B[t][c], din);
      if (!valid[t] || lg >= 2) continue;
      const int pt = p[t];
      if (a.fbar) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = 4 * lg + j;
          if (k < 6) a.fbar[(size_t)pt * 6 + k] = din[j] * N.is_in[k];
        }
      } else {
        // the feature adjoints: rho/rho0 gets fbar[0], the divergence pass at v0 gets
        // dt fbar[1], the boundary density fbar[3], kappa / kappa_v of the previous step
        // fbar[4] / fbar[5] (plus the pass-through when the solve did not overwrite
        // them); lane group 0 holds features 0..3, group 1 features 4 and 5
        const LearnedAdjSink& S = a.sink;
        if (lg == 0) {
          S.rhob[pt] += din[0] * N.is_in[0];
          S.dbar[pt] = S.dt * (din[1] * N.is_in[1]);
          S.rbb[pt] = din[3] * N.is_in[3];
        } else {
          S.kb[pt] = din[0] * N.is_in[4] + (S.keep_kappa ? S.kb[pt] : 0.f);
          S.kvb[pt] = din[1] * N.is_in[5] + (S.keep_kappav ? S.kvb[pt] : 0.f);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Adam (Kingma & Ba 2015) on the FP32 master parameters.
// ---------------------------------------------------------------------------
__global__ void k_mlp_adam(float* __restrict__ P, float* __restrict__ G, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= MLP_NP) return;
  const float g = G[k];
  G[k] = 0.f;
  const float mk = b1 * m[k] + (1.f - b1) * g;
  const float vk = b2 * v[k] + (1.f - b2) * g * g;
  m[k] = mk;
  v[k] = vk;
  P[k] -= lr * (mk * c1) / (sqrtf(vk * c2) + eps);
}

// Random batch of dataset rows (SplitMix64-style hash of (iteration, slot)).
__global__ void k_mlp_batch(uint32_t* __restrict__ idx, int batch, uint64_t seed, uint32_t size) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(b + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  idx[b] = (uint32_t)(z % (uint64_t)size);
}

// ---------------------------------------------------------------------------
// Dataset capture.  At the start of a step, particles selected by a hash of
// (step, particle id) append their features and index; at the end of the step
// the converged stiffness (kappa_v, kappa after the solves) becomes the target.
// ---------------------------------------------------------------------------
__global__ void k_mlp_capture_features(MlpFields s, const uint32_t* __restrict__ id, int n, uint64_t salt,
                                       uint32_t threshold, float* __restrict__ feat, uint32_t* __restrict__ rows,
                                       uint32_t* __restrict__ count, uint32_t cap) {
  const uint32_t i = gid();
  if (i >= (uint32_t)n) return;
  uint64_t z = salt ^ (0x9E3779B97F4A7C15ull * (uint64_t)(id[i] + 1));
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  if ((uint32_t)(z >> 32) >= threshold) return;
  const uint32_t slot = atomicAdd(count, 1u);
  if (slot >= cap) return;
  float f[6];
  features_from_fields(s, i, f);
  for (int k = 0; k < 6; ++k) feat[(size_t)slot * 6 + k] = f[k];
  rows[slot] = i;
}

__global__ void k_mlp_capture_targets(const float* __restrict__ kappa, const float* __restrict__ kappav,
                                      const uint32_t* __restrict__ rows, const uint32_t* __restrict__ first,
                                      const uint32_t* __restrict__ count, uint32_t cap, float* __restrict__ targ) {
  const uint32_t s = *first + blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t end = min(*count, cap);
  if (s >= end) return;
  const uint32_t i = rows[s];
  targ[(size_t)s * 2 + 0] = kappav[i];
  targ[(size_t)s * 2 + 1] = kappa[i];
}

// ---------------------------------------------------------------------------
static inline unsigned nb64(int n) { return (unsigned)((n + 63) / 64); }

void launch_mlp_pack(const float* P, void* frag, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_NFRAG_ALL * 64 + 255) / 256), dim3(256), 0, s, P,
                     static_cast<bf16x8*>(frag));
}

size_t mlp_frag_bytes() { return sizeof(bf16x8) * (size_t)MLP_NFRAG_ALL * 64; }
// Grid of the chained network: enough wavefronts for every group of T tiles,
// at most as many workgroups as the device holds at once (each stages its
// fragments once and loops over groups).
template <bool IG> static inline unsigned nb_chain(int n) {
  static long resident = 0;
  if (resident == 0) {
    int dev = 0, cus = 256, per_cu = 0;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mlp_chain<IG>, chain_block<IG>(),
                                                     ChainLds<IG>::BYTES) != hipSuccess || per_cu < 1)
      per_cu = 1;
    resident = (long)per_cu * cus;
  }
  const long groups = (n + 16 * chain_tiles<IG>() - 1) / (16 * chain_tiles<IG>());
  const long need = (groups + chain_block<IG>() / 64 - 1) / (chain_block<IG>() / 64);
  return (unsigned)std::max(1L, std::min(need, resident));
}

static MlpArgs base_args(const void* frag, const float* P, const MlpNorm& norm, int n) {
  MlpArgs a{};
  a.frag = static_cast<const bf16x8*>(frag);
  a.P = P;
  a.norm = norm;
  a.n = n;
  return a;
}

void launch_mlp_infer_array(const void* frag, const float* P, const MlpNorm& norm, const float* feat, int n,
                            float* out, hipStream_t s) {
  if (n <= 0) return;
  MlpArgs a = base_args(frag, P, norm, n);
  a.feat = feat;
  a.out = out;
  hipLaunchKernelGGL(k_mlp_chain<false>, dim3(nb_chain<false>(n)), dim3(chain_block<false>()),
                     ChainLds<false>::BYTES, s, a);
}

void launch_mlp_train_batch(const void* frag, const float* P, const MlpNorm& norm, const float* feat,
                            const float* targ, const uint32_t* index, int batch, float* grad, double* loss,
                            hipStream_t s) {
  if (batch <= 0) return;
  MlpArgs a = base_args(frag, P, norm, batch);
  a.feat = feat;
  a.targ = targ;
  a.index = index;
  a.loss_scale = 1.f / (float)batch;
  a.grad = grad;
  a.loss = loss;
  hipLaunchKernelGGL(k_mlp<MLP_TRAIN>, dim3(nb64(batch)), dim3(64), MlpLds::TRAIN_BYTES, s, a);
}

void launch_mlp_input_grad(const void* frag, const float* P, const MlpNorm& norm, const float* feat,
                           const MlpFields* fields, int n, const float* ybar, float* fbar, hipStream_t s) {
  if (n <= 0) return;
  MlpArgs a = base_args(frag, P, norm, n);
  a.feat = feat;
  if (fields) a.fields = *fields;
  a.ybar = ybar;
  a.fbar = fbar;
  hipLaunchKernelGGL(k_mlp_chain<true>, dim3(nb_chain<true>(n)), dim3(chain_block<true>()),
                     ChainLds<true>::BYTES, s, a);
}

void launch_mlp_adam(float* P, float* G, float* m, float* v, float lr, int t, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f / (1.f - powf(b1, (float)t)), c2 = 1.f / (1.f - powf(b2, (float)t));
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, P, G, m, v, lr, b1, b2, 1e-8f, c1, c2);
}

void launch_mlp_batch(uint32_t* idx, int batch, uint64_t seed, uint32_t size, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_batch, dim3((batch + 255) / 256), dim3(256), 0, s, idx, batch, seed, size);
}

void launch_mlp_capture_features(const MlpFields& f, const uint32_t* id, int n, uint64_t salt, uint32_t threshold,
                                 float* feat, uint32_t* rows, uint32_t* count, uint32_t cap, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mlp_capture_features, dim3((n + DFR_BLOCK - 1) / DFR_BLOCK), dim3(DFR_BLOCK), 0, s, f, id, n,
                     salt, threshold, feat, rows, count, cap);
}

void launch_mlp_capture_targets(const float* kappa, const float* kappav, const uint32_t* rows, const uint32_t* first,
                                uint32_t max_new, const uint32_t* count, uint32_t cap, float* targ, hipStream_t s) {
  if (max_new == 0) return;
  hipLaunchKernelGGL(k_mlp_capture_targets, dim3((max_new + 255) / 256), dim3(256), 0, s, kappa, kappav, rows, first,
                     count, cap, targ);
}

void launch_mlp_infer_apply(const void* frag, const float* P, const MlpNorm& norm, const MlpFields& f, int n,
                            const float* d, float inv_dt, float* kappav, float* kpred, hipStream_t s) {
  if (n <= 0) return;
  MlpArgs a = base_args(frag, P, norm, n);
  a.fields = f;
  a.dmask = d;
  a.inv_dt = inv_dt;
  a.kappav = kappav;
  a.kpred = kpred;
  hipLaunchKernelGGL(k_mlp_chain<false>, dim3(nb_chain<false>(n)), dim3(chain_block<false>()),
                     ChainLds<false>::BYTES, s, a);
}

void launch_mlp_input_grad_apply(const void* frag, const float* P, const MlpNorm& norm, const MlpFields& f, int n,
                                 const float* ybar, const LearnedAdjSink& sink, hipStream_t s) {
  if (n <= 0) return;
  MlpArgs a = base_args(frag, P, norm, n);
  a.fields = f;
  a.ybar = ybar;
  a.sink = sink;
  hipLaunchKernelGGL(k_mlp_chain<true>, dim3(nb_chain<true>(n)), dim3(chain_block<true>()),
                     ChainLds<true>::BYTES, s, a);
}

void set_mlp_kernel_attributes() {
  (void)hipFuncSetAttribute((const void*)k_mlp<MLP_TRAIN>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            MlpLds::TRAIN_BYTES);
}

}  // namespace dfr
