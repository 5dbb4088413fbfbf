// Reverse pass of one recorded DFSPH fluid-rigid step: the exact discrete
// adjoint of the forward step of dfr_api.hip (TimeStepDiffDFSPH::step,
// TimeStepDiffDFSPH.cpp:526-652, and the rigid update, RigidBody3dBoundarySimulator.cpp:278-344).
//
// The forward gather sums are differentiated in gather form: the fluid-fluid
// neighbour relation is symmetric (the predicate depends on |x_i - x_j| only),
// so the adjoint of "particle i sums over its neighbours j" is again a sum over
// the neighbours of i, with W and its derivatives odd/even in r.  Each reverse
// pass therefore runs on the same brick tiles as the forward passes: the halo's
// positions and the neighbours' adjoint fields are staged in LDS, the own
// particle's adjoint is written once.  Boundary samples of dynamic bodies
// collect their adjoints with FP64 atomics and are reduced per body into the
// rigid adjoint (FP64).  Clamps, the |k| > eps guards and the 20-neighbour rule
// are held at the values the forward step took.
//
// Kernel derivatives: gradW(r) = c r and gradGradW(r) = c I + e r r^T with
// e = c'(|r|) / |r|; cubic spline (SPHKernels.h:16-176): for q <= 1/2,
// c = l (3q - 2) / h^2, e = 3 l / (h^3 |r|); for 1/2 < q <= 1,
// c = -l (1-q)^2 / (h |r|), e = l / h (2 (1-q) / (h |r|^2) + (1-q)^2 / |r|^3).
#include "dfr_rigid_math.h"
#include "dfr_tile.h"

namespace dfr {

constexpr int RB = PASS_BLOCK;
constexpr int PTR = per_thread<RB, TILE_CAP>();

struct PairCE {
  float c, e;
};
// Outer part with g = l f^2 / (h |r|): c = -g, e = (2 l f / h^2 + g) / |r|^2.
// `on` = false gives c = e = 0 (a pair whose every term vanishes), folded into
// the select the support test needs anyway.
__device__ inline PairCE cubic_ce(const KernelConst& K, float r2, bool on = true) {
  const float irl = __builtin_amdgcn_rsqf(r2);
  const float q = r2 * irl * K.inv_h;
  const float f = 1.f - q;
  const bool inner = q <= 0.5f;
  const float cin = fmaf(q, 3.f * K.l_ih2, -2.f * K.l_ih2);
  const float g = (f * f) * (K.l_ih * irl);
  const float cout = -g;
  const float ein = (3.f * K.l_ih2 * K.inv_h) * irl;
  const float eout = fmaf(2.f * K.l_ih2, f, g) * (irl * irl);
  const bool valid = on && r2 > 1.0e-10f && q <= 1.f;
  PairCE o;
  o.c = valid ? (inner ? cin : cout) : 0.f;
  o.e = valid ? (inner ? ein : eout) : 0.f;
  return o;
}

__device__ inline void atomic_add3(double* p, float x, float y, float z) {
  atomicAdd(p, (double)x);
  atomicAdd(p + 1, (double)y);
  atomicAdd(p + 2, (double)z);
}

// ---------------------------------------------------------------------------
// Reverse of the Jacobi velocity update with stiffness field k
//   v_i += dt V_f sum_j [|k_i + k_j| > eps] (k_i + k_j) gradW_ij
//        + [|k_i| > eps] dt k_i sum_b V_b gradW_ib,   f_b = -m_f k_i V_b gradW_ib
// (TimeStepDiffDFSPH.cpp:1024-1055, :1101-1197, :1753-1784, :1826-1902), followed
// by the reverse of how k was formed for this update:
//   MODE 0: k = d alpha / dt          (divergence iteration; out = adjoint of delta)
//   MODE 1: k = (rho* - 1) alpha / dt^2, rho* = max(rho/rho0 + dt delta, 1)
//           (pressure iteration; out = adjoint of delta, rho_bar += adjoint of rho/rho0)
//   MODE 2: k = kappa_v warm start    (out = adjoint of the incoming kappa_v)
//   MODE 3: k = kappa warm start      (out = adjoint of the incoming kappa)
//   MODE 4/5: learned kappa_v / kappa warm start k = [d > 0] y_v / dt,
//           [rho* > 1] y_p / dt^2 (out[2i] / out[2i+1] = adjoint of the prediction)
// vbar is the adjoint of the updated velocity (unchanged: the update adds to v).
// LDS per slot: x, y, z, vbar (3), k.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ void __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_adj_update_t(AdjPass a, const float* __restrict__ k, const float4* __restrict__ vbar,
               const float* kacc, float kacc_scale, const float* __restrict__ bval,
               const float* __restrict__ kprev, const float* __restrict__ alpha, float* out,
               float* __restrict__ alpha_bar, float* __restrict__ rho_bar, float4* __restrict__ xbar) {
  __shared__ TileHdr<RB> T;
  extern __shared__ float4 lds4[];
  const StepConst& c = a.c;
  const uint32_t brick = xcd_block();
  const int cap = a.cap;
  BrickView B;
  if (!brick_open<RB>(a.g, a.fstart, a.thdr, brick, cap, T, B)) return;
  const uint32_t total = B.total;
  const bool in_lds = B.in_lds;
  float4* LA = lds4;        // x, y, z, k
  float4* LB = lds4 + cap;  // vbar
  const int t = threadIdx.x;
  if (in_lds) {
    uint32_t js[PTR];
    map_indices<RB>(a.tmap, brick, total, js);
    float4 q[PTR], w[PTR];
    float kk[PTR];
#pragma unroll
    for (int r = 0; r < PTR; ++r)
      if (t + r * RB < (int)total) { q[r] = a.x[js[r]]; w[r] = vbar[js[r]]; kk[r] = k[js[r]]; }
#pragma unroll
    for (int r = 0; r < PTR; ++r) {
      const uint32_t s = t + r * RB;
      if (s < total) {
        LA[s] = make_float4(q[r].x, q[r].y, q[r].z, kk[r]);
        LB[s] = make_float4(w[r].x, w[r].y, w[r].z, w[r].w);
      }
    }
  }
  __syncthreads();
  const size_t stride = (size_t)c.nf;
  for (uint32_t i = B.own_begin + t; i < B.own_end; i += RB) {
    const float ki = k[i];
    const float4 p = a.x[i];
    const float4 vb = vbar[i];
    const uint32_t cc = a.cnt[i];
    const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
    // fluid pairs, without the common factor dt V_f
    float kbf = 0.f, xfx = 0.f, xfy = 0.f, xfz = 0.f;
    // m = 0 for pairs with |k_i + k_j| <= eps and for slots that are not live
    auto pair = [&](float qx, float qy, float qz, float wx, float wy, float wz, float kj, bool live) {
      const float kSum = ki + kj;
      const float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
      const PairCE ce = cubic_ce(c.K, rx * rx + ry * ry + rz * rz, live && fabsf(kSum) > c.eps);
      const float ux = vb.x - wx, uy = vb.y - wy, uz = vb.z - wz;
      const float ru = rx * ux + ry * uy + rz * uz;
      kbf = fmaf(ce.c, ru, kbf);
      const float a = ce.c * kSum, b = (ce.e * ru) * kSum;
      xfx = fmaf(a, ux, fmaf(b, rx, xfx));
      xfy = fmaf(a, uy, fmaf(b, ry, xfy));
      xfz = fmaf(a, uz, fmaf(b, rz, xfz));
    };
    if (in_lds) {
      for_fluid_groups<2>(a.nbr8, stride, i, nF, [&](uint32_t s) { return Slot8{LA[s], LB[s]}; },
                          [&](const Slot8& q, bool live) {
                            pair(q.a.x, q.a.y, q.a.z, q.b.x, q.b.y, q.b.z, q.a.w, live);
                          });
    } else {
      for_fluid_slots_rolled(a.nbr8, stride, i, nF, [&](uint32_t s) {
        const uint32_t j = tile_global<RB>(T, s);
        const float4 q = a.x[j];
        const float4 w = vbar[j];
        pair(q.x, q.y, q.z, w.x, w.y, w.z, k[j], true);
      });
    }
    const float dtVf = c.dt * c.V_f;
    float kb = dtVf * kbf;
    float xbx = dtVf * xfx, xby = dtVf * xfy, xbz = dtVf * xfz;
    if (fabsf(ki) > c.eps) {
      auto load = [&](uint32_t j) { return a.bx[j]; };
      for_boundary_groups<2>(a.nbrB, stride, i, nB, load, [&](const float4& q, uint32_t j) {
        const float Vb = q.w;
        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
        const PairCE ce = cubic_ce(c.K, rx * rx + ry * ry + rz * rz);
        const float rv = rx * vb.x + ry * vb.y + rz * vb.z;
        kb += c.dt * Vb * ce.c * rv;
        const float hvx = ce.c * vb.x + ce.e * rv * rx, hvy = ce.c * vb.y + ce.e * rv * ry,
                    hvz = ce.c * vb.z + ce.e * rv * rz;
        const float s1 = c.dt * ki * Vb;
        xbx += s1 * hvx; xby += s1 * hvy; xbz += s1 * hvz;
        if ((int)j >= a.nb_static) {
          const AdjBodyDev& B = a.body[(int)a.bx0[j].w];
          const uint32_t sd = j - (uint32_t)a.nb_static;
          float sx = -s1 * hvx, sy = -s1 * hvy, sz = -s1 * hvz;
          // f = -m_f k_i V_b gradW;  F += f;  T += (x_b - x_rb) x f
          const float lvx = (float)((double)q.x - B.xr[0]), lvy = (float)((double)q.y - B.xr[1]),
                      lvz = (float)((double)q.z - B.xr[2]);
          const float Fx = (float)B.Fb[0], Fy = (float)B.Fb[1], Fz = (float)B.Fb[2];
          const float Tx = (float)B.Tb[0], Ty = (float)B.Tb[1], Tz = (float)B.Tb[2];
          const float fbx = Fx + (Ty * lvz - Tz * lvy), fby = Fy + (Tz * lvx - Tx * lvz),
                      fbz = Fz + (Tx * lvy - Ty * lvx);
          const float rf = rx * fbx + ry * fby + rz * fbz;
          kb -= c.m_f * Vb * ce.c * rf;
          const float s2 = c.m_f * ki * Vb;
          const float hfx = ce.c * fbx + ce.e * rf * rx, hfy = ce.c * fby + ce.e * rf * ry,
                      hfz = ce.c * fbz + ce.e * rf * rz;
          xbx -= s2 * hfx; xby -= s2 * hfy; xbz -= s2 * hfz;
          sx += s2 * hfx; sy += s2 * hfy; sz += s2 * hfz;
          const float fx = -s2 * ce.c * rx, fy = -s2 * ce.c * ry, fz = -s2 * ce.c * rz;
          sx += fy * Tz - fz * Ty; sy += fz * Tx - fx * Tz; sz += fx * Ty - fy * Tx;
          atomic_add3(a.sxb + 3 * (size_t)sd, sx, sy, sz);
        }
      });
    }
    if (kacc) kb += kacc_scale * kacc[i];
    float4 xo = xbar[i];
    xo.x += xbx; xo.y += xby; xo.z += xbz;
    xbar[i] = xo;
    const float b = bval[i];
    if (MODE == ADJ_PRESSURE_ITER) {
      const float invH2 = 1.f / (c.dt * c.dt);
      alpha_bar[i] += kb * (b - 1.f) * invH2;
      const float bb = kb * alpha[i] * invH2;
      if (b > 1.f) { rho_bar[i] += bb; out[i] = c.dt * bb; }
      else out[i] = 0.f;
    } else if (MODE == ADJ_DIV_ITER) {
      const float invH = 1.f / c.dt;
      alpha_bar[i] += kb * b * invH;
      out[i] = (b > 0.f) ? kb * alpha[i] * invH : 0.f;
    } else if (MODE == ADJ_PRESSURE_WARM) {
      const float invH2 = 1.f / (c.dt * c.dt);
      out[i] = (b > 1.f && kprev[i] > -0.00025f) ? kb * 0.5f * invH2 : 0.f;
    } else if (MODE == ADJ_PRESSURE_WARM_LEARNED) {
      const float invH2 = 1.f / (c.dt * c.dt);
      out[2 * (size_t)i + 1] = (b > 1.f) ? kb * invH2 : 0.f;
    } else if (MODE == ADJ_DIV_WARM_LEARNED) {
      out[2 * (size_t)i] = (b > 0.f) ? kb / c.dt : 0.f;
    } else {
      const float invH = 1.f / c.dt;
      out[i] = (b > 0.f && kprev[i] > -0.5f) ? kb * 0.5f * invH : 0.f;
    }
  }
}

// ---------------------------------------------------------------------------
// Reverse of delta_i = sum_j V_f (v_i - v_j).gradW_ij + sum_b V_b (v_i - v_b).gradW_ib
// (computeDensityChange :1977-2041 / computeDensityAdv :1926-1975) at velocity v:
//   vbar_i += V_f sum_j (d_i + d_j) gradW_ij + d_i sum_b V_b gradW_ib
//   xbar_i += V_f sum_j (d_i + d_j) H_ij (v_i - v_j) + d_i sum_b V_b H_ib (v_i - v_b)
// LDS per slot: x, y, z, v (3), d.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_adj_delta_t(AdjPass a, const float4* __restrict__ v, const float* __restrict__ dbar, float4* __restrict__ vbar,
              float4* __restrict__ xbar) {
  __shared__ TileHdr<RB> T;
  extern __shared__ float4 lds4[];
  const StepConst& c = a.c;
  const uint32_t brick = xcd_block();
  const int cap = a.cap;
  BrickView B;
  if (!brick_open<RB>(a.g, a.fstart, a.thdr, brick, cap, T, B)) return;
  const uint32_t total = B.total;
  const bool in_lds = B.in_lds;
  float4* LA = lds4;        // x, y, z, dbar
  float4* LB = lds4 + cap;  // v
  const int t = threadIdx.x;
  if (in_lds) {
    uint32_t js[PTR];
    map_indices<RB>(a.tmap, brick, total, js);
    float4 q[PTR], w[PTR];
    float dd[PTR];
#pragma unroll
    for (int r = 0; r < PTR; ++r)
      if (t + r * RB < (int)total) { q[r] = a.x[js[r]]; w[r] = v[js[r]]; dd[r] = dbar[js[r]]; }
#pragma unroll
    for (int r = 0; r < PTR; ++r) {
      const uint32_t s = t + r * RB;
      if (s < total) {
        LA[s] = make_float4(q[r].x, q[r].y, q[r].z, dd[r]);
        LB[s] = make_float4(w[r].x, w[r].y, w[r].z, w[r].w);
      }
    }
  }
  __syncthreads();
  const size_t stride = (size_t)c.nf;
  for (uint32_t i = B.own_begin + t; i < B.own_end; i += RB) {
    const float di = dbar[i];
    const float4 p = a.x[i];
    const float4 vi = v[i];
    const uint32_t cc = a.cnt[i];
    const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
    float gx = 0.f, gy = 0.f, gz = 0.f, hx = 0.f, hy = 0.f, hz = 0.f;  // without the factor V_f
    // slots that are not live get c = e = 0
    auto pair = [&](float qx, float qy, float qz, float wx, float wy, float wz, float dj, bool live) {
      const float s = di + dj;
      const float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
      const PairCE ce = cubic_ce(c.K, rx * rx + ry * ry + rz * rz, live);
      const float sc = s * ce.c;
      gx += sc * rx; gy += sc * ry; gz += sc * rz;
      const float ux = vi.x - wx, uy = vi.y - wy, uz = vi.z - wz;
      const float se = s * ce.e * (rx * ux + ry * uy + rz * uz);
      hx += sc * ux + se * rx; hy += sc * uy + se * ry; hz += sc * uz + se * rz;
    };
    if (in_lds) {
      for_fluid_groups<2>(a.nbr8, stride, i, nF, [&](uint32_t s) { return Slot8{LA[s], LB[s]}; },
                          [&](const Slot8& q, bool live) {
                            pair(q.a.x, q.a.y, q.a.z, q.b.x, q.b.y, q.b.z, q.a.w, live);
                          });
    } else {
      for_fluid_slots_rolled(a.nbr8, stride, i, nF, [&](uint32_t s) {
        const uint32_t j = tile_global<RB>(T, s);
        const float4 q = a.x[j];
        const float4 w = v[j];
        pair(q.x, q.y, q.z, w.x, w.y, w.z, dbar[j], true);
      });
    }
    float vbx = c.V_f * gx, vby = c.V_f * gy, vbz = c.V_f * gz;
    float xbx = c.V_f * hx, xby = c.V_f * hy, xbz = c.V_f * hz;
    if (di != 0.f) {
      auto load = [&](uint32_t j) { return Slot8{a.bx[j], a.bv[j]}; };
      for_boundary_groups<2>(a.nbrB, stride, i, nB, load, [&](const Slot8& s8, uint32_t j) {
        const float4 q = s8.a, vbj = s8.b;
        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
        const PairCE ce = cubic_ce(c.K, rx * rx + ry * ry + rz * rz);
        const float sc = di * q.w * ce.c;
        const float ux = vi.x - vbj.x, uy = vi.y - vbj.y, uz = vi.z - vbj.z;
        const float se = di * q.w * ce.e * (rx * ux + ry * uy + rz * uz);
        const float gbx = sc * rx, gby = sc * ry, gbz = sc * rz;
        const float tbx = sc * ux + se * rx, tby = sc * uy + se * ry, tbz = sc * uz + se * rz;
        vbx += gbx; vby += gby; vbz += gbz;
        xbx += tbx; xby += tby; xbz += tbz;
        if ((int)j >= a.nb_static) {
          const uint32_t sd = j - (uint32_t)a.nb_static;
          atomic_add3(a.svb + 3 * (size_t)sd, -gbx, -gby, -gbz);
          atomic_add3(a.sxb + 3 * (size_t)sd, -tbx, -tby, -tbz);
        }
      });
    }
    float4 vo = vbar[i];
    vo.x += vbx; vo.y += vby; vo.z += vbz;
    vbar[i] = vo;
    float4 xo = xbar[i];
    xo.x += xbx; xo.y += xby; xo.z += xbz;
    xbar[i] = xo;
  }
}

// ---------------------------------------------------------------------------
// Reverse of the density (TimeStep.cpp:147-202, as rho/rho0) and of the DFSPH
// factor alpha_i = -1/S_i, S_i = sum_j |g_ij|^2 + |G_i|^2, g_ij = -V_f gradW_ij,
// G_i = sum_j g_ij + sum_b g_ib (computeDFSPHFactor, :883-962), with
// Sb_i = alpha_bar_i alpha_i^2 and rb_i the adjoint of rho_i/rho0:
//   xbar_i += V_f sum_j (rb_i + rb_j) gradW_ij + rb_i sum_b V_b gradW_ib
//           + 2 V_f sum_j H_ij (-(Sb_i + Sb_j) g_ij - Sb_i G_i + Sb_j G_j)
//           - 2 Sb_i sum_b V_b H_ib G_i
// LDS per slot: x, y, z, Sb, rb, G (3).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(2, 8)))
k_adj_factor_t(AdjPass a, const float* __restrict__ alpha, const float* __restrict__ alpha_bar,
               const float* __restrict__ rho_bar, const float4* __restrict__ G, float4* __restrict__ xbar,
               const float* __restrict__ rbb) {
  __shared__ TileHdr<RB> T;
  extern __shared__ float4 lds4[];
  const StepConst& c = a.c;
  const uint32_t brick = xcd_block();
  const int cap = a.cap;
  BrickView B;
  if (!brick_open<RB>(a.g, a.fstart, a.thdr, brick, cap, T, B)) return;
  const uint32_t total = B.total;
  const bool in_lds = B.in_lds;
  float4* LA = lds4;        // x, y, z, Sb
  float4* LB = lds4 + cap;  // rb, G
  const int t = threadIdx.x;
  auto sbar = [&](uint32_t j) { const float al = alpha[j]; return alpha_bar[j] * al * al; };
  if (in_lds) {
    uint32_t js[PTR];
    map_indices<RB>(a.tmap, brick, total, js);
    float4 q[PTR], g[PTR];
    float sb[PTR], rb[PTR];
#pragma unroll
    for (int r = 0; r < PTR; ++r)
      if (t + r * RB < (int)total) { q[r] = a.x[js[r]]; g[r] = G[js[r]]; sb[r] = sbar(js[r]); rb[r] = rho_bar[js[r]]; }
#pragma unroll
    for (int r = 0; r < PTR; ++r) {
      const uint32_t s = t + r * RB;
      if (s < total) {
        LA[s] = make_float4(q[r].x, q[r].y, q[r].z, sb[r]);
        LB[s] = make_float4(rb[r], g[r].x, g[r].y, g[r].z);
      }
    }
  }
  __syncthreads();
  const size_t stride = (size_t)c.nf;
  for (uint32_t i = B.own_begin + t; i < B.own_end; i += RB) {
    const float Si = sbar(i), ri = rho_bar[i];
    const float4 p = a.x[i];
    const float4 Gi = G[i];
    const uint32_t cc = a.cnt[i];
    const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
    float dx = 0.f, dy = 0.f, dz = 0.f;  // density part, without V_f
    float fx = 0.f, fy = 0.f, fz = 0.f;  // factor part, without 2 V_f
    // slots that are not live get c = e = 0 (every term carries one of them)
    auto pair = [&](float qx, float qy, float qz, float Sj, float rj, float Gx, float Gy, float Gz, bool live) {
      const float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
      const PairCE ce = cubic_ce(c.K, rx * rx + ry * ry + rz * rz, live);
      const float sr = (ri + rj) * ce.c;
      dx += sr * rx; dy += sr * ry; dz += sr * rz;
      // u = (Sb_i + Sb_j) V_f gradW - Sb_i G_i + Sb_j G_j
      const float sg = (Si + Sj) * c.V_f * ce.c;
      const float ux = sg * rx - Si * Gi.x + Sj * Gx, uy = sg * ry - Si * Gi.y + Sj * Gy,
                  uz = sg * rz - Si * Gi.z + Sj * Gz;
      const float er = ce.e * (rx * ux + ry * uy + rz * uz);
      fx += ce.c * ux + er * rx; fy += ce.c * uy + er * ry; fz += ce.c * uz + er * rz;
    };
    if (in_lds) {
      for_fluid_groups<2>(a.nbr8, stride, i, nF, [&](uint32_t s) { return Slot8{LA[s], LB[s]}; },
                          [&](const Slot8& q, bool live) {
                            pair(q.a.x, q.a.y, q.a.z, q.a.w, q.b.x, q.b.y, q.b.z, q.b.w, live);
                          });
    } else {
      for_fluid_slots_rolled(a.nbr8, stride, i, nF, [&](uint32_t s) {
        const uint32_t j = tile_global<RB>(T, s);
        const float4 q = a.x[j];
        const float4 g = G[j];
        pair(q.x, q.y, q.z, sbar(j), rho_bar[j], g.x, g.y, g.z, true);
      });
    }
    const float tv = 2.f * c.V_f;
    float xbx = c.V_f * dx + tv * fx, xby = c.V_f * dy + tv * fy, xbz = c.V_f * dz + tv * fz;
    // rbb: adjoint of the boundary part of the density alone (a warm-start feature)
    const float rib = ri + (rbb ? rbb[i] : 0.f);
    if (Si != 0.f || rib != 0.f) {
      auto load = [&](uint32_t j) { return a.bx[j]; };
      for_boundary_groups<2>(a.nbrB, stride, i, nB, load, [&](const float4& q, uint32_t j) {
        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
        const PairCE ce = cubic_ce(c.K, rx * rx + ry * ry + rz * rz);
        const float sr = rib * q.w * ce.c;
        const float rG = rx * Gi.x + ry * Gi.y + rz * Gi.z;
        const float s2 = 2.f * Si * q.w;
        const float hx = s2 * (ce.c * Gi.x + ce.e * rG * rx), hy = s2 * (ce.c * Gi.y + ce.e * rG * ry),
                    hz = s2 * (ce.c * Gi.z + ce.e * rG * rz);
        const float tx = sr * rx - hx, ty = sr * ry - hy, tz = sr * rz - hz;
        xbx += tx; xby += ty; xbz += tz;
        if ((int)j >= a.nb_static) atomic_add3(a.sxb + 3 * (size_t)(j - (uint32_t)a.nb_static), -tx, -ty, -tz);
      });
    }
    float4 xo = xbar[i];
    xo.x += xbx; xo.y += xby; xo.z += xbz;
    xbar[i] = xo;
  }
}

__global__ void k_adj_position(int n, float dt, const float4* __restrict__ xbar, float4* __restrict__ vbar) {
  const uint32_t i = gid();
  if (i >= (uint32_t)n) return;
  const float4 x = xbar[i];
  float4 v = vbar[i];
  v.x += dt * x.x; v.y += dt * x.y; v.z += dt * x.z;
  vbar[i] = v;
}

// The step's counting sort gathered sorted[k] = previous[perm[k]]; its adjoint
// scatters back into the previous step's order.
__global__ void k_adj_unpermute(int n, const uint32_t* __restrict__ perm, const float4* __restrict__ xb,
                                const float4* __restrict__ vb, const float* __restrict__ kb,
                                const float* __restrict__ kvb, float4* __restrict__ xb2, float4* __restrict__ vb2,
                                float* __restrict__ kb2, float* __restrict__ kvb2) {
  const uint32_t k = gid();
  if (k >= (uint32_t)n) return;
  const uint32_t i = perm[k];
  xb2[i] = xb[k];
  vb2[i] = vb[k];
  kb2[i] = kb[k];
  kvb2[i] = kvb[k];
}

// Boundary transform at the start of the step (SimulatorBase.cpp:1827-1858):
// x_b = R(q) x0 + x_rb, v_b = omega x (R(q) x0) + v.  Per dynamic sample
//   xbar_rb += sx,  qbar += dR(q)x0/dq^T (sx + sv x omega),  vbar += sv,
//   omegabar += (x_b - x_rb) x sv,
// reduced per wavefront in FP64 and added to the body's adjoint.
__global__ void k_adj_boundary_reduce(const double* __restrict__ sxb, const double* __restrict__ svb,
                                      const float4* __restrict__ bx, const float4* __restrict__ bx0, int nb_static,
                                      int first, int n, AdjBodyDev* __restrict__ ab, int body) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double v[13];
#pragma unroll
  for (int m = 0; m < 13; ++m) v[m] = 0.0;
  AdjBodyDev& B = ab[body];
  if (k < n) {
    const int sd = first + k;
    const int j = nb_static + sd;
    const double sx[3] = {sxb[3 * sd], sxb[3 * sd + 1], sxb[3 * sd + 2]};
    const double sv[3] = {svb[3 * sd], svb[3 * sd + 1], svb[3 * sd + 2]};
    const double w[3] = {B.wpre[0], B.wpre[1], B.wpre[2]};
    const double y[3] = {sx[0] + (sv[1] * w[2] - sv[2] * w[1]), sx[1] + (sv[2] * w[0] - sv[0] * w[2]),
                         sx[2] + (sv[0] * w[1] - sv[1] * w[0])};
    const float4 x0 = bx0[j];
    const double pp[3] = {(double)x0.x, (double)x0.y, (double)x0.z};
    const double qw = B.q0[0], u[3] = {B.q0[1], B.q0[2], B.q0[3]};
    // d/dq of p + 2w (u x p) + 2 u x (u x p), applied transposed to y
    const double uxp[3] = {u[1] * pp[2] - u[2] * pp[1], u[2] * pp[0] - u[0] * pp[2], u[0] * pp[1] - u[1] * pp[0]};
    const double pxy[3] = {pp[1] * y[2] - pp[2] * y[1], pp[2] * y[0] - pp[0] * y[2], pp[0] * y[1] - pp[1] * y[0]};
    const double up = u[0] * pp[0] + u[1] * pp[1] + u[2] * pp[2];
    const double uy = u[0] * y[0] + u[1] * y[1] + u[2] * y[2];
    const double py = pp[0] * y[0] + pp[1] * y[1] + pp[2] * y[2];
    v[0] = sx[0]; v[1] = sx[1]; v[2] = sx[2];
    v[3] = 2.0 * (uxp[0] * y[0] + uxp[1] * y[1] + uxp[2] * y[2]);
    for (int m = 0; m < 3; ++m) v[4 + m] = 2.0 * qw * pxy[m] + 2.0 * up * y[m] + 2.0 * uy * pp[m] - 4.0 * py * u[m];
    v[7] = sv[0]; v[8] = sv[1]; v[9] = sv[2];
    const float4 xb = bx[j];
    const double rl[3] = {(double)xb.x - B.xr[0], (double)xb.y - B.xr[1], (double)xb.z - B.xr[2]};
    v[10] = rl[1] * sv[2] - rl[2] * sv[1];
    v[11] = rl[2] * sv[0] - rl[0] * sv[2];
    v[12] = rl[0] * sv[1] - rl[1] * sv[0];
  }
#pragma unroll
  for (int m = 0; m < 13; ++m) v[m] = wave_sum_d(v[m]);
  if ((threadIdx.x & 63) == 0) {
    for (int m = 0; m < 3; ++m) atomicAdd(&B.x[m], v[m]);
    for (int m = 0; m < 4; ++m) atomicAdd(&B.q[m], v[3 + m]);
    for (int m = 0; m < 3; ++m) atomicAdd(&B.v[m], v[7 + m]);
    for (int m = 0; m < 3; ++m) atomicAdd(&B.w[m], v[10 + m]);
  }
}

// ---------------------------------------------------------------------------
// Reverse of the rigid update of one step (k_rigid_step: v += dt F/m,
// omega += dt I^-1 ((I omega) x omega + T), v += dt g, x += dt v,
// q = normalize(q + dt/2 (0, omega) q), I = R I0 R^T), of the torque lever arm
// and of the release-window override; one thread per body, FP64.
// ab[b].x/q/v/w hold the adjoint after the step on entry and the adjoint of the
// state before the step's override on exit (boundary terms are added later).
// ---------------------------------------------------------------------------
__global__ void k_adj_rigid(RigidStepConst rc, AdjStepTable tb, const RigidDev* __restrict__ forced,
                            const RigidDev* __restrict__ pre, const RigidDev* __restrict__ post,
                            AdjBodyDev* __restrict__ ab, int nbodies) {
  const int bi = threadIdx.x;
  if (bi >= nbodies) return;
  AdjBodyDev& A = ab[bi];
  const RigidDev& b = forced[bi];
  A.dynamic = b.dynamic;
  for (int k = 0; k < 3; ++k) { A.xr[k] = b.x[k]; A.wpre[k] = pre[bi].w[k]; A.Fb[k] = 0.0; A.Tb[k] = 0.0; }
  for (int k = 0; k < 4; ++k) A.q0[k] = b.q[k];
  if (!b.dynamic) return;
  const bool anim = tb.animated[bi] != 0;
  const double dt = rc.dt;
  const M3d I = ld3(b.I), Iinv = ld3(b.Iinv), I0 = ld3(b.I0);
  const V3d w = {b.w[0], b.w[1], b.w[2]};
  const V3d Tq = {post[bi].Tlast[0], post[bi].Tlast[1], post[bi].Tlast[2]};
  const V3d Fq = {post[bi].Flast[0], post[bi].Flast[1], post[bi].Flast[2]};
  // forward recomputation
  V3d L = {0, 0, 0}, rhs = {0, 0, 0}, w1 = w;
  if (!anim) {
    L = mv(I, w);
    rhs = Tq;
    if (rc.gyroscopic) { const V3d lw = cross(L, w); rhs = {lw.x + Tq.x, lw.y + Tq.y, lw.z + Tq.z}; }
    const V3d dw = mv(Iinv, rhs);
    w1 = {w.x + dw.x * dt, w.y + dw.y * dt, w.z + dw.z * dt};
  }
  const double wq[4] = {0.0, w1.x, w1.y, w1.z};
  double d[4];
  quat_mul(wq, b.q, d);
  double qt[4];
  for (int k = 0; k < 4; ++k) qt[k] = b.q[k] + dt * 0.5 * d[k];
  const double nn = sqrt(qt[1] * qt[1] + qt[2] * qt[2] + qt[3] * qt[3] + qt[0] * qt[0]);
  double q1[4];
  for (int k = 0; k < 4; ++k) q1[k] = qt[k] / nn;
  // q1 = qt / |qt|
  const double dq = q1[0] * A.q[0] + q1[1] * A.q[1] + q1[2] * A.q[2] + q1[3] * A.q[3];
  double qtb[4];
  for (int k = 0; k < 4; ++k) qtb[k] = (A.q[k] - q1[k] * dq) / nn;
  // qt = q + dt/2 (0, w1) (x) q
  double qb[4] = {qtb[0], qtb[1], qtb[2], qtb[3]};
  const double hh = dt * 0.5;
  const double dbw = hh * qtb[0];
  const V3d dbv = {hh * qtb[1], hh * qtb[2], hh * qtb[3]};
  const V3d qv = {b.q[1], b.q[2], b.q[3]};
  const V3d qxd = cross(qv, dbv);
  V3d w1b = {A.w[0] - qv.x * dbw + dbv.x * b.q[0] + qxd.x, A.w[1] - qv.y * dbw + dbv.y * b.q[0] + qxd.y,
             A.w[2] - qv.z * dbw + dbv.z * b.q[0] + qxd.z};
  qb[0] += w1.x * dbv.x + w1.y * dbv.y + w1.z * dbv.z;
  const V3d dxw = cross(dbv, w1);
  qb[1] += -w1.x * dbw + dxw.x; qb[2] += -w1.y * dbw + dxw.y; qb[3] += -w1.z * dbw + dxw.z;
  // x1 = x + dt v2
  const V3d xbp = {A.x[0], A.x[1], A.x[2]};
  const V3d v2b = {A.v[0] + A.x[0] * dt, A.v[1] + A.x[1] * dt, A.v[2] + A.x[2] * dt};
  V3d wbp = w1b;
  M3d Ib = M3d::zero(), Iinvb = M3d::zero();
  if (!anim) {
    for (int k = 0; k < 3; ++k) A.Fb[k] = (&v2b.x)[k] * (b.invMass * dt);
    const V3d rhsb = mv(Iinv.tr(), V3d{w1b.x * dt, w1b.y * dt, w1b.z * dt});
    Iinvb = outer(w1b, rhs) * dt;
    A.Tb[0] = rhsb.x; A.Tb[1] = rhsb.y; A.Tb[2] = rhsb.z;
    if (rc.gyroscopic) {
      const V3d Lb = cross(w, rhsb);
      const V3d t1 = cross(rhsb, L);
      const V3d t2 = mv(I.tr(), Lb);
      wbp = {wbp.x + t1.x + t2.x, wbp.y + t1.y + t2.y, wbp.z + t1.z + t2.z};
      Ib = outer(Lb, w);
    }
  }
  // I = R I0 R^T, Iinv = I^-1, both functions of q
  const M3d Ibt = Ib - mm(mm(Iinv.tr(), Iinvb), Iinv.tr());
  const M3d R = quat_to_R(b.q);
  const M3d Rb = mm(mm(Ibt + Ibt.tr(), R), I0);
  {
    const double qw = b.q[0], x = b.q[1], y = b.q[2], z = b.q[3];
    qb[0] += 2.0 * (-z * Rb(0, 1) + y * Rb(0, 2) + z * Rb(1, 0) - x * Rb(1, 2) - y * Rb(2, 0) + x * Rb(2, 1));
    qb[1] += 2.0 * (y * Rb(0, 1) + z * Rb(0, 2) + y * Rb(1, 0) - qw * Rb(1, 2) + z * Rb(2, 0) + qw * Rb(2, 1)) -
             4.0 * x * (Rb(1, 1) + Rb(2, 2));
    qb[2] += 2.0 * (x * Rb(0, 1) + qw * Rb(0, 2) + x * Rb(1, 0) + z * Rb(1, 2) - qw * Rb(2, 0) + z * Rb(2, 1)) -
             4.0 * y * (Rb(0, 0) + Rb(2, 2));
    qb[3] += 2.0 * (-qw * Rb(0, 1) + x * Rb(0, 2) + qw * Rb(1, 0) + y * Rb(1, 2) + x * Rb(2, 0) + y * Rb(2, 1)) -
             4.0 * z * (Rb(0, 0) + Rb(1, 1));
  }
  // torque lever arm: T = sum (x_b - x_rb) x f  ->  xbar_rb -= F x Tbar
  const V3d Tbar = {A.Tb[0], A.Tb[1], A.Tb[2]};
  const V3d FxT = cross(Fq, Tbar);
  A.x[0] = xbp.x - FxT.x; A.x[1] = xbp.y - FxT.y; A.x[2] = xbp.z - FxT.z;
  for (int k = 0; k < 4; ++k) A.q[k] = qb[k];
  double vb[3] = {v2b.x, v2b.y, v2b.z}, wb[3] = {wbp.x, wbp.y, wbp.z};
  // release window: v, omega were overwritten by factor (v0, omega0)
  if (tb.ramp[bi]) {
    const double f = tb.factor[bi];
    for (int k = 0; k < 3; ++k) { A.v0[k] += f * vb[k]; A.w0[k] += f * wb[k]; vb[k] = 0.0; wb[k] = 0.0; }
  }
  for (int k = 0; k < 3; ++k) { A.v[k] = vb[k]; A.w[k] = wb[k]; }
}

// ---------------------------------------------------------------------------
static inline unsigned nblk(size_t n) { return (unsigned)((n + DFR_BLOCK - 1) / DFR_BLOCK); }
// dynamic LDS of the reverse passes: two 16-byte records per staged halo slot
static inline size_t shm_slots(const AdjPass& a, int floats) { return sizeof(float) * (size_t)a.cap * floats; }

void launch_adj_update(const AdjPass& a, int mode, const float* k, const float4* vbar, const float* kacc,
                       float kacc_scale, const float* bval, const float* kprev, const float* alpha, float* out,
                       float* alpha_bar, float* rho_bar, float4* xbar, hipStream_t s) {
  const dim3 grid(nbricks(a.g)), blk(RB);
  const size_t SHM8 = shm_slots(a, 8);
  switch (mode) {
    case ADJ_DIV_ITER:
      hipLaunchKernelGGL(k_adj_update_t<ADJ_DIV_ITER>, grid, blk, SHM8, s, a, k, vbar, kacc, kacc_scale, bval, kprev,
                         alpha, out, alpha_bar, rho_bar, xbar);
      break;
    case ADJ_PRESSURE_ITER:
      hipLaunchKernelGGL(k_adj_update_t<ADJ_PRESSURE_ITER>, grid, blk, SHM8, s, a, k, vbar, kacc, kacc_scale, bval,
                         kprev, alpha, out, alpha_bar, rho_bar, xbar);
      break;
    case ADJ_DIV_WARM:
      hipLaunchKernelGGL(k_adj_update_t<ADJ_DIV_WARM>, grid, blk, SHM8, s, a, k, vbar, kacc, kacc_scale, bval, kprev,
                         alpha, out, alpha_bar, rho_bar, xbar);
      break;
    case ADJ_PRESSURE_WARM:
      hipLaunchKernelGGL(k_adj_update_t<ADJ_PRESSURE_WARM>, grid, blk, SHM8, s, a, k, vbar, kacc, kacc_scale, bval,
                         kprev, alpha, out, alpha_bar, rho_bar, xbar);
      break;
    case ADJ_DIV_WARM_LEARNED:
      hipLaunchKernelGGL(k_adj_update_t<ADJ_DIV_WARM_LEARNED>, grid, blk, SHM8, s, a, k, vbar, kacc, kacc_scale, bval,
                         kprev, alpha, out, alpha_bar, rho_bar, xbar);
      break;
    default:
      hipLaunchKernelGGL(k_adj_update_t<ADJ_PRESSURE_WARM_LEARNED>, grid, blk, SHM8, s, a, k, vbar, kacc, kacc_scale,
                         bval, kprev, alpha, out, alpha_bar, rho_bar, xbar);
  }
}

// ---------------------------------------------------------------------------
// Reverse of the advection with the step's viscosity, v' = v + dt (g + a_i),
// at the velocities after the divergence solve (Viscosity_Standard.cpp:262-268,
// Viscosity_XSPH.cpp:62-68).  abar = dt vbar' of every particle (staged apart,
// since vbar itself is updated); per particle, over its fluid neighbours j with
// r = x_i - x_j, u = v_i - v_j, G = c r = gradW, H = c I + e r r^T:
//   Standard (c_k = 10 nu m / rho_k, D = |r|^2 + 0.01 h^2, q = u.r,
//   b = c_j abar_i - c_i abar_j):
//     vbar_i += (b.G / D) r
//     xbar_i += (b.G / D) u + (q / D) H b - 2 q (b.G) / D^2 r
//     rhobar_i += rho0 c_i q (abar_j.G) / (D rho_i)
//   XSPH (s = nu / dt, b = (m / rho_j) abar_i - (m / rho_i) abar_j):
//     vbar_i += -s W b,  xbar_i += -s (b.u) G,  rhobar_i += -rho0 s W (m / rho_i^2) (abar_j.u)
// rhobar is the adjoint of rho / rho0.  LDS per slot: x, y, z, rho | v, abar_x | abar_y, abar_z.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ void __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_adj_viscosity_t(AdjPass a, const float4* __restrict__ v, const float* __restrict__ rho,
                  const float4* __restrict__ abar, ForceConst f, float4* __restrict__ vbar, float4* __restrict__ xbar,
                  float* __restrict__ rho_bar) {
  __shared__ TileHdr<RB> T;
  extern __shared__ float4 lds4[];
  const StepConst& c = a.c;
  const uint32_t brick = xcd_block();
  const int cap = a.cap;
  BrickView B;
  if (!brick_open<RB>(a.g, a.fstart, a.thdr, brick, cap, T, B)) return;
  const uint32_t total = B.total;
  const bool in_lds = B.in_lds;
  float4* LA = lds4;
  float4* LB = lds4 + cap;
  float2* LC = reinterpret_cast<float2*>(lds4 + 2 * cap);
  const int t = threadIdx.x;
  if (in_lds) {
    uint32_t js[PTR];
    map_indices<RB>(a.tmap, brick, total, js);
    float4 q[PTR], w[PTR], ab[PTR];
    float rr[PTR];
#pragma unroll
    for (int r = 0; r < PTR; ++r)
      if (t + r * RB < (int)total) { q[r] = a.x[js[r]]; w[r] = v[js[r]]; ab[r] = abar[js[r]]; rr[r] = rho[js[r]]; }
#pragma unroll
    for (int r = 0; r < PTR; ++r) {
      const uint32_t s = t + r * RB;
      if (s < total) {
        LA[s] = make_float4(q[r].x, q[r].y, q[r].z, rr[r]);
        LB[s] = make_float4(w[r].x, w[r].y, w[r].z, ab[r].x);
        LC[s] = make_float2(ab[r].y, ab[r].z);
      }
    }
  }
  __syncthreads();
  const size_t stride = (size_t)c.nf;
  const float eps = 0.01f * (f.h * f.h);
  const float sx = f.visc / c.dt;  // XSPH's nu / dt
  for (uint32_t i = B.own_begin + t; i < B.own_end; i += RB) {
    const float4 p = a.x[i];
    const float4 vi = v[i];
    const float4 ai = abar[i];
    const float ri = rho[i];
    const float ci = (MODE == 1 ? 10.f * f.visc : 1.f) * (c.m_f / ri);
    const uint32_t nF = a.cnt[i] & 0xffffu;
    float vbx = 0.f, vby = 0.f, vbz = 0.f, xbx = 0.f, xby = 0.f, xbz = 0.f, rb = 0.f;
    auto pair = [&](float qx, float qy, float qz, float rj, float wx, float wy, float wz, float ajx, float ajy,
                    float ajz) {
      const float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
      const float ux = vi.x - wx, uy = vi.y - wy, uz = vi.z - wz;
      const float r2 = rx * rx + ry * ry + rz * rz;
      const float cj = (MODE == 1 ? 10.f * f.visc : 1.f) * (c.m_f / rj);
      const float bx = cj * ai.x - ci * ajx, by = cj * ai.y - ci * ajy, bz = cj * ai.z - ci * ajz;
      if (MODE == 1) {
        const PairCE ce = cubic_ce(c.K, r2);
        const float iD = 1.f / (r2 + eps);
        const float qd = (ux * rx + uy * ry + uz * rz) * iD;       // q / D
        const float br = bx * rx + by * ry + bz * rz;
        const float bG = ce.c * br * iD;                           // b.G / D
        vbx += bG * rx; vby += bG * ry; vbz += bG * rz;
        const float sr = qd * ce.e * br - 2.f * qd * bG;           // r coefficient of (q/D) H b - 2 q b.G / D^2 r
        const float sb = qd * ce.c;
        xbx += bG * ux + sb * bx + sr * rx;
        xby += bG * uy + sb * by + sr * ry;
        xbz += bG * uz + sb * bz + sr * rz;
        rb += ci * qd * ce.c * (ajx * rx + ajy * ry + ajz * rz);
      } else {
        const float W = cubic_W(c.K, sqrtf(r2));
        const float sw = -sx * W;
        vbx += sw * bx; vby += sw * by; vbz += sw * bz;
        const float gc = -sx * (bx * ux + by * uy + bz * uz) * cubic_grad_c(c.K, r2);
        xbx += gc * rx; xby += gc * ry; xbz += gc * rz;
        rb += sw * ci * (ajx * ux + ajy * uy + ajz * uz);
      }
    };
    if (in_lds) {
      for_fluid_slots(a.nbr8, stride, i, nF, [&](uint32_t s) {
        const float4 A = LA[s], Bv = LB[s];
        const float2 Cn = LC[s];
        pair(A.x, A.y, A.z, A.w, Bv.x, Bv.y, Bv.z, Bv.w, Cn.x, Cn.y);
      });
    } else {
      for_fluid_slots_rolled(a.nbr8, stride, i, nF, [&](uint32_t s) {
        const uint32_t j = tile_global<RB>(T, s);
        const float4 q = a.x[j], w = v[j], aj = abar[j];
        pair(q.x, q.y, q.z, rho[j], w.x, w.y, w.z, aj.x, aj.y, aj.z);
      });
    }
    float4 vo = vbar[i];
    vo.x += vbx; vo.y += vby; vo.z += vbz;
    vbar[i] = vo;
    float4 xo = xbar[i];
    xo.x += xbx; xo.y += xby; xo.z += xbz;
    xbar[i] = xo;
    rho_bar[i] += c.rho0 * rb / ri;
  }
}

// abar = dt vbar (the adjoint of the acceleration of the advection).
__global__ void k_adj_scale4(int n, float dt, const float4* __restrict__ vbar, float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint32_t)n) return;
  const float4 w = vbar[i];
  out[i] = make_float4(dt * w.x, dt * w.y, dt * w.z, 0.f);
}

void launch_adj_delta(const AdjPass& a, const float4* v, const float* dbar, float4* vbar, float4* xbar,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_adj_delta_t, dim3(nbricks(a.g)), dim3(RB), shm_slots(a, 8), s, a, v, dbar, vbar, xbar);
}

void launch_adj_viscosity(const AdjPass& a, int mode, const float4* v, const float* rho, const float4* vbar_in,
                          float4* abar, const ForceConst& fc, float4* vbar, float4* xbar, float* rho_bar,
                          hipStream_t s) {
  if (a.c.nf <= 0) return;
  hipLaunchKernelGGL(k_adj_scale4, dim3(nblk(a.c.nf)), dim3(DFR_BLOCK), 0, s, a.c.nf, a.c.dt, vbar_in, abar);
  const size_t shm = (2 * sizeof(float4) + sizeof(float2)) * (size_t)a.cap;
  if (mode == 1)
    hipLaunchKernelGGL(k_adj_viscosity_t<1>, dim3(nbricks(a.g)), dim3(RB), shm, s, a, v, rho, abar, fc, vbar, xbar,
                       rho_bar);
  else
    hipLaunchKernelGGL(k_adj_viscosity_t<2>, dim3(nbricks(a.g)), dim3(RB), shm, s, a, v, rho, abar, fc, vbar, xbar,
                       rho_bar);
}

void launch_adj_factor(const AdjPass& a, const float* alpha, const float* alpha_bar, const float* rho_bar,
                       const float4* G, float4* xbar, const float* rbb, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_factor_t, dim3(nbricks(a.g)), dim3(RB), shm_slots(a, 8), s, a, alpha, alpha_bar, rho_bar, G,
                     xbar, rbb);
}

void launch_adj_position(int n, float dt, const float4* xbar, float4* vbar, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_adj_position, dim3(nblk(n)), dim3(DFR_BLOCK), 0, s, n, dt, xbar, vbar);
}

void launch_adj_unpermute(int n, const uint32_t* perm, const float4* xb, const float4* vb, const float* kb,
                          const float* kvb, float4* xb2, float4* vb2, float* kb2, float* kvb2, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_adj_unpermute, dim3(nblk(n)), dim3(DFR_BLOCK), 0, s, n, perm, xb, vb, kb, kvb, xb2, vb2, kb2,
                     kvb2);
}

void launch_adj_boundary_reduce(const AdjPass& a, int body, int first, int n, AdjBodyDev* ab, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_adj_boundary_reduce, dim3(nblk(n)), dim3(DFR_BLOCK), 0, s, a.sxb, a.svb, a.bx, a.bx0,
                     a.nb_static, first, n, ab, body);
}

void launch_adj_rigid(const RigidStepConst& rc, const AdjStepTable& t, const RigidDev* forced, const RigidDev* pre,
                      const RigidDev* post, AdjBodyDev* ab, int nbodies, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_rigid, dim3(1), dim3(64), 0, s, rc, t, forced, pre, post, ab, nbodies);
}

void set_adjoint_kernel_attributes() {
  const int s10 = (int)((2 * sizeof(float4) + sizeof(float2)) * TILE_CAP);
  (void)hipFuncSetAttribute((const void*)k_adj_viscosity_t<1>, hipFuncAttributeMaxDynamicSharedMemorySize, s10);
  (void)hipFuncSetAttribute((const void*)k_adj_viscosity_t<2>, hipFuncAttributeMaxDynamicSharedMemorySize, s10);
  const int s8 = (int)(sizeof(float) * TILE_CAP * 8);
  (void)hipFuncSetAttribute((const void*)k_adj_update_t<ADJ_DIV_ITER>, hipFuncAttributeMaxDynamicSharedMemorySize, s8);
  (void)hipFuncSetAttribute((const void*)k_adj_update_t<ADJ_PRESSURE_ITER>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            s8);
  (void)hipFuncSetAttribute((const void*)k_adj_update_t<ADJ_DIV_WARM>, hipFuncAttributeMaxDynamicSharedMemorySize, s8);
  (void)hipFuncSetAttribute((const void*)k_adj_update_t<ADJ_PRESSURE_WARM>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            s8);
  (void)hipFuncSetAttribute((const void*)k_adj_update_t<ADJ_DIV_WARM_LEARNED>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, s8);
  (void)hipFuncSetAttribute((const void*)k_adj_update_t<ADJ_PRESSURE_WARM_LEARNED>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, s8);
  (void)hipFuncSetAttribute((const void*)k_adj_delta_t, hipFuncAttributeMaxDynamicSharedMemorySize, s8);
  (void)hipFuncSetAttribute((const void*)k_adj_factor_t, hipFuncAttributeMaxDynamicSharedMemorySize, s8);
}

}  // namespace dfr
