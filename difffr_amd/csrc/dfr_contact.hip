// Rigid-rigid contact between the bodies' boundary samples, the penalty form of
// RigidContactSolver (SPlisHSPlasH/InterlinkedSPH/RigidContactSolver.cpp; the
// solver's constructor sets use_penalty_force, so solveRigidContact :347-366
// runs beforePenaltyInitialize :307-345 and solveRigidContactPenalty :419-555).
//
// Every boundary sample r of body R has a rest volume and density fixed by its
// own body (constructor, :140-170): vol0 = gamma / (W(0) + sum_own W) and
// density0 = vol0 W(0) + sum_own vol0_k W.  In a step, a sample of a dynamic
// body with a neighbour on another body has density vol0 W(0) + sum_all vol0_k W;
// where density / density0 > 1 it pushes its body with
//   f = -beta (density / density0 - 1) n - mu |f_n| normalized(v_rel),
//   n = x_r - (sum_own x_k W) / (sum_own W),
//   v_rel = v_r - (sum_other v_k W) / (sum_own W),
// and torque (x_r - x_R) x f; W is the cubic spline of radius
// rigidContactSupportRadiusFactor r, the neighbours those of the simulation's
// search (radius h).  Sample velocities are v_R + w_R x (x_r - x_R) after
// gravity (dynamic and animated bodies; zero for static ones).
//
// One wavefront per sample of a dynamic body gathers its candidates from the
// 27 cells of the boundary bins; the forces are then applied body by body in
// sample order (k_contact_apply), one Dynamic3dRigidBody::addForce/addTorque
// per contact sample as the reference calls them, in FP64.
#include "dfr_device.h"
#include "dfr_kernels.h"

namespace dfr {

constexpr int CW = 4;  // wavefronts per workgroup

// Candidate enumeration shared by the two kernels: the wave's 27 cells of
// boundary bins, lane t of each round visiting candidate base + t.
struct CellSpan {
  uint32_t st[27], pre[28];
};

__device__ inline void cell_spans(const GridDesc& g, const uint32_t* __restrict__ bstart, const float4& p, int lane,
                                  CellSpan& S) {
  if (lane < 27) {
    int cx, cy, cz;
    cell_coords(g, p.x, p.y, p.z, cx, cy, cz);
    const int x = cx + lane % 3 - 1, y = cy + (lane / 3) % 3 - 1, z = cz + lane / 9 - 1;
    uint32_t st = 0, cnt = 0;
    if (x >= 0 && y >= 0 && z >= 0 && x < g.n[0] && y < g.n[1] && z < g.n[2]) {
      const uint32_t c = cell_key(g, x, y, z);
      st = bstart[c];
      cnt = bstart[c + 1] - st;
    }
    S.st[lane] = st;
    S.pre[lane + 1] = cnt;
  }
  __syncthreads();
  if (lane == 0) {
    S.pre[0] = 0;
    for (int c = 1; c <= 27; ++c) S.pre[c] += S.pre[c - 1];
  }
  __syncthreads();
}

// Sorted slot of candidate t (S.pre[c] <= t < S.pre[c + 1]).
__device__ inline uint32_t span_slot(const CellSpan& S, uint32_t t) {
  int lo = 0, hi = 26;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (S.pre[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return S.st[lo] + (t - S.pre[lo]);
}

// Rest volume (PASS 0) and rest density (PASS 1) of every sample.
template <int PASS>
__global__ void __launch_bounds__(64 * CW)
k_contact_rest(const float4* __restrict__ bpos, const float4* __restrict__ bx0, int nb, GridDesc g,
               const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bperm, float h2, KernelConst Kc,
               float gamma, float* __restrict__ vol0, float* __restrict__ dens0) {
  __shared__ CellSpan SS[CW];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t j = blockIdx.x * CW + w;
  const bool valid = j < (uint32_t)nb;  // uniform over the wavefront
  const float4 p = valid ? bpos[j] : make_float4(0.f, 0.f, 0.f, 0.f);
  cell_spans(g, bstart, p, lane, SS[w]);
  if (!valid) return;
  const float own = bx0[j].w;
  const uint32_t total = SS[w].pre[27];
  float s = 0.f;
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t t = base + (uint32_t)lane;
    if (t < total) {
      const uint32_t k = bperm[span_slot(SS[w], t)];
      const float4 q = bpos[k];
      if (k != j && bx0[k].w == own && dist2_exact(p.x, p.y, p.z, q.x, q.y, q.z) < h2) {
        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
        const float W = cubic_W(Kc, sqrtf(rx * rx + ry * ry + rz * rz));
        s += PASS == 0 ? W : vol0[k] * W;
      }
    }
  }
  s = wave_sum_f(s);
  if (lane == 0) {
    if (PASS == 0) vol0[j] = gamma / (Kc.w0 + s);
    else dens0[j] = vol0[j] * Kc.w0 + s;
  }
}

// Sample velocity after gravity: v_R + g dt + w_R x (x - x_R) for dynamic and
// animated bodies, zero otherwise.
__device__ inline float3 sample_vel(const RigidDev& b, const float4& x, const double* gdt) {
  if (!(b.dynamic || b.animated)) return make_float3(0.f, 0.f, 0.f);
  const double vx = b.v[0] + (b.dynamic ? gdt[0] : 0.0), vy = b.v[1] + (b.dynamic ? gdt[1] : 0.0),
               vz = b.v[2] + (b.dynamic ? gdt[2] : 0.0);
  const double rx = (double)x.x - b.x[0], ry = (double)x.y - b.x[1], rz = (double)x.z - b.x[2];
  return make_float3((float)(vx + (b.w[1] * rz - b.w[2] * ry)), (float)(vy + (b.w[2] * rx - b.w[0] * rz)),
                     (float)(vz + (b.w[0] * ry - b.w[1] * rx)));
}

// Penalty force and torque of every sample of a dynamic body:
// cf[d] = (f, 1) where the sample pushes its body, (0, 0) otherwise; ct[d] = torque.
__global__ void __launch_bounds__(64 * CW)
k_contact_forces(const float4* __restrict__ bpos, const float4* __restrict__ bx0, const uint32_t* __restrict__ dyn_idx,
                 int nd, GridDesc g, const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bperm, float h2,
                 KernelConst Kc, ContactConst cc, const float* __restrict__ vol0, const float* __restrict__ dens0,
                 const RigidDev* __restrict__ rig, float4* __restrict__ cf, float4* __restrict__ ct) {
  __shared__ CellSpan SS[CW];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t d = blockIdx.x * CW + w;
  const bool valid = d < (uint32_t)nd;
  const uint32_t j = valid ? dyn_idx[d] : 0u;
  const float4 p = valid ? bpos[j] : make_float4(0.f, 0.f, 0.f, 0.f);
  cell_spans(g, bstart, p, lane, SS[w]);
  if (!valid) return;
  const int R = (int)bx0[j].w;
  const RigidDev& B = rig[R];
  if (!B.dynamic) {
    if (lane == 0) { cf[d] = make_float4(0.f, 0.f, 0.f, 0.f); ct[d] = make_float4(0.f, 0.f, 0.f, 0.f); }
    return;
  }
  const uint32_t total = SS[w].pre[27];
  float dens = 0.f, sw = 0.f, sx = 0.f, sy = 0.f, sz = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
  int other = 0;
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t t = base + (uint32_t)lane;
    if (t < total) {
      const uint32_t k = bperm[span_slot(SS[w], t)];
      const float4 q = bpos[k];
      if (k != j && dist2_exact(p.x, p.y, p.z, q.x, q.y, q.z) < h2) {
        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
        const float W = cubic_W(Kc, sqrtf(rx * rx + ry * ry + rz * rz));
        dens += vol0[k] * W;
        const int Rk = (int)bx0[k].w;
        if (Rk == R) {
          sx += rx * W; sy += ry * W; sz += rz * W; sw += W;  // offsets: n = sum (x_r - x_k) W / sum W
        } else {
          other = 1;
          const float3 vk = sample_vel(rig[Rk], q, cc.gdt);
          vx += vk.x * W; vy += vk.y * W; vz += vk.z * W;
        }
      }
    }
  }
  const bool contact = __ballot(other) != 0ull;
  dens = wave_sum_f(dens); sw = wave_sum_f(sw);
  sx = wave_sum_f(sx); sy = wave_sum_f(sy); sz = wave_sum_f(sz);
  vx = wave_sum_f(vx); vy = wave_sum_f(vy); vz = wave_sum_f(vz);
  if (lane != 0) return;
  float4 fo = make_float4(0.f, 0.f, 0.f, 0.f), to = make_float4(0.f, 0.f, 0.f, 0.f);
  if (contact) {
    const float ratio = (vol0[j] * Kc.w0 + dens) / dens0[j];
    if (ratio > 1.f) {
      const float isw = 1.f / sw;
      const float nx = sx * isw, ny = sy * isw, nz = sz * isw;
      const float3 vr = sample_vel(B, p, cc.gdt);
      float ux = vr.x - vx * isw, uy = vr.y - vy * isw, uz = vr.z - vz * isw;
      const float s = -cc.beta * fmaxf(ratio - 1.f, 0.f);
      const float fnx = s * nx, fny = s * ny, fnz = s * nz;
      const float u2 = ux * ux + uy * uy + uz * uz;
      if (u2 > 0.f) {  // Eigen's normalized() leaves a zero vector as it is
        const float iu = 1.f / sqrtf(u2);
        ux *= iu; uy *= iu; uz *= iu;
      }
      const float ff = -cc.mu * sqrtf(fnx * fnx + fny * fny + fnz * fnz);
      const float fx = fnx + ff * ux, fy = fny + ff * uy, fz = fnz + ff * uz;
      const float rx = (float)((double)p.x - B.x[0]), ry = (float)((double)p.y - B.x[1]),
                  rz = (float)((double)p.z - B.x[2]);
      fo = make_float4(fx, fy, fz, 1.f);
      to = make_float4(ry * fz - rz * fy, rz * fx - rx * fz, rx * fy - ry * fx, 0.f);
    }
  }
  cf[d] = fo;
  ct[d] = to;
}

// Gravity, then one addForce / addTorque per contact sample in sample order
// (Dynamic3dRigidBody.h:113-142, 228-233), one thread per body.
__global__ void k_contact_apply(ContactConst cc, int nbodies, const float4* __restrict__ bx0,
                                const uint32_t* __restrict__ dyn_idx, int nd, const float4* __restrict__ cf,
                                const float4* __restrict__ ct, RigidDev* __restrict__ rig) {
  const int bi = threadIdx.x;
  if (bi >= nbodies) return;
  RigidDev& b = rig[bi];
  if (!b.dynamic) return;
  const double dt = cc.dt;
  for (int k = 0; k < 3; ++k) b.v[k] += (cc.g[k] * b.mass) * (b.invMass * dt);
  for (int d = 0; d < nd; ++d) {
    if ((int)bx0[dyn_idx[d]].w != bi) continue;
    const float4 f = cf[d];
    if (f.w == 0.f) continue;
    const float4 t = ct[d];
    b.v[0] += (double)f.x * (b.invMass * dt);
    b.v[1] += (double)f.y * (b.invMass * dt);
    b.v[2] += (double)f.z * (b.invMass * dt);
    double rx = t.x, ry = t.y, rz = t.z;
    if (cc.gyroscopic) {
      const double Lx = b.I[0] * b.w[0] + b.I[1] * b.w[1] + b.I[2] * b.w[2];
      const double Ly = b.I[3] * b.w[0] + b.I[4] * b.w[1] + b.I[5] * b.w[2];
      const double Lz = b.I[6] * b.w[0] + b.I[7] * b.w[1] + b.I[8] * b.w[2];
      rx += Ly * b.w[2] - Lz * b.w[1];
      ry += Lz * b.w[0] - Lx * b.w[2];
      rz += Lx * b.w[1] - Ly * b.w[0];
    }
    const double dwx = b.Iinv[0] * rx + b.Iinv[1] * ry + b.Iinv[2] * rz;
    const double dwy = b.Iinv[3] * rx + b.Iinv[4] * ry + b.Iinv[5] * rz;
    const double dwz = b.Iinv[6] * rx + b.Iinv[7] * ry + b.Iinv[8] * rz;
    b.w[0] += dwx * dt; b.w[1] += dwy * dt; b.w[2] += dwz * dt;
  }
}

static inline unsigned waves_blocks(int n) { return (unsigned)((n + CW - 1) / CW); }

void launch_contact_rest(const BoundaryBufs& b, const GridDesc& g, const GridBufs& gb, float h2, const KernelConst& Kc,
                         float gamma, float* vol0, float* dens0, hipStream_t s) {
  if (b.n <= 0) return;
  hipLaunchKernelGGL(k_contact_rest<0>, dim3(waves_blocks(b.n)), dim3(64 * CW), 0, s, b.pos, b.x0, b.n, g, gb.bstart,
                     gb.bperm, h2, Kc, gamma, vol0, dens0);
  hipLaunchKernelGGL(k_contact_rest<1>, dim3(waves_blocks(b.n)), dim3(64 * CW), 0, s, b.pos, b.x0, b.n, g, gb.bstart,
                     gb.bperm, h2, Kc, gamma, vol0, dens0);
}

void launch_contact(const BoundaryBufs& b, const GridDesc& g, const GridBufs& gb, float h2, const KernelConst& Kc,
                    const ContactConst& cc, const float* vol0, const float* dens0, float4* cf, float4* ct,
                    RigidDev* rig, int nbodies, hipStream_t s) {
  if (b.nd > 0)
    hipLaunchKernelGGL(k_contact_forces, dim3(waves_blocks(b.nd)), dim3(64 * CW), 0, s, b.pos, b.x0, b.dyn_idx, b.nd,
                       g, gb.bstart, gb.bperm, h2, Kc, cc, vol0, dens0, rig, cf, ct);
  hipLaunchKernelGGL(k_contact_apply, dim3(1), dim3(64), 0, s, cc, nbodies, b.x0, b.dyn_idx, b.nd, cf, ct, rig);
}

}  // namespace dfr
