// Gather passes of the DFSPH solver: density + factor, the Jacobi velocity
// update with Akinci coupling, the density recompute + residual, warm starts,
// advection and position update.  One thread per fluid particle, neighbours
// read from the transposed ELL lists (coalesced index stream), particle data in
// cell order so neighbour gathers hit L2.
#include "dfr_kernels.h"

namespace dfr {

__device__ inline uint32_t xcd_block2() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t per = nb / 8u, rem = nb % 8u;
  const uint32_t x = b % 8u, k = b / 8u;
  return x * per + min(x, rem) + k;
}
__device__ inline uint32_t pid() { return xcd_block2() * blockDim.x + threadIdx.x; }

// Block-level FP64 sum of one value per thread, atomically added to *out.
__device__ inline void block_add_double(double v, double* out) {
  __shared__ double part[DFR_BLOCK / 64];
  v = wave_sum_d(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < DFR_BLOCK / 64; ++w) t += part[w];
    if (t != 0.0) atomicAdd(out, t);
  }
}

// TimeStep::computeDensities (TimeStep.cpp:147-202) fused with
// TimeStepDiffDFSPH::computeDFSPHFactor (TimeStepDiffDFSPH.cpp:883-962) and, when
// the divergence solver runs, with the first loop of warmstartDivergenceSolve
// (:1720-1732: computeDensityChange + kappa_v warm start).  All read the same
// neighbour positions, so one pass produces rho, alpha, sum_grad_p_k and d.
__global__ void __launch_bounds__(DFR_BLOCK)
k_density_factor(StepConst c, const float4* __restrict__ pos, const float4* __restrict__ vel,
                 const float4* __restrict__ bpos, const float4* __restrict__ bvel, const uint32_t* __restrict__ nbr,
                 const uint32_t* __restrict__ cnt, float* __restrict__ rho, float* __restrict__ alpha,
                 float4* __restrict__ s_out, float* __restrict__ badv, float* __restrict__ kappav, int div_warm,
                 int divergence_on) {
  const uint32_t i = pid();
  if (i >= (uint32_t)c.nf) return;
  const float4 p = pos[i];
  float4 vi = make_float4(0.f, 0.f, 0.f, 0.f);
  if (divergence_on) vi = vel[i];
  const uint32_t cc = cnt[i];
  const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
  const size_t stride = (size_t)c.nf;
  float dens = c.V_f * c.K.w0;
  float sumsq = 0.f, gx = 0.f, gy = 0.f, gz = 0.f, delta = 0.f;
  for (uint32_t k = 0; k < nF; ++k) {
    const uint32_t j = nbr[(size_t)k * stride + i];
    const float4 q = pos[j];
    const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
    const float rl = sqrtf(rx * rx + ry * ry + rz * rz);
    dens += c.V_f * cubic_W(c.K, rl);
    const float sc = cubic_gradW_scale(c.K, rl);
    const float wx = sc * rx, wy = sc * ry, wz = sc * rz;
    const float px = -c.V_f * wx, py = -c.V_f * wy, pz = -c.V_f * wz;
    sumsq += px * px + py * py + pz * pz;
    gx -= px; gy -= py; gz -= pz;
    if (divergence_on) {
      const float4 vj = vel[j];
      delta += c.V_f * ((vi.x - vj.x) * wx + (vi.y - vj.y) * wy + (vi.z - vj.z) * wz);
    }
  }
  for (uint32_t k = nF; k < nF + nB; ++k) {
    const uint32_t j = nbr[(size_t)k * stride + i];
    const float4 q = bpos[j];
    const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
    const float rl = sqrtf(rx * rx + ry * ry + rz * rz);
    dens += q.w * cubic_W(c.K, rl);
    const float sc = cubic_gradW_scale(c.K, rl);
    const float wx = sc * rx, wy = sc * ry, wz = sc * rz;
    gx += q.w * wx; gy += q.w * wy; gz += q.w * wz;
    if (divergence_on) {
      const float4 vb = bvel[j];
      delta += q.w * ((vi.x - vb.x) * wx + (vi.y - vb.y) * wy + (vi.z - vb.z) * wz);
    }
  }
  rho[i] = dens * c.rho0;
  s_out[i] = make_float4(-gx, -gy, -gz, 0.f);
  sumsq += gx * gx + gy * gy + gz * gz;
  alpha[i] = (sumsq > c.eps) ? -1.f / sumsq : 0.f;
  if (divergence_on) {
    float d = fmaxf(delta, 0.f);
    if (nF + nB < 20u) d = 0.f;
    badv[i] = d;
    if (div_warm) {
      const float invH = 1.f / c.dt;
      kappav[i] = (d > 0.f) ? 0.5f * fmaxf(kappav[i], -0.5f) * invH : 0.f;
    }
  }
}

// Jacobi velocity update shared by the two warm starts and every iteration of
// both solves (TimeStepDiffDFSPH.cpp:1024-1055, :1101-1197, :1753-1784,
// :1826-1902): v_i -= dt sum_j (k_i + k_j) grad_p_ij over fluid pairs with
// |k_i + k_j| > eps; when |k_i| > eps each boundary neighbour b changes v_i by
// -dt k_i grad_p_ib and receives f = -m_i dv / dt (BoundaryModel::addForce,
// BoundaryModel.h:46-58), summed here in FP64 per dynamic body.
__global__ void __launch_bounds__(DFR_BLOCK)
k_velocity_update(StepConst c, const float4* __restrict__ pos, float4* __restrict__ vel,
                  const float4* __restrict__ bpos, const float4* __restrict__ bx0, int nb_static,
                  const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ cnt, const float* __restrict__ kfield,
                  float* __restrict__ kappa_acc, RigidDev* __restrict__ rig) {
  const uint32_t i = pid();
  const bool active = i < (uint32_t)c.nf;
  int cb = -1;
  double F0 = 0, F1 = 0, F2 = 0, T0 = 0, T1 = 0, T2 = 0;
  if (active) {
    const float ki = kfield[i];
    if (kappa_acc) kappa_acc[i] += ki;
    const float4 p = pos[i];
    float4 vi = vel[i];
    const uint32_t cc = cnt[i];
    const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
    const size_t stride = (size_t)c.nf;
    for (uint32_t k = 0; k < nF; ++k) {
      const uint32_t j = nbr[(size_t)k * stride + i];
      const float kSum = ki + kfield[j];
      if (fabsf(kSum) > c.eps) {
        const float4 q = pos[j];
        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
        const float rl = sqrtf(rx * rx + ry * ry + rz * rz);
        const float sc = cubic_gradW_scale(c.K, rl);
        const float gpx = -c.V_f * (sc * rx), gpy = -c.V_f * (sc * ry), gpz = -c.V_f * (sc * rz);
        const float a = c.dt * kSum;
        vi.x -= a * gpx; vi.y -= a * gpy; vi.z -= a * gpz;
      }
    }
    if (fabsf(ki) > c.eps) {
      const float invH = 1.f / c.dt;
      for (uint32_t k = nF; k < nF + nB; ++k) {
        const uint32_t j = nbr[(size_t)k * stride + i];
        const float4 q = bpos[j];
        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
        const float rl = sqrtf(rx * rx + ry * ry + rz * rz);
        const float sc = cubic_gradW_scale(c.K, rl);
        const float gpx = -q.w * (sc * rx), gpy = -q.w * (sc * ry), gpz = -q.w * (sc * rz);
        const float a = -c.dt * 1.f * ki;
        const float dvx = a * gpx, dvy = a * gpy, dvz = a * gpz;
        vi.x += dvx; vi.y += dvy; vi.z += dvz;
        if ((int)j >= nb_static) {
          const int body = (int)bx0[j].w;
          if (rig[body].dynamic) {
            if (body != cb && cb >= 0) {
              atomicAdd(&rig[cb].F[0], F0); atomicAdd(&rig[cb].F[1], F1); atomicAdd(&rig[cb].F[2], F2);
              atomicAdd(&rig[cb].T[0], T0); atomicAdd(&rig[cb].T[1], T1); atomicAdd(&rig[cb].T[2], T2);
              F0 = F1 = F2 = T0 = T1 = T2 = 0.0;
            }
            cb = body;
            const float fx = -c.m_f * dvx * invH, fy = -c.m_f * dvy * invH, fz = -c.m_f * dvz * invH;
            const double rxd = (double)q.x - rig[body].x[0], ryd = (double)q.y - rig[body].x[1],
                         rzd = (double)q.z - rig[body].x[2];
            F0 += fx; F1 += fy; F2 += fz;
            T0 += ryd * fz - rzd * fy; T1 += rzd * fx - rxd * fz; T2 += rxd * fy - ryd * fx;
          }
        }
      }
    }
    vel[i] = vi;
  }
  // Wave-level FP64 reduction of the coupling partials, one body at a time.
  while (true) {
    const unsigned long long m = __ballot(cb >= 0);
    if (m == 0ull) break;
    const int src = __ffsll((long long)m) - 1;
    const int b0 = __shfl(cb, src, 64);
    const bool mine = (cb == b0);
    double a0 = wave_sum_d(mine ? F0 : 0.0), a1 = wave_sum_d(mine ? F1 : 0.0), a2 = wave_sum_d(mine ? F2 : 0.0);
    double a3 = wave_sum_d(mine ? T0 : 0.0), a4 = wave_sum_d(mine ? T1 : 0.0), a5 = wave_sum_d(mine ? T2 : 0.0);
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&rig[b0].F[0], a0); atomicAdd(&rig[b0].F[1], a1); atomicAdd(&rig[b0].F[2], a2);
      atomicAdd(&rig[b0].T[0], a3); atomicAdd(&rig[b0].T[1], a4); atomicAdd(&rig[b0].T[2], a5);
    }
    if (mine) cb = -1;
  }
}

// TimeStepDiffDFSPH::computeDensityChange (:1977-2041) or computeDensityAdv
// (:1926-1975) for every particle, followed by the residual sum of
// divergenceSolveIteration / pressureSolveIteration (:1906-1916, :1203-1213) and
// k_i = b_i alpha_i kcoef for the next Jacobi update.
__global__ void __launch_bounds__(DFR_BLOCK)
k_recompute(StepConst c, int mode, float kcoef, const float4* __restrict__ pos, const float4* __restrict__ vel,
            const float4* __restrict__ bpos, const float4* __restrict__ bvel, const uint32_t* __restrict__ nbr,
            const uint32_t* __restrict__ cnt, const float* __restrict__ rho, const float* __restrict__ alpha,
            float* __restrict__ badv, float* __restrict__ kf, double* __restrict__ err) {
  const uint32_t i = pid();
  double e = 0.0;
  if (i < (uint32_t)c.nf) {
    const float4 p = pos[i];
    const float4 vi = vel[i];
    const uint32_t cc = cnt[i];
    const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
    const size_t stride = (size_t)c.nf;
    float delta = 0.f;
    for (uint32_t k = 0; k < nF; ++k) {
      const uint32_t j = nbr[(size_t)k * stride + i];
      const float4 q = pos[j];
      const float4 vj = vel[j];
      const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
      const float rl = sqrtf(rx * rx + ry * ry + rz * rz);
      const float sc = cubic_gradW_scale(c.K, rl);
      delta += c.V_f * ((vi.x - vj.x) * (sc * rx) + (vi.y - vj.y) * (sc * ry) + (vi.z - vj.z) * (sc * rz));
    }
    for (uint32_t k = nF; k < nF + nB; ++k) {
      const uint32_t j = nbr[(size_t)k * stride + i];
      const float4 q = bpos[j];
      const float4 vb = bvel[j];
      const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
      const float rl = sqrtf(rx * rx + ry * ry + rz * rz);
      const float sc = cubic_gradW_scale(c.K, rl);
      delta += q.w * ((vi.x - vb.x) * (sc * rx) + (vi.y - vb.y) * (sc * ry) + (vi.z - vb.z) * (sc * rz));
    }
    float b;
    if (mode == RECOMPUTE_DIV) {
      b = fmaxf(delta, 0.f);
      if (nF + nB < 20u) b = 0.f;
      kf[i] = b * alpha[i] * kcoef;
      e = (double)(c.rho0 * b);
    } else {
      b = fmaxf(rho[i] / c.rho0 + c.dt * delta, 1.f);
      kf[i] = (b - 1.f) * alpha[i] * kcoef;
      e = (double)(c.rho0 * b - c.rho0);
    }
    badv[i] = b;
  }
  block_add_double(e, err);
}

// First loop of warmstartPressureSolve (TimeStepDiffDFSPH.cpp:987-999):
// rho* and kappa_i <- 0.5 max(kappa_i, -0.00025) / dt^2 where rho* > 1.
__global__ void __launch_bounds__(DFR_BLOCK)
k_pressure_warm(StepConst c, const float4* __restrict__ pos, const float4* __restrict__ vel,
                const float4* __restrict__ bpos, const float4* __restrict__ bvel, const uint32_t* __restrict__ nbr,
                const uint32_t* __restrict__ cnt, const float* __restrict__ rho, float* __restrict__ badv,
                float* __restrict__ kappa) {
  const uint32_t i = pid();
  if (i >= (uint32_t)c.nf) return;
  const float4 p = pos[i];
  const float4 vi = vel[i];
  const uint32_t cc = cnt[i];
  const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
  const size_t stride = (size_t)c.nf;
  float delta = 0.f;
  for (uint32_t k = 0; k < nF; ++k) {
    const uint32_t j = nbr[(size_t)k * stride + i];
    const float4 q = pos[j];
    const float4 vj = vel[j];
    const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
    const float rl = sqrtf(rx * rx + ry * ry + rz * rz);
    const float sc = cubic_gradW_scale(c.K, rl);
    delta += c.V_f * ((vi.x - vj.x) * (sc * rx) + (vi.y - vj.y) * (sc * ry) + (vi.z - vj.z) * (sc * rz));
  }
  for (uint32_t k = nF; k < nF + nB; ++k) {
    const uint32_t j = nbr[(size_t)k * stride + i];
    const float4 q = bpos[j];
    const float4 vb = bvel[j];
    const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
    const float rl = sqrtf(rx * rx + ry * ry + rz * rz);
    const float sc = cubic_gradW_scale(c.K, rl);
    delta += q.w * ((vi.x - vb.x) * (sc * rx) + (vi.y - vb.y) * (sc * ry) + (vi.z - vb.z) * (sc * rz));
  }
  const float b = fmaxf(rho[i] / c.rho0 + c.dt * delta, 1.f);
  badv[i] = b;
  const float invH2 = 1.f / (c.dt * c.dt);
  kappa[i] = (b > 1.f) ? 0.5f * fmaxf(kappa[i], -0.00025f) * invH2 : 0.f;
}

// clearAccelerations + velocity advection (TimeStep.cpp:66-81,
// TimeStepDiffDFSPH.cpp:588-605) with the end-of-divergence-solve
// kappa_v *= dt (:870-880).
__global__ void k_advect(StepConst c, float4* __restrict__ vel, float* __restrict__ kappav, float kv_scale) {
  const uint32_t i = pid();
  if (i >= (uint32_t)c.nf) return;
  float4 v = vel[i];
  v.x += c.dt * c.g[0]; v.y += c.dt * c.g[1]; v.z += c.dt * c.g[2];
  vel[i] = v;
  if (kv_scale != 1.f) kappav[i] *= kv_scale;
}

// Position update (TimeStepDiffDFSPH.cpp:613-634) with the end-of-pressure-solve
// kappa *= dt^2 (:752-767).
__global__ void k_position(StepConst c, float4* __restrict__ pos, const float4* __restrict__ vel,
                           float* __restrict__ kappa, float k_scale) {
  const uint32_t i = pid();
  if (i >= (uint32_t)c.nf) return;
  float4 p = pos[i];
  const float4 v = vel[i];
  p.x += c.dt * v.x; p.y += c.dt * v.y; p.z += c.dt * v.z;
  pos[i] = p;
  if (k_scale != 1.f) kappa[i] *= k_scale;
}

// BoundaryModel_Akinci2012::computeBoundaryVolume (:257-284) driven by
// Simulation::updateBoundaryVolume (Simulation.cpp:831-902): each sample sums W
// over the samples of its own group (all static bodies form one group, each
// dynamic body its own).
struct GroupTable { int group[DFR_MAX_BODIES]; };
__global__ void k_boundary_volume(const float4* __restrict__ bpos_in, const float4* __restrict__ bx0, int nb,
                                  GroupTable gt, KernelConst K, float h2, GridDesc g,
                                  const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bperm,
                                  float4* __restrict__ bpos_out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= (uint32_t)nb) return;
  const float4 p = bpos_in[j];
  const int grp = gt.group[(int)bx0[j].w];
  int cx, cy, cz;
  cell_coords(g, p.x, p.y, p.z, cx, cy, cz);
  float delta = K.w0;
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = cz + dz;
    if (z < 0 || z >= g.n[2]) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = cy + dy;
      if (y < 0 || y >= g.n[1]) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int x = cx + dx;
        if (x < 0 || x >= g.n[0]) continue;
        const uint32_t c = cell_key(g, x, y, z);
        for (uint32_t k = bstart[c]; k < bstart[c + 1]; ++k) {
          const uint32_t m = bperm[k];
          if (m == j) continue;
          if (gt.group[(int)bx0[m].w] != grp) continue;
          const float4 q = bpos_in[m];
          if (dist2_exact(p.x, p.y, p.z, q.x, q.y, q.z) < h2) {
            const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
            delta += cubic_W(K, sqrtf(rx * rx + ry * ry + rz * rz));
          }
        }
      }
    }
  }
  float4 o = p;
  o.w = 1.f / delta;
  bpos_out[j] = o;
}

// ---------------------------------------------------------------------------
static inline unsigned nblk(size_t n) { return (unsigned)((n + DFR_BLOCK - 1) / DFR_BLOCK); }

void launch_density_factor(const StepConst& c, FluidBufs& f, BoundaryBufs& b, NbrBufs& nb, int div_warm,
                           int divergence_on, hipStream_t s) {
  hipLaunchKernelGGL(k_density_factor, dim3(nblk(c.nf)), dim3(DFR_BLOCK), 0, s, c, f.pos, f.vel, b.pos, b.vel, nb.nbr,
                     nb.cnt, f.rho, f.alpha, f.s, f.badv, f.kappav, div_warm, divergence_on);
}

void launch_velocity_update(const StepConst& c, FluidBufs& f, BoundaryBufs& b, NbrBufs& nb, RigidDev* rig,
                            const float* kfield, float* kappa_acc, hipStream_t s) {
  hipLaunchKernelGGL(k_velocity_update, dim3(nblk(c.nf)), dim3(DFR_BLOCK), 0, s, c, f.pos, f.vel, b.pos, b.x0,
                     b.n - b.nd, nb.nbr, nb.cnt, kfield, kappa_acc, rig);
}

void launch_recompute(const StepConst& c, FluidBufs& f, BoundaryBufs& b, NbrBufs& nb, int mode, float kcoef,
                      double* err_slot, hipStream_t s) {
  hipLaunchKernelGGL(k_recompute, dim3(nblk(c.nf)), dim3(DFR_BLOCK), 0, s, c, mode, kcoef, f.pos, f.vel, b.pos,
                     b.vel, nb.nbr, nb.cnt, f.rho, f.alpha, f.badv, f.kf, err_slot);
}

void launch_pressure_warm(const StepConst& c, FluidBufs& f, BoundaryBufs& b, NbrBufs& nb, hipStream_t s) {
  hipLaunchKernelGGL(k_pressure_warm, dim3(nblk(c.nf)), dim3(DFR_BLOCK), 0, s, c, f.pos, f.vel, b.pos, b.vel, nb.nbr,
                     nb.cnt, f.rho, f.badv, f.kappa);
}

void launch_advect(const StepConst& c, FluidBufs& f, float kappav_scale, hipStream_t s) {
  hipLaunchKernelGGL(k_advect, dim3(nblk(c.nf)), dim3(DFR_BLOCK), 0, s, c, f.vel, f.kappav, kappav_scale);
}

void launch_position(const StepConst& c, FluidBufs& f, float kappa_scale, hipStream_t s) {
  hipLaunchKernelGGL(k_position, dim3(nblk(c.nf)), dim3(DFR_BLOCK), 0, s, c, f.pos, f.vel, f.kappa, kappa_scale);
}

void launch_boundary_volume_impl(const float4* bpos_in, const float4* bx0, int nb, const int* groups, int nbodies,
                                 KernelConst K, float h2, const GridDesc& g, GridBufs& gb, float4* bpos_out,
                                 hipStream_t s) {
  GroupTable gt;
  for (int k = 0; k < DFR_MAX_BODIES; ++k) gt.group[k] = (k < nbodies) ? groups[k] : -2;
  hipLaunchKernelGGL(k_boundary_volume, dim3(nblk(nb)), dim3(DFR_BLOCK), 0, s, bpos_in, bx0, nb, gt, K, h2, g,
                     gb.bstart, gb.bperm, bpos_out);
}

}  // namespace dfr
