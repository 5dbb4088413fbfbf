// Elementwise passes of the DFSPH step (advection, position update) and the
// Akinci boundary volume.  The neighbour-gather passes live in dfr_tiled.hip.
#include <algorithm>
#include "dfr_kernels.h"

namespace dfr {

// clearAccelerations + velocity advection (TimeStep.cpp:66-81,
// TimeStepDiffDFSPH.cpp:588-605) with the end-of-divergence-solve
// kappa_v *= dt (:870-880).
__global__ void k_advect(StepConst c, float4* __restrict__ vel, float* __restrict__ kappav, float kv_scale) {
  const uint32_t i = gid();
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
  const uint32_t i = gid();
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

// One Jacobi iteration's convergence test (divergenceSolve :816-828,
// pressureSolve :700-712): continue while (avg > eta or it + 1 < min) and
// it + 1 < max.  probe is host-visible memory.
// The residual is the sum of the recompute launch's per-brick partials, in a
// fixed order (every run sums the same way).
constexpr int CHECK_BLOCK = 1024;
__global__ void __launch_bounds__(CHECK_BLOCK)
k_solve_check(const double* __restrict__ part, int nparts, double* __restrict__ err, int n, double eta, int it,
              int min_it, int max_it, int* __restrict__ stop, SolveProbe* probe) {
  if (stop && *stop) return;
  // four independent running sums per thread keep four loads in flight; the
  // order of every addition is fixed by nparts alone
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int k = threadIdx.x;
  for (; k + 3 * CHECK_BLOCK < nparts; k += 4 * CHECK_BLOCK) {
    v0 += part[k];
    v1 += part[k + CHECK_BLOCK];
    v2 += part[k + 2 * CHECK_BLOCK];
    v3 += part[k + 3 * CHECK_BLOCK];
  }
  for (; k < nparts; k += CHECK_BLOCK) v0 += part[k];
  double v = (v0 + v1) + (v2 + v3);
  __shared__ double w[CHECK_BLOCK / 64];
  v = wave_sum_d(v);
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double sum = 0.0;
  for (int k = 0; k < CHECK_BLOCK / 64; ++k) sum += w[k];
  *err = sum;
  if (!stop) return;
  const double avg = sum / (double)n;
  const bool chk = avg <= eta;
  const int done = ((chk && it + 1 >= min_it) || it + 1 >= max_it) ? 1 : 0;
  *stop = done;
  probe->avg = avg;
  probe->done = done;
  __threadfence_system();
}

// One segment per blockIdx.y: 16-byte words, then the 4-byte tail.
__global__ void k_multi_copy(CopyList L, const int* __restrict__ stop) {
  if (stop && *stop) return;
  const CopySeg sg = L.seg[blockIdx.y];
  const uint64_t n16 = sg.bytes >> 4;
  uint4* d = static_cast<uint4*>(sg.dst);
  const uint4* s = static_cast<const uint4*>(sg.src);
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += (uint64_t)gridDim.x * blockDim.x)
    d[k] = s ? s[k] : make_uint4(0u, 0u, 0u, 0u);
  const uint32_t tail = (uint32_t)((sg.bytes & 15u) >> 2);
  if (blockIdx.x == 0 && threadIdx.x < tail) {
    uint32_t* dw = static_cast<uint32_t*>(sg.dst) + n16 * 4;
    dw[threadIdx.x] = s ? static_cast<const uint32_t*>(sg.src)[n16 * 4 + threadIdx.x] : 0u;
  }
}

// ---------------------------------------------------------------------------
static inline unsigned nblk(size_t n) { return (unsigned)((n + DFR_BLOCK - 1) / DFR_BLOCK); }

void launch_multi_copy(const CopyList& list, hipStream_t s, const int* stop) {
  if (list.n <= 0) return;
  uint64_t most = 0;
  for (int k = 0; k < list.n; ++k) most = std::max<uint64_t>(most, list.seg[k].bytes >> 4);
  const unsigned bx = (unsigned)std::min<uint64_t>(std::max<uint64_t>((most + DFR_BLOCK - 1) / DFR_BLOCK, 1), 2048);
  hipLaunchKernelGGL(k_multi_copy, dim3(bx, (unsigned)list.n), dim3(DFR_BLOCK), 0, s, list, stop);
}

void launch_solve_check(const double* part, int nparts, double* err, int n, double eta, int it, int min_it,
                        int max_it, int* stop, SolveProbe* probe, hipStream_t s) {
  hipLaunchKernelGGL(k_solve_check, dim3(1), dim3(CHECK_BLOCK), 0, s, part, nparts, err, n, eta, it, min_it, max_it,
                     stop, probe);
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
