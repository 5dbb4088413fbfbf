// Brick-tiled neighbour build and gather passes.
//
// A workgroup owns one brick of 4x4x4 cells.  After the counting sort the
// brick's particles are contiguous, and every neighbour of them lies in the
// brick's 6x6x6-cell halo.  The workgroup stages the halo's particle data in
// LDS with coalesced loads (cell ranges are contiguous), so the per-neighbour
// gathers of all passes read LDS instead of L2.  Fluid neighbour lists store
// 16-bit slots into the halo ("tile slots"; slot order = halo cells in
// (z, y, x) order, particles of a cell in sorted order), eight slots per
// 16-byte pack, packs stored pack-major (nbr8[p * N + i]) so that a wavefront
// reads and writes 1 KB per instruction.  Boundary neighbours keep 32-bit
// sample indices.
//
// A halo larger than the LDS capacity is served from global memory through the
// same slot -> particle mapping (binary search over the halo cell offsets).
#include "dfr_tile.h"

namespace dfr {

// ---------------------------------------------------------------------------
// Neighbour build (CompactNSearch::find_neighbors replacement).
// ---------------------------------------------------------------------------
constexpr int BB = PASS_BLOCK;  // build workgroup: one particle of the brick per thread

// Fluid neighbour slots of one particle as they are found: a 16-entry ring per
// thread in LDS (entry e of thread t at ring[e * BB + t], so a wavefront's
// writes are consecutive), one 16-bit write per accepted slot; whenever eight
// are pending (a chunk of candidates adds at most 4, so at most 11 are) they
// leave as one pack of the pack-major list.
constexpr int RING = 16;
struct PackWriter {
  uint16_t* ring;  // this block's ring
  int t;
  uint32_t n = 0;   // slots accepted
  uint32_t np = 0;  // packs written
  __device__ inline void push(uint32_t s) {
    ring[(n & (RING - 1)) * BB + t] = (uint16_t)s;
    ++n;
  }
  // Branch-free push: the slot is written at the next free entry either way
  // (an entry past the accepted ones, which the next push overwrites; at most
  // 11 entries are pending, so it never holds one) and counted if accepted.
  __device__ inline void push_if(uint32_t s, bool accept) {
    ring[(n & (RING - 1)) * BB + t] = (uint16_t)s;
    n += accept ? 1u : 0u;
  }
  __device__ inline void store(uint4* __restrict__ nbr8, size_t stride, uint32_t i, uint32_t count) {
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t e0 = np * PACK + 2 * k, e1 = e0 + 1;
      const uint32_t lo = (2 * k < (int)count) ? ring[(e0 & (RING - 1)) * BB + t] : 0u;
      const uint32_t hi = (2 * k + 1 < (int)count) ? ring[(e1 & (RING - 1)) * BB + t] : 0u;
      w[k] = lo | (hi << 16);
    }
    if (np < (uint32_t)NPACK) nbr8[(size_t)np * stride + i] = make_uint4(w[0], w[1], w[2], w[3]);
    ++np;
  }
  __device__ inline void flush_full(uint4* __restrict__ nbr8, size_t stride, uint32_t i) {
    if (n - np * PACK >= (uint32_t)PACK) store(nbr8, stride, i, PACK);
  }
  __device__ inline uint32_t finish(uint4* __restrict__ nbr8, size_t stride, uint32_t i) {
    const uint32_t rest = n - np * PACK;
    if (rest > 0u) store(nbr8, stride, i, rest);
    return n;
  }
};

// Dynamic LDS: x, y, z of the halo in three arrays of a fixed stride, so one
// address serves a chunk of four candidates (immediate offsets); a chunk reads
// up to three slots past its row's end, which stay inside the arrays because
// a halo is staged only if it has at most BSTRIDE - 4 slots (three workgroups
// of 640 threads still fit a CU's LDS).
constexpr int BSTRIDE = 2304;
__global__ void __launch_bounds__(BB)
k_build_tiled(const float4* __restrict__ pos, int nf, GridDesc g, const uint32_t* __restrict__ fstart,
              const float4* __restrict__ bpos, const uint32_t* __restrict__ bstart,
              const uint32_t* __restrict__ bperm, float h2, int maxb, int cap, uint4* __restrict__ nbr8,
              uint32_t* __restrict__ nbrB, uint32_t* __restrict__ cnt, uint32_t* __restrict__ overflow,
              uint4* __restrict__ thdr, uint32_t* __restrict__ tmap) {
  __shared__ TileHdr<BB> T;
  __shared__ uint32_t bcs[HALO], bcn[HALO];  // boundary sample range of each halo cell
  __shared__ uint16_t ring[RING * BB];
  extern __shared__ float lds[];
  const uint32_t brick = xcd_block();
  if (brick_empty(fstart, brick)) {
    if (threadIdx.x == 0) thdr[brick] = make_uint4(0u, 0u, 0u, 0u);
    return;
  }
  tile_setup<BB>(g, fstart, brick, T);
  if (threadIdx.x == 0) thdr[brick] = make_uint4(T.own_begin, T.own_end, T.total, 0u);
  if (T.own_begin == T.own_end) return;
  const uint32_t total = T.total;
  const int t = threadIdx.x;
  if (t == 0) atomicMax(&overflow[3], total);
  if (total > 65535u) {
    // tile slots are 16-bit: leave empty lists (safe for every pass) and report
    for (uint32_t i = T.own_begin + t; i < T.own_end; i += BB) cnt[i] = 0u;
    if (t == 0) atomicMax(&overflow[0], total);
    return;
  }
  if (t < HALO) {
    const int cx = 4 * T.bx - 1 + t % 6, cy = 4 * T.by - 1 + (t / 6) % 6, cz = 4 * T.bz - 1 + t / 36;
    uint32_t st = 0, n = 0;
    if (cx >= 0 && cy >= 0 && cz >= 0 && cx < g.n[0] && cy < g.n[1] && cz < g.n[2]) {
      const uint32_t key = cell_key(g, cx, cy, cz);
      st = bstart[key];
      n = bstart[key + 1] - st;
    }
    bcs[t] = st;
    bcn[t] = n;
  }
  const bool in_lds = total <= (uint32_t)min(cap, BSTRIDE - 4);
  float* lx = lds; float* ly = lds + BSTRIDE; float* lz = lds + 2 * BSTRIDE;
  // the halo's slot -> particle map, for every slot a pass could stage
  constexpr int PTB = per_thread<BB>();
  uint32_t js[PTB];
  {
    const uint32_t mapped = min(total, (uint32_t)TILE_CAP);
    tile_indices<BB>(T, mapped, js);
#pragma unroll
    for (int r = 0; r < PTB; ++r) {
      const uint32_t s = t + r * BB;
      if (s < mapped) tmap[(size_t)brick * TILE_CAP + s] = js[r];
    }
  }
  if (in_lds) {
    float4 q[PTB];
#pragma unroll
    for (int r = 0; r < PTB; ++r)
      if (t + r * BB < (int)total) q[r] = pos[js[r]];
#pragma unroll
    for (int r = 0; r < PTB; ++r) {
      const uint32_t s = t + r * BB;
      if (s < total) { lx[s] = q[r].x; ly[s] = q[r].y; lz[s] = q[r].z; }
    }
  }
  __syncthreads();
  const size_t stride = (size_t)nf;
  for (uint32_t i = T.own_begin + t; i < T.own_end; i += BB) {
    const float4 p = pos[i];
    int cx, cy, cz;
    cell_coords(g, p.x, p.y, p.z, cx, cy, cz);
    const int hc0 = halo_cell<BB>(T, cx, cy, cz, 0, 0, 0);
    const uint32_t self = T.coff[hc0] + (i - T.cstart[hc0]);
    PackWriter pw{ring, t};
    // the three x-neighbour cells of a (dz, dy) row are consecutive halo
    // cells, so their candidates form one contiguous slot range; the centre
    // row goes as the two ranges either side of the particle's own slot
    // (ranges 4 and 5), which keeps the candidate order and needs no
    // per-candidate self test
    for (int row = 0; row < 10; ++row) {
      const int rr = row < 5 ? row : row - 1;
      const int dz = rr / 3 - 1, dy = rr % 3 - 1;
      const int hm = halo_cell<BB>(T, cx, cy, cz, -1, dy, dz);
      uint32_t r0 = T.coff[hm], r1 = T.coff[hm + 3];
      if (row == 4) r1 = self;
      if (row == 5) r0 = self + 1;
      if (in_lds) {
        for (uint32_t s0 = r0; s0 < r1; s0 += 4) {
          float qx[4], qy[4], qz[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t s = s0 + u;  // <= r1 + 2 <= total + 2 < BSTRIDE
            qx[u] = lx[s]; qy[u] = ly[s]; qz[u] = lz[s];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t s = s0 + u;
            pw.push_if(s, (s < r1) & (dist2_exact(p.x, p.y, p.z, qx[u], qy[u], qz[u]) < h2));
          }
          pw.flush_full(nbr8, stride, i);
        }
      } else {
        for (uint32_t s = r0; s < r1; ++s) {
          const float4 q = pos[tile_global<BB>(T, s)];
          if (dist2_exact(p.x, p.y, p.z, q.x, q.y, q.z) < h2) pw.push(s);
          pw.flush_full(nbr8, stride, i);
        }
      }
    }
    uint32_t nF = pw.finish(nbr8, stride, i);
    uint32_t nB = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int hc = halo_cell<BB>(T, cx, cy, cz, dx, dy, dz);
          const uint32_t n = bcn[hc];
          if (n == 0) continue;
          const uint32_t st = bcs[hc];
          // four samples at a time: their indices, then their positions, then the tests in order
          for (uint32_t k0 = 0; k0 < n; k0 += 4) {
            uint32_t jj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) jj[u] = bperm[st + min(k0 + u, n - 1)];
            float4 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = bpos[jj[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (k0 + u < n && dist2_exact(p.x, p.y, p.z, q[u].x, q[u].y, q[u].z) < h2) {
                if (nB < (uint32_t)maxb) nbrB[(size_t)nB * stride + i] = jj[u];
                ++nB;
              }
          }
        }
    if (nF > (uint32_t)MAXF) { atomicMax(&overflow[0], nF); nF = MAXF; }
    if (nB > (uint32_t)maxb) { atomicMax(&overflow[2], nB); nB = maxb; }
    cnt[i] = nF | (nB << 16);
  }
}

// ---------------------------------------------------------------------------
// Gather passes: PASS_BLOCK threads per brick.  Dynamic LDS per slot:
// x, y, z (+ per-pass fields).
// ---------------------------------------------------------------------------
constexpr int PB = PASS_BLOCK;
constexpr int PT = per_thread<PB>();
// The velocity update stages 4 floats per slot, so LDS would fit four 640-thread
// workgroups but the CU holds at most 32 wavefronts: 512 threads keep four.
constexpr int VB = 512;
constexpr int VT = per_thread<VB>();

// TimeStep::computeDensities (TimeStep.cpp:147-202) + computeDFSPHFactor
// (TimeStepDiffDFSPH.cpp:883-962) + the first loop of warmstartDivergenceSolve
// (:1720-1732), one pass.
__global__ void __launch_bounds__(PB) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_density_factor_t(PassArgs a, float* __restrict__ rho, float* __restrict__ alpha, float4* __restrict__ s_out,
                   float* __restrict__ badv, float* __restrict__ kappav, int div_warm, int divergence_on,
                   float2* __restrict__ aux) {
  __shared__ TileHdr<PB> T;
  extern __shared__ float4 lds4[];
  const StepConst& c = a.c;
  const uint32_t brick = xcd_block();
  const int cap = a.cap;
  BrickView B;
  if (!brick_open<PB>(a.g, a.fstart, a.thdr, brick, cap, T, B)) return;
  const uint32_t total = B.total;
  const bool in_lds = B.in_lds;
  float4* LA = lds4;                                          // x, y, z, vx
  float2* LB = reinterpret_cast<float2*>(lds4 + cap);         // vy, vz
  const int t = threadIdx.x;
  if (in_lds) {
    uint32_t js[PT];
    map_indices<PB>(a.tmap, brick, total, js);
    float4 q[PT], v[PT];
#pragma unroll
    for (int r = 0; r < PT; ++r)
      if (t + r * PB < (int)total) {
        q[r] = a.pos[js[r]];
        v[r] = divergence_on ? a.vel[js[r]] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
    for (int r = 0; r < PT; ++r) {
      const uint32_t s = t + r * PB;
      if (s < total) {
        LA[s] = make_float4(q[r].x, q[r].y, q[r].z, v[r].x);
        LB[s] = make_float2(v[r].y, v[r].z);
      }
    }
  }
  __syncthreads();
  const size_t stride = (size_t)c.nf;
  for (uint32_t i = B.own_begin + t; i < B.own_end; i += PB) {
    const float4 p = a.pos[i];
    float4 vi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (divergence_on) vi = a.vel[i];
    const uint32_t cc = a.cnt[i];
    const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
    float dens = c.V_f * c.K.w0;
    float sumsq = 0.f, gx = 0.f, gy = 0.f, gz = 0.f, delta = 0.f;
    // sums over fluid neighbours without the common factor V_f (applied once)
    float wsum = 0.f, sq = 0.f, fgx = 0.f, fgy = 0.f, fgz = 0.f, fdelta = 0.f;
    // a slot that is not live contributes W = 0 and gradW = 0 (exact no-ops)
    auto pair = [&](float qx, float qy, float qz, float vx, float vy, float vz, bool live) {
      const float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
      const PairKernel kw = cubic_pair(c.K, rx * rx + ry * ry + rz * rz, live);
      const float kc = kw.c;
      wsum += kw.W;
      const float wx = kc * rx, wy = kc * ry, wz = kc * rz;
      sq += wx * wx + wy * wy + wz * wz;
      fgx += wx; fgy += wy; fgz += wz;
      if (divergence_on) fdelta += (vi.x - vx) * wx + (vi.y - vy) * wy + (vi.z - vz) * wz;
    };
    if (in_lds) {
      for_fluid_groups<2>(a.nbr8, stride, i, nF, [&](uint32_t s) { return Slot6{LA[s], LB[s]}; },
                          [&](const Slot6& q, bool live) {
                            pair(q.a.x, q.a.y, q.a.z, q.a.w, q.b.x, q.b.y, live);
                          });
    } else {
      for_fluid_slots_rolled(a.nbr8, stride, i, nF, [&](uint32_t s) {
        const uint32_t j = tile_global<PB>(T, s);
        const float4 q = a.pos[j];
        const float4 v = divergence_on ? a.vel[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        pair(q.x, q.y, q.z, v.x, v.y, v.z, true);
      });
    }
    dens += c.V_f * wsum;
    sumsq = c.V_f * c.V_f * sq;
    gx = c.V_f * fgx; gy = c.V_f * fgy; gz = c.V_f * fgz;
    delta = c.V_f * fdelta;
    float bdens = 0.f;  // boundary part of the density sum (a feature of the learned warm start)
    for_boundary_groups(
        a.nbrB, stride, i, nB,
        [&](uint32_t j) { return Slot8{a.bpos[j], divergence_on ? a.bvel[j] : make_float4(0.f, 0.f, 0.f, 0.f)}; },
        [&](const Slot8& s, uint32_t) {
          const float4 q = s.a, vb = s.b;
          const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
          const PairKernel kw = cubic_pair(c.K, rx * rx + ry * ry + rz * rz);
          dens += q.w * kw.W;
          bdens += q.w * kw.W;
          const float wx = kw.c * rx, wy = kw.c * ry, wz = kw.c * rz;
          gx += q.w * wx; gy += q.w * wy; gz += q.w * wz;
          if (divergence_on) delta += q.w * ((vi.x - vb.x) * wx + (vi.y - vb.y) * wy + (vi.z - vb.z) * wz);
        });
    rho[i] = dens * c.rho0;
    if (aux) aux[i] = make_float2(delta, bdens);
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
}

// Jacobi velocity update shared by the two warm starts and every iteration of
// both solves (TimeStepDiffDFSPH.cpp:1024-1055, :1101-1197, :1753-1784,
// :1826-1902): v_i -= dt sum_j (k_i + k_j) grad_p_ij over fluid pairs with
// |k_i + k_j| > eps; when |k_i| > eps each boundary neighbour b changes v_i by
// -dt k_i grad_p_ib and receives f = -m_i dv / dt (BoundaryModel::addForce,
// BoundaryModel.h:46-58), summed here in FP64 per dynamic body.
// LDS per slot: x, y, z, k.
__global__ void __launch_bounds__(VB) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_velocity_update_t(PassArgs a, const float* __restrict__ kfield, float* __restrict__ kappa_acc,
                    RigidDev* __restrict__ rig, const int* __restrict__ stop, IterTapeOut tape) {
  __shared__ TileHdr<VB> T;
  extern __shared__ float4 lds4[];
  const StepConst& c = a.c;
  if (stop && *stop) return;
  const uint32_t brick = xcd_block();
  const int cap = a.cap;
  BrickView B;
  if (!brick_open<VB>(a.g, a.fstart, a.thdr, brick, cap, T, B)) return;
  const uint32_t total = B.total;
  const bool in_lds = B.in_lds;
  float4* L = lds4;  // x, y, z, k
  const int t = threadIdx.x;
  if (in_lds) {
    uint32_t js[VT];
    map_indices<VB>(a.tmap, brick, total, js);
    float4 q[VT];
    float kk[VT];
#pragma unroll
    for (int r = 0; r < VT; ++r)
      if (t + r * VB < (int)total) { q[r] = a.pos[js[r]]; kk[r] = kfield[js[r]]; }
#pragma unroll
    for (int r = 0; r < VT; ++r) {
      const uint32_t s = t + r * VB;
      if (s < total) L[s] = make_float4(q[r].x, q[r].y, q[r].z, kk[r]);
    }
  }
  __syncthreads();
  const size_t stride = (size_t)c.nf;
  const float invH = 1.f / c.dt;
  // The particle loop runs the same number of times on every lane of a wave, so
  // the coupling partials of each round are reduced over the wave at its end
  // (and are not live across the next round's fluid pair loop).
  for (uint32_t base = B.own_begin; base < B.own_end; base += VB) {
    const uint32_t i = base + t;
    int cb = -1;
    double F0 = 0, F1 = 0, F2 = 0, T0 = 0, T1 = 0, T2 = 0;
    if (i < B.own_end) {
      const float ki = kfield[i];
      if (kappa_acc) kappa_acc[i] += ki;
      const float4 p = a.pos[i];
      float4 vi = a.vel[i];
      if (tape.v) tape.v[i] = vi;
      if (tape.k) tape.k[i] = ki;
      if (tape.b) tape.b[i] = tape.bsrc[i];
      const uint32_t cc = a.cnt[i];
      const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
      // v_i -= dt (k_i + k_j) grad_p_ij = dt V_f (k_i + k_j) c_ij r_ij; dt V_f applied once
      float ux = 0.f, uy = 0.f, uz = 0.f;
      // pairs with |k_i + k_j| <= eps (and slots that are not live) add s = 0
      auto pair = [&](float qx, float qy, float qz, float kj, bool live) {
        const float kSum = ki + kj;
        const float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
        const float s = kSum * cubic_grad_c(c.K, rx * rx + ry * ry + rz * rz, live && fabsf(kSum) > c.eps);
        ux += s * rx; uy += s * ry; uz += s * rz;
      };
      if (in_lds) {
        for_fluid_groups<2>(a.nbr8, stride, i, nF, [&](uint32_t s) { return L[s]; },
                            [&](const float4& q, bool live) { pair(q.x, q.y, q.z, q.w, live); });
      } else {
        for_fluid_slots_rolled(a.nbr8, stride, i, nF, [&](uint32_t s) {
          const uint32_t j = tile_global<VB>(T, s);
          const float4 q = a.pos[j];
          pair(q.x, q.y, q.z, kfield[j], true);
        });
      }
      {
        const float dv = c.dt * c.V_f;
        vi.x += dv * ux; vi.y += dv * uy; vi.z += dv * uz;
      }
      if (fabsf(ki) > c.eps) {
        auto load = [&](uint32_t j) { return a.bpos[j]; };
        for_boundary_groups<2>(a.nbrB, stride, i, nB, load, [&](const float4& q, uint32_t j) {
          const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
          const float sc = cubic_grad_c(c.K, rx * rx + ry * ry + rz * rz);
          const float gpx = -q.w * (sc * rx), gpy = -q.w * (sc * ry), gpz = -q.w * (sc * rz);
          const float am = -c.dt * 1.f * ki;
          const float dvx = am * gpx, dvy = am * gpy, dvz = am * gpz;
          vi.x += dvx; vi.y += dvy; vi.z += dvz;
          if ((int)j >= a.nb_static) {
            const int body = (int)a.bx0[j].w;
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
        });
      }
      a.vel[i] = vi;
    }
    // Wave-level FP64 reduction of the round's coupling partials, one body at a time.
    while (true) {
      const unsigned long long m = __ballot(cb >= 0);
      if (m == 0ull) break;
      const int src = __ffsll((long long)m) - 1;
      const int b0 = __shfl(cb, src, 64);
      const bool mine = (cb == b0);
      const double a0 = wave_sum_d(mine ? F0 : 0.0), a1 = wave_sum_d(mine ? F1 : 0.0),
                   a2 = wave_sum_d(mine ? F2 : 0.0);
      const double a3 = wave_sum_d(mine ? T0 : 0.0), a4 = wave_sum_d(mine ? T1 : 0.0),
                   a5 = wave_sum_d(mine ? T2 : 0.0);
      if ((threadIdx.x & 63) == 0) {
        atomicAdd(&rig[b0].F[0], a0); atomicAdd(&rig[b0].F[1], a1); atomicAdd(&rig[b0].F[2], a2);
        atomicAdd(&rig[b0].T[0], a3); atomicAdd(&rig[b0].T[1], a4); atomicAdd(&rig[b0].T[2], a5);
      }
      if (mine) cb = -1;
    }
  }
}

// TimeStepDiffDFSPH::computeDensityChange (:1977-2041) or computeDensityAdv
// (:1926-1975) for every particle, followed by the residual sum of
// divergenceSolveIteration / pressureSolveIteration (:1906-1916, :1203-1213)
// and k_i = b_i alpha_i kcoef for the next Jacobi update (mode 0 divergence,
// 1 pressure); mode 2 is the first loop of warmstartPressureSolve (:987-999).
// LDS per slot: x, y, z, vx, vy, vz.
__global__ void __launch_bounds__(PB) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_recompute_t(PassArgs a, int mode, float kcoef, const float* __restrict__ rho, const float* __restrict__ alpha,
              float* __restrict__ badv, float* __restrict__ kf, float* __restrict__ kappa, double* __restrict__ err,
              const float* __restrict__ kpred, const int* __restrict__ stop) {
  __shared__ TileHdr<PB> T;
  extern __shared__ float4 lds4[];
  const StepConst& c = a.c;
  if (stop && *stop) return;
  const uint32_t brick = xcd_block();
  const int cap = a.cap;
  const bool residual = err && (mode == RECOMPUTE_DIV || mode == RECOMPUTE_PRESSURE);
  BrickView B;
  if (!brick_open<PB>(a.g, a.fstart, a.thdr, brick, cap, T, B)) {
    if (residual && threadIdx.x == 0) err[brick] = 0.0;
    return;
  }
  const uint32_t total = B.total;
  const bool in_lds = B.in_lds;
  float4* LA = lds4;                                          // x, y, z, vx
  float2* LB = reinterpret_cast<float2*>(lds4 + cap);         // vy, vz
  const int t = threadIdx.x;
  if (in_lds) {
    uint32_t js[PT];
    map_indices<PB>(a.tmap, brick, total, js);
    float4 q[PT], v[PT];
#pragma unroll
    for (int r = 0; r < PT; ++r)
      if (t + r * PB < (int)total) { q[r] = a.pos[js[r]]; v[r] = a.vel[js[r]]; }
#pragma unroll
    for (int r = 0; r < PT; ++r) {
      const uint32_t s = t + r * PB;
      if (s < total) {
        LA[s] = make_float4(q[r].x, q[r].y, q[r].z, v[r].x);
        LB[s] = make_float2(v[r].y, v[r].z);
      }
    }
  }
  __syncthreads();
  const size_t stride = (size_t)c.nf;
  double e = 0.0;
  for (uint32_t i = B.own_begin + t; i < B.own_end; i += PB) {
    const float4 p = a.pos[i];
    const float4 vi = a.vel[i];
    const uint32_t cc = a.cnt[i];
    const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
    float delta = 0.f;
    float fdelta = 0.f;  // fluid part without the common factor V_f
    auto pair = [&](float qx, float qy, float qz, float vx, float vy, float vz, bool live) {
      const float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
      const float sc = cubic_grad_c(c.K, rx * rx + ry * ry + rz * rz, live);
      fdelta += sc * ((vi.x - vx) * rx + (vi.y - vy) * ry + (vi.z - vz) * rz);
    };
    if (in_lds) {
      for_fluid_groups<4>(a.nbr8, stride, i, nF, [&](uint32_t s) { return Slot6{LA[s], LB[s]}; },
                          [&](const Slot6& q, bool live) {
                            pair(q.a.x, q.a.y, q.a.z, q.a.w, q.b.x, q.b.y, live);
                          });
    } else {
      for_fluid_slots_rolled(a.nbr8, stride, i, nF, [&](uint32_t s) {
        const uint32_t j = tile_global<PB>(T, s);
        const float4 q = a.pos[j];
        const float4 v = a.vel[j];
        pair(q.x, q.y, q.z, v.x, v.y, v.z, true);
      });
    }
    delta = c.V_f * fdelta;
    for_boundary_groups(a.nbrB, stride, i, nB, [&](uint32_t j) { return Slot8{a.bpos[j], a.bvel[j]}; },
                        [&](const Slot8& s, uint32_t) {
                          const float4 q = s.a, vb = s.b;
                          const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
                          const float sc = cubic_grad_c(c.K, rx * rx + ry * ry + rz * rz);
                          delta += q.w * sc * ((vi.x - vb.x) * rx + (vi.y - vb.y) * ry + (vi.z - vb.z) * rz);
                        });
    float b;
    if (mode == 0) {
      b = fmaxf(delta, 0.f);
      if (nF + nB < 20u) b = 0.f;
      kf[i] = b * alpha[i] * kcoef;
      e += (double)(c.rho0 * b);
    } else {
      b = fmaxf(rho[i] / c.rho0 + c.dt * delta, 1.f);
      if (mode == 1) {
        kf[i] = (b - 1.f) * alpha[i] * kcoef;
        e += (double)(c.rho0 * b - c.rho0);
      } else if (mode == PRESSURE_WARMSTART) {
        const float invH2 = 1.f / (c.dt * c.dt);
        kappa[i] = (b > 1.f) ? 0.5f * fmaxf(kappa[i], -0.00025f) * invH2 : 0.f;
      } else {  // learned warm start: the predicted stiffness (stored units)
        const float invH2 = 1.f / (c.dt * c.dt);
        kappa[i] = (b > 1.f) ? kpred[i] * invH2 : 0.f;
      }
    }
    badv[i] = b;
  }
  if (residual) block_sum_double<PB>(e, err + brick);
}

// Expands tile slots back to sorted fluid indices (diagnostics).
__global__ void __launch_bounds__(BB)
k_expand_lists(GridDesc g, const uint32_t* __restrict__ fstart, int nf, const uint4* __restrict__ nbr8,
               const uint32_t* __restrict__ cnt, uint32_t* __restrict__ out) {
  __shared__ TileHdr<BB> T;
  const uint32_t brick = blockIdx.x;
  tile_setup<BB>(g, fstart, brick, T);
  if (T.own_begin == T.own_end) return;
  const size_t stride = (size_t)nf;
  for (uint32_t i = T.own_begin + threadIdx.x; i < T.own_end; i += BB) {
    const uint32_t nF = cnt[i] & 0xffffu;
    uint32_t k = 0;
    for_fluid_slots(nbr8, stride, i, nF, [&](uint32_t s) { out[(size_t)(k++) * stride + i] = tile_global<BB>(T, s); });
  }
}

// ---------------------------------------------------------------------------

void launch_build_tiled(FluidBufs& f, BoundaryBufs& b, const GridDesc& g, GridBufs& gb, NbrBufs& nb, float h2,
                        hipStream_t s) {
  const size_t shm = sizeof(float) * 3 * (size_t)BSTRIDE;
  hipLaunchKernelGGL(k_build_tiled, dim3(nbricks(g)), dim3(BB), shm, s, f.pos, f.n, g, gb.fstart, b.pos, gb.bstart,
                     gb.bperm, h2, nb.maxb, nb.cap, nb.nbr8, nb.nbrB, nb.cnt, nb.overflow, nb.thdr, nb.tmap);
}

void launch_density_factor_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                             NbrBufs& nb, int div_warm, int divergence_on, float2* aux, hipStream_t s) {
  const size_t shm = sizeof(float) * (size_t)nb.cap * 6;
  hipLaunchKernelGGL(k_density_factor_t, dim3(nbricks(g)), dim3(PB), shm, s, make_args(c, g, gb, f, b, nb), f.rho,
                     f.alpha, f.s, f.badv, f.kappav, div_warm, divergence_on, aux);
}

void launch_velocity_update_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                              NbrBufs& nb, RigidDev* rig, const float* kfield, float* kappa_acc, hipStream_t s,
                              const int* stop, IterTapeOut tape) {
  const size_t shm = sizeof(float) * (size_t)nb.cap * 4;
  hipLaunchKernelGGL(k_velocity_update_t, dim3(nbricks(g)), dim3(VB), shm, s, make_args(c, g, gb, f, b, nb), kfield,
                     kappa_acc, rig, stop, tape);
}

void launch_recompute_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                        NbrBufs& nb, int mode, float kcoef, double* err_slot, hipStream_t s, const float* kpred,
                        const int* stop) {
  const size_t shm = sizeof(float) * (size_t)nb.cap * 6;
  hipLaunchKernelGGL(k_recompute_t, dim3(nbricks(g)), dim3(PB), shm, s, make_args(c, g, gb, f, b, nb), mode, kcoef,
                     f.rho, f.alpha, f.badv, f.kf, f.kappa, err_slot, kpred, stop);
}

void launch_expand_lists(const GridDesc& g, GridBufs& gb, FluidBufs& f, NbrBufs& nb, uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_expand_lists, dim3(nbricks(g)), dim3(BB), 0, s, g, gb.fstart, f.n, nb.nbr8, nb.cnt, out);
}

void set_tiled_kernel_attributes() {
  const int shm6 = (int)(sizeof(float) * TILE_CAP * 6), shm4 = (int)(sizeof(float) * TILE_CAP * 4);
  (void)hipFuncSetAttribute((const void*)k_density_factor_t, hipFuncAttributeMaxDynamicSharedMemorySize, shm6);
  (void)hipFuncSetAttribute((const void*)k_recompute_t, hipFuncAttributeMaxDynamicSharedMemorySize, shm6);
  (void)hipFuncSetAttribute((const void*)k_velocity_update_t, hipFuncAttributeMaxDynamicSharedMemorySize, shm4);
  (void)hipFuncSetAttribute((const void*)k_build_tiled, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(float) * 3 * BSTRIDE));
}

}  // namespace dfr
