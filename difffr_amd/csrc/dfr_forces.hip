// Non-pressure forces and the CFL step size, on the brick tiles of the gather
// passes (one workgroup per brick, halo staged in LDS, 16-bit slot lists).
//
//   Simulation::computeNonPressureForces (Simulation.cpp:618-631) after
//   clearAccelerations (a = g): surface tension, then viscosity;
//   Simulation::updateTimeStepSizeCFL (Simulation.cpp:544-616).
//
// Surface tension is Akinci et al. 2013 (SurfaceTension_Akinci2013.cpp): a
// normal pass n_i = h sum_j (m_j / rho_j) gradW_ij, then per fluid pair
// K_ij (-k m_j C(|r|) r/|r| - k (n_i - n_j)) with K_ij = 2 rho0 / (rho_i + rho_j)
// and the cohesion kernel C (SPHKernels.h:437-505, including its constant term
// h^6/64 taken without the kernel's factor, as the reference evaluates it), plus
// boundary adhesion -kb rho0 V_b A(|r|) r/|r| (SPHKernels.h:508-560; no reaction
// force on the body, as in the reference).  He et al. 2014
// (SurfaceTension_He2014.cpp:29-160) is a colour field c_i = (m / rho_i) W(0) +
// sum_j (m / rho_j) W_ij + sum_b V_b W_ib, gradC2_i = |(1 / c_i) sum_j (m / rho_j)
// c_j gradW_ij|^2, and per fluid pair (k / 4 rho_i) (m / rho_j) (gradC2_i + gradC2_j)
// gradW_ij plus per boundary sample (kb / 4 rho_i) V_b gradC2_i gradW_ib, whose
// reaction -m a goes to the body (BoundaryModel::addForce).  Viscosity is Standard
// (Viscosity_Standard.cpp:233-269, the reference fluid's default:
// FluidModel.cpp:100): 10 nu (m_j / rho_j) (v_ij . r_ij) / (|r_ij|^2 + 0.01 h^2)
// gradW_ij, or XSPH (Viscosity_XSPH.cpp:35-108): -(1/dt) nu (m_j / rho_j)
// (v_i - v_j) W_ij, over fluid pairs.  Densities are the step's (computeDensities), velocities those
// after the divergence solve.
#include "dfr_tile.h"

namespace dfr {

constexpr int FB = PASS_BLOCK;
constexpr int FT = per_thread<FB>();

// Akinci 2013 cohesion kernel C(|r|) for r2 <= h^2 (zero beyond).
__device__ inline float cohesion_w(const ForceConst& f, float r2) {
  if (!(r2 <= f.h * f.h)) return 0.f;
  const float r1 = sqrtf(r2);
  const float r3 = r2 * r1;
  const float d = f.h - r1;
  const float d3 = d * d * d;
  return (r1 > 0.5f * f.h) ? f.ck * d3 * r3 : f.ck * 2.f * d3 * r3 - f.cc;
}
// Akinci 2013 adhesion kernel A(|r|): nonzero for h/2 < |r| <= h.
__device__ inline float adhesion_w(const ForceConst& f, float r2) {
  if (!(r2 <= f.h * f.h)) return 0.f;
  const float rl = sqrtf(r2);
  if (!(rl > 0.5f * f.h)) return 0.f;
  const float x = -4.f * r2 / f.h + 6.f * rl - 2.f * f.h;
  return f.ak * sqrtf(sqrtf(fmaxf(x, 0.f)));
}

// Akinci 2013 normals.  LDS per slot: x, y, z, rho.
__global__ void __launch_bounds__(FB) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_normals_t(PassArgs a, const float* __restrict__ rho, float4* __restrict__ nrm) {
  __shared__ TileHdr<FB> T;
  extern __shared__ float4 lds4[];
  const StepConst& c = a.c;
  const uint32_t brick = xcd_block();
  BrickView B;
  if (!brick_open<FB>(a.g, a.fstart, a.thdr, brick, a.cap, T, B)) return;
  const uint32_t total = B.total;
  const bool in_lds = B.in_lds;
  float4* LA = lds4;
  const int t = threadIdx.x;
  if (in_lds) {
    uint32_t js[FT];
    map_indices<FB>(a.tmap, brick, total, js);
    float4 q[FT];
    float rr[FT];
#pragma unroll
    for (int r = 0; r < FT; ++r)
      if (t + r * FB < (int)total) { q[r] = a.pos[js[r]]; rr[r] = rho[js[r]]; }
#pragma unroll
    for (int r = 0; r < FT; ++r) {
      const uint32_t s = t + r * FB;
      if (s < total) LA[s] = make_float4(q[r].x, q[r].y, q[r].z, rr[r]);
    }
  }
  __syncthreads();
  const size_t stride = (size_t)c.nf;
  for (uint32_t i = B.own_begin + t; i < B.own_end; i += FB) {
    const float4 p = a.pos[i];
    const uint32_t nF = a.cnt[i] & 0xffffu;
    float nx = 0.f, ny = 0.f, nz = 0.f;
    auto pair = [&](float qx, float qy, float qz, float rj, bool live) {
      const float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
      const float s = (c.m_f / rj) * cubic_grad_c(c.K, rx * rx + ry * ry + rz * rz, live);
      nx += s * rx; ny += s * ry; nz += s * rz;
    };
    if (in_lds) {
      for_fluid_groups<4>(a.nbr8, stride, i, nF, [&](uint32_t s) { return LA[s]; },
                          [&](const float4& q, bool live) { pair(q.x, q.y, q.z, live ? q.w : 1.f, live); });
    } else {
      for_fluid_slots_rolled(a.nbr8, stride, i, nF, [&](uint32_t s) {
        const uint32_t j = tile_global<FB>(T, s);
        const float4 q = a.pos[j];
        pair(q.x, q.y, q.z, rho[j], true);
      });
    }
    const float h = c.K.h;
    nrm[i] = make_float4(h * nx, h * ny, h * nz, 0.f);
  }
}

// He 2014 colour field (PASS 0) and |grad c|^2 (PASS 1).  LDS per slot: x, y, z
// and m / rho_j (PASS 0) or (m / rho_j) c_j (PASS 1).
template <int PASS>
__global__ void __launch_bounds__(FB) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_he_color_t(PassArgs a, const float* __restrict__ rho, const float* __restrict__ color, float* __restrict__ out) {
  __shared__ TileHdr<FB> T;
  extern __shared__ float4 lds4[];
  const StepConst& c = a.c;
  const uint32_t brick = xcd_block();
  BrickView B;
  if (!brick_open<FB>(a.g, a.fstart, a.thdr, brick, a.cap, T, B)) return;
  const uint32_t total = B.total;
  const bool in_lds = B.in_lds;
  float4* LA = lds4;
  const int t = threadIdx.x;
  auto weight = [&](uint32_t j) { return PASS == 0 ? c.m_f / rho[j] : (c.m_f / rho[j]) * color[j]; };
  if (in_lds) {
    uint32_t js[FT];
    map_indices<FB>(a.tmap, brick, total, js);
    float4 q[FT];
    float ww[FT];
#pragma unroll
    for (int r = 0; r < FT; ++r)
      if (t + r * FB < (int)total) { q[r] = a.pos[js[r]]; ww[r] = weight(js[r]); }
#pragma unroll
    for (int r = 0; r < FT; ++r) {
      const uint32_t s = t + r * FB;
      if (s < total) LA[s] = make_float4(q[r].x, q[r].y, q[r].z, ww[r]);
    }
  }
  __syncthreads();
  const size_t stride = (size_t)c.nf;
  for (uint32_t i = B.own_begin + t; i < B.own_end; i += FB) {
    const float4 p = a.pos[i];
    const uint32_t cc = a.cnt[i];
    const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
    float sw = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
    auto pair = [&](float qx, float qy, float qz, float w) {
      const float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
      const float r2 = rx * rx + ry * ry + rz * rz;
      if (PASS == 0) {
        sw += w * cubic_W(c.K, sqrtf(r2));
      } else {
        const float s = w * cubic_grad_c(c.K, r2);
        gx += s * rx; gy += s * ry; gz += s * rz;
      }
    };
    if (in_lds) {
      for_fluid_slots(a.nbr8, stride, i, nF, [&](uint32_t s) {
        const float4 q = LA[s];
        pair(q.x, q.y, q.z, q.w);
      });
    } else {
      for_fluid_slots_rolled(a.nbr8, stride, i, nF, [&](uint32_t s) {
        const uint32_t j = tile_global<FB>(T, s);
        const float4 q = a.pos[j];
        pair(q.x, q.y, q.z, weight(j));
      });
    }
    if (PASS == 0) {
      float sb = 0.f;
      for_boundary_groups(a.nbrB, stride, i, nB, [&](uint32_t j) { return a.bpos[j]; },
                          [&](const float4& q, uint32_t) {
                            const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
                            sb += q.w * cubic_W(c.K, sqrtf(rx * rx + ry * ry + rz * rz));
                          });
      out[i] = ((c.m_f / rho[i]) * c.K.w0 + sw) + sb;
    } else {
      const float ic = 1.f / color[i];
      gx *= ic; gy *= ic; gz *= ic;
      out[i] = gx * gx + gy * gy + gz * gz;
    }
  }
}

// Accelerations a = g + surface tension + viscosity.  LDS per slot: x, y, z,
// rho | vx, vy, vz, nx (Akinci 2013) or gradC2 (He 2014) | ny, nz.
__global__ void __launch_bounds__(FB) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_forces_t(PassArgs a, const float* __restrict__ rho, const float4* __restrict__ nrm, const float* __restrict__ gc2,
           float4* __restrict__ acc, RigidDev* __restrict__ rig, ForceConst f) {
  __shared__ TileHdr<FB> T;
  extern __shared__ float4 lds4[];
  const StepConst& c = a.c;
  const uint32_t brick = xcd_block();
  const int cap = a.cap;
  BrickView B;
  if (!brick_open<FB>(a.g, a.fstart, a.thdr, brick, cap, T, B)) return;
  const uint32_t total = B.total;
  const bool in_lds = B.in_lds;
  float4* LA = lds4;
  float4* LB = lds4 + cap;
  float2* LC = reinterpret_cast<float2*>(lds4 + 2 * cap);
  const int t = threadIdx.x;
  if (in_lds) {
    uint32_t js[FT];
    map_indices<FB>(a.tmap, brick, total, js);
    float4 q[FT], v[FT], n[FT];
    float rr[FT];
#pragma unroll
    for (int r = 0; r < FT; ++r)
      if (t + r * FB < (int)total) {
        q[r] = a.pos[js[r]];
        v[r] = a.vel[js[r]];
        rr[r] = rho[js[r]];
        n[r] = f.st_method == 2 ? nrm[js[r]]
               : make_float4(f.st_method == 3 ? gc2[js[r]] : 0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
    for (int r = 0; r < FT; ++r) {
      const uint32_t s = t + r * FB;
      if (s < total) {
        LA[s] = make_float4(q[r].x, q[r].y, q[r].z, rr[r]);
        LB[s] = make_float4(v[r].x, v[r].y, v[r].z, n[r].x);
        LC[s] = make_float2(n[r].y, n[r].z);
      }
    }
  }
  __syncthreads();
  const size_t stride = (size_t)c.nf;
  const float invH = 1.f / c.dt;
  const bool he = f.st_method == 3;
  // the particle loop runs the same number of rounds on every lane of a wave, so
  // He 2014's boundary reactions are reduced over the wave at the end of a round
  for (uint32_t base = B.own_begin; base < B.own_end; base += FB) {
    const uint32_t i = base + t;
    int cb = -1;
    double F0 = 0, F1 = 0, F2 = 0, T0 = 0, T1 = 0, T2 = 0;
    if (i < B.own_end) {
    const float4 p = a.pos[i];
    const float4 vi = a.vel[i];
    const float ri = rho[i];
    const float4 ni = f.st_method == 2 ? nrm[i]
                      : make_float4(he ? gc2[i] : 0.f, 0.f, 0.f, 0.f);
    const float hf = 0.25f * f.st_k / ri;  // He 2014's factor k / (4 rho_i)
    const uint32_t cc = a.cnt[i];
    const uint32_t nF = cc & 0xffffu, nB = cc >> 16;
    float sx = 0.f, sy = 0.f, sz = 0.f;  // surface tension
    float vx = 0.f, vy = 0.f, vz = 0.f;  // viscosity
    auto pair = [&](float qx, float qy, float qz, float rj, float wx, float wy, float wz, float njx, float njy,
                    float njz) {
      const float rx = p.x - qx, ry = p.y - qy, rz = p.z - qz;
      const float r2 = rx * rx + ry * ry + rz * rz;
      if (he) {
        // njx holds gradC2_j
        const float s = hf * (c.m_f / rj) * (ni.x + njx) * cubic_grad_c(c.K, r2);
        sx += s * rx; sy += s * ry; sz += s * rz;
      } else if (f.st_method == 2) {
        const float K_ij = 2.f * c.rho0 / (ri + rj);
        float ax = 0.f, ay = 0.f, az = 0.f;
        if (r2 > 1.0e-9f) {
          const float il = 1.f / sqrtf(r2);
          const float s = f.st_k * c.m_f * cohesion_w(f, r2);
          ax -= s * (rx * il); ay -= s * (ry * il); az -= s * (rz * il);
        }
        ax -= f.st_k * (ni.x - njx); ay -= f.st_k * (ni.y - njy); az -= f.st_k * (ni.z - njz);
        sx += K_ij * ax; sy += K_ij * ay; sz += K_ij * az;
      }
      if (f.visc_mode == 2) {
        const float w = cubic_W(c.K, sqrtf(r2));
        const float s = invH * f.visc * (c.m_f / rj) * w;
        vx -= s * (vi.x - wx); vy -= s * (vi.y - wy); vz -= s * (vi.z - wz);
      } else if (f.visc_mode == 1) {
        // d nu (m_j / rho_j) (v_ij . r_ij) / (|r_ij|^2 + 0.01 h^2) gradW_ij, d = 10
        const float dot = (vi.x - wx) * rx + (vi.y - wy) * ry + (vi.z - wz) * rz;
        const float s = (10.f * f.visc * (c.m_f / rj) * dot) / (r2 + 0.01f * (f.h * f.h));
        const float gc = s * cubic_grad_c(c.K, r2);
        vx += gc * rx; vy += gc * ry; vz += gc * rz;
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
        const uint32_t j = tile_global<FB>(T, s);
        const float4 q = a.pos[j], w = a.vel[j];
        const float4 nj = f.st_method == 2 ? nrm[j] : make_float4(he ? gc2[j] : 0.f, 0.f, 0.f, 0.f);
        pair(q.x, q.y, q.z, rho[j], w.x, w.y, w.z, nj.x, nj.y, nj.z);
      });
    }
    float bx = 0.f, by = 0.f, bz = 0.f;  // boundary adhesion (Akinci 2013) or boundary term (He 2014)
    if (he && f.st_kb != 0.f) {
      const float hb = 0.25f * f.st_kb / ri * ni.x;  // kb gradC2_i / (4 rho_i)
      for_boundary_groups(a.nbrB, stride, i, nB, [&](uint32_t j) { return a.bpos[j]; },
                          [&](const float4& q, uint32_t j) {
                            const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
                            const float s = hb * q.w * cubic_grad_c(c.K, rx * rx + ry * ry + rz * rz);
                            const float ax = s * rx, ay = s * ry, az = s * rz;
                            bx += ax; by += ay; bz += az;
                            if ((int)j >= a.nb_static) {
                              const int body = (int)a.bx0[j].w;
                              if (rig[body].dynamic) {
                                if (body != cb && cb >= 0) {
                                  atomicAdd(&rig[cb].F[0], F0); atomicAdd(&rig[cb].F[1], F1);
                                  atomicAdd(&rig[cb].F[2], F2);
                                  atomicAdd(&rig[cb].T[0], T0); atomicAdd(&rig[cb].T[1], T1);
                                  atomicAdd(&rig[cb].T[2], T2);
                                  F0 = F1 = F2 = T0 = T1 = T2 = 0.0;
                                }
                                cb = body;
                                const float fx = -c.m_f * ax, fy = -c.m_f * ay, fz = -c.m_f * az;
                                const double rxd = (double)q.x - rig[body].x[0], ryd = (double)q.y - rig[body].x[1],
                                             rzd = (double)q.z - rig[body].x[2];
                                F0 += fx; F1 += fy; F2 += fz;
                                T0 += ryd * fz - rzd * fy; T1 += rzd * fx - rxd * fz; T2 += rxd * fy - ryd * fx;
                              }
                            }
                          });
    }
    if (f.st_method == 2 && f.st_kb != 0.f) {
      for_boundary_groups(a.nbrB, stride, i, nB, [&](uint32_t j) { return a.bpos[j]; },
                          [&](const float4& q, uint32_t) {
                            const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
                            const float r2 = rx * rx + ry * ry + rz * rz;
                            if (r2 > 1.0e-9f) {
                              const float il = 1.f / sqrtf(r2);
                              const float s = f.st_kb * c.rho0 * q.w * adhesion_w(f, r2);
                              bx -= s * (rx * il); by -= s * (ry * il); bz -= s * (rz * il);
                            }
                          });
    }
    acc[i] = make_float4(((c.g[0] + sx) + bx) + vx, ((c.g[1] + sy) + by) + vy, ((c.g[2] + sz) + bz) + vz, 0.f);
    }
    if (he && f.st_kb != 0.f) {
      // wave-level FP64 reduction of the round's reactions, one body at a time
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
}

// CFL speed bound: max over fluid particles of |v + a dt|^2 (a = acc, or g
// when acc is null) and over boundary samples of |v_b|^2 (samples of static
// bodies have v_b = 0 and never raise the bound), as float bits (non-negative
// floats order as unsigned integers); *out holds the bits of 0.1 beforehand.
__global__ void k_cfl_max(StepConst c, const float4* __restrict__ vel, const float4* __restrict__ acc, int nb,
                          const float4* __restrict__ bvel, uint32_t* __restrict__ out) {
  float m = 0.f;
  const uint32_t n = (uint32_t)c.nf + (uint32_t)nb;
  for (uint32_t k = gid(); k < n; k += gridDim.x * blockDim.x) {
    float s;
    if (k < (uint32_t)c.nf) {
      const float4 v = vel[k];
      const float4 g = acc ? acc[k] : make_float4(c.g[0], c.g[1], c.g[2], 0.f);
      const float x = v.x + g.x * c.dt, y = v.y + g.y * c.dt, z = v.z + g.z * c.dt;
      s = x * x + y * y + z * z;
    } else {
      const float4 v = bvel[k - (uint32_t)c.nf];
      s = v.x * v.x + v.y * v.y + v.z * v.z;
    }
    m = fmaxf(m, s);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63u) == 0u) atomicMax(out, __float_as_uint(m));
}

// v += dt a (the advection of the velocities with the non-pressure
// accelerations), with the divergence warm start's end scaling of kappa_v.
__global__ void k_advect_acc(StepConst c, float4* __restrict__ vel, const float4* __restrict__ acc,
                             float* __restrict__ kappav, float kv_scale) {
  const uint32_t i = gid();
  if (i >= (uint32_t)c.nf) return;
  float4 v = vel[i];
  const float4 a = acc[i];
  v.x += c.dt * a.x; v.y += c.dt * a.y; v.z += c.dt * a.z;
  vel[i] = v;
  if (kv_scale != 1.f) kappav[i] *= kv_scale;
}

static inline unsigned nblk_f(int n) { return (unsigned)((n + DFR_BLOCK - 1) / DFR_BLOCK); }

__global__ void k_set_u32(uint32_t* __restrict__ p, uint32_t v) { *p = v; }

void launch_normals_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                      NbrBufs& nb, float4* nrm, hipStream_t s) {
  hipLaunchKernelGGL(k_normals_t, dim3(nbricks(g)), dim3(FB), sizeof(float4) * (size_t)nb.cap, s,
                     make_args(c, g, gb, f, b, nb), f.rho, nrm);
}

void launch_he_color_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                       NbrBufs& nb, float* color, float* gc2, hipStream_t s) {
  const PassArgs a = make_args(c, g, gb, f, b, nb);
  const size_t shm = sizeof(float4) * (size_t)nb.cap;
  hipLaunchKernelGGL(k_he_color_t<0>, dim3(nbricks(g)), dim3(FB), shm, s, a, f.rho, nullptr, color);
  hipLaunchKernelGGL(k_he_color_t<1>, dim3(nbricks(g)), dim3(FB), shm, s, a, f.rho, color, gc2);
}

void launch_forces_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                     NbrBufs& nb, const float4* nrm, const float* gc2, float4* acc, RigidDev* rig,
                     const ForceConst& fc, hipStream_t s) {
  const size_t shm = (2 * sizeof(float4) + sizeof(float2)) * (size_t)nb.cap;
  hipLaunchKernelGGL(k_forces_t, dim3(nbricks(g)), dim3(FB), shm, s, make_args(c, g, gb, f, b, nb), f.rho, nrm, gc2,
                     acc, rig, fc);
}

void launch_cfl_max(const StepConst& c, const float4* vel, const float4* acc, int nb, const float4* bvel,
                    uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, out, 0x3DCCCCCDu);  // 0.1f, Simulation.cpp:552
  const int n = c.nf + nb;
  const int blocks = std::max(1, std::min((n + 255) / 256, 4096));
  hipLaunchKernelGGL(k_cfl_max, dim3(blocks), dim3(256), 0, s, c, vel, acc, nb, bvel, out);
}

void launch_advect_acc(const StepConst& c, FluidBufs& f, const float4* acc, float kappav_scale, hipStream_t s) {
  hipLaunchKernelGGL(k_advect_acc, dim3(nblk_f(c.nf)), dim3(DFR_BLOCK), 0, s, c, f.vel, acc, f.kappav, kappav_scale);
}

void set_forces_kernel_attributes() {
  (void)hipFuncSetAttribute((const void*)k_forces_t, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)((2 * sizeof(float4) + sizeof(float2)) * TILE_CAP));
}

}  // namespace dfr
