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
//
// With the manager across bodies (useRigidGradientManager) each contact sample
// also leaves a record of its gradient ingredients (k_contact_forces), from
// which k_contact_apply forms the solver's grad_net_{force,torque}_to_*
// sums (solveRigidContactPenalty :471-553) and hands them to the manager in
// place of the step's (update_rigid_body_gradient_manager, :1341-1370).
#include "dfr_device.h"
#include "dfr_kernels.h"
#include "dfr_rigid_math.h"

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
  // gradient ingredients over the other bodies' neighbours: d density / d x_R with the
  // simulation's kernel gradient, d density / d q_Rk, the weighted r_k, the last body
  float gx = 0.f, gy = 0.f, gz = 0.f, gk[4] = {0.f, 0.f, 0.f, 0.f}, ax = 0.f, ay = 0.f, az = 0.f;
  int rr_last = -1;
  const bool grads = cc.grec != nullptr;
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
          if (grads) {
            const RigidDev& bk = rig[Rk];
            rr_last = max(rr_last, Rk);
            const float gc = vol0[k] * cubic_grad_c(cc.Ks, rx * rx + ry * ry + rz * rz);
            const float wx = gc * rx, wy = gc * ry, wz = gc * rz;
            gx += wx; gy += wy; gz += wz;
            const float4 x0k = bx0[k];
            const float qk[4] = {(float)bk.q[0], (float)bk.q[1], (float)bk.q[2], (float)bk.q[3]};
            const DM<float, 3, 4> Jk = grad_Rqp<float>(qk, DV3<float>{x0k.x, x0k.y, x0k.z}, 1.f);
#pragma unroll
            for (int c = 0; c < 4; ++c) gk[c] -= wx * Jk(0, c) + wy * Jk(1, c) + wz * Jk(2, c);
            ax += (float)((double)q.x - bk.x[0]) * W;
            ay += (float)((double)q.y - bk.x[1]) * W;
            az += (float)((double)q.z - bk.x[2]) * W;
          }
        }
      }
    }
  }
  const bool contact = __ballot(other) != 0ull;
  dens = wave_sum_f(dens); sw = wave_sum_f(sw);
  sx = wave_sum_f(sx); sy = wave_sum_f(sy); sz = wave_sum_f(sz);
  vx = wave_sum_f(vx); vy = wave_sum_f(vy); vz = wave_sum_f(vz);
  if (grads) {
    gx = wave_sum_f(gx); gy = wave_sum_f(gy); gz = wave_sum_f(gz);
#pragma unroll
    for (int c = 0; c < 4; ++c) gk[c] = wave_sum_f(gk[c]);
    ax = wave_sum_f(ax); ay = wave_sum_f(ay); az = wave_sum_f(az);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) rr_last = max(rr_last, __shfl_xor(rr_last, off, 64));
  }
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
      if (grads) {
        const size_t nd_ = (size_t)nd;
        cc.grec[0 * nd_ + d] = make_float4(nx, ny, nz, (float)(rr_last < 0 ? R : rr_last));
        cc.grec[1 * nd_ + d] = make_float4(vr.x - vx * isw, vr.y - vy * isw, vr.z - vz * isw, 0.f);
        cc.grec[2 * nd_ + d] = make_float4(gx, gy, gz, 0.f);
        cc.grec[3 * nd_ + d] = make_float4(gk[0], gk[1], gk[2], gk[3]);
        cc.grec[4 * nd_ + d] = make_float4(ax * isw, ay * isw, az * isw, 0.f);
        cc.grec[5 * nd_ + d] = make_float4(fnx, fny, fnz, 0.f);
      }
    }
  }
  cf[d] = fo;
  ct[d] = to;
}

// The contact solver's gradient sums of one contact sample (solveRigidContactPenalty,
// RigidContactSolver.cpp:471-553), added to the rows of body R: (R, R) and, when its
// other-body neighbours end on body Q != R, (R, Q).  Layout of RigidDev::net.
__device__ void add_contact_gradients(double* NRR, double* NRQ, const RigidDev& b, const float4* rec,
                                      const float4& fx, const float4& p, const float4& x0, float beta, float mu) {
  const V3d n = {rec[0].x, rec[0].y, rec[0].z}, vr = {rec[1].x, rec[1].y, rec[1].z};
  const V3d gdx = {rec[2].x, rec[2].y, rec[2].z}, avg = {rec[4].x, rec[4].y, rec[4].z};
  const V3d fn = {rec[5].x, rec[5].y, rec[5].z}, f = {fx.x, fx.y, fx.z};
  const double gdqk[4] = {rec[3].x, rec[3].y, rec[3].z, rec[3].w};
  const V3d r = {(double)p.x - b.x[0], (double)p.y - b.x[1], (double)p.z - b.x[2]};
  const V3d gfn = {-(double)beta * n.x, -(double)beta * n.y, -(double)beta * n.z};
  const double fnn = sqrt(fn.x * fn.x + fn.y * fn.y + fn.z * fn.z);
  const double vn = sqrt(vr.x * vr.x + vr.y * vr.y + vr.z * vr.z);
  const V3d u = vn > 0.0 ? V3d{vr.x / vn, vr.y / vn, vr.z / vn} : vr;
  const double sff = -(double)mu * (fn.x * gfn.x + fn.y * gfn.y + fn.z * gfn.z) / fnn;
  const V3d gf = {gfn.x + u.x * sff, gfn.y + u.y * sff, gfn.z + u.z * sff};
  const M34d Jq = grad_Rqp<double>(b.q, V3d{(double)x0.x, (double)x0.y, (double)x0.z}, 1.0);
  double gdq[4];
  for (int c = 0; c < 4; ++c) gdq[c] = gdx.x * Jq(0, c) + gdx.y * Jq(1, c) + gdx.z * Jq(2, c);
  const double gfa[3] = {gf.x, gf.y, gf.z}, gxa[3] = {gdx.x, gdx.y, gdx.z};
  M3d Gx;
  M34d Gq, Gqk;
  for (int i = 0; i < 3; ++i) {
    for (int c = 0; c < 3; ++c) Gx(i, c) = gfa[i] * gxa[c];
    for (int c = 0; c < 4; ++c) { Gq(i, c) = gfa[i] * gdq[c]; Gqk(i, c) = gfa[i] * gdqk[c]; }
  }
  const double ua[3] = {u.x, u.y, u.z};
  const double sv = -(double)mu * fnn / vn;
  M3d Gvr;
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) Gvr(i, c) = (((i == c) ? 1.0 : 0.0) - ua[i] * ua[c]) * sv;
  const M3d Sr = skew(r);
  const M3d Gw = mm(Gvr, Sr.tr());
  const M3d GwK = mm(Gvr, skew(avg).tr()) * -1.0;
  const M3d SGx = mm(Sr, Gx);
  const M34d Tq = mm(Sr, Gq) + mm(skew(f).tr(), Jq);
  auto add3 = [](double* o, const M3d& m) { for (int k = 0; k < 9; ++k) o[k] += m.a[k]; };
  auto add34 = [](double* o, const M34d& m) { for (int k = 0; k < 12; ++k) o[k] += m.a[k]; };
  add3(NRR + 0, Gvr); add3(NRR + 9, Gx); add3(NRR + 18, Gw); add3(NRR + 36, SGx);
  add34(NRR + 54, Gq); add34(NRR + 66, Tq);
  if (NRQ) {
    add3(NRQ + 0, Gvr * -1.0); add3(NRQ + 18, GwK); add3(NRQ + 9, Gx * -1.0); add3(NRQ + 36, SGx * -1.0);
    add34(NRQ + 54, Gqk); add34(NRQ + 66, mm(Sr, Gqk));
  }
}

// Gravity, then one addForce / addTorque per contact sample in sample order
// (Dynamic3dRigidBody.h:113-142, 228-233), one thread per body.
__global__ void k_contact_apply(ContactConst cc, int nbodies, const float4* __restrict__ bpos,
                                const float4* __restrict__ bx0, const uint32_t* __restrict__ dyn_idx, int nd,
                                const float4* __restrict__ cf, const float4* __restrict__ ct,
                                RigidDev* __restrict__ rig) {
  const int bi = threadIdx.x;
  if (bi >= nbodies) return;
  RigidDev& b = rig[bi];
  if (!b.dynamic) return;
  const double dt = cc.dt;
  double* Nrow = cc.mgr_net ? cc.mgr_net + (size_t)bi * nbodies * 78 : nullptr;
  if (Nrow)  // the manager's sums of this body's row are replaced by the contact solver's
    for (int Q = 0; Q < nbodies; ++Q)
      if (rig[Q].dynamic)
        for (int k = 0; k < 78; ++k) Nrow[(size_t)Q * 78 + k] = 0.0;
  for (int k = 0; k < 3; ++k) b.v[k] += (cc.g[k] * b.mass) * (b.invMass * dt);
  for (int d = 0; d < nd; ++d) {
    if ((int)bx0[dyn_idx[d]].w != bi) continue;
    const float4 f = cf[d];
    if (f.w == 0.f) continue;
    if (Nrow) {
      float4 rec[6];
      for (int k = 0; k < 6; ++k) rec[k] = cc.grec[(size_t)k * nd + d];
      const int Q = (int)rec[0].w;
      const uint32_t j = dyn_idx[d];
      add_contact_gradients(Nrow + (size_t)bi * 78, Q != bi ? Nrow + (size_t)Q * 78 : nullptr, b, rec, f, bpos[j],
                            bx0[j], cc.beta, cc.mu);
    }
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
  hipLaunchKernelGGL(k_contact_apply, dim3(1), dim3(64), 0, s, cc, nbodies, b.pos, b.x0, b.dyn_idx, b.nd, cf, ct,
                     rig);
}

}  // namespace dfr
