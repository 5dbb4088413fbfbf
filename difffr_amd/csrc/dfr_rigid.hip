// Rigid side of the step: DiffFR's reference-mode pair Jacobians, the per-body
// forward-mode chain rule, rigid integration and the boundary transform.
// Pair terms are evaluated in FP32 (as in the reference); every sum over pairs,
// the 6x6 chain products and the rigid state are FP64.
#include "dfr_rigid_math.h"

namespace dfr {

// ---------------------------------------------------------------------------
// Pair Jacobians: TimeStepDiffDFSPH::computeRigidBodyGradient (:1226-1279) and
// computeGradient (:1281-1482) with its helpers (:1485-1694).  One thread per
// sample j of a dynamic, non-animated body; j walks its fluid neighbours i.
// d(b_i)/d(v_i) = (dt) * sum_j V_j gradW_ij equals -(dt) * sum_grad_p_k_i over
// the same neighbour set, so it is read from the factor pass instead of being
// re-summed for every pair.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(DFR_BLOCK)
k_gradient(StepConst c, GradConst gc, int k_first, int k_count, int body, const uint32_t* __restrict__ dyn_idx,
           int nd, const uint32_t* __restrict__ bf, const uint32_t* __restrict__ bfcnt, const float4* __restrict__ pos,
           const float4* __restrict__ vel, const float* __restrict__ badv, const float* __restrict__ alpha,
           const float4* __restrict__ s, const float4* __restrict__ bpos, const float4* __restrict__ bvel,
           const float4* __restrict__ bx0, RigidDev* __restrict__ rig, const int* __restrict__ stop) {
  // an iteration enqueued after the solve converged (device-side loop) adds nothing
  if (stop && *stop) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[78];
#pragma unroll
  for (int k = 0; k < 78; ++k) acc[k] = 0.0;
  if (t < k_count) {
    const int kk = k_first + t;
    const uint32_t j = dyn_idx[kk];
    const float4 bj = bpos[j];
    const float4 vjb = bvel[j];
    const float4 x0 = bx0[j];
    const DV3<float> xj = {bj.x, bj.y, bj.z};
    const float Vj = bj.w;
    const RigidDev& rb = rig[body];
    const float qf[4] = {(float)rb.q[0], (float)rb.q[1], (float)rb.q[2], (float)rb.q[3]};
    const DV3<float> rj = {(float)((double)bj.x - rb.x[0]), (float)((double)bj.y - rb.x[1]),
                           (float)((double)bj.z - rb.x[2])};
    const DV3<float> om = {(float)rb.w[0], (float)rb.w[1], (float)rb.w[2]};
    const DM<float, 3, 4> Jq = grad_Rqp<float>(qf, DV3<float>{x0.x, x0.y, x0.z}, 1.f);
    const DM3<float> Srj = skew(rj);
    const float invH = 1.f / c.dt;
    const uint32_t nn = bfcnt[kk];
    for (uint32_t n = 0; n < nn; ++n) {
      const uint32_t i = bf[(size_t)n * nd + kk];
      const float bb = badv[i];
      const float b_i = (gc.mode == 0) ? bb - 1.f : bb;
      const float fac = alpha[i];
      const float ki = b_i * fac * gc.coeff_trunc;  // truncated coefficient, :1266
      const float4 pi = pos[i];
      const float4 vi4 = vel[i];
      const float4 si4 = s[i];
      const float rx = pi.x - xj.x, ry = pi.y - xj.y, rz = pi.z - xj.z;
      const float rl = sqrtf(rx * rx + ry * ry + rz * rz);
      const float sc = cubic_gradW_scale(c.K, rl);
      const DV3<float> gW = {sc * rx, sc * ry, sc * rz};
      const DV3<float> gp = {-Vj * gW.x, -Vj * gW.y, -Vj * gW.z};  // grad_p_j
      float ga, gb;
      cubic_gradGradW_ab(c.K, rl, ga, gb);
      DM3<float> GG = outer(DV3<float>{rx, ry, rz}, DV3<float>{rx, ry, rz}) * gb;
      GG(0, 0) += ga; GG(1, 1) += ga; GG(2, 2) += ga;
      const DM3<float> H = GG * Vj;  // grad2_p_j_to_xj
      const float a = -c.dt * 1.f * ki;
      const DV3<float> velChange = {a * gp.x, a * gp.y, a * gp.z};
      const DV3<float> force = {-c.m_f * velChange.x * invH, -c.m_f * velChange.y * invH, -c.m_f * velChange.z * invH};
      const DV3<float> vij = {vi4.x - vjb.x, vi4.y - vjb.y, vi4.z - vjb.z};
      const DV3<float> GGv = mv(GG, vij);
      DV3<float> gbx = {0.f, 0.f, 0.f}, gbv = {0.f, 0.f, 0.f}, gbvi = {0.f, 0.f, 0.f};
      const bool on = (gc.mode == 0) ? (bb > 1.f) : (bb > 0.f);
      if (on) {
        if (gc.mode == 0) {
          gbx = {-Vj * gW.x / c.rho0 - c.dt * Vj * GGv.x, -Vj * gW.y / c.rho0 - c.dt * Vj * GGv.y,
                 -Vj * gW.z / c.rho0 - c.dt * Vj * GGv.z};
          gbv = {-c.dt * Vj * gW.x, -c.dt * Vj * gW.y, -c.dt * Vj * gW.z};
          gbvi = {-si4.x * c.dt, -si4.y * c.dt, -si4.z * c.dt};
        } else {
          gbx = {-(GGv.x * Vj), -(GGv.y * Vj), -(GGv.z * Vj)};
          gbv = {-Vj * gW.x, -Vj * gW.y, -Vj * gW.z};
          gbvi = {-si4.x, -si4.y, -si4.z};
        }
      }
      DV3<float> gfx = {0.f, 0.f, 0.f};
      if (fac < 0.f) {
        const DV3<float> Hs = mv(H, DV3<float>{si4.x, si4.y, si4.z});
        const float f2 = 2.f * fac * fac;
        gfx = {Hs.x * f2, Hs.y * f2, Hs.z * f2};
      }
      const float co = gc.coeff;
      const DV3<float> gkx = {(gbx.x * fac + gfx.x * b_i) * co, (gbx.y * fac + gfx.y * b_i) * co,
                              (gbx.z * fac + gfx.z * b_i) * co};
      const DM3<float> dvc_dxj = (outer(gp, gkx) + H * ki) * -1.f;
      const DV3<float> gkv = {gbv.x * fac * co, gbv.y * fac * co, gbv.z * fac * co};
      const DM3<float> dvc_dvj = outer(gp, gkv) * -1.f;
      const DV3<float> gkvi = {gbvi.x * fac * co, gbvi.y * fac * co, gbvi.z * fac * co};
      const DM3<float> dvc_dvi = outer(gp, gkvi) * -1.f;
      const DM3<float> gFx = dvc_dxj * (-c.m_f);
      const DM3<float> gFv = mm(inverse3(DM3<float>::eye() - dvc_dvi * c.dt), dvc_dvj * (-c.m_f));
#pragma unroll
      for (int k = 0; k < 9; ++k) { acc[k] += gFv.a[k]; acc[9 + k] += gFx.a[k]; }
      if (gc.optimize_rotation) {
        const DM<float, 3, 4> gFq = mm(gFx, Jq) + mm(mm(gFv, skew(om)), Jq);
        const DM3<float> gFw = mm(gFv, Srj.tr());
        const DM<float, 3, 4> gTq = mm(Srj, gFq) + mm(skew(force).tr(), Jq);
        const DM3<float> gTw = mm(Srj, gFw);
        const DM3<float> gTx = mm(Srj, gFx);
        const DM3<float> gTv = mm(Srj, gFv);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          acc[18 + k] += gFw.a[k]; acc[27 + k] += gTw.a[k]; acc[36 + k] += gTx.a[k]; acc[45 + k] += gTv.a[k];
        }
#pragma unroll
        for (int k = 0; k < 12; ++k) { acc[54 + k] += gFq.a[k]; acc[66 + k] += gTq.a[k]; }
      }
    }
  }
  const bool any = __ballot(t < k_count) != 0ull;
  if (!any) return;
#pragma unroll
  for (int k = 0; k < 78; ++k) {
    const double v = wave_sum_d(acc[k]);
    if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(&rig[body].net[k], v);
  }
}

// ---------------------------------------------------------------------------
// Per-body chain rule (BoundaryModel_Akinci2012::perform_chain_rule, :543-889),
// rigid velocity step (BoundarySimulator::updateBoundaryForces,
// BoundarySimulator.cpp:10-36; Dynamic3dRigidBody::addForce/addTorque/addGravity,
// Dynamic3dRigidBody.h:113-142, :228-233) and position step
// (Dynamic3dRigidBody::animate, :143-154; updateInertia, :215-219).  FP64.
// ---------------------------------------------------------------------------

__device__ inline M3d grad_inertia_v(const RigidDev& b, const M3d& R, const M3d& I0, V3d v, const M43d& Jq) {
  const V3d u = mv(I0, mv(R.tr(), v));
  const M3d a = mm(grad_Rqp<double>(b.q, u, 1.0), Jq);
  const M3d c = mm(mm(R, I0), mm(grad_Rqp<double>(b.q, v, -1.0), Jq));
  return a + c;
}

__device__ void chain_rule(RigidDev& b, const RigidStepConst& rc) {
  const double dt = rc.dt;
  const int mode = rc.gradient_mode;
  const double invMass = 1.0 / (b.mass + 1e-10);
  const M3d I = ld3(b.I), Iinv = ld3(b.Iinv), I0 = ld3(b.I0);
  const M3d R = quat_to_R(b.q);
  const V3d om = {b.w[0], b.w[1], b.w[2]};
  M3d Jxv = ld3(b.J + 0), Jxw = ld3(b.J + 9), Jvv = ld3(b.J + 18), Jvw = ld3(b.J + 27), Jwv = ld3(b.J + 36),
      Jww = ld3(b.J + 45);
  M43d Jqv = ld43(b.J + 54), Jqw = ld43(b.J + 66);
  const M3d nFv = ld3(b.net + 0), nFx = ld3(b.net + 9), nFw = ld3(b.net + 18), nTw = ld3(b.net + 27),
            nTx = ld3(b.net + 36), nTv = ld3(b.net + 45);
  const M34d nFq = ld34(b.net + 54), nTq = ld34(b.net + 66);
  V3d temp = {0, 0, 0};
  const V3d Iom = mv(I, om);
  if (mode == 0 || mode == 1) {
    M3d dFv0, dFw0;
    if (mode == 0) {
      dFv0 = mm(nFx, Jxv) + mm(nFv, Jvv) + mm(nFq, Jqv) + mm(nFw, Jwv);
      dFw0 = mm(nFx, Jxw) + mm(nFv, Jvw) + mm(nFq, Jqw) + mm(nFw, Jww);
    } else {
      dFv0 = mm(nFv, Jvv) + mm(nFw, Jwv);
      dFw0 = mm(nFv, Jvw) + mm(nFw, Jww);
    }
    Jvv = Jvv + dFv0 * (dt * invMass);
    Jvw = Jvw + dFw0 * (dt * invMass);
    Jxv = Jxv + Jvv * dt;
    Jxw = Jxw + Jvw * dt;
    const V3d T = {b.T[0], b.T[1], b.T[2]};
    M3d dTw0, dTv0;
    if (mode == 0) {
      dTw0 = mm(nTx, Jxw) + mm(nTv, Jvw) + mm(nTq, Jqw) + mm(nTw, Jww);
      dTv0 = mm(nTx, Jxv) + mm(nTv, Jvv) + mm(nTq, Jqv) + mm(nTw, Jwv);
    } else {
      dTw0 = mm(nTv, Jvw) + mm(nTw, Jww);
      dTv0 = mm(nTv, Jvv) + mm(nTw, Jwv);
    }
    M3d dTauw0, dTauv0;
    if (rc.gyroscopic) {
      const V3d lw = cross(Iom, om);
      temp = {lw.x + T.x, lw.y + T.y, lw.z + T.z};
      const M3d dLw_w0 = mm(skew(Iom), Jww) + mm(skew(om).tr(), mm(I, Jww) + grad_inertia_v(b, R, I0, om, Jqw));
      const M3d dLw_v0 = mm(skew(Iom), Jwv) + mm(skew(om).tr(), mm(I, Jwv) + grad_inertia_v(b, R, I0, om, Jqv));
      dTauw0 = dLw_w0 + dTw0;
      dTauv0 = dLw_v0 + dTv0;
    } else {
      temp = T;
      dTauw0 = dTw0;
      dTauv0 = dTv0;
    }
    // "+" on the d(I^-1) term as in the reference (:641-647)
    const V3d it = mv(Iinv, temp);
    const M3d gIw = mm(Iinv, grad_inertia_v(b, R, I0, it, Jqw));
    const M3d gIv = mm(Iinv, grad_inertia_v(b, R, I0, it, Jqv));
    Jww = Jww + (gIw + mm(Iinv, dTauw0)) * dt;
    Jwv = Jwv + (gIv + mm(Iinv, dTauv0)) * dt;
  } else {
    Jxv = Jxv + Jvv * dt;
    Jxw = Jxw + Jvw * dt;
    if (rc.gyroscopic) {
      temp = cross(Iom, om);
      const V3d it = mv(Iinv, temp);
      const M3d gIw = mm(Iinv, grad_inertia_v(b, R, I0, it, Jqw));
      const M3d gIv = mm(Iinv, grad_inertia_v(b, R, I0, it, Jqv));
      const M3d dLw_w0 = mm(skew(Iom), Jww) + mm(skew(om).tr(), mm(I, Jww) + grad_inertia_v(b, R, I0, om, Jqw));
      const M3d dLw_v0 = mm(skew(Iom), Jwv) + mm(skew(om).tr(), mm(I, Jwv) + grad_inertia_v(b, R, I0, om, Jqv));
      Jww = Jww + (gIw + mm(Iinv, dLw_w0)) * dt;
      Jwv = Jwv + (gIv + mm(Iinv, dLw_v0)) * dt;
    }
  }
  if (rc.optimize_rotation) {
    const V3d it = mv(Iinv, temp);
    const double nw[3] = {om.x + it.x * dt, om.y + it.y * dt, om.z + it.z * dt};
    const double p[4] = {0.0, nw[0], nw[1], nw[2]};
    double pq[4];
    quat_mul(p, b.q, pq);
    double nq[4];
    for (int k = 0; k < 4; ++k) nq[k] = b.q[k] + dt * 0.5 * pq[k];
    const double nn = sqrt(nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3] + nq[0] * nq[0]);
    double qh[4];
    for (int k = 0; k < 4; ++k) qh[k] = nq[k] / nn;
    M4d Nq;
    for (int r = 0; r < 4; ++r)
      for (int cc = 0; cc < 4; ++cc) Nq(r, cc) = ((r == cc) ? 1.0 : 0.0) - qh[r] * qh[cc];
    Nq = Nq * (1.0 / nn);
    M4d Pq;
    const double pv[16] = {p[0], -p[1], -p[2], -p[3], p[1], p[0], -p[3], p[2],
                           p[2], p[3], p[0], -p[1], p[3], -p[2], p[1], p[0]};
    for (int k = 0; k < 16; ++k) Pq.a[k] = pv[k];
    M43d Pw;
    const double* q = b.q;
    const double wv[12] = {-q[1], -q[2], -q[3], q[0], q[3], -q[2], -q[3], q[0], q[1], q[2], -q[1], q[0]};
    for (int k = 0; k < 12; ++k) Pw.a[k] = wv[k];
    Jqw = mm(Nq, Jqw + (mm(Pw, Jww) + mm(Pq, Jqw)) * (dt / 2.0));
    Jqv = mm(Nq, Jqv + (mm(Pw, Jwv) + mm(Pq, Jqv)) * (dt / 2.0));
  }
  st3(b.J + 0, Jxv); st3(b.J + 9, Jxw); st3(b.J + 18, Jvv); st3(b.J + 27, Jvw); st3(b.J + 36, Jwv);
  st3(b.J + 45, Jww); st43(b.J + 54, Jqv); st43(b.J + 66, Jqw);
}

// RigidBodyGradientManager (single-body entries R == RR; no rigid contact).
// perform_force_and_torque_chain_rule (:96-156): d(F, T)/d(v0, omega0) from the
// v and omega columns of the handed sums and the Jacobians before this step's
// update; perform_velocity_chain_rule (:276-365): the velocity chain with the
// gyroscopic term and the reference's "+" on the d(I^-1) term (:351-357).
__device__ void manager_velocity_chain(RigidDev& b, const RigidStepConst& rc) {
  const double dt = rc.dt;
  const M3d I = ld3(b.I), Iinv = ld3(b.Iinv), I0 = ld3(b.I0);
  const M3d R = quat_to_R(b.q);
  const V3d om = {b.w[0], b.w[1], b.w[2]};
  M3d Jvv = ld3(b.Jm + 18), Jvw = ld3(b.Jm + 27), Jwv = ld3(b.Jm + 36), Jww = ld3(b.Jm + 45);
  const M43d Jqv = ld43(b.Jm + 54), Jqw = ld43(b.Jm + 66);
  const M3d nFv = ld3(b.mnet + 0), nFw = ld3(b.mnet + 18), nTw = ld3(b.mnet + 27), nTv = ld3(b.mnet + 45);
  const M3d dFv0 = mm(nFv, Jvv) + mm(nFw, Jwv), dFw0 = mm(nFv, Jvw) + mm(nFw, Jww);
  const M3d dTv0 = mm(nTv, Jvv) + mm(nTw, Jwv), dTw0 = mm(nTv, Jvw) + mm(nTw, Jww);
  const V3d T = {b.T[0], b.T[1], b.T[2]};
  Jvv = Jvv + dFv0 * (dt * b.invMass);
  Jvw = Jvw + dFw0 * (dt * b.invMass);
  V3d temp = T;
  M3d dTauw0 = dTw0, dTauv0 = dTv0;
  if (rc.gyroscopic) {
    const V3d Iom = mv(I, om);
    const V3d lw = cross(Iom, om);
    temp = {lw.x + T.x, lw.y + T.y, lw.z + T.z};
    dTauw0 = mm(skew(Iom), Jww) + mm(skew(om).tr(), mm(I, Jww) + grad_inertia_v(b, R, I0, om, Jqw)) + dTw0;
    dTauv0 = mm(skew(Iom), Jwv) + mm(skew(om).tr(), mm(I, Jwv) + grad_inertia_v(b, R, I0, om, Jqv)) + dTv0;
  }
  const V3d it = mv(Iinv, temp);
  Jww = Jww + (mm(Iinv, grad_inertia_v(b, R, I0, it, Jqw)) + mm(Iinv, dTauw0)) * dt;
  Jwv = Jwv + (mm(Iinv, grad_inertia_v(b, R, I0, it, Jqv)) + mm(Iinv, dTauv0)) * dt;
  st3(b.Jm + 18, Jvv); st3(b.Jm + 27, Jvw); st3(b.Jm + 36, Jwv); st3(b.Jm + 45, Jww);
}

// perform_position_and_rotation_chain_rule (:367-423): after the velocity
// update (the new omega), before the position update (the current q).
__device__ void position_chain(double* Jm, const RigidDev& b, const RigidStepConst& rc) {
  const double dt = rc.dt;
  M3d Jxv = ld3(Jm + 0), Jxw = ld3(Jm + 9);
  const M3d Jvv = ld3(Jm + 18), Jvw = ld3(Jm + 27), Jwv = ld3(Jm + 36), Jww = ld3(Jm + 45);
  M43d Jqv = ld43(Jm + 54), Jqw = ld43(Jm + 66);
  Jxv = Jxv + Jvv * dt;
  Jxw = Jxw + Jvw * dt;
  const double p[4] = {0.0, b.w[0], b.w[1], b.w[2]};
  double pq[4];
  quat_mul(p, b.q, pq);
  double nq[4];
  for (int k = 0; k < 4; ++k) nq[k] = b.q[k] + dt * 0.5 * pq[k];
  const double nn = sqrt(nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3] + nq[0] * nq[0]);
  double qh[4];
  for (int k = 0; k < 4; ++k) qh[k] = nq[k] / nn;
  M4d Nq;
  for (int r = 0; r < 4; ++r)
    for (int cc = 0; cc < 4; ++cc) Nq(r, cc) = (((r == cc) ? 1.0 : 0.0) - qh[r] * qh[cc]) / nn;
  M4d Pq;
  const double pv[16] = {p[0], -p[1], -p[2], -p[3], p[1], p[0], -p[3], p[2],
                         p[2], p[3], p[0], -p[1], p[3], -p[2], p[1], p[0]};
  for (int k = 0; k < 16; ++k) Pq.a[k] = pv[k];
  M43d Pw;
  const double* q = b.q;
  const double wv[12] = {-q[1], -q[2], -q[3], q[0], q[3], -q[2], -q[3], q[0], q[1], q[2], -q[1], q[0]};
  for (int k = 0; k < 12; ++k) Pw.a[k] = wv[k];
  Jqv = mm(Nq, Jqv + (mm(Pw, Jwv) + mm(Pq, Jqv)) * (dt / 2.0));
  Jqw = mm(Nq, Jqw + (mm(Pw, Jww) + mm(Pq, Jqw)) * (dt / 2.0));
  st3(Jm + 0, Jxv); st3(Jm + 9, Jxw); st43(Jm + 54, Jqv); st43(Jm + 66, Jqw);
}
__device__ void manager_position_chain(RigidDev& b, const RigidStepConst& rc) { position_chain(b.Jm, b, rc); }

__device__ void rigid_forces(RigidDev& b, const RigidStepConst& rc);
__device__ void rigid_animate(RigidDev& b, const RigidStepConst& rc);

__global__ void k_rigid_step(RigidStepConst rc, RigidDev* __restrict__ rig) {
  const int bi = threadIdx.x;
  if (bi >= rc.nbodies) return;
  RigidDev& b = rig[bi];
  if (!b.dynamic) return;
  const int ph = rc.phases ? rc.phases : (RS_HAND | RS_MGR_VEL | RS_FORCES | RS_MGR_POS | RS_ANIMATE);
  if ((ph & RS_HAND) && rc.compute_gradient && !b.animated) {
    if (rc.use_manager) {
      // update_rigid_body_gradient_manager (:952-967): the (R, R) entry
      double* dst = rc.mgr_net ? rc.mgr_net + ((size_t)bi * rc.nbodies + bi) * 78 : b.mnet;
      for (int k = 0; k < 78; ++k) dst[k] = b.net[k];
    } else {
      chain_rule(b, rc);
    }
  }
  // the manager's chain runs for every dynamic body (SimulatorBase.cpp:1159-1166)
  if ((ph & RS_MGR_VEL) && rc.compute_gradient && rc.use_manager) manager_velocity_chain(b, rc);
  if (ph & RS_FORCES) rigid_forces(b, rc);
  if ((ph & RS_MGR_POS) && rc.compute_gradient && rc.use_manager) manager_position_chain(b, rc);
  if (ph & RS_ANIMATE) rigid_animate(b, rc);
}

// updateBoundaryForces and gravity (or, after the contact solver, its later
// updateBoundaryForces calls).
__device__ void rigid_forces(RigidDev& b, const RigidStepConst& rc) {
  const double dt = rc.dt;
  const V3d F = {b.F[0], b.F[1], b.F[2]}, T = {b.T[0], b.T[1], b.T[2]};
  if (!b.animated) {
    for (int k = 0; k < 3; ++k) b.v[k] += b.F[k] * (b.invMass * dt);
    const M3d I = ld3(b.I), Iinv = ld3(b.Iinv);
    const V3d om = {b.w[0], b.w[1], b.w[2]};
    V3d rhs = T;
    if (rc.gyroscopic) {
      const V3d L = mv(I, om);
      const V3d lw = cross(L, om);
      rhs = {lw.x + T.x, lw.y + T.y, lw.z + T.z};
    }
    const V3d dw = mv(Iinv, rhs);
    b.w[0] += dw.x * dt; b.w[1] += dw.y * dt; b.w[2] += dw.z * dt;
  }
  b.Flast[0] = F.x; b.Flast[1] = F.y; b.Flast[2] = F.z;
  b.Tlast[0] = T.x; b.Tlast[1] = T.y; b.Tlast[2] = T.z;
  for (int k = 0; k < 3; ++k) { b.F[k] = 0.0; b.T[k] = 0.0; }
  if (rc.contact) {
    // RigidContactSolver::solveRigidContact calls updateBoundaryForces once per
    // boundary model (RigidContactSolver.cpp:354-357): the later calls add zero
    // torques, which with the gyroscopic mode still add I^-1 ((I w) x w) dt
    if (rc.gyroscopic && !b.animated) {
      const M3d I = ld3(b.I), Iinv = ld3(b.Iinv);
      for (int call = 1; call < rc.nbodies; ++call) {
        const V3d om = {b.w[0], b.w[1], b.w[2]};
        const V3d dw = mv(Iinv, cross(mv(I, om), om));
        b.w[0] += dw.x * dt; b.w[1] += dw.y * dt; b.w[2] += dw.z * dt;
      }
    }
  } else if (!b.animated) {
    for (int k = 0; k < 3; ++k) b.v[k] += (rc.g[k] * b.mass) * (b.invMass * dt);
  }
}

// Dynamic3dRigidBody::animate and updateInertia.
__device__ void rigid_animate(RigidDev& b, const RigidStepConst& rc) {
  const double dt = rc.dt;
  for (int k = 0; k < 3; ++k) b.x[k] += b.v[k] * dt;
  const double wq[4] = {0.0, b.w[0], b.w[1], b.w[2]};
  double d[4];
  quat_mul(wq, b.q, d);
  for (int k = 0; k < 4; ++k) b.q[k] += dt * 0.5 * d[k];
  const double nn = sqrt(b.q[1] * b.q[1] + b.q[2] * b.q[2] + b.q[3] * b.q[3] + b.q[0] * b.q[0]);
  for (int k = 0; k < 4; ++k) b.q[k] /= nn;
  const M3d R = quat_to_R(b.q);
  const M3d In = mm(mm(R, ld3(b.I0)), R.tr());
  st3(b.I, In);
  st3(b.Iinv, inverse3(In));
  for (int k = 0; k < 78; ++k) { b.netlast[k] = b.net[k]; b.net[k] = 0.0; }
}

// ---------------------------------------------------------------------------
// RigidBodyGradientManager across bodies (RigidBodyGradientManager.cpp:96-423),
// run with the rigid contact solver: one thread per pair (R, Q) of one
// workgroup.  The force/torque chain reads the other pairs' Jacobians, so every
// thread forms its sums before any Jacobian changes.
// perform_force_and_torque_chain_rule (:96-156: v and omega terms) or _Rigid
// (:158-239: the x terms too, and the q terms of the omega0 sums; the q terms
// written after the v0 sums are statements of their own there and do not
// enter them), then perform_velocity_chain_rule (:276-365) with the body's
// current force and torque (cleared once updateBoundaryForces ran).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_manager_multi_velocity(RigidStepConst rc, RigidDev* __restrict__ rig,
                                                                double* __restrict__ J, const double* __restrict__ N,
                                                                int nb, int rigid) {
  const int t = threadIdx.x;
  const int R = t / nb, Q = t % nb;
  const bool on = t < nb * nb && rig[R].dynamic && rig[Q].dynamic;
  M3d dFv0 = M3d::zero(), dFw0 = M3d::zero(), dTv0 = M3d::zero(), dTw0 = M3d::zero();
  if (on) {
    for (int K = 0; K < nb; ++K) {
      if (!rig[K].dynamic) continue;
      const double* n = N + ((size_t)R * nb + K) * 78;
      const double* j = J + ((size_t)K * nb + Q) * 78;
      const M3d Fv = ld3(n + 0), Fx = ld3(n + 9), Fw = ld3(n + 18), Tw = ld3(n + 27), Tx = ld3(n + 36),
                Tv = ld3(n + 45);
      const M34d Fq = ld34(n + 54), Tq = ld34(n + 66);
      const M3d Jxv = ld3(j + 0), Jxw = ld3(j + 9), Jvv = ld3(j + 18), Jvw = ld3(j + 27), Jwv = ld3(j + 36),
                Jww = ld3(j + 45);
      const M43d Jqw = ld43(j + 66);
      if (rigid) {
        dFv0 = dFv0 + (mm(Fv, Jvv) + mm(Fw, Jwv) + mm(Fx, Jxv));
        dTv0 = dTv0 + (mm(Tv, Jvv) + mm(Tw, Jwv) + mm(Tx, Jxv));
        dFw0 = dFw0 + (mm(Fv, Jvw) + mm(Fw, Jww) + mm(Fx, Jxw) + mm(Fq, Jqw));
        dTw0 = dTw0 + (mm(Tv, Jvw) + mm(Tw, Jww) + mm(Tx, Jxw) + mm(Tq, Jqw));
      } else {
        dFv0 = dFv0 + (mm(Fv, Jvv) + mm(Fw, Jwv));
        dTv0 = dTv0 + (mm(Tv, Jvv) + mm(Tw, Jwv));
        dFw0 = dFw0 + (mm(Fv, Jvw) + mm(Fw, Jww));
        dTw0 = dTw0 + (mm(Tv, Jvw) + mm(Tw, Jww));
      }
    }
  }
  __syncthreads();
  if (!on) return;
  const RigidDev& b = rig[R];
  const double dt = rc.dt;
  double* j = J + ((size_t)R * nb + Q) * 78;
  const M3d I = ld3(b.I), Iinv = ld3(b.Iinv), I0 = ld3(b.I0);
  const M3d Rm = quat_to_R(b.q);
  const V3d om = {b.w[0], b.w[1], b.w[2]};
  M3d Jvv = ld3(j + 18), Jvw = ld3(j + 27), Jwv = ld3(j + 36), Jww = ld3(j + 45);
  const M43d Jqv = ld43(j + 54), Jqw = ld43(j + 66);
  const V3d T = {b.T[0], b.T[1], b.T[2]};
  Jvv = Jvv + dFv0 * (dt * b.invMass);
  Jvw = Jvw + dFw0 * (dt * b.invMass);
  V3d temp = T;
  M3d dTauw0 = dTw0, dTauv0 = dTv0;
  if (rc.gyroscopic) {
    const V3d Iom = mv(I, om);
    const V3d lw = cross(Iom, om);
    temp = {lw.x + T.x, lw.y + T.y, lw.z + T.z};
    dTauw0 = mm(skew(Iom), Jww) + mm(skew(om).tr(), mm(I, Jww) + grad_inertia_v(b, Rm, I0, om, Jqw)) + dTw0;
    dTauv0 = mm(skew(Iom), Jwv) + mm(skew(om).tr(), mm(I, Jwv) + grad_inertia_v(b, Rm, I0, om, Jqv)) + dTv0;
  }
  const V3d it = mv(Iinv, temp);
  Jww = Jww + (mm(Iinv, grad_inertia_v(b, Rm, I0, it, Jqw)) + mm(Iinv, dTauw0)) * dt;
  Jwv = Jwv + (mm(Iinv, grad_inertia_v(b, Rm, I0, it, Jqv)) + mm(Iinv, dTauv0)) * dt;
  st3(j + 18, Jvv); st3(j + 27, Jvw); st3(j + 36, Jwv); st3(j + 45, Jww);
}

// perform_position_and_rotation_chain_rule (:367-423) for every dynamic pair.
__global__ void __launch_bounds__(256) k_manager_multi_position(RigidStepConst rc, RigidDev* __restrict__ rig,
                                                                double* __restrict__ J, int nb) {
  const int t = threadIdx.x;
  if (t >= nb * nb) return;
  const int R = t / nb, Q = t % nb;
  if (!rig[R].dynamic || !rig[Q].dynamic) return;
  position_chain(J + ((size_t)R * nb + Q) * 78, rig[R], rc);
}

void launch_manager_multi_velocity(const RigidStepConst& rc, RigidDev* rig, double* J, const double* N, int nb,
                                   int rigid, hipStream_t s) {
  hipLaunchKernelGGL(k_manager_multi_velocity, dim3(1), dim3(256), 0, s, rc, rig, J, N, nb, rigid);
}

void launch_manager_multi_position(const RigidStepConst& rc, RigidDev* rig, double* J, int nb, hipStream_t s) {
  hipLaunchKernelGGL(k_manager_multi_position, dim3(1), dim3(256), 0, s, rc, rig, J, nb);
}

// SimulatorBase::updateBoundaryParticles (SimulatorBase.cpp:1827-1858).
__global__ void k_boundary_transform(const uint32_t* __restrict__ idx, int n, float4* __restrict__ bpos,
                                     float4* __restrict__ bvel, const float4* __restrict__ bx0,
                                     const RigidDev* __restrict__ rig) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t j = idx ? idx[k] : (uint32_t)k;
  const float4 x0 = bx0[j];
  const RigidDev& b = rig[(int)x0.w];
  const V3d xr = quat_rotate(b.q, V3d{(double)x0.x, (double)x0.y, (double)x0.z});
  float4 p = bpos[j];
  p.x = (float)(xr.x + b.x[0]); p.y = (float)(xr.y + b.x[1]); p.z = (float)(xr.z + b.x[2]);
  bpos[j] = p;
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  if (b.dynamic || b.animated) {
    const V3d r = {(double)p.x - b.x[0], (double)p.y - b.x[1], (double)p.z - b.x[2]};
    const V3d w = {b.w[0], b.w[1], b.w[2]};
    const V3d c = cross(w, r);
    v = make_float4((float)(c.x + b.v[0]), (float)(c.y + b.v[1]), (float)(c.z + b.v[2]), 0.f);
  }
  bvel[j] = v;
}

// TimeStepDiffDFSPH::beginStep (:353-430): during the release window the body
// velocity is overwritten by factor * (v0, omega0).
struct BeginStepTable { int ramp[DFR_MAX_BODIES]; int animated[DFR_MAX_BODIES]; double factor[DFR_MAX_BODIES]; };
__global__ void k_begin_step(BeginStepTable t, RigidDev* __restrict__ rig, int nbodies) {
  const int b = threadIdx.x;
  if (b >= nbodies) return;
  if (!rig[b].dynamic) return;
  if (t.ramp[b]) {
    for (int k = 0; k < 3; ++k) { rig[b].v[k] = t.factor[b] * rig[b].v0[k]; rig[b].w[k] = t.factor[b] * rig[b].w0[k]; }
  }
  rig[b].animated = t.animated[b];
}

// ---------------------------------------------------------------------------
void launch_gradient(const StepConst& c, const GradConst& gc, FluidBufs& f, BoundaryBufs& b, NbrBufs& nb,
                     RigidDev* rig, int k_first, int k_count, int body, hipStream_t s, const int* stop) {
  if (k_count <= 0) return;
  const unsigned blocks = (unsigned)((k_count + DFR_BLOCK - 1) / DFR_BLOCK);
  hipLaunchKernelGGL(k_gradient, dim3(blocks), dim3(DFR_BLOCK), 0, s, c, gc, k_first, k_count, body, b.dyn_idx, b.nd,
                     nb.bf, nb.bfcnt, f.pos, f.vel, f.badv, f.alpha, f.s, b.pos, b.vel, b.x0, rig, stop);
}

void launch_rigid_step(const RigidStepConst& rc, RigidDev* rig, hipStream_t s) {
  hipLaunchKernelGGL(k_rigid_step, dim3(1), dim3(64), 0, s, rc, rig);
}

void launch_boundary_transform(const uint32_t* idx, int n, BoundaryBufs& b, const RigidDev* rig, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_boundary_transform, dim3((n + DFR_BLOCK - 1) / DFR_BLOCK), dim3(DFR_BLOCK), 0, s, idx, n,
                     b.pos, b.vel, b.x0, rig);
}

void launch_begin_step(RigidDev* rig, int nbodies, const int* ramp, const double* factors, const int* animated,
                       hipStream_t s) {
  BeginStepTable t;
  for (int k = 0; k < DFR_MAX_BODIES; ++k) {
    t.ramp[k] = (k < nbodies) ? ramp[k] : 0;
    t.animated[k] = (k < nbodies) ? animated[k] : 0;
    t.factor[k] = (k < nbodies) ? factors[k] : 0.0;
  }
  hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(64), 0, s, t, rig, nbodies);
}

}  // namespace dfr
