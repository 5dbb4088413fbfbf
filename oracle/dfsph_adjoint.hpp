// Exact discrete adjoint of Sim::step (ORACLE: test infrastructure only).
//
// Reverse-mode derivative of the recorded steps with respect to the rigid
// bodies' release velocities (init_v_rb, init_omega_rb, applied by beginStep,
// TimeStepDiffDFSPH.cpp:353-430), seeded with dL/d(x_rb, q) after the last
// step.  Every operation of the forward step is differentiated as executed:
// neighbour lists, clamps (max(., 0), max(., 1), the |k| > eps guards of the
// Jacobi update, the 20-neighbour rule) and iteration counts are held at their
// recorded values, so the result is the derivative of the piecewise-smooth map
// the step computes.  The operations and their order follow
// TimeStepDiffDFSPH::step (:526-652): densities (TimeStep.cpp:147-202), factor
// (:883-962), divergence solve (:770-881, :1698-1922), advection, pressure
// solve (:654-768, :964-1220), position update (:613-634), then the rigid
// update (RigidBody3dBoundarySimulator.cpp:278-344) and the boundary transform
// (SimulatorBase.cpp:1827-1858).
//
// The loops scatter (pair i -> j) in double precision into per-thread
// accumulators (OpenMP), merged after every pass; the GPU reverse gathers
// instead, so the two share no arithmetic order.
#pragma once
#include "dfsph_oracle.hpp"

namespace orc {

template <typename RF, typename RR> struct Adjoint {
  using A = double;
  using V = V3<A>;
  const Sim<RF, RR>& S;
  // adjoint of the fluid state (the oracle never permutes particles)
  std::vector<V> xb, vb;
  std::vector<A> kb, kvb;
  struct BodyAdj {
    V x, v, w, v0, w0;
    A q[4] = {0, 0, 0, 0};  // (w, x, y, z)
  };
  std::vector<BodyAdj> body;
  // per-step scratch
  std::vector<std::vector<V>> sx, sv;  // adjoints of boundary sample positions / velocities
  std::vector<V> Fb, Tb;               // adjoints of the step's force / torque per body
  // per-thread accumulators of a scatter pass (kept zero between passes)
  struct Acc {
    std::vector<V> x, v;
    std::vector<A> k;
    std::vector<std::vector<V>> sx, sv;
  };
  std::vector<Acc> acc;

  explicit Adjoint(const Sim<RF, RR>& s) : S(s) {}

  // Runs f(i, acc) for every particle in parallel, then adds the per-thread
  // accumulators into xb, vb, kbuf (if given), sx, sv and clears them.
  template <class F> void scatter(long N, std::vector<A>* kbuf, F&& f) {
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    if (int(acc.size()) != nt) acc.assign(nt, Acc());
    for (auto& o : acc)
      if (long(o.x.size()) != N || o.sx.size() != sx.size()) {
        o.x.assign(N, V()); o.v.assign(N, V()); o.k.assign(N, A(0));
        o.sx.assign(sx.size(), {}); o.sv.assign(sv.size(), {});
        for (size_t b = 0; b < sx.size(); ++b) { o.sx[b].assign(sx[b].size(), V()); o.sv[b].assign(sv[b].size(), V()); }
      }
#pragma omp parallel
    {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      Acc& o = acc[tid];
#pragma omp for schedule(dynamic, 512)
      for (long i = 0; i < N; ++i) f(i, o);
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i)
      for (auto& o : acc) {
        xb[i] += o.x[i]; vb[i] += o.v[i];
        o.x[i] = V(); o.v[i] = V();
        if (kbuf) (*kbuf)[i] += o.k[i];
        o.k[i] = A(0);
      }
    for (auto& o : acc)
      for (size_t b = 0; b < sx.size(); ++b)
        for (size_t j = 0; j < sx[b].size(); ++j) {
          sx[b][j] += o.sx[b][j]; sv[b][j] += o.sv[b][j];
          o.sx[b][j] = V(); o.sv[b][j] = V();
        }
  }

  static V cv(const V3<RF>& a) { return V(A(a.x), A(a.y), A(a.z)); }
  V gw(const V3<RF>& r) const { return cv(S.K.gradW(r)); }
  // gradGradW(r) w
  V hv(const V3<RF>& r, const V& w) const {
    const M3<RF> H = S.K.gradGradW(r);
    V o;
    for (int a = 0; a < 3; ++a) o[a] = A(H(a, 0)) * w.x + A(H(a, 1)) * w.y + A(H(a, 2)) * w.z;
    return o;
  }

  // d(R(q) p)/dq^T y for R(q) p = p + 2w (u x p) + 2 u x (u x p), u = (x, y, z)
  // (Eigen's _transformVector and toRotationMatrix share this polynomial).
  static void rot_vjp(const Quat<A>& q, const V& p, const V& y, A* qb) {
    const V u(q.x, q.y, q.z);
    qb[0] += A(2) * u.cross(p).dot(y);
    const V g = p.cross(y) * (A(2) * q.w) + y * (A(2) * u.dot(p)) + p * (A(2) * u.dot(y)) - u * (A(4) * p.dot(y));
    qb[1] += g.x; qb[2] += g.y; qb[3] += g.z;
  }
  // d<Rb, R(q)>/dq for toRotationMatrix.
  static void rotmat_vjp(const Quat<A>& q, const M3<A>& Rb, A* qb) {
    const A w = q.w, x = q.x, y = q.y, z = q.z;
    auto R = [&](int a, int b) { return Rb(a, b); };
    qb[0] += A(2) * (-z * R(0, 1) + y * R(0, 2) + z * R(1, 0) - x * R(1, 2) - y * R(2, 0) + x * R(2, 1));
    qb[1] += A(2) * (y * R(0, 1) + z * R(0, 2) + y * R(1, 0) - w * R(1, 2) + z * R(2, 0) + w * R(2, 1)) -
             A(4) * x * (R(1, 1) + R(2, 2));
    qb[2] += A(2) * (x * R(0, 1) + w * R(0, 2) + x * R(1, 0) + z * R(1, 2) - w * R(2, 0) + z * R(2, 1)) -
             A(4) * y * (R(0, 0) + R(2, 2));
    qb[3] += A(2) * (-w * R(0, 1) + x * R(0, 2) + w * R(1, 0) + y * R(1, 2) + x * R(2, 0) + y * R(2, 1)) -
             A(4) * z * (R(0, 0) + R(1, 1));
  }

  // ---- reverse of the Jacobi velocity update with stiffness field k
  // (Sim::velocity_update): vb holds the adjoint of the updated velocity and
  // is unchanged (the update adds to v); adds into kbuf, xb, sx.
  void reverse_update(const StepTape<RF, RR>& T, const std::vector<RF>& k, std::vector<A>& kbuf) {
    const RF eps = RF(1e-5);
    const A Vf = A(S.V_f), mf = A(S.m_f), dt = A(S.dt);
    const long N = long(T.x.size());
    scatter(N, &kbuf, [&](long i, Acc& o) {
      auto& xb = o.x;
      auto& kbuf = o.k;
      auto& sx = o.sx;
      const V vbi = vb[i];
      for (uint32_t j : T.nFF[i]) {
        const RF kSum = k[i] + k[j];
        if (!(std::fabs(kSum) > eps)) continue;
        const V3<RF> r = T.x[i] - T.x[j];
        const V g = gw(r);
        const A s = dt * Vf * vbi.dot(g);  // v_i += dt kSum V_f gradW
        kbuf[i] += s;
        kbuf[j] += s;
        const V t = hv(r, vbi) * (dt * A(kSum) * Vf);
        xb[i] += t;
        xb[j] -= t;
      }
      if (!(std::fabs(k[i]) > eps)) return;
      for (size_t bi = 0; bi < S.bodies.size(); ++bi) {
        const auto& B = S.bodies[bi];
        const auto& bt = T.bodies[bi];
        for (uint32_t jb : T.nFB[bi][i]) {
          const V3<RF> r = T.x[i] - bt.xb[jb];
          const V g = gw(r);
          const A Vb = A(B.V[jb]);
          kbuf[i] += dt * Vb * vbi.dot(g);  // v_i += dt k_i V_b gradW
          const V t = hv(r, vbi) * (dt * A(k[i]) * Vb);
          xb[i] += t;
          if (!B.dynamic) continue;
          sx[bi][jb] -= t;
          // f = -m_f k_i V_b gradW;  F += f;  T += (x_b - x_rb) x f
          const V rl = V(V3<RR>(bt.xb[jb]) - bt.xrb);
          const V fbar = Fb[bi] + Tb[bi].cross(rl);
          kbuf[i] -= mf * Vb * g.dot(fbar);
          const V tf = hv(r, fbar) * (mf * A(k[i]) * Vb);
          xb[i] -= tf;
          sx[bi][jb] += tf;
          const V f = g * (-mf * A(k[i]) * Vb);
          sx[bi][jb] += f.cross(Tb[bi]);
        }
      }
    });
  }

  // ---- reverse of delta_i = sum_j V_f (v_i - v_j).gradW_ij + sum_b V_b (v_i - v_b).gradW_ib
  // (computeDensityAdv / computeDensityChange) evaluated at velocity v.
  void reverse_delta(const StepTape<RF, RR>& T, const std::vector<V3<RF>>& v, const std::vector<A>& db) {
    const A Vf = A(S.V_f);
    const long N = long(T.x.size());
    scatter(N, nullptr, [&](long i, Acc& o) {
      auto& xb = o.x;
      auto& vb = o.v;
      auto& sx = o.sx;
      auto& sv = o.sv;
      const A d = db[i];
      if (d == A(0)) return;
      for (uint32_t j : T.nFF[i]) {
        const V3<RF> r = T.x[i] - T.x[j];
        const V g = gw(r) * (d * Vf);
        vb[i] += g;
        vb[j] -= g;
        const V t = hv(r, cv(v[i]) - cv(v[j])) * (d * Vf);
        xb[i] += t;
        xb[j] -= t;
      }
      for (size_t bi = 0; bi < S.bodies.size(); ++bi) {
        const auto& B = S.bodies[bi];
        const auto& bt = T.bodies[bi];
        for (uint32_t jb : T.nFB[bi][i]) {
          const V3<RF> r = T.x[i] - bt.xb[jb];
          const A Vb = A(B.V[jb]);
          const V g = gw(r) * (d * Vb);
          vb[i] += g;
          const V t = hv(r, cv(v[i]) - cv(bt.vb[jb])) * (d * Vb);
          xb[i] += t;
          if (B.dynamic) { sv[bi][jb] -= g; sx[bi][jb] -= t; }
        }
      }
    });
  }

  // ---- reverse of alpha_i = -1 / S_i, S_i = sum_j |g_ij|^2 + |G_i|^2 (computeDFSPHFactor)
  void reverse_factor(const StepTape<RF, RR>& T, const std::vector<A>& ab) {
    const A Vf = A(S.V_f);
    const long N = long(T.x.size());
    scatter(N, nullptr, [&](long i, Acc& o) {
      auto& xb = o.x;
      auto& sx = o.sx;
      const A al = A(T.alpha[i]);
      if (al == A(0) || ab[i] == A(0)) return;
      const A Sb = ab[i] * al * al;  // d(-1/S)/dS = 1/S^2 = alpha^2
      const V G = cv(T.G[i]);
      for (uint32_t j : T.nFF[i]) {
        const V3<RF> r = T.x[i] - T.x[j];
        const V g = gw(r) * (-Vf);
        const V t = hv(r, g + G) * (A(2) * Sb * Vf);
        xb[i] -= t;
        xb[j] += t;
      }
      for (size_t bi = 0; bi < S.bodies.size(); ++bi) {
        const auto& B = S.bodies[bi];
        const auto& bt = T.bodies[bi];
        for (uint32_t jb : T.nFB[bi][i]) {
          const V3<RF> r = T.x[i] - bt.xb[jb];
          const V t = hv(r, G) * (A(2) * Sb * A(B.V[jb]));
          xb[i] -= t;
          if (B.dynamic) sx[bi][jb] += t;
        }
      }
    });
  }

  // ---- reverse of rho_i / rho0 = V_f W0 + sum_j V_f W_ij + sum_b V_b W_ib, and
  // of its boundary part sum_b V_b W_ib (adjoint rbb, learned warm start only)
  void reverse_density(const StepTape<RF, RR>& T, const std::vector<A>& rb, const std::vector<A>* rbb = nullptr) {
    const A Vf = A(S.V_f);
    const long N = long(T.x.size());
    scatter(N, nullptr, [&](long i, Acc& o) {
      auto& xb = o.x;
      auto& sx = o.sx;
      const A rbi = rb[i] + (rbb ? (*rbb)[i] : A(0));
      if (rb[i] == A(0) && rbi == A(0)) return;
      for (uint32_t j : T.nFF[i]) {
        const V t = gw(T.x[i] - T.x[j]) * (rb[i] * Vf);
        xb[i] += t;
        xb[j] -= t;
      }
      for (size_t bi = 0; bi < S.bodies.size(); ++bi) {
        const auto& B = S.bodies[bi];
        const auto& bt = T.bodies[bi];
        for (uint32_t jb : T.nFB[bi][i]) {
          const V t = gw(T.x[i] - bt.xb[jb]) * (rbi * A(B.V[jb]));
          xb[i] += t;
          if (B.dynamic) sx[bi][jb] -= t;
        }
      }
    });
  }

  // ---- reverse of v' = v + dt (g + a_visc(v, x, rho)) (the advection with the
  // viscosity of the step: Viscosity_Standard.cpp:262-268 or Viscosity_XSPH.cpp:62-68)
  // at the velocities after the divergence solve.  vb holds the adjoint of v' on
  // entry and of v on exit; rhob receives the adjoint of rho / rho0.  Standard,
  // per pair: T_ij = c_j (u.r) / D gradW(r), c_j = 10 nu m / rho_j, D = |r|^2 + 0.01 h^2;
  // XSPH: T_ij = -(nu / dt) (m / rho_j) u W(r); u = v_i - v_j, r = x_i - x_j.
  void reverse_viscosity(const StepTape<RF, RR>& T, std::vector<A>& rhob) {
    const long N = long(T.x.size());
    const A dt = A(S.dt), nu = A(RF(S.P.viscosity)), mf = A(S.m_f), rho0 = A(S.rho0);
    const A eps = 0.01 * A(S.h * S.h);
    std::vector<V> abar(N);
    for (long i = 0; i < N; ++i) abar[i] = vb[i] * dt;
    const int method = S.P.viscosityMethod;
    scatter(N, &rhob, [&](long i, Acc& o) {
      const V ai = abar[i];
      const V vi = cv(T.vDiv[i]);
      for (uint32_t j : T.nFF[i]) {
        const V3<RF> rr = T.x[i] - T.x[j];
        const V r = cv(rr);
        const V u = vi - cv(T.vDiv[j]);
        const A rj = A(T.rho[j]);
        if (method == 1) {
          const A c = A(10) * nu * mf / rj;
          const A D = A(rr.squaredNorm()) + eps;
          const A q = u.dot(r), aG = ai.dot(gw(rr));
          const V gu = r * (c * aG / D);
          o.v[i] += gu;
          o.v[j] -= gu;
          const V gr = u * (c * aG / D) + hv(rr, ai) * (c * q / D) - r * (A(2) * c * q * aG / (D * D));
          o.x[i] += gr;
          o.x[j] -= gr;
          o.k[j] -= (c * q * aG / D) / (rj / rho0);
        } else {
          const A c = nu / dt * mf / rj;
          const A W = A(S.K.W(rr)), au = ai.dot(u);
          const V gu = ai * (-c * W);
          o.v[i] += gu;
          o.v[j] -= gu;
          const V gr = gw(rr) * (-c * au);
          o.x[i] += gr;
          o.x[j] -= gr;
          o.k[j] += (c * au * W) / (rj / rho0);
        }
      }
    });
  }

  // ---- fluid part of one step, given Fb/Tb
  void fluid_reverse(const StepTape<RF, RR>& T) {
    const long N = long(T.x.size());
    const A dt = A(S.dt);
    const A invH = A(RF(1.0) / S.dt), invH2 = A(RF(1.0) / (S.dt * S.dt)), h2 = A(S.dt * S.dt);
    std::vector<A> rhob(N, 0), alb(N, 0), kbuf(N), db(N), kacc(N, 0);
    const bool learned = !T.feat.empty();
    std::vector<A> yb(learned ? 2 * N : 0, A(0));  // adjoint of the network's predictions
    for (long i = 0; i < N; ++i) vb[i] += xb[i] * dt;  // x' = x + dt v
    // pressure solve
    const bool warmP = S.P.usePressureWarmstart != 0;
    if (warmP) for (long i = 0; i < N; ++i) kacc[i] = kb[i] * h2;  // kappa' = (kappa_w + sum k) dt^2
    for (int m = T.Ip - 1; m >= 0; --m) {
      kbuf = kacc;
      reverse_update(T, T.kP[m], kbuf);
      for (long i = 0; i < N; ++i) {  // k = (rho* - 1) alpha / dt^2, rho* = max(rho/rho0 + dt delta, 1)
        const A b = A(T.bP[m][i]);
        alb[i] += kbuf[i] * (b - A(1)) * invH2;
        const A bb = kbuf[i] * A(T.alpha[i]) * invH2;
        if (T.bP[m][i] > RF(1)) { rhob[i] += bb; db[i] = dt * bb; }
        else db[i] = 0;
      }
      reverse_delta(T, T.vP[m], db);
    }
    if (warmP) {
      kbuf = kacc;
      reverse_update(T, T.kW, kbuf);
      for (long i = 0; i < N; ++i) {
        if (learned) yb[2 * i + 1] = (T.rsW[i] > RF(1)) ? kbuf[i] * invH2 : A(0);  // kappa_w = y1 / dt^2
        else kb[i] = (T.rsW[i] > RF(1) && T.kPrev[i] > RF(-0.00025)) ? kbuf[i] * A(0.5) * invH2 : A(0);
      }
    }
    // advection with the step's viscosity
    if (S.P.viscosityMethod != 0) reverse_viscosity(T, rhob);
    // divergence solve
    if (S.P.enableDivergenceSolver) {
      const bool warmV = S.P.useDivergenceWarmstart != 0;
      for (long i = 0; i < N; ++i) kacc[i] = warmV ? kvb[i] * dt : A(0);
      for (int m = T.Iv - 1; m >= 0; --m) {
        kbuf = kacc;
        reverse_update(T, T.kD[m], kbuf);
        for (long i = 0; i < N; ++i) {  // k = d alpha / dt, d = max(delta, 0) (0 below 20 neighbours)
          alb[i] += kbuf[i] * A(T.bD[m][i]) * invH;
          db[i] = (T.bD[m][i] > RF(0)) ? kbuf[i] * A(T.alpha[i]) * invH : A(0);
        }
        reverse_delta(T, T.vD[m], db);
      }
      if (warmV) {
        kbuf = kacc;
        reverse_update(T, T.kvW, kbuf);
        for (long i = 0; i < N; ++i) {
          if (learned) yb[2 * i] = (T.d0[i] > RF(0)) ? kbuf[i] * invH : A(0);  // kappa_v,w = y0 / dt
          else kvb[i] = (T.d0[i] > RF(0) && T.kvPrev[i] > RF(-0.5)) ? kbuf[i] * A(0.5) * invH : A(0);
        }
      }
    }
    std::vector<A> rbb;
    if (learned) {
      // through the network to its features: rho/rho0 - 1, dt delta(v0),
      // neighbour count (constant), boundary density, kappa_prev, kappa_v_prev
      const bool divWarm = S.P.enableDivergenceSolver && S.P.useDivergenceWarmstart;
      rbb.assign(N, A(0));
#pragma omp parallel for schedule(static)
      for (long i = 0; i < N; ++i) {
        double fb[6];
        S.net.input_grad(&T.feat[6 * i], &yb[2 * i], fb);
        kb[i] = fb[4] + (warmP ? A(0) : kb[i]);
        kvb[i] = fb[5] + (divWarm ? A(0) : kvb[i]);
        rhob[i] += fb[0];
        db[i] = dt * fb[1];
        rbb[i] = fb[3];
      }
      if (S.P.enableDivergenceSolver) reverse_delta(T, T.v0, db);
    }
    reverse_factor(T, alb);
    reverse_density(T, rhob, learned ? &rbb : nullptr);
  }

  // ---- rigid update of body k (velocityTimeStep + positionTimeStep), in place:
  // body[k] holds the adjoint after the step on entry and the adjoint of the
  // state after the release-window override on exit; sets Fb[k], Tb[k].
  void rigid_reverse(const StepTape<RF, RR>& T, int k) {
    const auto& bt = T.bodies[k];
    const auto& B = S.bodies[k];
    BodyAdj& a = body[k];
    const A dt = A(RR(S.dt));
    const Quat<A> q(A(bt.q.w), A(bt.q.x), A(bt.q.y), A(bt.q.z));
    M3<A> I, Iinv, I0;
    for (int c = 0; c < 9; ++c) { I.a[c] = A(bt.I.a[c]); Iinv.a[c] = A(bt.Iinv.a[c]); I0.a[c] = A(B.I0.a[c]); }
    const V w(A(bt.w.x), A(bt.w.y), A(bt.w.z));
    const V Tq(A(bt.T.x), A(bt.T.y), A(bt.T.z));
    // forward recomputation
    V L, rhs, w1 = w;
    if (!bt.animated) {
      L = I * w;
      rhs = S.P.gyroscopic ? L.cross(w) + Tq : Tq;
      w1 = w + (Iinv * rhs) * dt;
    }
    const Quat<A> wq(0, w1.x, w1.y, w1.z);
    const Quat<A> d = wq * q;
    const A hh = dt * A(0.5);
    const Quat<A> qt(q.w + hh * d.w, q.x + hh * d.x, q.y + hh * d.y, q.z + hh * d.z);
    const A n = qt.norm();
    const A q1[4] = {qt.w / n, qt.x / n, qt.y / n, qt.z / n};
    // q1 = qt / |qt|
    const A dq = q1[0] * a.q[0] + q1[1] * a.q[1] + q1[2] * a.q[2] + q1[3] * a.q[3];
    A qtb[4];
    for (int c = 0; c < 4; ++c) qtb[c] = (a.q[c] - q1[c] * dq) / n;
    // qt = q + dt/2 (0, w1) (x) q
    A qb[4] = {qtb[0], qtb[1], qtb[2], qtb[3]};
    const A dbw = hh * qtb[0];
    const V dbv(hh * qtb[1], hh * qtb[2], hh * qtb[3]);
    const V qv(q.x, q.y, q.z);
    V w1b = a.w - qv * dbw + dbv * q.w + qv.cross(dbv);
    qb[0] += w1.dot(dbv);
    const V qvb = w1 * (-dbw) + dbv.cross(w1);
    qb[1] += qvb.x; qb[2] += qvb.y; qb[3] += qvb.z;
    // x1 = x + dt v2
    const V xbp = a.x;
    const V v2b = a.v + a.x * dt;
    V vbp, wbp;
    M3<A> Ib, Iinvb;
    if (!bt.animated) {
      vbp = v2b;
      Fb[k] = v2b * (A(B.invMass) * dt);
      wbp = w1b;
      const V rhsb = (Iinv.transpose() * w1b) * dt;
      Iinvb = outer(w1b, rhs) * dt;
      Tb[k] = rhsb;
      if (S.P.gyroscopic) {
        const V Lb = w.cross(rhsb);
        wbp += rhsb.cross(L);
        Ib = outer(Lb, w);
        wbp += I.transpose() * Lb;
      }
    } else {
      vbp = v2b;
      wbp = w1b;
      Fb[k] = V();
      Tb[k] = V();
    }
    // I = R I0 R^T, Iinv = I^-1, both from q
    const M3<A> Ibt = Ib - Iinv.transpose() * Iinvb * Iinv.transpose();
    const M3<A> R = q.toRotationMatrix();
    const M3<A> Rb = (Ibt + Ibt.transpose()) * R * I0;
    rotmat_vjp(q, Rb, qb);
    a.x = xbp;
    a.v = vbp;
    a.w = wbp;
    for (int c = 0; c < 4; ++c) a.q[c] = qb[c];
  }

  // ---- one full step in reverse
  void step_reverse(const StepTape<RF, RR>& T) {
    const size_t nb = S.bodies.size();
    Fb.assign(nb, V());
    Tb.assign(nb, V());
    for (size_t k = 0; k < nb; ++k) {
      if (!S.bodies[k].dynamic) continue;
      rigid_reverse(T, int(k));
      // torque lever arm: T = sum (x_b - x_rb) x f  ->  x_rb gets -F x Tbar
      const V F(A(T.bodies[k].F.x), A(T.bodies[k].F.y), A(T.bodies[k].F.z));
      body[k].x -= F.cross(Tb[k]);
      // release window: v, omega overwritten by factor (v0, omega0)
      if (T.bodies[k].ramp) {
        const A f = A(T.bodies[k].factor);
        body[k].v0 += body[k].v * f;
        body[k].w0 += body[k].w * f;
        body[k].v = V();
        body[k].w = V();
      }
    }
    sx.assign(nb, {});
    sv.assign(nb, {});
    for (size_t k = 0; k < nb; ++k) { sx[k].assign(T.bodies[k].xb.size(), V()); sv[k].assign(T.bodies[k].xb.size(), V()); }
    fluid_reverse(T);
    // boundary transform at the start of the step: x_b = R(q) x0 + x_rb,
    // v_b = omega_pre x (R(q) x0) + v_pre
    for (size_t k = 0; k < nb; ++k) {
      const auto& B = S.bodies[k];
      if (!B.dynamic) continue;
      const auto& bt = T.bodies[k];
      const Quat<A> q(A(bt.q.w), A(bt.q.x), A(bt.q.y), A(bt.q.z));
      const V wpre(A(bt.wpre.x), A(bt.wpre.y), A(bt.wpre.z));
      for (size_t j = 0; j < bt.xb.size(); ++j) {
        const V y = sx[k][j] + sv[k][j].cross(wpre);
        body[k].x += sx[k][j];
        rot_vjp(q, V(V3<RR>(B.x0[j])), y, body[k].q);
        body[k].v += sv[k][j];
        const V rl = V(V3<RR>(bt.xb[j]) - bt.xrb);
        body[k].w += rl.cross(sv[k][j]);
      }
    }
  }

  // seed: per body (x_rb[3], q[4] as (w,x,y,z)) adjoint after the last tape;
  // out: per body dL/dv0[3], dL/domega0[3].
  void run(const double* seed, double* out) {
    const long N = long(S.x.size());
    const size_t nb = S.bodies.size();
    xb.assign(N, V()); vb.assign(N, V()); kb.assign(N, 0); kvb.assign(N, 0);
    body.assign(nb, BodyAdj());
    for (size_t k = 0; k < nb; ++k) {
      body[k].x = V(seed[7 * k], seed[7 * k + 1], seed[7 * k + 2]);
      for (int c = 0; c < 4; ++c) body[k].q[c] = seed[7 * k + 3 + c];
    }
    for (long s = long(S.tapes.size()) - 1; s >= 0; --s) step_reverse(S.tapes[s]);
    for (size_t k = 0; k < nb; ++k) {
      for (int c = 0; c < 3; ++c) { out[6 * k + c] = body[k].v0[c]; out[6 * k + 3 + c] = body[k].w0[c]; }
    }
  }
};

}  // namespace orc
