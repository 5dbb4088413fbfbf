// ORACLE C entry points (test infrastructure: loaded by tests/, smoke() and
// bench.py's cpu_baseline leg through ctypes, never by the product).
//
// Precision variants: 0 = <float,float> (reference precision),
//                     1 = <double,double> (FP64 oracle),
//                     2 = <float,double> (FP32 bulk, FP64 rigid).
#include <cstring>
#include <memory>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include "dfsph_adjoint.hpp"

using namespace orc;

struct OracleBase {
  virtual ~OracleBase() = default;
  virtual int add_body(size_t n, const double* x0, int dyn, const double* xrb, const double* q, double dens) = 0;
  virtual void set_fluid(size_t n, const double* x, const double* v) = 0;
  virtual void set_init_velocity(int body, const double* v0, const double* w0) = 0;
  virtual void init() = 0;
  virtual void step() = 0;
  virtual void reset_bodies() = 0;
  virtual size_t n_fluid() const = 0;
  virtual size_t n_body_particles(int body) const = 0;
  virtual int get_fluid(double* x, double* v) const = 0;
  virtual int get_field(const char* name, double* out) const = 0;
  virtual int get_body_field(int body, const char* name, double* out) const = 0;
  virtual int get_rigid(int body, double* out) const = 0;
  virtual int get_jacobians(int body, double* out) const = 0;
  virtual int get_net_gradients(int body, double* out) const = 0;
  virtual int get_mass_props(int body, double* out) const = 0;
  virtual void get_stats(int* it, int* itv, double* err, double* errv) const = 0;
  virtual int neighbor_counts(int kind, int body, unsigned* out) const = 0;
  virtual int neighbor_list(int kind, int body, size_t i, unsigned* out, int cap) const = 0;
  virtual int set_state(const char* name, const double* in) = 0;
  virtual void stage(int which) = 0;
  virtual void set_record(int on) = 0;
  virtual void set_fixed_iterations(int n, const int* iv, const int* ip) = 0;
  virtual int recorded_iterations(int* iv, int* ip, int cap) const = 0;
  virtual void adjoint(const double* seed, double* out) = 0;
  virtual int tape_field(int step, const char* name, int iter, double* out) const = 0;
  virtual void set_warm_net(const float* params, const float* norm, int bf16) = 0;
  virtual int pair_eval(int body, long j, long i, int mode, const double* in, double* out) = 0;
  virtual int pair_jacobians(int body, long j, long i, int mode, double* out) = 0;
  virtual int velocity_update_probe(int which, double* dv, double* ft) = 0;
  virtual int set_force_model(int body, int start, const double* f0t0, const double* A, const double* B, int flags) = 0;
  virtual int model_step(int n) = 0;
  virtual int get_manager(int body, double* jac, double* net) const = 0;
  virtual int get_manager_pair(int R, int Q, double* jac, double* net) const = 0;
  virtual int contact_probe(int obs, int pert, const double* dx, const double* dq, const double* dv,
                            const double* dw, double* out) const = 0;
};

template <typename RF, typename RR> struct Oracle : OracleBase {
  Sim<RF, RR> sim;
  explicit Oracle(const Params& p) : sim(p) {}
  int add_body(size_t n, const double* x0, int dyn, const double* xrb, const double* q, double dens) override {
    return sim.add_body(n, x0, dyn, xrb, q, dens);
  }
  void set_fluid(size_t n, const double* x, const double* v) override { sim.set_fluid(n, x, v); }
  void set_init_velocity(int body, const double* v0, const double* w0) override {
    auto& b = sim.bodies[body];
    b.v0 = V3<RR>(RR(v0[0]), RR(v0[1]), RR(v0[2]));
    b.w0 = V3<RR>(RR(w0[0]), RR(w0[1]), RR(w0[2]));
  }
  void init() override { sim.init(); }
  void step() override { sim.step(); }
  void reset_bodies() override { sim.reset_bodies(); }
  size_t n_fluid() const override { return sim.x.size(); }
  size_t n_body_particles(int body) const override { return sim.bodies[body].x.size(); }
  int get_fluid(double* x, double* v) const override {
    for (size_t i = 0; i < sim.x.size(); ++i)
      for (int k = 0; k < 3; ++k) {
        if (x) x[3 * i + k] = double(sim.x[i][k]);
        if (v) v[3 * i + k] = double(sim.v[i][k]);
      }
    return 0;
  }
  int get_field(const char* name, double* out) const override {
    const std::string n(name);
    const size_t N = sim.x.size();
    const std::vector<RF>* f = nullptr;
    if (n == "density") f = &sim.rho;
    else if (n == "factor") f = &sim.alpha;
    else if (n == "density_adv") f = &sim.bAdv;
    else if (n == "kappa") f = &sim.kappa;
    else if (n == "kappa_v") f = &sim.kappaV;
    else if (n == "color") f = &sim.color;     // He 2014 (zero when it did not run)
    else if (n == "gradC2") f = &sim.gradC2;
    if (f) { for (size_t i = 0; i < N; ++i) out[i] = i < f->size() ? double((*f)[i]) : 0.0; return 0; }
    if (n == "sum_grad_p_k") {
      for (size_t i = 0; i < N; ++i) for (int k = 0; k < 3; ++k) out[3 * i + k] = double(sim.s[i][k]);
      return 0;
    }
    // accelerations of the last step (gravity + non-pressure forces), Akinci 2013 normals
    const std::vector<V3<RF>>* vf = n == "accel" ? &sim.accel : n == "normal" ? &sim.normal : nullptr;
    if (vf) {
      for (size_t i = 0; i < N; ++i)
        for (int k = 0; k < 3; ++k) out[3 * i + k] = i < vf->size() ? double((*vf)[i][k]) : 0.0;
      return 0;
    }
    if (n == "dt") { out[0] = double(sim.dt); return 0; }  // the step size the next step starts with
    if (n == "time") { out[0] = double(sim.time); return 0; }
    return -1;
  }
  int get_body_field(int body, const char* name, double* out) const override {
    const std::string n(name);
    const auto& b = sim.bodies[body];
    if (n == "volume") { for (size_t j = 0; j < b.V.size(); ++j) out[j] = double(b.V[j]); return 0; }
    if (n == "contact_vol0" || n == "contact_density0") {  // zero when the contact solver is off
      const auto& f = n == "contact_vol0" ? b.cvol0 : b.cdens0;
      for (size_t j = 0; j < b.x.size(); ++j) out[j] = j < f.size() ? double(f[j]) : 0.0;
      return 0;
    }
    if (n == "contact_force" || n == "contact_torque") {  // sums over the last step's contact samples
      const auto& f = n == "contact_force" ? b.cF : b.cT;
      for (int k = 0; k < 3; ++k) out[k] = double(f[k]);
      return 0;
    }
    if (n == "x") { for (size_t j = 0; j < b.x.size(); ++j) for (int k = 0; k < 3; ++k) out[3 * j + k] = double(b.x[j][k]); return 0; }
    if (n == "v") { for (size_t j = 0; j < b.v.size(); ++j) for (int k = 0; k < 3; ++k) out[3 * j + k] = double(b.v[j][k]); return 0; }
    return -1;
  }
  // x(3) q(w,x,y,z)(4) v(3) omega(3) F_last(3) T_last(3) = 19
  int get_rigid(int body, double* out) const override {
    const auto& b = sim.bodies[body];
    for (int k = 0; k < 3; ++k) out[k] = double(b.xrb[k]);
    out[3] = double(b.q.w); out[4] = double(b.q.x); out[5] = double(b.q.y); out[6] = double(b.q.z);
    for (int k = 0; k < 3; ++k) { out[7 + k] = double(b.vrb[k]); out[10 + k] = double(b.wrb[k]); }
    for (int k = 0; k < 3; ++k) { out[13 + k] = double(b.F_last[k]); out[16 + k] = double(b.T_last[k]); }
    return 0;
  }
  // Jxv Jxw Jvv Jvw Jwv Jww (row-major 3x3 each) then Jqv Jqw (4x3 each) = 78
  int get_jacobians(int body, double* out) const override {
    const auto& b = sim.bodies[body];
    const M3<RR>* m3[6] = {&b.Jxv, &b.Jxw, &b.Jvv, &b.Jvw, &b.Jwv, &b.Jww};
    int o = 0;
    for (auto m : m3) for (int i = 0; i < 9; ++i) out[o++] = double(m->a[i]);
    for (int i = 0; i < 12; ++i) out[o++] = double(b.Jqv.a[i]);
    for (int i = 0; i < 12; ++i) out[o++] = double(b.Jqw.a[i]);
    return 0;
  }
  // Fv Fx Fw Tw Tx Tv (3x3) then Fq Tq (3x4) = 78 (last step's per-body sums)
  int get_net_gradients(int body, double* out) const override {
    const auto& b = sim.bodies[body];
    const M3<RR>* m3[6] = {&b.nFv, &b.nFx, &b.nFw, &b.nTw, &b.nTx, &b.nTv};
    int o = 0;
    for (auto m : m3) for (int i = 0; i < 9; ++i) out[o++] = double(m->a[i]);
    for (int i = 0; i < 12; ++i) out[o++] = double(b.nFq.a[i]);
    for (int i = 0; i < 12; ++i) out[o++] = double(b.nTq.a[i]);
    return 0;
  }
  // mass, I0 (9), I (9), Iinv (9) = 28
  int get_mass_props(int body, double* out) const override {
    const auto& b = sim.bodies[body];
    out[0] = double(b.mass);
    for (int i = 0; i < 9; ++i) { out[1 + i] = double(b.I0.a[i]); out[10 + i] = double(b.I.a[i]); out[19 + i] = double(b.Iinv.a[i]); }
    return 0;
  }
  void get_stats(int* it, int* itv, double* err, double* errv) const override {
    *it = sim.stats.iterations; *itv = sim.stats.iterationsV;
    *err = sim.stats.avgDensityErr; *errv = sim.stats.avgDivErr;
  }
  const std::vector<std::vector<uint32_t>>* lists(int kind, int body) const {
    if (kind == 0) return &sim.nFF;
    if (kind == 1) return body < int(sim.nFB.size()) ? &sim.nFB[body] : nullptr;
    if (kind == 2) return body < int(sim.nBF.size()) ? &sim.nBF[body] : nullptr;
    return nullptr;
  }
  int neighbor_counts(int kind, int body, unsigned* out) const override {
    auto L = lists(kind, body);
    if (!L) return -1;
    for (size_t i = 0; i < L->size(); ++i) out[i] = unsigned((*L)[i].size());
    return int(L->size());
  }
  int neighbor_list(int kind, int body, size_t i, unsigned* out, int cap) const override {
    auto L = lists(kind, body);
    if (!L || i >= L->size()) return -1;
    const auto& l = (*L)[i];
    const int n = int(l.size());
    for (int k = 0; k < n && k < cap; ++k) out[k] = l[k];
    return n;
  }
  int set_state(const char* name, const double* in) override {
    const std::string n(name);
    const size_t N = sim.x.size();
    if (n == "kappa") { for (size_t i = 0; i < N; ++i) sim.kappa[i] = RF(in[i]); return 0; }
    if (n == "kappa_v") { for (size_t i = 0; i < N; ++i) sim.kappaV[i] = RF(in[i]); return 0; }
    if (n == "time") { sim.time = RF(in[0]); return 0; }
    if (n == "dt") { sim.dt = RF(in[0]); return 0; }
    if (n == "x" || n == "v") {  // fluid positions or velocities [N][3]
      auto& f = (n == "x") ? sim.x : sim.v;
      for (size_t i = 0; i < N; ++i) f[i] = V3<RF>(RF(in[3 * i]), RF(in[3 * i + 1]), RF(in[3 * i + 2]));
      return 0;
    }
    // "jac<b>" / "mjac<b>": forward-mode Jacobians of body b / of the gradient
    // manager's (b, b) entry, in get_jacobians' layout (78)
    const bool mj = n.rfind("mjac", 0) == 0;
    if (mj || n.rfind("jac", 0) == 0) {
      const int b = std::atoi(n.c_str() + (mj ? 4 : 3));
      if (b < 0 || b >= int(sim.bodies.size())) return -1;
      auto& B = sim.bodies[b];
      M3<RR>* m3[6] = {&B.Jxv, &B.Jxw, &B.Jvv, &B.Jvw, &B.Jwv, &B.Jww};
      M3<RR>* mm3[6] = {&B.mJxv, &B.mJxw, &B.mJvv, &B.mJvw, &B.mJwv, &B.mJww};
      M43<RR>* q2[2] = {mj ? &B.mJqv : &B.Jqv, mj ? &B.mJqw : &B.Jqw};
      int o = 0;
      for (int k = 0; k < 6; ++k) for (int i = 0; i < 9; ++i) (mj ? mm3 : m3)[k]->a[i] = RR(in[o++]);
      for (int k = 0; k < 2; ++k) for (int i = 0; i < 12; ++i) q2[k]->a[i] = RR(in[o++]);
      return 0;
    }
    if (n.rfind("rigid", 0) == 0) {  // "rigid<b>": x(3), q (w,x,y,z)(4), v(3), omega(3) of body b
      const int b = std::atoi(n.c_str() + 5);
      if (b < 0 || b >= int(sim.bodies.size())) return -1;
      auto& B = sim.bodies[b];
      B.xrb = V3<RR>(RR(in[0]), RR(in[1]), RR(in[2]));
      B.q = Quat<RR>(RR(in[3]), RR(in[4]), RR(in[5]), RR(in[6]));
      B.vrb = V3<RR>(RR(in[7]), RR(in[8]), RR(in[9]));
      B.wrb = V3<RR>(RR(in[10]), RR(in[11]), RR(in[12]));
      sim.update_inertia(B);
      sim.update_boundary_particles(true);
      return 0;
    }
    return -1;
  }
  void set_record(int on) override {
    sim.record = on != 0;
    sim.tapes.clear();
  }
  void set_fixed_iterations(int n, const int* iv, const int* ip) override {
    sim.fixIv.assign(iv, iv + n);
    sim.fixIp.assign(ip, ip + n);
    sim.fixPos = 0;
  }
  int recorded_iterations(int* iv, int* ip, int cap) const override {
    const int n = int(sim.tapes.size());
    for (int s = 0; s < n && s < cap; ++s) { iv[s] = sim.tapes[s].Iv; ip[s] = sim.tapes[s].Ip; }
    return n;
  }
  // Fields of a recorded step (for tests): d0, kvW, rsW, kW (warm starts),
  // bD/kD/bP/kP (iteration iter), nn (fluid + boundary neighbour count).
  int tape_field(int step, const char* name, int iter, double* out) const override {
    if (step < 0 || step >= int(sim.tapes.size())) return -1;
    const auto& T = sim.tapes[step];
    const std::string n(name);
    const std::vector<RF>* f = nullptr;
    if (n == "d0") f = &T.d0; else if (n == "kvW") f = &T.kvW; else if (n == "rsW") f = &T.rsW;
    else if (n == "kW") f = &T.kW; else if (n == "alpha") f = &T.alpha; else if (n == "rho") f = &T.rho;
    else if (n == "bD" && iter < int(T.bD.size())) f = &T.bD[iter];
    else if (n == "kD" && iter < int(T.kD.size())) f = &T.kD[iter];
    else if (n == "bP" && iter < int(T.bP.size())) f = &T.bP[iter];
    else if (n == "kP" && iter < int(T.kP.size())) f = &T.kP[iter];
    if (f) { for (size_t i = 0; i < f->size(); ++i) out[i] = double((*f)[i]); return int(f->size()); }
    if (n == "kPrev") f = &T.kPrev; else if (n == "kvPrev") f = &T.kvPrev;
    if (f) { for (size_t i = 0; i < f->size(); ++i) out[i] = double((*f)[i]); return int(f->size()); }
    if (n == "vDiv" || n == "x") {
      const auto& w = (n == "x") ? T.x : T.vDiv;
      for (size_t i = 0; i < w.size(); ++i) for (int k = 0; k < 3; ++k) out[3 * i + k] = double(w[i][k]);
      return int(3 * w.size());
    }
    if (n == "feat") {
      for (size_t i = 0; i < T.feat.size(); ++i) out[i] = T.feat[i];
      return int(T.feat.size());
    }
    if (n == "nn") {
      for (size_t i = 0; i < T.nFF.size(); ++i) {
        size_t c = T.nFF[i].size();
        for (const auto& l : T.nFB) c += l[i].size();
        out[i] = double(c);
      }
      return int(T.nFF.size());
    }
    return -1;
  }
  // params == nullptr: plain warm starts again
  void set_warm_net(const float* params, const float* norm, int bf16) override {
    if (params) sim.net.set(params, norm, bf16);
    else sim.net.enabled = false;
  }
  void adjoint(const double* seed, double* out) override {
    // The adjoint covers the viscosity forces (both methods); surface tension
    // and the CFL step size are forward-only.
    if (sim.P.surfaceTensionMethod != 0 || sim.P.cflMethod != 0)
      throw std::runtime_error("adjoint: surface tension and the CFL step size are forward-only");
    Adjoint<RF, RR> adj(sim);
    adj.run(seed, out);
  }
  // ---- probes for the pinning tests --------------------------------------
  // The (boundary sample j of `body`, fluid particle i) force term of one
  // Jacobi iteration as a function of x_j, v_j, v_i (in: x_j, v_j, v_i and the
  // body origin x_rb, 12 values), with the neighbour lists, the other particles
  // and samples held: recomputes rho_i, alpha_i, s_i and the solver's b_i
  // (mode 0: rho* - 1, mode 1: d), then k_i = b_i alpha_i c with the real
  // c = 1/dt^2 or 1/dt, g = -V_j gradW(x_i - x_j), the velocity-change rate
  // -k_i g, the force on the body f = m k_i g and its torque (x_j - x_rb) x f.
  // out: b, alpha, k, dvc(3), f(3), tau(3), V_j W(x_i - x_j), rho_i = 14.
  // Restores every value it changed.
  int pair_eval(int body, long j, long i, int mode, const double* in, double* out) override {
    auto& S = sim;
    if (body < 0 || body >= int(S.bodies.size())) return -1;
    auto& b = S.bodies[body];
    if (j < 0 || j >= long(b.x.size()) || i < 0 || i >= long(S.x.size())) return -1;
    const V3<RF> xj0 = b.x[j], vj0 = b.v[j], vi0 = S.v[i], si0 = S.s[i];
    const RF rhoi0 = S.rho[i], ai0 = S.alpha[i], bi0 = S.bAdv[i];
    b.x[j] = V3<RF>(RF(in[0]), RF(in[1]), RF(in[2]));
    b.v[j] = V3<RF>(RF(in[3]), RF(in[4]), RF(in[5]));
    S.v[i] = V3<RF>(RF(in[6]), RF(in[7]), RF(in[8]));
    const V3<RR> xrb{RR(in[9]), RR(in[10]), RR(in[11])};
    S.density_of(i);
    S.factor_of(i);
    if (mode == 0) S.compute_density_adv(i); else S.compute_density_change(i);
    const RF bi = (mode == 0) ? S.bAdv[i] - RF(1) : S.bAdv[i];
    const RF c = (mode == 0) ? RF(1) / S.dt / S.dt : RF(1) / S.dt;
    const RF k = bi * S.alpha[i] * c;
    const V3<RF> g = S.K.gradW(S.x[i] - b.x[j]) * (-b.V[j]);
    const V3<RF> dvc = g * (-k);
    const V3<RF> f = dvc * (-S.m_f);
    const V3<RR> tau = (V3<RR>(b.x[j]) - xrb).cross(V3<RR>(f));
    out[0] = double(bi); out[1] = double(S.alpha[i]); out[2] = double(k);
    for (int q = 0; q < 3; ++q) { out[3 + q] = double(dvc[q]); out[6 + q] = double(f[q]); out[9 + q] = double(tau[q]); }
    out[12] = double(b.V[j] * S.K.W(S.x[i] - b.x[j]));
    out[13] = double(S.rho[i]);
    b.x[j] = xj0; b.v[j] = vj0; S.v[i] = vi0; S.s[i] = si0;
    S.rho[i] = rhoi0; S.alpha[i] = ai0; S.bAdv[i] = bi0;
    return 0;
  }
  // The reference-mode pair Jacobians of that term at the current state
  // (gradient_pair): F_v F_x F_w T_w T_x T_v (3x3) then F_q T_q (3x4) = 78.
  int pair_jacobians(int body, long j, long i, int mode, double* out) override {
    auto& S = sim;
    if (body < 0 || body >= int(S.bodies.size())) return -1;
    auto& b = S.bodies[body];
    if (j < 0 || j >= long(b.x.size()) || i < 0 || i >= long(S.x.size())) return -1;
    S.density_of(i);
    S.factor_of(i);
    if (mode == 0) S.compute_density_adv(i); else S.compute_density_change(i);
    auto zero = [&] {
      b.gFv[j].setZero(); b.gFx[j].setZero(); b.gFw[j].setZero(); b.gFq[j].setZero();
      b.gTw[j].setZero(); b.gTv[j].setZero(); b.gTx[j].setZero(); b.gTq[j].setZero();
    };
    zero();
    S.gradient_pair(b, i, uint32_t(j), mode);
    const M3<RR>* m3[6] = {&b.gFv[j], &b.gFx[j], &b.gFw[j], &b.gTw[j], &b.gTx[j], &b.gTv[j]};
    int o = 0;
    for (auto m : m3) for (int q = 0; q < 9; ++q) out[o++] = double(m->a[q]);
    for (int q = 0; q < 12; ++q) out[o++] = double(b.gFq[j].a[q]);
    for (int q = 0; q < 12; ++q) out[o++] = double(b.gTq[j].a[q]);
    zero();
    return 0;
  }
  // One Jacobi velocity update of every fluid particle from the current
  // positions and velocities (neighbour search, densities and factors first):
  // which 0 = pressure form k_i = (rho*_i - 1) alpha_i / dt^2, 1 = divergence
  // form k_i = d_i alpha_i / dt.  dv: velocity change [N][3]; ft: per body the
  // force and torque the update accumulated (6 per body).  Velocities and
  // body accumulators are restored.
  int velocity_update_probe(int which, double* dv, double* ft) override {
    auto& S = sim;
    S.neighborhood_search();
    S.compute_densities();
    S.compute_dfsph_factor();
    const long N = long(S.x.size());
    std::vector<RF> kf(N);
    for (long i = 0; i < N; ++i) {
      if (which == 0) { S.compute_density_adv(i); kf[i] = (S.bAdv[i] - RF(1)) * S.alpha[i] / (S.dt * S.dt); }
      else { S.compute_density_change(i); kf[i] = S.bAdv[i] * S.alpha[i] / S.dt; }
    }
    const auto v0 = S.v;
    for (auto& b : S.bodies) for (int t = 0; t < S.nthreads; ++t) { b.Fth[t] = V3<RR>(); b.Tth[t] = V3<RR>(); }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) S.velocity_update(i, kf[i], kf);
    for (long i = 0; i < N; ++i) for (int k = 0; k < 3; ++k) dv[3 * i + k] = double(S.v[i][k]) - double(v0[i][k]);
    for (size_t bi = 0; bi < S.bodies.size(); ++bi) {
      V3<RR> F, T;
      S.force_and_torque(S.bodies[bi], F, T);
      for (int k = 0; k < 3; ++k) { ft[6 * bi + k] = double(F[k]); ft[6 * bi + 3 + k] = double(T[k]); }
      for (int t = 0; t < S.nthreads; ++t) { S.bodies[bi].Fth[t] = V3<RR>(); S.bodies[bi].Tth[t] = V3<RR>(); }
    }
    S.v = v0;
    return 0;
  }
  // RigidBodyGradientManager entries (body, body): Jacobians in get_jacobians'
  // layout, handed sums in get_net_gradients' layout.
  // Entry (R, RR) of the manager's table across bodies (with the rigid contact
  // solver and the manager); the layout of get_manager.
  int get_manager_pair(int R, int Q, double* jac, double* net) const override {
    const int nb = int(sim.bodies.size());
    if (R < 0 || Q < 0 || R >= nb || Q >= nb || sim.MJ.size() != size_t(nb) * nb) return -1;
    const auto& J = sim.MJ[size_t(R) * nb + Q];
    const auto& N = sim.MN[size_t(R) * nb + Q];
    const M3<RR>* j3[6] = {&J.xv, &J.xw, &J.vv, &J.vw, &J.wv, &J.ww};
    int o = 0;
    for (auto m : j3) for (int i = 0; i < 9; ++i) jac[o++] = double(m->a[i]);
    for (int i = 0; i < 12; ++i) jac[o++] = double(J.qv.a[i]);
    for (int i = 0; i < 12; ++i) jac[o++] = double(J.qw.a[i]);
    const M3<RR>* n3[6] = {&N.Fv, &N.Fx, &N.Fw, &N.Tw, &N.Tx, &N.Tv};
    o = 0;
    for (auto m : n3) for (int i = 0; i < 9; ++i) net[o++] = double(m->a[i]);
    for (int i = 0; i < 12; ++i) net[o++] = double(N.Fq.a[i]);
    for (int i = 0; i < 12; ++i) net[o++] = double(N.Tq.a[i]);
    return 0;
  }
  int contact_probe(int obs, int pert, const double* dx, const double* dq, const double* dv, const double* dw,
                    double* out) const override {
    return sim.contact_probe(obs, pert, dx, dq, dv, dw, out);
  }
  int get_manager(int body, double* jac, double* net) const override {
    if (body < 0 || body >= int(sim.bodies.size())) return -1;
    const auto& b = sim.bodies[body];
    const M3<RR>* j3[6] = {&b.mJxv, &b.mJxw, &b.mJvv, &b.mJvw, &b.mJwv, &b.mJww};
    int o = 0;
    for (auto m : j3) for (int i = 0; i < 9; ++i) jac[o++] = double(m->a[i]);
    for (int i = 0; i < 12; ++i) jac[o++] = double(b.mJqv.a[i]);
    for (int i = 0; i < 12; ++i) jac[o++] = double(b.mJqw.a[i]);
    const M3<RR>* n3[6] = {&b.mFv, &b.mFx, &b.mFw, &b.mTw, &b.mTx, &b.mTv};
    o = 0;
    for (auto m : n3) for (int i = 0; i < 9; ++i) net[o++] = double(m->a[i]);
    for (int i = 0; i < 12; ++i) net[o++] = double(b.mFq.a[i]);
    for (int i = 0; i < 12; ++i) net[o++] = double(b.mTq.a[i]);
    return 0;
  }
  // flags: 1 = record the base run's inverse inertia per step (clears the
  // record), 2 = run the frozen-inverse map (Sim::ForceModel).
  int set_force_model(int body, int start, const double* f0t0, const double* A, const double* B, int flags) override {
    auto& m = sim.fm;
    if (body < 0 || body >= int(sim.bodies.size())) return -1;
    m.body = body; m.start = start;
    for (int k = 0; k < 3; ++k) { m.F0[k] = f0t0[k]; m.T0[k] = f0t0[3 + k]; }
    for (int k = 0; k < 39; ++k) { m.A[k] = A[k]; m.B[k] = B[k]; }
    m.recordFrozen = (flags & 1) != 0;
    m.useFrozen = (flags & 2) != 0;
    if (m.recordFrozen) m.frozen.clear();
    return 0;
  }
  int model_step(int n) override {
    if (sim.fm.body < 0) return -1;
    for (int s = 0; s < n; ++s) sim.model_step();
    return 0;
  }

  // Partial steps for per-kernel parity: 0 = neighbour search + density + factor.
  void stage(int which) override {
    if (which == 0) {
      sim.neighborhood_search();
      sim.compute_densities();
      sim.compute_dfsph_factor();
    }
  }
};

static thread_local std::string g_err;

extern "C" {

void* orc_create(int variant, const Params* p) {
  try {
    switch (variant) {
      case 0: return static_cast<OracleBase*>(new Oracle<float, float>(*p));
      case 1: return static_cast<OracleBase*>(new Oracle<double, double>(*p));
      case 2: return static_cast<OracleBase*>(new Oracle<float, double>(*p));
    }
  } catch (const std::exception& e) { g_err = e.what(); }
  return nullptr;
}
void orc_destroy(void* h) { delete static_cast<OracleBase*>(h); }
int orc_params_size() { return int(sizeof(Params)); }
int orc_set_fluid(void* h, size_t n, const double* x, const double* v) { static_cast<OracleBase*>(h)->set_fluid(n, x, v); return 0; }
int orc_add_body(void* h, size_t n, const double* x0, int dyn, const double* xrb, const double* q, double dens) {
  return static_cast<OracleBase*>(h)->add_body(n, x0, dyn, xrb, q, dens);
}
int orc_set_init_velocity(void* h, int body, const double* v0, const double* w0) { static_cast<OracleBase*>(h)->set_init_velocity(body, v0, w0); return 0; }
int orc_init(void* h) { static_cast<OracleBase*>(h)->init(); return 0; }
int orc_step(void* h, int nsteps) { for (int s = 0; s < nsteps; ++s) static_cast<OracleBase*>(h)->step(); return 0; }
int orc_reset_bodies(void* h) { static_cast<OracleBase*>(h)->reset_bodies(); return 0; }
int orc_stage(void* h, int which) { static_cast<OracleBase*>(h)->stage(which); return 0; }
size_t orc_n_fluid(void* h) { return static_cast<OracleBase*>(h)->n_fluid(); }
size_t orc_n_body_particles(void* h, int body) { return static_cast<OracleBase*>(h)->n_body_particles(body); }
int orc_get_fluid(void* h, double* x, double* v) { return static_cast<OracleBase*>(h)->get_fluid(x, v); }
int orc_get_field(void* h, const char* name, double* out) { return static_cast<OracleBase*>(h)->get_field(name, out); }
int orc_get_body_field(void* h, int body, const char* name, double* out) { return static_cast<OracleBase*>(h)->get_body_field(body, name, out); }
int orc_get_rigid(void* h, int body, double* out) { return static_cast<OracleBase*>(h)->get_rigid(body, out); }
int orc_get_jacobians(void* h, int body, double* out) { return static_cast<OracleBase*>(h)->get_jacobians(body, out); }
int orc_get_net_gradients(void* h, int body, double* out) { return static_cast<OracleBase*>(h)->get_net_gradients(body, out); }
int orc_get_mass_props(void* h, int body, double* out) { return static_cast<OracleBase*>(h)->get_mass_props(body, out); }
int orc_get_stats(void* h, int* it, int* itv, double* err, double* errv) { static_cast<OracleBase*>(h)->get_stats(it, itv, err, errv); return 0; }
int orc_neighbor_counts(void* h, int kind, int body, unsigned* out) { return static_cast<OracleBase*>(h)->neighbor_counts(kind, body, out); }
int orc_neighbor_list(void* h, int kind, int body, size_t i, unsigned* out, int cap) { return static_cast<OracleBase*>(h)->neighbor_list(kind, body, i, out, cap); }
int orc_set_state(void* h, const char* name, const double* in) { return static_cast<OracleBase*>(h)->set_state(name, in); }

// Exact adjoint: record tapes while stepping, optionally with fixed iteration
// counts, then run the reverse pass (dfsph_adjoint.hpp).
int orc_set_record(void* h, int on) { static_cast<OracleBase*>(h)->set_record(on); return 0; }
int orc_set_fixed_iterations(void* h, int n, const int* iv, const int* ip) {
  static_cast<OracleBase*>(h)->set_fixed_iterations(n, iv, ip);
  return 0;
}
int orc_recorded_iterations(void* h, int* iv, int* ip, int cap) {
  return static_cast<OracleBase*>(h)->recorded_iterations(iv, ip, cap);
}
int orc_tape_field(void* h, int step, const char* name, int iter, double* out) {
  return static_cast<OracleBase*>(h)->tape_field(step, name, iter, out);
}
int orc_get_manager_pair(void* h, int R, int Q, double* jac, double* net) {
  return static_cast<OracleBase*>(h)->get_manager_pair(R, Q, jac, net);
}
int orc_contact_probe(void* h, int obs, int pert, const double* dx, const double* dq, const double* dv,
                      const double* dw, double* out) {
  return static_cast<OracleBase*>(h)->contact_probe(obs, pert, dx, dq, dv, dw, out);
}
int orc_adjoint(void* h, const double* seed, double* out) {
  try {
    static_cast<OracleBase*>(h)->adjoint(seed, out);
  } catch (const std::exception& e) { g_err = e.what(); return -1; }
  return 0;
}

// Closed-form probes of the cubic kernel for the kernel tests.
// The Akinci 2013 cohesion and adhesion kernels of the oracle (FP32 or FP64).
void orc_cohesion_adhesion_eval(int variant_double, double radius, const double* r, int n, double* Cw, double* Aw) {
  auto run = [&](auto zero) {
    using T = decltype(zero);
    Cohesion<T> KC;
    Adhesion<T> KA;
    KC.set(T(radius));
    KA.set(T(radius));
    for (int i = 0; i < n; ++i) {
      const V3<T> v{T(r[3 * i]), T(r[3 * i + 1]), T(r[3 * i + 2])};
      Cw[i] = double(KC.W(v));
      Aw[i] = double(KA.W(v));
    }
  };
  if (variant_double) run(double(0)); else run(float(0));
}

void orc_kernel_eval(int variant_double, double radius, const double* r, int n, double* W, double* gW, double* ggW) {
  if (variant_double) {
    Cubic<double> K; K.set(radius);
    for (int i = 0; i < n; ++i) {
      V3<double> v(r[3 * i], r[3 * i + 1], r[3 * i + 2]);
      W[i] = K.W(v);
      auto g = K.gradW(v); for (int k = 0; k < 3; ++k) gW[3 * i + k] = g[k];
      auto H = K.gradGradW(v); for (int k = 0; k < 9; ++k) ggW[9 * i + k] = H.a[k];
    }
  } else {
    Cubic<float> K; K.set(float(radius));
    for (int i = 0; i < n; ++i) {
      V3<float> v(float(r[3 * i]), float(r[3 * i + 1]), float(r[3 * i + 2]));
      W[i] = K.W(v);
      auto g = K.gradW(v); for (int k = 0; k < 3; ++k) gW[3 * i + k] = g[k];
      auto H = K.gradGradW(v); for (int k = 0; k < 9; ++k) ggW[9 * i + k] = H.a[k];
    }
  }
}

// GradientUtils probes (double) for the finite-difference tests.
void orc_grad_utils(const double* q, const double* p, double* Rqp, double* RqTp, double* PQ, double* WQ) {
  Quat<double> qq(q[0], q[1], q[2], q[3]);
  V3<double> pp(p[0], p[1], p[2]);
  auto a = grad_Rqp_to_q(qq, pp); for (int i = 0; i < 12; ++i) Rqp[i] = a.a[i];
  auto b = grad_RqTp_to_q(qq, pp); for (int i = 0; i < 12; ++i) RqTp[i] = b.a[i];
  auto c = grad_p_q_product_to_q(qq); for (int i = 0; i < 16; ++i) PQ[i] = c.a[i];
  auto d = grad_omega_q_product_to_omega(qq); for (int i = 0; i < 12; ++i) WQ[i] = d.a[i];
}

// Learned warm start: network for the simulation (params == nullptr: off), and
// the network alone on n feature rows.
int orc_set_warm_net(void* h, const float* params, const float* norm, int bf16) {
  static_cast<OracleBase*>(h)->set_warm_net(params, norm, bf16);
  return 0;
}
int orc_net_predict(const float* params, const float* norm, int bf16, long n, const double* feat, double* out) {
  WarmNet net;
  net.set(params, norm, bf16);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) net.forward(feat + 6 * i, out + 2 * i);
  return 0;
}
int orc_net_input_grad(const float* params, const float* norm, int bf16, long n, const double* feat,
                       const double* ybar, double* out) {
  WarmNet net;
  net.set(params, norm, bf16);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) net.input_grad(feat + 6 * i, ybar + 2 * i, out + 6 * i);
  return 0;
}

int orc_pair_eval(void* h, int body, long j, long i, int mode, const double* in, double* out) {
  return static_cast<OracleBase*>(h)->pair_eval(body, j, i, mode, in, out);
}
int orc_pair_jacobians(void* h, int body, long j, long i, int mode, double* out) {
  return static_cast<OracleBase*>(h)->pair_jacobians(body, j, i, mode, out);
}
int orc_velocity_update_probe(void* h, int which, double* dv, double* ft) {
  return static_cast<OracleBase*>(h)->velocity_update_probe(which, dv, ft);
}
int orc_set_force_model(void* h, int body, int start, const double* f0t0, const double* A, const double* B, int flags) {
  return static_cast<OracleBase*>(h)->set_force_model(body, start, f0t0, A, B, flags);
}
int orc_model_step(void* h, int n) { return static_cast<OracleBase*>(h)->model_step(n); }
int orc_get_manager(void* h, int body, double* jac, double* net) {
  return static_cast<OracleBase*>(h)->get_manager(body, jac, net);
}
const char* orc_last_error() { return g_err.c_str(); }
}
