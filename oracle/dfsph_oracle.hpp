// ORACLE -- CPU restatement of DiffFR's differentiable DFSPH fluid-rigid step.
//
// This is test infrastructure: only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg use it, and only as the checker (or as the timed
// CPU baseline).  The product (difffr_amd/, libdfr_hip.so) never links it.
//
// It follows the reference loop by loop.  Every function cites the reference
// file:line it restates (paths relative to the DiffFR repository root).
// Templated on RF (bulk particle arithmetic) and RR (rigid state, force/torque
// partials, gradient accumulators, chain rule):
//   <float, float>   = the reference's own precision (Real = float,
//                      SPlisHSPlasH/Common.h:16);
//   <double, double> = the FP64 oracle run;
//   <float, double>  = the GPU design point (FP32 bulk, FP64 rigid).
//
// Neighbour predicate (CompactNSearch's own tie rule is not available offline,
// so this is a documented choice): j is a neighbour of i iff the squared
// distance, evaluated in RF as dx*dx + dy*dy + dz*dz without fused multiply-add,
// is strictly below h*h (also evaluated in RF).  Self pairs are excluded.
// W, gradW and gradGradW vanish at q = 1, so the rule only decides neighbour
// counts (the 20-neighbour rule of computeDensityChange).
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle_math.hpp"
#include "warmnet.hpp"

namespace orc {

struct Params {
  double radius = 0.025;         // particle radius r; support radius h = 4r
  double dt = 1e-3;              // fixed time step (cflMethod 0)
  double rho0 = 1000.0;          // FluidModel.cpp:45
  double gravity[3] = {0.0, -9.81, 0.0};
  double maxError = 0.01;        // percent, TimeStep.cpp:24-64
  double maxErrorV = 0.1;        // percent, TimeStepDiffDFSPH.cpp:110
  int minIterations = 2;
  int maxIterations = 100;
  int maxIterationsV = 100;
  int usePressureWarmstart = 1;  // TimeStepDiffDFSPH.cpp:102
  int useDivergenceWarmstart = 1;// TimeStepDiffDFSPH.cpp:103
  int enableDivergenceSolver = 1;// TimeStepDiffDFSPH.cpp:97
  int gradientMode = 1;          // 0 Complete, 1 Incomplete, 2 RigidGradOnly (SceneLoader.cpp:86)
  int optimizeRotation = 1;      // TimeStepDiffDFSPH.cpp:108
  int gyroscopic = 1;            // RigidBodyMode::WithGyroscopic, Simulation.cpp:114
  int computeGradient = 1;       // the BACKWARD build of the reference
  double uniformAccTime = 0.0;   // uniform_acc_rb_time
  double targetTime = 1e30;
  int releaseMode = 0;           // useReleaseRigidBodyMode, Simulation.h:444
  int useGradientManager = 0;    // useRigidGradientManager, Simulation.cpp:123
  // Non-pressure forces (Simulation::computeNonPressureForces, Simulation.cpp:618-631;
  // method indices in registration order, NonPressureForceRegistration.cpp:33-60)
  int surfaceTensionMethod = 0;  // 0 none, 2 Akinci et al. 2013, 3 He et al. 2014 (NonPressureForceRegistration.cpp:42-47)
  double surfaceTension = 0.05;  // SurfaceTensionBase.cpp:12
  double surfaceTensionBoundary = 0.01;
  int viscosityMethod = 0;       // 0 none, 1 Standard (FluidModel.cpp:100 makes it a fluid's default), 2 XSPH
  double viscosity = 0.01;       // ViscosityBase.cpp:12
  double viscosityBoundary = 0.0;  // Viscosity_XSPH.cpp:17
  // Time step size (Simulation::updateTimeStepSize, Simulation.cpp:524-616): 0 fixed, 1 CFL
  int cflMethod = 0;
  double cflFactor = 0.5;        // Simulation.cpp:56-60
  double cflMinTimeStepSize = 0.0001;
  double cflMaxTimeStepSize = 0.005;
  // Rigid-rigid contact (RigidContactSolver, penalty form; SceneLoader.cpp:58-71 defaults)
  int useRigidContact = 0;              // useRigidContactSolver
  double rigidContactBeta = 1.0;        // rigidContactBeta (penalty stiffness)
  double rigidContactGamma = 0.7;       // rigidContactGamma (rest volume factor)
  double rigidContactSupportRadiusFactor = 4.0;  // contact kernel radius / particle radius
  double rigidContactFriction = 0.0;    // rigidContactFrictionCoeff
};

// Akinci et al. 2013 cohesion kernel (SPHKernels.h:437-505) and adhesion
// kernel (SPHKernels.h:508-560), both with the support radius as radius.
template <typename T> struct Cohesion {
  T radius, k, c;
  void set(T r) {
    radius = r;
    k = T(32.0) / (T(M_PI) * std::pow(radius, T(9.0)));
    c = std::pow(radius, T(6.0)) / T(64.0);
  }
  T W(const V3<T>& r) const {
    T res = 0;
    const T r2 = r.squaredNorm();
    if (r2 <= radius * radius) {
      const T r1 = std::sqrt(r2);
      const T r3 = r2 * r1;
      if (r1 > T(0.5) * radius) res = k * std::pow(radius - r1, T(3.0)) * r3;
      else res = k * T(2.0) * std::pow(radius - r1, T(3.0)) * r3 - c;
    }
    return res;
  }
};
template <typename T> struct Adhesion {
  T radius, k;
  void set(T r) {
    radius = r;
    k = T(0.007) / std::pow(radius, T(3.25));
  }
  T W(const V3<T>& r) const {
    T res = 0;
    const T r2 = r.squaredNorm();
    if (r2 <= radius * radius) {
      const T rl = std::sqrt(r2);
      if (rl > T(0.5) * radius)
        res = k * std::pow(-T(4.0) * r2 / radius + T(6.0) * rl - T(2.0) * radius, T(0.25));
    }
    return res;
  }
};

// Eval of the analytic cubic spline kernel: SPlisHSPlasH/SPHKernels.h:16-176.
template <typename T> struct Cubic {
  T h, k, l, w0;
  void set(T radius) {
    h = radius;
    const T pi = T(M_PI);
    const T h3 = h * h * h;
    k = T(8.0) / (pi * h3);
    l = T(48.0) / (pi * h3);
    w0 = W(T(0));
  }
  T W(T r) const {  // SPHKernels.h:37-54
    T res = 0;
    const T q = r / h;
    if (q <= T(1)) {
      if (q <= T(0.5)) {
        const T q2 = q * q, q3 = q2 * q;
        res = k * (T(6) * q3 - T(6) * q2 + T(1));
      } else {
        res = k * (T(2) * std::pow(T(1) - q, T(3)));
      }
    }
    return res;
  }
  T W(const V3<T>& r) const { return W(r.norm()); }
  V3<T> gradW(const V3<T>& r) const {  // SPHKernels.h:81-101
    const T rl = r.norm();
    const T q = rl / h;
    if ((rl > T(1.0e-5)) && (q <= T(1))) {
      const V3<T> gradq = r * (T(1) / (rl * h));
      if (q <= T(0.5)) return gradq * (l * q * (T(3) * q - T(2)));
      const T f = T(1) - q;
      return gradq * (l * (-f * f));
    }
    return V3<T>::Zero();
  }
  M3<T> gradGradW(const V3<T>& r) const {  // SPHKernels.h:122-149
    const T rl = r.norm();
    const T q = rl / h;
    if ((rl > T(1.0e-5)) && (q <= T(1))) {
      const V3<T> gradq = r * (T(1) / (rl * h));
      const M3<T> ggq = (M3<T>::Identity() - outer(r, r) / (rl * rl)) * (T(1) / (rl * h));
      // (s * gradq) gradq^T, the evaluation order of the Eigen expression
      if (q <= T(0.5))
        return ggq * (l * q * (T(3) * q - T(2))) + outer(gradq * (l * (T(6) * q - T(2))), gradq);
      const T f = T(1) - q;
      return ggq * (l * (-f * f)) + outer(gradq * (l * T(2) * f), gradq);
    }
    return M3<T>::Zero();
  }
};

// Squared distance without contraction, shared with the GPU predicate.
template <typename T> inline T dist2_nofma(const V3<T>& a, const V3<T>& b) {
  volatile T dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  volatile T xx = dx * dx, yy = dy * dy, zz = dz * dz;
  volatile T s = xx + yy;
  return s + zz;
}

// Uniform grid used only to enumerate candidates; the predicate above decides.
template <typename RF> struct Grid {
  RF cs;
  V3<RF> lo;
  int nx = 1, ny = 1, nz = 1;
  std::vector<uint32_t> start, idx;
  void build(const std::vector<V3<RF>>& pts, RF cell) {
    cs = cell;
    V3<RF> mn(1e30f, 1e30f, 1e30f), mx(-1e30f, -1e30f, -1e30f);
    for (auto& p : pts) {
      mn.x = std::min(mn.x, p.x); mn.y = std::min(mn.y, p.y); mn.z = std::min(mn.z, p.z);
      mx.x = std::max(mx.x, p.x); mx.y = std::max(mx.y, p.y); mx.z = std::max(mx.z, p.z);
    }
    if (pts.empty()) { mn = V3<RF>(0, 0, 0); mx = mn; }
    lo = mn - V3<RF>(cs, cs, cs);
    nx = int((mx.x - lo.x) / cs) + 3; ny = int((mx.y - lo.y) / cs) + 3; nz = int((mx.z - lo.z) / cs) + 3;
    const size_t nc = size_t(nx) * ny * nz;
    start.assign(nc + 1, 0);
    idx.resize(pts.size());
    std::vector<uint32_t> key(pts.size());
    for (size_t i = 0; i < pts.size(); ++i) { key[i] = cell_of(pts[i]); start[key[i] + 1]++; }
    for (size_t c = 0; c < nc; ++c) start[c + 1] += start[c];
    std::vector<uint32_t> cur(start.begin(), start.end() - 1);
    for (size_t i = 0; i < pts.size(); ++i) idx[cur[key[i]]++] = uint32_t(i);  // stable by index
  }
  void coords(const V3<RF>& p, int& cx, int& cy, int& cz) const {
    cx = int(std::floor((p.x - lo.x) / cs)); cy = int(std::floor((p.y - lo.y) / cs)); cz = int(std::floor((p.z - lo.z) / cs));
    cx = std::min(std::max(cx, 0), nx - 1); cy = std::min(std::max(cy, 0), ny - 1); cz = std::min(std::max(cz, 0), nz - 1);
  }
  uint32_t cell_of(const V3<RF>& p) const { int cx, cy, cz; coords(p, cx, cy, cz); return uint32_t((size_t(cz) * ny + cy) * nx + cx); }
  // Neighbours of q among pts (indices into pts), ascending index order within
  // each cell, cells in (dz, dy, dx) order; finally sorted ascending.
  void query(const V3<RF>& q, const std::vector<V3<RF>>& pts, RF h2, int64_t self, std::vector<uint32_t>& out) const {
    out.clear();
    int cx, cy, cz;
    coords(q, cx, cy, cz);
    for (int dz = -1; dz <= 1; ++dz) for (int dy = -1; dy <= 1; ++dy) for (int dx = -1; dx <= 1; ++dx) {
      const int x = cx + dx, y = cy + dy, z = cz + dz;
      if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) continue;
      const size_t c = (size_t(z) * ny + y) * nx + x;
      for (uint32_t k = start[c]; k < start[c + 1]; ++k) {
        const uint32_t j = idx[k];
        if (int64_t(j) == self) continue;
        if (dist2_nofma(q, pts[j]) < h2) out.push_back(j);
      }
    }
    std::sort(out.begin(), out.end());
  }
};

template <typename RF, typename RR> struct Body {
  bool dynamic = false, animated = false;
  std::vector<V3<RF>> x0, x, v;  // local samples, world positions, velocities
  std::vector<RF> V;             // Akinci volumes
  std::vector<RF> cvol0, cdens0; // rigid-contact rest volume and density (RigidContactSolver.cpp:140-170)
  V3<RR> cF, cT;                 // sums of the contact forces / torques handed to the body in the last step
  RR mass = 0, invMass = 0;
  M3<RR> I0, I, Iinv;
  V3<RR> xrb, vrb, wrb, x_init, v_init, w_init;  // x_init etc. = state restored by reset
  Quat<RR> q, q_init;
  V3<RR> v0, w0;                 // init_v_rb / init_omega_rb of the driver
  std::vector<V3<RR>> Fth, Tth;  // per-thread partials, BoundaryModel.h:46-58
  V3<RR> F_last, T_last;         // force/torque of the last step (getForce/getTorque)
  // per-boundary-particle Jacobians, BoundaryModel_Akinci2012.h:42-50
  std::vector<M3<RR>> gFv, gFx, gFw, gTw, gTx, gTv;
  std::vector<M34<RR>> gFq, gTq;
  // per-body sums, BoundaryModel_Akinci2012.cpp:453-497
  M3<RR> nFv, nFx, nFw, nTw, nTx, nTv;
  M34<RR> nFq, nTq;
  // forward-mode rigid Jacobians, BoundaryModel_Akinci2012.cpp:94-106
  M3<RR> Jvv, Jvw, Jwv, Jww, Jxv, Jxw;
  M43<RR> Jqv, Jqw;
  // RigidBodyGradientManager entries (R, R) of this body: the sums it was last
  // handed (update_rigid_body_gradient_manager, BoundaryModel_Akinci2012.cpp:952-967)
  // and its Jacobians (RigidBodyGradientManager.h:58-65)
  M3<RR> mFv, mFx, mFw, mTw, mTx, mTv;
  M34<RR> mFq, mTq;
  M3<RR> mJvv, mJvw, mJwv, mJww, mJxv, mJxw;
  M43<RR> mJqv, mJqw;
  void reset_jacobians() {
    Jvv.setIdentity(); Jww.setIdentity();
    Jvw.setZero(); Jwv.setZero(); Jxv.setZero(); Jxw.setZero();
    Jqv.setZero(); Jqw.setZero();
    // RigidBodyGradientManager::reset, RigidBodyGradientManager.cpp:474-523
    mJvv.setIdentity(); mJww.setIdentity();
    mJvw.setZero(); mJwv.setZero(); mJxv.setZero(); mJxw.setZero();
    mJqv.setZero(); mJqw.setZero();
    mFv.setZero(); mFx.setZero(); mFw.setZero(); mTw.setZero(); mTx.setZero(); mTv.setZero();
    mFq.setZero(); mTq.setZero();
  }
};

struct Stats {
  int iterations = 0, iterationsV = 0;
  double avgDensityErr = 0, avgDivErr = 0;
};

// What the exact adjoint reads back for one step: the values every operation
// of the step consumed, recorded as the step runs (dfsph_adjoint.hpp).
template <typename RF, typename RR> struct BodyTape {
  std::vector<V3<RF>> xb, vb;        // boundary samples at the start of the step
  V3<RR> xrb, vpre, wpre, v, w, F, T;  // v, w after the release-window override
  Quat<RR> q;
  M3<RR> I, Iinv;
  int ramp = 0;
  RR factor = 0;
  bool animated = false;
};
template <typename RF, typename RR> struct StepTape {
  std::vector<V3<RF>> x, G;          // positions during the step, sum_grad_p_k
  std::vector<std::vector<uint32_t>> nFF;
  std::vector<std::vector<std::vector<uint32_t>>> nFB;
  std::vector<RF> rho, alpha;
  std::vector<RF> kvPrev, d0, kvW;   // divergence warm start: kappa_v in, d, kappa_v used
  std::vector<RF> kPrev, rsW, kW;    // pressure warm start: kappa in, rho*, kappa used
  std::vector<std::vector<V3<RF>>> vD, vP;   // velocity at the start of each iteration
  std::vector<V3<RF>> vDiv;          // velocity after the divergence solve
  std::vector<std::vector<RF>> kD, bD, kP, bP;  // k_i and d / rho* of each iteration
  std::vector<BodyTape<RF, RR>> bodies;
  int Iv = 0, Ip = 0;
  // learned warm start: network features [N][6] and the velocity delta was taken at
  std::vector<double> feat;
  std::vector<V3<RF>> v0;
};

template <typename RF, typename RR> class Sim {
 public:
  Params P;
  Cubic<RF> K;
  Cubic<RF> KRC;  // rigid-contact kernel, radius rigidContactSupportRadiusFactor r (RigidContactSolver.cpp:1437-1455)
  // RigidBodyGradientManager across bodies, entries (R, RR) (RigidBodyGradientManager.h:58-65), run when the
  // rigid contact solver couples the bodies: J = grad_{x,v,omega,q}_n_to_{v0,omega0}[R][RR] and the sums
  // N = grad_net_{force,torque}_to_{v,x,omega,q}_n[R][RR] it was last handed.
  struct MgrJ { M3<RR> vv, vw, wv, ww, xv, xw; M43<RR> qv, qw; };
  struct MgrN { M3<RR> Fv, Fx, Fw, Tw, Tx, Tv; M34<RR> Fq, Tq; };
  std::vector<MgrJ> MJ;
  std::vector<MgrN> MN;
  Cohesion<RF> KC;
  Adhesion<RF> KA;
  RF h, dt, rho0, V_f, m_f;
  V3<RF> g;
  int nthreads = 1;
  RF time = 0;
  int stepCount = 0;
  // fluid (reference AoS layout, FluidModel.h:110-151)
  std::vector<V3<RF>> x, v, s;
  std::vector<V3<RF>> accel, normal;  // FluidModel accelerations, Akinci 2013 normals
  std::vector<RF> color, gradC2;      // He 2014 colour field and |grad c|^2
  std::vector<RF> rho, alpha, bAdv, kappa, kappaV;
  std::vector<Body<RF, RR>> bodies;
  // neighbour lists per point-set pair
  std::vector<std::vector<uint32_t>> nFF;                // F -> F
  std::vector<std::vector<std::vector<uint32_t>>> nFB;   // [body][i] F -> B
  std::vector<std::vector<std::vector<uint32_t>>> nBF;   // [body][j] B -> F (dynamic/animated bodies)
  Stats stats;
  // Exact-adjoint support: per-step tapes, and optional fixed Jacobi iteration
  // counts per step (so that finite differences see the same iteration counts).
  bool record = false;
  std::vector<StepTape<RF, RR>> tapes;
  StepTape<RF, RR>* tp = nullptr;
  std::vector<int> fixIv, fixIp;
  int fixPos = 0;
  // Learned warm start (when net.enabled): the network's predictions from the
  // features of the step's start replace the halved previous stiffness of
  // both warm starts.
  WarmNet net;
  std::vector<double> pred;  // [N][2]: kappa_v dt, kappa dt^2

  explicit Sim(const Params& p) : P(p) {
    h = RF(4.0) * RF(P.radius);  // Simulation.cpp:195-216, support radius = 4r
    K.set(h);
    KRC.set(RF(P.rigidContactSupportRadiusFactor) * RF(P.radius));
    KC.set(h);  // Simulation.cpp:397-398
    KA.set(h);
    dt = RF(P.dt);
    rho0 = RF(P.rho0);
    const RF diam = RF(2.0) * RF(P.radius);
    V_f = RF(0.8) * diam * diam * diam;  // FluidModel.cpp:255-275
    m_f = V_f * rho0;
    g = V3<RF>(RF(P.gravity[0]), RF(P.gravity[1]), RF(P.gravity[2]));
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
  }

  void set_fluid(size_t n, const double* xs, const double* vs) {
    x.resize(n); v.resize(n); s.assign(n, V3<RF>());
    rho.assign(n, 0); alpha.assign(n, 0); bAdv.assign(n, 0); kappa.assign(n, 0); kappaV.assign(n, 0);
    for (size_t i = 0; i < n; ++i) {
      x[i] = V3<RF>(RF(xs[3 * i]), RF(xs[3 * i + 1]), RF(xs[3 * i + 2]));
      v[i] = vs ? V3<RF>(RF(vs[3 * i]), RF(vs[3 * i + 1]), RF(vs[3 * i + 2])) : V3<RF>();
    }
  }

  // Dynamic3dRigidBody::determineMassProperties, Dynamic3dRigidBody.h:184-213:
  // each sample carries (4/3) pi r^3 rho_body; inertia about the mesh origin.
  int add_body(size_t n, const double* x0s, int dynamic, const double* xrb, const double* qwxyz, double density) {
    Body<RF, RR> b;
    b.dynamic = dynamic != 0;
    b.x0.resize(n);
    for (size_t j = 0; j < n; ++j) b.x0[j] = V3<RF>(RF(x0s[3 * j]), RF(x0s[3 * j + 1]), RF(x0s[3 * j + 2]));
    b.x = b.x0; b.v.assign(n, V3<RF>()); b.V.assign(n, 0);
    b.xrb = V3<RR>(RR(xrb[0]), RR(xrb[1]), RR(xrb[2]));
    b.q = Quat<RR>(RR(qwxyz[0]), RR(qwxyz[1]), RR(qwxyz[2]), RR(qwxyz[3]));
    b.x_init = b.xrb; b.q_init = b.q;
    const RR r = RR(P.radius);
    const RR volume = RR(4.0 / 3.0 * M_PI) * r * r * r;
    const RR dm = volume * RR(density);
    b.mass = 0; b.I0.setZero();
    for (size_t j = 0; j < n; ++j) {
      b.mass += dm;
      const V3<RR> p(b.x0[j]);
      b.I0 += (M3<RR>::Identity() * p.dot(p) - outer(p, p)) * dm;
    }
    b.invMass = RR(1) / b.mass;
    update_inertia(b);
    b.Fth.assign(nthreads, V3<RR>()); b.Tth.assign(nthreads, V3<RR>());
    b.gFv.assign(n, M3<RR>()); b.gFx.assign(n, M3<RR>()); b.gFw.assign(n, M3<RR>());
    b.gTw.assign(n, M3<RR>()); b.gTx.assign(n, M3<RR>()); b.gTv.assign(n, M3<RR>());
    b.gFq.assign(n, M34<RR>()); b.gTq.assign(n, M34<RR>());
    b.reset_jacobians();
    bodies.push_back(b);
    return int(bodies.size()) - 1;
  }

  // Sets the boundary transform, then the Akinci volumes
  // (RigidBody3dBoundarySimulator::deferredInit, RigidBody3dBoundarySimulator.cpp:249-265).
  void init() {
    update_boundary_particles(true);
    update_boundary_volume();
    nFB.assign(bodies.size(), {});
    nBF.assign(bodies.size(), {});
    if (P.useRigidContact) contact_rest();
    reset_multi_manager();
  }
  bool multi_manager() const { return P.useRigidContact && P.computeGradient && P.useGradientManager; }
  // RigidBodyGradientManager::reset (RigidBodyGradientManager.cpp:474-523) for every pair.
  void reset_multi_manager() {
    const size_t nb = bodies.size();
    MJ.assign(nb * nb, MgrJ{});
    MN.assign(nb * nb, MgrN{});
    for (size_t R = 0; R < nb; ++R) { MJ[R * nb + R].vv.setIdentity(); MJ[R * nb + R].ww.setIdentity(); }
  }

  // ---- rigid-rigid contact (RigidContactSolver, penalty form) -------------
  // Neighbours of sample j of body bi among every body's samples: the
  // simulation's search radius h and predicate, the sample itself left out
  // (CompactNSearch over the boundary point sets, activated by the solver).
  struct SampleGrid {
    std::vector<V3<RF>> pts;
    std::vector<std::pair<uint32_t, uint32_t>> id;  // (body, sample)
    Grid<RF> g;
  };
  void sample_grid(SampleGrid& G) const {
    G.pts.clear(); G.id.clear();
    for (size_t bi = 0; bi < bodies.size(); ++bi)
      for (size_t j = 0; j < bodies[bi].x.size(); ++j) { G.pts.push_back(bodies[bi].x[j]); G.id.push_back({uint32_t(bi), uint32_t(j)}); }
    G.g.build(G.pts, h * RF(1.001));
  }
  // RigidContactSolver constructor (RigidContactSolver.cpp:140-170): per
  // sample, vol0 = gamma / (W(0) + sum over its own body's neighbours of W),
  // then density0 = vol0 W(0) + sum of the own neighbours' vol0 W.
  void contact_rest() {
    SampleGrid G;
    sample_grid(G);
    const RF W0 = KRC.w0, gamma = RF(P.rigidContactGamma);
    std::vector<std::vector<uint32_t>> nb(G.pts.size());
    for (size_t a = 0; a < G.pts.size(); ++a) G.g.query(G.pts[a], G.pts, h * h, int64_t(a), nb[a]);
    std::vector<RF> vol(G.pts.size());
    for (size_t a = 0; a < G.pts.size(); ++a) {
      RF sw = W0;
      for (uint32_t k : nb[a])
        if (G.id[k].first == G.id[a].first) sw += KRC.W(G.pts[a] - G.pts[k]);
      vol[a] = gamma / sw;
    }
    for (auto& b : bodies) { b.cvol0.assign(b.x.size(), RF(0)); b.cdens0.assign(b.x.size(), RF(0)); }
    for (size_t a = 0; a < G.pts.size(); ++a) {
      RF d = vol[a] * W0;
      for (uint32_t k : nb[a])
        if (G.id[k].first == G.id[a].first) d += vol[k] * KRC.W(G.pts[a] - G.pts[k]);
      auto& b = bodies[G.id[a].first];
      b.cvol0[G.id[a].second] = vol[a];
      b.cdens0[G.id[a].second] = d;
    }
  }
  // Dynamic3dRigidBody::addForce / addTorque (Dynamic3dRigidBody.h:113-142):
  // each call changes the velocity at once; with the gyroscopic mode every
  // torque call adds I^-1 ((I w) x w) dt again.
  void body_add_force(Body<RF, RR>& b, const V3<RR>& f, RR dtR) const { b.vrb += f * (b.invMass * dtR); }
  void body_add_torque(Body<RF, RR>& b, const V3<RR>& t, RR dtR) const {
    if (P.gyroscopic) {
      const V3<RR> L = b.I * b.wrb;
      b.wrb += (b.Iinv * (L.cross(b.wrb) + t)) * dtR;
    } else {
      b.wrb += (b.Iinv * t) * dtR;
    }
  }
  // RigidBody3dBoundarySimulator::velocityTimeStep with the contact solver
  // (RigidBody3dBoundarySimulator.cpp:284-298): gravity for every dynamic body,
  // then RigidContactSolver::solveRigidContact in its penalty form
  // (RigidContactSolver.cpp:347-366, beforePenaltyInitialize :307-345,
  // solveRigidContactPenalty :419-555): each sample with a neighbour on another
  // body gets density vol0 W(0) + sum over all its neighbours of vol0_k W;
  // where density / density0 > 1 it pushes its (dynamic) body with
  //   f = -beta (density / density0 - 1) n - mu |f_n| normalized(v_rel),
  //   n = x - (sum_own x_k W) / (sum_own W),
  //   v_rel = v - (sum_other v_k W) / (sum_own W)   (the own-body weight sum, as the reference divides),
  // torque (x - x_rb) x f, one addForce/addTorque per sample in body and sample
  // order; then updateBoundaryForces once per boundary model: the first call
  // applies the fluid force and torque (dynamic bodies that are not animated)
  // and clears them, the later ones apply zeros (BoundarySimulator.cpp:10-36).
  void contact_velocity_step(RR dtR) {
    const V3<RR> gR(RR(P.gravity[0]), RR(P.gravity[1]), RR(P.gravity[2]));
    for (auto& b : bodies)
      if (b.dynamic) body_add_force(b, gR * b.mass, dtR);  // addGravity (Dynamic3dRigidBody.h:228-233)
    // sample velocities after gravity (beforePenaltyInitialize :326-329); zero for static bodies
    std::vector<std::vector<V3<RF>>> vel(bodies.size());
    for (size_t bi = 0; bi < bodies.size(); ++bi) {
      const auto& b = bodies[bi];
      vel[bi].assign(b.x.size(), V3<RF>());
      if (b.dynamic || b.animated)
        for (size_t j = 0; j < b.x.size(); ++j) vel[bi][j] = V3<RF>(b.vrb + b.wrb.cross(V3<RR>(b.x[j]) - b.xrb));
    }
    SampleGrid G;
    sample_grid(G);
    const RF W0 = KRC.w0, beta = RF(P.rigidContactBeta), mu = RF(P.rigidContactFriction);
    for (auto& b : bodies) { b.cF = V3<RR>(); b.cT = V3<RR>(); }
    // grad_net_{force,torque}_to_*_before_contact[R][RR] (solveRigidContactPenalty :429-441, :471-553)
    const size_t NBD = bodies.size();
    const bool grads = multi_manager();
    std::vector<MgrN> CN(grads ? NBD * NBD : 0);
    std::vector<uint32_t> nb;
    for (size_t a = 0; a < G.pts.size(); ++a) {
      const uint32_t R = G.id[a].first, r = G.id[a].second;
      auto& b = bodies[R];
      if (!b.dynamic) continue;  // StaticRigidBody::addForce does nothing
      G.g.query(G.pts[a], G.pts, h * h, int64_t(a), nb);
      bool contact = false;
      for (uint32_t k : nb) contact |= G.id[k].first != R;
      if (!contact) continue;
      const V3<RF>& xr = b.x[r];
      RF dens = b.cvol0[r] * W0;
      V3<RF> sx, sv;
      RF sw = 0;
      // gradient sums over the other bodies' neighbours, in RR (double): d density / d x_R
      // (the simulation's kernel gradient), / d q_R, / d q_Rk, and the weighted r_k;
      // RR = the body of the last such neighbour (point sets in body order)
      V3<RR> gdx, avg_rk;
      RR gdq[4] = {0, 0, 0, 0}, gdqk[4] = {0, 0, 0, 0};
      uint32_t RRb = R;
      for (uint32_t k : nb) {
        const auto& bk = bodies[G.id[k].first];
        const uint32_t kk = G.id[k].second;
        const RF w = KRC.W(xr - bk.x[kk]);
        dens += bk.cvol0[kk] * w;
        if (G.id[k].first == R) { sx += bk.x[kk] * w; sw += w; }
        else {
          sv += vel[G.id[k].first][kk] * w;
          if (grads) {
            RRb = G.id[k].first;
            const V3<RR> gw = V3<RR>(K.gradW(xr - bk.x[kk])) * RR(bk.cvol0[kk]);
            gdx += gw;
            avg_rk += (V3<RR>(bk.x[kk]) - bk.xrb) * RR(w);
            const M34<RR> Jq = grad_Rqp_to_q(b.q, V3<RR>(b.x0[r]));
            const M34<RR> Jqk = grad_Rqp_to_q(bk.q, V3<RR>(bk.x0[kk]));
            for (int c = 0; c < 4; ++c) {
              gdq[c] += gw.x * Jq(0, c) + gw.y * Jq(1, c) + gw.z * Jq(2, c);
              gdqk[c] -= gw.x * Jqk(0, c) + gw.y * Jqk(1, c) + gw.z * Jqk(2, c);
            }
          }
        }
      }
      const RF ratio = dens / b.cdens0[r];
      if (!(ratio > RF(1))) continue;
      const V3<RF> n = xr - sx * (RF(1) / sw);
      const V3<RF> vrel = vel[R][r] - sv * (RF(1) / sw);
      const V3<RF> fn = n * (-beta * std::max(ratio - RF(1), RF(0)));
      const RF vn2 = vrel.squaredNorm();
      const V3<RF> unit = vn2 > RF(0) ? vrel * (RF(1) / std::sqrt(vn2)) : vrel;  // Eigen normalized()
      const V3<RF> f = fn + unit * (-mu * fn.norm());
      const V3<RF> rr = V3<RF>(V3<RR>(xr) - b.xrb);
      body_add_force(b, V3<RR>(f), dtR);
      body_add_torque(b, V3<RR>(rr.cross(f)), dtR);
      b.cF += V3<RR>(f);
      b.cT += V3<RR>(rr.cross(f));
      if (grads) {
        // :471-553: every density derivative through f_n = -beta (rho / rho0 - 1) n with
        // d f_n / d rho taken as -beta n, the friction through |f_n| and normalized(v_rel)
        const V3<RR> nR(n), fR(f), fnR(fn), urR(unit), vr(vrel), rrR(rr);
        const RR muR = RR(mu);
        avg_rk = avg_rk * (RR(1) / RR(sw));
        const V3<RR> gfn = nR * (-RR(beta));
        const RR fnn = fnR.norm();
        const V3<RR> gff = urR * (-muR * fnR.dot(gfn) / fnn);
        const V3<RR> gf = gfn + gff;
        M3<RR> Gx;
        M34<RR> Gq, Gqk;
        for (int i = 0; i < 3; ++i) {
          for (int c = 0; c < 3; ++c) Gx(i, c) = gf[i] * gdx[c];
          for (int c = 0; c < 4; ++c) { Gq(i, c) = gf[i] * gdq[c]; Gqk(i, c) = gf[i] * gdqk[c]; }
        }
        const M3<RR> Gvr = (M3<RR>::Identity() - outer(urR, urR)) * (-muR * fnn / vr.norm());
        const M3<RR> Gw = Gvr * skew(rrR).transpose();
        const M3<RR> GwK = -(Gvr * skew(avg_rk).transpose());
        MgrN& o = CN[R * NBD + R];
        o.Fx += Gx;
        o.Tx += skew(rrR) * Gx;
        o.Fq += Gq;
        o.Tq += skew(rrR) * Gq + skew(fR).transpose() * grad_Rqp_to_q(b.q, V3<RR>(b.x0[r]));
        o.Fv += Gvr;
        o.Fw += Gw;
        if (RRb != R) {
          MgrN& c = CN[R * NBD + RRb];
          c.Fv += -Gvr;
          c.Fw += GwK;
          c.Fx += -Gx;
          c.Tx += -(skew(rrR) * Gx);
          c.Fq += Gqk;
          c.Tq += skew(rrR) * Gqk;
        }
      }
    }
    for (size_t call = 0; call < bodies.size(); ++call)
      for (auto& b : bodies) {
        if (!b.dynamic) continue;
        V3<RR> F, T;
        if (call == 0) {
          force_and_torque(b, F, T);
          b.F_last = F; b.T_last = T;
          for (int t = 0; t < nthreads; ++t) { b.Fth[t] = V3<RR>(); b.Tth[t] = V3<RR>(); }
        }
        if (!b.animated) { body_add_force(b, F, dtR); body_add_torque(b, T, dtR); }
      }
    // update_rigid_body_gradient_manager (:1341-1370): the manager's sums of every
    // dynamic pair are replaced by the contact solver's
    if (grads)
      for (size_t R = 0; R < NBD; ++R)
        for (size_t Q = 0; Q < NBD; ++Q)
          if (bodies[R].dynamic && bodies[Q].dynamic) MN[R * NBD + Q] = CN[R * NBD + Q];
  }

  static void update_inertia(Body<RF, RR>& b) {  // Dynamic3dRigidBody.h:215-219
    const M3<RR> R = b.q.toRotationMatrix();
    b.I = R * b.I0 * R.transpose();
    b.Iinv = inverse(b.I);
  }

  // SimulatorBase::updateBoundaryParticles, SimulatorBase.cpp:1827-1858.
  void update_boundary_particles(bool force) {
    for (auto& b : bodies) {
      const bool moving = b.dynamic || b.animated;
      if (!(moving || force)) continue;
#pragma omp parallel for schedule(static)
      for (long j = 0; j < long(b.x0.size()); ++j) {
        const V3<RR> xr = b.q.rotate(V3<RR>(b.x0[j])) + b.xrb;
        b.x[j] = V3<RF>(xr);
        if (moving) b.v[j] = V3<RF>(b.wrb.cross(V3<RR>(b.x[j]) - b.xrb) + b.vrb);
        else b.v[j] = V3<RF>();
      }
    }
  }

  // Simulation::updateBoundaryVolume, Simulation.cpp:831-902, with
  // BoundaryModel_Akinci2012::computeBoundaryVolume, :257-284:
  // V_b = 1 / (W0 + sum W) over all static samples for a static body, over the
  // body's own samples for a dynamic one.
  void update_boundary_volume() {
    const RF h2 = h * h;
    std::vector<V3<RF>> statics;
    std::vector<int> sbody; std::vector<uint32_t> sidx;
    for (size_t bi = 0; bi < bodies.size(); ++bi)
      if (!bodies[bi].dynamic && !bodies[bi].animated)
        for (size_t j = 0; j < bodies[bi].x.size(); ++j) { statics.push_back(bodies[bi].x[j]); sbody.push_back(int(bi)); sidx.push_back(uint32_t(j)); }
    auto volumes = [&](const std::vector<V3<RF>>& pts, std::vector<RF>& out) {
      Grid<RF> gr; gr.build(pts, h * RF(1.001));
      out.assign(pts.size(), 0);
#pragma omp parallel
      {
        std::vector<uint32_t> nb;
#pragma omp for schedule(static)
        for (long i = 0; i < long(pts.size()); ++i) {
          gr.query(pts[i], pts, h2, i, nb);
          RF delta = K.w0;
          for (uint32_t k : nb) delta += K.W(pts[i] - pts[k]);
          out[i] = RF(1.0) / delta;
        }
      }
    };
    std::vector<RF> vs;
    volumes(statics, vs);
    for (size_t k = 0; k < statics.size(); ++k) bodies[sbody[k]].V[sidx[k]] = vs[k];
    for (auto& b : bodies)
      if (b.dynamic || b.animated) { std::vector<RF> vb; volumes(b.x, vb); b.V = vb; }
  }

  // CompactNSearch::find_neighbors as wrapped by Simulation::performNeighborhoodSearch
  // (Simulation.cpp:743-748); point-set pairs F->F, F->B and B->F as activated at
  // Simulation.cpp:883-899.  B->F is kept for dynamic/animated bodies only (the
  // gradient loop visits no other).
  void neighborhood_search() {
    const RF h2 = h * h;
    const RF cell = h * RF(1.001);
    const long N = long(x.size());
    Grid<RF> gf; gf.build(x, cell);
    nFF.assign(N, {});
#pragma omp parallel
    {
      std::vector<uint32_t> nb;
#pragma omp for schedule(static)
      for (long i = 0; i < N; ++i) { gf.query(x[i], x, h2, i, nb); nFF[i] = nb; }
    }
    nFB.assign(bodies.size(), {});
    nBF.assign(bodies.size(), {});
    for (size_t bi = 0; bi < bodies.size(); ++bi) {
      auto& b = bodies[bi];
      Grid<RF> gb; gb.build(b.x, cell);
      nFB[bi].assign(N, {});
#pragma omp parallel
      {
        std::vector<uint32_t> nb;
#pragma omp for schedule(static)
        for (long i = 0; i < N; ++i) { gb.query(x[i], b.x, h2, -1, nb); nFB[bi][i] = nb; }
      }
      if (b.dynamic || b.animated) {
        nBF[bi].assign(b.x.size(), {});
#pragma omp parallel
        {
          std::vector<uint32_t> nb;
#pragma omp for schedule(static)
          for (long j = 0; j < long(b.x.size()); ++j) { gf.query(b.x[j], x, h2, -1, nb); nBF[bi][j] = nb; }
        }
      }
    }
  }

  // TimeStepDiffDFSPH::beginStep, TimeStepDiffDFSPH.cpp:353-430.
  void begin_step() {
    stepCount++;
    const RF uat = RF(P.uniformAccTime);
    for (size_t k = 0; k < bodies.size(); ++k) {
      auto& b = bodies[k];
      if (!b.dynamic) continue;
      if (P.releaseMode) {  // :381-403: only boundary model 1 is driven
        if (k != 1) continue;
        RR factor = 1;
        if (time <= uat + dt) {
          if (uat > RF(1e-3)) b.animated = true;
          factor = 0;
        } else {
          b.animated = false;
        }
        b.vrb = b.v0 * factor;
        b.wrb = b.w0 * factor;
        if (tp) { tp->bodies[k].ramp = 1; tp->bodies[k].factor = factor; }
        continue;
      }
      if (time <= uat + dt) {  // :405-426
        RF factor = 1;
        if (uat > RF(1e-3)) {
          factor = (time / uat) > RF(1) ? RF(1) : (time / uat);
          b.animated = true;
        }
        b.vrb = b.v0 * RR(factor);
        b.wrb = b.w0 * RR(factor);
        if (tp) { tp->bodies[k].ramp = 1; tp->bodies[k].factor = RR(factor); }
      } else {
        b.animated = false;
      }
    }
  }

  // TimeStep::computeDensities, TimeStep.cpp:147-202 (scalar form).
  void density_of(long i) {
    RF d = V_f * K.w0;
    for (uint32_t j : nFF[i]) d += V_f * K.W(x[i] - x[j]);
    for (size_t bi = 0; bi < bodies.size(); ++bi)
      for (uint32_t j : nFB[bi][i]) d += bodies[bi].V[j] * K.W(x[i] - bodies[bi].x[j]);
    rho[i] = d * rho0;
  }
  void compute_densities() {
    const long N = long(x.size());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) density_of(i);
  }

  // TimeStepDiffDFSPH::computeDFSPHFactor, TimeStepDiffDFSPH.cpp:883-962.
  void compute_dfsph_factor() {
    const long N = long(x.size());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) factor_of(i);
  }
  void factor_of(long i) {
    const RF eps = RF(1e-5);
    {
      RF sum_grad_p_k = 0;
      V3<RF> grad_p_i;
      for (uint32_t j : nFF[i]) {
        const V3<RF> grad_p_j = K.gradW(x[i] - x[j]) * (-V_f);
        sum_grad_p_k += grad_p_j.squaredNorm();
        grad_p_i -= grad_p_j;
      }
      for (size_t bi = 0; bi < bodies.size(); ++bi)
        for (uint32_t j : nFB[bi][i]) {
          const V3<RF> grad_p_j = K.gradW(x[i] - bodies[bi].x[j]) * (-bodies[bi].V[j]);
          grad_p_i -= grad_p_j;
        }
      s[i] = -grad_p_i;
      sum_grad_p_k += grad_p_i.squaredNorm();
      alpha[i] = (sum_grad_p_k > eps) ? -RF(1.0) / sum_grad_p_k : RF(0);
    }
  }

  // TimeStepDiffDFSPH::computeDensityAdv, TimeStepDiffDFSPH.cpp:1926-1975.
  void compute_density_adv(long i) {
    RF delta = 0;
    for (uint32_t j : nFF[i]) delta += V_f * (v[i] - v[j]).dot(K.gradW(x[i] - x[j]));
    for (size_t bi = 0; bi < bodies.size(); ++bi)
      for (uint32_t j : nFB[bi][i]) delta += bodies[bi].V[j] * (v[i] - bodies[bi].v[j]).dot(K.gradW(x[i] - bodies[bi].x[j]));
    RF da = rho[i] / rho0 + dt * delta;
    bAdv[i] = std::max(da, RF(1.0));
  }

  // Density change rate before the clamp (TimeStepDiffDFSPH.cpp:1977-2041),
  // and the neighbour count.
  RF density_change_raw(long i, unsigned& nn) const {
    RF da = 0;
    nn = unsigned(nFF[i].size());
    for (uint32_t j : nFF[i]) da += V_f * (v[i] - v[j]).dot(K.gradW(x[i] - x[j]));
    for (size_t bi = 0; bi < bodies.size(); ++bi) {
      nn += unsigned(nFB[bi][i].size());
      for (uint32_t j : nFB[bi][i]) da += bodies[bi].V[j] * (v[i] - bodies[bi].v[j]).dot(K.gradW(x[i] - bodies[bi].x[j]));
    }
    return da;
  }

  // TimeStepDiffDFSPH::computeDensityChange, TimeStepDiffDFSPH.cpp:1977-2041.
  void compute_density_change(long i) {
    unsigned nn = 0;
    RF da = std::max(density_change_raw(i, nn), RF(0.0));
    if (nn < 20) da = 0;
    bAdv[i] = da;
  }

  // Features of the learned warm start at the step's start (after densities
  // and factors): rho/rho0 - 1, dt * (density change rate before the clamp;
  // 0 without the divergence solver), neighbour count, boundary part of
  // rho/rho0, previous kappa and kappa_v (stored units), then the network's
  // predictions of (kappa_v dt, kappa dt^2).
  void learned_predict() {
    const long N = long(x.size());
    pred.assign(2 * N, 0.0);
    std::vector<double> feat(6 * N);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      unsigned nn = 0;
      const RF da = density_change_raw(i, nn);
      RF bd = 0;
      for (size_t bi = 0; bi < bodies.size(); ++bi)
        for (uint32_t j : nFB[bi][i]) bd += bodies[bi].V[j] * K.W(x[i] - bodies[bi].x[j]);
      double* f = &feat[6 * i];
      f[0] = double(rho[i] / rho0 - RF(1));
      f[1] = P.enableDivergenceSolver ? double(dt * da) : 0.0;
      f[2] = double(nn);
      f[3] = double(bd);
      f[4] = double(kappa[i]);
      f[5] = double(kappaV[i]);
      net.forward(f, &pred[2 * i]);
    }
    if (tp) { tp->feat = std::move(feat); tp->v0 = v; }
  }

  // BoundaryModel::addForce, BoundaryModel.h:46-58 (per-thread partials).
  void add_force(Body<RF, RR>& b, const V3<RF>& pos, const V3<RF>& f) {
    if (!b.dynamic) return;
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    b.Fth[tid] += V3<RR>(f);
    b.Tth[tid] += (V3<RR>(pos) - b.xrb).cross(V3<RR>(f));
  }

  // The velocity update shared by the warm starts and the Jacobi iterations:
  // fluid pairs with |kSum| > eps, boundary pairs when |k_i| > eps, with the
  // Akinci reaction force on the boundary (TimeStepDiffDFSPH.cpp:1024-1055,
  // :1123-1196, :1753-1784, :1847-1901).
  void velocity_update(long i, RF ki, const std::vector<RF>& kfield) {
    const RF eps = RF(1e-5);
    V3<RF>& vi = v[i];
    for (uint32_t j : nFF[i]) {
      const RF kj = kfield[j];
      const RF kSum = ki + kj;  // density0 ratio is 1 for a single fluid
      if (std::fabs(kSum) > eps) {
        const V3<RF> grad_p_j = K.gradW(x[i] - x[j]) * (-V_f);
        vi -= grad_p_j * (dt * kSum);
      }
    }
    if (std::fabs(ki) > eps) {
      const RF invH = RF(1.0) / dt;
      for (size_t bi = 0; bi < bodies.size(); ++bi) {
        auto& b = bodies[bi];
        for (uint32_t j : nFB[bi][i]) {
          const V3<RF> grad_p_j = K.gradW(x[i] - b.x[j]) * (-b.V[j]);
          const V3<RF> velChange = grad_p_j * (-dt * RF(1.0) * ki);
          vi += velChange;
          const V3<RF> force = velChange * (-m_f) * invH;
          add_force(b, b.x[j], force);
        }
      }
    }
  }

  // TimeStepDiffDFSPH::warmstartDivergenceSolve, TimeStepDiffDFSPH.cpp:1698-1788.
  void warmstart_divergence_solve() {
    const long N = long(x.size());
    const RF invH = RF(1.0) / dt;
    if (tp) tp->kvPrev = kappaV;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      compute_density_change(i);
      if (bAdv[i] > RF(0.0))
        kappaV[i] = net.enabled ? RF(pred[2 * i]) * invH : RF(0.5) * std::max(kappaV[i], RF(-0.5)) * invH;
      else kappaV[i] = 0;
    }
    if (tp) { tp->d0 = bAdv; tp->kvW = kappaV; }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) velocity_update(i, kappaV[i], kappaV);
  }

  // TimeStepDiffDFSPH::warmstartPressureSolve, TimeStepDiffDFSPH.cpp:964-1059.
  void warmstart_pressure_solve() {
    const long N = long(x.size());
    const RF invH2 = RF(1.0) / (dt * dt);
    if (tp) tp->kPrev = kappa;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      compute_density_adv(i);
      if (bAdv[i] > RF(1.0))
        kappa[i] = net.enabled ? RF(pred[2 * i + 1]) * invH2 : RF(0.5) * std::max(kappa[i], RF(-0.00025)) * invH2;
      else kappa[i] = 0;
    }
    if (tp) { tp->rsW = bAdv; tp->kW = kappa; }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) velocity_update(i, kappa[i], kappa);
  }

  // TimeStepDiffDFSPH::pressureSolveIteration, TimeStepDiffDFSPH.cpp:1061-1220.
  RF pressure_solve_iteration() {
    const long N = long(x.size());
    const RF h2 = dt * dt;
    const RF invH2 = RF(1.0) / h2;
    compute_rigid_body_gradient(0);
    std::vector<RF> kfield(N);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) kfield[i] = (bAdv[i] - RF(1.0)) * alpha[i] * invH2;
    if (tp) { tp->vP.push_back(v); tp->kP.push_back(kfield); tp->bP.push_back(bAdv); }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      const RF ki = kfield[i];
      if (P.usePressureWarmstart) kappa[i] += ki;
      velocity_update(i, ki, kfield);
    }
    RF err_f = 0;
#pragma omp parallel for reduction(+ : err_f) schedule(static)
    for (long i = 0; i < N; ++i) {
      compute_density_adv(i);
      err_f += rho0 * bAdv[i] - rho0;
    }
    return err_f / RF(N);
  }

  // TimeStepDiffDFSPH::divergenceSolveIteration, TimeStepDiffDFSPH.cpp:1790-1922.
  RF divergence_solve_iteration() {
    const long N = long(x.size());
    const RF invH = RF(1.0) / dt;
    compute_rigid_body_gradient(1);
    std::vector<RF> kfield(N);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) kfield[i] = bAdv[i] * alpha[i] * invH;
    if (tp) { tp->vD.push_back(v); tp->kD.push_back(kfield); tp->bD.push_back(bAdv); }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      const RF ki = kfield[i];
      if (P.useDivergenceWarmstart) kappaV[i] += ki;
      velocity_update(i, ki, kfield);
    }
    RF err_f = 0;
#pragma omp parallel for reduction(+ : err_f) schedule(static)
    for (long i = 0; i < N; ++i) {
      compute_density_change(i);
      err_f += rho0 * bAdv[i];
    }
    return err_f / RF(N);
  }

  // TimeStepDiffDFSPH::divergenceSolve, TimeStepDiffDFSPH.cpp:770-881.
  void divergence_solve() {
    const long N = long(x.size());
    if (P.useDivergenceWarmstart) warmstart_divergence_solve();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) compute_density_change(i);
    int it = 0;
    bool chk = false;
    RF avg = 0;
    const RF eta = (RF(1.0) / dt) * RF(P.maxErrorV) * RF(0.01) * rho0;
    const int fixed = fixPos < int(fixIv.size()) ? fixIv[fixPos] : -1;
    while (fixed >= 0 ? it < fixed : ((!chk || it < 1) && it < P.maxIterationsV)) {
      avg = divergence_solve_iteration();
      chk = avg <= eta;
      it++;
    }
    stats.iterationsV = it;
    stats.avgDivErr = avg;
    if (P.useDivergenceWarmstart)
      for (long i = 0; i < N; ++i) kappaV[i] *= dt;
  }

  // TimeStepDiffDFSPH::pressureSolve, TimeStepDiffDFSPH.cpp:654-768.
  void pressure_solve() {
    const long N = long(x.size());
    if (P.usePressureWarmstart) warmstart_pressure_solve();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) compute_density_adv(i);
    int it = 0;
    bool chk = false;
    RF avg = 0;
    const RF eta = RF(P.maxError) * RF(0.01) * rho0;
    const int fixed = fixPos < int(fixIp.size()) ? fixIp[fixPos] : -1;
    while (fixed >= 0 ? it < fixed : ((!chk || it < P.minIterations) && it < P.maxIterations)) {
      avg = pressure_solve_iteration();
      chk = avg <= eta;
      it++;
    }
    stats.iterations = it;
    stats.avgDensityErr = avg;
    if (P.usePressureWarmstart) {
      const RF h2 = dt * dt;
      for (long i = 0; i < N; ++i) kappa[i] *= h2;
    }
  }

  // --------------------------------------------------------------------------
  // Reference-mode gradient: computeRigidBodyGradient + computeGradient
  // (TimeStepDiffDFSPH.cpp:1226-1482) and helpers (:1485-1694).
  // --------------------------------------------------------------------------
  V3<RF> grad_b_to_fluid_v(long i, int mode) const {  // :1632-1694
    const RF bi = bAdv[i];
    if (mode == 0 ? bi <= RF(1.0) : bi <= RF(0.0)) return V3<RF>();
    V3<RF> gsum;
    for (uint32_t j : nFF[i]) gsum += K.gradW(x[i] - x[j]) * V_f;
    for (size_t bi2 = 0; bi2 < bodies.size(); ++bi2)
      for (uint32_t j : nFB[bi2][i]) gsum += K.gradW(x[i] - bodies[bi2].x[j]) * bodies[bi2].V[j];
    if (mode == 0) gsum *= dt;
    return gsum;
  }

  void compute_gradient(Body<RF, RR>& b, long i, uint32_t j, RF b_i, RF ki, const V3<RF>& force, int mode) {
    const RF invH = RF(1.0) / dt;
    const RF invH2 = RF(1.0) / dt / dt;
    const V3<RF>& xi = x[i];
    const V3<RF>& xj = b.x[j];
    const RF Vj = b.V[j];
    const V3<RF> grad_p_j = K.gradW(xi - xj) * (-Vj);
    const M3<RF> H = K.gradGradW(xi - xj) * Vj;  // grad2_p_j_to_xj
    const RF factor = alpha[i];
    const V3<RF> vij = v[i] - b.v[j];
    // :1518-1577
    V3<RF> grad_b_xj;
    if (mode == 0) {
      if (bAdv[i] > RF(1.0)) grad_b_xj = K.gradW(xi - xj) * (-Vj) / rho0 - (K.gradGradW(xi - xj) * vij) * (dt * Vj);
    } else {
      if (bAdv[i] > RF(0.0)) grad_b_xj = -((K.gradGradW(xi - xj) * vij) * Vj);
    }
    // :1485-1516
    V3<RF> grad_factor_xj;
    if (factor < RF(0.)) grad_factor_xj = (H * s[i]) * (RF(2.) * factor * factor);
    const RF coeff = (mode == 0) ? invH2 : invH;
    const V3<RF> grad_ki_xj = (grad_b_xj * factor + grad_factor_xj * b_i) * coeff;
    const M3<RF> dvc_dxj = -(outer(grad_p_j, grad_ki_xj) + H * ki);
    // :1579-1628
    V3<RF> grad_b_vj;
    if (mode == 0) {
      if (bAdv[i] > RF(1.0)) grad_b_vj = K.gradW(xi - xj) * (-dt * Vj);
    } else {
      if (bAdv[i] > RF(0.0)) grad_b_vj = K.gradW(xi - xj) * (-Vj);
    }
    const V3<RF> grad_ki_vj = grad_b_vj * factor * coeff;
    const M3<RF> dvc_dvj = -outer(grad_p_j, grad_ki_vj);
    const V3<RF> grad_b_vi = grad_b_to_fluid_v(i, mode);
    const V3<RF> grad_ki_vi = grad_b_vi * factor * coeff;
    const M3<RF> dvc_dvi = -outer(grad_p_j, grad_ki_vi);
    const M3<RF> gFx = dvc_dxj * (-m_f);
    const M3<RF> gFv = inverse(M3<RF>::Identity() - dvc_dvi * dt) * (dvc_dvj * (-m_f));
    b.gFx[j] += M3<RR>(gFx);
    b.gFv[j] += M3<RR>(gFv);
    if (P.optimizeRotation) {  // :1425-1480
      const V3<RF> rj = V3<RF>(V3<RR>(xj) - b.xrb);
      const Quat<RF> q(b.q);
      const M34<RF> Jq = grad_Rqp_to_q(q, b.x0[j]);
      const V3<RF> om(b.wrb);
      const M34<RF> gFq = gFx * Jq + gFv * skew(om) * Jq;
      const M3<RF> gFw = gFv * skew(rj).transpose();
      const M34<RF> gTq = skew(rj) * gFq + skew(force).transpose() * Jq;
      const M3<RF> gTw = skew(rj) * gFw;
      b.gFw[j] += M3<RR>(gFw);
      b.gFq[j] += M34<RR>(gFq);
      b.gTw[j] += M3<RR>(gTw);
      b.gTq[j] += M34<RR>(gTq);
      b.gTv[j] += M3<RR>(skew(rj) * gFv);
      b.gTx[j] += M3<RR>(skew(rj) * gFx);
    }
  }

  void compute_rigid_body_gradient(int mode) {  // :1226-1279
    if (!P.computeGradient) return;
    for (size_t bi = 0; bi < bodies.size(); ++bi) {
      auto& b = bodies[bi];
      if (!(b.dynamic && !b.animated)) continue;
#pragma omp parallel for schedule(static)
      for (long j = 0; j < long(b.x.size()); ++j)
        for (uint32_t i : nBF[bi][j]) gradient_pair(b, long(i), uint32_t(j), mode);
    }
  }

  // One (boundary sample j, fluid particle i) term of computeRigidBodyGradient
  // (:1255-1275): k_i from the current b field, then computeGradient.
  void gradient_pair(Body<RF, RR>& b, long i, uint32_t j, int mode) {
    const RF invH = RF(1.0) / dt;
    const RF invH2 = RF(1.0) / dt / dt;
    const RF b_i = (mode == 0) ? bAdv[i] - RF(1.0) : bAdv[i];
    // quirk: the coefficient is truncated to an unsigned int (:1266)
    const unsigned int coeff = (mode == 0) ? (unsigned int)invH2 : (unsigned int)invH;
    const RF ki = b_i * alpha[i] * RF(coeff);
    const V3<RF> grad_p_j = K.gradW(x[i] - b.x[j]) * (-b.V[j]);
    const V3<RF> velChange = grad_p_j * (-dt * RF(1.0) * ki);
    const V3<RF> force = velChange * (-m_f) * invH;
    compute_gradient(b, i, j, b_i, ki, force, mode);
  }

  // BoundaryModel_Akinci2012::accumulate_and_reset_gradient, :453-497.
  void accumulate_and_reset_gradient(Body<RF, RR>& b) {
    b.nFv.setZero(); b.nFx.setZero(); b.nFw.setZero(); b.nFq.setZero();
    b.nTw.setZero(); b.nTq.setZero(); b.nTv.setZero(); b.nTx.setZero();
    for (size_t j = 0; j < b.x.size(); ++j) {
      b.nFv += b.gFv[j]; b.nFx += b.gFx[j]; b.nFw += b.gFw[j]; b.nFq += b.gFq[j];
      b.nTw += b.gTw[j]; b.nTq += b.gTq[j]; b.nTv += b.gTv[j]; b.nTx += b.gTx[j];
      b.gFv[j].setZero(); b.gFx[j].setZero(); b.gFw[j].setZero(); b.gFq[j].setZero();
      b.gTw[j].setZero(); b.gTv[j].setZero(); b.gTx[j].setZero(); b.gTq[j].setZero();
    }
  }

  void force_and_torque(const Body<RF, RR>& b, V3<RR>& F, V3<RR>& T) const {  // BoundaryModel.cpp:38-53
    F = V3<RR>(); T = V3<RR>();
    for (int t = 0; t < nthreads; ++t) { F += b.Fth[t]; T += b.Tth[t]; }
  }

  // compute_grad_inertia_v_to_{omega0,v0}, BoundaryModel_Akinci2012.cpp:524-541
  // (the manager's, RigidBodyGradientManager.cpp:263-279, is the same).
  static M3<RR> grad_inertia_v(const Body<RF, RR>& b, const V3<RR>& vv, const M43<RR>& Jq) {
    const M3<RR> R = b.q.toRotationMatrix();
    const M3<RR> a = grad_Rqp_to_q(b.q, b.I0 * (R.transpose() * vv)) * Jq;
    const M3<RR> c = R * b.I0 * (grad_RqTp_to_q(b.q, vv) * Jq);
    return a + c;
  }

  // BoundaryModel_Akinci2012::perform_chain_rule, :543-889.
  void perform_chain_rule(Body<RF, RR>& b, RR dtR) {
    const int mode = P.gradientMode;
    const RR invMass = RR(1.0) / (b.mass + RR(1e-10));
    const M3<RR> I = b.I, Iinv = b.Iinv;
    const V3<RR> om = b.wrb;
    V3<RR> temp_v;
    if (mode == 0 || mode == 1) {
      M3<RR> dF_v0, dF_w0;
      if (mode == 0) {
        dF_v0 = b.nFx * b.Jxv + b.nFv * b.Jvv + b.nFq * b.Jqv + b.nFw * b.Jwv;
        dF_w0 = b.nFx * b.Jxw + b.nFv * b.Jvw + b.nFq * b.Jqw + b.nFw * b.Jww;
      } else {
        dF_v0 = b.nFv * b.Jvv + b.nFw * b.Jwv;
        dF_w0 = b.nFv * b.Jvw + b.nFw * b.Jww;
      }
      b.Jvv = b.Jvv + dF_v0 * (dtR * invMass);
      b.Jvw = b.Jvw + dF_w0 * (dtR * invMass);
      b.Jxv = b.Jxv + b.Jvv * dtR;
      b.Jxw = b.Jxw + b.Jvw * dtR;
      V3<RR> F, T;
      force_and_torque(b, F, T);
      M3<RR> dT_w0, dT_v0;
      if (mode == 0) {
        dT_w0 = b.nTx * b.Jxw + b.nTv * b.Jvw + b.nTq * b.Jqw + b.nTw * b.Jww;
        dT_v0 = b.nTx * b.Jxv + b.nTv * b.Jvv + b.nTq * b.Jqv + b.nTw * b.Jwv;
      } else {
        dT_w0 = b.nTv * b.Jvw + b.nTw * b.Jww;
        dT_v0 = b.nTv * b.Jvv + b.nTw * b.Jwv;
      }
      M3<RR> dTau_w0, dTau_v0;
      if (P.gyroscopic) {
        temp_v = (I * om).cross(om) + T;
        const M3<RR> dLw_w0 = skew(I * om) * b.Jww + skew(om).transpose() * (I * b.Jww + grad_inertia_v(b, om, b.Jqw));
        const M3<RR> dLw_v0 = skew(I * om) * b.Jwv + skew(om).transpose() * (I * b.Jwv + grad_inertia_v(b, om, b.Jqv));
        dTau_w0 = dLw_w0 + dT_w0;
        dTau_v0 = dLw_v0 + dT_v0;
      } else {
        temp_v = T;
        dTau_w0 = dT_w0;
        dTau_v0 = dT_v0;
      }
      // quirk: "+" on the d(I^-1) term (:641-647)
      const M3<RR> gIi_w0 = Iinv * grad_inertia_v(b, Iinv * temp_v, b.Jqw);
      const M3<RR> gIi_v0 = Iinv * grad_inertia_v(b, Iinv * temp_v, b.Jqv);
      b.Jww = b.Jww + (gIi_w0 + Iinv * dTau_w0) * dtR;
      b.Jwv = b.Jwv + (gIi_v0 + Iinv * dTau_v0) * dtR;
    } else {  // RigidGradOnly, :674-709
      b.Jxv = b.Jxv + b.Jvv * dtR;
      b.Jxw = b.Jxw + b.Jvw * dtR;
      if (P.gyroscopic) {
        temp_v = (I * om).cross(om);
        const M3<RR> gIi_w0 = Iinv * grad_inertia_v(b, Iinv * temp_v, b.Jqw);
        const M3<RR> gIi_v0 = Iinv * grad_inertia_v(b, Iinv * temp_v, b.Jqv);
        const M3<RR> dLw_w0 = skew(I * om) * b.Jww + skew(om).transpose() * (I * b.Jww + grad_inertia_v(b, om, b.Jqw));
        const M3<RR> dLw_v0 = skew(I * om) * b.Jwv + skew(om).transpose() * (I * b.Jwv + grad_inertia_v(b, om, b.Jqv));
        b.Jww = b.Jww + (gIi_w0 + Iinv * dLw_w0) * dtR;
        b.Jwv = b.Jwv + (gIi_v0 + Iinv * dLw_v0) * dtR;
      }
      // temp_v stays zero without the gyroscopic term
    }
    if (P.optimizeRotation) {  // :827-871
      const Quat<RR> q = b.q;
      const V3<RR> nw = om + (Iinv * temp_v) * dtR;
      const Quat<RR> p(0, nw.x, nw.y, nw.z);
      Quat<RR> nq = q;
      const Quat<RR> pq = p * nq;
      const RR hh = dtR * RR(0.5);
      nq.x += hh * pq.x; nq.y += hh * pq.y; nq.z += hh * pq.z; nq.w += hh * pq.w;
      const Quat<RR> nqn = nq.normalized();
      const RR qv[4] = {nqn.w, nqn.x, nqn.y, nqn.z};
      M4<RR> Nq;
      const RR nn = nq.norm();
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) Nq(r, c) = ((r == c ? RR(1) : RR(0)) - qv[r] * qv[c]) / nn;
      const M4<RR> Pq = grad_p_q_product_to_q(p);
      const M43<RR> Pw = grad_omega_q_product_to_omega(q);
      b.Jqw = Nq * (b.Jqw + (Pw * b.Jww + Pq * b.Jqw) * (dtR / RR(2.)));
      b.Jqv = Nq * (b.Jqv + (Pw * b.Jwv + Pq * b.Jqv) * (dtR / RR(2.)));
    }
  }

  // TimeStepDiffDFSPH::backwardPerStep, :493-524.
  void backward_per_step() {
    if (!P.computeGradient) return;
    for (auto& b : bodies)
      if (b.dynamic && !b.animated) {
        accumulate_and_reset_gradient(b);
        if (P.useGradientManager) {
          hand_to_manager(b);
          if (multi_manager()) {  // the (R, R) entry of the manager's table
            const size_t nb = bodies.size(), R = size_t(&b - &bodies[0]);
            MgrN& n = MN[R * nb + R];
            n.Fv = b.nFv; n.Fx = b.nFx; n.Fw = b.nFw; n.Fq = b.nFq;
            n.Tv = b.nTv; n.Tx = b.nTx; n.Tw = b.nTw; n.Tq = b.nTq;
          }
        } else {
          perform_chain_rule(b, RR(dt));
        }
      }
  }

  // BoundaryModel_Akinci2012::update_rigid_body_gradient_manager, :952-967.
  static void hand_to_manager(Body<RF, RR>& b) {
    b.mFv = b.nFv; b.mFx = b.nFx; b.mFw = b.nFw; b.mFq = b.nFq;
    b.mTv = b.nTv; b.mTx = b.nTx; b.mTw = b.nTw; b.mTq = b.nTq;
  }

  // RigidBodyGradientManager (entries R == RR; no rigid contact solver), as
  // SimulatorBase::timeStepNoGUI runs it (SimulatorBase.cpp:1159-1166) for every
  // dynamic body: perform_force_and_torque_chain_rule (:96-156), whose
  // d(F, T)/d(v0, omega0) use the v and omega sums and the Jacobians before this
  // step's update, then perform_velocity_chain_rule (:276-365) ...
  void manager_velocity_chain(Body<RF, RR>& b) {
    const RR dtR = RR(dt);
    const M3<RR> dF_v0 = b.mFv * b.mJvv + b.mFw * b.mJwv;
    const M3<RR> dF_w0 = b.mFv * b.mJvw + b.mFw * b.mJww;
    const M3<RR> dT_v0 = b.mTv * b.mJvv + b.mTw * b.mJwv;
    const M3<RR> dT_w0 = b.mTv * b.mJvw + b.mTw * b.mJww;
    b.mJvv = b.mJvv + dF_v0 * (dtR * b.invMass);
    b.mJvw = b.mJvw + dF_w0 * (dtR * b.invMass);
    const M3<RR> I = b.I, Iinv = b.Iinv;
    const V3<RR> om = b.wrb;
    V3<RR> F, T;
    force_and_torque(b, F, T);
    V3<RR> temp_v = T;
    M3<RR> dTau_w0 = dT_w0, dTau_v0 = dT_v0;
    if (P.gyroscopic) {
      temp_v = (I * om).cross(om) + T;
      dTau_w0 = skew(I * om) * b.mJww + skew(om).transpose() * (I * b.mJww + grad_inertia_v(b, om, b.mJqw)) + dT_w0;
      dTau_v0 = skew(I * om) * b.mJwv + skew(om).transpose() * (I * b.mJwv + grad_inertia_v(b, om, b.mJqv)) + dT_v0;
    }
    // the same "+" on the d(I^-1) term as perform_chain_rule (:351-357)
    const M3<RR> gIi_w0 = Iinv * grad_inertia_v(b, Iinv * temp_v, b.mJqw);
    const M3<RR> gIi_v0 = Iinv * grad_inertia_v(b, Iinv * temp_v, b.mJqv);
    b.mJww = b.mJww + (gIi_w0 + Iinv * dTau_w0) * dtR;
    b.mJwv = b.mJwv + (gIi_v0 + Iinv * dTau_v0) * dtR;
  }

  // ... and, after the velocity update (the new omega) and before the position
  // update (the current q), perform_position_and_rotation_chain_rule (:367-423).
  void manager_position_chain(Body<RF, RR>& b) {
    const RR dtR = RR(dt);
    b.mJxv = b.mJxv + b.mJvv * dtR;
    b.mJxw = b.mJxw + b.mJvw * dtR;
    const Quat<RR> q = b.q;
    const Quat<RR> p(0, b.wrb.x, b.wrb.y, b.wrb.z);
    Quat<RR> nq = q;
    const Quat<RR> pq = p * nq;
    const RR hh = dtR * RR(0.5);
    nq.x += hh * pq.x; nq.y += hh * pq.y; nq.z += hh * pq.z; nq.w += hh * pq.w;
    const Quat<RR> nqn = nq.normalized();
    const RR qv[4] = {nqn.w, nqn.x, nqn.y, nqn.z};
    M4<RR> Nq;
    const RR nn = nq.norm();
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) Nq(r, c) = ((r == c ? RR(1) : RR(0)) - qv[r] * qv[c]) / nn;
    const M4<RR> Pq = grad_p_q_product_to_q(p);
    const M43<RR> Pw = grad_omega_q_product_to_omega(q);
    b.mJqv = Nq * (b.mJqv + (Pw * b.mJwv + Pq * b.mJqv) * (dtR / RR(2.)));
    b.mJqw = Nq * (b.mJqw + (Pw * b.mJww + Pq * b.mJqw) * (dtR / RR(2.)));
  }

  // Test hook that pins the contact solver's gradient terms by finite
  // differences; not part of the reference's step.  At the first step's state
  // (positions as they are, body velocities v0 + g dt and omega0), the contact
  // samples of body `obs` (a neighbour on another body, density ratio > 1), with
  // their n, |f_n| and weights held, and body `pert` moved by dx, its raw
  // quaternion by dq (not renormalised: the rotation then in the homogeneous form
  // the reference differentiates), its v by dv and omega by dw:
  //   out[0:3] = sum -beta n density,  out[3:6] = sum r x (-beta n density) (r held),
  //   out[6:9] = sum -mu |f_n| normalized(v_rel).
  int contact_probe(int obs, int pert, const double* dx, const double* dq, const double* dv, const double* dw,
                    double* out) const {
    const size_t nb = bodies.size();
    if (obs < 0 || pert < 0 || size_t(obs) >= nb || size_t(pert) >= nb) return -1;
    const V3<RR> gR(RR(P.gravity[0]), RR(P.gravity[1]), RR(P.gravity[2]));
    const RR dtR = RR(dt);
    auto state = [&](bool perturbed, std::vector<std::vector<V3<RF>>>& X, std::vector<std::vector<V3<RF>>>& Vv) {
      X.assign(nb, {}); Vv.assign(nb, {});
      for (size_t k = 0; k < nb; ++k) {
        const auto& b = bodies[k];
        Quat<RR> q = b.q;
        V3<RR> xr = b.xrb, v = b.v0 + gR * dtR, w = b.w0;
        if (!b.dynamic) { v = V3<RR>(); w = V3<RR>(); }
        if (perturbed && int(k) == pert) {
          q.w += dq[0]; q.x += dq[1]; q.y += dq[2]; q.z += dq[3];
          xr += V3<RR>(dx[0], dx[1], dx[2]);
          v += V3<RR>(dv[0], dv[1], dv[2]);
          w += V3<RR>(dw[0], dw[1], dw[2]);
        }
        X[k].resize(b.x0.size()); Vv[k].resize(b.x0.size());
        for (size_t j = 0; j < b.x0.size(); ++j) {
          // the rotation as the homogeneous form (w^2 - |u|^2) p + 2 (u.p) u + 2 w u x p,
          // which get_grad_Rqp_to_q differentiates (GradientUtils.cpp:9-20); equal to
          // the rotation on unit quaternions
          const V3<RR> p0(b.x0[j]), u = q.vec();
          const V3<RR> rp = perturbed && int(k) == pert
                                ? p0 * (q.w * q.w - u.dot(u)) + u * (RR(2) * u.dot(p0)) + u.cross(p0) * (RR(2) * q.w)
                                : q.rotate(p0);
          const V3<RR> xj = rp + xr;
          X[k][j] = V3<RF>(xj);
          Vv[k][j] = V3<RF>(v + w.cross(xj - xr));
        }
      }
    };
    std::vector<std::vector<V3<RF>>> X0, V0, X1, V1;
    state(false, X0, V0);
    state(true, X1, V1);
    std::vector<V3<RF>> pts;
    std::vector<std::pair<uint32_t, uint32_t>> id;
    for (size_t k = 0; k < nb; ++k)
      for (size_t j = 0; j < X0[k].size(); ++j) { pts.push_back(X0[k][j]); id.push_back({uint32_t(k), uint32_t(j)}); }
    Grid<RF> G;
    G.build(pts, h * RF(1.001));
    const RF W0 = KRC.w0, beta = RF(P.rigidContactBeta), mu = RF(P.rigidContactFriction);
    const auto& b = bodies[obs];
    V3<RR> S, ST, SF;
    std::vector<uint32_t> nbl;
    for (size_t a = 0; a < pts.size(); ++a) {
      if (id[a].first != uint32_t(obs)) continue;
      const uint32_t r = id[a].second;
      G.query(pts[a], pts, h * h, int64_t(a), nbl);
      bool contact = false;
      for (uint32_t k : nbl) contact |= id[k].first != uint32_t(obs);
      if (!contact) continue;
      RF dens = b.cvol0[r] * W0, sw = 0;
      V3<RF> sx, sv;
      for (uint32_t k : nbl) {
        const RF w = KRC.W(pts[a] - pts[k]);
        dens += bodies[id[k].first].cvol0[id[k].second] * w;
        if (id[k].first == uint32_t(obs)) { sx += pts[k] * w; sw += w; }
        else sv += V0[id[k].first][id[k].second] * w;
      }
      const RF ratio = dens / b.cdens0[r];
      if (!(ratio > RF(1))) continue;
      const V3<RF> n = pts[a] - sx * (RF(1) / sw);
      const RF fnn = (n * (-beta * (ratio - RF(1)))).norm();
      // perturbed density and relative velocity, weights of the unperturbed state
      const V3<RF> xr1 = X1[obs][r];
      RF dens1 = b.cvol0[r] * W0;
      V3<RF> sv1;
      for (uint32_t k : nbl) {
        const RF vk = bodies[id[k].first].cvol0[id[k].second];
        if (id[k].first == uint32_t(obs)) {
          dens1 += vk * KRC.W(pts[a] - pts[k]);  // the own body's part is held: it is invariant under the body's motion
        } else {
          dens1 += vk * KRC.W(xr1 - X1[id[k].first][id[k].second]);
          sv1 += V1[id[k].first][id[k].second] * KRC.W(pts[a] - pts[k]);
        }
      }
      const V3<RR> fd = V3<RR>(n) * (-RR(beta) * RR(dens1));
      S += fd;
      ST += (V3<RR>(pts[a]) - b.xrb).cross(fd);
      const V3<RF> vrel = V1[obs][r] - sv1 * (RF(1) / sw);
      const RF vn = vrel.norm();
      SF += V3<RR>(vrel * (-mu * fnn / vn));
    }
    for (int c = 0; c < 3; ++c) { out[c] = double(S[c]); out[3 + c] = double(ST[c]); out[6 + c] = double(SF[c]); }
    return 0;
  }

  // ---- RigidBodyGradientManager across bodies (with the rigid contact solver) ----
  // perform_force_and_torque_chain_rule (RigidBodyGradientManager.cpp:96-156: the v
  // and omega terms) or, after the contact solver, _Rigid (:158-239: the x terms too,
  // and the q terms of the omega0 sums; the q terms written after the v0 sums,
  // :195-196 and :200-201, are statements of their own and do not enter them), over
  // every dynamic R, RR and Rk; then perform_velocity_chain_rule (:276-365) with
  // the body's current force and torque (cleared once updateBoundaryForces ran).
  struct MgrD { M3<RR> Fv0, Fw0, Tv0, Tw0; };
  void multi_velocity_chain(bool rigid) {
    const size_t nb = bodies.size();
    const RR dtR = RR(dt);
    std::vector<MgrD> d(nb * nb);
    for (size_t R = 0; R < nb; ++R) {
      if (!bodies[R].dynamic) continue;
      for (size_t Q = 0; Q < nb; ++Q) {
        if (!bodies[Q].dynamic) continue;
        MgrD& o = d[R * nb + Q];
        for (size_t K = 0; K < nb; ++K) {
          if (!bodies[K].dynamic) continue;
          const MgrN& n = MN[R * nb + K];
          const MgrJ& J = MJ[K * nb + Q];
          if (rigid) {
            o.Fv0 += n.Fv * J.vv + n.Fw * J.wv + n.Fx * J.xv;
            o.Tv0 += n.Tv * J.vv + n.Tw * J.wv + n.Tx * J.xv;
            o.Fw0 += n.Fv * J.vw + n.Fw * J.ww + n.Fx * J.xw + n.Fq * J.qw;
            o.Tw0 += n.Tv * J.vw + n.Tw * J.ww + n.Tx * J.xw + n.Tq * J.qw;
          } else {
            o.Fv0 += n.Fv * J.vv + n.Fw * J.wv;
            o.Tv0 += n.Tv * J.vv + n.Tw * J.wv;
            o.Fw0 += n.Fv * J.vw + n.Fw * J.ww;
            o.Tw0 += n.Tv * J.vw + n.Tw * J.ww;
          }
        }
      }
    }
    for (size_t R = 0; R < nb; ++R) {
      auto& b = bodies[R];
      if (!b.dynamic) continue;
      const M3<RR> I = b.I, Iinv = b.Iinv;
      const V3<RR> om = b.wrb;
      V3<RR> F, T;
      force_and_torque(b, F, T);
      const V3<RR> temp_v = P.gyroscopic ? (I * om).cross(om) + T : T;
      for (size_t Q = 0; Q < nb; ++Q) {
        if (!bodies[Q].dynamic) continue;
        MgrJ& J = MJ[R * nb + Q];
        const MgrD& o = d[R * nb + Q];
        J.vv = J.vv + o.Fv0 * (dtR * b.invMass);
        J.vw = J.vw + o.Fw0 * (dtR * b.invMass);
        M3<RR> dTau_w0 = o.Tw0, dTau_v0 = o.Tv0;
        if (P.gyroscopic) {
          dTau_w0 = skew(I * om) * J.ww + skew(om).transpose() * (I * J.ww + grad_inertia_v(b, om, J.qw)) + o.Tw0;
          dTau_v0 = skew(I * om) * J.wv + skew(om).transpose() * (I * J.wv + grad_inertia_v(b, om, J.qv)) + o.Tv0;
        }
        const M3<RR> gIi_w0 = Iinv * grad_inertia_v(b, Iinv * temp_v, J.qw);
        const M3<RR> gIi_v0 = Iinv * grad_inertia_v(b, Iinv * temp_v, J.qv);
        J.ww = J.ww + (gIi_w0 + Iinv * dTau_w0) * dtR;
        J.wv = J.wv + (gIi_v0 + Iinv * dTau_v0) * dtR;
      }
    }
  }
  // perform_position_and_rotation_chain_rule (:367-423) for every dynamic pair.
  void multi_position_chain() {
    const size_t nb = bodies.size();
    const RR dtR = RR(dt);
    for (size_t R = 0; R < nb; ++R) {
      auto& b = bodies[R];
      if (!b.dynamic) continue;
      const Quat<RR> q = b.q;
      const Quat<RR> p(0, b.wrb.x, b.wrb.y, b.wrb.z);
      Quat<RR> nq = q;
      const Quat<RR> pq = p * nq;
      const RR hh = dtR * RR(0.5);
      nq.x += hh * pq.x; nq.y += hh * pq.y; nq.z += hh * pq.z; nq.w += hh * pq.w;
      const Quat<RR> nqn = nq.normalized();
      const RR qv[4] = {nqn.w, nqn.x, nqn.y, nqn.z};
      M4<RR> Nq;
      const RR nn = nq.norm();
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) Nq(r, c) = ((r == c ? RR(1) : RR(0)) - qv[r] * qv[c]) / nn;
      const M4<RR> Pq = grad_p_q_product_to_q(p);
      const M43<RR> Pw = grad_omega_q_product_to_omega(q);
      for (size_t Q = 0; Q < nb; ++Q) {
        if (!bodies[Q].dynamic) continue;
        MgrJ& J = MJ[R * nb + Q];
        J.xv = J.xv + J.vv * dtR;
        J.xw = J.xw + J.vw * dtR;
        J.qv = Nq * (J.qv + (Pw * J.wv + Pq * J.qv) * (dtR / RR(2.)));
        J.qw = Nq * (J.qw + (Pw * J.ww + Pq * J.qw) * (dtR / RR(2.)));
      }
    }
  }

  // RigidBody3dBoundarySimulator::velocityTimeStep (:278-320) with
  // BoundarySimulator::updateBoundaryForces (BoundarySimulator.cpp:10-36) and
  // Dynamic3dRigidBody::addForce/addTorque/addGravity (Dynamic3dRigidBody.h:113-142, 228-233).
  void rigid_velocity_step() {
    const RR dtR = RR(dt);
    if (P.useRigidContact) { contact_velocity_step(dtR); return; }
    const V3<RR> gR(RR(P.gravity[0]), RR(P.gravity[1]), RR(P.gravity[2]));
    for (size_t k = 0; k < bodies.size(); ++k) {
      auto& b = bodies[k];
      if (!b.dynamic) continue;
      V3<RR> F, T;
      force_and_torque(b, F, T);
      // the force-model hook's frozen-inverse map (see ForceModel); I^-1 otherwise
      const M3<RR> Mi = (fm.useFrozen && int(k) == fm.body) ? fm.frozen.at(stepCount - 1) * RR(2) - b.Iinv : b.Iinv;
      if (!b.animated) {
        b.vrb += F * (b.invMass * dtR);
        if (P.gyroscopic) {
          const V3<RR> L = b.I * b.wrb;
          b.wrb += (Mi * (L.cross(b.wrb) + T)) * dtR;
        } else {
          b.wrb += (Mi * T) * dtR;
        }
      }
      b.F_last = F; b.T_last = T;
      for (int t = 0; t < nthreads; ++t) { b.Fth[t] = V3<RR>(); b.Tth[t] = V3<RR>(); }
      if (!b.animated) b.vrb += (gR * b.mass) * (b.invMass * dtR);
    }
  }

  // RigidBody3dBoundarySimulator::positionTimeStep (:322-344) with
  // Dynamic3dRigidBody::animate (Dynamic3dRigidBody.h:143-154).
  void rigid_position_step() {
    const RR dtR = RR(dt);
    for (auto& b : bodies) {
      if (!b.dynamic) continue;
      b.xrb += b.vrb * dtR;
      const Quat<RR> wq(0, b.wrb.x, b.wrb.y, b.wrb.z);
      const Quat<RR> d = wq * b.q;
      const RR hh = dtR * RR(0.5);
      b.q.x += hh * d.x; b.q.y += hh * d.y; b.q.z += hh * d.z; b.q.w += hh * d.w;
      b.q.normalize();
      update_inertia(b);
    }
    update_boundary_particles(false);
  }

  // clearAccelerations (a = g) + Simulation::computeNonPressureForces
  // (Simulation.cpp:618-631: surface tension, then viscosity) at the step's
  // densities and the velocities after the divergence solve.  Boundary
  // viscosity (which would push the bodies) is not part of this restatement.
  void compute_nonpressure() {
    const long N = long(x.size());
    accel.assign(N, g);
    if (P.surfaceTensionMethod == 2) surface_tension_akinci2013();
    if (P.surfaceTensionMethod == 3) surface_tension_he2014();
    if (P.viscosityMethod == 1) viscosity_standard();
    if (P.viscosityMethod == 2) viscosity_xsph();
  }
  // Viscosity_Standard::step (Viscosity_Standard.cpp:233-269), fluid part (3D: d = 10):
  // a_i += d nu (m_j / rho_j) ((v_i - v_j) . r_ij) / (|r_ij|^2 + 0.01 h^2) gradW_ij,
  // the scalar formed in double as the reference's 0.01 literal makes it.
  void viscosity_standard() {
    const long N = long(x.size());
    const RF d = RF(10.0), nu = RF(P.viscosity);
    const RF h2 = h * h;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      V3<RF>& ai = accel[i];
      for (uint32_t j : nFF[i]) {
        const V3<RF> xixj = x[i] - x[j];
        const double s = double(d * nu * (m_f / rho[j]) * (v[i] - v[j]).dot(xixj)) /
                         (double(xixj.squaredNorm()) + 0.01 * double(h2));
        ai += K.gradW(xixj) * RF(s);
      }
    }
  }
  // SurfaceTension_Akinci2013::computeNormals / step (SurfaceTension_Akinci2013.cpp:26-147).
  void surface_tension_akinci2013() {
    const long N = long(x.size());
    normal.assign(N, V3<RF>());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      V3<RF> ni;
      for (uint32_t j : nFF[i]) ni += K.gradW(x[i] - x[j]) * (m_f / rho[j]);
      normal[i] = ni * h;
    }
    const RF k = RF(P.surfaceTension), kb = RF(P.surfaceTensionBoundary);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      V3<RF>& ai = accel[i];
      for (uint32_t j : nFF[i]) {
        const RF K_ij = RF(2.0) * rho0 / (rho[i] + rho[j]);
        V3<RF> a;
        V3<RF> xixj = x[i] - x[j];
        const RF length2 = xixj.squaredNorm();
        if (length2 > RF(1.0e-9)) {
          xixj = xixj * (RF(1.0) / std::sqrt(length2));
          a -= xixj * (k * m_f) * KC.W(x[i] - x[j]);
        }
        a -= (normal[i] - normal[j]) * k;
        ai += a * K_ij;
      }
      for (size_t bi = 0; bi < bodies.size(); ++bi)
        for (uint32_t j : nFB[bi][i]) {
          V3<RF> xixj = x[i] - bodies[bi].x[j];
          const RF length2 = xixj.squaredNorm();
          if (length2 > RF(1.0e-9)) {
            xixj = xixj * (RF(1.0) / std::sqrt(length2));
            ai -= xixj * (kb * rho0 * bodies[bi].V[j]) * KA.W(x[i] - bodies[bi].x[j]);
          }
        }
    }
  }
  // SurfaceTension_He2014::step (SurfaceTension_He2014.cpp:29-160) with Akinci
  // 2012 boundaries: colour c_i = (m / rho_i) W(0) + sum_j (m / rho_j) W_ij +
  // sum_b V_b W_ib; its gradient (1 / c_i) sum_j (m / rho_j) c_j gradW_ij and
  // gradC2_i = |.|^2; then a_i += (k / 4 rho_i) sum_j (m / rho_j) (gradC2_i +
  // gradC2_j) gradW_ij + (kb / 4 rho_i) sum_b V_b gradC2_i gradW_ib, the body
  // receiving -m a of each boundary term (BoundaryModel::addForce).
  void surface_tension_he2014() {
    const long N = long(x.size());
    color.assign(N, RF(0));
    gradC2.assign(N, RF(0));
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      RF ci = m_f / rho[i] * K.w0;
      for (uint32_t j : nFF[i]) ci += m_f / rho[j] * K.W(x[i] - x[j]);
      for (size_t bi = 0; bi < bodies.size(); ++bi)
        for (uint32_t j : nFB[bi][i]) ci += bodies[bi].V[j] * K.W(x[i] - bodies[bi].x[j]);
      color[i] = ci;
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      V3<RF> g;
      for (uint32_t j : nFF[i]) g += K.gradW(x[i] - x[j]) * (m_f / rho[j] * color[j]);
      g = g * (RF(1.0) / color[i]);
      gradC2[i] = g.squaredNorm();
    }
    const RF k = RF(P.surfaceTension), kb = RF(P.surfaceTensionBoundary);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      V3<RF>& ai = accel[i];
      const RF factor = RF(0.25) * k / rho[i], factorb = RF(0.25) * kb / rho[i];
      for (uint32_t j : nFF[i])
        ai += K.gradW(x[i] - x[j]) * (factor * m_f / rho[j] * (gradC2[i] + gradC2[j]));
      for (size_t bi = 0; bi < bodies.size(); ++bi) {
        auto& b = bodies[bi];
        for (uint32_t j : nFB[bi][i]) {
          const V3<RF> a = K.gradW(x[i] - b.x[j]) * (factorb * b.V[j] * gradC2[i]);
          ai += a;
          add_force(b, b.x[j], a * (-m_f));
        }
      }
    }
  }
  // Viscosity_XSPH::step (Viscosity_XSPH.cpp:35-108), fluid part.
  void viscosity_xsph() {
    const long N = long(x.size());
    const RF invH = RF(1.0) / dt;
    const RF nu = RF(P.viscosity);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      V3<RF>& ai = accel[i];
      for (uint32_t j : nFF[i]) ai -= (v[i] - v[j]) * (invH * nu * (m_f / rho[j])) * K.W(x[i] - x[j]);
    }
  }
  // Simulation::updateTimeStepSizeCFL (Simulation.cpp:544-616): the squared
  // speed bound starts at 0.1, fluid particles count with v + a h, samples of
  // dynamic bodies with their velocity.  Returns the new step size.
  RF cfl_step_size() const {
    RF maxVel = RF(0.1);
    const RF diameter = RF(2.0) * RF(P.radius);
    for (size_t i = 0; i < x.size(); ++i) {
      const RF velMag = (v[i] + accel[i] * dt).squaredNorm();
      if (velMag > maxVel) maxVel = velMag;
    }
    for (const auto& b : bodies)
      if (b.dynamic || b.animated)
        for (const auto& vb : b.v) {
          const RF velMag = vb.squaredNorm();
          if (velMag > maxVel) maxVel = velMag;
        }
    RF hn = RF(P.cflFactor) * RF(0.4) * (diameter / std::sqrt(maxVel));
    hn = std::min(hn, RF(P.cflMaxTimeStepSize));
    hn = std::max(hn, RF(P.cflMinTimeStepSize));
    return hn;
  }
  bool nonpressure_on() const { return P.surfaceTensionMethod == 2 || P.surfaceTensionMethod == 3 || P.viscosityMethod != 0 || P.cflMethod == 1; }

  // One full step in SimulatorBase::timeStepNoGUI order (SimulatorBase.cpp:1142-1200)
  // around TimeStepDiffDFSPH::step (TimeStepDiffDFSPH.cpp:526-652).
  void step() {
    tp = nullptr;
    if (record) {
      tapes.emplace_back();
      tp = &tapes.back();
      tp->bodies.resize(bodies.size());
      for (size_t k = 0; k < bodies.size(); ++k) {
        auto& t = tp->bodies[k];
        const auto& b = bodies[k];
        t.xb = b.x; t.vb = b.v; t.vpre = b.vrb; t.wpre = b.wrb;
      }
    }
    begin_step();
    if (tp)
      for (size_t k = 0; k < bodies.size(); ++k) {
        auto& t = tp->bodies[k];
        const auto& b = bodies[k];
        t.xrb = b.xrb; t.q = b.q; t.v = b.vrb; t.w = b.wrb; t.I = b.I; t.Iinv = b.Iinv; t.animated = b.animated;
      }
    neighborhood_search();
    compute_densities();
    compute_dfsph_factor();
    if (tp) { tp->x = x; tp->nFF = nFF; tp->nFB = nFB; tp->rho = rho; tp->alpha = alpha; tp->G = s; }
    if (net.enabled) {
      if (tp) { tp->kPrev = kappa; tp->kvPrev = kappaV; }
      learned_predict();
    }
    if (P.enableDivergenceSolver) divergence_solve();
    else stats.iterationsV = 0;
    if (tp) tp->vDiv = v;
    const long N = long(x.size());
    // the step size of the step's start (TimeStepDiffDFSPH.cpp:535) advects
    // the velocities and the positions and advances the time; the CFL update
    // (:584-586) sets the one the pressure solve, the rigid bodies and the
    // next step read
    const RF h_old = dt;
    if (nonpressure_on()) {
      compute_nonpressure();
      if (P.cflMethod == 1) dt = cfl_step_size();
#pragma omp parallel for schedule(static)
      for (long i = 0; i < N; ++i) v[i] += accel[i] * h_old;  // :588-605
    } else {
#pragma omp parallel for schedule(static)
      for (long i = 0; i < N; ++i) v[i] += g * dt;  // clearAccelerations + advect (:588-605)
    }
    pressure_solve();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) x[i] += v[i] * h_old;  // :613-634
    time = time + h_old;
    backward_per_step();
    const bool mgr = P.computeGradient && P.useGradientManager;
    if (multi_manager()) {
      // SimulatorBase::timeStepNoGUI (SimulatorBase.cpp:1159-1170): after the fluid
      // coupling (with fluid), the boundary velocity step, after the rigid coupling
      // (RigidBodyGradientManager.cpp:455-468)
      if (!x.empty()) multi_velocity_chain(false);
      rigid_velocity_step();
      multi_velocity_chain(true);
      multi_position_chain();
    } else {
      if (mgr) for (auto& b : bodies) if (b.dynamic) manager_velocity_chain(b);
      rigid_velocity_step();
      if (mgr) for (auto& b : bodies) if (b.dynamic) manager_position_chain(b);
    }
    if (tp) {
      for (size_t k = 0; k < bodies.size(); ++k) { tp->bodies[k].F = bodies[k].F_last; tp->bodies[k].T = bodies[k].T_last; }
      tp->Iv = stats.iterationsV; tp->Ip = stats.iterations;
    }
    rigid_position_step();
    tp = nullptr;
    ++fixPos;
  }

  // --------------------------------------------------------------------------
  // Test hook that pins the chain rule (perform_chain_rule) by finite
  // differences; not part of the reference's step.  A prescribed affine
  // force/torque law on body `body` replaces the fluid coupling (no fluid
  // step runs), from step `start` + 1 on (after the two-step release window):
  //   F = F0 + A (x, v, omega, q),
  //   T = T0 + B (x', v', omega, q),
  // columns x(3) v(3) omega(3) q(w,x,y,z)(4), with x' = x + dt v' and
  // v' = v + dt F / m the position and velocity after this step's force update:
  // the reference's torque chain reads J_x and J_v after their update
  // (BoundaryModel_Akinci2012.cpp:581-613), so its derivative is that of a
  // torque evaluated there (with the gradient manager, x' = x and v' = v).  The per-body sums grad_net_* are the law's
  // matrices (A -> F_x, F_v, F_w, F_q; B -> T_x, T_v, T_w, T_q).
  // useFrozen: the angular update uses (2 Ibar^-1 - I(q)^-1) with Ibar^-1 the
  // world inverse inertia of a base run at the same step (recordFrozen).  At
  // the base run's state it equals I^-1, and its derivative in q is
  // +I^-1 (dI) I^-1, the sign the reference's chain rule uses
  // (BoundaryModel_Akinci2012.cpp:641-647), so central differences of this map
  // are what the forward-mode chain accumulates.
  // --------------------------------------------------------------------------
  struct ForceModel {
    int body = -1, start = 2;
    double F0[3] = {0, 0, 0}, T0[3] = {0, 0, 0};
    double A[39] = {}, B[39] = {};  // 3 x 13, row-major
    bool recordFrozen = false, useFrozen = false;
    std::vector<M3<RR>> frozen;
  } fm;

  static V3<RR> affine(const double* c, const double* M, const V3<RR>& x, const V3<RR>& v, const V3<RR>& w,
                       const Quat<RR>& q) {
    const RR s[13] = {x.x, x.y, x.z, v.x, v.y, v.z, w.x, w.y, w.z, q.w, q.x, q.y, q.z};
    V3<RR> out{RR(c[0]), RR(c[1]), RR(c[2])};
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 13; ++k) out[r] += RR(M[13 * r + k]) * s[k];
    return out;
  }

  void model_step() {
    begin_step();
    time = time + dt;
    if (fm.recordFrozen) {
      if (int(fm.frozen.size()) < stepCount) fm.frozen.resize(stepCount);
      fm.frozen[stepCount - 1] = bodies[fm.body].Iinv;
    }
    for (size_t k = 0; k < bodies.size(); ++k) {
      auto& b = bodies[k];
      for (int t = 0; t < nthreads; ++t) { b.Fth[t] = V3<RR>(); b.Tth[t] = V3<RR>(); }
      if (int(k) != fm.body || stepCount <= fm.start) {
        b.nFv.setZero(); b.nFx.setZero(); b.nFw.setZero(); b.nFq.setZero();
        b.nTw.setZero(); b.nTq.setZero(); b.nTv.setZero(); b.nTx.setZero();
      } else {
        const V3<RR> F = affine(fm.F0, fm.A, b.xrb, b.vrb, b.wrb, b.q);
        const V3<RR> v1 = b.vrb + F * (b.invMass * RR(dt));
        const V3<RR> x1 = b.xrb + v1 * RR(dt);
        // the manager's torque chain reads J_x, J_v before their update: its
        // derivative is that of a torque evaluated at the current state
        const V3<RR> T = P.useGradientManager ? affine(fm.T0, fm.B, b.xrb, b.vrb, b.wrb, b.q)
                                              : affine(fm.T0, fm.B, x1, v1, b.wrb, b.q);
        b.Fth[0] = F; b.Tth[0] = T;
        for (int r = 0; r < 3; ++r) {
          for (int c = 0; c < 3; ++c) {
            b.nFx(r, c) = RR(fm.A[13 * r + c]); b.nFv(r, c) = RR(fm.A[13 * r + 3 + c]); b.nFw(r, c) = RR(fm.A[13 * r + 6 + c]);
            b.nTx(r, c) = RR(fm.B[13 * r + c]); b.nTv(r, c) = RR(fm.B[13 * r + 3 + c]); b.nTw(r, c) = RR(fm.B[13 * r + 6 + c]);
          }
          for (int c = 0; c < 4; ++c) { b.nFq(r, c) = RR(fm.A[13 * r + 9 + c]); b.nTq(r, c) = RR(fm.B[13 * r + 9 + c]); }
        }
      }
      if (P.computeGradient && b.dynamic && !b.animated) {
        if (P.useGradientManager) hand_to_manager(b);
        else perform_chain_rule(b, RR(dt));
      }
    }
    const bool mgr = P.computeGradient && P.useGradientManager;
    if (mgr) for (auto& b : bodies) if (b.dynamic) manager_velocity_chain(b);
    rigid_velocity_step();
    if (mgr) for (auto& b : bodies) if (b.dynamic) manager_position_chain(b);
    rigid_position_step();
  }

  // Reset of the rigid bodies and gradients between trajectories
  // (Dynamic3dRigidBody::reset, BoundaryModel_Akinci2012::reset).
  void reset_bodies() {
    for (auto& b : bodies) {
      b.xrb = b.x_init; b.q = b.q_init; b.vrb = V3<RR>(); b.wrb = V3<RR>();
      update_inertia(b);
      b.reset_jacobians();
    }
    reset_multi_manager();
    update_boundary_particles(true);
    time = 0; stepCount = 0;
  }
};

}  // namespace orc
