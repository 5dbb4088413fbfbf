// Context, host-side step orchestration and the C ABI of libdfr_hip.so.
//
// The step follows SimulatorBase::timeStepNoGUI (Simulator/SimulatorBase.cpp:1142-1200)
// around TimeStepDiffDFSPH::step (SPlisHSPlasH/DiffDFSPH/TimeStepDiffDFSPH.cpp:526-652):
// begin step (release window), neighbour search, density + factor, divergence
// solve, advection, pressure solve, position update, DiffFR's per-step chain
// rule, rigid velocity/position update and the boundary transform.  The host
// decides when a Jacobi solve has converged from the FP64 residual the
// recompute kernel leaves in device memory.
#include <algorithm>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dfr.h"
#include "dfr_kernels.h"

using namespace dfr;

namespace {

enum TimerClass {
  TC_SORT = 0, TC_NBR, TC_DENSITY, TC_VUPDATE, TC_RECOMPUTE, TC_PWARM, TC_ELEMWISE, TC_GRADIENT, TC_RIGID,
  TC_ADJ_UPDATE, TC_ADJ_DELTA, TC_ADJ_FACTOR, TC_ADJ_MISC, TC_CHECKPOINT, TC_MLP, TC_ADJ_MLP, TC_COUNT
};
const char* kTimerNames[TC_COUNT] = {"grid_sort", "neighbor_build", "density_factor", "velocity_update",
                                     "density_recompute", "pressure_warmstart", "advect_position",
                                     "gradient_pairs", "rigid_step", "adj_velocity_update",
                                     "adj_density_change", "adj_density_factor", "adj_rigid_misc",
                                     "checkpoint_copy", "mlp_warmstart", "adj_mlp_warmstart"};

struct BodyInfo {
  bool dynamic = false;
  int64_t n = 0;
  int64_t offset = 0;       // first sample in the concatenated boundary arrays
  int64_t dyn_first = -1;   // first index in dyn_idx (dynamic bodies)
  double x_rb[3] = {0, 0, 0}, q[4] = {1, 0, 0, 0}, density = 1000.0;
  double v0[3] = {0, 0, 0}, w0[3] = {0, 0, 0};
  std::vector<float> x0;    // local samples
  bool animated = false;
};

// State of the simulation at the start of a step (what dfr_step reads), in one
// device block: fluid in its current order, boundary samples, rigid bodies.
struct Snapshot {
  void* block = nullptr;
  float4 *pos = nullptr, *vel = nullptr, *bpos = nullptr, *bvel = nullptr;
  float *kappa = nullptr, *kappav = nullptr;
  uint32_t* id = nullptr;
  RigidDev* rig = nullptr;
  float time = 0.f;
  float dt = 0.f;  // step size the next step starts with (changes with the CFL rule)
  int64_t step_count = 0;
  std::vector<uint8_t> animated;
};

struct ArenaFull : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// A neighbour list outgrew its capacity (DFR_ERR_CAPACITY).
struct CapacityError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Bump allocator over large device chunks for per-step tapes (kept across
// rollouts, freed by dfr_adjoint_end / dfr_destroy).
struct Arena {
  std::vector<std::pair<char*, size_t>> chunks;
  size_t cur = 0, off = 0;
  size_t limit = 0;  // > 0: bytes the arena may hand out (dfr_adjoint_arena_limit)
  static constexpr size_t CHUNK = (size_t)2 << 30;
  void* alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (limit && used() + bytes > limit)
      throw ArenaFull("adjoint tape arena limit reached (dfr_adjoint_arena_limit)");
    while (cur < chunks.size() && chunks[cur].second - off < bytes) { ++cur; off = 0; }
    if (cur == chunks.size()) {
      const size_t sz = std::max(bytes, CHUNK);
      void* p = nullptr;
      const hipError_t e = hipMalloc(&p, sz);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        throw ArenaFull("out of device memory for the adjoint tape "
                        "(use checkpointing: dfr_adjoint_begin with checkpoint_every >= 1)");
      }
      chunks.push_back({static_cast<char*>(p), sz});
      off = 0;
    }
    void* p = chunks[cur].first + off;
    off += bytes;
    return p;
  }
  size_t used() const {
    size_t u = 0;
    for (size_t k = 0; k < cur && k < chunks.size(); ++k) u += chunks[k].second;
    return u + off;
  }
  void reset() { cur = 0; off = 0; }
  std::pair<size_t, size_t> mark() const { return {cur, off}; }
  void rewind(std::pair<size_t, size_t> m) { cur = m.first; off = m.second; }
  void release() {
    for (auto& ch : chunks) (void)hipFree(ch.first);
    chunks.clear();
    reset();
  }
};

// What the reverse pass of one step reads: the values that the step
// overwrote while it ran, and the step's structures (neighbour lists, grid,
// permutation, density, factor, sum_grad_p_k), which are the live buffers
// right after a replayed step or copies kept in the arena.
struct StepTapeDev {
  float4* x = nullptr;                    // fluid positions during the step (sorted order)
  float4* v0 = nullptr;                   // fluid velocities after the sort (learned warm-start feature)
  float4* vDiv = nullptr;                 // velocities after the divergence solve (the viscosity's input)
  float *kvPrev = nullptr, *d0 = nullptr, *kvW = nullptr;  // divergence warm start
  float *kPrev = nullptr, *rsW = nullptr, *kW = nullptr;   // pressure warm start
  float4 *bpos = nullptr, *bvel = nullptr;  // boundary samples at the start of the step
  RigidDev *rig_pre = nullptr, *rig_forced = nullptr;  // before / after the release-window override
  std::vector<float4*> vD, vP;            // velocity at the start of each Jacobi iteration
  std::vector<float*> kD, bD, kP, bP;     // k_i and d / rho* of each iteration
  int Iv = 0, Ip = 0;
  AdjStepTable tbl{};
  Arena* arena = nullptr;                 // set: per-step buffers come from the arena
  // structures of the step
  const uint32_t *fstart = nullptr, *perm = nullptr, *nbrB = nullptr, *cnt = nullptr;
  const uint4* nbr8 = nullptr;
  const uint4* thdr = nullptr;            // brick headers and halo slot maps
  const uint32_t* tmap = nullptr;
  const float *rho = nullptr, *alpha = nullptr;
  const float4* s = nullptr;
  const float2* aux = nullptr;
  const RigidDev* rig_post = nullptr;
};

// Learned kappa warm start: parameters, Adam state, captured training pairs.
struct MlpState {
  float *P = nullptr, *G = nullptr, *m = nullptr, *v = nullptr;  // FP32 parameters, gradient, Adam moments
  void* frag = nullptr;                    // bf16 MFMA fragments of the parameters
  MlpNorm norm{};
  bool has_params = false, norm_set = false, enabled = false;
  int adam_t = 0;
  float* kpred = nullptr;                  // [n] predicted kappa (stored units)
  float2* aux = nullptr;                   // [n] (velocity divergence, boundary density)
  // training pairs
  bool capturing = false;
  int64_t cap = 0;
  uint32_t threshold = 0;
  uint64_t salt = 0;
  float *feat = nullptr, *targ = nullptr;
  uint32_t *rows = nullptr, *count = nullptr;  // count[0] total, count[1] at the start of the step
  uint32_t* batch = nullptr;
  int batch_cap = 0;
  double* losses = nullptr;
  int losses_cap = 0;
};

// Exact adjoint of a rollout: checkpoints every k steps of the primal run and
// the reverse pass's working buffers.
struct AdjState {
  bool active = false;
  int k = 1;
  int64_t steps = 0;                       // primal steps since dfr_adjoint_begin
  std::vector<Snapshot> ckpt;              // state before steps 0, k, 2k, ... (k >= 1); blocks reused
  size_t nckpt = 0;                        // checkpoints of the current rollout
  std::vector<StepTapeDev> tapes;          // k == 0: every step's tape, kept in HBM
  Arena arena;
  // kept tapes: the steps write their structures straight into arena buffers;
  // the context's own buffers are swapped back in when the rollout ends
  struct Live {
    uint4* nbr8; uint4* thdr; uint32_t *tmap, *nbrB, *cnt, *fstart, *perm;
    float *rho, *alpha;
    float4* s;
    float2* aux;
  };
  bool swapped = false;
  Live own{};
  std::vector<Snapshot> seg;               // states inside the segment being reversed
  Snapshot final_state;
  std::vector<int> iv, ip;                 // iteration counts of the primal steps
  bool bufs = false;
  float4 *xb = nullptr, *vb = nullptr, *xb2 = nullptr, *vb2 = nullptr;
  float *kb = nullptr, *kvb = nullptr, *kb2 = nullptr, *kvb2 = nullptr;
  float *dbar = nullptr, *alb = nullptr, *rhob = nullptr;
  float *ybar = nullptr, *rbb = nullptr;  // learned warm start: prediction adjoint, boundary-density adjoint
  double *sxb = nullptr, *svb = nullptr;
  AdjBodyDev* body = nullptr;
  StepTapeDev tape;
  std::vector<double> start_grad;          // dL/d(x, q, v, omega) at the window start, 13 per body
  int64_t seg_cap = -1;                    // bytes one segment's kept tapes may take (-1: free memory)
  bool light = false;                      // this backward pass re-runs every step twice (tapes did not fit)
  std::vector<Snapshot> user;              // states kept by dfr_state_save
};

}  // namespace

struct dfr_ctx {
  dfr_params P{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool initialized = false;

  // derived constants
  float h = 0, dtf = 0;       // dtf: the TimeManager's step size (the next step's start)
  float4* acc = nullptr;      // non-pressure accelerations (g + forces) of the current step
  float *he_color = nullptr, *he_gc2 = nullptr;  // He 2014 colour field and |grad c|^2
  // rigid-rigid contact: rest volume / density per boundary sample, per-sample force and torque
  float *cvol0 = nullptr, *cdens0 = nullptr;
  float4 *cforce = nullptr, *ctorque = nullptr;
  bool contact_ready = false;  // rest values computed (invariant under the bodies' rigid motion)
  // RigidBodyGradientManager across bodies (contact + manager): [nb][nb][78] Jacobians and handed
  // sums, and the contact samples' gradient records ([6][nd])
  double *mgrJ = nullptr, *mgrN = nullptr;
  float4* cgrec = nullptr;
  KernelConst Kc{};            // cubic spline of radius rigid_contact_support_radius_factor r
  float4* nrm = nullptr;      // Akinci 2013 normals
  uint32_t* d_cflmax = nullptr;  // CFL squared-speed bound (float bits)
  ForceConst fc{};
  StepConst sc{};
  GridDesc grid{};

  // host staging
  std::vector<float> fx, fv;     // initial fluid (AoS)
  std::vector<BodyInfo> bodies;
  std::vector<int> body_order;   // bodies in boundary-array order (static first)

  FluidBufs f;
  BoundaryBufs b;
  GridBufs gb;
  NbrBufs nb;
  RigidDev* d_rig = nullptr;
  std::vector<RigidDev> h_rig_init;  // state restored by dfr_reset
  double* d_err = nullptr;
  double* h_err = nullptr;           // pinned
  uint32_t* h_overflow = nullptr;    // pinned [4], as NbrBufs::overflow
  int halo_cap_forced = -1;          // >= 0: LDS halo capacity fixed by dfr_set_halo_capacity
  uint32_t halo_max = 0;             // largest halo of the last build
  int err_slots = 0;
  // Jacobi solves without a host round trip per iteration: stop flags
  // (divergence, pressure), per-iteration probes in mapped pinned memory,
  // and two events the host waits on one iteration late
  int* d_stop = nullptr;
  double* d_part = nullptr;          // per-brick residual partials of the current recompute launch
  int nparts = 0;
  SolveProbe* h_probe = nullptr;
  SolveProbe* d_probe = nullptr;
  int probe_cap = 0;
  hipEvent_t pipe_ev[2] = {nullptr, nullptr};

  float time = 0.f;
  int64_t step_count = 0;
  dfr_stats stats{};

  bool profiling = false;
  uint32_t prof_mask = 0;  // timer classes recorded while profiling
  struct Rec { int cls; hipEvent_t a, b; };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> event_pool;
  double t_sum[TC_COUNT] = {0};
  int64_t t_cnt[TC_COUNT] = {0};

  std::vector<void*> allocations;

  // exact adjoint
  AdjState adj;
  StepTapeDev* rec = nullptr;     // step being recorded for the reverse pass
  bool replay = false;            // re-running steps for the reverse pass
  int force_iv = -1, force_ip = -1;  // iteration counts imposed on the next step
  std::vector<int> fixed_iv, fixed_ip;  // dfr_set_fixed_iterations
  size_t fixed_pos = 0;
  MlpState mlp;
};

// ---------------------------------------------------------------------------
namespace {

struct HipError {
  hipError_t e;
  std::string where;
};

#define HIPCHK(expr)                                                       \
  do {                                                                     \
    hipError_t _e = (expr);                                                \
    if (_e != hipSuccess) throw HipError{_e, std::string(#expr)};          \
  } while (0)

template <typename T> T* dalloc(dfr_ctx* c, size_t n) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  c->allocations.push_back(p);
  return static_cast<T*>(p);
}

int fail(dfr_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

template <typename F> int guarded(dfr_ctx* c, F&& fn) {
  if (!c) return DFR_ERR_ARG;
  try {
    return fn();
  } catch (const HipError& e) {
    return fail(c, DFR_ERR_HIP, std::string("HIP error ") + hipGetErrorString(e.e) + " in " + e.where);
  } catch (const CapacityError& e) {
    return fail(c, DFR_ERR_CAPACITY, e.what());
  } catch (const ArenaFull& e) {
    return fail(c, DFR_ERR_HIP, e.what());
  } catch (const std::exception& e) {
    return fail(c, DFR_ERR_ARG, e.what());
  }
}

hipEvent_t take_event(dfr_ctx* c) {
  if (!c->event_pool.empty()) { hipEvent_t e = c->event_pool.back(); c->event_pool.pop_back(); return e; }
  hipEvent_t e;
  HIPCHK(hipEventCreate(&e));
  return e;
}

// Times one launch class with events on the context stream (profiling only).
struct Timed {
  dfr_ctx* c; int cls; hipEvent_t a = nullptr;
  Timed(dfr_ctx* c_, int cls_) : c(c_), cls(cls_) {
    if (c->profiling && ((c->prof_mask >> cls) & 1u)) { a = take_event(c); HIPCHK(hipEventRecord(a, c->stream)); }
  }
  ~Timed() {
    if (c->profiling && a) {
      hipEvent_t e = nullptr;
      if (!c->event_pool.empty()) { e = c->event_pool.back(); c->event_pool.pop_back(); }
      else if (hipEventCreate(&e) != hipSuccess) return;
      if (hipEventRecord(e, c->stream) == hipSuccess) c->pending.push_back({cls, a, e});
    }
  }
};

void drain_timers(dfr_ctx* c) {
  for (auto& r : c->pending) {
    float ms = 0.f;
    HIPCHK(hipEventSynchronize(r.b));
    HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
    if (r.cls >= 0) {  // launches of a Jacobi iteration that did not run are not counted
      c->t_sum[r.cls] += ms;
      c->t_cnt[r.cls] += 1;
    }
    c->event_pool.push_back(r.a);
    c->event_pool.push_back(r.b);
  }
  c->pending.clear();
}

void sync_check(dfr_ctx* c) {
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
}

// Grid covering the initial scene with a margin; particles outside are clamped
// into border cells, which keeps the search exact (the predicate decides) and
// only costs time.
GridDesc make_grid(const float lo_in[3], const float hi_in[3], float h) {
  GridDesc g{};
  const float cs = h * 1.001f;
  const float margin = 4.f * cs;
  for (int k = 0; k < 3; ++k) {
    g.lo[k] = lo_in[k] - margin;
    int n = (int)std::ceil((hi_in[k] + margin - g.lo[k]) / cs) + 1;
    n = (n + 3) & ~3;
    g.n[k] = std::max(n, 4);
    g.bn[k] = g.n[k] / 4;
  }
  g.inv_cs = 1.f / cs;
  g.nkeys = (uint32_t)((uint64_t)g.bn[0] * g.bn[1] * g.bn[2] * 64ull);
  return g;
}

void fill_rigid_from_body(const dfr_ctx* c, const BodyInfo& bi, RigidDev& r) {
  std::memset(&r, 0, sizeof(r));
  for (int k = 0; k < 3; ++k) { r.x[k] = bi.x_rb[k]; r.v0[k] = bi.v0[k]; r.w0[k] = bi.w0[k]; }
  for (int k = 0; k < 4; ++k) r.q[k] = bi.q[k];
  r.dynamic = bi.dynamic ? 1 : 0;
  // Dynamic3dRigidBody::determineMassProperties (Dynamic3dRigidBody.h:184-213)
  const double rad = c->P.radius;
  const double dm = 4.0 / 3.0 * M_PI * rad * rad * rad * bi.density;
  double I0[9] = {0};
  for (int64_t j = 0; j < bi.n; ++j) {
    const double p[3] = {bi.x0[3 * j], bi.x0[3 * j + 1], bi.x0[3 * j + 2]};
    const double pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    for (int a = 0; a < 3; ++a)
      for (int bb = 0; bb < 3; ++bb) I0[3 * a + bb] += dm * ((a == bb ? pp : 0.0) - p[a] * p[bb]);
  }
  r.mass = dm * (double)bi.n;
  r.invMass = 1.0 / r.mass;
  for (int k = 0; k < 9; ++k) r.I0[k] = I0[k];
  // updateInertia (Dynamic3dRigidBody.h:215-219): I = R I0 R^T, Iinv = I^-1
  const double w = r.q[0], x = r.q[1], y = r.q[2], z = r.q[3];
  const double tx = 2 * x, ty = 2 * y, tz = 2 * z, twx = tx * w, twy = ty * w, twz = tz * w;
  const double txx = tx * x, txy = ty * x, txz = tz * x, tyy = ty * y, tyz = tz * y, tzz = tz * z;
  const double R[9] = {1 - (tyy + tzz), txy - twz, txz + twy, txy + twz, 1 - (txx + tzz), tyz - twx,
                       txz - twy, tyz + twx, 1 - (txx + tyy)};
  double RI[9] = {0}, I[9] = {0};
  for (int a = 0; a < 3; ++a) for (int bb = 0; bb < 3; ++bb) for (int k = 0; k < 3; ++k) RI[3 * a + bb] += R[3 * a + k] * I0[3 * k + bb];
  for (int a = 0; a < 3; ++a) for (int bb = 0; bb < 3; ++bb) for (int k = 0; k < 3; ++k) I[3 * a + bb] += RI[3 * a + k] * R[3 * bb + k];
  for (int k = 0; k < 9; ++k) r.I[k] = I[k];
  auto cof = [&](int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return I[3 * i1 + j1] * I[3 * i2 + j2] - I[3 * i1 + j2] * I[3 * i2 + j1];
  };
  const double det = cof(0, 0) * I[0] + cof(1, 0) * I[3] + cof(2, 0) * I[6];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.Iinv[3 * i + j] = det != 0.0 ? cof(j, i) / det : 0.0;
  // Jvv = Jww = I, all else 0 (BoundaryModel_Akinci2012.cpp:94-106,
  // RigidBodyGradientManager::reset :474-523)
  for (int k = 0; k < 3; ++k) {
    r.J[18 + 4 * k] = 1.0; r.J[45 + 4 * k] = 1.0;
    r.Jm[18 + 4 * k] = 1.0; r.Jm[45 + 4 * k] = 1.0;
  }
}

void upload_fluid_initial(dfr_ctx* c) {
  const int64_t n = c->f.n;
  std::vector<float4> p(n), v(n);
  std::vector<uint32_t> id(n);
  for (int64_t i = 0; i < n; ++i) {
    p[i] = make_float4(c->fx[3 * i], c->fx[3 * i + 1], c->fx[3 * i + 2], 0.f);
    v[i] = make_float4(c->fv[3 * i], c->fv[3 * i + 1], c->fv[3 * i + 2], 0.f);
    id[i] = (uint32_t)i;
  }
  HIPCHK(hipMemcpyAsync(c->f.pos, p.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->f.vel, v.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->f.id, id.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->f.kappa, 0, n * sizeof(float), c->stream));
  HIPCHK(hipMemsetAsync(c->f.kappav, 0, n * sizeof(float), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

// RigidBodyGradientManager::reset (RigidBodyGradientManager.cpp:474-523) of the table across bodies.
void reset_manager_multi(dfr_ctx* c) {
  if (!c->mgrJ) return;
  const size_t nb = c->bodies.size();
  std::vector<double> J(nb * nb * 78, 0.0);
  for (size_t R = 0; R < nb; ++R)
    for (int d = 0; d < 3; ++d) { J[(R * nb + R) * 78 + 18 + 4 * d] = 1.0; J[(R * nb + R) * 78 + 45 + 4 * d] = 1.0; }
  HIPCHK(hipMemcpyAsync(c->mgrJ, J.data(), J.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->mgrN, 0, J.size() * sizeof(double), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

void place_boundary_and_volumes(dfr_ctx* c) {
  reset_manager_multi(c);
  const int nbodies = (int)c->bodies.size();
  HIPCHK(hipMemcpyAsync(c->d_rig, c->h_rig_init.data(), sizeof(RigidDev) * nbodies, hipMemcpyHostToDevice, c->stream));
  launch_boundary_transform(nullptr, c->b.n, c->b, c->d_rig, c->stream);
  launch_sort_boundary(c->b, c->grid, c->gb, c->stream);
  std::vector<int> groups(nbodies);
  for (int k = 0; k < nbodies; ++k) groups[k] = c->bodies[k].dynamic ? k + 1 : 0;
  float4* tmp = c->f.pos2;  // large enough scratch only if nf >= nb; use a dedicated buffer otherwise
  float4* own = nullptr;
  if (c->f.n < c->b.n) {
    HIPCHK(hipMalloc(&own, sizeof(float4) * std::max<int64_t>(c->b.n, 1)));
    tmp = own;
  }
  launch_boundary_volume_impl(c->b.pos, c->b.x0, c->b.n, groups.data(), nbodies, c->sc.K, c->h * c->h, c->grid,
                              c->gb, tmp, c->stream);
  HIPCHK(hipMemcpyAsync(c->b.pos, tmp, sizeof(float4) * c->b.n, hipMemcpyDeviceToDevice, c->stream));
  sync_check(c);
  if (own) HIPCHK(hipFree(own));
}

void begin_step(dfr_ctx* c) {
  // TimeStepDiffDFSPH::beginStep (:379-429).  A driven body gets
  // (v, omega) = factor (v0, omega0) on the device (ramp flag).
  const int nb = (int)c->bodies.size();
  std::vector<int> ramp(nb, 0), anim(nb, 0);
  std::vector<double> fac(nb, 0.0);
  const float uat = (float)c->P.uniform_acc_time;
  for (int k = 0; k < nb; ++k) {
    BodyInfo& bi = c->bodies[k];
    if (!bi.dynamic) continue;
    if (c->P.release_mode) {
      // release branch (:381-403): only boundary model 1 is driven; it rests
      // (animated when uat > 1e-3) through uat + dt, then takes (v0, omega0)
      // at the start of every later step
      if (k == 1) {
        ramp[k] = 1;
        if (c->time <= uat + c->dtf) {
          if (uat > 1e-3f) bi.animated = true;
          fac[k] = 0.0;
        } else {
          fac[k] = 1.0;
          bi.animated = false;
        }
      }
      anim[k] = bi.animated ? 1 : 0;
      continue;
    }
    // passive-body branch (:405-426)
    if (c->time <= uat + c->dtf) {
      float factor = 1.f;
      if (uat > 1e-3f) {
        factor = (c->time / uat) > 1.f ? 1.f : (c->time / uat);
        bi.animated = true;
      }
      ramp[k] = 1;
      fac[k] = (double)factor;
    } else {
      bi.animated = false;
    }
    anim[k] = bi.animated ? 1 : 0;
  }
  launch_begin_step(c->d_rig, nb, ramp.data(), fac.data(), anim.data(), c->stream);
  if (c->rec) {
    AdjStepTable& t = c->rec->tbl;
    for (int k = 0; k < DFR_MAX_BODIES; ++k) {
      t.ramp[k] = k < nb ? ramp[k] : 0;
      t.animated[k] = k < nb ? anim[k] : 0;
      t.factor[k] = k < nb ? fac[k] : 0.0;
    }
  }
}

// The gather passes stage halos of up to nb.cap slots in LDS; the capacity
// follows the largest halo of the last build (64-slot margin, rounded to 64),
// so a launch reserves no more LDS than the halos need and more workgroups fit
// on a CU.  A halo that outgrows it before the next update is served from
// global memory with the same arithmetic.
void update_halo_cap(dfr_ctx* c) {
  if (c->h_overflow[3]) c->halo_max = c->h_overflow[3];
  if (c->halo_cap_forced >= 0) {
    c->nb.cap = std::min(c->halo_cap_forced, TILE_CAP);
  } else if (c->halo_max) {
    c->nb.cap = (int)std::min<uint32_t>((uint32_t)TILE_CAP, (c->halo_max + 64u + 63u) & ~63u);
  }
}

void check_overflow(dfr_ctx* c) {
  update_halo_cap(c);
  if (c->h_overflow[0] || c->h_overflow[1] || c->h_overflow[2]) {
    char buf[320];
    std::snprintf(buf, sizeof(buf),
                  "neighbour list capacity exceeded (fluid neighbours: %u > %d, boundary neighbours: %u > "
                  "max_neighbors=%d, fluid neighbours of a dynamic sample: %u > max_boundary_neighbors=%d)",
                  c->h_overflow[0], MAXF, c->h_overflow[2], c->P.max_neighbors, c->h_overflow[1],
                  c->P.max_boundary_neighbors);
    c->h_overflow[0] = c->h_overflow[1] = c->h_overflow[2] = 0;
    throw CapacityError(buf);
  }
}

void run_gradient(dfr_ctx* c, int mode, const int* stop) {
  if (!c->P.compute_gradient || c->replay) return;
  GradConst gc{};
  gc.mode = mode;
  const float invH = 1.f / c->dtf;
  const float invH2 = 1.f / c->dtf / c->dtf;
  // quirk of the reference: the coefficient is an unsigned int (:1266)
  gc.coeff_trunc = (mode == 0) ? (float)(unsigned int)invH2 : (float)(unsigned int)invH;
  gc.coeff = (mode == 0) ? invH2 : invH;
  gc.optimize_rotation = c->P.optimize_rotation;
  for (size_t k = 0; k < c->bodies.size(); ++k) {
    const BodyInfo& bi = c->bodies[k];
    if (!bi.dynamic || bi.animated || bi.n == 0) continue;
    Timed t(c, TC_GRADIENT);
    launch_gradient(c->sc, gc, c->f, c->b, c->nb, c->d_rig, (int)bi.dyn_first, (int)bi.n, (int)k, c->stream, stop);
  }
}

template <typename T> void d2d(dfr_ctx* c, T* dst, const T* src, size_t n) {
  if (n > 0) HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
}

// Device copies and zero fills issued at one point of the step, sent as one
// k_multi_copy launch (pieces that are not 16-byte aligned go through the
// runtime's own copy / fill).
struct Batch {
  dfr_ctx* c;
  CopyList L{};
  explicit Batch(dfr_ctx* c_) : c(c_) { L.n = 0; }
  template <typename T> void copy(T* dst, const T* src, size_t n) { add(dst, src, n * sizeof(T)); }
  template <typename T> void zero(T* dst, size_t n) { add(dst, nullptr, n * sizeof(T)); }
  void add(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    if (((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u) != 0 || (bytes & 3u) != 0) {
      if (src) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
      else HIPCHK(hipMemsetAsync(dst, 0, bytes, c->stream));
      return;
    }
    if (L.n == COPY_SEGS) flush();
    L.seg[L.n++] = CopySeg{src, dst, (uint64_t)bytes};
  }
  void flush(const int* stop = nullptr) {
    launch_multi_copy(L, c->stream, stop);
    L.n = 0;
  }
};

// Per-iteration tape buffers, grown on first use.
template <typename T> T* talloc(dfr_ctx* c, StepTapeDev* rec, size_t n) {
  return rec && rec->arena ? static_cast<T*>(rec->arena->alloc(sizeof(T) * std::max<size_t>(n, 1))) : dalloc<T>(c, n);
}

// Tape slots of Jacobi iteration m (grown on first use); the iteration's
// velocity update fills them.
IterTapeOut tape_iteration(dfr_ctx* c, std::vector<float4*>& v, std::vector<float*>& k, std::vector<float*>& b,
                           int m) {
  const size_t n = (size_t)std::max(c->f.n, 1);
  while ((int)v.size() <= m) {
    v.push_back(talloc<float4>(c, c->rec, n));
    k.push_back(talloc<float>(c, c->rec, n));
    b.push_back(talloc<float>(c, c->rec, n));
  }
  IterTapeOut o;
  o.v = v[m];
  o.k = k[m];
  o.b = b[m];
  o.bsrc = c->f.badv;
  return o;
}

// The Jacobi iterations of one solve.  fixed >= 0: exactly that many (replay,
// imposed counts).  Otherwise the reference's loop condition decides
// (divergenceSolve :816-828, pressureSolve :700-712), evaluated on the device
// by k_solve_check: every launch of an iteration is guarded by the solve's
// stop flag, and the host keeps one iteration enqueued beyond the one whose
// probe it waits for, so the stream never drains between iterations.  The
// iteration enqueued after the last one is a no-op: its tape slot is given
// back and its timers are not counted.  With the reference-mode gradient on,
// each iteration waits for its residual instead (those launches are not
// guarded).
struct IterTape {
  std::vector<float4*>* v;
  std::vector<float*>* k;
  std::vector<float*>* b;
};

template <class Iter>
int run_solve(dfr_ctx* c, int solve, int fixed, int min_it, int max_it, double eta, double* err0, int n, Iter&& iter,
              IterTape* tape, double& avg) {
  if (fixed >= 0) {
    for (int it = 0; it < fixed; ++it) iter(it, nullptr);
    return fixed;
  }
  int* stop = c->d_stop + solve;
  SolveProbe* hp = c->h_probe + (size_t)solve * c->probe_cap;
  SolveProbe* dp = c->d_probe + (size_t)solve * c->probe_cap;
  Arena* arena = (c->rec && c->rec->arena) ? c->rec->arena : nullptr;
  for (int it = 0;; ++it) {
    hp[it].done = -1;
    const std::pair<size_t, size_t> mark = arena ? arena->mark() : std::pair<size_t, size_t>{0, 0};
    const size_t tsz = tape ? tape->v->size() : 0;
    const size_t timers = c->pending.size();
    iter(it, stop);
    launch_solve_check(c->d_part, c->nparts, err0 + std::min(it, max_it - 1), n, eta, it, min_it, max_it, stop,
                       dp + it, c->stream);
    HIPCHK(hipEventRecord(c->pipe_ev[it & 1], c->stream));
    if (it >= 1) {
      HIPCHK(hipEventSynchronize(c->pipe_ev[(it - 1) & 1]));
      check_overflow(c);
      const SolveProbe pr = hp[it - 1];
      if (pr.done != 0 && pr.done != 1) throw std::runtime_error("Jacobi solve: iteration probe not written");
      if (pr.done) {
        for (size_t r = timers; r < c->pending.size(); ++r) c->pending[r].cls = -1;
        if (tape && arena && tape->v->size() > tsz) {
          tape->v->resize(tsz);
          tape->k->resize(tsz);
          tape->b->resize(tsz);
          arena->rewind(mark);
        }
        avg = pr.avg;
        return it;
      }
    }
    if (it + 1 >= max_it) {  // the last iteration the loop allows, and the previous one did not end the solve
      HIPCHK(hipEventSynchronize(c->pipe_ev[it & 1]));
      check_overflow(c);
      avg = hp[it].avg;
      return it + 1;
    }
  }
}

// Iteration count imposed on the current step (-1: run to convergence).
int imposed(dfr_ctx* c, bool pressure) {
  const int forced = pressure ? c->force_ip : c->force_iv;
  if (forced >= 0) return forced;
  const std::vector<int>& v = pressure ? c->fixed_ip : c->fixed_iv;
  return c->fixed_pos < v.size() ? v[c->fixed_pos] : -1;
}

// The TimeManager's step size: what the next step starts with (divergence
// solve, non-pressure forces, advection, positions, time) and, within a step
// after the CFL rule, what the pressure solve and the rigid bodies use.
void set_step_size(dfr_ctx* c, float dt) {
  c->dtf = dt;
  c->sc.dt = dt;
  c->stats.dt = dt;
}

void do_step(dfr_ctx* c) {
  const float dt = c->dtf;
  const int nf = c->f.n;
  StepTapeDev* rec = c->rec;
  if (rec) {
    Batch bt(c);
    bt.copy(rec->rig_pre, c->d_rig, c->bodies.size());
    bt.copy(rec->bpos, c->b.pos, c->b.n);
    bt.copy(rec->bvel, c->b.vel, c->b.n);
    bt.flush();
  }
  begin_step(c);
  {
    Timed t(c, TC_SORT);
    if (nf > 0) launch_sort_fluid_and_boundary(c->f, c->b, c->grid, c->gb, c->stream);
    else launch_sort_boundary(c->b, c->grid, c->gb, c->stream);
  }
  {
    // tape of the step's start (the sort leaves the rigid state alone) and the step's counters
    Batch bt(c);
    if (rec) {
      bt.copy(rec->rig_forced, c->d_rig, c->bodies.size());
      bt.copy(rec->x, c->f.pos, nf);
      if (c->mlp.enabled) bt.copy(rec->v0, c->f.vel, nf);
      bt.copy(rec->kvPrev, c->f.kappav, nf);
      bt.copy(rec->kPrev, c->f.kappa, nf);
    }
    bt.zero(c->nb.overflow, 4);
    bt.zero(c->d_err, c->err_slots);
    bt.zero(c->d_stop, 2);
    bt.flush();
  }
  {
    Timed t(c, TC_NBR);
    if (nf > 0) {
      launch_build_tiled(c->f, c->b, c->grid, c->gb, c->nb, c->h * c->h, c->stream);
      launch_build_bf(c->f, c->b, c->grid, c->gb, c->nb, c->h * c->h, c->stream);
    }
    HIPCHK(hipMemcpyAsync(c->h_overflow, c->nb.overflow, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  }
  const bool divOn = c->P.enable_divergence_solver != 0;
  const bool warmV = c->P.use_divergence_warmstart != 0;
  const bool warmP = c->P.use_pressure_warmstart != 0;
  MlpState& ml = c->mlp;
  const bool learned = ml.enabled;
  if (nf > 0) {
    Timed t(c, TC_DENSITY);
    launch_density_factor_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, (divOn && warmV && !learned) ? 1 : 0,
                            divOn ? 1 : 0, (learned || ml.capturing) ? ml.aux : nullptr, c->stream);
  }
  const MlpFields fields{c->f.rho, ml.aux, c->nb.cnt, c->f.kappa, c->f.kappav, c->sc.rho0, dt};
  if (ml.capturing && nf > 0) {
    d2d(c, ml.count + 1, ml.count, 1);
    launch_mlp_capture_features(fields, c->f.id, nf, ml.salt ^ (uint64_t)c->step_count * 0xD1B54A32D192ED03ull,
                                ml.threshold, ml.feat, ml.rows, ml.count, (uint32_t)ml.cap, c->stream);
  }
  if (learned && nf > 0) {
    // learned warm starts replace warmstartDivergenceSolve's and warmstartPressureSolve's halved kappa
    Timed t(c, TC_MLP);
    launch_mlp_infer_apply(ml.frag, ml.P, ml.norm, fields, nf, (divOn && warmV) ? c->f.badv : nullptr, 1.f / dt,
                           c->f.kappav, ml.kpred, c->stream);
  }
  // the divergence warm start's inputs go on the tape: written by its velocity
  // update when that runs, copied otherwise
  const bool vwarm_runs = divOn && warmV && nf > 0;
  if (rec && nf > 0 && !vwarm_runs) {
    Batch bt(c);
    bt.copy(rec->d0, c->f.badv, nf);
    bt.copy(rec->kvW, c->f.kappav, nf);
    bt.flush();
  }
  const int fixV = imposed(c, false), fixP = imposed(c, true);
  // ---- divergence solve (TimeStepDiffDFSPH::divergenceSolve, :770-881)
  int itv = 0;
  double avgv = 0.0;
  if (divOn && nf > 0) {
    const float invH = 1.f / dt;
    if (warmV) {
      Timed t(c, TC_VUPDATE);
      IterTapeOut to;
      if (rec) { to.k = rec->kvW; to.b = rec->d0; to.bsrc = c->f.badv; }
      launch_velocity_update_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, c->d_rig, c->f.kappav, nullptr, c->stream,
                               nullptr, to);
    }
    {
      Timed t(c, TC_RECOMPUTE);
      launch_recompute_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, RECOMPUTE_DIV, invH, nullptr, c->stream);
    }
    const double eta = (1.0f / dt) * (float)c->P.max_error_v * 0.01f * c->sc.rho0;
    const int maxv = c->P.max_iterations_v;
    IterTape tv{rec ? &rec->vD : nullptr, rec ? &rec->kD : nullptr, rec ? &rec->bD : nullptr};
    itv = run_solve(
        c, 0, fixV, 1, maxv, eta, c->d_err, nf,
        [&](int m, const int* stop) {
          run_gradient(c, 1, stop);
          const IterTapeOut to = rec ? tape_iteration(c, rec->vD, rec->kD, rec->bD, m) : IterTapeOut{};
          {
            Timed t(c, TC_VUPDATE);
            launch_velocity_update_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, c->d_rig, c->f.kf,
                                     warmV ? c->f.kappav : nullptr, c->stream, stop, to);
          }
          Timed t(c, TC_RECOMPUTE);
          launch_recompute_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, RECOMPUTE_DIV, invH, c->d_part, c->stream,
                             nullptr, stop);
        },
        rec ? &tv : nullptr, avgv);
  }
  {
    // non-pressure forces and the step size (TimeStepDiffDFSPH.cpp:581-586): the
    // accelerations use the step size of the step's start, which also advects the
    // velocities (:588-605) and the positions and advances the time; the CFL rule
    // sets the one of the pressure solve, the rigid bodies and the next step
    Timed t(c, TC_ELEMWISE);
    const bool forces = c->acc != nullptr;
    float h_new = dt;
    if (rec && rec->vDiv && nf > 0) d2d(c, rec->vDiv, c->f.vel, nf);
    if (forces && nf > 0) {
      if (c->nrm) launch_normals_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, c->nrm, c->stream);
      if (c->he_color)
        launch_he_color_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, c->he_color, c->he_gc2, c->stream);
      launch_forces_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, c->nrm, c->he_gc2, c->acc, c->d_rig, c->fc,
                      c->stream);
    }
    if (c->d_cflmax) {
      launch_cfl_max(c->sc, c->f.vel, forces ? c->acc : nullptr, c->b.n, c->b.vel, c->d_cflmax, c->stream);
      uint32_t bits = 0;
      HIPCHK(hipMemcpyAsync(&bits, c->d_cflmax, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      float maxVel;
      std::memcpy(&maxVel, &bits, sizeof(float));
      // Simulation::updateTimeStepSizeCFL (Simulation.cpp:608-612), in the reference's Real = float
      const float diameter = 2.f * (float)c->P.radius;
      float hn = (float)c->P.cfl_factor * 0.4f * (diameter / std::sqrt(maxVel));
      hn = std::min(hn, (float)c->P.cfl_max_time_step_size);
      hn = std::max(hn, (float)c->P.cfl_min_time_step_size);
      h_new = hn;
    }
    if (nf > 0) {
      if (forces) launch_advect_acc(c->sc, c->f, c->acc, (divOn && warmV) ? dt : 1.f, c->stream);
      else launch_advect(c->sc, c->f, (divOn && warmV) ? dt : 1.f, c->stream);
    }
    if (h_new != dt) set_step_size(c, h_new);
  }
  const float dtp = c->dtf;  // step size of the pressure solve and the rigid bodies
  // ---- pressure solve (TimeStepDiffDFSPH::pressureSolve, :654-768)
  int it = 0;
  double avg = 0.0;
  if (nf > 0) {
    const float invH2 = 1.f / (dtp * dtp);
    if (warmP) {
      {
        Timed t(c, TC_PWARM);
        launch_recompute_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb,
                           learned ? PRESSURE_WARMSTART_LEARNED : PRESSURE_WARMSTART, 0.f, nullptr, c->stream,
                           ml.kpred);
      }
      Timed t(c, TC_VUPDATE);
      IterTapeOut to;  // the pressure warm start's inputs go on the tape
      if (rec) { to.k = rec->kW; to.b = rec->rsW; to.bsrc = c->f.badv; }
      launch_velocity_update_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, c->d_rig, c->f.kappa, nullptr, c->stream,
                               nullptr, to);
    }
    {
      Timed t(c, TC_RECOMPUTE);
      launch_recompute_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, RECOMPUTE_PRESSURE, invH2, nullptr, c->stream);
    }
    const double eta = (float)c->P.max_error * 0.01f * c->sc.rho0;
    const int base = c->P.max_iterations_v, maxp = c->P.max_iterations;
    IterTape tp{rec ? &rec->vP : nullptr, rec ? &rec->kP : nullptr, rec ? &rec->bP : nullptr};
    it = run_solve(
        c, 1, fixP, c->P.min_iterations, maxp, eta, c->d_err + base, nf,
        [&](int m, const int* stop) {
          run_gradient(c, 0, stop);
          const IterTapeOut to = rec ? tape_iteration(c, rec->vP, rec->kP, rec->bP, m) : IterTapeOut{};
          {
            Timed t(c, TC_VUPDATE);
            launch_velocity_update_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, c->d_rig, c->f.kf,
                                     warmP ? c->f.kappa : nullptr, c->stream, stop, to);
          }
          Timed t(c, TC_RECOMPUTE);
          launch_recompute_t(c->sc, c->grid, c->gb, c->f, c->b, c->nb, RECOMPUTE_PRESSURE, invH2, c->d_part,
                             c->stream, nullptr, stop);
        },
        rec ? &tp : nullptr, avg);
    Timed t(c, TC_ELEMWISE);
    StepConst sco = c->sc;  // positions advance with the step's start size (:613-634)
    sco.dt = dt;
    launch_position(sco, c->f, warmP ? dtp * dtp : 1.f, c->stream);
    if (ml.capturing)  // the converged stiffness of this step is the training target
      launch_mlp_capture_targets(c->f.kappa, c->f.kappav, ml.rows, ml.count + 1, (uint32_t)nf, ml.count,
                                 (uint32_t)ml.cap, ml.targ, c->stream);
  }
  c->time = c->time + dt;
  {
    // backwardPerStep (:493-524) + velocityTimeStep + positionTimeStep
    Timed t(c, TC_RIGID);
    RigidStepConst rc{};
    rc.dt = (double)dtp;
    for (int k = 0; k < 3; ++k) rc.g[k] = (double)(float)c->P.gravity[k];
    rc.gradient_mode = c->P.gradient_mode;
    rc.optimize_rotation = c->P.optimize_rotation;
    rc.gyroscopic = c->P.gyroscopic;
    rc.compute_gradient = c->replay ? 0 : c->P.compute_gradient;
    rc.nbodies = (int)c->bodies.size();
    rc.use_manager = c->P.use_gradient_manager;
    if (c->P.use_rigid_contact) {
      // RigidBody3dBoundarySimulator::velocityTimeStep with the contact solver (:284-298)
      const float hh = c->h * c->h;
      if (!c->contact_ready) {
        launch_contact_rest(c->b, c->grid, c->gb, hh, c->Kc, (float)c->P.rigid_contact_gamma, c->cvol0, c->cdens0,
                            c->stream);
        c->contact_ready = true;
      }
      // SimulatorBase::timeStepNoGUI (SimulatorBase.cpp:1155-1170): the step's gradient sums go to the
      // manager (or the per-body chain rule runs) before the boundary velocity step; with the manager
      // its chain after the fluid coupling, the contact solver, then its chain after the rigid coupling
      const bool multi = c->mgrJ && rc.compute_gradient && rc.use_manager;
      if (rc.compute_gradient) {
        RigidStepConst r1 = rc;
        r1.phases = RS_HAND;
        r1.mgr_net = multi ? c->mgrN : nullptr;
        launch_rigid_step(r1, c->d_rig, c->stream);
      }
      if (multi && nf > 0) launch_manager_multi_velocity(rc, c->d_rig, c->mgrJ, c->mgrN, rc.nbodies, 0, c->stream);
      ContactConst cc{};
      cc.dt = (double)dtp;
      for (int k = 0; k < 3; ++k) { cc.g[k] = rc.g[k]; cc.gdt[k] = rc.g[k] * (double)dtp; }
      cc.beta = (float)c->P.rigid_contact_beta;
      cc.mu = (float)c->P.rigid_contact_friction;
      cc.gyroscopic = c->P.gyroscopic;
      cc.Ks = c->sc.K;
      cc.grec = multi ? c->cgrec : nullptr;
      cc.mgr_net = multi ? c->mgrN : nullptr;
      launch_contact(c->b, c->grid, c->gb, hh, c->Kc, cc, c->cvol0, c->cdens0, c->cforce, c->ctorque, c->d_rig,
                     rc.nbodies, c->stream);
      rc.contact = 1;
      RigidStepConst r2 = rc;
      r2.phases = RS_FORCES;
      launch_rigid_step(r2, c->d_rig, c->stream);
      if (multi) {
        launch_manager_multi_velocity(rc, c->d_rig, c->mgrJ, c->mgrN, rc.nbodies, 1, c->stream);
        launch_manager_multi_position(rc, c->d_rig, c->mgrJ, rc.nbodies, c->stream);
      }
      rc.phases = RS_ANIMATE;
    }
    launch_rigid_step(rc, c->d_rig, c->stream);
    launch_boundary_transform(c->b.dyn_idx, c->b.nd, c->b, c->d_rig, c->stream);
  }
  c->step_count++;
  if (rec) { rec->Iv = itv; rec->Ip = it; }
  if (c->fixed_pos < c->fixed_iv.size()) c->fixed_pos++;
  c->stats.iterations = it;
  c->stats.iterations_v = itv;
  c->stats.avg_density_err = avg;
  c->stats.avg_div_err = avgv;
  c->stats.total_iterations += it;
  c->stats.total_iterations_v += itv;
  c->stats.steps_since_reset += 1;
  c->stats.time = c->time;
  c->stats.step_count = c->step_count;
  c->stats.trajectory_finished =
      (c->time >= (float)c->P.target_time + (float)c->P.uniform_acc_time) ? 1 : 0;  // endStep (:448)
}

template <typename T>
void download_unpermute(const dfr_ctx* c, const T* dsrc, int comps_in, int comps_out, float* out) {
  const int64_t n = c->f.n;
  std::vector<T> buf(n);
  std::vector<uint32_t> id(n);
  HIPCHK(hipMemcpy(buf.data(), dsrc, n * sizeof(T), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(id.data(), c->f.id, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  for (int64_t k = 0; k < n; ++k) {
    const float* v = reinterpret_cast<const float*>(&buf[k]);
    for (int q = 0; q < comps_out; ++q) out[(int64_t)id[k] * comps_out + q] = v[q];
  }
  (void)comps_in;
}

// ---------------------------------------------------------------------------
// Exact adjoint: checkpoints, record-replay and the reverse pass of a step.
// ---------------------------------------------------------------------------
size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

void snapshot_alloc(dfr_ctx* c, Snapshot& S) {
  if (S.block) return;
  const size_t nf = (size_t)std::max(c->f.n, 1), nb = (size_t)std::max(c->b.n, 1),
               nr = (size_t)std::max<size_t>(c->bodies.size(), 1);
  const size_t sz[7] = {align256(nf * 16), align256(nf * 16), align256(nf * 4), align256(nf * 4),
                        align256(nf * 4), align256(nb * 16) * 2, align256(nr * sizeof(RigidDev))};
  size_t total = 0;
  for (size_t v : sz) total += v;
  HIPCHK(hipMalloc(&S.block, total));
  char* p = static_cast<char*>(S.block);
  S.pos = reinterpret_cast<float4*>(p); p += sz[0];
  S.vel = reinterpret_cast<float4*>(p); p += sz[1];
  S.kappa = reinterpret_cast<float*>(p); p += sz[2];
  S.kappav = reinterpret_cast<float*>(p); p += sz[3];
  S.id = reinterpret_cast<uint32_t*>(p); p += sz[4];
  S.bpos = reinterpret_cast<float4*>(p); p += sz[5] / 2;
  S.bvel = reinterpret_cast<float4*>(p); p += sz[5] / 2;
  S.rig = reinterpret_cast<RigidDev*>(p);
}

void snapshot_free(Snapshot& S) {
  if (S.block) (void)hipFree(S.block);
  S = Snapshot{};
}

void snapshot_save(dfr_ctx* c, Snapshot& S) {
  Timed t(c, TC_CHECKPOINT);
  snapshot_alloc(c, S);
  const int nf = c->f.n;
  Batch bt(c);
  bt.copy(S.pos, c->f.pos, nf); bt.copy(S.vel, c->f.vel, nf);
  bt.copy(S.kappa, c->f.kappa, nf); bt.copy(S.kappav, c->f.kappav, nf); bt.copy(S.id, c->f.id, nf);
  bt.copy(S.bpos, c->b.pos, c->b.n); bt.copy(S.bvel, c->b.vel, c->b.n);
  bt.copy(S.rig, c->d_rig, c->bodies.size());
  bt.flush();
  S.time = c->time;
  S.dt = c->dtf;
  S.step_count = c->step_count;
  S.animated.resize(c->bodies.size());
  for (size_t k = 0; k < c->bodies.size(); ++k) S.animated[k] = c->bodies[k].animated ? 1 : 0;
}

void snapshot_restore(dfr_ctx* c, const Snapshot& S) {
  Timed t(c, TC_CHECKPOINT);
  const int nf = c->f.n;
  Batch bt(c);
  bt.copy(c->f.pos, (const float4*)S.pos, nf); bt.copy(c->f.vel, (const float4*)S.vel, nf);
  bt.copy(c->f.kappa, (const float*)S.kappa, nf); bt.copy(c->f.kappav, (const float*)S.kappav, nf);
  bt.copy(c->f.id, (const uint32_t*)S.id, nf);
  bt.copy(c->b.pos, (const float4*)S.bpos, c->b.n); bt.copy(c->b.vel, (const float4*)S.bvel, c->b.n);
  bt.copy(c->d_rig, (const RigidDev*)S.rig, c->bodies.size());
  bt.flush();
  c->time = S.time;
  set_step_size(c, S.dt);
  c->step_count = S.step_count;
  for (size_t k = 0; k < c->bodies.size(); ++k) c->bodies[k].animated = S.animated[k] != 0;
}

void adjoint_buffers(dfr_ctx* c) {
  AdjState& A = c->adj;
  if (A.bufs) return;
  const size_t nf = (size_t)std::max(c->f.n, 1), nb = (size_t)std::max(c->b.n, 1),
               nd = (size_t)std::max(c->b.nd, 1), nr = (size_t)std::max<size_t>(c->bodies.size(), 1);
  A.xb = dalloc<float4>(c, nf); A.vb = dalloc<float4>(c, nf); A.xb2 = dalloc<float4>(c, nf); A.vb2 = dalloc<float4>(c, nf);
  A.kb = dalloc<float>(c, nf); A.kvb = dalloc<float>(c, nf); A.kb2 = dalloc<float>(c, nf); A.kvb2 = dalloc<float>(c, nf);
  A.dbar = dalloc<float>(c, nf); A.alb = dalloc<float>(c, nf); A.rhob = dalloc<float>(c, nf);
  A.ybar = dalloc<float>(c, 2 * nf); A.rbb = dalloc<float>(c, nf);
  A.sxb = dalloc<double>(c, 3 * nd); A.svb = dalloc<double>(c, 3 * nd);
  A.body = dalloc<AdjBodyDev>(c, nr);
  StepTapeDev& T = A.tape;
  T.x = dalloc<float4>(c, nf);
  T.v0 = dalloc<float4>(c, nf);
  T.vDiv = c->P.viscosity_method != 0 ? dalloc<float4>(c, nf) : nullptr;
  T.kvPrev = dalloc<float>(c, nf); T.d0 = dalloc<float>(c, nf); T.kvW = dalloc<float>(c, nf);
  T.kPrev = dalloc<float>(c, nf); T.rsW = dalloc<float>(c, nf); T.kW = dalloc<float>(c, nf);
  T.bpos = dalloc<float4>(c, nb); T.bvel = dalloc<float4>(c, nb);
  T.rig_pre = dalloc<RigidDev>(c, nr); T.rig_forced = dalloc<RigidDev>(c, nr);
  set_adjoint_kernel_attributes();
  A.bufs = true;
}

// A stored step's own buffers, from the arena.
void tape_alloc_fixed(dfr_ctx* c, StepTapeDev& T) {
  Arena& R = c->adj.arena;
  T.arena = &R;
  const size_t nf = (size_t)std::max(c->f.n, 1), nb = (size_t)std::max(c->b.n, 1),
               nr = (size_t)std::max<size_t>(c->bodies.size(), 1);
  auto fa = [&](size_t n) { return static_cast<float*>(R.alloc(sizeof(float) * n)); };
  T.x = static_cast<float4*>(R.alloc(sizeof(float4) * nf));
  T.v0 = c->mlp.enabled ? static_cast<float4*>(R.alloc(sizeof(float4) * nf)) : nullptr;
  T.vDiv = c->P.viscosity_method != 0 ? static_cast<float4*>(R.alloc(sizeof(float4) * nf)) : nullptr;
  T.kvPrev = fa(nf); T.d0 = fa(nf); T.kvW = fa(nf); T.kPrev = fa(nf); T.rsW = fa(nf); T.kW = fa(nf);
  T.bpos = static_cast<float4*>(R.alloc(sizeof(float4) * nb));
  T.bvel = static_cast<float4*>(R.alloc(sizeof(float4) * nb));
  T.rig_pre = static_cast<RigidDev*>(R.alloc(sizeof(RigidDev) * nr));
  T.rig_forced = static_cast<RigidDev*>(R.alloc(sizeof(RigidDev) * nr));
}

// The reverse pass right after a replayed step reads the live structures.
void live_view(dfr_ctx* c, StepTapeDev& T) {
  T.fstart = c->gb.fstart; T.perm = c->gb.perm;
  T.nbr8 = c->nb.nbr8; T.thdr = c->nb.thdr; T.tmap = c->nb.tmap; T.nbrB = c->nb.nbrB; T.cnt = c->nb.cnt;
  T.rho = c->f.rho; T.alpha = c->f.alpha; T.s = c->f.s; T.aux = c->mlp.aux;
  T.rig_post = c->d_rig;
}

template <typename T0> const T0* arena_copy(dfr_ctx* c, const T0* src, size_t n) {
  T0* dst = static_cast<T0*>(c->adj.arena.alloc(sizeof(T0) * std::max<size_t>(n, 1)));
  d2d(c, dst, src, n);
  return dst;
}

// Points the context's per-step structures (neighbour lists, grid offsets,
// permutation, density, factor, sum_grad_p_k, network features) at fresh arena
// buffers, so the step about to run leaves them in place as part of its tape.
void tape_swap_in(dfr_ctx* c, StepTapeDev& T) {
  AdjState& A = c->adj;
  Arena& R = A.arena;
  const size_t NF = (size_t)std::max(c->f.n, 1);
  // every buffer first (an allocation that fails leaves the context as it was)
  AdjState::Live L{};
  L.nbr8 = static_cast<uint4*>(R.alloc(sizeof(uint4) * NPACK * NF));
  const size_t NB = (size_t)c->grid.bn[0] * c->grid.bn[1] * c->grid.bn[2];
  L.thdr = static_cast<uint4*>(R.alloc(sizeof(uint4) * NB));
  L.tmap = static_cast<uint32_t*>(R.alloc(sizeof(uint32_t) * NB * TILE_CAP));
  L.nbrB = static_cast<uint32_t*>(R.alloc(sizeof(uint32_t) * (size_t)c->nb.maxb * NF));
  L.cnt = static_cast<uint32_t*>(R.alloc(sizeof(uint32_t) * NF));
  L.fstart = static_cast<uint32_t*>(R.alloc(sizeof(uint32_t) * ((size_t)c->grid.nkeys + 1)));
  L.perm = static_cast<uint32_t*>(R.alloc(sizeof(uint32_t) * NF));
  L.rho = static_cast<float*>(R.alloc(sizeof(float) * NF));
  L.alpha = static_cast<float*>(R.alloc(sizeof(float) * NF));
  L.s = static_cast<float4*>(R.alloc(sizeof(float4) * NF));
  L.aux = c->mlp.aux ? static_cast<float2*>(R.alloc(sizeof(float2) * NF)) : nullptr;
  if (!A.swapped) {
    A.own = {c->nb.nbr8, c->nb.thdr, c->nb.tmap, c->nb.nbrB, c->nb.cnt, c->gb.fstart, c->gb.perm, c->f.rho,
             c->f.alpha, c->f.s, c->mlp.aux};
    A.swapped = true;
  }
  c->nb.nbr8 = L.nbr8; c->nb.thdr = L.thdr; c->nb.tmap = L.tmap; c->nb.nbrB = L.nbrB; c->nb.cnt = L.cnt;
  c->gb.fstart = L.fstart; c->gb.perm = L.perm;
  c->f.rho = L.rho; c->f.alpha = L.alpha; c->f.s = L.s;
  if (c->mlp.aux) c->mlp.aux = L.aux;
  T.nbr8 = c->nb.nbr8; T.thdr = c->nb.thdr; T.tmap = c->nb.tmap; T.nbrB = c->nb.nbrB; T.cnt = c->nb.cnt;
  T.fstart = c->gb.fstart; T.perm = c->gb.perm;
  T.rho = c->f.rho; T.alpha = c->f.alpha; T.s = c->f.s; T.aux = c->mlp.aux;
}

// Ends the swap: the context's own buffers receive the last step's structures.
void tape_swap_out(dfr_ctx* c) {
  AdjState& A = c->adj;
  if (!A.swapped) return;
  const int nf = c->f.n;
  const auto& o = A.own;
  d2d(c, o.nbr8, (const uint4*)c->nb.nbr8, (size_t)NPACK * nf);
  {
    const size_t NB = (size_t)c->grid.bn[0] * c->grid.bn[1] * c->grid.bn[2];
    d2d(c, o.thdr, (const uint4*)c->nb.thdr, NB);
    d2d(c, o.tmap, (const uint32_t*)c->nb.tmap, NB * TILE_CAP);
  }
  d2d(c, o.nbrB, (const uint32_t*)c->nb.nbrB, (size_t)c->nb.maxb * nf);
  d2d(c, o.cnt, (const uint32_t*)c->nb.cnt, (size_t)nf);
  d2d(c, o.fstart, (const uint32_t*)c->gb.fstart, (size_t)c->grid.nkeys + 1);
  d2d(c, o.perm, (const uint32_t*)c->gb.perm, (size_t)nf);
  d2d(c, o.rho, (const float*)c->f.rho, (size_t)nf);
  d2d(c, o.alpha, (const float*)c->f.alpha, (size_t)nf);
  d2d(c, o.s, (const float4*)c->f.s, (size_t)nf);
  if (o.aux) d2d(c, o.aux, (const float2*)c->mlp.aux, (size_t)nf);
  c->nb.nbr8 = o.nbr8; c->nb.thdr = o.thdr; c->nb.tmap = o.tmap; c->nb.nbrB = o.nbrB; c->nb.cnt = o.cnt;
  c->gb.fstart = o.fstart; c->gb.perm = o.perm;
  c->f.rho = o.rho; c->f.alpha = o.alpha; c->f.s = o.s; c->mlp.aux = o.aux;
  A.swapped = false;
}

// Ends the current rollout; device blocks (checkpoints, tape arena) stay
// allocated for the next one until adjoint_free.
void adjoint_release(dfr_ctx* c) {
  AdjState& A = c->adj;
  c->rec = nullptr;
  tape_swap_out(c);
  A.nckpt = 0;
  A.iv.clear();
  A.ip.clear();
  A.tapes.clear();
  A.arena.reset();
  A.steps = 0;
  A.active = false;
}

void adjoint_free(dfr_ctx* c) {
  AdjState& A = c->adj;
  adjoint_release(c);
  for (auto& S : A.ckpt) snapshot_free(S);
  for (auto& S : A.seg) snapshot_free(S);
  snapshot_free(A.final_state);
  A.ckpt.clear();
  A.seg.clear();
  A.arena.release();
}

// Reverse pass of the step just replayed with c->rec = &tape: the adjoint
// fields (sorted order of that step) become the adjoint of the state before
// the step (order of the previous step), and the rigid adjoints move back one
// step.
void reverse_step(dfr_ctx* c, const StepTapeDev& T) {
  AdjState& A = c->adj;
  const int nf = c->f.n;
  const float dt = c->dtf;
  const int nbodies = (int)c->bodies.size();
  AdjPass a{};
  a.c = c->sc; a.g = c->grid; a.fstart = T.fstart; a.x = T.x;
  a.nbr8 = T.nbr8; a.thdr = T.thdr; a.tmap = T.tmap; a.nbrB = T.nbrB; a.cnt = T.cnt;
  a.bx = T.bpos; a.bv = T.bvel; a.bx0 = c->b.x0; a.nb_static = c->b.n - c->b.nd;
  a.cap = c->nb.cap;
  a.sxb = A.sxb; a.svb = A.svb; a.body = A.body;
  {
    Timed t(c, TC_ADJ_MISC);
    RigidStepConst rc{};
    rc.dt = (double)dt;
    rc.gyroscopic = c->P.gyroscopic;
    rc.nbodies = nbodies;
    launch_adj_rigid(rc, T.tbl, T.rig_forced, T.rig_pre, T.rig_post, A.body, nbodies, c->stream);
    Batch bt(c);
    if (c->b.nd > 0) {
      bt.zero(A.sxb, 3 * (size_t)c->b.nd);
      bt.zero(A.svb, 3 * (size_t)c->b.nd);
    }
    if (nf > 0) {
      bt.zero(A.alb, (size_t)nf);
      bt.zero(A.rhob, (size_t)nf);
      if (c->mlp.enabled) bt.zero(A.ybar, 2 * (size_t)nf);
    }
    bt.flush();
  }
  if (nf > 0) {
    {
      Timed t(c, TC_ADJ_MISC);
      launch_adj_position(nf, dt, A.xb, A.vb, c->stream);
    }
    const bool warmP = c->P.use_pressure_warmstart != 0;
    const bool learned = c->mlp.enabled;
    const float h2 = dt * dt;
    for (int m = T.Ip - 1; m >= 0; --m) {
      {
        Timed t(c, TC_ADJ_UPDATE);
        launch_adj_update(a, ADJ_PRESSURE_ITER, T.kP[m], A.vb, warmP ? A.kb : nullptr, h2, T.bP[m], nullptr,
                          T.alpha, A.dbar, A.alb, A.rhob, A.xb, c->stream);
      }
      Timed t(c, TC_ADJ_DELTA);
      launch_adj_delta(a, T.vP[m], A.dbar, A.vb, A.xb, c->stream);
    }
    if (warmP) {
      Timed t(c, TC_ADJ_UPDATE);
      if (learned)
        launch_adj_update(a, ADJ_PRESSURE_WARM_LEARNED, T.kW, A.vb, A.kb, h2, T.rsW, nullptr, T.alpha, A.ybar,
                          A.alb, A.rhob, A.xb, c->stream);
      else
        launch_adj_update(a, ADJ_PRESSURE_WARM, T.kW, A.vb, A.kb, h2, T.rsW, T.kPrev, T.alpha, A.kb, A.alb,
                          A.rhob, A.xb, c->stream);
    }
    if (c->P.viscosity_method != 0) {
      // the advection with the step's viscosity, at the velocities after the divergence solve
      Timed t(c, TC_ADJ_MISC);
      launch_adj_viscosity(a, c->P.viscosity_method, T.vDiv, T.rho, A.vb, A.vb2, c->fc, A.vb, A.xb, A.rhob,
                           c->stream);
    }
    if (c->P.enable_divergence_solver) {
      const bool warmV = c->P.use_divergence_warmstart != 0;
      for (int m = T.Iv - 1; m >= 0; --m) {
        {
          Timed t(c, TC_ADJ_UPDATE);
          launch_adj_update(a, ADJ_DIV_ITER, T.kD[m], A.vb, warmV ? A.kvb : nullptr, dt, T.bD[m], nullptr,
                            T.alpha, A.dbar, A.alb, A.rhob, A.xb, c->stream);
        }
        Timed t(c, TC_ADJ_DELTA);
        launch_adj_delta(a, T.vD[m], A.dbar, A.vb, A.xb, c->stream);
      }
      if (warmV) {
        Timed t(c, TC_ADJ_UPDATE);
        if (learned)
          launch_adj_update(a, ADJ_DIV_WARM_LEARNED, T.kvW, A.vb, A.kvb, dt, T.d0, nullptr, T.alpha, A.ybar,
                            A.alb, A.rhob, A.xb, c->stream);
        else
          launch_adj_update(a, ADJ_DIV_WARM, T.kvW, A.vb, A.kvb, dt, T.d0, T.kvPrev, T.alpha, A.kvb, A.alb,
                            A.rhob, A.xb, c->stream);
      }
    }
    if (learned) {
      // through the network to its features, then to the state they were computed from
      Timed t(c, TC_ADJ_MLP);
      const MlpFields fields{T.rho, T.aux, T.cnt, T.kPrev, T.kvPrev, c->sc.rho0, dt};
      const bool divWarm = c->P.enable_divergence_solver && c->P.use_divergence_warmstart;
      const LearnedAdjSink sink{A.kb, A.kvb, A.rhob, A.dbar, A.rbb, warmP ? 0 : 1, divWarm ? 0 : 1, dt};
      launch_mlp_input_grad_apply(c->mlp.frag, c->mlp.P, c->mlp.norm, fields, nf, A.ybar, sink, c->stream);
      if (c->P.enable_divergence_solver) launch_adj_delta(a, T.v0, A.dbar, A.vb, A.xb, c->stream);
    }
    {
      Timed t(c, TC_ADJ_FACTOR);
      launch_adj_factor(a, T.alpha, A.alb, A.rhob, T.s, A.xb, learned ? A.rbb : nullptr, c->stream);
    }
  }
  Timed t(c, TC_ADJ_MISC);
  for (int k = 0; k < nbodies; ++k) {
    const BodyInfo& bi = c->bodies[k];
    if (bi.dynamic && bi.n > 0) launch_adj_boundary_reduce(a, k, (int)bi.dyn_first, (int)bi.n, A.body, c->stream);
  }
  if (nf > 0) {
    launch_adj_unpermute(nf, T.perm, A.xb, A.vb, A.kb, A.kvb, A.xb2, A.vb2, A.kb2, A.kvb2, c->stream);
    std::swap(A.xb, A.xb2); std::swap(A.vb, A.vb2); std::swap(A.kb, A.kb2); std::swap(A.kvb, A.kvb2);
  }
}

// One recorded step s of the rollout: impose the primal's iteration counts.
void replay_step(dfr_ctx* c, int64_t s, bool record) {
  c->force_iv = c->adj.iv[s];
  c->force_ip = c->adj.ip[s];
  c->rec = record ? &c->adj.tape : nullptr;
  do_step(c);
  c->rec = nullptr;
  c->force_iv = c->force_ip = -1;
}

// Reverse of the segment s0..s1-1 (checkpoint `seg`) with every step's tape
// kept in the arena: the segment is re-run once from its checkpoint with
// recording, then its steps are reversed from the kept tapes.  Returns false,
// with the arena rewound and nothing reversed, when the first step's tape shows
// that the segment's tapes would not fit in device memory.
bool reverse_segment_kept(dfr_ctx* c, int64_t seg, int64_t s0, int64_t s1) {
  AdjState& A = c->adj;
  if (A.seg_cap == 0 || A.light) return false;
  A.arena.reset();
  A.tapes.clear();
  A.tapes.resize((size_t)(s1 - s0));
  snapshot_restore(c, A.ckpt[seg]);
  for (int64_t s = s0; s < s1; ++s) {
    StepTapeDev& T = A.tapes[(size_t)(s - s0)];
    try {
      tape_alloc_fixed(c, T);
      tape_swap_in(c, T);
      c->force_iv = A.iv[s];
      c->force_ip = A.ip[s];
      c->rec = &T;
      do_step(c);
      c->rec = nullptr;
      c->force_iv = c->force_ip = -1;
      T.rig_post = arena_copy(c, (const RigidDev*)c->d_rig, c->bodies.size());
    } catch (const ArenaFull&) {
      // a later step's tape outgrew the estimate: this segment and the rest
      // of the pass go the memory-light way (the checkpoint is restored before
      // the re-run); the arena's chunks are freed for the states that way keeps
      c->rec = nullptr;
      c->force_iv = c->force_ip = -1;
      (void)hipGetLastError();
      HIPCHK(hipStreamSynchronize(c->stream));
      tape_swap_out(c);
      HIPCHK(hipStreamSynchronize(c->stream));
      A.tapes.clear();
      A.arena.release();
      A.light = true;
      return false;
    }
    if (s == s0 && s1 - s0 > 1) {
      // room for the rest of the segment, with a margin for steps that iterate more
      size_t cap = 0;
      for (const auto& ch : A.arena.chunks) cap += ch.second;
      const size_t need = A.arena.used() / 2 * 3 * (size_t)(s1 - s0);
      size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      const bool over = A.seg_cap > 0 ? need > (size_t)A.seg_cap : (need > cap && need - cap + ((size_t)2 << 30) > fr);
      if (over) {
        tape_swap_out(c);
        HIPCHK(hipStreamSynchronize(c->stream));
        A.tapes.clear();
        A.arena.release();
        A.light = true;
        return false;
      }
    }
  }
  for (int64_t s = s1 - 1; s >= s0; --s) reverse_step(c, A.tapes[(size_t)(s - s0)]);
  A.tapes.clear();
  return true;
}

// Device scratch for the host-array entry points below.
// ---------------------------------------------------------------------------
// Allocation-free network loss/gradient (dfr_mlp_loss_grad_noalloc).  Its
// staging (features [rows][6], targets [rows][2]) and accumulators (gradient
// [MLP_NP] floats, loss one double) live in `scratch`, one fixed array per GPU.
// In host code `scratch` names no device memory: its address on a device comes
// from hipGetSymbolAddress, taken once per device at context creation with
// that device current, and only that address is passed to kernels and copies.
// ---------------------------------------------------------------------------
constexpr size_t NA_ROWS = DFR_NOALLOC_MAX_ROWS;
constexpr size_t NA_FEAT = 0;                                     // bytes
constexpr size_t NA_TARG = NA_FEAT + sizeof(float) * 6 * NA_ROWS;
constexpr size_t NA_GRAD = NA_TARG + sizeof(float) * 2 * NA_ROWS;
constexpr size_t NA_LOSS = (NA_GRAD + sizeof(float) * MLP_NP + 255) / 256 * 256;
constexpr size_t NA_BYTES = NA_LOSS + 256;
constexpr int NA_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[NA_MAX_DEVICES] = {};  // per device, from hipGetSymbolAddress

// Called with `device` current.
void noalloc_setup(int device) {
  if (device < 0 || device >= NA_MAX_DEVICES) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_dev[device] = static_cast<unsigned char*>(p);
}

unsigned char* noalloc_base(int device) {
  if (device < 0 || device >= NA_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_dev[device];
}

// Clears the accumulators with a kernel (no runtime memset), 4-byte words.
__global__ void k_clear_words(uint32_t* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

template <typename T> struct DevTmp {
  T* p = nullptr;
  explicit DevTmp(size_t n) { if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) throw HipError{hipErrorOutOfMemory, "DevTmp"}; }
  ~DevTmp() { if (p) (void)hipFree(p); }
};

// ---------------------------------------------------------------------------
// Learned warm start: allocation, initialisation, training loop.
// ---------------------------------------------------------------------------
void mlp_alloc(dfr_ctx* c) {
  MlpState& M = c->mlp;
  if (M.P) return;
  M.P = dalloc<float>(c, MLP_NP);
  M.G = dalloc<float>(c, MLP_NP);
  M.m = dalloc<float>(c, MLP_NP);
  M.v = dalloc<float>(c, MLP_NP);
  HIPCHK(hipMemsetAsync(M.G, 0, sizeof(float) * MLP_NP, c->stream));
  HIPCHK(hipMemsetAsync(M.m, 0, sizeof(float) * MLP_NP, c->stream));
  HIPCHK(hipMemsetAsync(M.v, 0, sizeof(float) * MLP_NP, c->stream));
  M.frag = dalloc<char>(c, mlp_frag_bytes());
  const size_t nf = (size_t)std::max(c->f.n, 1);
  M.kpred = dalloc<float>(c, nf);
  M.aux = dalloc<float2>(c, nf);
  M.count = dalloc<uint32_t>(c, 2);
  HIPCHK(hipMemsetAsync(M.count, 0, sizeof(uint32_t) * 2, c->stream));
  set_mlp_kernel_attributes();
  for (int k = 0; k < 6; ++k) { M.norm.mu_in[k] = 0.f; M.norm.is_in[k] = 1.f; }
  for (int k = 0; k < 2; ++k) { M.norm.mu_out[k] = 0.f; M.norm.sd_out[k] = 1.f; }
}

void mlp_upload_params(dfr_ctx* c, const std::vector<float>& P) {
  MlpState& M = c->mlp;
  HIPCHK(hipMemcpyAsync(M.P, P.data(), sizeof(float) * MLP_NP, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(M.m, 0, sizeof(float) * MLP_NP, c->stream));
  HIPCHK(hipMemsetAsync(M.v, 0, sizeof(float) * MLP_NP, c->stream));
  launch_mlp_pack(M.P, M.frag, c->stream);
  M.adam_t = 0;
  M.has_params = true;
}

// Uniform init with bound sqrt(6 / fan_in) (He et al. 2015) for the SiLU
// layers, a tenth of it for the linear output layer; zero biases.
void mlp_init_params(dfr_ctx* c, uint64_t seed) {
  static constexpr int O[4] = {32, 32, 32, 2}, I[4] = {6, 32, 32, 32}, OW[4] = {0, 224, 1280, 2336};
  static_assert(OW[3] + O[3] * I[3] + O[3] == MLP_NP, "layer table and parameter count disagree");
  std::vector<float> P(MLP_NP, 0.f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  auto uni = [&]() {
    st ^= st >> 12; st ^= st << 25; st ^= st >> 27;
    return (double)((st * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int L = 0; L < 4; ++L) {
    const double b = std::sqrt(6.0 / I[L]) * (L == 3 ? 0.1 : 1.0);
    for (int k = 0; k < O[L] * I[L]; ++k) P[OW[L] + k] = (float)((2.0 * uni() - 1.0) * b);
  }
  mlp_upload_params(c, P);
}

uint32_t mlp_dataset_size(dfr_ctx* c) {
  MlpState& M = c->mlp;
  if (!M.count) return 0;
  uint32_t n = 0;
  HIPCHK(hipMemcpyAsync(&n, M.count, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return std::min<uint32_t>(n, (uint32_t)M.cap);
}

// Feature and target mean / standard deviation over the captured pairs.
void mlp_fit_norm(dfr_ctx* c, uint32_t n) {
  MlpState& M = c->mlp;
  std::vector<float> f((size_t)n * 6), t((size_t)n * 2);
  HIPCHK(hipMemcpy(f.data(), M.feat, sizeof(float) * f.size(), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(t.data(), M.targ, sizeof(float) * t.size(), hipMemcpyDeviceToHost));
  auto stats = [&](const std::vector<float>& a, int w, int k, float& mu, float& sd) {
    double s1 = 0, s2 = 0;
    for (uint32_t r = 0; r < n; ++r) { const double x = a[(size_t)r * w + k]; s1 += x; s2 += x * x; }
    const double m = s1 / n, var = std::max(s2 / n - m * m, 0.0);
    mu = (float)m;
    sd = var > 1e-30 ? (float)std::sqrt(var) : 1.f;
  };
  for (int k = 0; k < 6; ++k) { float sd; stats(f, 6, k, M.norm.mu_in[k], sd); M.norm.is_in[k] = 1.f / sd; }
  for (int k = 0; k < 2; ++k) stats(t, 2, k, M.norm.mu_out[k], M.norm.sd_out[k]);
  M.norm_set = true;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void dfr_default_params(dfr_params* p, double radius, double dt) {
  std::memset(p, 0, sizeof(*p));
  p->radius = radius;
  p->dt = dt;
  p->rho0 = 1000.0;
  p->gravity[0] = 0.0; p->gravity[1] = -9.81; p->gravity[2] = 0.0;
  p->max_error = 0.01;
  p->max_error_v = 0.1;
  p->min_iterations = 2;
  p->max_iterations = 100;
  p->max_iterations_v = 100;
  p->use_pressure_warmstart = 1;
  p->use_divergence_warmstart = 1;
  p->enable_divergence_solver = 1;
  p->gradient_mode = 1;
  p->optimize_rotation = 1;
  p->gyroscopic = 1;
  p->compute_gradient = 1;
  p->uniform_acc_time = 0.0;
  p->target_time = 1e30;
  p->max_neighbors = 64;
  p->max_boundary_neighbors = 128;
  p->release_mode = 0;
  p->use_gradient_manager = 0;
  p->surface_tension_method = 0;        // SurfaceTensionBase.cpp:12-13
  p->surface_tension = 0.05;
  p->surface_tension_boundary = 0.01;
  p->viscosity_method = 0;              // ViscosityBase.cpp:12, Viscosity_XSPH.cpp:17
  p->viscosity = 0.01;
  p->viscosity_boundary = 0.0;
  p->cfl_method = 0;                    // Simulation.cpp:56-60
  p->cfl_factor = 0.5;
  p->cfl_min_time_step_size = 0.0001;
  p->cfl_max_time_step_size = 0.005;
  p->use_rigid_contact = 0;             // SceneLoader.cpp:58-71
  p->rigid_contact_beta = 1.0;
  p->rigid_contact_gamma = 0.7;
  p->rigid_contact_support_radius_factor = 4.0;
  p->rigid_contact_friction = 0.0;
}

dfr_ctx* dfr_create(const dfr_params* p, int device) {
  if (!p || p->radius <= 0.0 || p->dt <= 0.0) return nullptr;
  // built: surface tension 0/2/3 (none, Akinci 2013, He 2014), viscosity 0/1/2 (none, Standard, XSPH)
  // without boundary viscosity, step size 0/1 (fixed, CFL)
  if ((p->surface_tension_method != 0 && p->surface_tension_method != 2 && p->surface_tension_method != 3) ||
      p->viscosity_method < 0 || p->viscosity_method > 2 || p->viscosity_boundary != 0.0 ||
      (p->cfl_method != 0 && p->cfl_method != 1) ||
      (p->cfl_method == 1 && !(p->cfl_min_time_step_size > 0.0 && p->cfl_max_time_step_size >= p->cfl_min_time_step_size)) ||
      (p->use_rigid_contact != 0 && !(p->rigid_contact_support_radius_factor > 0.0 && p->rigid_contact_gamma > 0.0)))
    return nullptr;
  dfr_ctx* c = new dfr_ctx();
  c->P = *p;
  if (c->P.max_neighbors <= 0) c->P.max_neighbors = 64;
  if (c->P.max_boundary_neighbors <= 0) c->P.max_boundary_neighbors = 128;
  if (c->P.max_neighbors > 65535) c->P.max_neighbors = 65535;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return nullptr;
  }
  noalloc_setup(device);
  c->h = 4.f * (float)p->radius;  // Simulation.cpp:195-216
  c->dtf = (float)p->dt;
  const float pi = (float)M_PI;
  const float h3 = c->h * c->h * c->h;
  c->sc.K.h = c->h;
  c->sc.K.k = 8.f / (pi * h3);
  c->sc.K.l = 48.f / (pi * h3);
  c->sc.K.w0 = c->sc.K.k;  // W(0) = k (6*0 - 6*0 + 1)
  c->sc.K.inv_h = 1.f / c->h;
  c->sc.K.l_ih = c->sc.K.l / c->h;
  c->sc.K.l_ih2 = c->sc.K.l / (c->h * c->h);
  {
    const float hc = (float)(p->rigid_contact_support_radius_factor * p->radius);
    const float hc3 = hc * hc * hc;
    c->Kc.h = hc;
    c->Kc.k = 8.f / (pi * hc3);
    c->Kc.l = 48.f / (pi * hc3);
    c->Kc.w0 = c->Kc.k;
    c->Kc.inv_h = 1.f / hc;
    c->Kc.l_ih = c->Kc.l / hc;
    c->Kc.l_ih2 = c->Kc.l / (hc * hc);
  }
  c->sc.dt = c->dtf;
  c->stats.dt = c->dtf;
  c->sc.rho0 = (float)p->rho0;
  const float diam = 2.f * (float)p->radius;
  c->sc.V_f = 0.8f * diam * diam * diam;  // FluidModel.cpp:255-275
  c->sc.m_f = c->sc.V_f * c->sc.rho0;
  for (int k = 0; k < 3; ++k) c->sc.g[k] = (float)p->gravity[k];
  c->sc.eps = 1e-5f;  // TimeStepDiffDFSPH.h:26
  // non-pressure forces: Akinci 2013 cohesion / adhesion kernels of radius h
  // (CohesionKernel::setRadius, AdhesionKernel::setRadius, SPHKernels.h:446-521)
  c->fc.h = c->h;
  c->fc.st_method = p->surface_tension_method;
  c->fc.st_k = (float)p->surface_tension;
  c->fc.st_kb = (float)p->surface_tension_boundary;
  c->fc.visc_mode = p->viscosity_method;
  c->fc.visc = (float)p->viscosity;
  c->fc.ck = 32.f / ((float)M_PI * std::pow(c->h, 9.f));
  c->fc.cc = std::pow(c->h, 6.f) / 64.f;
  c->fc.ak = 0.007f / std::pow(c->h, 3.25f);
  return c;
}

int dfr_destroy(dfr_ctx* c) {
  if (!c) return DFR_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  try {
    adjoint_free(c);
  } catch (...) {
  }
  for (auto& S : c->adj.user) snapshot_free(S);
  for (auto& r : c->pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  for (void* p : c->allocations) (void)hipFree(p);
  if (c->h_err) (void)hipHostFree(c->h_err);
  if (c->h_overflow) (void)hipHostFree(c->h_overflow);
  if (c->h_probe) (void)hipHostFree(c->h_probe);
  for (auto e : c->pipe_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return DFR_OK;
}

const char* dfr_last_error(const dfr_ctx* c) { return c ? c->err.c_str() : "null context"; }

int dfr_set_fluid(dfr_ctx* c, int64_t n, const float* x, const float* v) {
  return guarded(c, [&]() {
    if (c->initialized) return fail(c, DFR_ERR_STATE, "dfr_set_fluid after dfr_init");
    if (n < 0 || (n > 0 && !x)) return fail(c, DFR_ERR_ARG, "dfr_set_fluid: bad arguments");
    if (n > (int64_t)0x7fffffff) return fail(c, DFR_ERR_ARG, "dfr_set_fluid: too many particles");
    c->fx.assign(x, x + 3 * n);
    c->fv.assign(3 * n, 0.f);
    if (v) c->fv.assign(v, v + 3 * n);
    c->f.n = (int)n;
    return DFR_OK;
  });
}

int dfr_add_boundary(dfr_ctx* c, int64_t n, const float* x0, int is_dynamic, const double x_rb[3],
                     const double q[4], double density) {
  return guarded(c, [&]() {
    if (c->initialized) return fail(c, DFR_ERR_STATE, "dfr_add_boundary after dfr_init");
    if (n < 0 || (n > 0 && !x0) || !x_rb || !q) return fail(c, DFR_ERR_ARG, "dfr_add_boundary: bad arguments");
    if ((int)c->bodies.size() >= DFR_MAX_BODIES) return fail(c, DFR_ERR_ARG, "too many bodies");
    BodyInfo bi;
    bi.dynamic = is_dynamic != 0;
    bi.n = n;
    bi.x0.assign(x0, x0 + 3 * n);
    for (int k = 0; k < 3; ++k) bi.x_rb[k] = x_rb[k];
    for (int k = 0; k < 4; ++k) bi.q[k] = q[k];
    bi.density = density;
    c->bodies.push_back(bi);
    return (int)c->bodies.size() - 1;
  });
}

int dfr_set_init_velocity(dfr_ctx* c, int body, const double v0[3], const double w0[3]) {
  return guarded(c, [&]() {
    if (body < 0 || body >= (int)c->bodies.size() || !v0 || !w0) return fail(c, DFR_ERR_ARG, "bad body");
    BodyInfo& bi = c->bodies[body];
    for (int k = 0; k < 3; ++k) { bi.v0[k] = v0[k]; bi.w0[k] = w0[k]; }
    if (c->initialized) {
      adjoint_release(c);
      HIPCHK(hipMemcpy(&c->d_rig[body].v0[0], v0, 3 * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(&c->d_rig[body].w0[0], w0, 3 * sizeof(double), hipMemcpyHostToDevice));
      for (int k = 0; k < 3; ++k) { c->h_rig_init[body].v0[k] = v0[k]; c->h_rig_init[body].w0[k] = w0[k]; }
    }
    return DFR_OK;
  });
}

int dfr_init(dfr_ctx* c) {
  return guarded(c, [&]() {
    if (c->initialized) return fail(c, DFR_ERR_STATE, "dfr_init called twice");
    if (c->P.use_rigid_contact && c->P.use_gradient_manager && c->bodies.size() > 16)
      return fail(c, DFR_ERR_ARG, "the gradient manager across bodies (rigid contact) holds at most 16 bodies");
    HIPCHK(hipSetDevice(c->device));
    const int nbodies = (int)c->bodies.size();
    // boundary arrays: static bodies first, then dynamic bodies
    int64_t off = 0, nd = 0;
    for (int pass = 0; pass < 2; ++pass)
      for (int k = 0; k < nbodies; ++k) {
        BodyInfo& bi = c->bodies[k];
        if ((pass == 0) == bi.dynamic) continue;
        bi.offset = off;
        off += bi.n;
        if (bi.dynamic) { bi.dyn_first = nd; nd += bi.n; }
        c->body_order.push_back(k);
      }
    c->b.n = (int)off;
    c->b.nd = (int)nd;
    const int nf = c->f.n;
    c->sc.nf = nf;
    c->sc.nb = c->b.n;
    c->sc.nbodies = nbodies;
    // grid bounds from the initial scene
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int64_t i = 0; i < nf; ++i)
      for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], c->fx[3 * i + k]); hi[k] = std::max(hi[k], c->fx[3 * i + k]); }
    std::vector<float4> x0(std::max<int64_t>(off, 1));
    for (int k = 0; k < nbodies; ++k) {
      const BodyInfo& bi = c->bodies[k];
      double R[9];
      {
        const double w = bi.q[0], x = bi.q[1], y = bi.q[2], z = bi.q[3];
        const double tx = 2 * x, ty = 2 * y, tz = 2 * z, twx = tx * w, twy = ty * w, twz = tz * w;
        const double txx = tx * x, txy = ty * x, txz = tz * x, tyy = ty * y, tyz = tz * y, tzz = tz * z;
        const double Rv[9] = {1 - (tyy + tzz), txy - twz, txz + twy, txy + twz, 1 - (txx + tzz), tyz - twx,
                              txz - twy, tyz + twx, 1 - (txx + tyy)};
        std::memcpy(R, Rv, sizeof(R));
      }
      for (int64_t j = 0; j < bi.n; ++j) {
        const float* p = &bi.x0[3 * j];
        x0[bi.offset + j] = make_float4(p[0], p[1], p[2], (float)k);
        for (int a = 0; a < 3; ++a) {
          const double wpos = R[3 * a] * p[0] + R[3 * a + 1] * p[1] + R[3 * a + 2] * p[2] + bi.x_rb[a];
          lo[a] = std::min(lo[a], (float)wpos);
          hi[a] = std::max(hi[a], (float)wpos);
        }
      }
    }
    if (lo[0] > hi[0]) { for (int k = 0; k < 3; ++k) { lo[k] = 0.f; hi[k] = 1.f; } }
    c->grid = make_grid(lo, hi, c->h);
    // allocations
    const size_t NF = (size_t)std::max(nf, 1);
    c->f.pos = dalloc<float4>(c, NF); c->f.vel = dalloc<float4>(c, NF);
    c->f.pos2 = dalloc<float4>(c, NF); c->f.vel2 = dalloc<float4>(c, NF);
    c->f.kappa = dalloc<float>(c, NF); c->f.kappav = dalloc<float>(c, NF);
    c->f.kappa2 = dalloc<float>(c, NF); c->f.kappav2 = dalloc<float>(c, NF);
    c->f.id = dalloc<uint32_t>(c, NF); c->f.id2 = dalloc<uint32_t>(c, NF);
    c->f.rho = dalloc<float>(c, NF); c->f.alpha = dalloc<float>(c, NF);
    c->f.s = dalloc<float4>(c, NF); c->f.badv = dalloc<float>(c, NF); c->f.kf = dalloc<float>(c, NF);
    if (c->P.surface_tension_method != 0 || c->P.viscosity_method != 0) {
      c->acc = dalloc<float4>(c, NF);
      if (c->P.surface_tension_method == 2) c->nrm = dalloc<float4>(c, NF);
      if (c->P.surface_tension_method == 3) { c->he_color = dalloc<float>(c, NF); c->he_gc2 = dalloc<float>(c, NF); }
      set_forces_kernel_attributes();
    }
    if (c->P.cfl_method == 1) c->d_cflmax = dalloc<uint32_t>(c, 1);
    const size_t NB = (size_t)std::max<int64_t>(off, 1);
    if (c->P.use_rigid_contact) {
      c->cvol0 = dalloc<float>(c, NB); c->cdens0 = dalloc<float>(c, NB);
      c->cforce = dalloc<float4>(c, std::max<int64_t>(nd, 1)); c->ctorque = dalloc<float4>(c, std::max<int64_t>(nd, 1));
      if (c->P.use_gradient_manager) {
        const size_t nbb = (size_t)nbodies * nbodies * 78;
        c->mgrJ = dalloc<double>(c, nbb); c->mgrN = dalloc<double>(c, nbb);
        c->cgrec = dalloc<float4>(c, 6 * (size_t)std::max<int64_t>(nd, 1));
      }
    }
    c->b.pos = dalloc<float4>(c, NB); c->b.vel = dalloc<float4>(c, NB); c->b.x0 = dalloc<float4>(c, NB);
    c->b.dyn_idx = dalloc<uint32_t>(c, std::max<int64_t>(nd, 1));
    const uint32_t nk = c->grid.nkeys;
    c->gb.fcount = dalloc<uint32_t>(c, nk); c->gb.fstart = dalloc<uint32_t>(c, nk + 1);
    c->gb.bcount = dalloc<uint32_t>(c, nk); c->gb.bstart = dalloc<uint32_t>(c, nk + 1);
    c->gb.key = dalloc<uint32_t>(c, NF); c->gb.slot = dalloc<uint32_t>(c, NF); c->gb.perm = dalloc<uint32_t>(c, NF);
    c->gb.bkey = dalloc<uint32_t>(c, NB); c->gb.bslot = dalloc<uint32_t>(c, NB); c->gb.bperm = dalloc<uint32_t>(c, NB);
    c->gb.tile_sums = dalloc<uint32_t>(c, 2 * (scan_tile_count(nk) + 1));
    // the sorts expect zero counts and leave them zero
    HIPCHK(hipMemsetAsync(c->gb.fcount, 0, sizeof(uint32_t) * nk, c->stream));
    HIPCHK(hipMemsetAsync(c->gb.bcount, 0, sizeof(uint32_t) * nk, c->stream));
    c->nb.maxb = c->P.max_neighbors;
    c->nb.maxbf = c->P.max_boundary_neighbors;
    c->nb.nbr8 = dalloc<uint4>(c, (size_t)NPACK * NF);
    {
      const size_t NB = (size_t)c->grid.bn[0] * c->grid.bn[1] * c->grid.bn[2];
      c->nb.thdr = dalloc<uint4>(c, NB);
      c->nb.tmap = dalloc<uint32_t>(c, NB * TILE_CAP);
    }
    c->nb.nbrB = dalloc<uint32_t>(c, (size_t)c->nb.maxb * NF);
    c->nb.cnt = dalloc<uint32_t>(c, NF);
    c->nb.bf = dalloc<uint32_t>(c, (size_t)c->nb.maxbf * std::max<int64_t>(nd, 1));
    c->nb.bfcnt = dalloc<uint32_t>(c, std::max<int64_t>(nd, 1));
    c->nb.overflow = dalloc<uint32_t>(c, 4);
    c->sc.maxn = MAXF;
    set_tiled_kernel_attributes();
    c->sc.maxbf = c->nb.maxbf;
    c->d_rig = dalloc<RigidDev>(c, std::max(nbodies, 1));
    c->err_slots = c->P.max_iterations_v + c->P.max_iterations + 2;
    c->d_err = dalloc<double>(c, c->err_slots);
    HIPCHK(hipHostMalloc((void**)&c->h_err, sizeof(double) * c->err_slots, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&c->h_overflow, sizeof(uint32_t) * 4, hipHostMallocDefault));
    c->h_overflow[0] = c->h_overflow[1] = c->h_overflow[2] = c->h_overflow[3] = 0;
    c->d_stop = dalloc<int>(c, 2);
    c->nparts = (int)(c->grid.bn[0] * c->grid.bn[1] * c->grid.bn[2]);
    c->d_part = dalloc<double>(c, c->nparts);
    c->probe_cap = std::max(c->P.max_iterations, c->P.max_iterations_v) + 1;
    HIPCHK(hipHostMalloc((void**)&c->h_probe, sizeof(SolveProbe) * 2 * c->probe_cap,
                         hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer((void**)&c->d_probe, c->h_probe, 0));
    for (auto& e : c->pipe_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // uploads
    HIPCHK(hipMemcpy(c->b.x0, x0.data(), sizeof(float4) * off, hipMemcpyHostToDevice));
    std::vector<uint32_t> dyn(std::max<int64_t>(nd, 1));
    for (int64_t k = 0; k < nd; ++k) dyn[k] = (uint32_t)(off - nd + k);
    HIPCHK(hipMemcpy(c->b.dyn_idx, dyn.data(), sizeof(uint32_t) * dyn.size(), hipMemcpyHostToDevice));
    c->h_rig_init.assign(std::max(nbodies, 1), RigidDev{});
    for (int k = 0; k < nbodies; ++k) fill_rigid_from_body(c, c->bodies[k], c->h_rig_init[k]);
    if (nf > 0) upload_fluid_initial(c);
    place_boundary_and_volumes(c);
    c->initialized = true;
    c->time = 0.f;
    c->step_count = 0;
    return DFR_OK;
  });
}

int dfr_step(dfr_ctx* c, int nsteps) {
  return guarded(c, [&]() {
    if (!c->initialized) return fail(c, DFR_ERR_STATE, "dfr_step before dfr_init");
    HIPCHK(hipSetDevice(c->device));
    AdjState& A = c->adj;
    // A step that fails ends the recording of the rollout (its tapes and
    // checkpoints no longer describe the state), and leaves no step marked as
    // being recorded.
    struct StepGuard {
      dfr_ctx* c;
      bool ok = false;
      ~StepGuard() {
        c->rec = nullptr;
        c->force_iv = c->force_ip = -1;
        if (!ok && c->adj.active) {
          (void)hipStreamSynchronize(c->stream);
          try { adjoint_release(c); } catch (...) {}
        }
      }
    } guard{c};
    for (int s = 0; s < nsteps; ++s) {
      if (A.active && A.k == 0) {  // keep every step's tape in HBM
        A.tapes.emplace_back();
        StepTapeDev& T = A.tapes.back();
        tape_alloc_fixed(c, T);
        tape_swap_in(c, T);
        c->rec = &T;
        do_step(c);
        c->rec = nullptr;
        T.rig_post = arena_copy(c, (const RigidDev*)c->d_rig, c->bodies.size());
      } else {
        if (A.active && A.steps % A.k == 0) {
          if (A.nckpt == A.ckpt.size()) A.ckpt.emplace_back();
          snapshot_save(c, A.ckpt[A.nckpt++]);
        }
        do_step(c);
      }
      if (A.active) {
        A.iv.push_back(c->stats.iterations_v);
        A.ip.push_back(c->stats.iterations);
        A.steps++;
      }
    }
    sync_check(c);
    check_overflow(c);
    guard.ok = true;
    if (c->profiling) drain_timers(c);
    return DFR_OK;
  });
}

int dfr_reset(dfr_ctx* c) {
  return guarded(c, [&]() {
    if (!c->initialized) return fail(c, DFR_ERR_STATE, "dfr_reset before dfr_init");
    adjoint_release(c);
    HIPCHK(hipSetDevice(c->device));
    for (auto& bi : c->bodies) bi.animated = false;
    if (c->f.n > 0) upload_fluid_initial(c);
    place_boundary_and_volumes(c);
    c->time = 0.f;
    set_step_size(c, (float)c->P.dt);  // SimulatorBase::reset (SimulatorBase.cpp:908-910)
    c->step_count = 0;
    c->stats.time = 0.0;
    c->stats.step_count = 0;
    return DFR_OK;
  });
}

static int check_body(const dfr_ctx* c, int body) {
  return (c && c->initialized && body >= 0 && body < (int)c->bodies.size()) ? 0 : 1;
}

int dfr_get_rigid(const dfr_ctx* c, int body, double out[19]) {
  if (check_body(c, body) || !out) return DFR_ERR_ARG;
  RigidDev r;
  if (hipMemcpy(&r, &c->d_rig[body], sizeof(RigidDev), hipMemcpyDeviceToHost) != hipSuccess) return DFR_ERR_HIP;
  for (int k = 0; k < 3; ++k) { out[k] = r.x[k]; out[7 + k] = r.v[k]; out[10 + k] = r.w[k]; out[13 + k] = r.Flast[k]; out[16 + k] = r.Tlast[k]; }
  for (int k = 0; k < 4; ++k) out[3 + k] = r.q[k];
  return DFR_OK;
}

int dfr_set_rigid(dfr_ctx* c, int body, const double in[13]) {
  if (check_body(c, body) || !in) return DFR_ERR_ARG;
  return guarded(c, [&]() {
    adjoint_release(c);
    RigidDev r;
    HIPCHK(hipMemcpy(&r, &c->d_rig[body], sizeof(RigidDev), hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; ++k) { r.x[k] = in[k]; r.v[k] = in[7 + k]; r.w[k] = in[10 + k]; }
    for (int k = 0; k < 4; ++k) r.q[k] = in[3 + k];
    // keep the world inertia consistent with the new rotation
    BodyInfo tmp = c->bodies[body];
    for (int k = 0; k < 3; ++k) tmp.x_rb[k] = r.x[k];
    for (int k = 0; k < 4; ++k) tmp.q[k] = r.q[k];
    RigidDev fresh;
    fill_rigid_from_body(c, tmp, fresh);
    std::memcpy(r.I, fresh.I, sizeof(r.I));
    std::memcpy(r.Iinv, fresh.Iinv, sizeof(r.Iinv));
    HIPCHK(hipMemcpy(&c->d_rig[body], &r, sizeof(RigidDev), hipMemcpyHostToDevice));
    launch_boundary_transform(nullptr, c->b.n, c->b, c->d_rig, c->stream);
    sync_check(c);
    return DFR_OK;
  });
}

int dfr_get_body(const dfr_ctx* c, int body, double* mass, int* dynamic, int* animated) {
  if (check_body(c, body)) return DFR_ERR_ARG;
  if (mass && hipMemcpy(mass, &c->d_rig[body].mass, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return DFR_ERR_HIP;
  if (dynamic) *dynamic = c->bodies[body].dynamic ? 1 : 0;
  if (animated) *animated = c->bodies[body].animated ? 1 : 0;
  return DFR_OK;
}

int dfr_set_gradient_mode(dfr_ctx* c, int mode) {
  if (!c || mode < 0 || mode > 2) return DFR_ERR_ARG;
  c->P.gradient_mode = mode;
  return DFR_OK;
}

int dfr_set_compute_gradient(dfr_ctx* c, int on) {
  if (!c) return DFR_ERR_ARG;
  c->P.compute_gradient = on ? 1 : 0;
  return DFR_OK;
}

int dfr_get_jacobians(const dfr_ctx* c, int body, double out[78]) {
  if (check_body(c, body) || !out) return DFR_ERR_ARG;
  return hipMemcpy(out, &c->d_rig[body].J[0], 78 * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess ? DFR_OK
                                                                                                       : DFR_ERR_HIP;
}

int dfr_get_gradient_manager(const dfr_ctx* c, int R, int RR, double jac[78], double net[78]) {
  if (check_body(c, R) || check_body(c, RR)) return DFR_ERR_ARG;
  if (jac) std::fill(jac, jac + 78, 0.0);
  if (net) std::fill(net, net + 78, 0.0);
  if (!c->bodies[R].dynamic || !c->bodies[RR].dynamic) return DFR_OK;  // a static body has no entries
  if (c->mgrJ) {  // the table across bodies (rigid contact solver with the manager)
    const size_t o = ((size_t)R * c->bodies.size() + RR) * 78;
    if (jac && hipMemcpy(jac, c->mgrJ + o, 78 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return DFR_ERR_HIP;
    if (net && hipMemcpy(net, c->mgrN + o, 78 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return DFR_ERR_HIP;
    return DFR_OK;
  }
  // entries between different bodies stay zero without the rigid contact solver
  if (R != RR) return DFR_OK;
  if (jac && hipMemcpy(jac, &c->d_rig[R].Jm[0], 78 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return DFR_ERR_HIP;
  if (net && hipMemcpy(net, &c->d_rig[R].mnet[0], 78 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return DFR_ERR_HIP;
  return DFR_OK;
}

int dfr_get_net_gradients(const dfr_ctx* c, int body, double out[78]) {
  if (check_body(c, body) || !out) return DFR_ERR_ARG;
  return hipMemcpy(out, &c->d_rig[body].netlast[0], 78 * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess
             ? DFR_OK
             : DFR_ERR_HIP;
}

int dfr_reset_gradient(dfr_ctx* c) {
  return guarded(c, [&]() {
    if (!c->initialized) return fail(c, DFR_ERR_STATE, "not initialised");
    for (int k = 0; k < (int)c->bodies.size(); ++k) {
      double J[78] = {0}, Z[78] = {0};
      for (int d = 0; d < 3; ++d) { J[18 + 4 * d] = 1.0; J[45 + 4 * d] = 1.0; }
      HIPCHK(hipMemcpy(&c->d_rig[k].J[0], J, sizeof(J), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(&c->d_rig[k].Jm[0], J, sizeof(J), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(&c->d_rig[k].mnet[0], Z, sizeof(Z), hipMemcpyHostToDevice));
    }
    reset_manager_multi(c);
    return DFR_OK;
  });
}

int dfr_get_fluid(const dfr_ctx* c, float* x, float* v) {
  if (!c || !c->initialized) return DFR_ERR_STATE;
  try {
    if (x) download_unpermute<float4>(c, c->f.pos, 4, 3, x);
    if (v) download_unpermute<float4>(c, c->f.vel, 4, 3, v);
  } catch (const HipError&) {
    return DFR_ERR_HIP;
  }
  return DFR_OK;
}

int dfr_get_field(const dfr_ctx* c, const char* name, float* out) {
  if (!c || !c->initialized || !name || !out) return DFR_ERR_ARG;
  const std::string n(name);
  try {
    if (n == "density") download_unpermute<float>(c, c->f.rho, 1, 1, out);
    else if (n == "factor") download_unpermute<float>(c, c->f.alpha, 1, 1, out);
    else if (n == "density_adv") download_unpermute<float>(c, c->f.badv, 1, 1, out);
    else if (n == "kappa") download_unpermute<float>(c, c->f.kappa, 1, 1, out);
    else if (n == "kappa_v") download_unpermute<float>(c, c->f.kappav, 1, 1, out);
    else if (n == "sum_grad_p_k") download_unpermute<float4>(c, c->f.s, 4, 3, out);
    else if (n == "accel" && c->acc) download_unpermute<float4>(c, c->acc, 4, 3, out);
    else if (n == "normal" && c->nrm) download_unpermute<float4>(c, c->nrm, 4, 3, out);
    else if (n == "color" && c->he_color) download_unpermute<float>(c, c->he_color, 1, 1, out);
    else if (n == "gradC2" && c->he_gc2) download_unpermute<float>(c, c->he_gc2, 1, 1, out);
    else return DFR_ERR_ARG;
  } catch (const HipError&) {
    return DFR_ERR_HIP;
  }
  return DFR_OK;
}

int dfr_set_field(dfr_ctx* c, const char* name, const float* in) {
  if (!c || !c->initialized || !name || !in) return DFR_ERR_ARG;
  return guarded(c, [&]() {
    adjoint_release(c);
    const std::string n(name);
    const int64_t nf = c->f.n;
    std::vector<uint32_t> id(nf);
    HIPCHK(hipMemcpy(id.data(), c->f.id, nf * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (n == "x" || n == "v") {  // 3 per particle; the fourth component is kept
      float4* d4 = (n == "x") ? c->f.pos : c->f.vel;
      std::vector<float4> b4(nf);
      HIPCHK(hipMemcpy(b4.data(), d4, nf * sizeof(float4), hipMemcpyDeviceToHost));
      for (int64_t k = 0; k < nf; ++k) {
        const float* s = in + 3 * (size_t)id[k];
        b4[k].x = s[0]; b4[k].y = s[1]; b4[k].z = s[2];
      }
      HIPCHK(hipMemcpy(d4, b4.data(), nf * sizeof(float4), hipMemcpyHostToDevice));
      return DFR_OK;
    }
    float* dst = nullptr;
    if (n == "kappa") dst = c->f.kappa;
    else if (n == "kappa_v") dst = c->f.kappav;
    else return fail(c, DFR_ERR_ARG, "unknown field");
    std::vector<float> buf(nf);
    for (int64_t k = 0; k < nf; ++k) buf[k] = in[id[k]];
    HIPCHK(hipMemcpy(dst, buf.data(), nf * sizeof(float), hipMemcpyHostToDevice));
    return DFR_OK;
  });
}

int dfr_set_time(dfr_ctx* c, double time, int64_t step_count) {
  if (!c || !c->initialized || step_count < 0) return DFR_ERR_ARG;
  return guarded(c, [&]() {
    adjoint_release(c);
    c->time = (float)time;
    c->step_count = step_count;
    c->stats.time = (double)c->time;
    c->stats.step_count = step_count;
    return DFR_OK;
  });
}

int dfr_get_boundary_field(const dfr_ctx* c, int body, const char* name, float* out) {
  if (check_body(c, body) || !name || !out) return DFR_ERR_ARG;
  const BodyInfo& bi = c->bodies[body];
  std::vector<float4> buf(std::max<int64_t>(bi.n, 1));
  const std::string n(name);
  const float4* src = (n == "v") ? c->b.vel : c->b.pos;
  if (bi.n > 0 &&
      hipMemcpy(buf.data(), src + bi.offset, bi.n * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess)
    return DFR_ERR_HIP;
  if (n == "volume") {
    for (int64_t j = 0; j < bi.n; ++j) out[j] = buf[j].w;
  } else if (n == "x" || n == "v") {
    for (int64_t j = 0; j < bi.n; ++j) { out[3 * j] = buf[j].x; out[3 * j + 1] = buf[j].y; out[3 * j + 2] = buf[j].z; }
  } else {
    return DFR_ERR_ARG;
  }
  return DFR_OK;
}

int dfr_get_stats(const dfr_ctx* c, dfr_stats* out) {
  if (!c || !out) return DFR_ERR_ARG;
  *out = c->stats;
  return DFR_OK;
}

int dfr_reset_stats(dfr_ctx* c) {
  if (!c) return DFR_ERR_ARG;
  c->stats.total_iterations = 0;
  c->stats.total_iterations_v = 0;
  c->stats.steps_since_reset = 0;
  for (int k = 0; k < TC_COUNT; ++k) { c->t_sum[k] = 0; c->t_cnt[k] = 0; }
  return DFR_OK;
}

int dfr_set_halo_capacity(dfr_ctx* c, int slots) {
  if (!c) return DFR_ERR_ARG;
  c->halo_cap_forced = slots < 0 ? -1 : std::min(slots, TILE_CAP);
  c->nb.cap = slots < 0 ? TILE_CAP : c->halo_cap_forced;
  if (slots < 0 && c->h_overflow) update_halo_cap(c);
  return DFR_OK;
}

int dfr_halo_stats(const dfr_ctx* c, int64_t out[3]) {
  if (!c || !out) return DFR_ERR_ARG;
  out[0] = c->halo_max;
  out[1] = c->nb.cap;
  out[2] = TILE_CAP;
  return DFR_OK;
}

int dfr_neighbor_stats(const dfr_ctx* c, double out[3]) {
  if (!c || !c->initialized || !out) return DFR_ERR_ARG;
  const int64_t nf = c->f.n, nd = c->b.nd;
  std::vector<uint32_t> cnt(std::max<int64_t>(nf, 1)), bfc(std::max<int64_t>(nd, 1));
  if (nf && hipMemcpy(cnt.data(), c->nb.cnt, nf * 4, hipMemcpyDeviceToHost) != hipSuccess) return DFR_ERR_HIP;
  if (nd && hipMemcpy(bfc.data(), c->nb.bfcnt, nd * 4, hipMemcpyDeviceToHost) != hipSuccess) return DFR_ERR_HIP;
  double sf = 0, sb = 0, sbf = 0;
  for (int64_t i = 0; i < nf; ++i) { sf += cnt[i] & 0xffff; sb += cnt[i] >> 16; }
  for (int64_t k = 0; k < nd; ++k) sbf += bfc[k];
  out[0] = nf ? sf / nf : 0.0;
  out[1] = nf ? sb / nf : 0.0;
  out[2] = nd ? sbf / nd : 0.0;
  return DFR_OK;
}

int64_t dfr_count_1ring(const dfr_ctx* c) {
  if (!c || !c->initialized) return DFR_ERR_ARG;
  const int64_t nf = c->f.n, nb = c->b.n;
  if (nf == 0 || c->step_count == 0) return 0;  // no lists since dfr_init / dfr_reset
  // owner of each global sample index that belongs to a dynamic body (-1: static)
  std::vector<int32_t> owner(std::max<int64_t>(nb, 1), -1);
  int ndyn = 0;
  for (size_t k = 0; k < c->bodies.size(); ++k) {
    const BodyInfo& bi = c->bodies[k];
    if (!bi.dynamic) continue;
    for (int64_t s = 0; s < bi.n; ++s) owner[bi.offset + s] = ndyn;
    ++ndyn;
  }
  if (ndyn == 0) return 0;
  const int slots = c->nb.maxb;
  std::vector<uint32_t> cnt(nf), nbr((size_t)slots * nf);
  if (hipMemcpy(cnt.data(), c->nb.cnt, nf * 4, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(nbr.data(), c->nb.nbrB, (size_t)slots * nf * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return DFR_ERR_HIP;
  std::vector<int64_t> seen(ndyn, -1);  // last particle that counted each body
  int64_t pairs = 0;
  for (int64_t i = 0; i < nf; ++i) {
    const uint32_t n = cnt[i] >> 16;
    for (uint32_t s = 0; s < n; ++s) {
      const int32_t o = owner[nbr[(size_t)s * nf + i]];
      if (o >= 0 && seen[o] != i) { seen[o] = i; ++pairs; }
    }
  }
  return pairs;
}

int64_t dfr_get_neighbors(const dfr_ctx* c, int kind, uint32_t* counts, uint32_t* lists, int64_t cap) {
  if (!c || !c->initialized || !counts || !lists) return DFR_ERR_ARG;
  const int64_t nf = c->f.n, nd = c->b.nd;
  std::vector<uint32_t> id(std::max<int64_t>(nf, 1));
  if (nf && hipMemcpy(id.data(), c->f.id, nf * 4, hipMemcpyDeviceToHost) != hipSuccess) return DFR_ERR_HIP;
  if (kind == 0 || kind == 1) {
    std::vector<uint32_t> cnt(std::max<int64_t>(nf, 1));
    const int slots = (kind == 0) ? MAXF : c->nb.maxb;
    std::vector<uint32_t> nbr((size_t)slots * std::max<int64_t>(nf, 1));
    if (nf && hipMemcpy(cnt.data(), c->nb.cnt, nf * 4, hipMemcpyDeviceToHost) != hipSuccess) return DFR_ERR_HIP;
    if (nf && kind == 0) {
      uint32_t* tmp = nullptr;
      if (hipMalloc(&tmp, (size_t)slots * nf * 4) != hipSuccess) return DFR_ERR_HIP;
      dfr_ctx* cc = const_cast<dfr_ctx*>(c);
      launch_expand_lists(cc->grid, cc->gb, cc->f, cc->nb, tmp, cc->stream);
      const bool ok = hipStreamSynchronize(c->stream) == hipSuccess &&
                      hipMemcpy(nbr.data(), tmp, (size_t)slots * nf * 4, hipMemcpyDeviceToHost) == hipSuccess;
      (void)hipFree(tmp);
      if (!ok) return DFR_ERR_HIP;
    } else if (nf) {
      if (hipMemcpy(nbr.data(), c->nb.nbrB, (size_t)slots * nf * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return DFR_ERR_HIP;
    }
    // rows in particle-id order
    std::vector<int64_t> row_of(nf);
    for (int64_t k = 0; k < nf; ++k) row_of[id[k]] = k;
    int64_t w = 0;
    for (int64_t r = 0; r < nf; ++r) {
      const int64_t k = row_of[r];
      const uint32_t n = (kind == 0) ? (cnt[k] & 0xffff) : (cnt[k] >> 16);
      counts[r] = n;
      for (uint32_t s = 0; s < n; ++s) {
        if (w >= cap) return DFR_ERR_CAPACITY;
        const uint32_t j = nbr[(size_t)s * nf + k];
        lists[w++] = (kind == 0) ? id[j] : j;
      }
    }
    return nf;
  }
  if (kind == 2) {
    std::vector<uint32_t> cnt(std::max<int64_t>(nd, 1));
    std::vector<uint32_t> bf((size_t)c->nb.maxbf * std::max<int64_t>(nd, 1));
    if (nd && hipMemcpy(cnt.data(), c->nb.bfcnt, nd * 4, hipMemcpyDeviceToHost) != hipSuccess) return DFR_ERR_HIP;
    if (nd && hipMemcpy(bf.data(), c->nb.bf, (size_t)c->nb.maxbf * nd * 4, hipMemcpyDeviceToHost) != hipSuccess)
      return DFR_ERR_HIP;
    int64_t w = 0;
    for (int64_t k = 0; k < nd; ++k) {
      counts[k] = cnt[k];
      for (uint32_t s = 0; s < cnt[k]; ++s) {
        if (w >= cap) return DFR_ERR_CAPACITY;
        lists[w++] = id[bf[(size_t)s * nd + k]];
      }
    }
    return nd;
  }
  return DFR_ERR_ARG;
}

int dfr_set_profiling(dfr_ctx* c, int on) {
  if (!c) return DFR_ERR_ARG;
  c->profiling = on != 0;
  c->prof_mask = on ? 0xffffffffu : 0u;
  return DFR_OK;
}

int dfr_set_profiling_classes(dfr_ctx* c, uint32_t mask) {
  if (!c) return DFR_ERR_ARG;
  c->prof_mask = mask;
  c->profiling = mask != 0;
  return DFR_OK;
}

int dfr_kernel_times(const dfr_ctx* c, int n, char names[][32], double* avg_ms, int64_t* launches) {
  if (!c) return DFR_ERR_ARG;
  for (int k = 0; k < TC_COUNT && k < n; ++k) {
    if (names) { std::strncpy(names[k], kTimerNames[k], 31); names[k][31] = 0; }
    if (avg_ms) avg_ms[k] = c->t_cnt[k] ? c->t_sum[k] / (double)c->t_cnt[k] : 0.0;
    if (launches) launches[k] = c->t_cnt[k];
  }
  return TC_COUNT;
}

int dfr_adjoint_begin(dfr_ctx* c, int checkpoint_every) {
  return guarded(c, [&]() {
    if (!c->initialized) return fail(c, DFR_ERR_STATE, "dfr_adjoint_begin before dfr_init");
    if (c->P.use_rigid_contact)
      return fail(c, DFR_ERR_STATE,
                  "the exact adjoint does not cover the rigid contact solver (its reference-mode gradient runs)");
    if (c->P.surface_tension_method != 0 || c->P.cfl_method != 0)
      return fail(c, DFR_ERR_STATE,
                  "the exact adjoint covers viscosity but no other non-pressure force and no CFL step size "
                  "(surface_tension_method and cfl_method must be 0)");
    HIPCHK(hipSetDevice(c->device));
    adjoint_release(c);
    adjoint_buffers(c);
    c->adj.k = checkpoint_every > 0 ? checkpoint_every : 0;
    c->adj.active = true;
    return DFR_OK;
  });
}

int dfr_adjoint_backward(dfr_ctx* c, const double* seed, double* grad) {
  return guarded(c, [&]() {
    AdjState& A = c->adj;
    if (!A.active || A.steps == 0) return fail(c, DFR_ERR_STATE, "dfr_adjoint_backward without a recorded rollout");
    if (!seed || !grad) return fail(c, DFR_ERR_ARG, "dfr_adjoint_backward: null seed or grad");
    HIPCHK(hipSetDevice(c->device));
    const int nbodies = (int)c->bodies.size();
    const dfr_stats saved_stats = c->stats;
    const bool stored = A.k == 0;
    A.light = false;
    if (!stored) snapshot_save(c, A.final_state);
    std::vector<AdjBodyDev> hb(std::max(nbodies, 1));
    std::memset(hb.data(), 0, sizeof(AdjBodyDev) * hb.size());
    for (int b = 0; b < nbodies; ++b) {
      for (int k = 0; k < 3; ++k) hb[b].x[k] = seed[7 * b + k];
      for (int k = 0; k < 4; ++k) hb[b].q[k] = seed[7 * b + 3 + k];
    }
    HIPCHK(hipMemcpyAsync(A.body, hb.data(), sizeof(AdjBodyDev) * nbodies, hipMemcpyHostToDevice, c->stream));
    const size_t nf = (size_t)c->f.n;
    HIPCHK(hipMemsetAsync(A.xb, 0, sizeof(float4) * nf, c->stream));
    HIPCHK(hipMemsetAsync(A.vb, 0, sizeof(float4) * nf, c->stream));
    HIPCHK(hipMemsetAsync(A.kb, 0, sizeof(float) * nf, c->stream));
    HIPCHK(hipMemsetAsync(A.kvb, 0, sizeof(float) * nf, c->stream));
    struct ReplayFlag {
      dfr_ctx* c;
      explicit ReplayFlag(dfr_ctx* x) : c(x) { c->replay = true; }
      ~ReplayFlag() { c->replay = false; c->rec = nullptr; c->force_iv = c->force_ip = -1; }
    } flag(c);
    const int64_t S = A.steps, k = std::max(A.k, 1);
    if (stored) {
      for (int64_t s = S - 1; s >= 0; --s) reverse_step(c, A.tapes[s]);
      sync_check(c);
    }
    for (int64_t seg = stored ? -1 : (S - 1) / k; seg >= 0; --seg) {
      const int64_t s0 = seg * k, s1 = std::min(s0 + k, S);
      if (reverse_segment_kept(c, seg, s0, s1)) {
        sync_check(c);
        check_overflow(c);
        continue;
      }
      // the segment's tapes do not fit: keep its states instead, and re-run
      // every step once more with recording right before reversing it
      if ((int64_t)A.seg.size() < k) A.seg.resize(k);
      if (s1 - s0 > 1) {  // states at the start of steps s0+1 .. s1-1
        snapshot_restore(c, A.ckpt[seg]);
        for (int64_t s = s0; s < s1 - 1; ++s) {
          replay_step(c, s, false);
          snapshot_save(c, A.seg[s - s0 + 1]);
        }
      }
      for (int64_t s = s1 - 1; s >= s0; --s) {
        snapshot_restore(c, s == s0 ? A.ckpt[seg] : A.seg[s - s0]);
        replay_step(c, s, true);
        live_view(c, A.tape);
        reverse_step(c, A.tape);
      }
      sync_check(c);
      check_overflow(c);
    }
    if (!stored) {
      tape_swap_out(c);
      A.arena.reset();
      snapshot_restore(c, A.final_state);
    }
    HIPCHK(hipMemcpyAsync(hb.data(), A.body, sizeof(AdjBodyDev) * nbodies, hipMemcpyDeviceToHost, c->stream));
    sync_check(c);
    if (c->profiling) drain_timers(c);
    bool finite = true;
    A.start_grad.assign(13 * (size_t)nbodies, 0.0);
    for (int b = 0; b < nbodies; ++b) {
      double* o = A.start_grad.data() + 13 * (size_t)b;
      for (int q = 0; q < 3; ++q) {
        grad[6 * b + q] = hb[b].v0[q];
        grad[6 * b + 3 + q] = hb[b].w0[q];
        o[q] = hb[b].x[q]; o[7 + q] = hb[b].v[q]; o[10 + q] = hb[b].w[q];
      }
      for (int q = 0; q < 4; ++q) o[3 + q] = hb[b].q[q];
      for (int q = 0; q < 13; ++q) finite = finite && std::isfinite(o[q]);
      for (int q = 0; q < 6; ++q) finite = finite && std::isfinite(grad[6 * b + q]);
    }
    c->stats = saved_stats;
    if (!finite)
      return fail(c, DFR_ERR_RANGE,
                  "the exact gradient left the FP32 range: it grows exponentially once fluid and body interact "
                  "strongly; use a shorter window or the forward-mode gradient");
    return DFR_OK;
  });
}

int dfr_state_save(dfr_ctx* c, int slot) {
  if (!c || slot < 0 || slot > 4095) return DFR_ERR_ARG;
  return guarded(c, [&]() {
    if (!c->initialized) return fail(c, DFR_ERR_STATE, "dfr_state_save before dfr_init");
    HIPCHK(hipSetDevice(c->device));
    AdjState& A = c->adj;
    if ((int)A.user.size() <= slot) A.user.resize(slot + 1);
    snapshot_save(c, A.user[slot]);
    sync_check(c);
    return DFR_OK;
  });
}

int dfr_state_load(dfr_ctx* c, int slot) {
  if (!c || slot < 0) return DFR_ERR_ARG;
  return guarded(c, [&]() {
    AdjState& A = c->adj;
    if (slot >= (int)A.user.size() || !A.user[slot].block)
      return fail(c, DFR_ERR_ARG, "dfr_state_load: no state saved under this slot");
    HIPCHK(hipSetDevice(c->device));
    adjoint_release(c);
    snapshot_restore(c, A.user[slot]);
    sync_check(c);
    c->stats.time = c->time;
    c->stats.step_count = c->step_count;
    return DFR_OK;
  });
}

int dfr_state_free(dfr_ctx* c, int slot) {
  if (!c || slot < 0) return DFR_ERR_ARG;
  return guarded(c, [&]() {
    AdjState& A = c->adj;
    if (slot < (int)A.user.size() && A.user[slot].block) {
      HIPCHK(hipSetDevice(c->device));
      HIPCHK(hipStreamSynchronize(c->stream));
      snapshot_free(A.user[slot]);
    }
    return DFR_OK;
  });
}

int dfr_adjoint_segment_memory(dfr_ctx* c, int64_t bytes) {
  if (!c || bytes < -1) return DFR_ERR_ARG;
  c->adj.seg_cap = bytes;
  return DFR_OK;
}

int dfr_adjoint_arena_limit(dfr_ctx* c, int64_t bytes) {
  if (!c || bytes < 0) return DFR_ERR_ARG;
  c->adj.arena.limit = (size_t)bytes;
  return DFR_OK;
}

int64_t dfr_adjoint_tape_bytes(const dfr_ctx* c) {
  if (!c) return DFR_ERR_ARG;
  return (int64_t)c->adj.arena.used();
}

int dfr_adjoint_start_gradient(const dfr_ctx* c, double* out) {
  if (!c || !out) return DFR_ERR_ARG;
  if (c->adj.start_grad.size() != 13 * c->bodies.size()) return DFR_ERR_STATE;
  std::memcpy(out, c->adj.start_grad.data(), sizeof(double) * c->adj.start_grad.size());
  return DFR_OK;
}

int dfr_adjoint_end(dfr_ctx* c) {
  return guarded(c, [&]() {
    HIPCHK(hipSetDevice(c->device));
    if (c->stream) HIPCHK(hipStreamSynchronize(c->stream));
    adjoint_free(c);
    return DFR_OK;
  });
}

int64_t dfr_adjoint_iterations(const dfr_ctx* c, int* iv, int* ip, int64_t cap) {
  if (!c) return DFR_ERR_ARG;
  const int64_t n = c->adj.steps;
  for (int64_t s = 0; s < n && s < cap; ++s) {
    if (iv) iv[s] = c->adj.iv[s];
    if (ip) ip[s] = c->adj.ip[s];
  }
  return n;
}

int dfr_set_fixed_iterations(dfr_ctx* c, int n, const int* iv, const int* ip) {
  if (!c || n < 0 || (n > 0 && (!iv || !ip))) return DFR_ERR_ARG;
  c->fixed_iv.assign(iv, iv + n);
  c->fixed_ip.assign(ip, ip + n);
  c->fixed_pos = 0;
  return DFR_OK;
}

int dfr_mlp_capture(dfr_ctx* c, int64_t capacity, double fraction, uint64_t seed) {
  return guarded(c, [&]() {
    if (!c->initialized) return fail(c, DFR_ERR_STATE, "dfr_mlp_capture before dfr_init");
    if (capacity < 0 || fraction < 0.0 || fraction > 1.0) return fail(c, DFR_ERR_ARG, "dfr_mlp_capture: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    mlp_alloc(c);
    MlpState& M = c->mlp;
    if (capacity == 0) { M.capturing = false; return DFR_OK; }
    if (capacity > (int64_t)0xffffffffu) return fail(c, DFR_ERR_ARG, "dfr_mlp_capture: capacity too large");
    if (capacity != M.cap) {
      M.capturing = false;
      HIPCHK(hipStreamSynchronize(c->stream));
      for (void* p : {(void*)M.feat, (void*)M.targ, (void*)M.rows})
        if (p) { (void)hipFree(p); c->allocations.erase(std::find(c->allocations.begin(), c->allocations.end(), p)); }
      M.feat = dalloc<float>(c, (size_t)capacity * 6);
      M.targ = dalloc<float>(c, (size_t)capacity * 2);
      M.rows = dalloc<uint32_t>(c, (size_t)capacity);
      M.cap = capacity;
    }
    // a capture already running with this capacity keeps its rows (appends)
    if (!(M.capturing && capacity == M.cap)) HIPCHK(hipMemsetAsync(M.count, 0, sizeof(uint32_t) * 2, c->stream));
    M.threshold = fraction >= 1.0 ? 0xffffffffu : (uint32_t)(fraction * 4294967296.0);
    M.salt = seed * 0xBF58476D1CE4E5B9ull + 1;
    M.capturing = true;
    return DFR_OK;
  });
}

int64_t dfr_mlp_dataset_size(dfr_ctx* c) {
  if (!c) return DFR_ERR_ARG;
  return guarded(c, [&]() { return (int)mlp_dataset_size(c); });
}

int64_t dfr_mlp_dataset(dfr_ctx* c, float* feat, float* targ, int64_t cap) {
  if (!c) return DFR_ERR_ARG;
  return guarded(c, [&]() {
    const uint32_t n = mlp_dataset_size(c);
    const size_t k = (size_t)std::min<int64_t>(n, cap);
    if (feat && k) HIPCHK(hipMemcpy(feat, c->mlp.feat, sizeof(float) * 6 * k, hipMemcpyDeviceToHost));
    if (targ && k) HIPCHK(hipMemcpy(targ, c->mlp.targ, sizeof(float) * 2 * k, hipMemcpyDeviceToHost));
    return (int)n;
  });
}

int dfr_mlp_init_params(dfr_ctx* c, uint64_t seed) {
  return guarded(c, [&]() {
    if (!c->initialized) return fail(c, DFR_ERR_STATE, "dfr_mlp_init_params before dfr_init");
    HIPCHK(hipSetDevice(c->device));
    mlp_alloc(c);
    mlp_init_params(c, seed);
    sync_check(c);
    return DFR_OK;
  });
}

int dfr_mlp_set_params(dfr_ctx* c, const float* params, const float* norm) {
  return guarded(c, [&]() {
    if (!c->initialized || !params || !norm) return fail(c, DFR_ERR_ARG, "dfr_mlp_set_params: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    mlp_alloc(c);
    MlpNorm& N = c->mlp.norm;
    for (int k = 0; k < 6; ++k) { N.mu_in[k] = norm[k]; N.is_in[k] = 1.f / norm[6 + k]; }
    for (int k = 0; k < 2; ++k) { N.mu_out[k] = norm[12 + k]; N.sd_out[k] = norm[14 + k]; }
    c->mlp.norm_set = true;
    mlp_upload_params(c, std::vector<float>(params, params + MLP_NP));
    sync_check(c);
    return DFR_OK;
  });
}

int dfr_mlp_get_params(dfr_ctx* c, float* params, float* norm) {
  return guarded(c, [&]() {
    if (!c->mlp.has_params) return fail(c, DFR_ERR_STATE, "no network parameters");
    if (params) HIPCHK(hipMemcpy(params, c->mlp.P, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost));
    if (norm) {
      const MlpNorm& N = c->mlp.norm;
      for (int k = 0; k < 6; ++k) { norm[k] = N.mu_in[k]; norm[6 + k] = 1.f / N.is_in[k]; }
      for (int k = 0; k < 2; ++k) { norm[12 + k] = N.mu_out[k]; norm[14 + k] = N.sd_out[k]; }
    }
    return DFR_OK;
  });
}

int dfr_mlp_train(dfr_ctx* c, int iters, int batch, double lr, uint64_t seed, double* losses) {
  return guarded(c, [&]() {
    if (!c->initialized || iters <= 0 || batch <= 0 || lr <= 0.0) return fail(c, DFR_ERR_ARG, "dfr_mlp_train: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    mlp_alloc(c);
    MlpState& M = c->mlp;
    const uint32_t n = mlp_dataset_size(c);
    if (n == 0) return fail(c, DFR_ERR_STATE, "dfr_mlp_train: no captured training pairs");
    if (!M.norm_set) mlp_fit_norm(c, n);
    if (!M.has_params) mlp_init_params(c, seed);
    if (M.batch_cap < batch) { M.batch = dalloc<uint32_t>(c, batch); M.batch_cap = batch; }
    if (M.losses_cap < iters) { M.losses = dalloc<double>(c, iters); M.losses_cap = iters; }
    HIPCHK(hipMemsetAsync(M.losses, 0, sizeof(double) * iters, c->stream));
    for (int it = 0; it < iters; ++it) {
      Timed t(c, TC_MLP);
      launch_mlp_batch(M.batch, batch, seed * 0x9E3779B97F4A7C15ull + (uint64_t)(M.adam_t + 1) * 0x632BE59BD9B4E019ull, n,
                       c->stream);
      launch_mlp_train_batch(M.frag, M.P, M.norm, M.feat, M.targ, M.batch, batch, M.G, M.losses + it, c->stream);
      const double cosw = 0.5 * (1.0 + std::cos(M_PI * (double)it / (double)iters));
      launch_mlp_adam(M.P, M.G, M.m, M.v, (float)(lr * (0.1 + 0.9 * cosw)), ++M.adam_t, c->stream);
      launch_mlp_pack(M.P, M.frag, c->stream);
    }
    std::vector<double> h(iters);
    HIPCHK(hipMemcpyAsync(h.data(), M.losses, sizeof(double) * iters, hipMemcpyDeviceToHost, c->stream));
    sync_check(c);
    if (c->profiling) drain_timers(c);
    if (losses)
      for (int it = 0; it < iters; ++it) losses[it] = h[it] / (2.0 * batch);
    return DFR_OK;
  });
}

int dfr_mlp_enable(dfr_ctx* c, int on) {
  return guarded(c, [&]() {
    if (on && !c->mlp.has_params) return fail(c, DFR_ERR_STATE, "dfr_mlp_enable: no trained network");
    adjoint_release(c);
    c->mlp.enabled = on != 0;
    return DFR_OK;
  });
}

int dfr_mlp_predict(dfr_ctx* c, int64_t n, const float* feat, float* out) {
  return guarded(c, [&]() {
    if (!c->mlp.has_params || n < 0 || (n > 0 && (!feat || !out))) return fail(c, DFR_ERR_ARG, "dfr_mlp_predict: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    DevTmp<float> f(6 * n), o(2 * n);
    HIPCHK(hipMemcpy(f.p, feat, sizeof(float) * 6 * n, hipMemcpyHostToDevice));
    launch_mlp_infer_array(c->mlp.frag, c->mlp.P, c->mlp.norm, f.p, (int)n, o.p, c->stream);
    sync_check(c);
    HIPCHK(hipMemcpy(out, o.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost));
    return DFR_OK;
  });
}

int dfr_mlp_loss_grad(dfr_ctx* c, int64_t n, const float* feat, const float* targ, float* grad, double* loss) {
  return guarded(c, [&]() {
    if (!c->mlp.has_params || n <= 0 || !feat || !targ || !grad) return fail(c, DFR_ERR_ARG, "dfr_mlp_loss_grad: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    MlpState& M = c->mlp;
    DevTmp<float> f(6 * n), t(2 * n), g(MLP_NP);
    DevTmp<double> l(1);
    HIPCHK(hipMemcpy(f.p, feat, sizeof(float) * 6 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t.p, targ, sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(g.p, 0, sizeof(float) * MLP_NP, c->stream));
    HIPCHK(hipMemsetAsync(l.p, 0, sizeof(double), c->stream));
    launch_mlp_train_batch(M.frag, M.P, M.norm, f.p, t.p, nullptr, (int)n, g.p, l.p, c->stream);
    sync_check(c);
    HIPCHK(hipMemcpy(grad, g.p, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost));
    if (loss) HIPCHK(hipMemcpy(loss, l.p, sizeof(double), hipMemcpyDeviceToHost));
    return DFR_OK;
  });
}

int dfr_mlp_loss_grad_noalloc(dfr_ctx* c, int64_t n, const float* feat, const float* targ, float* grad,
                              double* loss) {
  return guarded(c, [&]() {
    if (!c->mlp.has_params || n <= 0 || !feat || !targ || !grad)
      return fail(c, DFR_ERR_ARG, "dfr_mlp_loss_grad_noalloc: bad arguments");
    if (n > (int64_t)NA_ROWS) {
      char buf[160];
      std::snprintf(buf, sizeof(buf), "dfr_mlp_loss_grad_noalloc: %lld rows exceed the fixed scratch of %d rows",
                    (long long)n, (int)DFR_NOALLOC_MAX_ROWS);
      return fail(c, DFR_ERR_CAPACITY, buf);
    }
    HIPCHK(hipSetDevice(c->device));
    unsigned char* base = noalloc_base(c->device);
    if (!base) return fail(c, DFR_ERR_STATE, "dfr_mlp_loss_grad_noalloc: no scratch address on this device");
    float* f = reinterpret_cast<float*>(base + NA_FEAT);
    float* t = reinterpret_cast<float*>(base + NA_TARG);
    float* g = reinterpret_cast<float*>(base + NA_GRAD);
    double* l = reinterpret_cast<double*>(base + NA_LOSS);
    const MlpState& M = c->mlp;
    HIPCHK(hipMemcpyAsync(f, feat, sizeof(float) * 6 * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(t, targ, sizeof(float) * 2 * n, hipMemcpyHostToDevice, c->stream));
    // the previous call's gradient and loss stay in the array: clear them
    hipLaunchKernelGGL(k_clear_words, dim3((MLP_NP + 255) / 256), dim3(256), 0, c->stream,
                       reinterpret_cast<uint32_t*>(g), MLP_NP);
    hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(64), 0, c->stream, reinterpret_cast<uint32_t*>(l), 2);
    launch_mlp_train_batch(M.frag, M.P, M.norm, f, t, nullptr, (int)n, g, l, c->stream);
    HIPCHK(hipMemcpyAsync(grad, g, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost, c->stream));
    if (loss) HIPCHK(hipMemcpyAsync(loss, l, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    sync_check(c);
    return DFR_OK;
  });
}

int dfr_mlp_input_grad(dfr_ctx* c, int64_t n, const float* feat, const float* ybar, float* fbar) {
  return guarded(c, [&]() {
    if (!c->mlp.has_params || n <= 0 || !feat || !ybar || !fbar) return fail(c, DFR_ERR_ARG, "dfr_mlp_input_grad: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    MlpState& M = c->mlp;
    DevTmp<float> f(6 * n), y(2 * n), o(6 * n);
    HIPCHK(hipMemcpy(f.p, feat, sizeof(float) * 6 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(y.p, ybar, sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    launch_mlp_input_grad(M.frag, M.P, M.norm, f.p, nullptr, (int)n, y.p, o.p, c->stream);
    sync_check(c);
    HIPCHK(hipMemcpy(fbar, o.p, sizeof(float) * 6 * n, hipMemcpyDeviceToHost));
    return DFR_OK;
  });
}

int64_t dfr_num_fluid(const dfr_ctx* c) { return c ? c->f.n : DFR_ERR_ARG; }
int64_t dfr_num_boundary(const dfr_ctx* c, int body) {
  if (!c || body < 0 || body >= (int)c->bodies.size()) return DFR_ERR_ARG;
  return c->bodies[body].n;
}
int dfr_num_bodies(const dfr_ctx* c) { return c ? (int)c->bodies.size() : DFR_ERR_ARG; }

}  // extern "C"
