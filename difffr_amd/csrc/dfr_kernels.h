// Buffer bundles and kernel launchers of libdfr_hip.so (internal header).
#pragma once
#include <utility>
#include "dfr_device.h"

namespace dfr {

// Brick tiles: largest halo staged in LDS (a 6x6x6-cell halo holds ~2160
// particles at rest density: V_f = 0.8 (2r)^3 packs 10 particles into a cell of
// edge h = 4r; the capacity in use, NbrBufs::cap, follows the largest halo of
// recent builds and never exceeds this) and fluid-neighbour slots per particle
// (16-bit tile slots).
constexpr int TILE_CAP = 2560;
constexpr int MAXF = 128;
constexpr int PACK = 8;               // 16-bit slots per 16-byte pack
constexpr int NPACK = MAXF / PACK;    // packs per particle
constexpr int PASS_BLOCK = 640;       // threads of a gather-pass workgroup (one brick: ~640 particles at rest density)

// Fluid state in HBM, structure-of-arrays, permuted into cell order every step.
struct FluidBufs {
  int n = 0;
  float4 *pos = nullptr, *vel = nullptr;       // xyz + unused w
  float *kappa = nullptr, *kappav = nullptr;   // warm-start stiffness (persist across steps)
  uint32_t* id = nullptr;                      // original particle id
  float4 *pos2 = nullptr, *vel2 = nullptr;     // permutation targets
  float *kappa2 = nullptr, *kappav2 = nullptr;
  uint32_t* id2 = nullptr;
  float* rho = nullptr;       // density
  float* alpha = nullptr;     // DFSPH factor
  float4* s = nullptr;        // sum_grad_p_k (xyz)
  float* badv = nullptr;      // rho* (pressure) or d (divergence): "densityAdv"
  float* kf = nullptr;        // k_i of the current Jacobi iteration
  void swap_sorted() {
    std::swap(pos, pos2); std::swap(vel, vel2); std::swap(kappa, kappa2);
    std::swap(kappav, kappav2); std::swap(id, id2);
  }
};

// Boundary samples of all bodies, concatenated body by body (fixed order).
struct BoundaryBufs {
  int n = 0;
  int nd = 0;                 // samples of dynamic/animated bodies
  float4* pos = nullptr;      // world xyz + volume in w
  float4* vel = nullptr;      // xyz
  float4* x0 = nullptr;       // local xyz + body index in w
  uint32_t* dyn_idx = nullptr;  // indices of dynamic/animated samples
};

struct GridBufs {
  uint32_t *fcount = nullptr, *fstart = nullptr, *bcount = nullptr, *bstart = nullptr;
  uint32_t *key = nullptr, *slot = nullptr, *perm = nullptr;
  uint32_t *bkey = nullptr, *bslot = nullptr, *bperm = nullptr;
  uint32_t* tile_sums = nullptr;
};

struct NbrBufs {
  int maxb = 0, maxbf = 0;
  uint4* nbr8 = nullptr;      // [NPACK][nf] fluid neighbours as tile slots, 8 per pack
  uint4* thdr = nullptr;      // [nbricks] own_begin, own_end, halo total of each brick (written by the build)
  uint32_t* tmap = nullptr;   // [nbricks][TILE_CAP] sorted particle index of each halo slot
  uint32_t* nbrB = nullptr;   // [maxb][nf] boundary neighbours (sample index)
  uint32_t* cnt = nullptr;    // [nf]  nF | nB << 16
  uint32_t* bf = nullptr;     // [maxbf][nd] fluid neighbours of dynamic samples (sorted index)
  uint32_t* bfcnt = nullptr;  // [nd]
  uint32_t* overflow = nullptr;  // [4] max count seen when a list overflowed (fluid, B->F, boundary),
                                 //     largest halo of the last build
  int cap = TILE_CAP;            // halo slots the passes stage in LDS; larger halos read global memory
};

// dfr_grid.hip
size_t scan_tile_count(uint32_t n);
// One index set of the cell sort: positions, per-cell counts (zero between
// sorts), cell key and slot of each index, cell starts (nkeys + 1), the
// permutation, the in-cell order key (nullptr: the index itself) and
// scan_tile_count(nkeys) + 1 words of scan scratch.
struct SortSet {
  const float4* pos;
  uint32_t n;
  uint32_t* count;
  uint32_t* key;
  uint32_t* slot;
  uint32_t* start;
  uint32_t* perm;
  const uint32_t* order;
  uint32_t* tile_sums;
};
struct SortSets {
  SortSet s[2];
  uint32_t split;  // blocks of set 0 in the launches over indices (set 1 takes the rest)
};
// The fluid (physically permuted into cell order) and the boundary samples
// (index list only) are sorted by the same launches.  gb.tile_sums holds
// 2 * (scan_tile_count(nkeys) + 1) words; the counts start and end at zero.
void launch_sort_fluid_and_boundary(FluidBufs& f, BoundaryBufs& b, const GridDesc& g, GridBufs& gb, hipStream_t s);
void launch_sort_boundary(BoundaryBufs& b, const GridDesc& g, GridBufs& gb, hipStream_t s);
void launch_build_bf(FluidBufs& f, BoundaryBufs& b, const GridDesc& g, GridBufs& gb, NbrBufs& nb, float h2,
                     hipStream_t s);

// dfr_forces.hip: non-pressure forces and the CFL bound
struct ForceConst {
  float h;              // support radius (the cohesion and adhesion kernels' radius)
  float st_k, st_kb;    // surface tension (cohesion, curvature) and boundary adhesion coefficients
  float visc;           // viscosity coefficient
  int st_method;        // 0 none, 2 Akinci 2013, 3 He 2014
  int visc_mode;        // 0 none, 1 Standard, 2 XSPH
  float ck, cc;         // cohesion kernel: 32 / (pi h^9), h^6 / 64
  float ak;             // adhesion kernel: 0.007 / h^3.25
};
void launch_normals_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                      NbrBufs& nb, float4* nrm, hipStream_t s);
// He 2014: the colour field, then |grad c|^2 (two gather passes).
void launch_he_color_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                       NbrBufs& nb, float* color, float* gc2, hipStream_t s);
void launch_forces_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                     NbrBufs& nb, const float4* nrm, const float* gc2, float4* acc, RigidDev* rig,
                     const ForceConst& fc, hipStream_t s);
void launch_cfl_max(const StepConst& c, const float4* vel, const float4* acc, int nb, const float4* bvel,
                    uint32_t* out, hipStream_t s);
void launch_advect_acc(const StepConst& c, FluidBufs& f, const float4* acc, float kappav_scale, hipStream_t s);
void set_forces_kernel_attributes();

// dfr_tiled.hip
void launch_build_tiled(FluidBufs& f, BoundaryBufs& b, const GridDesc& g, GridBufs& gb, NbrBufs& nb, float h2,
                        hipStream_t s);
// aux (optional): per particle (velocity divergence, boundary density sum_b V_b W_ib)
void launch_density_factor_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                             NbrBufs& nb, int div_warm, int divergence_on, float2* aux, hipStream_t s);
// Where a recorded Jacobi iteration (or warm start) keeps what it starts from:
// the velocity before the update, the stiffness it applies and the solve's b
// field (each null: not recorded).  The velocity update writes them for its
// own particles.
struct IterTapeOut {
  float4* v = nullptr;
  float* k = nullptr;
  float* b = nullptr;
  const float* bsrc = nullptr;
};
// stop (optional): device flag of a Jacobi solve; a set flag makes the launch a no-op
void launch_velocity_update_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                              NbrBufs& nb, RigidDev* rig, const float* kfield, float* kappa_acc, hipStream_t s,
                              const int* stop = nullptr, IterTapeOut tape = {});
// mode 0: divergence recompute, 1: pressure recompute, 2: pressure warm start,
// 3: pressure warm start from predicted stiffness kpred (learned)
void launch_recompute_t(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                        NbrBufs& nb, int mode, float kcoef, double* err_slot, hipStream_t s,
                        const float* kpred = nullptr, const int* stop = nullptr);
void launch_expand_lists(const GridDesc& g, GridBufs& gb, FluidBufs& f, NbrBufs& nb, uint32_t* out, hipStream_t s);
void set_tiled_kernel_attributes();

// dfr_solver.hip
enum RecomputeMode { RECOMPUTE_DIV = 0, RECOMPUTE_PRESSURE = 1, PRESSURE_WARMSTART = 2, PRESSURE_WARMSTART_LEARNED = 3 };
void launch_advect(const StepConst& c, FluidBufs& f, float kappav_scale, hipStream_t s);
void launch_position(const StepConst& c, FluidBufs& f, float kappa_scale, hipStream_t s);
// Several device copies / zero fills in one launch (16-byte aligned pointers,
// sizes a multiple of 4 bytes); src == nullptr fills dst with zeros.
struct CopySeg {
  const void* src;
  void* dst;
  uint64_t bytes;
};
constexpr int COPY_SEGS = 8;
struct CopyList {
  CopySeg seg[COPY_SEGS];
  int n;
};
// stop != nullptr: the copy is skipped when *stop is set (a Jacobi iteration
// enqueued after its solve ended).
void launch_multi_copy(const CopyList& list, hipStream_t s, const int* stop = nullptr);
// Convergence test of one Jacobi iteration, on the device: probe gets the
// average residual err / n and whether the solve ends after iteration it
// (the loop condition of the reference); stop is set when it does.  A launch
// after the solve ended changes nothing.
struct SolveProbe {
  double avg;
  int done;  // 1: the solve ends after this iteration, 0: it continues
  int pad;
};
// Sums the per-brick residual partials of a recompute launch in a fixed order
// into *err; with stop != nullptr also the iteration's convergence test
// (skipped when *stop is already set).
void launch_solve_check(const double* part, int nparts, double* err, int n, double eta, int it, int min_it,
                        int max_it, int* stop, SolveProbe* probe, hipStream_t s);
void launch_boundary_volume_impl(const float4* bpos_in, const float4* bx0, int nb, const int* groups, int nbodies,
                                 KernelConst K, float h2, const GridDesc& g, GridBufs& gb, float4* bpos_out,
                                 hipStream_t s);

// dfr_rigid.hip
struct GradConst {
  int mode;              // 0 pressure, 1 divergence
  float coeff_trunc;     // (unsigned int)(1/dt^2) or (unsigned int)(1/dt)
  float coeff;           // 1/dt^2 or 1/dt
  int optimize_rotation;
};
void launch_gradient(const StepConst& c, const GradConst& gc, FluidBufs& f, BoundaryBufs& b, NbrBufs& nb,
                     RigidDev* rig, int k_first, int k_count, int body, hipStream_t s, const int* stop = nullptr);
struct RigidStepConst {
  double dt;
  double g[3];
  int gradient_mode;     // 0 Complete, 1 Incomplete, 2 RigidGradOnly
  int optimize_rotation;
  int gyroscopic;
  int compute_gradient;
  int nbodies;
  int use_manager;       // useRigidGradientManager: the manager's chain instead of perform_chain_rule
  int contact;           // the rigid contact solver ran this step: gravity already applied, and the
                         // nbodies - 1 later updateBoundaryForces calls add their gyroscopic terms
  int phases;            // RS_* parts of the step to run (0: all)
  double* mgr_net;       // manager across bodies (with contact): [nb][nb][78] handed sums, else null
};
// Parts of k_rigid_step, run in this order: hand the step's sums to the manager (or the per-body
// chain rule), the single-body manager's velocity chain, the force/torque update, its position
// chain, animate.
enum RigidPhase { RS_HAND = 1, RS_MGR_VEL = 2, RS_FORCES = 4, RS_MGR_POS = 8, RS_ANIMATE = 16 };
void launch_rigid_step(const RigidStepConst& rc, RigidDev* rig, hipStream_t s);
// RigidBodyGradientManager across bodies (entries (R, RR), with the rigid contact solver):
// J, N = [nb][nb][78] Jacobians (layout of RigidDev::Jm) and handed sums (layout of RigidDev::net).
// The force/torque and velocity chain (rigid = 0 after the fluid coupling, 1 after the rigid
// coupling), then the position chain; nb <= 16.
void launch_manager_multi_velocity(const RigidStepConst& rc, RigidDev* rig, double* J, const double* N, int nb,
                                   int rigid, hipStream_t s);
void launch_manager_multi_position(const RigidStepConst& rc, RigidDev* rig, double* J, int nb, hipStream_t s);

// ---------------------------------------------------------------------------
// dfr_contact.hip: rigid-rigid contact (RigidContactSolver, penalty form).
// ---------------------------------------------------------------------------
struct ContactConst {
  double dt;
  double g[3];
  double gdt[3];     // g dt (the gravity step the sample velocities include)
  float beta, mu;    // rigidContactBeta, rigidContactFrictionCoeff
  int gyroscopic;
  KernelConst Ks;    // the simulation's kernel (its gradient enters the contact gradient terms)
  float4* grec;      // [6][nd] per-sample gradient records, or null (no manager across bodies)
  double* mgr_net;   // [nb][nb][78] the manager's handed sums, replaced by the contact terms, or null
};
// Rest volume and density of every boundary sample from its own body's neighbours.
void launch_contact_rest(const BoundaryBufs& b, const GridDesc& g, const GridBufs& gb, float h2, const KernelConst& Kc,
                         float gamma, float* vol0, float* dens0, hipStream_t s);
// Gravity and the contact forces of the step, applied to the rigid velocities.
void launch_contact(const BoundaryBufs& b, const GridDesc& g, const GridBufs& gb, float h2, const KernelConst& Kc,
                    const ContactConst& cc, const float* vol0, const float* dens0, float4* cf, float4* ct,
                    RigidDev* rig, int nbodies, hipStream_t s);
void launch_boundary_transform(const uint32_t* idx, int n, BoundaryBufs& b, const RigidDev* rig, hipStream_t s);
void launch_begin_step(RigidDev* rig, int nbodies, const int* ramp_flags, const double* factors,
                       const int* animated, hipStream_t s);

// ---------------------------------------------------------------------------
// dfr_adjoint.hip: reverse pass of one recorded step.
// ---------------------------------------------------------------------------
// Adjoint of one rigid body during the reverse of a step (FP64).
struct AdjBodyDev {
  double x[3], q[4], v[3], w[3];  // adjoint of the rigid state (x_rb, q (w,x,y,z), v, omega)
  double v0[3], w0[3];            // accumulated dL/d init_v_rb, dL/d init_omega_rb
  double Fb[3], Tb[3];            // adjoint of the step's force / torque sums
  double xr[3];                   // x_rb during the step (torque lever arm)
  double wpre[3];                 // omega that produced the boundary velocities of the step
  double q0[4];                   // q during the step
  int dynamic, pad;
};

// Per-body flags of the step being reversed (release window of beginStep).
struct AdjStepTable {
  int ramp[DFR_MAX_BODIES];
  int animated[DFR_MAX_BODIES];
  double factor[DFR_MAX_BODIES];
};

// What every reverse gather pass reads: the step's geometry (tape positions,
// neighbour lists, boundary state at the start of the step) and where the
// boundary-sample adjoints go.
struct AdjPass {
  StepConst c;
  GridDesc g;
  const uint32_t* fstart;
  const float4* x;     // fluid positions during the step (sorted order)
  const uint4* nbr8;
  const uint4* thdr;   // brick headers and halo slot maps of the step
  const uint32_t* tmap;
  const uint32_t* nbrB;
  const uint32_t* cnt;
  const float4* bx;    // boundary positions at the start of the step, w = volume
  const float4* bv;    // boundary velocities at the start of the step
  const float4* bx0;   // local sample, w = body index
  int nb_static;
  int cap;             // halo slots staged in LDS (larger halos read global memory)
  double* sxb;         // [3 * nd] adjoint of dynamic sample positions
  double* svb;         // [3 * nd] adjoint of dynamic sample velocities
  const AdjBodyDev* body;
};

enum AdjUpdateMode {
  ADJ_DIV_ITER = 0, ADJ_PRESSURE_ITER = 1, ADJ_DIV_WARM = 2, ADJ_PRESSURE_WARM = 3,
  ADJ_DIV_WARM_LEARNED = 4, ADJ_PRESSURE_WARM_LEARNED = 5
};
void launch_adj_update(const AdjPass& a, int mode, const float* k, const float4* vbar, const float* kacc,
                       float kacc_scale, const float* bval, const float* kprev, const float* alpha, float* out,
                       float* alpha_bar, float* rho_bar, float4* xbar, hipStream_t s);
void launch_adj_delta(const AdjPass& a, const float4* v, const float* dbar, float4* vbar, float4* xbar,
                      hipStream_t s);
// Reverse of the advection with Standard (mode 1) or XSPH (mode 2) viscosity:
// abar = dt vbar_in (scratch), then vbar, xbar and rho_bar (adjoint of rho / rho0)
// receive the force's terms.  v: velocities after the divergence solve.
void launch_adj_viscosity(const AdjPass& a, int mode, const float4* v, const float* rho, const float4* vbar_in,
                          float4* abar, const ForceConst& fc, float4* vbar, float4* xbar, float* rho_bar,
                          hipStream_t s);
void launch_adj_factor(const AdjPass& a, const float* alpha, const float* alpha_bar, const float* rho_bar,
                       const float4* G, float4* xbar, const float* rbb, hipStream_t s);
void launch_adj_position(int n, float dt, const float4* xbar, float4* vbar, hipStream_t s);
void launch_adj_unpermute(int n, const uint32_t* perm, const float4* xb, const float4* vb, const float* kb,
                          const float* kvb, float4* xb2, float4* vb2, float* kb2, float* kvb2, hipStream_t s);
void launch_adj_boundary_reduce(const AdjPass& a, int body, int first, int n, AdjBodyDev* ab, hipStream_t s);
void launch_adj_rigid(const RigidStepConst& rc, const AdjStepTable& t, const RigidDev* forced, const RigidDev* pre,
                      const RigidDev* post, AdjBodyDev* ab, int nbodies, hipStream_t s);
void set_adjoint_kernel_attributes();

// ---------------------------------------------------------------------------
// dfr_mlp.hip: learned kappa warm start (6 -> 64 -> 64 -> 64 -> 2, SiLU).
// Flat FP32 parameters: W1[64][6] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
// W4[2][64] b4[2].
// ---------------------------------------------------------------------------
constexpr int MLP_NP = 2402;  // 6 -> 32 -> 32 -> 32 -> 2

struct MlpNorm {
  float mu_in[6], is_in[6];    // feature mean and 1/std
  float mu_out[2], sd_out[2];  // target (kappa_v, kappa) mean and std
};

struct MlpFields {            // the step's fields (sorted order)
  const float* rho;
  const float2* aux;          // (velocity divergence, boundary density) from the density pass
  const uint32_t* cnt;
  const float* kappa;
  const float* kappav;
  float rho0, dt;
};

size_t mlp_frag_bytes();
void launch_mlp_pack(const float* P, void* frag, hipStream_t s);
void launch_mlp_infer_array(const void* frag, const float* P, const MlpNorm& norm, const float* feat, int n,
                            float* out, hipStream_t s);
void launch_mlp_train_batch(const void* frag, const float* P, const MlpNorm& norm, const float* feat,
                            const float* targ, const uint32_t* index, int batch, float* grad, double* loss,
                            hipStream_t s);
void launch_mlp_input_grad(const void* frag, const float* P, const MlpNorm& norm, const float* feat,
                           const MlpFields* fields, int n, const float* ybar, float* fbar, hipStream_t s);
void launch_mlp_adam(float* P, float* G, float* m, float* v, float lr, int t, hipStream_t s);
void launch_mlp_batch(uint32_t* idx, int batch, uint64_t seed, uint32_t size, hipStream_t s);
void launch_mlp_capture_features(const MlpFields& f, const uint32_t* id, int n, uint64_t salt, uint32_t threshold,
                                 float* feat, uint32_t* rows, uint32_t* count, uint32_t cap, hipStream_t s);
void launch_mlp_capture_targets(const float* kappa, const float* kappav, const uint32_t* rows, const uint32_t* first,
                                uint32_t max_new, const uint32_t* count, uint32_t cap, float* targ, hipStream_t s);
// Inference on the step's fields that applies the warm starts directly:
// kappa_v = [d > 0] y_v / dt (d == nullptr: kappa_v untouched), kpred = y_p.
void launch_mlp_infer_apply(const void* frag, const float* P, const MlpNorm& norm, const MlpFields& f, int n,
                            const float* d, float inv_dt, float* kappav, float* kpred, hipStream_t s);
// Input gradient on the step's fields that applies the feature adjoints
// directly: rho_bar += fbar[0], dbar = dt fbar[1], rbb = fbar[3],
// kb = fbar[4] (+ kb if keep_kappa), kvb = fbar[5] (+ kvb if keep_kappav).
struct LearnedAdjSink {
  float *kb, *kvb, *rhob, *dbar, *rbb;
  int keep_kappa, keep_kappav;
  float dt;
};
void launch_mlp_input_grad_apply(const void* frag, const float* P, const MlpNorm& norm, const MlpFields& f, int n,
                                 const float* ybar, const LearnedAdjSink& sink, hipStream_t s);
void set_mlp_kernel_attributes();

}  // namespace dfr
