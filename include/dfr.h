/*
 * libdfr_hip.so -- C ABI of the MI355X (gfx950) differentiable DFSPH
 * fluid-rigid time step.
 *
 * The library owns all device memory of one simulation context.  Host
 * pointers passed in are borrowed for the duration of the call only.  A
 * context is bound to one device and one HIP stream; it is not thread-safe.
 * Every function returns 0 on success and a negative DFR_ERR_* code on
 * failure; dfr_last_error() then holds a message.
 *
 * Reference interfaces replaced (paths relative to the DiffFR repository):
 *   SPlisHSPlasH/DiffDFSPH/TimeStepDiffDFSPH.{h,cpp}      TimeStepDiffDFSPH::step and getters
 *   SPlisHSPlasH/BoundaryModel_Akinci2012.{h,cpp}          boundary samples, coupling, gradient storage
 *   SPlisHSPlasH/Dynamic3dRigidBody.h                      rigid state and integration
 *   Simulator/SimulatorBase.cpp:1142-1200                  timeStepNoGUI step order
 */
#ifndef DFR_H
#define DFR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFR_OK 0
#define DFR_ERR_ARG -1        /* invalid argument or call order */
#define DFR_ERR_HIP -2        /* HIP runtime error */
#define DFR_ERR_CAPACITY -3   /* a neighbour list exceeded its capacity */
#define DFR_ERR_STATE -4      /* context not initialised / already initialised */
#define DFR_ERR_RANGE -5      /* an exact gradient left the FP32 range */

typedef struct dfr_ctx dfr_ctx;

/* Solver parameters.  Defaults (dfr_default_params) are the reference's:
 * TimeStep.cpp:24-64 (minIterations 2, maxIterations 100, maxError 0.01 %),
 * TimeStepDiffDFSPH.cpp:90-111 (maxIterationsV 100, maxErrorV 0.1 %, warm
 * starts and divergence solver on, optimize rotation on), Simulation.cpp:100-125
 * (gravity, rigid-body mode WithGyroscopic), SceneLoader.cpp:45-87
 * (gradientMode 1). */
typedef struct dfr_params {
  double radius;                 /* particle radius r; support radius h = 4r */
  double dt;                     /* fixed time step */
  double rho0;                   /* rest density */
  double gravity[3];
  double max_error;              /* pressure solve tolerance, percent of rho0 */
  double max_error_v;            /* divergence solve tolerance, percent of rho0 per second */
  int32_t min_iterations;
  int32_t max_iterations;
  int32_t max_iterations_v;
  int32_t use_pressure_warmstart;
  int32_t use_divergence_warmstart;
  int32_t enable_divergence_solver;
  int32_t gradient_mode;         /* 0 Complete, 1 Incomplete, 2 RigidGradOnly */
  int32_t optimize_rotation;
  int32_t gyroscopic;            /* 1 = RigidBodyMode::WithGyroscopic */
  int32_t compute_gradient;      /* accumulate DiffFR's forward-mode rigid gradient */
  double uniform_acc_time;       /* uniform_acc_rb_time */
  double target_time;
  int32_t max_neighbors;         /* boundary samples per fluid particle, default 64 (fluid
                                    neighbours: 128 per particle, fixed) */
  int32_t max_boundary_neighbors;/* fluid neighbours per dynamic sample, default 128 */
  int32_t release_mode;          /* useReleaseRigidBodyMode (Simulation.h:444): body 1 is held at rest
                                    until uniform_acc_time + dt, then set to (v0, omega0) at the start
                                    of every step; other bodies are not driven */
  int32_t use_gradient_manager;  /* useRigidGradientManager (Simulation.cpp:123): the per-step sums go to
                                    the RigidBodyGradientManager chain (dfr_get_gradient_manager) instead
                                    of perform_chain_rule; dfr_get_jacobians then stays at its reset value */
  /* Non-pressure forces after the divergence solve (Simulation::computeNonPressureForces,
     Simulation.cpp:618-631; method numbers as registered in NonPressureForceRegistration.cpp:33-60).
     The exact adjoint (dfr_adjoint_begin) does not cover them, nor CFL step sizes. */
  int32_t surface_tension_method;  /* 0 none, 2 Akinci et al. 2013 (SurfaceTension_Akinci2013.cpp),
                                      3 He et al. 2014 (SurfaceTension_He2014.cpp): the reference's
                                      numbering, NonPressureForceRegistration.cpp:42-47 */
  double surface_tension;          /* cohesion and curvature coefficient (default 0.05) */
  double surface_tension_boundary; /* adhesion coefficient (default 0.01) */
  int32_t viscosity_method;        /* 0 none, 1 Standard (Viscosity_Standard.cpp; the reference fluid's default,
                                      FluidModel.cpp:100), 2 XSPH (Viscosity_XSPH.cpp) */
  double viscosity;                /* default 0.01 */
  double viscosity_boundary;       /* 0 only: boundary viscosity (which pushes the bodies) is not built */
  int32_t cfl_method;              /* 0 fixed dt, 1 CFL (Simulation::updateTimeStepSizeCFL, :544-616) */
  double cfl_factor;               /* default 0.5 */
  double cfl_min_time_step_size;   /* default 1e-4 */
  double cfl_max_time_step_size;   /* default 5e-3 */
  /* Rigid-rigid contact between the bodies' boundary samples (RigidContactSolver, penalty form;
     Configuration keys of SceneLoader.cpp:58-71).  With the reference-mode gradient and
     use_gradient_manager the contact solver's gradient terms fill the manager's entries between
     bodies (dfr_get_gradient_manager(R, RR), at most 16 bodies); the exact adjoint refuses contact. */
  int32_t use_rigid_contact;       /* useRigidContactSolver, default 0 */
  double rigid_contact_beta;       /* rigidContactBeta, penalty stiffness (default 1) */
  double rigid_contact_gamma;      /* rigidContactGamma, rest volume factor (default 0.7) */
  double rigid_contact_support_radius_factor; /* rigidContactSupportRadiusFactor: contact kernel radius / r (4) */
  double rigid_contact_friction;   /* rigidContactFrictionCoeff (default 0) */
} dfr_params;

typedef struct dfr_stats {
  int32_t iterations;            /* pressure iterations of the last step (m_iterations) */
  int32_t iterations_v;          /* divergence iterations of the last step (m_iterationsV) */
  double avg_density_err;        /* kg/m^3, last pressure iteration */
  double avg_div_err;            /* kg/m^3/s, last divergence iteration */
  double time;                   /* simulated time (FP32 accumulation, as TimeManager) */
  int64_t step_count;
  int64_t total_iterations;      /* sums since the last dfr_reset_stats */
  int64_t total_iterations_v;
  int64_t steps_since_reset;
  int32_t trajectory_finished;   /* TimeStepDiffDFSPH::is_trajectory_finish_callback */
  double dt;                     /* step size the next step starts with (changes only with cfl_method 1) */
} dfr_stats;

/* Fills p with the reference defaults for a given particle radius and dt. */
void dfr_default_params(dfr_params* p, double radius, double dt);

/* Creates a context on HIP device `device`.  Replaces the construction of
 * Simulation + TimeStepDiffDFSPH (Simulation.cpp:195-216, 714-720). */
dfr_ctx* dfr_create(const dfr_params* p, int device);
int dfr_destroy(dfr_ctx* ctx);
const char* dfr_last_error(const dfr_ctx* ctx);

/* Fluid particles, AoS float[n*3] like FluidModel::m_x / m_v (FluidModel.h:110-151).
 * v may be NULL (zero velocity).  Must precede dfr_init. */
int dfr_set_fluid(dfr_ctx* ctx, int64_t n, const float* x, const float* v);

/* One boundary body of Akinci samples (BoundaryModel_Akinci2012::initModel,
 * :286-374) given in the body's local frame, placed at x_rb with rotation q
 * (w,x,y,z); mass properties follow Dynamic3dRigidBody::determineMassProperties
 * (Dynamic3dRigidBody.h:184-213).  Returns the body index (>= 0). */
int dfr_add_boundary(dfr_ctx* ctx, int64_t n, const float* x0, int is_dynamic, const double x_rb[3],
                     const double q_wxyz[4], double density);

/* init_v_rb / init_omega_rb of a dynamic body (TimeStepDiffDFSPH.h set_init_v_rb). */
int dfr_set_init_velocity(dfr_ctx* ctx, int body, const double v0[3], const double omega0[3]);

/* Uploads, places the boundary samples and computes the Akinci volumes
 * (Simulation::updateBoundaryVolume, Simulation.cpp:831-902). */
int dfr_init(dfr_ctx* ctx);

/* Runs nsteps full steps in SimulatorBase::timeStepNoGUI order: DFSPH step with
 * DiffFR's gradient accumulation (TimeStepDiffDFSPH::step, :526-652), rigid
 * velocity and position update and the boundary transform.  A neighbour list
 * that outgrows its capacity fails the step with DFR_ERR_CAPACITY.  A failed
 * step ends an adjoint recording (dfr_adjoint_begin) in progress and leaves
 * the state of a partial step; dfr_reset or dfr_state_load restores a whole
 * one. */
int dfr_step(dfr_ctx* ctx, int nsteps);

/* Restores the fluid and rigid bodies to their state at dfr_init, resets the
 * gradient chain and time (SimulatorBase::reset, SimulatorBase.cpp:887-934). */
int dfr_reset(dfr_ctx* ctx);

/* Rigid state out[19]: x(3), q(w,x,y,z)(4), v(3), omega(3), force(3), torque(3)
 * of the last step (BoundaryModel_Akinci2012::get_position_rb, ...
 * BoundaryModel_Akinci2012.h:169-190, BoundaryModel::getForce/getTorque). */
int dfr_get_rigid(const dfr_ctx* ctx, int body, double out[19]);

/* Overwrites the rigid state x, q, v, omega (in[13]). */
int dfr_set_rigid(dfr_ctx* ctx, int body, const double in[13]);

/* Keeps the current state (fluid positions, velocities and stiffness fields,
 * boundary samples, rigid bodies with their forward-mode Jacobians, time and
 * step count) in device memory under slot (0..4095) / puts it back.  The
 * in-memory counterpart of SimulatorBase::saveState and
 * loadStateWithRigidExisted (Simulator/SimulatorBase.cpp:1899-1960,
 * :2124-2290) for loops that restart from a state, e.g. the per-step finite
 * differences of experiments/others/python/finite-difference.py:55-110.
 * dfr_state_load ends an adjoint recording. */
int dfr_state_save(dfr_ctx* ctx, int slot);
int dfr_state_load(dfr_ctx* ctx, int slot);
/* Frees the device memory of a kept state (no effect on an empty slot). */
int dfr_state_free(dfr_ctx* ctx, int slot);

/* Body properties (any pointer may be NULL): the mass, (4/3) pi r^3 rho_body
 * per boundary sample for static bodies too (RigidBodyObject::getMass,
 * Dynamic3dRigidBody.h:59,184-198); whether the body is dynamic
 * (isDynamic); whether the last step drove it kinematically in the
 * uniform-acceleration window (isAnimated, set by beginStep,
 * TimeStepDiffDFSPH.cpp:379-429) -- pySPlisHSPlasH/RigidBodyModule.cpp:19-22. */
int dfr_get_body(const dfr_ctx* ctx, int body, double* mass, int* dynamic, int* animated);

/* Gradient mode read by the following steps: 0 Complete, 1 Incomplete,
 * 2 RigidGradOnly (Simulation::setGradientMode,
 * pySPlisHSPlasH/SimulationModule.cpp:208). */
int dfr_set_gradient_mode(dfr_ctx* ctx, int mode);

/* Switches DiffFR's forward-mode gradient accumulation (the BACKWARD path of
 * TimeStepDiffDFSPH::step, :526-652, and backwardPerStep, :493-524) on or off
 * for the following steps; dfr_params.compute_gradient sets it at creation. */
int dfr_set_compute_gradient(dfr_ctx* ctx, int on);

/* Forward-mode rigid Jacobians out[78]: Jxv, Jxw, Jvv, Jvw, Jwv, Jww (3x3
 * row-major), then Jqv, Jqw (4x3, rows w,x,y,z) -- get_grad_x_to_v0 etc.,
 * BoundaryModel_Akinci2012.h:209-323. */
int dfr_get_jacobians(const dfr_ctx* ctx, int body, double out[78]);

/* Pair-Jacobian sums of the last step out[78]: dF/dv, dF/dx, dF/domega,
 * dtau/domega, dtau/dx, dtau/dv (3x3), dF/dq, dtau/dq (3x4)
 * (grad_net_force_to_vn etc., BoundaryModel_Akinci2012.cpp:453-497). */
int dfr_get_net_gradients(const dfr_ctx* ctx, int body, double out[78]);

/* RigidBodyGradientManager entries of bodies (R, RR) (SimulationModule.cpp:406-413,
 * RigidBodyGradientManager.h:24-52): jac[78] in dfr_get_jacobians' layout
 * (get_grad_x_to_v0(R, RR), get_grad_q_to_v0(R, RR), ...), net[78] in
 * dfr_get_net_gradients' layout (get_grad_net_force_to_vn(R, RR), ...).  Either
 * pointer may be NULL.  Entries with R != RR are zero without the rigid contact
 * solver (with it, its gradient terms fill them: RigidContactSolver.cpp:471-553,
 * RigidBodyGradientManager.cpp:158-239); all entries of a static body, or without
 * use_gradient_manager, are zero. */
int dfr_get_gradient_manager(const dfr_ctx* ctx, int R, int RR, double jac[78], double net[78]);

/* Clears the forward-mode Jacobians to their initial values
 * (TimeStepDiffDFSPH reset_gradient / BoundaryModel_Akinci2012::reset_gradient). */
int dfr_reset_gradient(dfr_ctx* ctx);

/* Fluid state in particle-id order (the order given to dfr_set_fluid).  Any
 * pointer may be NULL. */
int dfr_get_fluid(const dfr_ctx* ctx, float* x, float* v);

/* A per-particle field in particle-id order: "density", "factor",
 * "density_adv", "kappa", "kappa_v" (n floats) or "sum_grad_p_k" (3n); with
 * non-pressure forces on, "accel" (the last step's FluidModel accelerations,
 * g + forces, 3n) and, with surface tension, "normal" (Akinci 2013, 3n) or "color"
 * and "gradC2" (He 2014, n each). */
int dfr_get_field(const dfr_ctx* ctx, const char* name, float* out);

/* Overwrites "kappa" or "kappa_v" (n floats) or "x" / "v" (3n floats: positions,
 * velocities), in particle-id order; loading a saved state
 * (SimulatorBase::loadStateWithRigidExisted, SimulatorBase.cpp:2124-2290, and
 * readFluidParticlesState, :2545-2600). */
int dfr_set_field(dfr_ctx* ctx, const char* name, const float* in);

/* Simulated time and step count, as TimeManager::loadState restores them
 * (TimeManager.cpp:63-67); the begin-step window and the trajectory end use
 * them. */
int dfr_set_time(dfr_ctx* ctx, double time, int64_t step_count);

/* Boundary sample data of one body: "x", "v" (3n floats) or "volume" (n). */
int dfr_get_boundary_field(const dfr_ctx* ctx, int body, const char* name, float* out);

int dfr_get_stats(const dfr_ctx* ctx, dfr_stats* out);
int dfr_reset_stats(dfr_ctx* ctx);

/* Neighbour-list diagnostics of the last step: mean fluid and boundary
 * neighbours per fluid particle and mean fluid neighbours per dynamic sample. */
int dfr_neighbor_stats(const dfr_ctx* ctx, double out[3]);

/* Number of (fluid particle, dynamic body) pairs in which the particle has at
 * least one of the body's samples among its neighbours in the last step's
 * lists (TimeStepDiffDFSPH::countNeighborDOF, TimeStepDiffDFSPH.cpp:283-330,
 * read by get_num_1ring_fluid_particle, :2242-2245).  Returns the count or a
 * negative error code. */
int64_t dfr_count_1ring(const dfr_ctx* ctx);

/* Halo tiles of the gather passes (no counterpart in the reference: a tuning
 * and test hook of this implementation).  Each pass stages a brick's halo of
 * up to `slots` particles in LDS; by default (slots < 0) the capacity follows
 * the largest halo of the last neighbour build.  A larger halo, or every halo
 * with slots = 0, is read from global memory with identical arithmetic.
 * dfr_halo_stats: out = {largest halo of the last build, capacity in use,
 * maximum capacity}. */
int dfr_set_halo_capacity(dfr_ctx* ctx, int slots);
int dfr_halo_stats(const dfr_ctx* ctx, int64_t out[3]);

/* Neighbour lists of the last step in particle-id / sample-index terms:
 * kind 0 fluid->fluid (ids), kind 1 fluid->boundary (global sample index of
 * `counts`-ordered bodies), kind 2 dynamic sample->fluid (ids).  counts gets
 * one entry per row; lists gets the concatenated rows (capacity `cap`
 * entries).  Returns the number of rows or a negative error. */
int64_t dfr_get_neighbors(const dfr_ctx* ctx, int kind, uint32_t* counts, uint32_t* lists, int64_t cap);

/* Per-kernel-class timing with HIP events on the context's stream
 * (0 = off).  dfr_kernel_times fills names/avg-ms/launch-count for up to n
 * classes and returns the number of classes. */
int dfr_set_profiling(dfr_ctx* ctx, int on);
/* Only the classes whose bit is set in mask (bit k = k-th class of
 * dfr_kernel_times), so that timing one kernel class adds no events around
 * the others; mask 0 = off. */
int dfr_set_profiling_classes(dfr_ctx* ctx, uint32_t mask);
int dfr_kernel_times(const dfr_ctx* ctx, int n, char names[][32], double* avg_ms, int64_t* launches);

/* Exact adjoint of a rollout (no counterpart in the reference, whose only
 * gradient is the forward-mode approximation above; this is the adjoint path a
 * differentiable step needs, cf. TimeStepDiffDFSPH::backwardPerStep,
 * TimeStepDiffDFSPH.cpp:493-524, which it replaces for exact gradients).
 *
 * dfr_adjoint_begin marks the current state as the start of a rollout; the
 * following dfr_step calls save a checkpoint (44 bytes per fluid particle, 32
 * per boundary sample) every `checkpoint_every` steps (<= 0: every step) and
 * log their iteration counts.  dfr_adjoint_backward then returns
 * grad[6*b..6*b+5] = dL/d(init_v_rb, init_omega_rb) of body b for the seed
 * seed[7*b..7*b+6] = dL/d(x_rb, q (w,x,y,z)) of every body after the last
 * step: each segment of checkpoint_every steps is re-run once from its
 * checkpoint, keeping its steps' tapes in HBM, and the reverse kernels run
 * through them.  When a segment's tapes would not fit (or exceed the bound
 * set by dfr_adjoint_segment_memory), the segment's states are kept instead
 * and every step is re-run once more with recording right before it is
 * reversed.  Clamps, the eps guards, the 20-neighbour rule and
 * iteration counts are held at their values in the run (the derivative of the
 * piecewise-smooth step).  The state after the last step is restored on
 * return, so backward can be called again with another seed.  dfr_reset,
 * dfr_set_rigid, dfr_set_field and dfr_set_init_velocity end the recording;
 * dfr_adjoint_end frees the checkpoints.
 *
 * Once fluid and body interact strongly, this derivative grows exponentially
 * with the length of the window (the step is discontinuous in the eps guards
 * and the 20-neighbour rule, and its a.e. derivative is large); when the
 * reverse pass leaves the FP32 range, backward writes grad as computed and
 * returns DFR_ERR_RANGE. */
int dfr_adjoint_begin(dfr_ctx* ctx, int checkpoint_every);
int dfr_adjoint_backward(dfr_ctx* ctx, const double* seed, double* grad);
int dfr_adjoint_end(dfr_ctx* ctx);
/* After dfr_adjoint_backward: out[13*b..13*b+12] = dL/d(x_rb, q (w,x,y,z), v,
 * omega) of body b at the start of the recorded window, i.e. before its first
 * step (for a window that starts after the release window of beginStep, the
 * gradient with respect to the rigid state the window starts from). */
int dfr_adjoint_start_gradient(const dfr_ctx* ctx, double* out);
/* Bytes one segment's kept tapes may occupy in the backward pass of a
 * checkpointed rollout: -1 (default) whatever device memory is free, 0 never
 * (every step re-run twice), > 0 that bound. */
int dfr_adjoint_segment_memory(dfr_ctx* ctx, int64_t bytes);
/* Test hook (no counterpart in the reference): bytes the tape arena may hand
 * out (0 = no limit beyond device memory).  Reaching it has the effect of
 * running out of device memory: a kept-tape rollout (checkpoint_every <= 0)
 * fails its step with DFR_ERR_HIP and ends the recording; a checkpointed
 * backward pass falls back to re-running every step twice, freeing the
 * arena's memory first, as when a segment's tapes outgrow free memory. */
int dfr_adjoint_arena_limit(dfr_ctx* ctx, int64_t bytes);
/* Device bytes the kept tapes of the current rollout occupy (checkpoint_every
 * <= 0), for sizing rollouts against device memory. */
int64_t dfr_adjoint_tape_bytes(const dfr_ctx* ctx);
/* Iteration counts (divergence, pressure) of the steps recorded since
 * dfr_adjoint_begin; returns the number of steps. */
int64_t dfr_adjoint_iterations(const dfr_ctx* ctx, int* iv, int* ip, int64_t cap);

/* Imposes exactly iv[s] divergence and ip[s] pressure iterations on the next
 * n steps (n = 0 clears), e.g. to replay another run's iteration counts. */
int dfr_set_fixed_iterations(dfr_ctx* ctx, int n, const int* iv, const int* ip);

/* Learned kappa warm start (the learned part of this library; no counterpart
 * in the reference, whose warm starts halve the previous step's stiffness,
 * TimeStepDiffDFSPH.cpp:964-1059 and :1698-1788).  A 6 -> 32 -> 32 -> 32 -> 2
 * MLP (SiLU, bf16 weights on MFMA, FP32 accumulation) maps per-particle
 * features (rho/rho0 - 1, dt * velocity divergence, neighbour count, boundary
 * density, previous kappa, previous kappa_v) to the stiffness the two solves
 * converge to; the Jacobi solves still iterate to the same tolerance.
 *
 * dfr_mlp_capture(capacity, fraction, seed) records, during the following
 * steps, (features, converged (kappa_v, kappa)) pairs of a hashed `fraction` of
 * the particles into a device buffer of `capacity` rows (capacity 0 stops).
 * Called again with the same capacity while capturing, it keeps the rows
 * captured so far and appends with the new fraction and seed (rollouts of
 * several drops from one saved state into one data set).
 * dfr_mlp_train runs `iters` Adam steps (batch rows, cosine-decayed lr) on the
 * captured pairs with the library's MFMA forward/backward kernels; losses[it]
 * (optional) gets the mean squared normalised error of each batch.
 * Parameters are a flat float[2402] (W1[32][6] b1[32] W2[32][32] b2[32]
 * W3[32][32] b3[32] W4[2][32] b4[2]); norm is float[16] (feature mean[6],
 * feature std[6], target mean[2], target std[2]).  dfr_mlp_enable switches the
 * learned warm start on or off for the following steps. */
int dfr_mlp_capture(dfr_ctx* ctx, int64_t capacity, double fraction, uint64_t seed);
int64_t dfr_mlp_dataset_size(dfr_ctx* ctx);
int64_t dfr_mlp_dataset(dfr_ctx* ctx, float* feat, float* targ, int64_t cap);
int dfr_mlp_init_params(dfr_ctx* ctx, uint64_t seed);
int dfr_mlp_set_params(dfr_ctx* ctx, const float* params, const float* norm);
int dfr_mlp_get_params(dfr_ctx* ctx, float* params, float* norm);
int dfr_mlp_train(dfr_ctx* ctx, int iters, int batch, double lr, uint64_t seed, double* losses);
int dfr_mlp_enable(dfr_ctx* ctx, int on);
/* The network on host arrays: predictions out[n][2] (kappa_v, kappa; stored
 * units) for features feat[n][6]; the summed squared normalised error and the
 * gradient of its mean over the n rows; dL/dfeatures for dL/dpredictions. */
int dfr_mlp_predict(dfr_ctx* ctx, int64_t n, const float* feat, float* out);
int dfr_mlp_loss_grad(dfr_ctx* ctx, int64_t n, const float* feat, const float* targ, float* grad, double* loss);

/* Rows dfr_mlp_loss_grad_noalloc accepts per call. */
#define DFR_NOALLOC_MAX_ROWS 65536

/* dfr_mlp_loss_grad without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight training or serving loops, harnesses that forbid
 * hipMalloc).  The feature/target staging and the gradient and loss
 * accumulators live in one fixed device array per GPU (about 2.1 MB, taken at
 * load and held until the process ends); n above DFR_NOALLOC_MAX_ROWS returns
 * DFR_ERR_CAPACITY.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (several threads or contexts) race.  Same results as
 * dfr_mlp_loss_grad up to the order of the gradient's float atomics. */
int dfr_mlp_loss_grad_noalloc(dfr_ctx* ctx, int64_t n, const float* feat, const float* targ, float* grad,
                              double* loss);
int dfr_mlp_input_grad(dfr_ctx* ctx, int64_t n, const float* feat, const float* ybar, float* fbar);

/* Number of fluid particles / boundary samples / bodies. */
int64_t dfr_num_fluid(const dfr_ctx* ctx);
int64_t dfr_num_boundary(const dfr_ctx* ctx, int body);
int dfr_num_bodies(const dfr_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* DFR_H */
