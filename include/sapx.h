/*
 * sapx — MI355X (gfx950) SAP compliant-contact step, C ABI.
 *
 * Plain pointers and sizes only.  Model descriptions are HOST memory (the
 * library keeps its own device copy); state / control / contact views are
 * caller-allocated DEVICE memory.  Every entry point returns an int status
 * (SAPX_OK = 0, negative = error class) and never throws; the message of the
 * last error on the calling thread is available from sapx_last_error().
 * All work is enqueued asynchronously on the stream passed in (a hipStream_t,
 * NULL = default stream); nothing synchronises unless stated.
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef SAPX_H
#define SAPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAPX_OK 0
#define SAPX_ERR_INVALID (-1)     /* bad argument / inconsistent model (ValueError)      */
#define SAPX_ERR_UNSUPPORTED (-2) /* feature outside the native path (NotImplementedError) */
#define SAPX_ERR_HIP (-3)         /* HIP runtime failure (RuntimeError)                   */
#define SAPX_ERR_CAPACITY (-4)    /* sizes exceed the compiled kernel limits              */

#define SAPX_LS_ARMIJO_DECAY 0
#define SAPX_LS_EXACT_ROOT 1
#define SAPX_LS_MONOTONE_DECAY 2 /* one-wave step kernel only */

#define SAPX_WEIGHT_BODY_INERTIA 0
#define SAPX_WEIGHT_DIAG_DELASSUS 1

#define SAPX_POINT_WITNESS 0
#define SAPX_POINT_MIDPOINT 1

#define SAPX_INTEGRATE_MIDPOINT 0
#define SAPX_INTEGRATE_SAP_EULER 1

typedef struct SapxModel SapxModel;
typedef struct SapxSolver SapxSolver;
typedef struct SapxPipeline SapxPipeline;

/*
 * Model arrays, field-for-field the reference's SapModel
 * (sim/sap_runtime.py:47-137).  Host pointers; transforms are 7 floats
 * [p, q(xyzw)], inertia 9 floats row-major.  Worlds must be equally shaped
 * and world-major (joints, dofs, coords, bodies, shapes of world w are the
 * w-th contiguous block); global shapes (world -1, e.g. the ground plane)
 * follow all per-world shapes.
 */
typedef struct SapxModelDesc {
    int32_t world_count;
    int32_t joint_count;
    int32_t dof_count;
    int32_t coord_count;
    int32_t body_count;
    int32_t shape_count;
    int32_t articulation_count;
    int32_t pair_count;

    const int32_t* joint_type;
    const int32_t* joint_parent;
    const int32_t* joint_child;
    const int32_t* joint_q_start;   /* joint_count + 1 */
    const int32_t* joint_qd_start;  /* joint_count + 1 */
    const int32_t* joint_dof_dim;   /* joint_count x 2 */
    const int32_t* joint_ancestor;
    const int32_t* articulation_start; /* articulation_count + 1 */
    const float* joint_axis;        /* dof_count x 3 */
    const float* joint_X_p;         /* joint_count x 7 */
    const float* joint_X_c;         /* joint_count x 7 */
    const float* joint_armature;    /* dof_count */

    const float* body_mass;
    const float* body_inertia;      /* body_count x 9 */
    const float* body_com;          /* body_count x 3 */
    const float* gravity;           /* world_count x 3 */

    const int32_t* joint_target_mode;
    const float* joint_target_ke;
    const float* joint_target_kd;
    const float* joint_effort_limit;
    const float* joint_limit_lower;
    const float* joint_limit_upper;
    const float* joint_limit_ke;
    const float* joint_limit_kd;

    const int32_t* shape_body;
    const int32_t* shape_type;
    const int32_t* shape_world;
    const float* shape_transform;   /* shape_count x 7 */
    const float* shape_scale;       /* shape_count x 3 */
    const float* shape_margin;
    const float* shape_gap;
    const float* shape_collision_radius;
    const float* shape_material_mu;
    const float* shape_material_ke;
    const float* shape_material_tau;
    const int32_t* shape_pairs;     /* pair_count x 2, explicit candidate pairs */
} SapxModelDesc;

/*
 * Solver knobs: SolverSAP.__init__ keyword arguments after preset expansion
 * (sim/solver_sap.py:749-968, presets 75-126).
 */
typedef struct SapxSolverOpts {
    int32_t max_rigid_contact;      /* per-world contact capacity C */
    double fallback_mu;
    double fallback_stiffness;
    double fallback_tau_d;          /* per-shape dissipation fallback */
    double contact_beta;
    double contact_sigma;
    double diag_shift;
    int32_t max_iterations;
    double optimality_abs_tol;
    double optimality_rel_tol;
    double cost_abs_tol;
    double cost_rel_tol;
    int32_t line_search;            /* SAPX_LS_* */
    int32_t line_search_max_iterations;
    double armijo_c;
    double rho;
    double line_search_relative_slop;
    int32_t contact_weight_mode;    /* SAPX_WEIGHT_* */
    int32_t contact_point_mode;     /* SAPX_POINT_* */
    int32_t position_integration;   /* SAPX_INTEGRATE_* */
    int32_t capture_snapshots;      /* keep per-step internal buffers for sapx_solver_read
                                       (capture_contact_jacobian_snapshots, sim/solver_sap.py:778) */
} SapxSolverOpts;

/* Public-order state (SapState, sim/sap_runtime.py:177-205), device memory. */
typedef struct SapxStateView {
    float* joint_q;     /* coord_count */
    float* joint_qd;    /* dof_count, public order */
    float* body_q;      /* body_count x 7 */
    float* body_qd;     /* body_count x 6, may be NULL */
    float* body_f;      /* body_count x 6 public [f, tau] at COM, may be NULL */
} SapxStateView;

/* Public-order control (SapControl, sim/sap_runtime.py:208-227), device memory; NULL = zeros. */
typedef struct SapxControlView {
    const float* joint_f;
    const float* joint_target_pos;
    const float* joint_target_vel;
    const float* joint_act;
} SapxControlView;

/*
 * Flat rigid contacts (SapContacts / collision Contacts,
 * sim/sap_runtime.py:230-264, sim/collision/contacts.py:14-114).
 * count: 1 int (live count).  world_start: world_count+1 ints, CSR of a
 * world-grouped list (written by sapx_collide); when NULL the step groups
 * the flat list itself (stable, flat order within a world).
 */
typedef struct SapxContactsView {
    int32_t* count;
    int32_t* world_start;
    int32_t* shape0;
    int32_t* shape1;
    float* point0;      /* x 3, body frame of shape0's body */
    float* point1;      /* x 3, body frame of shape1's body */
    float* normal;      /* x 3, world, from shape0 to shape1 */
    float* margin0;
    float* margin1;
    int32_t capacity;
} SapxContactsView;

/* Library version (major*10000 + minor*100 + patch). */
int32_t sapx_version(void);

/* Message of the last error raised on this thread (never NULL). */
const char* sapx_last_error(void);

/*
 * Build the device-resident model: uploads arrays, validates equal world
 * shapes, builds the per-world topology template and level schedule.
 * Replaces model upload + host schedules in SapFreeMotion.__init__
 * (sim/free_motion.py:1427-1626), SapContactJacobian.__init__
 * (sim/contact_jacobian.py:1233-1540) and SapContactSolve.__init__
 * (sim/contact_solve.py:3291-3651).
 */
int32_t sapx_model_create(const SapxModelDesc* desc, SapxModel** out);
void sapx_model_destroy(SapxModel* model);

/* Solver instance (SolverSAP.__init__, sim/solver_sap.py:749-968). */
int32_t sapx_solver_create(const SapxModel* model, const SapxSolverOpts* opts, SapxSolver** out);
void sapx_solver_destroy(SapxSolver* solver);

/* Clear the warm-start guess (SolverSAP.reset_runtime_state, sim/solver_sap.py:1004-1017). */
int32_t sapx_solver_reset(SapxSolver* solver, void* stream);

/*
 * One SAP timestep for every world: SolverSAP.step(state_in, state_out,
 * control, contacts, dt) (sim/solver_sap.py:1388-1494) — boundary conversion,
 * free motion, contact Jacobian, Newton solve with line search, integration
 * and public forward kinematics, fused per world.  dt_world (device, f64,
 * world_count) overrides dt when non-NULL (sim/solver_sap.py:975-992).
 * state_in and state_out may alias.
 */
int32_t sapx_step(SapxSolver* solver, const SapxStateView* state_in, const SapxStateView* state_out,
                  const SapxControlView* control, const SapxContactsView* contacts, double dt,
                  const double* dt_world, void* stream);

/*
 * Per-world statistics of the last step, copied to HOST arrays of
 * world_count entries (SapContactSolveResult.newton_iterations_env etc.,
 * sim/contact_solve.py:188-232).  Any pointer may be NULL.  Synchronises the
 * stream.  flags bit0 = converged, bit1 = optimality condition reached,
 * bit2 = cost criterion reached, bit3 = iteration cap hit, bits 8..15 = last
 * line-search status (two's complement, 0 = ok).
 */
int32_t sapx_get_stats(SapxSolver* solver, int32_t* newton_iterations, int32_t* line_search_iterations,
                       int32_t* contact_count, int32_t* flags, void* stream);

/*
 * Accumulated Newton-iteration histogram over every world-step since the last
 * reset, copied to HOST: hist[0..255] = number of world-steps that took that
 * many Newton iterations (255 = 255 or more), hist[256] = total line-search
 * iterations, hist[257] = world-steps counted, hist[258] / hist[259] =
 * learned warm starts taken / rejected (260 entries).  The reference has no
 * counterpart (its per-world counters, contact_solve.py:224-229, are
 * overwritten every step); it serves the iteration statistics of the
 * benchmark without a per-step synchronisation.  Synchronises the stream;
 * reset != 0 clears the accumulators afterwards.
 */
int32_t sapx_get_histogram(SapxSolver* solver, int64_t* hist, int32_t reset, void* stream);

/*
 * Copy an internal per-world buffer of the last step into caller DEVICE
 * memory (parity/debug; SapContactJacobian.capture_local_snapshots,
 * sim/contact_jacobian.py:1611-1626).  which: 0 = v (f64 dof_count, SAP
 * order), 1 = v_star (f64 dof_count), 2 = v0 (f64 dof_count), 3 = contact
 * impulses gamma (f64 world_count x C x 3), 4 = contact phi0 (f64 world_count
 * x C), 5 = contact weight (f64 world_count x C), 6 = dynamics matrix A
 * diagonal (f64 dof_count), 7 = dynamics matrix A = H + diag(armature) as the
 * Newton solve uses it (f64 world_count x n x n, n = dofs per world), 8 =
 * contact Jacobian rows t1, t2, n of each slot (f64 world_count x C x 3 x n,
 * zero past the world's contact count; contact_jacobian.py:51-84
 * contact_env_jac).  bytes must equal the buffer size.  Requires a
 * solver created with capture_snapshots != 0 (SAPX_ERR_INVALID otherwise).
 */
int32_t sapx_solver_read(SapxSolver* solver, int32_t which, void* dst, int64_t bytes, void* stream);

/* Collision pipeline (SapCollisionPipeline.__init__, sim/collision/pipeline.py:322-508). */
int32_t sapx_pipeline_create(const SapxModel* model, int32_t max_contacts_per_world, SapxPipeline** out);
void sapx_pipeline_destroy(SapxPipeline* pipeline);

/*
 * Collision pipeline with the all-pairs broadphase: SapCollisionPipeline(model,
 * broad_phase="nxn") (sim/collision/pipeline.py:433-442; SapBroadPhaseAllPairs,
 * sim/collision/broad_phase_nxn.py:122-325, world map and filters
 * broad_phase_common.py:18-313).  broad_phase="sap" (sweep and prune,
 * broad_phase_sap.py) yields the same candidate set and binds here too.
 * Each world's slice is its shapes that carry the COLLIDE_SHAPES flag (bit 1
 * of shape_flags; NULL: every shape) followed by the colliding global shapes;
 * its pairs are enumerated once, in lower-triangular order (row r, columns
 * c > r), canonicalised to (min, max), and kept when their worlds and
 * collision groups may collide and the pair is not in filter_pairs
 * (filter_count x 2, any order).  sapx_collide then tests the kept pairs'
 * AABBs every call, in that deterministic order.  Global-global pairs are not
 * generated.  All arrays are host memory; shape_collision_group and
 * shape_flags have shape_count entries.
 */
int32_t sapx_pipeline_create_all_pairs(const SapxModel* model, int32_t max_contacts_per_world,
                                       const int32_t* shape_collision_group, const int32_t* shape_flags,
                                       const int32_t* filter_pairs, int32_t filter_count, SapxPipeline** out);

/*
 * Rigid contact generation: SapCollisionPipeline.collide(state, contacts)
 * (sim/collision/pipeline.py:546-709) — shape AABBs, the candidate pairs' AABB
 * test, the primitive narrowphase and the MPR / GJK path for capsule-box and
 * box-box pairs, written world-grouped in candidate-pair order
 * (deterministic).  body_q: device, body_count x 7.  Two kernel launches
 * (contacts, then their placement); the pipeline's staging area (world_count x
 * max_contacts_per_world x 52 bytes) carries the contacts between them, so
 * calls on one pipeline must not overlap in time (different pipelines may run
 * concurrently).  Graph-capturable.
 */
int32_t sapx_collide(SapxPipeline* pipeline, const float* body_q, const SapxContactsView* contacts, void* stream);

/*
 * sapx_collide without allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same contacts as
 * sapx_collide, bit for bit, but the contact staging area and per-world counts
 * live in the library's static device scratch (55 MB, held by every device that
 * loads the library until the process ends) instead of the pipeline's buffers.
 * Calls no hipMalloc / hipFree.  Limits: at most 65536 worlds and 2^20 contact
 * slots (worlds x max_contacts_per_world) per call (SAPX_ERR_CAPACITY beyond).
 * Every call on one device shares that scratch: calls that overlap in time on
 * one device race (serialise them on one stream).  Each call rewrites every
 * count and staged contact it reads, so nothing needs clearing between calls;
 * the scratch's unsupported-pair counter is not reported (pair kinds are checked
 * when the pipeline is created).
 */
int32_t sapx_collide_noalloc(SapxPipeline* pipeline, const float* body_q, const SapxContactsView* contacts,
                             void* stream);

/* ---------------------------------------------------------------------------
 * Learned warm start.  No counterpart in the reference: an MLP predicts a
 * correction of SolverSAP's plain initial velocity (the previous step's
 * solution, or v0) from state and contact features; the step starts from the
 * prediction only where its SAP cost is lower than the plain start's, and the
 * Newton solve runs to the unchanged tolerance either way.  The network is
 * trained on the GPU by the library's own forward / backward / Adam kernels.
 * ------------------------------------------------------------------------- */

/* Input features per world (the layout is fixed by the model's topology). */
int32_t sapx_solver_feature_count(const SapxSolver* solver);

/*
 * Features of every world into features (device, world_count x feature_count
 * float32): free joints (height, gravity direction and velocities in the body
 * frame), per-dof q / qd / target errors, per-body contact count, minimum
 * signed distance and summed normal, then the previous step's correction
 * prev_target (device, world_count x dof_per_world, as target_out of
 * sapx_solver_set_warm_start_io wrote it; NULL = zeros).  contacts must be
 * world-grouped (world_start set), as sapx_collide writes them.
 */
int32_t sapx_features(SapxSolver* solver, const SapxStateView* state_in, const SapxControlView* control,
                      const SapxContactsView* contacts, const float* prev_target, float* features, void* stream);

/*
 * Buffers used by the following sapx_step calls (device, world_count x
 * dof_per_world float32, or NULL): start_correction is added to the plain
 * start (free-joint rows in the joint's body frame) and taken where it lowers
 * the SAP cost; target_out receives converged v minus the plain start in the
 * same convention (training data).  Requires the one-wave step kernel.
 * The histogram of sapx_get_histogram counts learned starts taken (hist[258])
 * and rejected (hist[259]).
 */
int32_t sapx_solver_set_warm_start_io(SapxSolver* solver, const float* start_correction, float* target_out);

/* ---------------------------------------------------------------------------
 * Work schedule of sapx_step.  No counterpart in the reference (it has one code
 * path): how the worlds of a step map onto kernels, chosen per solver by an
 * autotuner that times the candidates on the caller's batch
 * (sap_warp_amd/autotune.py) or set directly.  Every schedule computes the
 * same step.
 * ------------------------------------------------------------------------- */
#define SAPX_KERNEL_WAVE 0  /* one 64-lane wave per world (dofs <= 64, capacity <= 64)  */
#define SAPX_KERNEL_FUSED 1 /* one 256-thread workgroup per world                       */

typedef struct SapxSchedule {
    int32_t kernel;         /* SAPX_KERNEL_*                                                   */
    int32_t small_capacity; /* wave: contact slots (16, 32 or 64) of the variant every world
                               with at most that many contacts runs in; the 64-slot variant
                               takes the others                                                 */
    int32_t fork;           /* wave: launch the 64-slot variant on a second stream beside it    */
    int32_t world_order;    /* wave: dispatch worlds most-expensive-first                       */
} SapxSchedule;

int32_t sapx_solver_get_schedule(const SapxSolver* solver, SapxSchedule* out);
/* SAPX_ERR_UNSUPPORTED when the model does not fit the requested kernel / variant. */
int32_t sapx_solver_set_schedule(SapxSolver* solver, const SapxSchedule* schedule);

/* ---------------------------------------------------------------------------
 * Batch step.  No counterpart in the reference (its step needs equally shaped
 * worlds, sim/contact_jacobian.py:1273-1281): the worlds of up to four solvers
 * (world kinds of different dof counts, at most 48 dofs per world, each with its
 * own model) advance in one kernel launch per contact variant, one wave per
 * world, each group reading its own kernel parameters.  Same per-world result
 * as sapx_step on each solver (identity world order).
 * ------------------------------------------------------------------------- */
typedef struct SapxBatch SapxBatch;

int32_t sapx_batch_create(SapxSolver* const* solvers, int32_t count, SapxBatch** out);
void sapx_batch_destroy(SapxBatch* batch);
/*
 * Contact capacity of the launch over every world (16, 32 or 64); with a solver
 * capacity above it, a second launch of the 64-slot variant takes the worlds
 * with more contacts, on a second stream forked from the caller's (fork != 0)
 * or after the first.  Re-uploads the bound slots (synchronous).
 */
int32_t sapx_batch_set_schedule(SapxBatch* batch, int32_t small_capacity, int32_t fork);
/*
 * Binds slot 0 or 1 to one set of per-solver views (arrays of `count` entries,
 * device memory as for sapx_step; contacts world-grouped) and a timestep, and
 * uploads the kernel parameters (synchronous).  Two slots serve the usual
 * ping-pong of two states.
 */
int32_t sapx_batch_bind(SapxBatch* batch, int32_t slot, const SapxStateView* state_in,
                        const SapxStateView* state_out, const SapxControlView* control,
                        const SapxContactsView* contacts, double dt);
/*
 * One step of every world of every solver with the views bound to `slot`.
 * Graph-capturable once every solver has stepped (the first step after a reset
 * re-uploads the slot's warm-start flags, which a capture refuses).
 */
int32_t sapx_batch_step(SapxBatch* batch, int32_t slot, void* stream);

typedef struct SapxMlp SapxMlp;

/* buffer ids of sapx_mlp_write / sapx_mlp_read */
#define SAPX_MLP_PARAMS 0 /* flat float32 parameters: per layer W (out x in, row-major) then b */
#define SAPX_MLP_X_MEAN 1 /* input normalisation: x_n = (x - mean) / std, in_features each   */
#define SAPX_MLP_X_STD 2
#define SAPX_MLP_Y_MEAN 3 /* output normalisation: y = y_n * std + mean, out_features each  */
#define SAPX_MLP_Y_STD 4
#define SAPX_MLP_GRADS 5  /* gradients of the last sapx_mlp_gradients call (read only)      */

/*
 * ReLU MLP with `layers` linear layers of padded widths dims[0..layers]
 * (multiples of 64); in_features / out_features are the real widths.  Weights
 * start at zero: write them with sapx_mlp_write(SAPX_MLP_PARAMS).
 */
int32_t sapx_mlp_create(int32_t layers, const int32_t* dims, int32_t in_features, int32_t out_features,
                        SapxMlp** out);
void sapx_mlp_destroy(SapxMlp* mlp);
int64_t sapx_mlp_param_count(const SapxMlp* mlp);
/* Host -> device copy of a buffer; writing the parameters also resets the Adam moments. */
int32_t sapx_mlp_write(SapxMlp* mlp, int32_t which, const float* host, int64_t count);
/* Device -> host copy (synchronises the device). */
int32_t sapx_mlp_read(const SapxMlp* mlp, int32_t which, float* host, int64_t count);
/* Pre-size the workspace for batches up to `batch` rows (otherwise grown on first use). */
int32_t sapx_mlp_reserve(SapxMlp* mlp, int32_t batch);
/* y (device, batch x out_features) = network(x (device, batch x in_features)), de-normalised. */
int32_t sapx_mlp_forward(SapxMlp* mlp, const float* x, int32_t batch, float* y, void* stream);
/*
 * Gradients of the mean squared error (normalised output units, real outputs
 * only) against target (device, batch x out_features); loss (device float, or
 * NULL) receives the loss.
 */
int32_t sapx_mlp_gradients(SapxMlp* mlp, const float* x, const float* target, int32_t batch, float* loss,
                           void* stream);
/* One bias-corrected Adam step with the last gradients. */
int32_t sapx_mlp_adam(SapxMlp* mlp, float lr, float beta1, float beta2, float eps, void* stream);

/* ---------------------------------------------------------------------------
 * Blocked Cholesky of large symmetric positive definite matrices.
 * Replaces BlockCholeskySolverBatched (sim/blocked_cholesky.py:725-901), whose
 * factorisation runs one workgroup per matrix: here each block column's panel
 * and Schur-complement update spread over the whole GPU (64 x 64 blocks,
 * float32, MFMA updates), for the dynamics / Newton matrices of worlds far
 * above the step kernels' 256 dofs.  Rows from n to the next multiple of 64
 * are identity (the reference's padding, blocked_cholesky.py:299-313).
 * ------------------------------------------------------------------------- */
typedef struct SapxChol SapxChol;

/*
 * Workspace for batch_size matrices of order up to max_num_equations
 * (BlockCholeskySolverBatched.__init__, blocked_cholesky.py:726-779): the
 * factor, a kept copy of the input (for float64-residual refinement) and the
 * inverted diagonal blocks.  ceil(max_num_equations / 64) * batch_size <= 512.
 */
int32_t sapx_chol_create(int32_t max_num_equations, int32_t batch_size, SapxChol** out);
void sapx_chol_destroy(SapxChol* chol);
/*
 * Factor A (device float32; batch item b, row i at A + b * a_stride + i * lda;
 * the full symmetric n x n matrix) with n <= max_num_equations active
 * equations (factorize, blocked_cholesky.py:781-809).  Enqueued on stream; no
 * allocation, so it can be captured in a graph.
 */
int32_t sapx_chol_factorize(SapxChol* chol, const float* A, int32_t n, int32_t lda, int64_t a_stride,
                            void* stream);
/*
 * result = A^-1 rhs for every batch item (solve, blocked_cholesky.py:811-823);
 * rhs / result device float64, batch_size x n contiguous, not aliased.
 * refine = number of extra sweeps r = rhs - A result (float64), result += A^-1 r.
 */
int32_t sapx_chol_solve(SapxChol* chol, const double* rhs, double* result, int32_t refine, void* stream);
/*
 * Synchronises, copies the per-matrix status of the last factorisation into
 * info (host, batch_size entries: 0, or 1 + the first column whose pivot was
 * not positive), and fails if a solve's wait for a block timed out.
 */
int32_t sapx_chol_info(SapxChol* chol, int32_t* info);

/* ---------------------------------------------------------------------------
 * SAP Newton solve of one large world (thousands of dofs, tens of thousands of
 * contacts) spread over the whole GPU.  Replaces the contact solve of the
 * reference (SapContactSolve, contact_solve.py:6129-6285, with the projection
 * 743-921, optimality and cost tests 1970-2087, exact-root 2913-3140 and
 * Armijo 2670-2905 line searches) for a world too large for the per-world
 * step kernels; the problem comes from that world's free motion and contact
 * Jacobian (contact_jacobian.py:51-84 fields).  Contact rows are sparse: each
 * contact lists at most kmax dofs (those of the two articulations it touches).
 * All arrays are device memory.
 * ------------------------------------------------------------------------- */
typedef struct SapxLarge SapxLarge;

typedef struct SapxLargeProblem {
    int32_t n;                      /* dofs */
    int32_t contact_count;          /* nc */
    int32_t kmax;                   /* dofs listed per contact (the solver's kmax) */
    int32_t pad_;
    const float* A;                 /* n x n dynamics matrix, row-major, full symmetric */
    const double* v_star;           /* n: free-motion velocity */
    const double* v_guess;          /* n: initial velocity (v0 or a warm start) */
    const int32_t* participating;   /* n: 1 where the dof takes part, else it is held at v* */
    const int32_t* J_cols;          /* nc x kmax dof indices, -1 = unused */
    const float* J_vals;            /* nc x 3 x kmax: rows t1, t2, n of each contact */
    const double* phi0;             /* nc signed distances */
    const double* weight;           /* nc contact weights w */
    const double* stiffness;        /* nc combined stiffness k */
    const double* tau;              /* nc dissipation time scales */
    const double* mu;               /* nc friction coefficients */
    /* PD / limit terms (contact_solve.py:490-741), all NULL or all set */
    const int32_t* pd_active;
    const double* pd_a;
    const double* pd_gain;
    const double* pd_limit;
    const int32_t* lo_active;
    const int32_t* up_active;
    const double* lo_vhat;
    const double* up_vhat;
    const double* lo_r;
    const double* up_r;
    double dt;
} SapxLargeProblem;

typedef struct SapxLargeOpts {
    int32_t line_search;            /* SAPX_LS_EXACT_ROOT or SAPX_LS_ARMIJO_DECAY */
    int32_t max_iterations;
    int32_t ls_max_iterations;
    int32_t refine;                 /* float64-residual sweeps of each Newton solve */
    double optimality_abs_tol;
    double optimality_rel_tol;
    double cost_abs_tol;
    double cost_rel_tol;
    double armijo_c;
    double rho;
    double relative_slop;
    double beta;
    double sigma;
    double diag_shift;
} SapxLargeOpts;

typedef struct SapxLargeStats {
    int32_t iterations;
    int32_t ls_iterations;
    int32_t converged;
    int32_t optimality_reached;
    int32_t status;                 /* last line search: 0 or its failure code */
    int32_t pad_;
    double cost;
    double grad_norm;
    double opt_tol;
} SapxLargeStats;

/* Workspace for worlds up to max_dofs dofs and max_contacts contacts; kmax <= 256. */
int32_t sapx_large_create(int32_t max_dofs, int32_t max_contacts, int32_t kmax, SapxLarge** out);
void sapx_large_destroy(SapxLarge* large);
/*
 * Solve; v_out (n) receives v, gamma_out (nc x 3, may be NULL) the impulses at
 * the last evaluation, stats (host) the iteration record.  Synchronises the
 * stream once per Newton iteration (the host reads the convergence test).
 */
int32_t sapx_large_solve(SapxLarge* large, const SapxLargeProblem* problem, const SapxLargeOpts* opts, double* v_out,
                         double* gamma_out, SapxLargeStats* stats, void* stream);

/* -------------------------------------------------------------------------
 * One world of free rigid bodies on the whole GPU (a bin of a thousand boxes, spheres
 * and capsules): the collision of SapCollisionPipeline.collide (sim/collision/pipeline.py)
 * and the step of SolverSAP.step (sim/solver_sap.py:1388-1494) for a single large world,
 * every stage spread over the GPU; its Newton solve is sapx_large_solve.  The model
 * (sapx_model_create) must hold one world whose joints are all free joints rooted at the
 * world with identity joint frames, one per body, without drives or limits
 * (SAPX_ERR_UNSUPPORTED otherwise); opts are the solver knobs (line searches exact_root
 * and armijo_decay).
 * ------------------------------------------------------------------------- */
typedef struct SapxBigWorld SapxBigWorld;

int32_t sapx_bigworld_create(const SapxModel* model, const SapxSolverOpts* opts, SapxBigWorld** out);
void sapx_bigworld_destroy(SapxBigWorld* big);
/* Forget the warm start: the next step's solve starts from v0 (SolverSAP.reset_runtime_state,
 * sim/solver_sap.py:1004-1017); after each step the solve starts from the last solution. */
void sapx_bigworld_reset(SapxBigWorld* big);
/*
 * Contacts of the model's explicit candidate pairs at body poses body_q (device,
 * body_count x 7) into the view (capacity >= max_rigid_contact), in candidate order with
 * the per-world capacity rule of sapx_collide; asynchronous on stream.
 */
int32_t sapx_bigworld_collide(SapxBigWorld* big, const float* body_q, const SapxContactsView* contacts,
                              void* stream);
/*
 * One step of the world on those contacts: state_out receives joint_q, joint_qd and (when
 * set) body_q / body_qd; v_out (dof_count doubles, SAP order) and gamma_out (contacts x 3)
 * are optional device outputs; stats (host) the Newton record.  Synchronises the stream
 * (contact count, then once per Newton iteration).
 */
int32_t sapx_bigworld_step(SapxBigWorld* big, const SapxStateView* state_in, const SapxStateView* state_out,
                           const SapxControlView* control, const SapxContactsView* contacts, float dt, double* v_out,
                           double* gamma_out, SapxLargeStats* stats, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SAPX_H */
