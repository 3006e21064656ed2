/*
 * mps.h - C ABI of the MI355X (gfx950) implementation of the OpenMps
 * per-timestep MPS path (Computer::ForwardTime).
 *
 * Plain C: opaque context, plain pointers and sizes, int status codes.
 * Every call is synchronous at return.  One host thread per context; a context
 * owns one HIP stream and every device buffer it uses.  Host pointers passed
 * in are only borrowed for the duration of the call.
 *
 * Each declaration cites the reference interface it replaces (paths relative
 * to the OpenMps source tree, src/OpenMps/).
 */
#ifndef MPS_H_
#define MPS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPS_ABI_VERSION 1

/* Particle::Type, Particle.hpp:16-29 */
enum mps_particle_type { MPS_FLUID = 0, MPS_WALL = 1, MPS_DUMMY = 2, MPS_DISABLED = 3 };

/* Status codes.  The reference throws instead:
 *   MPS_ERR_CG_NOT_CONVERGED <- Computer::Exception, Computer.hpp:1424-1428
 *   MPS_ERR_CELL_OVERFLOW    <- Grid::Exception,     Grid.hpp:314-318      */
enum mps_status {
  MPS_OK = 0,
  MPS_ERR_CG_NOT_CONVERGED = 1,
  MPS_ERR_CELL_OVERFLOW = 2,
  MPS_ERR_HIP = 3,
  MPS_ERR_BAD_ARG = 4,
  MPS_ERR_NO_DEVICE = 5,
  MPS_ERR_STATE = 6
};

/* Environment (Environment.hpp:9-256).  The first block holds the constructor
 * arguments, the second the derived constants the constructor computes. */
typedef struct mps_env {
  double max_dt_arg, courant, g, rho, nu, r_e_by_l0, l0, min_x, min_z, max_x, max_z;
  double r_lap_by_l0; /* extension: Laplacian radius / l0 (== r_e_by_l0 in the reference) */
  double max_dt;          /* MaxDt,  Environment.hpp:138 */
  double max_dx;          /* MaxDx,  Environment.hpp:143 */
  double r_e;             /* R_e,    Environment.hpp:145 */
  double r_lap;           /* extension: Laplacian radius (== r_e in the reference) */
  double gx, gz;          /* G,      Environment.hpp:149 */
  double neighbor_length; /* NeighborLength, Environment.hpp:161 */
  double n0;              /* N0(),   Environment.hpp:164-209 */
  double n0_lap;          /* extension: lattice sum at r_lap (== n0 in the reference) */
} mps_env;

/* Environment::Environment(maxDt, courant, g, rho, nu, r_eByl_0, l_0, minX,
 * minZ, maxX, maxZ), Environment.hpp:101-216 (default macro set, DIM = 2). */
int mps_env_init(mps_env* env, double max_dt, double courant, double g, double rho, double nu,
                 double r_e_by_l0, double l0, double min_x, double min_z, double max_x,
                 double max_z);
/* Extension: separate Laplacian radius for the viscosity term and the PPE
 * matrix; recomputes r_lap, n0_lap and neighbor_length. */
int mps_env_set_laplacian_radius(mps_env* env, double r_lap_by_l0);

typedef struct mps_ctx mps_ctx;

/* Computer(allowableResidual, env, posWall, posWallPre), Computer.hpp:1674-1691
 * and CreateComputer, Computer.hpp:1800-1815.  Selects HIP device `device`. */
int mps_create(const mps_env* env, double allowable_residual, int device, mps_ctx** out);
/* ~Computer */
void mps_destroy(mps_ctx* ctx);
/* what() of the last error raised on this context ("" if none). */
const char* mps_last_error(const mps_ctx* ctx);
/* Computer::GetEnvironment, Computer.hpp:1786-1789 */
int mps_get_env(const mps_ctx* ctx, mps_env* out);
/* Grid shape, Grid.hpp:137-150: out = {nx, nz, cell capacity, max neighbours} */
int mps_grid_shape(const mps_ctx* ctx, int64_t* out4);

/* Computer::AddParticles, Computer.hpp:1754-1777 (appends; the appended
 * positions also become the static wall positions, Main.cpp:305-318). */
int mps_add_particles(mps_ctx* ctx, size_t count, const double* x, const double* z,
                      const double* u, const double* w, const double* p, const double* n,
                      const uint8_t* type);
/* Particles().size() */
size_t mps_size(const mps_ctx* ctx);
/* Computer::Particles, Computer.hpp:1780-1783.  NULL pointers are skipped. */
int mps_get_particles(mps_ctx* ctx, double* x, double* z, double* u, double* w, double* p,
                      double* n, uint8_t* type);
/* Overwrites particle state in place (test hook; the reference tests write
 * Particles() directly through friend access).  NULL pointers are skipped. */
int mps_set_particles(mps_ctx* ctx, const double* x, const double* z, const double* u,
                      const double* w, const double* p, const double* n, const uint8_t* type);
/* Environment::T / Dt, Environment.hpp:225-242 */
int mps_get_time(const mps_ctx* ctx, double* t, double* dt);
/* Sets the clock (the reference tests do Dt() = dt; SetNextT()). */
int mps_set_time(mps_ctx* ctx, double t, double dt);

/* positionWall(i, t, dt) / positionWallPre(t, dt), Computer.hpp:993 and
 * :1013-1019.  Without callbacks every non-fluid particle returns to its
 * position at mps_add_particles, as Main.cpp:312-318 does. */
typedef void (*mps_wall_pre_fn)(void* user, double t, double dt);
typedef void (*mps_wall_fn)(void* user, double t, double dt, size_t count, const uint32_t* index,
                            double* x_out, double* z_out);
int mps_set_wall_callbacks(mps_ctx* ctx, mps_wall_fn wall, mps_wall_pre_fn pre, void* user);

/* Device-side wall kinematics, in place of positionWall callbacks
 * (Computer.hpp:1013-1019) for prescribed motions; no counterpart in the
 * reference, whose walls move only through the callbacks.  Each step the
 * non-fluid particle i of group g = group[i] >= 1 is moved to its position
 * at mps_add_particles plus
 *   (vx t + ax sin(omega t + phase), vz t + az sin(omega t + phase))
 * of motions[g - 1], t the clock after the step's advance (the t
 * positionWall receives), and its velocity becomes the displacement over dt,
 * as in the reference; group 0 (and particles added later) stays put.
 * Clears the wall callbacks; n_motions = 0 switches the motion off.  group
 * has one entry per particle present (mps_size). */
typedef struct mps_wall_motion {
  double vx, vz, ax, az, omega, phase;
} mps_wall_motion;
int mps_set_wall_motion(mps_ctx* ctx, size_t n_motions, const mps_wall_motion* motions, const uint8_t* group);

/* Computer::ForwardTime(dt), Computer.hpp:1700-1742 */
int mps_forward_time(mps_ctx* ctx, double dt);
/* Computer::ForwardTime() with Computer::DetermineDt, Computer.hpp:1745-1751, :759-777 */
int mps_forward_time_cfl(mps_ctx* ctx, double* dt_out);
/* Computer::DetermineDt alone */
int mps_determine_dt(mps_ctx* ctx, double* dt_out);

/* ---- pressure solver selection (extension; default = the reference CG) ---- */
enum mps_solver_kind {
  MPS_SOLVER_CG_FP64 = 0,  /* Computer.hpp:1359-1429, FP64 throughout */
  MPS_SOLVER_MIXED_IR = 1  /* FP32 CG inside FP64 iterative refinement */
};
enum mps_stop_rule {
  MPS_STOP_INITIAL_RESIDUAL = 0, /* r.r < tol^2 r0.r0 (Computer.hpp:1386, :1408) */
  MPS_STOP_RHS_NORM = 1          /* |b - A x| <= tol |b|, residual in FP64 */
};
enum mps_warm_start {
  MPS_WARM_PREVIOUS = 0, /* x0 = p of the previous step (Computer.hpp:1217-1219) */
  MPS_WARM_ZERO = 1,
  MPS_WARM_LEARNED = 2   /* x0 from the learned per-particle model */
};
/* Preconditioner of the mixed solver's FP32 inner CG (the FP64 CG, the
 * reference algorithm, is never preconditioned). */
enum mps_precond {
  MPS_PRECOND_NONE = 0,
  MPS_PRECOND_JACOBI = 1,
  /* Jacobi plus additive piecewise-constant corrections on the neighbour
   * search cells and on 2x2 and 4x4 blocks of cells; needs a system assembled
   * from the current particle state and cells, and uses Jacobi otherwise */
  MPS_PRECOND_MULTILEVEL = 2,
  /* Jacobi plus a multigrid correction on the cells: one V-cycle on the
   * Galerkin operator of the cells (a 9-point stencil on the cell grid) and
   * of 2x2 blocks of cells below it, added to the damped Jacobi step; same
   * requirements as MPS_PRECOND_MULTILEVEL, and uses Jacobi otherwise */
  MPS_PRECOND_AMG = 3,
  /* the faster of MPS_PRECOND_MULTILEVEL and MPS_PRECOND_AMG for the flow at
   * hand (the first is cheaper per iteration, the second needs far fewer
   * iterations once the flow has developed): the first two solves time one
   * each (solve_ms), the faster is used and the other timed again after 16
   * solves, then after 32, 64, ... 1024 while it keeps losing; decided when
   * the system is assembled; mps_solve_stats.precond_used reports it */
  MPS_PRECOND_AUTO = 4
};
typedef struct mps_solver_opts {
  int kind;          /* mps_solver_kind */
  int stop_rule;     /* mps_stop_rule */
  int warm_start;    /* mps_warm_start */
  int precond;       /* mps_precond, mixed solver only */
  double tol;        /* <= 0: the allowable residual given to mps_create */
  int64_t max_iter;  /* <= 0: number of unknowns, as the reference */
  double inner_tol;  /* mixed: relative residual reduction per refinement sweep */
  int64_t inner_max_iter;
} mps_solver_opts;
int mps_set_solver(mps_ctx* ctx, const mps_solver_opts* opts);
int mps_get_solver(const mps_ctx* ctx, mps_solver_opts* out);

typedef struct mps_solve_stats {
  int64_t iterations;       /* CG iterations (inner ones summed for the mixed solver) */
  int64_t outer_iterations; /* refinement sweeps (mixed), 1 otherwise */
  double rr0;               /* r0.r0 of the warm start */
  double rr;                /* final r.r of the recursion */
  double rel_residual;      /* |b - A x| / |b| recomputed in FP64 after the solve */
  double initial_rel_residual; /* |b - A x0| / |b| */
  double solve_ms;          /* device time of the solve (HIP events), warm start included */
  int64_t learned_used;     /* 1: the learned guess replaced the previous-step guess */
  double prev_rel_residual; /* |b - A p_prev| / |b| when the learned guess was evaluated (A p_prev in
                               FP32 for systems assembled from particles) */
  int64_t fallback;         /* mixed: 1 when a refinement sweep stalled and the FP64 CG finished */
  int64_t precond_used;     /* mixed: the preconditioner the FP32 CG ran with (mps_precond) */
} mps_solve_stats;
/* The last solve's statistics.  solve_ms is read from the solve's HIP events
   on the first call after the solve (the solve does not wait for its end
   event), so this call can return MPS_ERR_HIP; the last error message is kept
   otherwise. */
int mps_get_solve_stats(const mps_ctx* ctx, mps_solve_stats* out);

/* ---- stages (the friend-class test hooks of Computer.hpp:437-460) ---- */
int mps_stage_search(mps_ctx* ctx);            /* SearchNeighbor, Computer.hpp:698-756 */
int mps_stage_density(mps_ctx* ctx);           /* ComputeNeighborDensities, :780-833 */
int mps_stage_error_correction(mps_ctx* ctx);  /* ComputeErrorCorrection, :877-910 */
int mps_stage_explicit(mps_ctx* ctx);          /* ComputeExplicitForces, :914-1021 */
int mps_stage_save_x(mps_ctx* ctx);            /* SaveX, :1025-1039 */
int mps_stage_ppe_assemble(mps_ctx* ctx);      /* SetPressurePoissonEquation, :1145-1356 */
int mps_stage_ppe_solve(mps_ctx* ctx);         /* SolvePressurePoissonEquation, :1359-1429 */
int mps_stage_pressure_writeback(mps_ctx* ctx);/* ComputeImplicitForces p = max(x,0), :1076-1097 */
int mps_stage_pressure_gradient(mps_ctx* ctx); /* ModifyByPressureGradient, :1433-1564 */
int mps_stage_dynamic_stabilize(mps_ctx* ctx); /* DynamicStabilize, :1566-1657 */
/* NeighborDensityVariationSpeed(i) for every i (HS), Computer.hpp:838-872 */
int mps_density_variation(mps_ctx* ctx, double* out);

/* ---- state getters for parity checks ---- */
/* neighbor[i][0], Computer.hpp:580-588 */
int mps_get_neighbor_counts(mps_ctx* ctx, uint32_t* counts);
/* neighbor[i][1..count], concatenated in particle order (CSR by the counts). */
int mps_get_neighbors(mps_ctx* ctx, uint32_t* ids);
enum mps_aux {
  MPS_AUX_N_WITHOUT_SPP = 0, /* nWithoutSpp, Computer.hpp:530-533 */
  MPS_AUX_ECS = 1,           /* ecs, :520-523 */
  MPS_AUX_SAVED_X = 2,       /* originalX[.][0], :525-528 */
  MPS_AUX_SAVED_Z = 3,       /* originalX[.][1] */
  MPS_AUX_DU_X = 4,          /* du[.][0], :515-518 */
  MPS_AUX_DU_Z = 5,          /* du[.][1] */
  MPS_AUX_PPE_B = 6,         /* ppe.b */
  MPS_AUX_PPE_X = 7          /* ppe.x */
};
int mps_get_aux(mps_ctx* ctx, int which, double* out);
int mps_set_aux(mps_ctx* ctx, int which, const double* in);
/* ppe.A as CSR with ascending columns per row, the layout viennacl::copy
 * produces (Computer.hpp:1352-1355). */
int mps_ppe_shape(mps_ctx* ctx, size_t* n, size_t* nnz); /* unknowns, stored entries */
int mps_get_ppe(mps_ctx* ctx, uint64_t* rowptr, uint32_t* col, double* val);
/* Installs an arbitrary system for the solve stage (the CG tests write
 * ppe.A, ppe.b and ppe.x directly, test_ComputerConjugateGradient.cpp:115-229). */
int mps_set_ppe(mps_ctx* ctx, size_t n, const uint64_t* rowptr, const uint32_t* col,
                const double* val, const double* b, const double* x0);

/* ---- learned warm start (no reference counterpart) ----
 * A per-particle MLP (132 -> 256 -> 256 -> 1, GELU, BF16 MFMA with FP32
 * accumulation) predicts p - p_prev from the features of the 32 nearest
 * stencil neighbours; see mps_mlp.hip.  Parameter layouts on the host:
 * w1 [256][132], b1 [256], w2 [256][256], b2 [256], w3 [256], b3 [1]. */
typedef struct mps_model mps_model;
typedef struct mps_dataset mps_dataset;
int mps_model_create(int device, int seed, mps_model** out);
void mps_model_destroy(mps_model* m);
int mps_model_get_params(mps_model* m, float* w1, float* b1, float* w2, float* b2, float* w3,
                         float* b3);
int mps_model_set_params(mps_model* m, const float* w1, const float* b1, const float* w2,
                         const float* b2, const float* w3, const float* b3);
/* y = MLP(x) for `rows` feature rows x [rows][132] (host arrays). */
int mps_model_forward(mps_model* m, size_t rows, const float* x, float* y);
/* Mean-squared-error loss of (x, t) and its gradient for every parameter. */
int mps_model_gradients(mps_model* m, size_t rows, const float* x, const float* t, float* loss,
                        float* g_w1, float* g_b1, float* g_w2, float* g_b2, float* g_w3,
                        float* g_b3);
/* Adam over shuffled mini-batches of the dataset; losses_out[epochs]. */
int mps_model_train(mps_model* m, mps_dataset* d, int epochs, int batch, double lr, uint64_t seed,
                    double* losses_out);
int mps_dataset_create(int device, size_t capacity, mps_dataset** out);
void mps_dataset_destroy(mps_dataset* d);
int mps_dataset_size(mps_dataset* d, size_t* n);
int mps_dataset_get(mps_dataset* d, size_t first, size_t count, float* x, float* t);
/* Use the model for MPS_WARM_LEARNED (NULL detaches). */
int mps_attach_model(mps_ctx* ctx, mps_model* m);
/* At every pressure solve, append (features, (p - p_prev) / s_d) of a
 * `fraction` of the alive rows to the dataset (NULL detaches). */
int mps_attach_dataset(mps_ctx* ctx, mps_dataset* d, double fraction, uint64_t seed);
/* Features of the current state (computes them): out [n][132].  Needs the pressure system
 * assembled from the particles (MPS_ERR_STATE otherwise). */
int mps_get_features(mps_ctx* ctx, float* out);

/* Storage order of the particles on the device (extension, no reference
 * counterpart; every getter and setter works in the caller's order either
 * way).  MPS_ORDER_CELL (default): each neighbour search stores the particles
 * in cell order (cells in key order, ascending caller index inside a cell), so
 * neighbours sit close together in memory whatever order the caller used.
 * MPS_ORDER_KEEP: the current order is kept (the caller's order unless a
 * search already permuted it).  Results are the same bit for bit up to the
 * solver's own rounding (its sums run in storage order); a particle that
 * leaves the grid reports an empty neighbour list in cell order (the reference
 * keeps its last list; no stage reads it). */
enum mps_storage_order { MPS_ORDER_CELL = 0, MPS_ORDER_KEEP = 1 };
int mps_set_storage_order(mps_ctx* ctx, int order);

/* ---- several scenes stepped together (extension, no reference counterpart) ----
 * A batch steps independent scenes (contexts on one device) together.  Each
 * step runs the stages before and after the pressure solve per scene on the
 * scenes' own streams from `host_threads` host threads (0: the calling
 * thread), and the solves of every scene on the mixed solver with the
 * multilevel preconditioner as one set of grouped launches: each kernel of
 * the solve is one launch over all scenes still iterating, every block
 * working on one scene with that scene's own alpha, beta and stop test
 * (block-diagonal CG with per-scene convergence).  Scenes on the learned warm
 * start share one inference launch over all their feature rows (they must
 * have the same model attached).  Scenes on other solvers solve on their own.
 * Results per scene are those of its own ForwardTime up to the solver's
 * rounding.  The batch borrows the contexts (destroy it first) and the
 * contexts must not be used by other threads during a batch call. */
typedef struct mps_batch mps_batch;
int mps_batch_create(int device, int host_threads, mps_batch** out);
void mps_batch_destroy(mps_batch* b);
const char* mps_batch_last_error(const mps_batch* b);
int mps_batch_add(mps_batch* b, mps_ctx* scene);
/* Computer::ForwardTime for every scene (Computer.hpp:1700-1751): dt > 0 the
 * same fixed step for all; dt <= 0 each scene's ForwardTime() with its own
 * DetermineDt step (:1745-1751, :759-777), written to dt_out[s] if dt_out is
 * not NULL.  An error names the scene ("scene k: ..."); the scenes without an
 * error have completed their step. */
int mps_batch_forward_time(mps_batch* b, double dt, double* dt_out);
/* The last batched solve: its duration (HIP events on the batch's stream) and
 * the number of grouped launches it made. */
int mps_batch_solve_info(const mps_batch* b, double* ms, int64_t* launches);

/* ---- measurement (no reference counterpart) ---- */
enum mps_kernel_id {
  MPS_K_CELL_KEYS = 0, MPS_K_CELL_SCAN, MPS_K_CELL_SCATTER, MPS_K_CELL_SORT, MPS_K_NEIGHBORS,
  MPS_K_DENSITY_ECS, MPS_K_EXPLICIT, MPS_K_UPDATE, MPS_K_DENSITY_SAVE, MPS_K_PPE_COUNT,
  MPS_K_PPE_FILL, MPS_K_CG_INIT, MPS_K_CG_SPMV, MPS_K_CG_UPDATE, MPS_K_CG_ITER32,
  MPS_K_CG_INIT32, MPS_K_IR_RESIDUAL, MPS_K_GRADIENT, MPS_K_DS, MPS_K_OTHER,
  MPS_K_FEATURES, MPS_K_MLP, MPS_K_ML_SETUP, MPS_K_ML_PREC, MPS_K_CG_SPMV32, MPS_K_CG_UPDATE32,
  MPS_K_AMG_CYCLE, MPS_K_COUNT
};
typedef struct mps_kernel_stat {
  const char* name;
  int64_t launches;
  double total_ms;    /* summed HIP-event durations on the context stream */
  double total_bytes; /* summed algorithmic HBM bytes of those launches */
} mps_kernel_stat;
/* 1: bracket every tracked launch with HIP events on the context stream. */
int mps_set_profiling(mps_ctx* ctx, int on);
int mps_get_kernel_stat(mps_ctx* ctx, int id, mps_kernel_stat* out);
int mps_reset_kernel_stats(mps_ctx* ctx);
/* Launches the last solve's SpMV kernel again `reps` times back to back on
 * the context stream, between two HIP events, with the operands that solve
 * left (id MPS_K_CG_SPMV32: the multilevel inner CG's, MPS_K_CG_SPMV: the
 * FP64 CG's), and returns the mean duration of one launch.  A measurement
 * hook with no counterpart in the reference: the search direction and its
 * product are overwritten (the next solve rebuilds them); the solution is
 * not touched.  MPS_ERR_STATE if the last solve did not run that kernel. */
int mps_replay_spmv(mps_ctx* ctx, int id, int reps, double* ms_per_launch);

/* Copies `bytes` bytes between two device buffers with a plain streaming
 * kernel and returns its duration: the known-size reference for scaling HBM
 * traffic counters (FETCH_SIZE / WRITE_SIZE) to bytes. */
int mps_device_copy_probe(size_t bytes, double* ms_out);

/* ---- allocation-free solve (no reference counterpart) ----
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Solves the
 * caller's system A x = b (CSR in device memory: rowptr[n + 1], col and val
 * [rowptr[n]], the diagonal stored as one of the row's entries) with the FP64
 * CG of mps_set_ppe + mps_stage_ppe_solve: the same kernels, the reference's
 * stop rule |r| < allowable_residual |r0| (Computer.hpp:1386, :1408), x the
 * initial guess on entry and the solution on return, at most max_iter
 * iterations (<= 0: n).  b, x, rowptr, col and val are device pointers; the
 * work runs on `stream` (a hipStream_t; NULL: the default stream) of `device`,
 * and the call returns once it has finished.  Allocates nothing: r, p, q, the
 * partial sums and the status live in a fixed array of the library that each
 * device keeps until the process ends (3 MPS_NOALLOC_MAX_ROWS + 3072 doubles,
 * 24 MiB), so n <= MPS_NOALLOC_MAX_ROWS (MPS_ERR_BAD_ARG above, with the
 * message in err).  Every call on a device shares that array: calls that
 * overlap on one device race.  `err` (may be NULL) receives the message of a
 * failed call; `iterations` (may be NULL) the CG iterations.  Returns
 * MPS_ERR_CG_NOT_CONVERGED as the stage does. */
#define MPS_NOALLOC_MAX_ROWS 1048576
int mps_cg_solve_csr_noalloc(int device, size_t n, const uint64_t* rowptr, const uint32_t* col,
                             const double* val, const double* b, double* x, double allowable_residual,
                             int64_t max_iter, void* stream, int64_t* iterations, char* err,
                             size_t err_len);

/* Device arithmetic self-check: out_sqrt[i] = sqrt(a[i]), out_div[i] = a[i]/b[i]
 * computed by the same code the kernels use (pins rounding against the host). */
int mps_device_math_check(size_t n, const double* a, const double* b, double* out_sqrt,
                          double* out_div);

#ifdef __cplusplus
}
#endif
#endif /* MPS_H_ */
