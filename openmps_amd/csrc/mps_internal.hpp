// Internal declarations of the MI355X MPS library: the context, device views
// and the launch helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mps.h"

namespace mps {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlock = 256;        // threads per block of the particle passes
constexpr int kRedBlocks = 1024;   // fixed partial-sum slots of the reduction kernels
constexpr int kRedOut = 16;        // reduced scalars (red_out, red_host)

// Particle type codes, Particle.hpp:16-29.
constexpr uint8_t kFluid = 0, kWall = 1, kDummy = 2, kDisabled = 3;
constexpr uint32_t kNoCell = 0xffffffffu;  // cell key of a particle outside the grid

// Device-side CG scalars.  rr is double-buffered by iteration parity so that
// block 0 of a kernel can publish the next value while the other blocks of
// the same kernel still read the previous one.
struct CgStatus {
  double rr[2];
  double gamma[2];    // FP32 inner solver: (r, M r) and alpha by iteration parity
  double alpha[2];
  double alpha_ml;    // multilevel inner solver: alpha of the running iteration
  double tol;
  double rr0;
  double b2;          // |b|^2 (stop rule on |b|)
  int converged;      // 1 once the stop rule holds
  int pad;
  long long iterations;
  long long stop_mark;  // multilevel inner solver: 0 running, k + 1 once it stopped at iteration k
};

// Mixed-precision refinement scalars.
struct IrStatus {
  double rr_outer;    // |b - A x|^2 in FP64
  double b2;
  double rr0;
  int done;
  int pad;
};

constexpr int kAmgMaxLevels = 20;  // grid levels of the cell-grid multigrid

constexpr int kFeat = 132;     // warm-start feature width
constexpr int kFeatPad = 160;  // padded to the MFMA K step (32)
constexpr int kHid = 256;      // hidden width
constexpr int kNbr = 32;       // neighbours per feature row

struct Dataset {
  size_t cap = 0;
  __bf16* x = nullptr;             // cap x kFeatPad
  float* t = nullptr;              // cap
  unsigned long long* count = nullptr;
  std::string err;
  void init(size_t capacity);
  size_t size() const;
  void release();
};

// MLP 132 -> 256 -> 256 -> 1 (GELU).  Parameters k = 0..5: w1 [256][160]
// (columns >= 132 zero), b1 [256], w2 [256][256], b2 [256], w3 [256], b3 [1],
// FP32 masters with Adam moments, BF16 copies of w1 / w2 for the GEMMs.
struct Model {
  hipStream_t stream = nullptr;
  float* w[6] = {};
  float* g[6] = {};
  float* m1[6] = {};
  float* m2[6] = {};
  __bf16 *w1b = nullptr, *w2b = nullptr;
  __bf16* w2p = nullptr;  // w2 in the K order of the fused inference kernel
  float* loss = nullptr;
  long long adam_t = 0;
  int cus = 0;     // compute units of the device (persistent inference grid)
  size_t cap = 0;  // batch work buffers
  __bf16 *xb = nullptr, *h1b = nullptr, *da2b = nullptr, *da1b = nullptr;
  float *tb = nullptr, *z1 = nullptr, *z2 = nullptr, *a2 = nullptr, *y = nullptr, *dy = nullptr,
        *dh1 = nullptr;
  uint32_t* idx = nullptr;
  std::string err;
  static size_t param_count(int k);
  void init(int seed);
  void release();
  void refresh_bf16();
  void ensure_batch(size_t m);
  void forward(hipStream_t s, size_t m, const __bf16* x, float* out, bool train);
  void backward(hipStream_t s, size_t m);
  void adam(hipStream_t s, float lr);
  void pack_host_rows(size_t m, const float* x_host);
  void train(Dataset& ds, int epochs, int batch, float lr, uint64_t seed, double* losses);
};

struct KStat {
  long long launches = 0;
  double ms = 0, bytes = 0;
};

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  int kid = 0;
  long long tag = 0;   // CG iteration of the launch (0 outside the solver)
  double bytes = 0;
};

struct Context {
  mps_env env{};
  double eps = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  double t = 0, dt = 0;

  // particles (SoA, FP64 as the reference)
  size_t n = 0, alloc_n = 0;
  double *x = nullptr, *z = nullptr, *u = nullptr, *w = nullptr, *p = nullptr, *nd = nullptr;
  uint8_t* type = nullptr;
  double *nws = nullptr, *ecs = nullptr, *ox = nullptr, *oz = nullptr;
  double *dux = nullptr, *duz = nullptr, *wall_x = nullptr, *wall_z = nullptr;
  double* hs_out = nullptr;

  // uniform grid of cells of side NeighborLength (Grid.hpp:137-150)
  long long gnx = 0, gnz = 0;
  uint32_t cap = 0, kmax = 0;  // cell capacity; list bound 9 cap (Computer.hpp:1767-1770)
  uint32_t kstride = 0;        // list slots per particle: kmax rounded up to a multiple of 4
  size_t ncells = 0;
  uint32_t *cell_count = nullptr, *cell_start = nullptr, *cell_fill = nullptr;
  uint32_t *key = nullptr, *sorted = nullptr;
  double2* pos_sorted = nullptr;  // (x, z) in cell order (the search's candidate ranges)
  uint64_t* in_cells = nullptr;    // particles in cells (device scalar, from the cell scan)
  uint64_t* scan_tmp = nullptr;   // block sums of the scans
  int* flags = nullptr;           // [0] cell overflow, [1] neighbour overflow
  int* flags_host = nullptr;      // pinned mirror

  // storage order: the particles are kept in cell order, permuted at every
  // neighbour search (cells in key order, ascending caller index inside a
  // cell, particles outside the grid last); orig[k] is the caller's index of
  // the particle stored at slot k.  Every per-particle array on the device is
  // in storage order; the API maps at the boundary (storage_order).
  uint32_t* orig = nullptr;
  long long order_serial = 0;       // permutations applied (0: storage order = caller order)
  bool keep_order = false;          // MPS_ORDER_KEEP: the search does not permute
  std::vector<uint32_t> orig_host;  // host copy of orig for order_host_serial
  long long order_host_serial = -1;
  // arrays the permutation writes into, then swapped with the live ones
  static constexpr int kMoved = 14;  // FP64 per-particle arrays the permutation moves
  double* spare_d[kMoved] = {};
  uint8_t *spare_type = nullptr, *spare_group = nullptr;
  uint32_t *spare_orig = nullptr, *spare_key = nullptr;

  // neighbour lists: slot k of particle i at nbr_slot(i, k, kstride) (groups
  // of four slots per particle, a slice's 64 particles side by side)
  uint32_t *nbr = nullptr, *ncount = nullptr;
  bool have_neighbors = false;
  long long search_serial = 0;  // neighbour searches run so far (cells of the last one in key / sorted)
  unsigned long long* nbr_total_dev = nullptr;
  unsigned long long* nbr_total_host = nullptr;  // pinned
  double nbr_total = 0;                          // sum of list lengths (byte accounting)

  // PPE in sliced ELL (SELL-64): diagonal kept dense, off-diagonals of row i
  // at slice_off[i/64] + k*64 + i%64 for k < slice_width[i/64]
  size_t nslices = 0, sell_cap = 0;
  uint32_t *row_len = nullptr, *slice_width = nullptr;
  uint64_t* slice_off = nullptr;
  double* sval = nullptr;
  float* sval32 = nullptr;
  uint32_t* scol = nullptr;
  double *diag = nullptr, *b = nullptr, *sol = nullptr;
  float* diag32 = nullptr;
  size_t ppe_n = 0;        // unknowns of the installed system
  size_t ppe_alloc = 0;    // capacity of the PPE / CG vectors
  bool have_ppe = false;
  bool have_ppe32 = false;
  // the system was assembled from the current particle positions, so the
  // FP32 solver may recompute off-diagonals from them (matrix-free)
  bool ppe_from_particles = false;
  bool ppe_col16 = false;  // the system's 16-bit column offsets (col16, slice_wide) are set: assembled by the fill
  long long ppe_search_serial = -1;  // search whose cells the assembled system matches
  // (x, z) and (u, w) as 16-byte pairs for the neighbour walks' gathers (one
  // load where two 8-byte ones would be): packed at the start of the step's
  // first walk (pack_state) and written through by the explicit update; the
  // PPE assembly reads them (the staged assembly packs first)
  double2* xz = nullptr;
  double2* uw = nullptr;
  float* ghost32 = nullptr;  // SPP ghost share of the diagonal (1 on rows without unknowns)
  uint2* qpos = nullptr;     // fixed-point positions of the matrix-free operator
  double qpos_inv_h = 0;     // their scale: q = (x - min) / h
  // compressed columns of the matrix-free FP32 SpMV: the fill stores every
  // column also as the int16 offset j - i in the same SELL slot; a slice with
  // an offset outside int16 ("wide") is read through the u32 columns
  uint8_t* slice_wide = nullptr;
  int16_t* col16 = nullptr;  // sell_cap entries, beside scol
  double mf_nnz_wide = 0;
  bool mf_wide_pending = false;  // mf_nnz_wide still in ppe_counters_host[2]
  // [0] off-diagonal nnz, [1] wide nnz, [2..5] the cell box of the rows with
  // off-diagonals (gnx - 1 - min x, max x, gnz - 1 - min z, max z: atomic maxima)
  unsigned long long* ppe_counters = nullptr;
  unsigned long long* ppe_counters_host = nullptr;  // pinned: nnz, slots, wide nnz, cell box [3..6]
  uint32_t amg_box[4] = {0, 0, 0, 0};  // cells x0, z0, nx, nz of that box (the multigrid's level 0)
  bool amg_box_ready = false;          // the box belongs to the installed system
  uint4* box_part = nullptr;           // per-block partial boxes
  double nnz_offdiag = 0, sell_slots = 0;

  // CG work vectors
  double *r = nullptr, *pa = nullptr, *q = nullptr;  // residual, direction, A direction
  double *partA = nullptr, *partB = nullptr, *partC = nullptr;
  double* red_out = nullptr;
  double* red_host = nullptr;  // pinned, kRedOut entries (as red_out)
  CgStatus* st = nullptr;
  CgStatus* st_host = nullptr;
  CgStatus* st32 = nullptr;
  CgStatus* st32_host = nullptr;
  // iterations of the mixed solver's FP32 sweeps in the last solve with each
  // preconditioner (mps_precond up to MPS_PRECOND_AMG; pred_kind: the
  // running solve's), by sweep
  static constexpr int kSweepPred = 8;
  static constexpr int kPredKinds = 4;
  long long sweep_pred[kPredKinds][kSweepPred] = {};
  int pred_kind = 0;
  int sweep_index = 0;
  long long& pred(int sweep) { return sweep_pred[pred_kind][sweep < kSweepPred - 1 ? sweep : kSweepPred - 1]; }
  double final_rr = -1, final_bb = 0;  // FP64 |b - A x|^2, |b|^2 of the solve's x (-1: not known)
  // the learned warm start's residual norms are on their way to red_host
  // (learned_resolve reads them at the solver's first synchronisation)
  bool learned_pending = false;
  // the feature rows and predictions live in a batch's buffers (mps_batch):
  // not owned, not reallocated here
  bool feats_borrowed = false;
  // work enqueued ahead on the auxiliary stream: the multilevel set-up, which
  // runs beside the learned warm start (fork after the assembly, join before
  // the solve)
  hipStream_t stream_aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool ml_setup_pending = false;
  // mixed-precision work vectors
  float *pa32 = nullptr, *d32 = nullptr;  // inner CG direction and correction
  float *r32 = nullptr, *z32 = nullptr, *q32 = nullptr;  // multilevel-preconditioned inner CG
  float* agg_inv = nullptr;  // 1 / (1^T A_kk 1) of the cells and 2x2, 4x4 cell blocks (0: no correction)
  double* part_ml = nullptr;  // partial sums of the per-cell-block kernel (2 x part_ml_cap)
  size_t part_ml_cap = 0;
  int replay_ready = 0;  // SpMV the last solve ran (MPS_K_CG_SPMV32 / MPS_K_CG_SPMV), 0: none
  CgStatus* st_replay = nullptr;
  bool ml_rowsums_ready = false;  // the last fill left the aggregate row sums in z32, r32, d32
  bool amg_rows_ready = false;    // the last fill left the multigrid's row direction sums in amg_rows
  // MPS_PRECOND_AUTO: the preconditioner chosen for the assembled system, the
  // latest solve time with each (multilevel, multigrid), the current choice,
  // and the solve count at which the other one is timed again
  int precond_now = MPS_PRECOND_MULTILEVEL;
  double auto_ms[2] = {-1.0, -1.0};
  long long auto_at[2] = {0, 0};  // solve count when auto_ms[k] was recorded
  int auto_best = MPS_PRECOND_MULTILEVEL;
  long long auto_solves = 0, auto_next = 0, auto_interval = 16;
  uint32_t* ml_flag = nullptr;  // 4x4 cell blocks holding particles: flags, then their scan
  uint32_t* ml_off = nullptr;
  uint32_t* ml_list = nullptr;  // ... and their indices, ascending
  size_t ml_blk_cap = 0;
  uint4* zrec = nullptr;  // multilevel SpMV records {qx, qz, z, 0} per row
  size_t zrec_cap = 0;
  uint32_t* ml_pid = nullptr;  // packed per-slot particle ids / Jacobi factors, per-lane scales
  float* ml_dv = nullptr;
  float4* ml_sc = nullptr;
  size_t ml_pack_cap = 0;      // lanes (64 per listed block)
  uint64_t* ml_active = nullptr;       // count of such blocks
  uint64_t* ml_active_host = nullptr;  // pinned
  size_t agg_cap = 0;
  size_t agg_off[4] = {0, 0, 0, 0};  // level l at agg_inv + agg_off[l]
  float* dinv32 = nullptr;
  float4 *g32a = nullptr, *g32b = nullptr;  // packed (r, w, s, M) of the inner CG, by parity
  double *part3a = nullptr, *part3b = nullptr;  // 3 x kRedBlocks partial sums, by parity
  // cell-grid multigrid (MPS_PRECOND_AMG): the level shapes (level 0 = the
  // neighbour-search cells, then 2x2 blocks), the FP64 stencils of the set-up,
  // the FP32 stencils, 1 / diagonal and per-point vectors of the cycle, the
  // coarsest level's inverse, and the rows' per-direction sums (set-up)
  int amg_levels = 0;
  uint32_t amg_nx[kAmgMaxLevels] = {}, amg_nz[kAmgMaxLevels] = {}, amg_off[kAmgMaxLevels] = {};
  size_t amg_pts = 0, amg_cap = 0, amg_rows_cap = 0;
  double* amg_S64 = nullptr;
  float *amg_S = nullptr, *amg_dinv = nullptr, *amg_r = nullptr, *amg_x = nullptr, *amg_e = nullptr;
  float* amg_inv = nullptr;
  double* amg_rows = nullptr;
  // the smoothed prolongation: its rows per packed slot (FP16) and pass 1's
  // direction sums per level-0 point (9 planes)
  uint4* amg_q9 = nullptr;      // P_s's rows per packed slot (16 bytes), then their ninth coefficients
  uint16_t* amg_q9c = nullptr;  // inside amg_q9
  float* amg_tq = nullptr;
  size_t amg_q9_cap = 0, amg_tq_cap = 0, amg_tq_stride = 0;
  IrStatus* ir = nullptr;
  IrStatus* ir_host = nullptr;
  long long last_iters = 0;

  mps_solver_opts solver{};
  mps_solve_stats stats{};

  // learned warm start and snapshot collection
  Model* model = nullptr;
  Dataset* dataset = nullptr;
  double collect_frac = 0;
  double collect_sd = 0;  // output scale of the snapshot's feature pass
  uint64_t collect_seed = 0;
  long long collect_step = 0;
  bool collect_ok = false;
  __bf16* feats = nullptr;  // n x kFeatPad of the last feature pass
  void* feat_rec = nullptr;  // n 32-byte neighbour records of the feature pass
  double* learn_part = nullptr;  // 5 partial sums per block of the feature pass's residual
  size_t learn_part_cap = 0;
  float* ypred = nullptr;
  size_t feats_cap = 0;

  // wall kinematics (Computer.hpp:993, :1013-1019)
  mps_wall_fn wall_fn = nullptr;
  mps_wall_pre_fn wall_pre = nullptr;
  void* wall_user = nullptr;
  uint8_t* wall_group = nullptr;        // device wall kinematics: group of each particle (alloc_n)
  mps_wall_motion* wall_motion = nullptr;  // its motions (device)
  size_t wall_motions = 0;

  // measurement
  bool profiling = false;
  KStat kstat[MPS_K_COUNT];
  std::vector<EventPair> events;
  size_t events_used = 0;
  hipEvent_t solve_a = nullptr, solve_b = nullptr;
  bool solve_pending = false;       // solve_b recorded, stats.solve_ms not read yet (settle_solve)
  bool solve_pending_auto = false;  // ... and the automatic choice records that time

  std::string err;
};

// ---- several scenes stepped together (mps_batch) -------------------------------
// The pressure solves of a batch run as grouped launches: one launch per
// kernel covers every scene still iterating, each block working in its own
// scene's grid (mps_cg.hip: batch_solve).  Buffers for the argument tables,
// block offsets, per-scene masks and statuses, grown to the scene count.
struct BatchWork {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t cap = 0;                       // scenes the buffers hold
  static constexpr int kKinds = 56;     // regions: argument tables, block offsets of each grouped launch
  static constexpr size_t kArgBytes = 512;  // bytes reserved per scene and table
  unsigned char* tab = nullptr;         // device: kKinds x cap x kArgBytes
  unsigned char* tab_host = nullptr;    // pinned staging
  uint32_t* first = nullptr;            // device: kKinds x (cap + 1) block offsets
  uint32_t* first_host = nullptr;       // pinned
  CgStatus* st = nullptr;               // device: inner-solve status per scene
  CgStatus* st_host = nullptr;          // pinned
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  double solve_ms = 0;                  // last batched solve, HIP events
  long long launches = 0;               // grouped launches of the last solve
  void ensure(size_t scenes);
  void release();
};
// The multilevel mixed solver's per-scene set-up ahead of a batched solve, on
// the scene's stream (synchronises it once for the block count).
void ml_prepare(Context& c);
// The cell-grid multigrid's per-scene set-up ahead of a batched solve (levels,
// operators, packed slots, the smoothed prolongation's rows), on the scene's
// stream (synchronises it once for the block count).
void amg_prepare(Context& c);
// The scene's solve can join a batched solve (mixed precision, the multilevel
// or the cell-grid multigrid preconditioner, a system assembled from the
// current particles and cells).
bool batch_eligible(const Context& c);
// ... with the multigrid (amg_prepare; otherwise ml_prepare)
bool batch_amg(const Context& c);
// Solves the assembled systems of the eligible scenes together on w.stream
// (which must already wait for each scene's assembly and ml_prepare); each
// scene's statistics, solution (c.sol) and final residual are set as its own
// solve would.  Scenes whose FP32 sweeps stall finish with the FP64 CG alone.
void batch_solve(std::vector<Context*>& scenes, BatchWork& w);

// mps_mlp.hip.  The learned warm start of several scenes whose feature rows
// (learned_features, run on each scene beforehand) are consecutive in
// `feats`: one inference over all rows into `ypred`, then per scene
// x0 = p + s_d y, its FP64 residual and the keep-or-revert decision, all on
// stream s; the residual norms go to each scene's red_host (read them with
// batch_learned_outcome once s has been synchronised).
void batch_learned_start(std::vector<Context*>& scenes, Model& m, hipStream_t s, const __bf16* feats, float* ypred,
                         size_t rows);
void batch_learned_outcome(Context& c);

// ---- error plumbing ----------------------------------------------------------
struct Failure {
  int code;
  std::string msg;
};

#define MPS_HIP(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::mps::Failure{MPS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

template <class T>
T* dalloc(size_t count) {
  void* p = nullptr;
  if (count == 0) count = 1;
  MPS_HIP(hipMalloc(&p, count * sizeof(T)));
  return static_cast<T*>(p);
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// ---- measured launches ---------------------------------------------------------
// Counts the launch and its algorithmic bytes; when profiling is on, also
// records a HIP event pair around it on the context stream (the statistics
// are then folded in by collect_events, which can drop solver launches that
// ran after convergence: those exit at once and do no work).
template <class F>
void tracked(Context& c, int kid, double bytes, F&& launch, long long tag = 0) {
  if (!c.profiling) {
    KStat& s = c.kstat[kid];
    s.launches++;
    s.bytes += bytes;
    launch();
    return;
  }
  if (c.events_used == c.events.size()) {
    EventPair e;
    MPS_HIP(hipEventCreate(&e.a));
    MPS_HIP(hipEventCreate(&e.b));
    c.events.push_back(e);
  }
  EventPair& e = c.events[c.events_used++];
  e.kid = kid;
  e.tag = tag;
  e.bytes = bytes;
  MPS_HIP(hipEventRecord(e.a, c.stream));
  launch();
  MPS_HIP(hipEventRecord(e.b, c.stream));
}
// Adds bytes to the last launch of a kernel once they are known on the host
// (sizes counted by the kernel itself).
inline void add_bytes(Context& c, int kid, double bytes) {
  if (c.profiling)
    for (size_t k = c.events_used; k-- > 0;)
      if (c.events[k].kid == kid) {
        c.events[k].bytes += bytes;
        return;
      }
  c.kstat[kid].bytes += bytes;
}
// Folds the recorded event pairs into the kernel statistics (stream must be
// idle); pairs tagged above max_tag are discarded.
void collect_events(Context& c, long long max_tag = 0x7fffffffffffffffLL);
// Removes `count` launches of `bytes` each from a kernel's statistics (solver
// launches found to have run after convergence, when not profiling).
void uncount(Context& c, int kid, long long count, double bytes);
void check_launch();

inline unsigned grid_for(size_t n, int block = kBlock) {
  return static_cast<unsigned>((n + block - 1) / block);
}

// ---- storage order <-> caller order (mps_api.cpp) ----------------------------------
// The caller's index of every stored particle (host copy of orig, refreshed
// after a permutation; the stream is synchronised when it is refreshed).
const std::vector<uint32_t>& storage_order(Context& c);
// host[orig[k]] = dev[k] for the c.n particles (synchronous)
void download_caller_order(Context& c, const void* dev, void* host, size_t elem);
// dev[k] = host[orig[k]] for the c.n particles (synchronous)
void upload_caller_order(Context& c, const void* host, void* dev, size_t elem);

// ---- stage entry points (defined in the .hip units) ------------------------------
void search(Context& c);                 // mps_search.hip
void density(Context& c, bool ecs, bool savex, bool dens);  // mps_particles.hip
void density_variation(Context& c, double* out_dev);
void explicit_forces(Context& c);
// density(c, true, false, true) followed by explicit_forces(c), the two list
// walks fused into one kernel (same results)
void density_and_explicit(Context& c);
void explicit_update(Context& c);  // wall targets and the integration of the explicit stage
void wall_targets_from_callback(Context& c);
void pressure_writeback(Context& c);
void pressure_gradient(Context& c);
void dynamic_stabilize(Context& c);
void ppe_assemble(Context& c);           // mps_ppe.hip
void pack_state(Context& c);             // mps_particles.hip: xz, uw from x, z, u, w
// density(c, false, true, true) followed by ppe_assemble(c), the density walk
// and the PPE count walk fused into one kernel (same results)
void density_save_and_assemble(Context& c);
void ppe_make_fp32(Context& c);
double mf_nnz_wide(Context& c);  // mps_ppe.hip: nnz of the wide slices; valid once the stream was synchronised
// mps_ppe.hip: the box of cells of the rows with off-diagonals (c.amg_box),
// computed now (one synchronisation) unless the assembly left it
void amg_box(Context& c);
// c.amg_rows holds 9 doubles for each of n rows (mps_cg.hip)
void ensure_amg_rows(Context& c, size_t n);
// The mixed solver's preconditioner for a system about to be assembled from
// the particles (MPS_PRECOND_AUTO picks it; mps_cg.hip), and the one in force.
int precond_for_system(Context& c);
int effective_precond(const Context& c);
double replay_spmv(Context& c, int kid, int reps);  // mps_cg.hip: mean ms of one launch
// exclusive scan of n u32 counts into out, grand total to *total (device); mps_search.hip
void scan_counts(Context& c, const uint32_t* in, size_t n, uint32_t* out, uint64_t* total);
void ppe_upload_csr(Context& c, size_t n, const uint64_t* rowptr, const uint32_t* col,
                    const double* val, const double* b, const double* x0);
void ppe_download_csr(Context& c, std::vector<uint64_t>& rowptr, std::vector<uint32_t>& col,
                      std::vector<double>& val);
void ensure_sell_capacity(Context& c, size_t nslices, size_t slots);
void ensure_ppe_vectors(Context& c, size_t n);  // mps_api.cpp
void ppe_solve(Context& c);              // mps_cg.hip
void settle_solve(Context& c);           // mps_cg.hip: the last solve's time into stats (first use)
void residual64(Context& c, double& rr, double& bb, bool track);
void reduce2(Context& c, const double* a, const double* b, int nparts);  // -> c.red_host[0..1]
void residual64_launch(Context& c);  // r = b - A x; |r|^2, |b|^2 -> c.red_out[0..1] (no sync)
struct FeatRec;
bool learned_residual32(Context& c, FeatRec* rec);  // mps_cg.hip: see there
void ml_setup_ahead(Context& c);  // mps_cg.hip: the multilevel set-up on the auxiliary stream
void join_aux(Context& c);        // the main stream waits for it (if pending)
// mps_mlp.hip.  learned_features enqueues (no synchronisation) the residual
// of x = p_prev, the scales and the feature rows; red_out then holds |r|^2,
// |b|^2 [0, 1], the alive sums of p^2, diag^2 and count [2..4], and
// [6..9] = |r_prev|^2, 1/s_p, 1/s_r and the output scale s_d.
void learned_features(Context& c);
// after learned_warm_start on a mixed solve: synchronises, records the
// statistics and returns whether r, rr, bb hold the residual of the solver's x
bool learned_resolve(Context& c, double& rr, double& bb);
void learned_warm_start(Context& c);
void collect_before(Context& c);
void collect_after(Context& c);
void model_forward_host(Model& m, size_t rows, const float* x, float* y);
void model_grads_host(Model& m, size_t rows, const float* x, const float* t, float* loss_out);
double max_speed_dt(Context& c);         // mps_particles.hip (DetermineDt)
void device_math_check(size_t n, const double* a, const double* b, double* s, double* d);
double device_copy_probe(size_t bytes);

}  // namespace mps

// The opaque handles of the C ABI.
struct mps_ctx {
  mps::Context c;
};
struct mps_model {
  int device = 0;
  mps::Model m;
};
struct mps_dataset {
  int device = 0;
  mps::Dataset d;
};
