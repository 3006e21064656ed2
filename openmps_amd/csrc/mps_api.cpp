// C ABI of the MI355X MPS library (declarations and reference citations in
// include/mps.h).  Owns the context: device buffers, the HIP stream, the
// ForwardTime orchestration and the error-code mapping.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <exception>
#include <new>

#include "mps_internal.hpp"


namespace mps {

void check_launch() {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Failure{MPS_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e)};
}

void collect_events(Context& c, long long max_tag) {
  for (size_t k = 0; k < c.events_used; k++) {
    const EventPair& e = c.events[k];
    if (e.tag > max_tag) continue;
    float ms = 0;
    MPS_HIP(hipEventElapsedTime(&ms, e.a, e.b));
    KStat& s = c.kstat[e.kid];
    s.ms += ms;
    s.launches++;
    s.bytes += e.bytes;
  }
  c.events_used = 0;
}

void uncount(Context& c, int kid, long long count, double bytes) {
  if (c.profiling || count <= 0) return;
  c.kstat[kid].launches -= count;
  c.kstat[kid].bytes -= count * bytes;
}

namespace {

// Particle::W, Particle.hpp:71-75
double weight_h(double r, double re) { return ((0 < r) && (r < re)) ? (re / r - 1) : 0; }

// Environment.hpp:164-209: n0 = sum over the lattice points (i, j) l0,
// i, j in [-R, R), R = ceil(r_e / l0), (0, 0) excluded, in that loop order.
double lattice_sum(double r_by_l0, double l0, double r) {
  const int range = static_cast<int>(std::ceil(r_by_l0));
  double s = 0;
  for (int i = -range; i < range; i++)
    for (int j = -range; j < range; j++) {
      if (i == 0 && j == 0) continue;
      const double px = i * l0, pz = j * l0;
      const double d = std::sqrt(std::fabs(px) * std::fabs(px) + std::fabs(pz) * std::fabs(pz));
      if (d < r) s += weight_h(d, r);
    }
  return s;
}

void derive(mps_env& e) {
  e.max_dt = std::min(e.max_dt_arg, std::sqrt(2 * (e.courant * e.l0) / e.g));  // :138
  e.max_dx = e.courant * e.l0;                                                 // :143
  e.r_e = e.r_e_by_l0 * e.l0;                                                  // :145
  e.gx = 0;
  e.gz = -e.g;                                                                 // :149
  const double rmax = std::max(e.r_e_by_l0, e.r_lap_by_l0);
  e.neighbor_length = rmax * e.l0 * (1 + e.courant * 2);                       // :161
  e.n0 = lattice_sum(e.r_e_by_l0, e.l0, e.r_e);                                // :164-209
  if (e.r_lap_by_l0 == e.r_e_by_l0) {
    e.r_lap = e.r_e;
    e.n0_lap = e.n0;
  } else {
    e.r_lap = e.r_lap_by_l0 * e.l0;
    e.n0_lap = lattice_sum(e.r_lap_by_l0, e.l0, e.r_lap);
  }
}

template <class T>
void grow(T*& p, size_t old_n, size_t new_n, hipStream_t s) {
  T* q = dalloc<T>(new_n);
  MPS_HIP(hipMemsetAsync(q, 0, new_n * sizeof(T), s));
  if (p && old_n) MPS_HIP(hipMemcpyAsync(q, p, old_n * sizeof(T), hipMemcpyDeviceToDevice, s));
  MPS_HIP(hipStreamSynchronize(s));
  dfree(p);
  p = q;
}

template <class F>
int guard(mps_ctx* ctx, F&& f) {
  if (!ctx) return MPS_ERR_BAD_ARG;
  try {
    f(ctx->c);
    ctx->c.err.clear();
    return MPS_OK;
  } catch (Failure& e) {
    ctx->c.err = e.msg;
    if (e.code == MPS_ERR_HIP) (void)hipGetLastError();
    return e.code;
  } catch (std::bad_alloc&) {
    ctx->c.err = "host allocation failed";
    return MPS_ERR_BAD_ARG;
  } catch (std::exception& e) {
    ctx->c.err = e.what();
    return MPS_ERR_BAD_ARG;
  }
}

void sync(Context& c) {
  MPS_HIP(hipStreamSynchronize(c.stream));
  collect_events(c);
}

void require_neighbors(const Context& c) {
  if (!c.have_neighbors && c.n > 0) throw Failure{MPS_ERR_STATE, "neighbour search has not run"};
}

void free_all(Context& c) {
  dfree(c.x); dfree(c.z); dfree(c.u); dfree(c.w); dfree(c.p); dfree(c.nd); dfree(c.type);
  dfree(c.nws); dfree(c.ecs); dfree(c.ox); dfree(c.oz); dfree(c.dux); dfree(c.duz);
  dfree(c.wall_x); dfree(c.wall_z); dfree(c.hs_out); dfree(c.wall_group); dfree(c.wall_motion);
  dfree(c.xz); dfree(c.uw); dfree(c.orig);
  for (auto& d : c.spare_d) dfree(d);
  dfree(c.spare_type); dfree(c.spare_group); dfree(c.spare_orig); dfree(c.spare_key);
  dfree(c.key); dfree(c.sorted); dfree(c.pos_sorted); dfree(c.in_cells); dfree(c.nbr); dfree(c.ncount);
  dfree(c.cell_count); dfree(c.cell_start); dfree(c.cell_fill); dfree(c.scan_tmp); dfree(c.flags);
  dfree(c.nbr_total_dev); dfree(c.ppe_counters);
  dfree(c.row_len); dfree(c.slice_width); dfree(c.slice_off); dfree(c.sval); dfree(c.sval32);
  dfree(c.scol); dfree(c.diag); dfree(c.diag32); dfree(c.b); dfree(c.sol);
  dfree(c.r); dfree(c.pa); dfree(c.q); dfree(c.partA); dfree(c.partB); dfree(c.partC);
  dfree(c.red_out); dfree(c.st); dfree(c.st32); dfree(c.ir);
  dfree(c.pa32); dfree(c.d32); dfree(c.dinv32);
  if (!c.feats_borrowed) {
    dfree(c.feats);
    dfree(c.ypred);
  }
  dfree(c.feat_rec); dfree(c.learn_part);
  dfree(c.ghost32); dfree(c.qpos); dfree(c.slice_wide); dfree(c.col16);
  dfree(c.g32a); dfree(c.g32b); dfree(c.part3a); dfree(c.part3b);
  dfree(c.r32); dfree(c.z32); dfree(c.q32); dfree(c.agg_inv); dfree(c.part_ml);
  dfree(c.ml_flag); dfree(c.ml_off); dfree(c.ml_list); dfree(c.ml_active); dfree(c.st_replay);
  dfree(c.ml_pid); dfree(c.ml_dv); dfree(c.ml_sc); dfree(c.zrec);
  dfree(c.amg_S64); dfree(c.amg_S); dfree(c.amg_dinv); dfree(c.amg_r); dfree(c.amg_x); dfree(c.amg_e);
  dfree(c.amg_inv); dfree(c.amg_rows); dfree(c.box_part); dfree(c.amg_q9); dfree(c.amg_tq);
  if (c.flags_host) (void)hipHostFree(c.flags_host);
  if (c.nbr_total_host) (void)hipHostFree(c.nbr_total_host);
  if (c.ppe_counters_host) (void)hipHostFree(c.ppe_counters_host);
  if (c.red_host) (void)hipHostFree(c.red_host);
  if (c.ml_active_host) (void)hipHostFree(c.ml_active_host);
  if (c.st_host) (void)hipHostFree(c.st_host);
  if (c.st32_host) (void)hipHostFree(c.st32_host);
  if (c.ir_host) (void)hipHostFree(c.ir_host);
  for (auto& e : c.events) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  c.events.clear();
  if (c.solve_a) (void)hipEventDestroy(c.solve_a);
  if (c.solve_b) (void)hipEventDestroy(c.solve_b);
  if (c.ev_fork) (void)hipEventDestroy(c.ev_fork);
  if (c.ev_join) (void)hipEventDestroy(c.ev_join);
  if (c.stream_aux) (void)hipStreamDestroy(c.stream_aux);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
}

template <class T>
void pinned(T*& p, size_t count) {
  void* q = nullptr;
  MPS_HIP(hipHostMalloc(&q, count * sizeof(T), hipHostMallocDefault));
  std::memset(q, 0, count * sizeof(T));
  p = static_cast<T*>(q);
}

double* aux_ptr(Context& c, int which, size_t& count) {
  count = c.n;
  switch (which) {
    case MPS_AUX_N_WITHOUT_SPP: return c.nws;
    case MPS_AUX_ECS: return c.ecs;
    case MPS_AUX_SAVED_X: return c.ox;
    case MPS_AUX_SAVED_Z: return c.oz;
    case MPS_AUX_DU_X: return c.dux;
    case MPS_AUX_DU_Z: return c.duz;
    case MPS_AUX_PPE_B: count = c.ppe_n; return c.b;
    case MPS_AUX_PPE_X: count = c.ppe_n; return c.sol;
    default: throw Failure{MPS_ERR_BAD_ARG, "unknown auxiliary array"};
  }
}

// CSR rows / columns of a particles' system between caller order and storage
// order: row k of `out` is row map_row[k] of `in`, columns mapped through
// map_col, ascending in each row (the diagonal is one of the row's entries).
void permute_csr(size_t n, const uint64_t* rp, const uint32_t* col, const double* val,
                 const std::vector<uint32_t>& map_row, const std::vector<uint32_t>& map_col,
                 std::vector<uint64_t>& orp, std::vector<uint32_t>& ocol, std::vector<double>& oval) {
  orp.assign(n + 1, 0);
  ocol.clear();
  oval.clear();
  std::vector<std::pair<uint32_t, double>> row;
  for (size_t k = 0; k < n; k++) {
    const uint32_t i = map_row[k];
    row.clear();
    for (uint64_t e = rp[i]; e < rp[i + 1]; e++) row.emplace_back(map_col[col[e]], val[e]);
    std::sort(row.begin(), row.end(),
              [](const std::pair<uint32_t, double>& a, const std::pair<uint32_t, double>& b) {
                return a.first < b.first;
              });
    for (auto& e : row) {
      ocol.push_back(e.first);
      oval.push_back(e.second);
    }
    orp[k + 1] = ocol.size();
  }
}

std::vector<uint32_t> inverse(const std::vector<uint32_t>& o) {
  std::vector<uint32_t> inv(o.size());
  for (size_t k = 0; k < o.size(); k++) inv[o[k]] = static_cast<uint32_t>(k);
  return inv;
}

// Rows of the installed system are the particles (in storage order).
bool ppe_over_particles(const Context& c) { return c.have_ppe && c.n > 0 && c.ppe_n == c.n; }

void forward(Context& c, double dt) {
  // Computer::ForwardTime(dt), Computer.hpp:1700-1742
  c.dt = dt;
  c.t += dt;
  search(c);
  // density + HS/ECS and the explicit accelerations in one list walk, then the update
  density_and_explicit(c);
  // second density pass + SaveX + PPE assembly (its count pass fused with the density pass)
  density_save_and_assemble(c);
  ppe_solve(c);
  pressure_writeback(c);
  pressure_gradient(c);
  dynamic_stabilize(c);
  sync(c);
}

const char* kKernelNames[MPS_K_COUNT] = {
    "cell_keys", "cell_scan", "cell_scatter", "cell_sort", "neighbors", "density_ecs", "explicit",
    "update", "density_save", "ppe_count", "ppe_fill", "cg_init", "cg_spmv", "cg_update",
    "cg_iter32", "cg_init32", "ir_residual", "gradient", "ds", "other", "features", "mlp",
    "ml_setup", "ml_prec", "cg_spmv32", "cg_update32", "amg_cycle"};

}  // namespace

const std::vector<uint32_t>& storage_order(Context& c) {
  if (c.order_host_serial != c.order_serial || c.orig_host.size() != c.n) {
    c.orig_host.resize(c.n);
    if (c.n) {
      MPS_HIP(hipMemcpyAsync(c.orig_host.data(), c.orig, c.n * 4, hipMemcpyDeviceToHost, c.stream));
      MPS_HIP(hipStreamSynchronize(c.stream));
    }
    c.order_host_serial = c.order_serial;
  }
  return c.orig_host;
}

void download_caller_order(Context& c, const void* dev, void* host, size_t elem) {
  const size_t n = c.n;
  if (n == 0) return;
  if (c.order_serial == 0) {
    MPS_HIP(hipMemcpyAsync(host, dev, n * elem, hipMemcpyDeviceToHost, c.stream));
    MPS_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  const std::vector<uint32_t>& o = storage_order(c);
  std::vector<unsigned char> tmp(n * elem);
  MPS_HIP(hipMemcpyAsync(tmp.data(), dev, n * elem, hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
  unsigned char* h = static_cast<unsigned char*>(host);
  for (size_t k = 0; k < n; k++) std::memcpy(h + o[k] * elem, tmp.data() + k * elem, elem);
}

void upload_caller_order(Context& c, const void* host, void* dev, size_t elem) {
  const size_t n = c.n;
  if (n == 0) return;
  if (c.order_serial == 0) {
    MPS_HIP(hipMemcpyAsync(dev, host, n * elem, hipMemcpyHostToDevice, c.stream));
    MPS_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  const std::vector<uint32_t>& o = storage_order(c);
  std::vector<unsigned char> tmp(n * elem);
  const unsigned char* h = static_cast<const unsigned char*>(host);
  for (size_t k = 0; k < n; k++) std::memcpy(tmp.data() + k * elem, h + o[k] * elem, elem);
  MPS_HIP(hipMemcpyAsync(dev, tmp.data(), n * elem, hipMemcpyHostToDevice, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
}

void ensure_ppe_vectors(Context& c, size_t n) {
  if (n <= c.ppe_alloc && c.r) return;
  const size_t m = (n + 63) / 64 * 64 + 64;
  const size_t ns = m / 64 + 1;
  dfree(c.row_len); dfree(c.slice_width); dfree(c.slice_off); dfree(c.diag); dfree(c.diag32);
  dfree(c.b); dfree(c.sol); dfree(c.r); dfree(c.pa); dfree(c.q);
  dfree(c.pa32); dfree(c.d32); dfree(c.dinv32);
  dfree(c.ghost32); dfree(c.qpos); dfree(c.slice_wide);
  dfree(c.g32a); dfree(c.g32b); dfree(c.r32); dfree(c.z32); dfree(c.q32);
  c.row_len = dalloc<uint32_t>(m);
  c.slice_width = dalloc<uint32_t>(ns);
  c.slice_off = dalloc<uint64_t>(ns + 1);
  c.diag = dalloc<double>(m);
  c.diag32 = dalloc<float>(m);
  c.b = dalloc<double>(m);
  c.sol = dalloc<double>(m);
  c.r = dalloc<double>(m);
  c.pa = dalloc<double>(m);
  c.q = dalloc<double>(m);
  c.pa32 = dalloc<float>(m);
  c.d32 = dalloc<float>(m);
  c.dinv32 = dalloc<float>(m);
  c.ghost32 = dalloc<float>(m);
  c.g32a = dalloc<float4>(m);
  c.g32b = dalloc<float4>(m);
  c.r32 = dalloc<float>(m);
  c.z32 = dalloc<float>(m);
  c.q32 = dalloc<float>(m);
  c.qpos = dalloc<uint2>(m);
  c.slice_wide = dalloc<uint8_t>(ns);
  MPS_HIP(hipMemsetAsync(c.sol, 0, m * 8, c.stream));
  MPS_HIP(hipMemsetAsync(c.pa, 0, m * 8, c.stream));
  MPS_HIP(hipMemsetAsync(c.pa32, 0, m * 4, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
  c.ppe_alloc = n;
  c.have_ppe = false;
  c.have_ppe32 = false;
  c.ppe_from_particles = false;
}

}  // namespace mps

using mps::Context;
using mps::Failure;

extern "C" {

int mps_env_init(mps_env* env, double max_dt, double courant, double g, double rho, double nu,
                 double r_e_by_l0, double l0, double min_x, double min_z, double max_x, double max_z) {
  if (!env || !(l0 > 0) || !(r_e_by_l0 > 0)) return MPS_ERR_BAD_ARG;
  mps_env e{};
  e.max_dt_arg = max_dt;
  e.courant = courant;
  e.g = g;
  e.rho = rho;
  e.nu = nu;
  e.r_e_by_l0 = r_e_by_l0;
  e.r_lap_by_l0 = r_e_by_l0;
  e.l0 = l0;
  e.min_x = min_x;
  e.min_z = min_z;
  e.max_x = max_x;
  e.max_z = max_z;
  mps::derive(e);
  *env = e;
  return MPS_OK;
}

int mps_env_set_laplacian_radius(mps_env* env, double r_lap_by_l0) {
  if (!env || !(r_lap_by_l0 > 0)) return MPS_ERR_BAD_ARG;
  env->r_lap_by_l0 = r_lap_by_l0;
  mps::derive(*env);
  return MPS_OK;
}

int mps_create(const mps_env* env, double allowable_residual, int device, mps_ctx** out) {
  if (!env || !out) return MPS_ERR_BAD_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return MPS_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= ndev) return MPS_ERR_BAD_ARG;
  mps_ctx* h = new (std::nothrow) mps_ctx;
  if (!h) return MPS_ERR_BAD_ARG;
  Context& c = h->c;
  try {
    MPS_HIP(hipSetDevice(device));
    c.device = device;
    c.env = *env;
    c.eps = allowable_residual;
    c.solver.kind = MPS_SOLVER_CG_FP64;
    c.solver.stop_rule = MPS_STOP_INITIAL_RESIDUAL;
    c.solver.warm_start = MPS_WARM_PREVIOUS;
    c.solver.inner_tol = 1e-4;
    MPS_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    // Grid.hpp:137-150
    const double NL = c.env.neighbor_length;
    c.gnx = static_cast<long long>(std::ceil((c.env.max_x - c.env.min_x) / NL)) + 2;
    c.gnz = static_cast<long long>(std::ceil((c.env.max_z - c.env.min_z) / NL)) + 2;
    const long long c1 = static_cast<long long>(std::ceil(NL / c.env.l0)) + 1;
    if (c.gnx <= 0 || c.gnz <= 0 || c1 <= 0 || static_cast<double>(c.gnx) * c.gnz > 4.0e9)
      throw Failure{MPS_ERR_BAD_ARG, "grid shape out of range"};
    c.cap = static_cast<uint32_t>(c1 * c1);
    c.kmax = c.cap * 9;
    c.kstride = (c.kmax + 3u) & ~3u;
    c.ncells = static_cast<size_t>(c.gnx * c.gnz);
    c.cell_count = mps::dalloc<uint32_t>(c.ncells);
    c.cell_start = mps::dalloc<uint32_t>(c.ncells);
    c.cell_fill = mps::dalloc<uint32_t>(c.ncells);
    c.scan_tmp = mps::dalloc<uint64_t>(c.ncells / 2048 + 4096);
    c.flags = mps::dalloc<int>(4);
    c.nbr_total_dev = mps::dalloc<unsigned long long>(1);
    c.in_cells = mps::dalloc<uint64_t>(1);
    c.ppe_counters = mps::dalloc<unsigned long long>(8);
    c.partA = mps::dalloc<double>(mps::kRedBlocks);
    c.partB = mps::dalloc<double>(mps::kRedBlocks);
    c.partC = mps::dalloc<double>(mps::kRedBlocks);
    c.part3a = mps::dalloc<double>(3 * mps::kRedBlocks);
    c.part3b = mps::dalloc<double>(3 * mps::kRedBlocks);
    c.red_out = mps::dalloc<double>(mps::kRedOut);
    c.st = mps::dalloc<mps::CgStatus>(1);
    c.st32 = mps::dalloc<mps::CgStatus>(1);
    c.ir = mps::dalloc<mps::IrStatus>(1);
    mps::pinned(c.flags_host, 4);
    mps::pinned(c.nbr_total_host, 1);
    mps::pinned(c.ppe_counters_host, 8);
    mps::pinned(c.red_host, mps::kRedOut);
    mps::pinned(c.ml_active_host, 1);
    mps::pinned(c.st_host, 1);
    mps::pinned(c.st32_host, 1);
    mps::pinned(c.ir_host, 1);
    MPS_HIP(hipEventCreate(&c.solve_a));
    MPS_HIP(hipEventCreate(&c.solve_b));
  } catch (Failure& f) {
    std::fprintf(stderr, "mps_create: %s\n", f.msg.c_str());
    mps::free_all(c);
    delete h;
    return f.code;
  }
  *out = h;
  return MPS_OK;
}

void mps_destroy(mps_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  mps::free_all(ctx->c);
  delete ctx;
}

const char* mps_last_error(const mps_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int mps_get_env(const mps_ctx* ctx, mps_env* out) {
  if (!ctx || !out) return MPS_ERR_BAD_ARG;
  *out = ctx->c.env;
  return MPS_OK;
}

int mps_grid_shape(const mps_ctx* ctx, int64_t* out4) {
  if (!ctx || !out4) return MPS_ERR_BAD_ARG;
  out4[0] = ctx->c.gnx;
  out4[1] = ctx->c.gnz;
  out4[2] = ctx->c.cap;
  out4[3] = ctx->c.kmax;
  return MPS_OK;
}

int mps_add_particles(mps_ctx* ctx, size_t count, const double* x, const double* z, const double* u,
                      const double* w, const double* p, const double* n, const uint8_t* type) {
  return mps::guard(ctx, [&](Context& c) {
    if (count == 0) return;
    if (!x || !z || !u || !w || !p || !n || !type) throw Failure{MPS_ERR_BAD_ARG, "null particle array"};
    for (size_t k = 0; k < count; k++)
      if (type[k] > MPS_DISABLED) throw Failure{MPS_ERR_BAD_ARG, "particle type out of range"};
    const size_t old = c.n, nn = c.n + count;
    if (nn > 0xffffffffull) throw Failure{MPS_ERR_BAD_ARG, "too many particles"};
    MPS_HIP(hipSetDevice(c.device));
    if (nn > c.alloc_n) {
      const size_t m = (nn + 63) / 64 * 64;  // whole slices
      const size_t mo = c.alloc_n;
      mps::grow(c.x, old, m, c.stream);
      mps::grow(c.z, old, m, c.stream);
      mps::grow(c.u, old, m, c.stream);
      mps::grow(c.w, old, m, c.stream);
      mps::grow(c.p, old, m, c.stream);
      mps::grow(c.nd, old, m, c.stream);
      mps::grow(c.type, old, m, c.stream);
      mps::grow(c.nws, old, m, c.stream);
      mps::grow(c.ecs, old, m, c.stream);
      mps::grow(c.ox, old, m, c.stream);
      mps::grow(c.oz, old, m, c.stream);
      mps::grow(c.dux, old, m, c.stream);
      mps::grow(c.duz, old, m, c.stream);
      mps::grow(c.wall_x, old, m, c.stream);
      mps::grow(c.wall_z, old, m, c.stream);
      if (c.wall_group) mps::grow(c.wall_group, old, m, c.stream);
      mps::grow(c.orig, old, m, c.stream);
      for (auto& d : c.spare_d) mps::grow(d, 0, m, c.stream);
      mps::grow(c.spare_type, 0, m, c.stream);
      if (c.wall_group) mps::grow(c.spare_group, 0, m, c.stream);
      mps::grow(c.spare_orig, 0, m, c.stream);
      mps::grow(c.spare_key, 0, m, c.stream);
      mps::grow(c.hs_out, 0, m, c.stream);
      mps::grow(c.key, 0, m, c.stream);
      mps::grow(c.sorted, 0, m, c.stream);
      mps::grow(c.pos_sorted, 0, m, c.stream);
      mps::grow(c.xz, 0, m, c.stream);
      mps::grow(c.uw, 0, m, c.stream);
      mps::grow(c.ncount, old, m, c.stream);
      // neighbour rows of existing particles are kept (boost::multi_array::resize)
      mps::grow(c.nbr, (mo / 64) * 64 * static_cast<size_t>(c.kstride), m * static_cast<size_t>(c.kstride),
                c.stream);
      c.alloc_n = m;
    }
    mps::ensure_ppe_vectors(c, nn);
    const hipMemcpyKind h2d = hipMemcpyHostToDevice;
    MPS_HIP(hipMemcpyAsync(c.x + old, x, count * 8, h2d, c.stream));
    MPS_HIP(hipMemcpyAsync(c.z + old, z, count * 8, h2d, c.stream));
    MPS_HIP(hipMemcpyAsync(c.u + old, u, count * 8, h2d, c.stream));
    MPS_HIP(hipMemcpyAsync(c.w + old, w, count * 8, h2d, c.stream));
    MPS_HIP(hipMemcpyAsync(c.p + old, p, count * 8, h2d, c.stream));
    MPS_HIP(hipMemcpyAsync(c.nd + old, n, count * 8, h2d, c.stream));
    MPS_HIP(hipMemcpyAsync(c.type + old, type, count, h2d, c.stream));
    MPS_HIP(hipMemcpyAsync(c.wall_x + old, x, count * 8, h2d, c.stream));
    MPS_HIP(hipMemcpyAsync(c.wall_z + old, z, count * 8, h2d, c.stream));
    // particles added while device kinematics are on stay put (group 0)
    if (c.wall_group) MPS_HIP(hipMemsetAsync(c.wall_group + old, 0, count, c.stream));
    // appended particles are stored after the others, in the caller's order
    std::vector<uint32_t> ids(count);
    for (size_t k = 0; k < count; k++) ids[k] = static_cast<uint32_t>(old + k);
    MPS_HIP(hipMemcpyAsync(c.orig + old, ids.data(), count * 4, h2d, c.stream));
    MPS_HIP(hipStreamSynchronize(c.stream));
    if (c.order_host_serial == c.order_serial && c.orig_host.size() == old)
      c.orig_host.insert(c.orig_host.end(), ids.begin(), ids.end());
    c.n = nn;
    c.have_ppe = false;
  });
}

size_t mps_size(const mps_ctx* ctx) { return ctx ? ctx->c.n : 0; }

int mps_get_particles(mps_ctx* ctx, double* x, double* z, double* u, double* w, double* p, double* n,
                      uint8_t* type) {
  return mps::guard(ctx, [&](Context& c) {
    if (c.n == 0) return;
    if (x) mps::download_caller_order(c, c.x, x, 8);
    if (z) mps::download_caller_order(c, c.z, z, 8);
    if (u) mps::download_caller_order(c, c.u, u, 8);
    if (w) mps::download_caller_order(c, c.w, w, 8);
    if (p) mps::download_caller_order(c, c.p, p, 8);
    if (n) mps::download_caller_order(c, c.nd, n, 8);
    if (type) mps::download_caller_order(c, c.type, type, 1);
  });
}

int mps_set_particles(mps_ctx* ctx, const double* x, const double* z, const double* u, const double* w,
                      const double* p, const double* n, const uint8_t* type) {
  return mps::guard(ctx, [&](Context& c) {
    const size_t N = c.n;
    if (N == 0) return;
    if (type)
      for (size_t k = 0; k < N; k++)
        if (type[k] > MPS_DISABLED) throw Failure{MPS_ERR_BAD_ARG, "particle type out of range"};
    if (x) mps::upload_caller_order(c, x, c.x, 8);
    if (z) mps::upload_caller_order(c, z, c.z, 8);
    if (u) mps::upload_caller_order(c, u, c.u, 8);
    if (w) mps::upload_caller_order(c, w, c.w, 8);
    if (p) mps::upload_caller_order(c, p, c.p, 8);
    if (n) mps::upload_caller_order(c, n, c.nd, 8);
    if (type) mps::upload_caller_order(c, type, c.type, 1);
    c.ppe_from_particles = false;  // positions no longer those the system was assembled from
  });
}

int mps_get_time(const mps_ctx* ctx, double* t, double* dt) {
  if (!ctx) return MPS_ERR_BAD_ARG;
  if (t) *t = ctx->c.t;
  if (dt) *dt = ctx->c.dt;
  return MPS_OK;
}

int mps_set_time(mps_ctx* ctx, double t, double dt) {
  if (!ctx) return MPS_ERR_BAD_ARG;
  ctx->c.t = t;
  ctx->c.dt = dt;
  return MPS_OK;
}

int mps_set_wall_motion(mps_ctx* ctx, size_t n_motions, const mps_wall_motion* motions, const uint8_t* group) {
  if (n_motions > 255 || (n_motions > 0 && (!motions || !group))) return MPS_ERR_BAD_ARG;
  return mps::guard(ctx, [&](Context& c) {
    mps::sync(c);
    c.wall_motions = 0;
    if (n_motions == 0) return;
    for (size_t i = 0; i < c.n; i++)
      if (group[i] > n_motions) throw Failure{MPS_ERR_BAD_ARG, "wall group without a motion"};
    if (!c.wall_group) {
      c.wall_group = mps::dalloc<uint8_t>(std::max<size_t>(c.alloc_n, 1));
      MPS_HIP(hipMemset(c.wall_group, 0, std::max<size_t>(c.alloc_n, 1)));
    }
    mps::dfree(c.wall_motion);
    c.wall_motion = mps::dalloc<mps_wall_motion>(n_motions);
    MPS_HIP(hipMemcpy(c.wall_motion, motions, n_motions * sizeof(mps_wall_motion), hipMemcpyHostToDevice));
    if (!c.spare_group) c.spare_group = mps::dalloc<uint8_t>(std::max<size_t>(c.alloc_n, 1));
    mps::upload_caller_order(c, group, c.wall_group, 1);
    c.wall_motions = n_motions;
    c.wall_fn = nullptr;
    c.wall_pre = nullptr;
    c.wall_user = nullptr;
  });
}

int mps_set_wall_callbacks(mps_ctx* ctx, mps_wall_fn wall, mps_wall_pre_fn pre, void* user) {
  if (!ctx) return MPS_ERR_BAD_ARG;
  ctx->c.wall_fn = wall;
  ctx->c.wall_pre = pre;
  ctx->c.wall_user = user;
  ctx->c.wall_motions = 0;  // the callbacks replace any device kinematics
  return MPS_OK;
}

int mps_forward_time(mps_ctx* ctx, double dt) {
  return mps::guard(ctx, [&](Context& c) {
    if (!(dt > 0)) throw Failure{MPS_ERR_BAD_ARG, "dt must be positive"};
    mps::forward(c, dt);
  });
}

int mps_determine_dt(mps_ctx* ctx, double* dt_out) {
  return mps::guard(ctx, [&](Context& c) {
    const double dt = mps::max_speed_dt(c);
    if (dt_out) *dt_out = dt;
  });
}

int mps_forward_time_cfl(mps_ctx* ctx, double* dt_out) {
  return mps::guard(ctx, [&](Context& c) {
    const double dt = mps::max_speed_dt(c);  // Computer::DetermineDt
    if (dt_out) *dt_out = dt;
    mps::forward(c, dt);
  });
}

int mps_set_solver(mps_ctx* ctx, const mps_solver_opts* opts) {
  if (!ctx || !opts) return MPS_ERR_BAD_ARG;
  if (opts->kind < 0 || opts->kind > MPS_SOLVER_MIXED_IR) return MPS_ERR_BAD_ARG;
  if (opts->stop_rule < 0 || opts->stop_rule > MPS_STOP_RHS_NORM) return MPS_ERR_BAD_ARG;
  if (opts->warm_start < 0 || opts->warm_start > MPS_WARM_LEARNED) return MPS_ERR_BAD_ARG;
  if (opts->precond < 0 || opts->precond > MPS_PRECOND_AUTO) return MPS_ERR_BAD_ARG;
  mps::Context& c = ctx->c;
  if (opts->precond == MPS_PRECOND_AUTO && c.solver.precond != MPS_PRECOND_AUTO) {
    // the automatic choice starts over: both preconditioners are timed again
    c.auto_ms[0] = c.auto_ms[1] = -1.0;
    c.auto_at[0] = c.auto_at[1] = 0;
    c.auto_best = MPS_PRECOND_MULTILEVEL;
    c.auto_solves = c.auto_next = 0;
    c.auto_interval = 16;
  }
  c.solver = *opts;
  return MPS_OK;
}

int mps_get_solver(const mps_ctx* ctx, mps_solver_opts* out) {
  if (!ctx || !out) return MPS_ERR_BAD_ARG;
  *out = ctx->c.solver;
  return MPS_OK;
}

int mps_get_solve_stats(const mps_ctx* ctx, mps_solve_stats* out) {
  if (!ctx || !out) return MPS_ERR_BAD_ARG;
  // the solve time is read from the solve's events on first use (the
  // statistics read changes no observable state; a last error stays set)
  Context& c = const_cast<mps_ctx*>(ctx)->c;
  try {
    mps::settle_solve(c);
  } catch (Failure& e) {
    c.err = e.msg;
    if (e.code == MPS_ERR_HIP) (void)hipGetLastError();
    return e.code;
  }
  *out = c.stats;
  return MPS_OK;
}

int mps_stage_search(mps_ctx* ctx) {
  return mps::guard(ctx, [&](Context& c) {
    mps::search(c);
    mps::sync(c);
  });
}
int mps_stage_density(mps_ctx* ctx) {
  return mps::guard(ctx, [&](Context& c) {
    mps::require_neighbors(c);
    mps::density(c, false, false, true);
    mps::sync(c);
  });
}
int mps_stage_error_correction(mps_ctx* ctx) {
  return mps::guard(ctx, [&](Context& c) {
    mps::require_neighbors(c);
    mps::density(c, true, false, false);
    mps::sync(c);
  });
}
int mps_stage_explicit(mps_ctx* ctx) {
  return mps::guard(ctx, [&](Context& c) {
    mps::require_neighbors(c);
    mps::explicit_forces(c);
    mps::sync(c);
  });
}
int mps_stage_save_x(mps_ctx* ctx) {
  return mps::guard(ctx, [&](Context& c) {
    mps::require_neighbors(c);
    mps::density(c, false, true, false);
    mps::sync(c);
  });
}
int mps_stage_ppe_assemble(mps_ctx* ctx) {
  return mps::guard(ctx, [&](Context& c) {
    mps::require_neighbors(c);
    mps::ppe_assemble(c);
    mps::sync(c);
  });
}
int mps_stage_ppe_solve(mps_ctx* ctx) {
  return mps::guard(ctx, [&](Context& c) {
    mps::ppe_solve(c);
    mps::sync(c);
  });
}
int mps_stage_pressure_writeback(mps_ctx* ctx) {
  return mps::guard(ctx, [&](Context& c) {
    if (c.ppe_n != c.n) throw Failure{MPS_ERR_STATE, "installed system does not match the particles"};
    mps::pressure_writeback(c);
    mps::sync(c);
  });
}
int mps_stage_pressure_gradient(mps_ctx* ctx) {
  return mps::guard(ctx, [&](Context& c) {
    mps::require_neighbors(c);
    mps::pressure_gradient(c);
    mps::sync(c);
  });
}
int mps_stage_dynamic_stabilize(mps_ctx* ctx) {
  return mps::guard(ctx, [&](Context& c) {
    mps::require_neighbors(c);
    mps::dynamic_stabilize(c);
    mps::sync(c);
  });
}

int mps_density_variation(mps_ctx* ctx, double* out) {
  return mps::guard(ctx, [&](Context& c) {
    mps::require_neighbors(c);
    if (c.n == 0) return;
    if (!out) throw Failure{MPS_ERR_BAD_ARG, "null output"};
    mps::density_variation(c, c.hs_out);
    mps::sync(c);
    mps::download_caller_order(c, c.hs_out, out, 8);
  });
}

int mps_get_neighbor_counts(mps_ctx* ctx, uint32_t* counts) {
  return mps::guard(ctx, [&](Context& c) {
    if (c.n == 0) return;
    mps::sync(c);
    mps::download_caller_order(c, c.ncount, counts, 4);
  });
}

int mps_get_neighbors(mps_ctx* ctx, uint32_t* ids) {
  return mps::guard(ctx, [&](Context& c) {
    const size_t N = c.n;
    if (N == 0) return;
    std::vector<uint32_t> cnt(N);
    const size_t slots = (N + 63) / 64 * 64 * static_cast<size_t>(c.kstride);
    std::vector<uint32_t> all(slots);
    MPS_HIP(hipMemcpyAsync(cnt.data(), c.ncount, N * 4, hipMemcpyDeviceToHost, c.stream));
    MPS_HIP(hipMemcpyAsync(all.data(), c.nbr, slots * 4, hipMemcpyDeviceToHost, c.stream));
    mps::sync(c);
    // particles in the caller's order, neighbours as the caller's indices
    const std::vector<uint32_t>& orig = mps::storage_order(c);
    const std::vector<uint32_t> at = mps::inverse(orig);
    size_t o = 0;
    for (size_t e = 0; e < N; e++) {
      const size_t i = at[e];
      for (uint32_t k = 0; k < cnt[i]; k++)
        // slot k of particle i (the device layout, mps_device.hpp nbr_slot)
        ids[o++] = orig[all[(((i >> 6) * (c.kstride >> 2) + (k >> 2)) * 64 + (i & 63)) * 4 + (k & 3)]];
    }
  });
}

int mps_get_aux(mps_ctx* ctx, int which, double* out) {
  return mps::guard(ctx, [&](Context& c) {
    size_t count = 0;
    double* p = mps::aux_ptr(c, which, count);
    if (count == 0) return;
    mps::sync(c);
    if (count == c.n)  // per particle (the PPE vectors of a particles' system too)
      mps::download_caller_order(c, p, out, 8);
    else
      MPS_HIP(hipMemcpy(out, p, count * 8, hipMemcpyDeviceToHost));
  });
}

int mps_set_aux(mps_ctx* ctx, int which, const double* in) {
  return mps::guard(ctx, [&](Context& c) {
    size_t count = 0;
    double* p = mps::aux_ptr(c, which, count);
    if (count == 0) return;
    mps::sync(c);
    if (count == c.n)
      mps::upload_caller_order(c, in, p, 8);
    else
      MPS_HIP(hipMemcpy(p, in, count * 8, hipMemcpyHostToDevice));
  });
}

int mps_ppe_shape(mps_ctx* ctx, size_t* n, size_t* nnz) {
  return mps::guard(ctx, [&](Context& c) {
    if (!c.have_ppe) throw Failure{MPS_ERR_STATE, "no pressure Poisson system assembled"};
    if (n) *n = c.ppe_n;
    if (nnz) *nnz = static_cast<size_t>(c.nnz_offdiag) + c.ppe_n;
  });
}

int mps_get_ppe(mps_ctx* ctx, uint64_t* rowptr, uint32_t* col, double* val) {
  return mps::guard(ctx, [&](Context& c) {
    if (!c.have_ppe) throw Failure{MPS_ERR_STATE, "no pressure Poisson system assembled"};
    std::vector<uint64_t> rp;
    std::vector<uint32_t> cl;
    std::vector<double> vl;
    mps::ppe_download_csr(c, rp, cl, vl);
    if (mps::ppe_over_particles(c) && c.order_serial != 0) {
      // rows and columns back to the caller's particle order
      const std::vector<uint32_t>& orig = mps::storage_order(c);
      std::vector<uint64_t> rp2;
      std::vector<uint32_t> cl2;
      std::vector<double> vl2;
      mps::permute_csr(c.n, rp.data(), cl.data(), vl.data(), mps::inverse(orig), orig, rp2, cl2, vl2);
      rp.swap(rp2);
      cl.swap(cl2);
      vl.swap(vl2);
    }
    std::memcpy(rowptr, rp.data(), rp.size() * 8);
    std::memcpy(col, cl.data(), cl.size() * 4);
    std::memcpy(val, vl.data(), vl.size() * 8);
  });
}

int mps_set_ppe(mps_ctx* ctx, size_t n, const uint64_t* rowptr, const uint32_t* col, const double* val,
                const double* b, const double* x0) {
  return mps::guard(ctx, [&](Context& c) {
    if (!rowptr || !b || !x0 || (rowptr[n] > 0 && (!col || !val)))
      throw Failure{MPS_ERR_BAD_ARG, "null system array"};
    if (n == c.n && n > 0 && c.order_serial != 0) {
      // a system over the particles, given in the caller's order: stored in
      // the particles' storage order
      const std::vector<uint32_t>& orig = mps::storage_order(c);
      std::vector<uint64_t> rp2;
      std::vector<uint32_t> cl2;
      std::vector<double> vl2;
      mps::permute_csr(n, rowptr, col, val, orig, mps::inverse(orig), rp2, cl2, vl2);
      std::vector<double> b2(n), x2(n);
      for (size_t k = 0; k < n; k++) {
        b2[k] = b[orig[k]];
        x2[k] = x0[orig[k]];
      }
      mps::ppe_upload_csr(c, n, rp2.data(), cl2.data(), vl2.data(), b2.data(), x2.data());
      return;
    }
    mps::ppe_upload_csr(c, n, rowptr, col, val, b, x0);
  });
}

int mps_set_storage_order(mps_ctx* ctx, int order) {
  if (order != MPS_ORDER_CELL && order != MPS_ORDER_KEEP) return MPS_ERR_BAD_ARG;
  return mps::guard(ctx, [&](Context& c) { c.keep_order = (order == MPS_ORDER_KEEP); });
}

int mps_set_profiling(mps_ctx* ctx, int on) {
  return mps::guard(ctx, [&](Context& c) {
    mps::sync(c);
    c.profiling = (on != 0);
  });
}

int mps_get_kernel_stat(mps_ctx* ctx, int id, mps_kernel_stat* out) {
  if (!ctx || !out || id < 0 || id >= MPS_K_COUNT) return MPS_ERR_BAD_ARG;
  const mps::KStat& s = ctx->c.kstat[id];
  out->name = mps::kKernelNames[id];
  out->launches = s.launches;
  out->total_ms = s.ms;
  out->total_bytes = s.bytes;
  return MPS_OK;
}

int mps_reset_kernel_stats(mps_ctx* ctx) {
  return mps::guard(ctx, [&](Context& c) {
    mps::sync(c);
    for (auto& s : c.kstat) s = mps::KStat{};
  });
}

int mps_replay_spmv(mps_ctx* ctx, int id, int reps, double* ms_per_launch) {
  if (!ms_per_launch || reps <= 0 || (id != MPS_K_CG_SPMV32 && id != MPS_K_CG_SPMV)) return MPS_ERR_BAD_ARG;
  return mps::guard(ctx, [&](Context& c) { *ms_per_launch = mps::replay_spmv(c, id, reps); });
}

// ---- learned warm start ----------------------------------------------------------

}  // extern "C"

namespace mps {
namespace {
template <class H, class F>
int guard_obj(H* h, F&& f) {
  if (!h) return MPS_ERR_BAD_ARG;
  try {
    MPS_HIP(hipSetDevice(h->device));
    f();
    return MPS_OK;
  } catch (Failure& e) {
    std::fprintf(stderr, "mps: %s\n", e.msg.c_str());
    if (e.code == MPS_ERR_HIP) (void)hipGetLastError();
    return e.code;
  } catch (std::exception& e) {
    std::fprintf(stderr, "mps: %s\n", e.what());
    return MPS_ERR_BAD_ARG;
  }
}
// host [256][132] <-> device [256][160] for w1; other parameters are dense
void params_to_device(mps::Model& m, const float* const* src) {
  for (int k = 0; k < 6; k++) {
    const size_t n = mps::Model::param_count(k);
    std::vector<float> h(n, 0.0f);
    if (k == 0) {
      for (int r = 0; r < mps::kHid; r++)
        for (int q = 0; q < mps::kFeat; q++) h[r * mps::kFeatPad + q] = src[0][r * mps::kFeat + q];
    } else {
      std::memcpy(h.data(), src[k], n * 4);
    }
    MPS_HIP(hipMemcpy(m.w[k], h.data(), n * 4, hipMemcpyHostToDevice));
  }
  m.refresh_bf16();
}
void params_to_host(float* const* base, float* const* dst) {
  for (int k = 0; k < 6; k++) {
    const size_t n = mps::Model::param_count(k);
    std::vector<float> h(n);
    MPS_HIP(hipMemcpy(h.data(), base[k], n * 4, hipMemcpyDeviceToHost));
    if (k == 0) {
      for (int r = 0; r < mps::kHid; r++)
        for (int q = 0; q < mps::kFeat; q++) dst[0][r * mps::kFeat + q] = h[r * mps::kFeatPad + q];
    } else {
      std::memcpy(dst[k], h.data(), n * 4);
    }
  }
}
}  // namespace
}  // namespace mps

extern "C" {

int mps_model_create(int device, int seed, mps_model** out) {
  if (!out) return MPS_ERR_BAD_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return MPS_ERR_NO_DEVICE;
  }
  mps_model* h = new (std::nothrow) mps_model;
  if (!h) return MPS_ERR_BAD_ARG;
  h->device = device;
  const int rc = mps::guard_obj(h, [&] { h->m.init(seed); });
  if (rc != MPS_OK) {
    h->m.release();
    delete h;
    return rc;
  }
  *out = h;
  return MPS_OK;
}

void mps_model_destroy(mps_model* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  m->m.release();
  delete m;
}

int mps_model_get_params(mps_model* m, float* w1, float* b1, float* w2, float* b2, float* w3, float* b3) {
  return mps::guard_obj(m, [&] {
    float* dst[6] = {w1, b1, w2, b2, w3, b3};
    for (float* p : dst)
      if (!p) throw Failure{MPS_ERR_BAD_ARG, "null parameter array"};
    mps::params_to_host(m->m.w, dst);
  });
}

int mps_model_set_params(mps_model* m, const float* w1, const float* b1, const float* w2, const float* b2,
                         const float* w3, const float* b3) {
  return mps::guard_obj(m, [&] {
    const float* src[6] = {w1, b1, w2, b2, w3, b3};
    for (const float* p : src)
      if (!p) throw Failure{MPS_ERR_BAD_ARG, "null parameter array"};
    mps::params_to_device(m->m, src);
  });
}

int mps_model_forward(mps_model* m, size_t rows, const float* x, float* y) {
  return mps::guard_obj(m, [&] {
    if (rows == 0) return;
    mps::model_forward_host(m->m, rows, x, y);
  });
}

int mps_model_gradients(mps_model* m, size_t rows, const float* x, const float* t, float* loss, float* g_w1,
                        float* g_b1, float* g_w2, float* g_b2, float* g_w3, float* g_b3) {
  return mps::guard_obj(m, [&] {
    if (rows == 0) throw Failure{MPS_ERR_BAD_ARG, "no rows"};
    mps::model_grads_host(m->m, rows, x, t, loss);
    float* dst[6] = {g_w1, g_b1, g_w2, g_b2, g_w3, g_b3};
    mps::params_to_host(m->m.g, dst);
  });
}

int mps_model_train(mps_model* m, mps_dataset* d, int epochs, int batch, double lr, uint64_t seed,
                    double* losses_out) {
  if (!d) return MPS_ERR_BAD_ARG;
  return mps::guard_obj(m, [&] { m->m.train(d->d, epochs, batch, static_cast<float>(lr), seed, losses_out); });
}

int mps_dataset_create(int device, size_t capacity, mps_dataset** out) {
  if (!out || capacity == 0) return MPS_ERR_BAD_ARG;
  *out = nullptr;
  mps_dataset* h = new (std::nothrow) mps_dataset;
  if (!h) return MPS_ERR_BAD_ARG;
  h->device = device;
  const int rc = mps::guard_obj(h, [&] { h->d.init(capacity); });
  if (rc != MPS_OK) {
    h->d.release();
    delete h;
    return rc;
  }
  *out = h;
  return MPS_OK;
}

void mps_dataset_destroy(mps_dataset* d) {
  if (!d) return;
  (void)hipSetDevice(d->device);
  d->d.release();
  delete d;
}

int mps_dataset_size(mps_dataset* d, size_t* n) {
  return mps::guard_obj(d, [&] { *n = d->d.size(); });
}

int mps_dataset_get(mps_dataset* d, size_t first, size_t count, float* x, float* t) {
  return mps::guard_obj(d, [&] {
    if (first + count > d->d.size()) throw Failure{MPS_ERR_BAD_ARG, "range beyond the dataset"};
    std::vector<__bf16> hx(count * mps::kFeatPad);
    MPS_HIP(hipMemcpy(hx.data(), d->d.x + first * mps::kFeatPad, hx.size() * 2, hipMemcpyDeviceToHost));
    if (t) MPS_HIP(hipMemcpy(t, d->d.t + first, count * 4, hipMemcpyDeviceToHost));
    if (x)
      for (size_t r = 0; r < count; r++)
        for (int q = 0; q < mps::kFeat; q++) x[r * mps::kFeat + q] = static_cast<float>(hx[r * mps::kFeatPad + q]);
  });
}

int mps_attach_model(mps_ctx* ctx, mps_model* m) {
  if (!ctx) return MPS_ERR_BAD_ARG;
  ctx->c.model = m ? &m->m : nullptr;
  return MPS_OK;
}

int mps_attach_dataset(mps_ctx* ctx, mps_dataset* d, double fraction, uint64_t seed) {
  if (!ctx || !(fraction >= 0 && fraction <= 1)) return MPS_ERR_BAD_ARG;
  ctx->c.dataset = d ? &d->d : nullptr;
  ctx->c.collect_frac = fraction;
  ctx->c.collect_seed = seed;
  ctx->c.collect_step = 0;
  return MPS_OK;
}

int mps_get_features(mps_ctx* ctx, float* out) {
  return mps::guard(ctx, [&](Context& c) {
    if (!c.have_ppe || c.ppe_n != c.n) throw Failure{MPS_ERR_STATE, "assemble the PPE first"};
    const size_t n = c.n;
    if (n == 0) return;
    std::vector<__bf16> h(n * mps::kFeatPad);
    mps::learned_features(c);
    MPS_HIP(hipMemcpyAsync(h.data(), c.feats, h.size() * 2, hipMemcpyDeviceToHost, c.stream));
    mps::sync(c);
    const std::vector<uint32_t>& orig = mps::storage_order(c);  // rows in the caller's order
    for (size_t r = 0; r < n; r++)
      for (int q = 0; q < mps::kFeat; q++)
        out[static_cast<size_t>(orig[r]) * mps::kFeat + q] = static_cast<float>(h[r * mps::kFeatPad + q]);
  });
}

int mps_device_copy_probe(size_t bytes, double* ms_out) {
  try {
    const double ms = mps::device_copy_probe(bytes);
    if (ms_out) *ms_out = ms;
  } catch (Failure& f) {
    std::fprintf(stderr, "mps_device_copy_probe: %s\n", f.msg.c_str());
    return f.code;
  }
  return MPS_OK;
}

int mps_device_math_check(size_t n, const double* a, const double* b, double* out_sqrt, double* out_div) {
  try {
    mps::device_math_check(n, a, b, out_sqrt, out_div);
  } catch (Failure& f) {
    std::fprintf(stderr, "mps_device_math_check: %s\n", f.msg.c_str());
    return f.code;
  }
  return MPS_OK;
}

}  // extern "C"
