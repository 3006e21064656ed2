// Pressure Poisson equation assembly: Computer::SetPressurePoissonEquation
// (Computer.hpp:1145-1356) rebuilt as two neighbour passes that write a
// device-resident sliced ELL matrix (SELL-64: one slice = one wavefront of
// 64 rows, slot k of row i at slice_off[i/64] + 64k + i%64).
//
//   k_ppe_count  RHS b_i (HS + ECS, FP64), warm start x0_i = p_i, the
//                off-diagonal count of row i and, per slice, the wavefront
//                maximum of those counts and whether the slice is wide (a
//                column j with j - i outside int16);
//   scan         slice widths -> slice offsets;
//   k_ppe_fill   HL off-diagonals 3 r_e / (n0 r^3) with ascending columns
//                (int16 offsets j - i, and u32 columns in the wide slices
//                only), zero padding, and the diagonal -sum(a_ij) including
//                the SPP ghost (Computer.hpp:1290-1292).
//
// The reference builds per-row vectors in parallel and scatters them serially
// into a uBLAS compressed matrix before copying to ViennaCL CSR
// (:1330-1354); nothing here is serial.
#include <algorithm>

#include "mps_device.hpp"
#include "mps_internal.hpp"

namespace mps {

void scan_slices(Context& c, const uint32_t* widths, size_t nslices, uint64_t* offsets);

namespace {

struct NV {
  const uint32_t* __restrict__ ids;
  const uint32_t* __restrict__ count;
  uint32_t kmax;
};

struct PpeConsts {
  double re, re2, rl, rl2, n0, n0l, rho, dt;
  int warm;             // mps_warm_start
  double qx0, qz0, qinv;  // fixed-point positions: q = (x - x0) * qinv mod 2^32
};

constexpr int kWalk = 4;  // neighbours whose loads are issued together (walk_list)

// Column j of row i fits the 16-bit offset j - i.
__device__ __forceinline__ bool offset16(uint32_t j, size_t i) {
  const long long off = static_cast<long long>(j) - static_cast<long long>(i);
  return off >= -32768 && off <= 32767;
}

// Per slice (one wavefront of rows, all lanes converged): the longest row,
// the stored count, and whether any row has a column beyond a 16-bit offset
// (then the fill stores the slice's 32-bit columns too).
__device__ __forceinline__ void count_slice(size_t n, size_t sl, uint32_t cnt, bool fits,
                                            uint32_t* __restrict__ slice_width, uint8_t* __restrict__ slice_wide,
                                            unsigned long long* nnz_total, unsigned long long* nnz_wide) {
  const unsigned wmax = wave_max_u32(cnt);
  const unsigned long long wsum = static_cast<unsigned long long>(wave_sum(static_cast<double>(cnt)));
  const unsigned long long far = __ballot(!fits);
  if ((threadIdx.x & 63) == 0 && sl * 64 < n) {
    slice_width[sl] = wmax;
    slice_wide[sl] = static_cast<uint8_t>(far != 0);
    atomicAdd(nnz_total, wsum);
    if (far) atomicAdd(nnz_wide, wsum);
  }
}

// b and x0 (Computer.hpp:1185-1221), row lengths and slice widths.
__global__ void __launch_bounds__(kBlock) k_ppe_count(size_t n, const double* __restrict__ x,
                                                      const double* __restrict__ z,
                                                      const double* __restrict__ u,
                                                      const double* __restrict__ w,
                                                      const double* __restrict__ p,
                                                      const uint8_t* __restrict__ type, NV nb,
                                                      PpeConsts k, const double* __restrict__ nws,
                                                      const double* __restrict__ ecs,
                                                      double* __restrict__ b, double* __restrict__ sol,
                                                      uint32_t* __restrict__ row_len,
                                                      uint32_t* __restrict__ slice_width,
                                                      uint8_t* __restrict__ slice_wide,
                                                      uint2* __restrict__ qpos,
                                                      unsigned long long* nnz_total,
                                                      unsigned long long* nnz_wide) {
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const size_t i = blk * static_cast<size_t>(blockDim.x) + threadIdx.x;
  uint32_t cnt = 0;
  bool fits = true;
  if (i < n) {
    // integer differences of q are exact, so r_ij in the FP32 operator
    // carries only the quantisation step (|domain| / 2^30)
    qpos[i] = make_uint2(static_cast<uint32_t>(static_cast<long long>(rint((x[i] - k.qx0) * k.qinv))),
                         static_cast<uint32_t>(static_cast<long long>(rint((z[i] - k.qz0) * k.qinv))));
    const uint8_t ti = type[i];
    if (ti == kDummy || ti == kDisabled) {
      b[i] = 0;
      sol[i] = 0;
    } else {
      const double xi = x[i], zi = z[i], ui = u[i], wi = w[i];
      double h = 0.0;
      struct F {
        double x, z, u, w;
        uint8_t t;
      };
      walk_list<kWalk, F>(static_cast<uint32_t>(i), nb.count[i], nb.ids, nb.kmax,
                          [&](uint32_t j) { return F{x[j], z[j], u[j], w[j], type[j]}; },
                          [&](uint32_t j, const F& f) {
                            const double dx = f.x - xi, dz = f.z - zi;
                            const double r2 = dx * dx + dz * dz;
                            if (r2 < k.re2) {
                              const double r = sqrt(r2);
                              const double du = f.u - ui, dw = f.w - wi;
                              h += (dx * du + dz * dw) / (r * r * r);
                            }
                            if (r2 < k.rl2 && f.t != kDummy) {
                              cnt++;
                              fits = fits && offset16(j, i);
                            }
                          });
      const double speed = (nws[i] < k.n0) ? 0.0 : -k.re * h;  // HS, :838-872
      b[i] = -k.rho / (k.n0 * k.dt) * (speed + ecs[i]);
      sol[i] = (k.warm == MPS_WARM_ZERO) ? 0.0 : p[i];
    }
    row_len[i] = cnt;
  }
  count_slice(n, i >> 6, cnt, fits, slice_width, slice_wide, nnz_total, nnz_wide);
}

// The second ComputeNeighborDensities pass with SaveX (Computer.hpp:780-833,
// :1025-1039) and k_ppe_count in one walk of the neighbour list: nothing
// moves between the two stages, so the RHS reads the positions, velocities
// and nWithoutSpp the density pass has in hand.  Every value is computed as
// the two separate kernels compute it (same expressions, same list order).
__global__ void __launch_bounds__(kBlock) k_density_ppe_count(size_t n, const double* __restrict__ x,
                                                              const double* __restrict__ z,
                                                              const double* __restrict__ u,
                                                              const double* __restrict__ w,
                                                              const double* __restrict__ p,
                                                              const uint8_t* __restrict__ type, NV nb,
                                                              PpeConsts k, double* __restrict__ nd,
                                                              double* __restrict__ nws,
                                                              const double* __restrict__ ecs,
                                                              double* __restrict__ ox, double* __restrict__ oz,
                                                              double* __restrict__ b, double* __restrict__ sol,
                                                              uint32_t* __restrict__ row_len,
                                                              uint32_t* __restrict__ slice_width,
                                                              uint8_t* __restrict__ slice_wide,
                                                              uint2* __restrict__ qpos,
                                                              unsigned long long* nnz_total,
                                                              unsigned long long* nnz_wide,
                                                              const double2* __restrict__ xz,
                                                              const double2* __restrict__ uw) {
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const size_t i = blk * static_cast<size_t>(blockDim.x) + threadIdx.x;
  uint32_t cnt = 0;
  bool fits = true;
  if (i < n) {
    const double xi = x[i], zi = z[i];
    qpos[i] = make_uint2(static_cast<uint32_t>(static_cast<long long>(rint((xi - k.qx0) * k.qinv))),
                         static_cast<uint32_t>(static_cast<long long>(rint((zi - k.qz0) * k.qinv))));
    const uint8_t ti = type[i];
    if (ti == kDummy || ti == kDisabled) {
      nws[i] = k.n0;
      b[i] = 0;
      sol[i] = 0;
    } else {
      ox[i] = xi;
      oz[i] = zi;
      const double ui = u[i], wi = w[i];
      double s = 0.0, h = 0.0;
      const uint32_t m = nb.count[i];
      auto use = [&](uint32_t j, double xj, double zj, double uj, double wj, uint8_t tj) {
        const double dx = xj - xi, dz = zj - zi;
        const double r2 = dx * dx + dz * dz;
        if (r2 < k.re2) {
          const double r = sqrt(r2);
          s += weight(r, k.re);
          const double du = uj - ui, dw = wj - wi;
          h += (dx * du + dz * dw) / (r * r * r);
        }
        if (r2 < k.rl2 && tj != kDummy) {
          cnt++;
          fits = fits && offset16(j, i);
        }
      };
      // kWalk neighbours per round trip: their ids, then all their fields,
      // loaded before any is used; the sums still run in list order
      uint32_t q = 0;
      for (; q + kWalk <= m; q += kWalk) {
        uint32_t jj[kWalk];
#pragma unroll
        for (int v = 0; v < kWalk; v++) jj[v] = nb.ids[nbr_slot(static_cast<uint32_t>(i), q + v, nb.kmax)];
        double2 pj[kWalk], vj[kWalk];
        uint8_t tj[kWalk];
#pragma unroll
        for (int v = 0; v < kWalk; v++) {
          pj[v] = xz[jj[v]];
          vj[v] = uw[jj[v]];
          tj[v] = type[jj[v]];
        }
#pragma unroll
        for (int v = 0; v < kWalk; v++) use(jj[v], pj[v].x, pj[v].y, vj[v].x, vj[v].y, tj[v]);
      }
      for (; q < m; q++) {
        const uint32_t j = nb.ids[nbr_slot(static_cast<uint32_t>(i), q, nb.kmax)];
        const double2 pj = xz[j], vj = uw[j];
        use(j, pj.x, pj.y, vj.x, vj.y, type[j]);
      }
      nws[i] = s;
      nd[i] = (s < k.n0) ? k.n0 : s;
      const double speed = (s < k.n0) ? 0.0 : -k.re * h;  // HS, :838-872
      b[i] = -k.rho / (k.n0 * k.dt) * (speed + ecs[i]);
      sol[i] = (k.warm == MPS_WARM_ZERO) ? 0.0 : p[i];
    }
    row_len[i] = cnt;
  }
  count_slice(n, i >> 6, cnt, fits, slice_width, slice_wide, nnz_total, nnz_wide);
}

// HL off-diagonals and diagonal (Computer.hpp:1232-1328).  ghost32 keeps the
// SPP ghost's share of the diagonal (diag = ghost - sum_j a_ij) for the
// matrix-free FP32 operator; rows without unknowns (dummy, disabled) get
// ghost32 = diag = 1 and no off-diagonals.
//
// The diagonal is summed in neighbour-list order (bit-exact with the
// reference); the off-diagonals are stored with ascending column.  A list is
// the concatenation of at most 9 ascending runs (one per cell, each cell's
// members ascending), so a 9-way merge orders it without a sort; the list
// walk stages the emitted columns in LDS, so the merge's chain of dependent
// reads stays in LDS.  Ascending columns make slot k of neighbouring rows
// point at neighbouring particles: the solver's gathers of a wavefront touch
// a few cache lines, not 64.  A row with more than kStage columns or more
// than 9 runs is stored in list order.
constexpr int kRuns = 9;

// Optional output of the fill for the multilevel preconditioner: per row i
// (0 for dummy / disabled rows and rows outside the grid), for the cell, the
// 2x2 and the 4x4 block of cells of i (cells of the neighbour search, `key`):
// S_l(i) = A_ii + sum of the stored A_ij with j in the same aggregate, summed
// in stored (ascending column) order.  s0 == nullptr: not wanted.
struct AggSums {
  const uint32_t* __restrict__ key;
  uint32_t gnz;
  float *s0, *s1, *s2;
  // the cell-grid multigrid's per-row sums (rows9 != nullptr; the fill's
  // ROWS9 instance): per row i with stored off-diagonals, not dummy or
  // disabled, in a cell: A_ii + the stored A_ij whose column is such a row
  // too, summed by the direction of the column's cell (d = 3 (dx + 1) +
  // dz + 1), in stored order; 0 elsewhere; FP64, rows9[9 i + d]
  double* rows9;
  const uint32_t* __restrict__ aggkey;  // ROWS9: the cell of each such row, kNoCell for the others
};

// The cell of each row that the multigrid aggregates (stored off-diagonals,
// not dummy or disabled, in a cell), kNoCell for the others: one gather per
// stored column in the fill instead of three.
__global__ void __launch_bounds__(kBlock) k_agg_key(size_t n, const uint32_t* __restrict__ key,
                                                    const uint8_t* __restrict__ type,
                                                    const uint32_t* __restrict__ row_len, uint32_t* __restrict__ out) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const uint8_t t = type[i];
  out[i] = (t != kDummy && t != kDisabled && row_len[i] > 0) ? key[i] : kNoCell;
}

constexpr int kStage = 40;  // emitted columns per row staged in LDS for the merge

// A stored value and the direction of its column's cell (-1: not summed).
struct DirA {
  int dir;
  double a;
};
__device__ __forceinline__ DirA make_pair_da(int dir, double a) { return DirA{dir, a}; }

// One row of k_ppe_fill; kk returns its stored non-zeros.
template <bool ROWS9>
__device__ __forceinline__ void fill_row(size_t n, size_t i, const double* __restrict__ x,
                                         const double* __restrict__ z, const uint8_t* __restrict__ type, NV nb,
                                         const PpeConsts& k, const double* __restrict__ nws,
                                         const uint64_t* __restrict__ slice_off,
                                         const uint32_t* __restrict__ slice_width, double* __restrict__ val,
                                         uint32_t* __restrict__ col, int16_t* __restrict__ col16,
                                         double* __restrict__ diag, float* __restrict__ ghost32, const AggSums& ag,
                                         uint32_t (*stage)[kBlock], bool wide, uint32_t& kk,
                                         const double2* __restrict__ xz) {
  (void)n;
  const int t = threadIdx.x;
  const size_t base = slice_off[i >> 6] + (i & 63);
  const uint32_t width = slice_width[i >> 6];
  const uint32_t self = static_cast<uint32_t>(i);
  const uint8_t ti = type[i];
  kk = 0;
  double d, g = 0.0;
  double a0 = 0, a1 = 0, a2 = 0;  // aggregate row sums (AggSums)
  double r9[9];                    // ROWS9
#pragma unroll
  for (int u = 0; u < 9; u++) r9[u] = 0.0;
  if (ti == kDummy || ti == kDisabled) {
    d = 1.0;
    g = 1.0;
  } else {
    const double xi = x[i], zi = z[i];
    double s = 0.0, sx = 0, sz = 0;
    const uint32_t m = nb.count[i];
    // the emitted columns (r < r_lap, not dummy) form at most 9 ascending
    // runs, one per cell: their starts, for the merge below
    uint32_t er[kRuns] = {};
    int nr = 0;
    uint32_t ne = 0, prev = 0;
    bool staged = true;
    struct F {
      double x, z;
      uint8_t t;
    };
    walk_list<kWalk, F>(self, m, nb.ids, nb.kmax,
                        [&](uint32_t j) {
                          const double2 pj = xz[j];
                          return F{pj.x, pj.y, type[j]};
                        },
                        [&](uint32_t j, const F& f) {
      const double dx = f.x - xi, dz = f.z - zi;
      const double r2 = dx * dx + dz * dz;
      if (r2 < k.rl2 && f.t != kDummy) {
        const double r = sqrt(r2);
        s += -(3 * k.rl / k.n0l / (r * r * r));
        if (ne == 0 || j < prev) {
          if (nr < kRuns) {
#pragma unroll
            for (int u = 0; u < kRuns; u++)
              if (u == nr) er[u] = ne;
            nr++;
          } else {
            staged = false;  // not a cell-ordered list: store in list order
          }
        }
        prev = j;
        if (ne < kStage) stage[ne][t] = j;
        else staged = false;
        ne++;
      }
      if (r2 < k.re2) {
        const double wg = weight(sqrt(r2), k.re);
        sx += wg * dx;
        sz += wg * dz;
      }
                        });
    double gxp, gzp;
    if (ghost_pos(xi, zi, sx, sz, nws[i], k.n0, k.re, gxp, gzp)) {
      const double rx = xi - gxp, rz = zi - gzp;
      const double r = sqrt(rx * rx + rz * rz);
      const double a = 3 * k.rl / k.n0l / (r * r * r);
      s += -a;
      g = -a;
    }
    d = s;
    const uint32_t ki = ag.s0 ? ag.key[i] : kNoCell;
    const uint32_t cx = ki / ag.gnz, cz = ki % ag.gnz;
    if (ki != kNoCell) a0 = a1 = a2 = d;
    // ROWS9: the row's cell when it takes part (its row length is known from the count pass)
    const uint32_t k9 = ROWS9 ? ag.aggkey[i] : kNoCell;
    const int c9x = static_cast<int>(k9 / ag.gnz), c9z = static_cast<int>(k9 % ag.gnz);
    // value of an emitted column j in slot kk (and its share of the aggregate row sums)
    auto emit = [&](uint32_t j) {
      const double2 pj = xz[j];
      const double dx = pj.x - xi, dz = pj.y - zi;
      const double r = sqrt(dx * dx + dz * dz);
      const size_t o = base + static_cast<size_t>(kk) * 64;
      const double a = 3 * k.rl / k.n0l / (r * r * r);
      val[o] = a;
      if (wide) col[o] = j;
      col16[o] = static_cast<int16_t>(static_cast<long long>(j) - static_cast<long long>(self));
      kk++;
      if (ki != kNoCell) {
        const uint32_t kj = ag.key[j];
        if (kj != kNoCell) {
          const uint32_t jx = kj / ag.gnz, jz = kj % ag.gnz;
          if ((jx >> 2) == (cx >> 2) && (jz >> 2) == (cz >> 2)) {
            a2 += a;
            if ((jx >> 1) == (cx >> 1) && (jz >> 1) == (cz >> 1)) {
              a1 += a;
              if (jx == cx && jz == cz) a0 += a;
            }
          }
        }
      }
      int dir = -1;
      if (ROWS9 && k9 != kNoCell) {
        const uint32_t kj = ag.aggkey[j];
        if (kj != kNoCell) dir = 3 * (static_cast<int>(kj / ag.gnz) - c9x + 1) + static_cast<int>(kj % ag.gnz) - c9z + 1;
      }
      return make_pair_da(dir, a);
    };
    // consecutive columns of one direction are summed in a run and the run
    // added to its direction's sum when the direction changes (in cell order
    // the ascending columns visit the directions in ascending order: one run
    // each); direction 4's first run starts from the diagonal
    int cur = -1;
    double run = 0.0;
    bool seen4 = false;
    auto flush9 = [&]() {
#pragma unroll
      for (int u = 0; u < 9; u++) r9[u] += (u == cur) ? run : 0.0;
    };
    auto add9 = [&](const DirA& e) {
      if (ROWS9 && e.dir >= 0) {
        if (e.dir != cur) {
          if (cur >= 0) flush9();
          cur = e.dir;
          run = (cur == 4 && !seen4) ? d : 0.0;
          seen4 = seen4 || cur == 4;
        }
        run += e.a;
      }
    };
    if (staged && nr <= 1) {
      // one ascending run (the particles in cell order: the list visits the
      // cells, and each cell's members, in ascending index order): the
      // staged columns in their order
#pragma unroll 4
      for (uint32_t v = 0; v < ne; v++) add9(emit(stage[v][t]));
    } else if (staged) {
      // 9-way merge of the staged runs (LDS only on its dependency chain),
      // the ascending columns written first, the values computed after
      uint32_t cur[kRuns], end[kRuns], hd[kRuns];
#pragma unroll
      for (int u = 0; u < kRuns; u++) {
        cur[u] = (u < nr) ? er[u] : 0u;
        end[u] = (u < nr) ? ne : 0u;
      }
#pragma unroll
      for (int u = 0; u + 1 < kRuns; u++)
        if (u + 1 < nr) end[u] = er[u + 1];
#pragma unroll
      for (int u = 0; u < kRuns; u++) hd[u] = (cur[u] < end[u]) ? stage[cur[u]][t] : 0xffffffffu;
      for (uint32_t v = 0; v < ne; v++) {
        uint32_t best = hd[0];
        int sel = 0;
#pragma unroll
        for (int u = 1; u < kRuns; u++)
          if (hd[u] < best) {
            best = hd[u];
            sel = u;
          }
        uint32_t nc = 0, nend = 0;
#pragma unroll
        for (int u = 0; u < kRuns; u++)
          if (u == sel) {
            nc = cur[u] + 1;
            nend = end[u];
          }
        const uint32_t nh = (nc < nend) ? stage[nc][t] : 0xffffffffu;
#pragma unroll
        for (int u = 0; u < kRuns; u++)
          if (u == sel) {
            cur[u] = nc;
            hd[u] = nh;
          }
        col[base + static_cast<size_t>(v) * 64] = best;
      }
#pragma unroll 4
      for (uint32_t v = 0; v < ne; v++) add9(emit(col[base + static_cast<size_t>(v) * 64]));
    } else {
      for (uint32_t q = 0; q < m; q++) {
        const uint32_t j = nb.ids[nbr_slot(self, q, nb.kmax)];
        const double2 pj = xz[j];
        const double dx = pj.x - xi, dz = pj.y - zi;
        if (dx * dx + dz * dz < k.rl2 && type[j] != kDummy) add9(emit(j));
      }
    }
    if (ROWS9 && k9 != kNoCell) {
      if (cur >= 0) flush9();
      if (!seen4) r9[4] = d;
    }
  }
  const uint32_t stored = kk;
  for (uint32_t pad = kk; pad < width; pad++) {
    const size_t o = base + static_cast<size_t>(pad) * 64;
    val[o] = 0.0;
    if (wide) col[o] = self;
    col16[o] = 0;
  }
  kk = stored;
  diag[i] = d;
  ghost32[i] = static_cast<float>(g);
  if (ag.s0) {
    ag.s0[i] = static_cast<float>(a0);
    ag.s1[i] = static_cast<float>(a1);
    ag.s2[i] = static_cast<float>(a2);
  }
  if (ROWS9) {
#pragma unroll
    for (int u = 0; u < 9; u++) ag.rows9[9 * i + u] = r9[u];
  }
}

// the stage array limits a CU to four blocks (four wavefronts per SIMD): the
// registers of that occupancy are available
template <bool ROWS9>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) k_ppe_fill(size_t n, const double* __restrict__ x,
                                                     const double* __restrict__ z,
                                                     const uint8_t* __restrict__ type, NV nb,
                                                     PpeConsts k, const double* __restrict__ nws,
                                                     const uint64_t* __restrict__ slice_off,
                                                     const uint32_t* __restrict__ slice_width,
                                                     double* __restrict__ val, uint32_t* __restrict__ col,
                                                     int16_t* __restrict__ col16,
                                                     const uint8_t* __restrict__ slice_wide, double* __restrict__ diag,
                                                     float* __restrict__ ghost32, AggSums ag,
                                                     const double2* __restrict__ xz) {
  // each thread's emitted columns in list order, [slot][thread] (conflict-free)
  __shared__ uint32_t stage[kStage][kBlock];
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const size_t i = blk * static_cast<size_t>(blockDim.x) + threadIdx.x;
  // int16 column offsets j - i in every slot; the u32 columns (same slots)
  // only in the slices the count pass flagged wide (an offset beyond int16)
  uint32_t kk = 0;
  if (i < n)
    fill_row<ROWS9>(n, i, x, z, type, nb, k, nws, slice_off, slice_width, val, col, col16, diag, ghost32, ag,
                    stage, slice_wide[i >> 6] != 0, kk, xz);
}

__global__ void __launch_bounds__(kBlock) k_to_f32(size_t n, const double* __restrict__ a,
                                                   float* __restrict__ b) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) b[i] = static_cast<float>(a[i]);
}

}  // namespace

void ensure_sell_capacity(Context& c, size_t nslices, size_t slots) {
  if (slots > c.sell_cap) {
    dfree(c.sval);
    dfree(c.scol);
    dfree(c.sval32);
    dfree(c.col16);
    const size_t cap = slots + slots / 8 + 64;
    c.sval = dalloc<double>(cap);
    c.scol = dalloc<uint32_t>(cap);
    c.col16 = dalloc<int16_t>(cap);
    c.sell_cap = cap;
    c.have_ppe32 = false;
  }
  (void)nslices;
}

namespace {

// Shapes, capacity and constants of the system about to be assembled.
PpeConsts ppe_prepare(Context& c) {
  const size_t n = c.n;
  c.ppe_n = n;
  c.nslices = (n + 63) / 64;
  c.have_ppe = true;
  c.have_ppe32 = false;
  // worst case: every row as long as the longest neighbour list
  if (n > 0) ensure_sell_capacity(c, c.nslices, c.nslices * 64 * static_cast<size_t>(c.kmax));
  const double extent = std::max(c.env.max_x - c.env.min_x, c.env.max_z - c.env.min_z);
  c.qpos_inv_h = (extent > 0) ? 1073741824.0 / extent : 1.0 / c.env.l0;
  PpeConsts k{c.env.r_e, c.env.r_e * c.env.r_e, c.env.r_lap, c.env.r_lap * c.env.r_lap, c.env.n0,
              c.env.n0_lap, c.env.rho, c.dt, c.solver.warm_start, c.env.min_x, c.env.min_z,
              c.qpos_inv_h};
  if (n > 0) MPS_HIP(hipMemsetAsync(c.ppe_counters, 0, 6 * sizeof(unsigned long long), c.stream));
  return k;
}

// The box of cells holding the rows with stored off-diagonals (the unknowns
// of the cell-grid multigrid), in two passes: per block of k_row_box the
// extremes (minima as maxima of gn - 1 - v; 0 in the second: no row), then
// k_row_box_final's maxima over the blocks into box[0..3].
constexpr int kBoxBlocks = 1024;
__global__ void __launch_bounds__(kBlock) k_row_box(size_t n, const uint32_t* __restrict__ row_len,
                                                    const uint32_t* __restrict__ key, uint32_t gnx, uint32_t gnz,
                                                    uint4* __restrict__ part) {
  __shared__ uint4 lds[kBlock / 64];
  uint4 m = make_uint4(0u, 0u, 0u, 0u);
  for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
       i += static_cast<size_t>(kBlock) * gridDim.x) {
    if (row_len[i] == 0) continue;
    const uint32_t k = key[i];
    if (k == kNoCell) continue;
    const uint32_t cx = k / gnz, cz = k % gnz;
    m.x = max(m.x, gnx - 1 - cx);
    m.y = max(m.y, cx + 1);
    m.z = max(m.z, gnz - 1 - cz);
    m.w = max(m.w, cz + 1);
  }
  m.x = wave_max_u32(m.x);
  m.y = wave_max_u32(m.y);
  m.z = wave_max_u32(m.z);
  m.w = wave_max_u32(m.w);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint4 t = lds[0];
    for (int k = 1; k < kBlock / 64; k++) {
      t.x = max(t.x, lds[k].x);
      t.y = max(t.y, lds[k].y);
      t.z = max(t.z, lds[k].z);
      t.w = max(t.w, lds[k].w);
    }
    part[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(256) k_row_box_final(const uint4* __restrict__ part, int nparts,
                                                       unsigned long long* __restrict__ box) {
  __shared__ uint4 lds[4];
  uint4 m = make_uint4(0u, 0u, 0u, 0u);
  for (int k = threadIdx.x; k < nparts; k += 256) {
    const uint4 t = part[k];
    m.x = max(m.x, t.x);
    m.y = max(m.y, t.y);
    m.z = max(m.z, t.z);
    m.w = max(m.w, t.w);
  }
  m.x = wave_max_u32(m.x);
  m.y = wave_max_u32(m.y);
  m.z = wave_max_u32(m.z);
  m.w = wave_max_u32(m.w);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint4 t = lds[0];
    for (int k = 1; k < 4; k++) {
      t.x = max(t.x, lds[k].x);
      t.y = max(t.y, lds[k].y);
      t.z = max(t.z, lds[k].z);
      t.w = max(t.w, lds[k].w);
    }
    box[0] = t.x;
    box[1] = t.y;
    box[2] = t.z;
    box[3] = t.w;
  }
}

// Enqueues the two passes (box into ppe_counters[2..5]).
void row_box_launch(Context& c, size_t n) {
  if (!c.box_part) c.box_part = dalloc<uint4>(kBoxBlocks);
  const int nb = static_cast<int>(std::min<size_t>(kBoxBlocks, std::max<size_t>(1, (n + kBlock - 1) / kBlock)));
  hipLaunchKernelGGL(k_row_box, dim3(nb), dim3(kBlock), 0, c.stream, n, static_cast<const uint32_t*>(c.row_len),
                     static_cast<const uint32_t*>(c.key), static_cast<uint32_t>(c.gnx),
                     static_cast<uint32_t>(c.gnz), c.box_part);
  hipLaunchKernelGGL(k_row_box_final, dim3(1), dim3(256), 0, c.stream, static_cast<const uint4*>(c.box_part), nb,
                     c.ppe_counters + 2);
}

// The box of cells of the rows with off-diagonals, from the counters k_row_box
// left in ppe_counters_host[3..6].
void amg_box_from_counters(Context& c) {
  const unsigned long long* h = c.ppe_counters_host + 3;
    if (h[1] == 0) {  // no row with off-diagonals: a one-cell level
      c.amg_box[0] = c.amg_box[1] = 0;
      c.amg_box[2] = c.amg_box[3] = 1;
    } else {
      const uint32_t x0 = static_cast<uint32_t>(c.gnx - 1 - h[0]), x1 = static_cast<uint32_t>(h[1] - 1);
      const uint32_t z0 = static_cast<uint32_t>(c.gnz - 1 - h[2]), z1 = static_cast<uint32_t>(h[3] - 1);
      c.amg_box[0] = x0;
      c.amg_box[1] = z0;
      c.amg_box[2] = x1 - x0 + 1;
      c.amg_box[3] = z1 - z0 + 1;
    }
  c.amg_box_ready = true;
}

// Slice offsets from the counted row lengths, then the fill.
void ppe_finish(Context& c, const PpeConsts& k) {
  const size_t n = c.n;
  const NV nb{c.nbr, c.ncount, c.kstride};
  scan_slices(c, c.slice_width, c.nslices, c.slice_off);
  const int pc = c.solver.kind == MPS_SOLVER_MIXED_IR ? precond_for_system(c) : MPS_PRECOND_NONE;
  const bool want_box = pc == MPS_PRECOND_AMG;
  if (want_box) {
    row_box_launch(c, n);
    MPS_HIP(hipMemcpyAsync(c.ppe_counters_host + 3, c.ppe_counters + 2, 4 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c.stream));
  }
  MPS_HIP(hipMemcpyAsync(c.ppe_counters_host, c.ppe_counters, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipMemcpyAsync(c.ppe_counters_host + 1, c.slice_off + c.nslices, sizeof(uint64_t),
                         hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipMemcpyAsync(c.ppe_counters_host + 2, c.ppe_counters + 1, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
  c.nnz_offdiag = c.ppe_counters_host[0];
  c.sell_slots = c.ppe_counters_host[1];
  // the wide slices' stored count (byte accounting only), from the count pass
  c.mf_nnz_wide = static_cast<double>(c.ppe_counters_host[2]);
  c.mf_wide_pending = false;
  c.amg_box_ready = false;
  if (want_box) amg_box_from_counters(c);
  // row sums of the multilevel preconditioner's aggregates, when that solver will run
  const bool want_sums = pc == MPS_PRECOND_MULTILEVEL;
  // the cell-grid multigrid's per-row direction sums, when that solver will run
  const bool want9 = want_box;
  if (want9) {
    // the aggregated rows' cells into spare_key (free between permutations)
    ensure_amg_rows(c, n);
    hipLaunchKernelGGL(k_agg_key, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n,
                       static_cast<const uint32_t*>(c.key), static_cast<const uint8_t*>(c.type),
                       static_cast<const uint32_t*>(c.row_len), c.spare_key);
  }
  const AggSums ag{c.key,  static_cast<uint32_t>(c.gnz),    want_sums ? c.z32 : nullptr, c.r32, c.d32,
                   c.amg_rows, static_cast<const uint32_t*>(c.spare_key)};
  // slots: value and 16-bit offset; the wide slices' stored entries also a 32-bit column
  const double bytes = n * (16.0 + 1 + 8 + 4 + 8 + 4) + 4.0 * c.nbr_total + c.sell_slots * 10.0 +
                       4.0 * c.mf_nnz_wide +
                       (want_sums ? n * 12.0 + c.nnz_offdiag * 4.0 : 0.0) +
                       (want9 ? n * (72.0 + 4) + c.nnz_offdiag * 4.0 : 0.0);  // k_agg_key's 13 N apart
  tracked(c, MPS_K_PPE_FILL, bytes, [&] {
    if (want9)
      hipLaunchKernelGGL(k_ppe_fill<true>, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z,
                         static_cast<const uint8_t*>(c.type), nb, k, static_cast<const double*>(c.nws),
                         static_cast<const uint64_t*>(c.slice_off),
                         static_cast<const uint32_t*>(c.slice_width), c.sval, c.scol, c.col16,
                         static_cast<const uint8_t*>(c.slice_wide), c.diag, c.ghost32, ag,
                         static_cast<const double2*>(c.xz));
    else
      hipLaunchKernelGGL(k_ppe_fill<false>, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z,
                         static_cast<const uint8_t*>(c.type), nb, k, static_cast<const double*>(c.nws),
                         static_cast<const uint64_t*>(c.slice_off),
                         static_cast<const uint32_t*>(c.slice_width), c.sval, c.scol, c.col16,
                         static_cast<const uint8_t*>(c.slice_wide), c.diag, c.ghost32, ag,
                         static_cast<const double2*>(c.xz));
  });
  check_launch();
  c.ml_rowsums_ready = want_sums;
  c.amg_rows_ready = want9;
  c.ppe_from_particles = true;
  c.ppe_col16 = true;
  c.ppe_search_serial = c.search_serial;
}

}  // namespace

void ppe_assemble(Context& c) {
  const PpeConsts k = ppe_prepare(c);
  const size_t n = c.n;
  if (n == 0) return;
  const NV nb{c.nbr, c.ncount, c.kstride};
  tracked(c, MPS_K_PPE_COUNT, n * (64.0 + 1 + 4) + 4.0 * c.nbr_total + n * 28.0, [&] {
    hipLaunchKernelGGL(k_ppe_count, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z, c.u, c.w,
                       c.p, static_cast<const uint8_t*>(c.type), nb, k, static_cast<const double*>(c.nws),
                       static_cast<const double*>(c.ecs), c.b, c.sol, c.row_len, c.slice_width,
                       c.slice_wide, c.qpos, c.ppe_counters, c.ppe_counters + 1);
  });
  check_launch();
  pack_state(c);  // the fill gathers the packed positions
  ppe_finish(c, k);
}

void density_save_and_assemble(Context& c) {
  const PpeConsts k = ppe_prepare(c);
  const size_t n = c.n;
  if (n == 0) return;
  const NV nb{c.nbr, c.ncount, c.kstride};
  // the second density pass's bytes and the count pass's (x, z, u, w, p, type,
  // ecs read; n, nWithoutSpp, X0, b, x0, row length, position written)
  tracked(c, MPS_K_DENSITY_SAVE, n * (40.0 + 1 + 8 + 4 + 16 + 16 + 16 + 4 + 8) + 4.0 * c.nbr_total, [&] {
    hipLaunchKernelGGL(k_density_ppe_count, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z, c.u,
                       c.w, c.p, static_cast<const uint8_t*>(c.type), nb, k, c.nd, c.nws,
                       static_cast<const double*>(c.ecs), c.ox, c.oz, c.b, c.sol, c.row_len, c.slice_width,
                       c.slice_wide, c.qpos, c.ppe_counters, c.ppe_counters + 1,
                       static_cast<const double2*>(c.xz),
                       static_cast<const double2*>(c.uw));
  });
  check_launch();
  ppe_finish(c, k);
}

void amg_box(Context& c) {
  if (c.amg_box_ready) return;
  row_box_launch(c, c.ppe_n);
  check_launch();
  MPS_HIP(hipMemcpyAsync(c.ppe_counters_host + 3, c.ppe_counters + 2, 4 * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
  amg_box_from_counters(c);
}

double mf_nnz_wide(Context& c) {
  if (c.mf_wide_pending) {
    c.mf_nnz_wide = static_cast<double>(c.ppe_counters_host[2]);
    c.mf_wide_pending = false;
  }
  return c.mf_nnz_wide;
}

// FP32 copies of an already assembled system (the mixed solver on a system
// installed through mps_set_ppe, or after switching solvers).
void ppe_make_fp32(Context& c) {
  if (c.have_ppe32 || c.ppe_n == 0) return;
  if (!c.sval32) c.sval32 = dalloc<float>(c.sell_cap);
  const size_t slots = static_cast<size_t>(c.sell_slots);
  if (slots > 0)  // a diagonal system stores no off-diagonal slot
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(slots)), dim3(kBlock), 0, c.stream, slots,
                       static_cast<const double*>(c.sval), c.sval32);
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(c.ppe_n)), dim3(kBlock), 0, c.stream, c.ppe_n,
                     static_cast<const double*>(c.diag), c.diag32);
  check_launch();
  c.have_ppe32 = true;
}

// CSR (any column order; diagonal entries may be absent) -> device SELL-64.
void ppe_upload_csr(Context& c, size_t n, const uint64_t* rowptr, const uint32_t* col,
                    const double* val, const double* b, const double* x0) {
  const size_t ns = (n + 63) / 64;
  std::vector<double> dg(n, 0.0);
  std::vector<uint32_t> len(n, 0), width(ns, 0);
  for (size_t i = 0; i < n; i++) {
    for (uint64_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
      if (col[k] >= n) throw Failure{MPS_ERR_BAD_ARG, "column index out of range"};
      if (col[k] == i) dg[i] += val[k];
      else len[i]++;
    }
    width[i / 64] = std::max(width[i / 64], len[i]);
  }
  std::vector<uint64_t> off(ns + 1, 0);
  for (size_t s = 0; s < ns; s++) off[s + 1] = off[s] + 64ull * width[s];
  const size_t slots = off[ns];
  std::vector<double> sv(slots, 0.0);
  std::vector<uint32_t> sc(slots, 0);
  for (size_t i = 0; i < n; i++) {
    const size_t base = off[i / 64] + (i % 64);
    uint32_t kk = 0;
    for (uint64_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
      if (col[k] == i) continue;
      sv[base + 64ull * kk] = val[k];
      sc[base + 64ull * kk] = col[k];
      kk++;
    }
    for (; kk < width[i / 64]; kk++) sc[base + 64ull * kk] = static_cast<uint32_t>(i);
  }
  ensure_ppe_vectors(c, n);
  ensure_sell_capacity(c, ns, slots);
  MPS_HIP(hipMemcpyAsync(c.sval, sv.data(), slots * 8, hipMemcpyHostToDevice, c.stream));
  MPS_HIP(hipMemcpyAsync(c.scol, sc.data(), slots * 4, hipMemcpyHostToDevice, c.stream));
  MPS_HIP(hipMemcpyAsync(c.diag, dg.data(), n * 8, hipMemcpyHostToDevice, c.stream));
  MPS_HIP(hipMemcpyAsync(c.row_len, len.data(), n * 4, hipMemcpyHostToDevice, c.stream));
  MPS_HIP(hipMemcpyAsync(c.slice_width, width.data(), ns * 4, hipMemcpyHostToDevice, c.stream));
  MPS_HIP(hipMemcpyAsync(c.slice_off, off.data(), (ns + 1) * 8, hipMemcpyHostToDevice, c.stream));
  MPS_HIP(hipMemcpyAsync(c.b, b, n * 8, hipMemcpyHostToDevice, c.stream));
  MPS_HIP(hipMemcpyAsync(c.sol, x0, n * 8, hipMemcpyHostToDevice, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
  c.ppe_n = n;
  c.nslices = ns;
  c.sell_slots = slots;
  uint64_t nnz = 0;
  for (size_t i = 0; i < n; i++) nnz += len[i];
  c.nnz_offdiag = nnz;
  c.have_ppe = true;
  c.have_ppe32 = false;
  c.ppe_col16 = false;
  c.ppe_from_particles = false;
}

// Device SELL-64 -> CSR with ascending columns, diagonal included.
void ppe_download_csr(Context& c, std::vector<uint64_t>& rowptr, std::vector<uint32_t>& col,
                      std::vector<double>& val) {
  const size_t n = c.ppe_n, ns = c.nslices;
  std::vector<double> dg(n), sv(c.sell_slots);
  std::vector<uint32_t> len(n), sc(c.sell_slots);
  std::vector<int16_t> s16(c.ppe_col16 ? c.sell_slots : 0);
  std::vector<uint8_t> wide(c.ppe_col16 ? ns : 0);
  std::vector<uint64_t> off(ns + 1);
  MPS_HIP(hipMemcpyAsync(dg.data(), c.diag, n * 8, hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipMemcpyAsync(len.data(), c.row_len, n * 4, hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipMemcpyAsync(off.data(), c.slice_off, (ns + 1) * 8, hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipMemcpyAsync(sv.data(), c.sval, c.sell_slots * 8, hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipMemcpyAsync(sc.data(), c.scol, c.sell_slots * 4, hipMemcpyDeviceToHost, c.stream));
  if (c.ppe_col16) {  // assembled by the fill: u32 columns in the wide slices only
    MPS_HIP(hipMemcpyAsync(s16.data(), c.col16, c.sell_slots * 2, hipMemcpyDeviceToHost, c.stream));
    MPS_HIP(hipMemcpyAsync(wide.data(), c.slice_wide, ns, hipMemcpyDeviceToHost, c.stream));
  }
  MPS_HIP(hipStreamSynchronize(c.stream));
  rowptr.assign(n + 1, 0);
  col.clear();
  val.clear();
  std::vector<std::pair<uint32_t, double>> row;
  for (size_t i = 0; i < n; i++) {
    row.clear();
    const size_t base = off[i / 64] + (i % 64);
    const bool narrow = c.ppe_col16 && !wide[i / 64];
    for (uint32_t k = 0; k < len[i]; k++) {
      const size_t o = base + 64ull * k;
      row.emplace_back(narrow ? static_cast<uint32_t>(static_cast<long long>(i) + s16[o]) : sc[o], sv[o]);
    }
    row.emplace_back(static_cast<uint32_t>(i), dg[i]);
    std::stable_sort(row.begin(), row.end(),
                     [](const std::pair<uint32_t, double>& a, const std::pair<uint32_t, double>& b) {
                       return a.first < b.first;
                     });
    for (auto& e : row) {
      col.push_back(e.first);
      val.push_back(e.second);
    }
    rowptr[i + 1] = col.size();
  }
}

}  // namespace mps
