// Neighbour search: Computer::SearchNeighbor (Computer.hpp:698-756) over the
// uniform grid of Grid.hpp, rebuilt for the GPU.
//
// The reference inserts particles serially into fixed-capacity cells and then
// walks, for each particle, the 9 surrounding cells in x-major order
// (Grid.hpp:393-401) with ascending particle index inside a cell.  Here:
//   1. k_cell_keys     cell key per particle (out-of-grid -> Disabled) and an
//                      atomic per-cell histogram, which also detects the
//                      "Too many particle in a block" overflow (Grid.hpp:314-318);
//   2. exclusive scan  of the histogram -> cell start offsets;
//   3. k_cell_scatter  counting-sort scatter (a one-digit radix sort whose digit
//                      is the whole cell key);
//   4. k_cell_sort     orders each cell's few entries by particle index, which
//                      restores exactly the serial insertion order; k_cell_pos
//                      stores the positions in that order;
//   5. k_neighbors     one thread per particle walks the 9 cells in the
//                      reference order (three contiguous ranges of ids and
//                      positions, one per column of cells) and writes its
//                      list into the sliced neighbour layout, four slots
//                      (16 bytes) per store.
#include <utility>

#include "mps_device.hpp"
#include "mps_internal.hpp"

namespace mps {
namespace {

constexpr uint32_t kNoKey = kNoCell;

struct GridParams {
  double min_x, min_z, nl;
  long long gnx, gnz;
  uint32_t cap;
  // squared-distance bounds around NL^2 (relative 2^-48): below lo a pair is
  // certainly within NL, above hi certainly not, between them R is computed
  double lo, hi;
};

// Grid::Block (Grid.hpp:250-254) with the range test of Grid::Store
// (Grid.hpp:278-290).  The floored quotient is range-checked as a double, so
// NaN / inf / huge coordinates fall outside the grid as on the reference host.
__device__ __forceinline__ bool cell_of(const GridParams& g, double px, double pz, long long& bx,
                                        long long& bz) {
  const double fx = floor((px - g.min_x) / g.nl);
  const double fz = floor((pz - g.min_z) / g.nl);
  if (!(fx >= 0 && fx < static_cast<double>(g.gnx) && fz >= 0 && fz < static_cast<double>(g.gnz)))
    return false;
  bx = static_cast<long long>(fx);
  bz = static_cast<long long>(fz);
  return true;
}

// atomicAdd(&counter[k], 1) for every valid lane of the wavefront, one
// atomic per run of equal keys in consecutive lanes (particles stored in
// lattice order put three or four neighbours of a wavefront in one cell):
// the run's first lane adds the run's length and hands the old value on.
// Returns old + the lane's position in its run, the value of one of the
// orders in which the single atomics could have run.  Every lane of the
// wavefront must call it; lanes with valid == false (key kNoKey) add nothing.
__device__ __forceinline__ uint32_t run_atomic_add(uint32_t* counter, uint32_t k, bool valid) {
  const int lane = threadIdx.x & 63;
  const uint32_t kk = valid ? k : kNoKey;
  const uint32_t prev = __shfl_up(kk, 1, 64);
  const bool head = lane == 0 || kk != prev || kk == kNoKey;
  const unsigned long long heads = __ballot(head);
  const unsigned long long le = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  const int start = 63 - __clzll(heads & le);  // lane 0 is a head: never empty
  const unsigned long long above = heads & ~le;
  const int end = above ? __ffsll(above) - 1 : 64;
  uint32_t old = 0;
  if (head && kk != kNoKey) old = atomicAdd(&counter[kk], static_cast<uint32_t>(end - lane));
  return __shfl(old, start, 64) + static_cast<uint32_t>(lane - start);
}

// The search's counters cleared in one launch: cell counts and fill
// positions, the flags and the list total.
__global__ void __launch_bounds__(kBlock) k_search_clear(size_t ncells, uint32_t* __restrict__ count,
                                                         uint32_t* __restrict__ fill, int* __restrict__ flags,
                                                         unsigned long long* __restrict__ total) {
  const size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x;
  if (i < ncells) {
    count[i] = 0;
    fill[i] = 0;
  }
  if (i < 4) flags[i] = 0;
  if (i == 0) *total = 0;
}

__global__ void __launch_bounds__(kBlock) k_cell_keys(GridParams g, size_t n, const double* __restrict__ x,
                                                      const double* __restrict__ z, uint8_t* __restrict__ type,
                                                      uint32_t* __restrict__ key, uint32_t* cell_count,
                                                      int* flags) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  uint32_t k = kNoKey;  // lanes past n take part in the wavefront's run count
  if (i < n && type[i] != kDisabled) {
    long long bx, bz;
    if (cell_of(g, x[i], z[i], bx, bz))
      k = static_cast<uint32_t>(bx * g.gnz + bz);
    else
      type[i] = kDisabled;  // Computer.hpp:711-715
  }
  const uint32_t l = run_atomic_add(cell_count, k, k != kNoKey);
  if (k != kNoKey && l >= g.cap) flags[0] = 1;
  if (i < n) key[i] = k;
}

// ---- exclusive scan of u32 counts (x mult) into u32 or u64 offsets --------------
constexpr int kScanItems = 8;
constexpr int kScanChunk = 256 * kScanItems;

__global__ void __launch_bounds__(256) k_scan_reduce(const uint32_t* __restrict__ in, size_t n,
                                                     uint64_t* __restrict__ bsum, uint64_t mult) {
  __shared__ uint64_t s[4];
  const size_t base = blockIdx.x * static_cast<size_t>(kScanChunk) + threadIdx.x * kScanItems;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++)
    if (base + k < n) acc += static_cast<uint64_t>(in[base + k]) * mult;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) bsum[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

// Inclusive scan of one value per thread across a 256-thread block.
__device__ __forceinline__ uint64_t block_inclusive_scan(uint64_t v, uint64_t* lds4) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) lds4[wv] = v;
  __syncthreads();
  uint64_t off = 0;
  for (int k = 0; k < wv; k++) off += lds4[k];
  __syncthreads();
  return v + off;
}

// Single block: exclusive scan of the block sums in place, grand total to *total.
__global__ void __launch_bounds__(256) k_scan_top(uint64_t* bsum, int nb, uint64_t* total) {
  __shared__ uint64_t lds4[4];
  __shared__ uint64_t tot;
  uint64_t carry = 0;
  for (int base = 0; base < nb; base += 256) {
    const int idx = base + threadIdx.x;
    const uint64_t v = (idx < nb) ? bsum[idx] : 0;
    const uint64_t inc = block_inclusive_scan(v, lds4);
    if (idx < nb) bsum[idx] = carry + inc - v;
    if (threadIdx.x == 255) tot = inc;
    __syncthreads();
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <class TO>
__global__ void __launch_bounds__(256) k_scan_apply(const uint32_t* __restrict__ in, size_t n,
                                                    const uint64_t* __restrict__ bsum,
                                                    TO* __restrict__ out, uint64_t mult) {
  __shared__ uint64_t lds4[4];
  const size_t base = blockIdx.x * static_cast<size_t>(kScanChunk) + threadIdx.x * kScanItems;
  uint64_t v[kScanItems];
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    v[k] = (base + k < n) ? static_cast<uint64_t>(in[base + k]) * mult : 0;
    acc += v[k];
  }
  const uint64_t inc = block_inclusive_scan(acc, lds4);
  uint64_t run = bsum[blockIdx.x] + inc - acc;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n) out[base + k] = static_cast<TO>(run);
    run += v[k];
  }
}

// Counting-sort scatter of the particles in cells; particles outside every
// cell (disabled) go after them, at in_cells + (running count in flags[2]),
// in no particular order: they take part in nothing.
__global__ void __launch_bounds__(kBlock) k_cell_scatter(size_t n, const uint32_t* __restrict__ key,
                                                         const uint32_t* __restrict__ cell_start,
                                                         uint32_t* cell_fill,
                                                         uint32_t* __restrict__ sorted,
                                                         const uint64_t* __restrict__ in_cells, int* flags) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  const uint32_t k = i < n ? key[i] : kNoKey;  // lanes past n take part in the run count
  const uint32_t slot = run_atomic_add(cell_fill, k, k != kNoKey);
  if (k != kNoKey)
    sorted[cell_start[k] + slot] = static_cast<uint32_t>(i);
  else if (i < n)
    sorted[in_cells[0] + static_cast<uint32_t>(atomicAdd(&flags[2], 1))] = static_cast<uint32_t>(i);
}

// Orders each cell's entries by the caller's index orig[] (the reference
// inserts serially in index order, Computer.hpp:704-717; the stored order
// is a permutation of it).  Overflowed cells are left alone: the search
// reports an error for them.
__global__ void __launch_bounds__(kBlock) k_cell_sort(size_t ncells, uint32_t cap,
                                                      const uint32_t* __restrict__ cell_start,
                                                      const uint32_t* __restrict__ cell_count,
                                                      const uint32_t* __restrict__ orig,
                                                      uint32_t* __restrict__ sorted) {
  const size_t c = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (c >= ncells) return;
  const uint32_t cnt = cell_count[c];
  if (cnt < 2 || cnt > cap) return;
  uint32_t* s = sorted + cell_start[c];
  if (cnt <= 32) {  // a bitonic network in registers over (caller index, slot) pairs
    uint64_t v[32];
#pragma unroll
    for (int e = 0; e < 32; e++) {
      const uint32_t j = (static_cast<uint32_t>(e) < cnt) ? s[e] : 0u;
      v[e] = (static_cast<uint32_t>(e) < cnt) ? ((static_cast<uint64_t>(orig[j]) << 32) | j)
                                              : ~0ull;  // padding sorts last
    }
#pragma unroll
    for (int k = 2; k <= 32; k <<= 1)
#pragma unroll
      for (int d = k >> 1; d > 0; d >>= 1)
#pragma unroll
        for (int e = 0; e < 32; e++) {
          const int l = e ^ d;
          if (l > e) {
            const uint64_t lo = v[e] < v[l] ? v[e] : v[l], hi = v[e] < v[l] ? v[l] : v[e];
            const bool up = (e & k) == 0;
            v[e] = up ? lo : hi;
            v[l] = up ? hi : lo;
          }
        }
#pragma unroll
    for (int e = 0; e < 32; e++)
      if (static_cast<uint32_t>(e) < cnt) s[e] = static_cast<uint32_t>(v[e]);
    return;
  }
  for (uint32_t a = 1; a < cnt; a++) {
    const uint32_t v = s[a], ov = orig[v];
    uint32_t b = a;
    while (b > 0 && orig[s[b - 1]] > ov) {
      s[b] = s[b - 1];
      b--;
    }
    s[b] = v;
  }
}

// The particle state in cell order: slot k takes the particle at old slot
// sorted[k] (then sorted[k] = k).  Every per-particle array moves, also the
// stage results a later stage skips for some particle types and so keeps
// (saved X, ECS, nWithoutSpp, du of dummy particles: the reference never
// overwrites them).  Particles outside the grid (the tail) get empty
// neighbour lists.
struct Permute {
  size_t n;
  uint32_t* sorted;
  const uint64_t* in_cells;
  const double* src[Context::kMoved];
  double* dst[Context::kMoved];
  const uint8_t* type_s;
  uint8_t* type_d;
  const uint8_t* grp_s;  // null: no device wall groups
  uint8_t* grp_d;
  const uint32_t* orig_s;
  uint32_t* orig_d;
  const uint32_t* key_s;
  uint32_t* key_d;
  uint32_t* ncount;
};
__global__ void __launch_bounds__(kBlock) k_permute(Permute a) {
  const size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (k >= a.n) return;
  const uint32_t j = a.sorted[k];
#pragma unroll
  for (int m = 0; m < Context::kMoved; m++) a.dst[m][k] = a.src[m][j];
  a.type_d[k] = a.type_s[j];
  if (a.grp_s) a.grp_d[k] = a.grp_s[j];
  a.orig_d[k] = a.orig_s[j];
  a.key_d[k] = a.key_s[j];
  a.sorted[k] = static_cast<uint32_t>(k);
  if (k >= a.in_cells[0]) a.ncount[k] = 0;
}

// The positions in cell order (slots below the number of particles in
// cells), so that the neighbour walk reads a cell's candidates as one
// contiguous range instead of gathering them by index.
__global__ void __launch_bounds__(kBlock) k_cell_pos(size_t n, const uint64_t* __restrict__ in_cells,
                                                     const uint32_t* __restrict__ sorted,
                                                     const double* __restrict__ x, const double* __restrict__ z,
                                                     double2* __restrict__ pos) {
  const size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (k >= n || k >= in_cells[0]) return;
  const uint32_t j = sorted[k];
  pos[k] = make_double2(x[j], z[j]);
}

// Computer::SearchNeighbor's per-particle walk, Computer.hpp:720-755: the 9
// cells in Grid.hpp:393-401 order, j != i, R(x_i, x_j) < NeighborLength.
// Disabled particles keep their previous list, as in the reference.
//
// The cells (cx, cz - 1), (cx, cz), (cx, cz + 1) are consecutive cell
// indices, so their members are one contiguous range of the cell-ordered
// slots, in the reference's visiting order (cells in turn, each cell's
// members by index): the walk is three ranges, one per column of cells.
// IDENT: the particles are stored in cell order (slot k holds particle k),
// so the candidate ids are the slots themselves and are not loaded.
//
// A lane's found neighbours form groups of four slots (16 bytes, the list
// layout's unit).  Stored by each lane as it completes them, the groups of a
// wavefront reach a cache line at different times and the lines are written
// back part-filled (2.2x the list bytes).  Instead each lane keeps its last
// kRing completed groups in a ring in LDS, and after each batch of
// candidates (the batch loop is uniform for the wavefront) the groups every
// lane has completed are stored together: one 1-KB coalesced store per group
// index.  A lane more than kRing groups ahead of the slowest stores its
// oldest ring group directly to make room; the end of the walk stores the
// rest.
constexpr int kRing = 4;
constexpr int kNbrBatch = 4;  // candidates per lane per round trip to memory
template <bool IDENT>
__global__ void __launch_bounds__(kBlock) k_neighbors(GridParams g, size_t n, const double* __restrict__ x,
                                                      const double* __restrict__ z,
                                                      const uint8_t* __restrict__ type,
                                                      const uint32_t* __restrict__ cell_start,
                                                      const uint32_t* __restrict__ cell_count,
                                                      const uint32_t* __restrict__ sorted,
                                                      const double2* __restrict__ pos, uint32_t kmax,
                                                      uint32_t kstride, uint32_t* __restrict__ nbr,
                                                      uint32_t* __restrict__ ncount, int* flags,
                                                      unsigned long long* total) {
  __shared__ uint4 ring[kRing][kBlock];
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const size_t i = blk * static_cast<size_t>(blockDim.x) + threadIdx.x;
  const bool active = i < n && type[i] != kDisabled;
  uint32_t cnt = 0;
  // complete groups of this lane (those whose last slot is below kmax);
  // groups below `flushed` (a wavefront value) are stored, and of the others
  // the last kRing are in ring[group % kRing], the older ones stored
  uint32_t ngrp = 0, flushed = 0;
  uint4 grp = make_uint4(0u, 0u, 0u, 0u);
  uint4* const out = reinterpret_cast<uint4*>(nbr);
  const uint32_t self = static_cast<uint32_t>(i);
  auto store = [&](uint32_t gi, const uint4& v) { out[nbr_slot(self, 4 * gi, kstride) >> 2] = v; };
  double xi = 0, zi = 0;
  long long bx = 0, bz = 0;
  if (active) {
    xi = x[i];
    zi = z[i];
    cell_of(g, xi, zi, bx, bz);  // in the grid: the others are disabled
  }
  const double nl = g.nl;
  auto test = [&](bool valid, uint32_t j, double xj, double zj) {
    if (!valid || j == i) return;
    const double rx = xi - xj, rz = zi - zj;
    // Computer::R, Computer.hpp:568-572: sqrt(r.r) < NL.  sqrt is correctly
    // rounded, so the test only depends on r.r near NL^2: outside the
    // narrow band [lo, hi] the comparison of r.r decides it exactly.
    const double q = rx * rx + rz * rz;
    bool within = q < g.lo;
    // the band is a few ulps of NL^2 wide: its square root is taken on its
    // own branch, which a wavefront almost never enters (evaluated beside
    // the comparisons it costs every candidate beyond lo twenty FP64
    // instructions)
    if (q >= g.lo && q <= g.hi) within = sqrt(q) < nl;
    if (within) {
      const uint32_t lane = cnt & 3u;
      grp.x = (lane == 0) ? j : grp.x;
      grp.y = (lane == 1) ? j : grp.y;
      grp.z = (lane == 2) ? j : grp.z;
      grp.w = (lane == 3) ? j : grp.w;
      if (lane == 3 && cnt < kmax) {
        // group ngrp is complete; a full ring gives up its oldest group
        if (ngrp >= flushed + kRing) store(ngrp - kRing, ring[ngrp % kRing][threadIdx.x]);
        ring[ngrp % kRing][threadIdx.x] = grp;
        ngrp++;
      }
      cnt++;
    }
  };
  for (int ox = -1; ox <= 1; ox++) {
    const long long cx = bx + ox;
    uint32_t s = 0, m = 0;
    if (active && cx >= 0 && cx < g.gnx) {
      const size_t c0 = static_cast<size_t>(cx * g.gnz + (bz > 0 ? bz - 1 : 0));
      const size_t c1 = static_cast<size_t>(cx * g.gnz + (bz + 1 < g.gnz ? bz + 1 : bz));
      s = cell_start[c0];
      m = cell_start[c1] + cell_count[c1] - s;
    }
    // a batch loop uniform for the wavefront (lanes past their range skip
    // the tests), so that the completed groups can be stored after a batch
    const uint32_t mw = wave_max_u32(m);
    for (uint32_t l = 0; l < mw; l += kNbrBatch) {
      uint32_t jj[kNbrBatch];
      double2 pj[kNbrBatch];
      bool ok[kNbrBatch];
#pragma unroll
      for (int v = 0; v < kNbrBatch; v++) {
        ok[v] = l + v < m;
        const uint32_t k = s + (ok[v] ? l + v : 0u);  // a valid slot (unused when !ok)
        jj[v] = IDENT ? k : (ok[v] ? sorted[k] : 0u);
        pj[v] = ok[v] ? pos[k] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int v = 0; v < kNbrBatch; v++) test(ok[v], jj[v], pj[v].x, pj[v].y);
      // the groups every active lane has completed
      const uint32_t upto = wave_min_u32(active ? ngrp : 0xffffffffu);
      if (upto != 0xffffffffu) {
        for (uint32_t gi = flushed; gi < upto; gi++)
          if (active && gi + kRing >= ngrp) store(gi, ring[gi % kRing][threadIdx.x]);
        flushed = upto > flushed ? upto : flushed;
      }
    }
  }
  // the rest: the complete groups still in the ring and the last, partial
  // one (its slots past the count are never read)
  const bool partial = active && (cnt & 3u) && cnt <= kmax;
  const uint32_t mine = active ? ngrp + (partial ? 1u : 0u) : 0u;
  const uint32_t last = wave_max_u32(mine);
  for (uint32_t gi = flushed; gi < last; gi++) {
    if (gi < ngrp && gi + kRing >= ngrp) store(gi, ring[gi % kRing][threadIdx.x]);
    if (partial && gi == ngrp) store(gi, grp);
  }
  if (active) {
    if (cnt > kmax) {
      flags[1] = 1;
      cnt = kmax;
    }
    ncount[i] = cnt;
  }
  const double ws = wave_sum(static_cast<double>(cnt));
  if ((threadIdx.x & 63) == 0 && ws > 0) atomicAdd(total, static_cast<unsigned long long>(ws));
}

}  // namespace

// Exclusive scan out[k] = sum_{m<k} in[m] * mult over n entries; optional total.
template <class TO>
static void scan_u32(Context& c, const uint32_t* in, size_t n, TO* out, uint64_t mult, uint64_t* total) {
  const int nb = static_cast<int>((n + kScanChunk - 1) / kScanChunk);
  if (nb == 0) return;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, c.stream, in, n, c.scan_tmp, mult);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, c.stream, c.scan_tmp, nb, total);
  hipLaunchKernelGGL(k_scan_apply<TO>, dim3(nb), dim3(256), 0, c.stream, in, n,
                     static_cast<const uint64_t*>(c.scan_tmp), out, mult);
}

void scan_counts(Context& c, const uint32_t* in, size_t n, uint32_t* out, uint64_t* total) {
  scan_u32<uint32_t>(c, in, n, out, 1, total);
}

void scan_slices(Context& c, const uint32_t* widths, size_t nslices, uint64_t* offsets) {
  scan_u32<uint64_t>(c, widths, nslices, offsets, 64, offsets + nslices);
}

// Moves the particle state into the cell order the sort left in c.sorted.
static void permute_to_cell_order(Context& c) {
  const size_t n = c.n;
  double** live[Context::kMoved] = {&c.x,   &c.z,   &c.u,      &c.w,      &c.p,  &c.nd, &c.wall_x,
                                    &c.wall_z, &c.nws, &c.ecs, &c.ox, &c.oz, &c.dux, &c.duz};
  Permute a{};
  a.n = n;
  a.sorted = c.sorted;
  a.in_cells = c.in_cells;
  for (int m = 0; m < Context::kMoved; m++) {
    a.src[m] = *live[m];
    a.dst[m] = c.spare_d[m];
  }
  a.type_s = c.type;
  a.type_d = c.spare_type;
  if (c.wall_group) {
    a.grp_s = c.wall_group;
    a.grp_d = c.spare_group;
  }
  a.orig_s = c.orig;
  a.orig_d = c.spare_orig;
  a.key_s = c.key;
  a.key_d = c.spare_key;
  a.ncount = c.ncount;
  // each moved array read (gathered) and written once, the sorted ids read and written
  const double bytes = 2.0 * n * (8 * Context::kMoved + 1 + 4 + 4 + (c.wall_group ? 1 : 0)) + 8.0 * n;
  tracked(c, MPS_K_CELL_SORT, bytes, [&] {
    hipLaunchKernelGGL(k_permute, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, a);
  });
  for (int m = 0; m < Context::kMoved; m++) std::swap(*live[m], c.spare_d[m]);
  std::swap(c.type, c.spare_type);
  if (c.wall_group) std::swap(c.wall_group, c.spare_group);
  std::swap(c.orig, c.spare_orig);
  std::swap(c.key, c.spare_key);
  c.order_serial++;
  // a system assembled over the particles refers to the previous order
  if (c.have_ppe && c.ppe_n == c.n) {
    c.have_ppe = false;
    c.have_ppe32 = false;
    c.ppe_from_particles = false;
  }
}

void search(Context& c) {
  const size_t n = c.n;
  c.search_serial++;
  const double nl2 = c.env.neighbor_length * c.env.neighbor_length;
  GridParams g{c.env.min_x, c.env.min_z, c.env.neighbor_length, c.gnx, c.gnz, c.cap,
               nl2 * (1.0 - 0x1p-48), nl2 * (1.0 + 0x1p-48)};
  hipLaunchKernelGGL(k_search_clear, dim3(grid_for(std::max<size_t>(c.ncells, 4))), dim3(kBlock), 0, c.stream,
                     c.ncells, c.cell_count, c.cell_fill, c.flags, c.nbr_total_dev);
  check_launch();
  if (n == 0) {
    c.have_neighbors = true;
    return;
  }
  tracked(c, MPS_K_CELL_KEYS, 26.0 * n, [&] {
    hipLaunchKernelGGL(k_cell_keys, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, g, n, c.x, c.z,
                       c.type, c.key, c.cell_count, c.flags);
  });
  tracked(c, MPS_K_CELL_SCAN, 8.0 * c.ncells, [&] {
    scan_u32<uint32_t>(c, c.cell_count, c.ncells, c.cell_start, 1, c.in_cells);
  });
  tracked(c, MPS_K_CELL_SCATTER, 16.0 * n, [&] {
    hipLaunchKernelGGL(k_cell_scatter, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n,
                       static_cast<const uint32_t*>(c.key), static_cast<const uint32_t*>(c.cell_start),
                       c.cell_fill, c.sorted, static_cast<const uint64_t*>(c.in_cells), c.flags);
  });
  // sorted ids read and written, their caller indices gathered
  tracked(c, MPS_K_CELL_SORT, 12.0 * n + 8.0 * c.ncells, [&] {
    hipLaunchKernelGGL(k_cell_sort, dim3(grid_for(c.ncells)), dim3(kBlock), 0, c.stream, c.ncells,
                       c.cap, static_cast<const uint32_t*>(c.cell_start),
                       static_cast<const uint32_t*>(c.cell_count), static_cast<const uint32_t*>(c.orig),
                       c.sorted);
  });
  if (!c.keep_order) permute_to_cell_order(c);
  // the sorted ids and x, z read, the positions in cell order written
  tracked(c, MPS_K_CELL_SORT, 36.0 * n, [&] {
    hipLaunchKernelGGL(k_cell_pos, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n,
                       static_cast<const uint64_t*>(c.in_cells), static_cast<const uint32_t*>(c.sorted),
                       static_cast<const double*>(c.x), static_cast<const double*>(c.z), c.pos_sorted);
  });
  // x, z, type of every particle, the cell tables, the sorted ids and
  // positions and the list lengths; the list itself (4 bytes a neighbour) is
  // added below
  tracked(c, MPS_K_NEIGHBORS, 41.0 * n + 8.0 * c.ncells, [&] {
    // in cell order (after the permutation) slot k holds particle k
    auto walk = c.keep_order ? k_neighbors<false> : k_neighbors<true>;
    hipLaunchKernelGGL(walk, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, g, n, c.x, c.z,
                       static_cast<const uint8_t*>(c.type), static_cast<const uint32_t*>(c.cell_start),
                       static_cast<const uint32_t*>(c.cell_count), static_cast<const uint32_t*>(c.sorted),
                       static_cast<const double2*>(c.pos_sorted), c.kmax, c.kstride, c.nbr, c.ncount, c.flags,
                       c.nbr_total_dev);
  });
  check_launch();
  MPS_HIP(hipMemcpyAsync(c.flags_host, c.flags, 4 * sizeof(int), hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipMemcpyAsync(c.nbr_total_host, c.nbr_total_dev, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
  c.nbr_total = *c.nbr_total_host;
  add_bytes(c, MPS_K_NEIGHBORS, 4.0 * c.nbr_total);
  c.have_neighbors = true;
  if (c.flags_host[0] || c.flags_host[1])
    throw Failure{MPS_ERR_CELL_OVERFLOW, "Too many particle in a block"};
}

}  // namespace mps
