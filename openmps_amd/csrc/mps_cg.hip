// Pressure solve: Computer::SolvePressurePoissonEquation (Computer.hpp:1359-1429)
// on the GPU, plus the mixed-precision iterative-refinement solver.
//
// CG iteration k is two kernels:
//   k_cg_spmv   reduces the r.r partials of the previous kernel, applies the
//               stop rule, computes A r (r is the only gathered vector) and
//               p_k = r + beta p_{k-1}, q_k = A r + beta q_{k-1} for its own
//               rows, and writes partial sums of p_k.q_k;
//   k_cg_update reduces the p.q partials, alpha = rr / pq, x += alpha p,
//               r -= alpha q, and writes partial sums of r.r.
// Partial sums live in fixed slots (one per block of a fixed-size grid); every
// block of the consuming kernel re-reduces them in the same order, so all
// blocks agree bitwise on alpha, beta and the stop decision without atomics or
// grid-wide fences.  The host enqueues iterations in chunks and reads the
// device status once per chunk; kernels after convergence exit immediately.
//
// Launch shape: at most 1024 blocks of 512 threads (8 waves per SIMD on 256
// CUs); block b owns one contiguous row range, and blocks are remapped so that
// each XCD walks a contiguous eighth of the rows (XCD-local L2 hits on the
// gathered vector entries).  The SELL row sum keeps four slots in flight per
// lane.  These kernels use explicit FMA: the CG is checked against the
// tolerance, not bit for bit.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>
#include <type_traits>

#include "mps_device.hpp"
#include "mps_internal.hpp"

namespace mps {
namespace {

constexpr int NT = 512;  // threads per block of the solver kernels
constexpr int kCgParts = 2;  // FP64 CG: partials per thread loaded ahead (1024 slots)
constexpr int kCgRows = 2;   // FP64 CG: rows per thread whose product waits in registers for beta

template <class T>
struct Sell {
  const uint64_t* __restrict__ off;
  const uint32_t* __restrict__ width;
  const T* __restrict__ val;
  const uint32_t* __restrict__ col;
  const T* __restrict__ diag;
  size_t n;
  size_t rpb;  // rows per block (multiple of NT)
  // optional: the columns as int16 offsets j - i (slices not flagged wide),
  // 2 bytes a slot instead of col's 4
  const int16_t* __restrict__ col16 = nullptr;
  const uint8_t* __restrict__ wide = nullptr;
};

// Column of slot o of row i: with the 16-bit offsets the 32-bit columns are
// stored for the wide slices only.
template <class T>
__device__ __forceinline__ uint32_t sell_col(const Sell<T>& A, size_t i, size_t o) {
  if (A.col16 && !A.wide[i >> 6]) return static_cast<uint32_t>(static_cast<int>(i) + A.col16[o]);
  return A.col[o];
}

// Where a block sits in its scene's grid.  A single-scene launch is the whole
// grid, its blocks taken in XCD-aware order; a grouped launch (Group below)
// concatenates the grids of several scenes, and each block runs as block
// `bid` of the `nb` blocks of its scene's grid.
struct Frame {
  unsigned bid, nb;
  bool remap;
};
__device__ __forceinline__ Frame whole_grid() { return Frame{blockIdx.x, gridDim.x, true}; }

struct Rows {
  size_t lo, hi;
};
__device__ __forceinline__ Rows block_rows(size_t n, size_t rpb, const Frame& f) {
  const unsigned blk = f.remap ? xcd_block(f.bid, f.nb) : f.bid;
  const size_t lo = blk * rpb;
  const size_t hi = (lo + rpb < n) ? lo + rpb : n;
  return Rows{lo, hi};
}
__device__ __forceinline__ Rows block_rows(size_t n, size_t rpb) { return block_rows(n, rpb, whole_grid()); }
__device__ __forceinline__ bool lead(const Frame& f) { return f.bid == 0 && threadIdx.x == 0; }
__device__ __forceinline__ bool lead() { return blockIdx.x == 0 && threadIdx.x == 0; }

// A grouped launch: the concatenated grids of `count` scenes, scene s owning
// blocks first[s] .. first[s + 1] - 1; scenes with active[s] == 0 (if active
// is given) exit at once.  Returns false for such a block; otherwise sets the
// block's scene and its frame in that scene's grid (block-uniform).
struct Group {
  const uint32_t* __restrict__ first;
  const int* __restrict__ active;
  int count;
};
__device__ __forceinline__ bool group_frame(const Group& g, int& s, Frame& f) {
  const unsigned b = blockIdx.x;
  int lo = 0, hi = g.count;  // first[lo] <= b < first[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (g.first[mid] <= b) lo = mid;
    else hi = mid;
  }
  s = lo;
  f = Frame{b - g.first[lo], g.first[lo + 1] - g.first[lo], false};
  return g.active == nullptr || g.active[lo] != 0;
}

// Block-uniform read of the convergence flag (thread 0 reads, LDS broadcast),
// so that a block either runs every barrier below or none.
__device__ __forceinline__ bool already_converged(const CgStatus* st) {
  __shared__ int s_conv;
  if (threadIdx.x == 0) s_conv = st->converged;
  __syncthreads();
  return s_conv != 0;
}

// s + sum_k val[k] * v(col[k]) over the stored off-diagonal slots of row i,
// four independent slots (loads, gathers) in flight per iteration.  The
// value and column streams are read once per launch: nontemporal loads keep
// them from evicting the gathered vector entries from the caches.
template <class T, class V>
__device__ __forceinline__ T sell_row(const Sell<T>& A, size_t i, T s, V v) {
  const size_t base = A.off[i >> 6] + (i & 63);
  const uint32_t wdt = A.width[i >> 6];
  const T* __restrict__ vp = A.val + base;
  uint32_t k = 0;
  if (A.col16 && !A.wide[i >> 6]) {
    // the same slots in the same order, columns from the 16-bit offsets
    const int16_t* __restrict__ cp = A.col16 + base;
    const int ii = static_cast<int>(i);
    for (; k + 4 <= wdt; k += 4) {
      const size_t o = static_cast<size_t>(k) * 64;
      const uint32_t j0 = static_cast<uint32_t>(ii + __builtin_nontemporal_load(cp + o)),
                     j1 = static_cast<uint32_t>(ii + __builtin_nontemporal_load(cp + o + 64)),
                     j2 = static_cast<uint32_t>(ii + __builtin_nontemporal_load(cp + o + 128)),
                     j3 = static_cast<uint32_t>(ii + __builtin_nontemporal_load(cp + o + 192));
      const T a0 = __builtin_nontemporal_load(vp + o), a1 = __builtin_nontemporal_load(vp + o + 64),
              a2 = __builtin_nontemporal_load(vp + o + 128), a3 = __builtin_nontemporal_load(vp + o + 192);
      const T x0 = v(j0), x1 = v(j1), x2 = v(j2), x3 = v(j3);
      s = fma(a0, x0, s);
      s = fma(a1, x1, s);
      s = fma(a2, x2, s);
      s = fma(a3, x3, s);
    }
    for (; k < wdt; k++) {
      const size_t o = static_cast<size_t>(k) * 64;
      s = fma(__builtin_nontemporal_load(vp + o),
              v(static_cast<uint32_t>(ii + __builtin_nontemporal_load(cp + o))), s);
    }
    return s;
  }
  const uint32_t* __restrict__ cp = A.col + base;
  for (; k + 4 <= wdt; k += 4) {
    const size_t o = static_cast<size_t>(k) * 64;
    const uint32_t j0 = __builtin_nontemporal_load(cp + o), j1 = __builtin_nontemporal_load(cp + o + 64),
                   j2 = __builtin_nontemporal_load(cp + o + 128),
                   j3 = __builtin_nontemporal_load(cp + o + 192);
    const T a0 = __builtin_nontemporal_load(vp + o), a1 = __builtin_nontemporal_load(vp + o + 64),
            a2 = __builtin_nontemporal_load(vp + o + 128), a3 = __builtin_nontemporal_load(vp + o + 192);
    const T x0 = v(j0), x1 = v(j1), x2 = v(j2), x3 = v(j3);
    s = fma(a0, x0, s);
    s = fma(a1, x1, s);
    s = fma(a2, x2, s);
    s = fma(a3, x3, s);
  }
  for (; k < wdt; k++) {
    const size_t o = static_cast<size_t>(k) * 64;
    s = fma(__builtin_nontemporal_load(vp + o), v(__builtin_nontemporal_load(cp + o)), s);
  }
  return s;
}

// diag_i x_i + sum_j a_ij x_j over the SELL row (the FP64 CG's product).
template <class V>
__device__ __forceinline__ double row_product(const Sell<double>& A, size_t i, double xi, V v) {
  return sell_row(A, i, A.diag[i] * xi, v);
}

// A caller's system in CSR (device pointers; the diagonal is one of the row's
// entries): the matrix of the allocation-free CG entry point.  Same block
// partition fields as Sell.
struct Csr {
  const uint64_t* __restrict__ rowptr;
  const uint32_t* __restrict__ col;
  const double* __restrict__ val;
  size_t n;
  size_t rpb;
};

// sum_k val[k] x(col[k]) over the row's entries in stored order, two entries
// in flight per iteration.
template <class V>
__device__ __forceinline__ double row_product(const Csr& A, size_t i, double, V v) {
  const uint64_t lo = A.rowptr[i], hi = A.rowptr[i + 1];
  double s = 0;
  uint64_t k = lo;
  for (; k + 2 <= hi; k += 2) {
    const uint32_t j0 = A.col[k], j1 = A.col[k + 1];
    const double a0 = A.val[k], a1 = A.val[k + 1];
    const double x0 = v(j0), x1 = v(j1);
    s = fma(a0, x0, s);
    s = fma(a1, x1, s);
  }
  if (k < hi) s = fma(A.val[k], v(A.col[k]), s);
  return s;
}

// The FP32 operator of a system assembled from particles, recomputed from
// positions instead of read back (matrix-free): a_ij = 3 r_lap / (n0_lap r_ij^3)
// (Computer.hpp:1290), r_ij from fixed-point positions with step h = extent /
// 2^30 (exact integer differences), the stored columns as int16 offsets j - i
// ("narrow" slices) or the SELL u32 columns ("wide" slices, slice-uniform
// branch), loops to each row's own length, and
//   (A p)_i = ghost_i p_i + sum_j a_ij (p_j - p_i),
// i.e. the assembled diagonal -sum_j a_ij plus the SPP ghost share.
struct MatFree {
  const uint64_t* __restrict__ off;
  const uint32_t* __restrict__ col;
  const uint64_t* __restrict__ off16;  // slice offsets of col16 (the SELL ones: same slots)
  const int16_t* __restrict__ col16;
  const uint8_t* __restrict__ wide;
  const uint32_t* __restrict__ len;
  const uint2* __restrict__ pos;
  const float* __restrict__ ghost;
  float c0;  // 3 r_lap / n0_lap / h^3 for r in fixed-point steps h
  size_t n;
  size_t rpb;
};

template <int UNROLL = 4, class V>
__device__ __forceinline__ float op_apply(const Sell<float>& A, size_t i, float pi, V v) {
  return sell_row(A, i, A.diag[i] * pi, v);
}

// a_ij (u_j - u_i) + s.  r^2 in fixed-point units is >= 1 for distinct
// positions (never denormal), so the hardware reciprocal square root is used
// without the denormal scaling rsqrtf adds; A.c0 carries the unit conversion.
template <class J, class V>
__device__ __forceinline__ float mf_term(const MatFree& A, uint2 qi, uint2 q, float vj, float pi, float s) {
  const float fx = static_cast<float>(static_cast<int>(q.x - qi.x));
  const float fz = static_cast<float>(static_cast<int>(q.y - qi.y));
  const float t = __builtin_amdgcn_rsqf(fmaf(fx, fx, fz * fz));
  return fmaf(A.c0 * (t * t * t), vj - pi, s);
}

// UNROLL stored columns (index loads, position and value gathers) in flight.
template <int UNROLL, class J, class V>
__device__ __forceinline__ float mf_sum(const MatFree& A, uint2 qi, uint32_t m, float pi, float s,
                                        J colk, V v) {
  uint32_t k = 0;
  for (; k + UNROLL <= m; k += UNROLL) {
    uint32_t j[UNROLL];
    uint2 q[UNROLL];
    float vj[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; u++) j[u] = colk(k + u);
#pragma unroll
    for (int u = 0; u < UNROLL; u++) q[u] = A.pos[j[u]];
#pragma unroll
    for (int u = 0; u < UNROLL; u++) vj[u] = v(j[u]);
#pragma unroll
    for (int u = 0; u < UNROLL; u++) s = mf_term<J, V>(A, qi, q[u], vj[u], pi, s);
  }
  for (; k < m; k++) {
    const uint32_t j = colk(k);
    s = mf_term<J, V>(A, qi, A.pos[j], v(j), pi, s);
  }
  return s;
}

template <int UNROLL = 4, class V>
__device__ __forceinline__ float op_apply(const MatFree& A, size_t i, float pi, V v) {
  const size_t sl = i >> 6;
  const uint32_t m = A.len[i];
  const uint2 qi = A.pos[i];
  const float s = A.ghost[i] * pi;
  if (A.wide[sl]) {
    const uint32_t* __restrict__ cp = A.col + A.off[sl] + (i & 63);
    return mf_sum<UNROLL>(A, qi, m, pi, s, [&](uint32_t k) { return cp[static_cast<size_t>(k) * 64]; }, v);
  }
  const int16_t* __restrict__ cp = A.col16 + A.off16[sl] + (i & 63);
  const int ii = static_cast<int>(i);
  return mf_sum<UNROLL>(A, qi, m, pi, s,
                        [&](uint32_t k) { return static_cast<uint32_t>(ii + cp[static_cast<size_t>(k) * 64]); },
                        v);
}

// The matrix-free operator reading each row's fixed-point position and the
// vector value it is applied to as one 16-byte record {qx, qz, v, 0}: one
// gather per stored column instead of two (the multilevel solver keeps its
// z in the records; the positions are set once per solve).
struct MatFreeRec {
  const uint64_t* __restrict__ off;
  const uint32_t* __restrict__ col;
  const uint64_t* __restrict__ off16;
  const int16_t* __restrict__ col16;
  const uint8_t* __restrict__ wide;
  const uint32_t* __restrict__ len;
  const uint4* __restrict__ rec;
  const float* __restrict__ ghost;
  float c0;
  size_t n;
  size_t rpb;
};

template <int UNROLL, class J>
__device__ __forceinline__ float rec_sum(const MatFreeRec& A, uint2 qi, uint32_t m, float pi, float s, J colk) {
  uint32_t k = 0;
  for (; k + UNROLL <= m; k += UNROLL) {
    uint32_t j[UNROLL];
    uint4 rj[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; u++) j[u] = colk(k + u);
#pragma unroll
    for (int u = 0; u < UNROLL; u++) rj[u] = A.rec[j[u]];
#pragma unroll
    for (int u = 0; u < UNROLL; u++) {
      const float fx = static_cast<float>(static_cast<int>(rj[u].x - qi.x));
      const float fz = static_cast<float>(static_cast<int>(rj[u].y - qi.y));
      const float t = __builtin_amdgcn_rsqf(fmaf(fx, fx, fz * fz));
      s = fmaf(A.c0 * (t * t * t), __uint_as_float(rj[u].z) - pi, s);
    }
  }
  for (; k < m; k++) {
    const uint4 rj = A.rec[colk(k)];
    const float fx = static_cast<float>(static_cast<int>(rj.x - qi.x));
    const float fz = static_cast<float>(static_cast<int>(rj.y - qi.y));
    const float t = __builtin_amdgcn_rsqf(fmaf(fx, fx, fz * fz));
    s = fmaf(A.c0 * (t * t * t), __uint_as_float(rj.z) - pi, s);
  }
  return s;
}

// (A v)_i with v in the records; vi returns v_i.
template <int UNROLL = 4>
__device__ __forceinline__ float rec_apply(const MatFreeRec& A, size_t i, float& vi) {
  const size_t sl = i >> 6;
  const uint32_t m = A.len[i];
  const uint4 ri = A.rec[i];
  const uint2 qi = make_uint2(ri.x, ri.y);
  const float pi = __uint_as_float(ri.z);
  vi = pi;
  const float s = A.ghost[i] * pi;
  if (A.wide[sl]) {
    const uint32_t* __restrict__ cp = A.col + A.off[sl] + (i & 63);
    return rec_sum<UNROLL>(A, qi, m, pi, s, [&](uint32_t k) { return cp[static_cast<size_t>(k) * 64]; });
  }
  const int16_t* __restrict__ cp = A.col16 + A.off16[sl] + (i & 63);
  const int ii = static_cast<int>(i);
  return rec_sum<UNROLL>(A, qi, m, pi, s,
                         [&](uint32_t k) { return static_cast<uint32_t>(ii + cp[static_cast<size_t>(k) * 64]); });
}

// {qx, qz, 0, 0} for every row (rows outside every cell keep z = 0).
// The records' positions (z = 0), and the FP32 r and d cleared: rows that
// no preconditioner pass visits keep r = d = 0.
__global__ void __launch_bounds__(256) k_rec_init(size_t n, const uint2* __restrict__ qpos, uint4* __restrict__ rec,
                                                  float* __restrict__ r, float* __restrict__ d) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) {
    const uint2 q = qpos[i];
    rec[i] = make_uint4(q.x, q.y, 0u, 0u);
    r[i] = 0.0f;
    d[i] = 0.0f;
  }
}

// r = b - A x, partial sums of r.r and b.b (Computer.hpp:1382-1386).
template <class M>
__device__ __forceinline__ void cg_init_body(const M& A, const double* __restrict__ b, const double* __restrict__ x,
                                             double* __restrict__ r, double* __restrict__ part_rr,
                                             double* __restrict__ part_bb, const Frame& f) {
  __shared__ double lds[NT / 64];
  const Rows rw = block_rows(A.n, A.rpb, f);
  double arr = 0, abb = 0;
  for (size_t i = rw.lo + threadIdx.x; i < rw.hi; i += NT) {
    const double s = row_product(A, i, x[i], [&](uint32_t j) { return x[j]; });
    const double ri = b[i] - s;
    r[i] = ri;
    arr = fma(ri, ri, arr);
    abb = fma(b[i], b[i], abb);
  }
  const double srr = block_sum<NT>(arr, lds);
  const double sbb = block_sum<NT>(abb, lds);
  if (threadIdx.x == 0) {
    part_rr[f.bid] = srr;
    part_bb[f.bid] = sbb;
  }
}
template <class M>
__global__ void __launch_bounds__(NT) k_cg_init(M A, const double* __restrict__ b,
                                                const double* __restrict__ x, double* __restrict__ r,
                                                double* __restrict__ part_rr,
                                                double* __restrict__ part_bb) {
  cg_init_body(A, b, x, r, part_rr, part_bb, whole_grid());
}

struct StopRule {
  int rule;      // mps_stop_rule
  double tol;    // eps (initial-residual rule) or relative tolerance on |b|
};

// One CG iteration, first half (see file comment).  `it` is the 1-based
// iteration number; FIRST marks it == 1, where p_1 = r_0.  The product is
// applied to r, the only vector gathered from other rows, and the search
// direction's product follows by linearity:
//   p_k = r_k + beta p_{k-1},  A p_k = A r_k + beta A p_{k-1},
// so p and q = A p are updated in place for the block's own rows.
template <bool FIRST, class M>
__global__ void __launch_bounds__(NT, 8) k_cg_spmv(M A, const double* __restrict__ r,
                                                   double* __restrict__ p, double* __restrict__ q,
                                                   const double* __restrict__ part_in,
                                                   const double* __restrict__ part_bb, int nparts,
                                                   double* __restrict__ part_out, CgStatus* st,
                                                   StopRule sr, long long it) {
  __shared__ double lds[2 * (NT / 64)];
  if (already_converged(st)) return;
  const int par = static_cast<int>(it & 1);
  // the partials are loaded first and summed after the products of the
  // block's first rows, which do not depend on beta (the loads of both are in
  // flight together); sums in the order of reduce_partials
  const bool want_bb = FIRST && sr.rule == MPS_STOP_RHS_NORM;
  double pa[kCgParts], pb[kCgParts];
#pragma unroll
  for (int u = 0; u < kCgParts; u++) {
    const int k = threadIdx.x + u * NT;
    pa[u] = (k < nparts) ? part_in[k] : 0.0;
    pb[u] = (want_bb && k < nparts) ? part_bb[k] : 0.0;
  }
  const Rows rw = block_rows(A.n, A.rpb);
  double rv[kCgRows], sv[kCgRows];
#pragma unroll
  for (int u = 0; u < kCgRows; u++) {
    const size_t i = rw.lo + threadIdx.x + static_cast<size_t>(u) * NT;
    rv[u] = sv[u] = 0.0;
    if (i < rw.hi) {
      rv[u] = r[i];
      sv[u] = row_product(A, i, rv[u], [&](uint32_t j) { return r[j]; });
    }
  }
  double sa = 0, sb = 0;
#pragma unroll
  for (int u = 0; u < kCgParts; u++) {
    sa += pa[u];
    sb += pb[u];
  }
  for (int k = threadIdx.x + kCgParts * NT; k < nparts; k += NT) {
    sa += part_in[k];
    if (want_bb) sb += part_bb[k];
  }
  double rr, bb = 0;
  block_sum2<NT>(sa, sb, lds, rr, bb);
  double beta = 0, tol;
  bool conv;
  if (FIRST) {
    if (sr.rule == MPS_STOP_RHS_NORM) {
      tol = (sr.tol * sr.tol) * bb;
      conv = (rr <= tol);
    } else {
      tol = rr * sr.tol * sr.tol;  // residual0, Computer.hpp:1386
      conv = (tol == 0);
    }
  } else {
    tol = st->tol;
    conv = (sr.rule == MPS_STOP_RHS_NORM) ? (rr <= tol) : (rr < tol);  // :1408
  }
  if (lead()) {
    st->rr[par] = rr;
    if (FIRST) {
      st->tol = tol;
      st->rr0 = rr;
      st->b2 = bb;
    }
    if (conv) {
      st->converged = 1;
      st->iterations = it - 1;
    }
  }
  if (conv) return;
  if (!FIRST) beta = rr / st->rr[par ^ 1];
  double acc = 0;
  auto finish = [&](size_t i, double ri, double s) {
    const double pi = FIRST ? ri : fma(beta, p[i], ri);
    const double qi = FIRST ? s : fma(beta, q[i], s);
    p[i] = pi;
    q[i] = qi;
    acc = fma(pi, qi, acc);
  };
#pragma unroll
  for (int u = 0; u < kCgRows; u++) {
    const size_t i = rw.lo + threadIdx.x + static_cast<size_t>(u) * NT;
    if (i < rw.hi) finish(i, rv[u], sv[u]);
  }
  for (size_t i = rw.lo + threadIdx.x + static_cast<size_t>(kCgRows) * NT; i < rw.hi; i += NT) {
    const double ri = r[i];
    finish(i, ri, row_product(A, i, ri, [&](uint32_t j) { return r[j]; }));
  }
  const double sp = block_sum<NT>(acc, lds);
  if (threadIdx.x == 0) part_out[blockIdx.x] = sp;
}

// One CG iteration, second half (Computer.hpp:1398-1404).
__global__ void __launch_bounds__(NT) k_cg_update(size_t n, size_t rpb, double* __restrict__ x,
                                                  double* __restrict__ r, const double* __restrict__ p,
                                                  const double* __restrict__ q,
                                                  const double* __restrict__ part_in, int nparts,
                                                  double* __restrict__ part_out, CgStatus* st,
                                                  long long it) {
  __shared__ double lds[NT / 64];
  if (already_converged(st)) return;
  const int par = static_cast<int>(it & 1);
  // the partials and the block's first rows are loaded together; sums in the
  // order of reduce_partials
  double pa[kCgParts];
#pragma unroll
  for (int u = 0; u < kCgParts; u++) {
    const int k = threadIdx.x + u * NT;
    pa[u] = (k < nparts) ? part_in[k] : 0.0;
  }
  const Rows rw = block_rows(n, rpb);
  double xv[kCgRows], pv[kCgRows], qv[kCgRows], rv[kCgRows];
#pragma unroll
  for (int u = 0; u < kCgRows; u++) {
    const size_t i = rw.lo + threadIdx.x + static_cast<size_t>(u) * NT;
    const bool in = i < rw.hi;
    xv[u] = in ? x[i] : 0.0;
    pv[u] = in ? p[i] : 0.0;
    qv[u] = in ? q[i] : 0.0;
    rv[u] = in ? r[i] : 0.0;
  }
  double sa = 0;
#pragma unroll
  for (int u = 0; u < kCgParts; u++) sa += pa[u];
  for (int k = threadIdx.x + kCgParts * NT; k < nparts; k += NT) sa += part_in[k];
  const double pq = block_sum<NT>(sa, lds);
  const double alpha = st->rr[par] / pq;
  double acc = 0;
#pragma unroll
  for (int u = 0; u < kCgRows; u++) {
    const size_t i = rw.lo + threadIdx.x + static_cast<size_t>(u) * NT;
    if (i < rw.hi) {
      x[i] = fma(alpha, pv[u], xv[u]);
      const double ri = fma(-alpha, qv[u], rv[u]);
      r[i] = ri;
      acc = fma(ri, ri, acc);
    }
  }
  for (size_t i = rw.lo + threadIdx.x + static_cast<size_t>(kCgRows) * NT; i < rw.hi; i += NT) {
    x[i] = fma(alpha, p[i], x[i]);
    const double ri = fma(-alpha, q[i], r[i]);
    r[i] = ri;
    acc = fma(ri, ri, acc);
  }
  const double s = block_sum<NT>(acc, lds);
  if (threadIdx.x == 0) part_out[blockIdx.x] = s;
  if (lead()) st->iterations = it;
}

// Stop test after the last permitted iteration.
__global__ void __launch_bounds__(NT) k_cg_check(const double* __restrict__ part_in, int nparts,
                                                 CgStatus* st, StopRule sr, long long it) {
  __shared__ double lds[NT / 64];
  if (already_converged(st)) return;
  const double rr = reduce_partials<NT>(part_in, nparts, lds);
  const bool conv = (sr.rule == MPS_STOP_RHS_NORM) ? (rr <= st->tol) : (rr < st->tol);
  if (threadIdx.x == 0) {
    st->rr[it & 1] = rr;
    if (conv) {
      st->converged = 1;
      st->iterations = it - 1;
    }
  }
}

// Final reduction of two partial arrays into out[0], out[1] (single block).
__global__ void __launch_bounds__(NT) k_reduce2(const double* __restrict__ a,
                                                const double* __restrict__ b, int nparts,
                                                double* __restrict__ out) {
  __shared__ double lds[NT / 64];
  const double sa = reduce_partials<NT>(a, nparts, lds);
  const double sb = reduce_partials<NT>(b, nparts, lds);
  if (threadIdx.x == 0) {
    out[0] = sa;
    out[1] = sb;
  }
}

// ---------------- mixed precision: FP32 CG inside FP64 refinement ----------------
//
// Each refinement sweep solves A d = r / |r| in FP32 with the Chronopoulos-Gear
// form of preconditioned CG (Jacobi or none): one global reduction phase per
// iteration, so one kernel per iteration.  Iteration i, given the reduced
// gamma_i = (r_i, u_i), delta_i = (w_i, u_i) with u = M r, w = A u:
//   beta_i  = gamma_i / gamma_{i-1},  alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1})
//   s_i = w_i + beta_i s_{i-1},  p_i = u_i + beta_i p_{i-1},  d += alpha_i p_i,
//   r_{i+1} = r_i - alpha_i s_i,  u_{i+1} = M r_{i+1},  w_{i+1} = A u_{i+1},
// then the partial sums of gamma_{i+1}, delta_{i+1} and r_{i+1}.r_{i+1}.  The
// SpMV gathers u_{i+1,j} of other blocks' rows by recomputing them from the
// previous iteration's (r, w, s, M) of row j, packed in one float4 that is
// double-buffered by iteration parity; p and d are only touched by their own
// row.  Dots accumulate in FP64.

struct Dots3 {
  double a, b, c;
};
// Deterministic block sums of three doubles per thread (one LDS round).
__device__ __forceinline__ Dots3 block_sum3(double a, double b, double c, double* lds) {
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    lds[wv] = a;
    lds[NT / 64 + wv] = b;
    lds[2 * (NT / 64) + wv] = c;
  }
  __syncthreads();
  Dots3 s{lds[0], lds[NT / 64], lds[2 * (NT / 64)]};
#pragma unroll
  for (int k = 1; k < NT / 64; k++) {
    s.a += lds[k];
    s.b += lds[NT / 64 + k];
    s.c += lds[2 * (NT / 64) + k];
  }
  __syncthreads();
  return s;
}
__device__ __forceinline__ Dots3 reduce_partials3(const double* __restrict__ part, int count,
                                                  double* lds) {
  double a = 0, b = 0, c = 0;
  for (int k = threadIdx.x; k < count; k += NT) {
    a += part[k];
    b += part[kRedBlocks + k];
    c += part[2 * kRedBlocks + k];
  }
  return block_sum3(a, b, c, lds);
}
__device__ __forceinline__ void write_partials3(double* __restrict__ part, Dots3 s) {
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s.a;
    part[kRedBlocks + blockIdx.x] = s.b;
    part[2 * kRedBlocks + blockIdx.x] = s.c;
  }
}

// r_0 = r64 / |r64|, u_0 = M r_0, w_0 = A u_0; G_out = (r_0, w_0, 0, M), p = d = 0,
// partials of gamma_0, delta_0, r_0.r_0.  Neighbour u_0,j straight from r64.
template <class Op>
__global__ void __launch_bounds__(NT, 8) k_cgf_init(Op A, const double* __restrict__ r64, double inv_norm,
                                                 const float* __restrict__ dinv, float4* __restrict__ gout,
                                                 float* __restrict__ p, float* __restrict__ d,
                                                 double* __restrict__ part_out) {
  __shared__ double lds[3 * (NT / 64)];
  const Rows rw = block_rows(A.n, A.rpb);
  double ag = 0, ad = 0, ar = 0;
  for (size_t i = rw.lo + threadIdx.x; i < rw.hi; i += NT) {
    const float dv = dinv ? dinv[i] : 1.0f;
    const float ri = static_cast<float>(r64[i] * inv_norm);
    const float ui = dv * ri;
    const float wi = op_apply(A, i, ui, [&](uint32_t j) {
      return (dinv ? dinv[j] : 1.0f) * static_cast<float>(r64[j] * inv_norm);
    });
    gout[i] = make_float4(ri, wi, 0.0f, dv);
    p[i] = 0.0f;
    d[i] = 0.0f;
    ag = fma(static_cast<double>(ri), static_cast<double>(ui), ag);
    ad = fma(static_cast<double>(wi), static_cast<double>(ui), ad);
    ar = fma(static_cast<double>(ri), static_cast<double>(ri), ar);
  }
  write_partials3(part_out, block_sum3(ag, ad, ar, lds));
}

// Columns in flight per lane in the iteration kernel (each gathers a float4
// and a position): 3 keeps it within 64 VGPRs, i.e. 8 waves per SIMD.
constexpr int kIterUnroll = 3;

// Iteration i (0-based): stop test on r_i.r_i < tol_rel^2 r_0.r_0, then the
// updates above.  `it` = i + 1 is the iteration number the launch completes.
template <class Op>
__global__ void __launch_bounds__(NT, 8) k_cgf_iter(Op A, const float4* __restrict__ gin,
                                                 float4* __restrict__ gout, float* __restrict__ p,
                                                 float* __restrict__ d, const double* __restrict__ part_in,
                                                 int nparts, double* __restrict__ part_out, CgStatus* st,
                                                 double tol_rel, long long it) {
  __shared__ double lds[3 * (NT / 64)];
  if (already_converged(st)) return;
  const Dots3 g = reduce_partials3(part_in, nparts, lds);  // gamma_i, delta_i, rr_i
  const int par = static_cast<int>(it & 1);
  const bool first = (it == 1);
  double tol, beta = 0.0, alpha;
  bool conv;
  if (first) {
    tol = g.c * tol_rel * tol_rel;
    conv = (tol == 0);
    alpha = g.a / g.b;
  } else {
    tol = st->tol;
    conv = g.c < tol;
    beta = g.a / st->gamma[par ^ 1];
    alpha = g.a / (g.b - beta * g.a / st->alpha[par ^ 1]);
  }
  if (lead()) {
    st->rr[par] = g.c;
    st->gamma[par] = g.a;
    st->alpha[par] = alpha;
    if (first) {
      st->tol = tol;
      st->rr0 = g.c;
    }
    if (conv) {
      st->converged = 1;
      st->iterations = it - 1;
    }
  }
  if (conv) return;
  const float fa = static_cast<float>(alpha), fb = static_cast<float>(beta);
  const Rows rw = block_rows(A.n, A.rpb);
  double ag = 0, ad = 0, ar = 0;
  for (size_t i = rw.lo + threadIdx.x; i < rw.hi; i += NT) {
    const float4 h = gin[i];
    const float si = fmaf(fb, h.z, h.y);
    const float pi = fmaf(fb, p[i], h.w * h.x);
    d[i] = fmaf(fa, pi, d[i]);
    p[i] = pi;
    const float ri = fmaf(-fa, si, h.x);
    const float ui = h.w * ri;
    const float wi = op_apply<kIterUnroll>(A, i, ui, [&](uint32_t j) {
      const float4 q = gin[j];
      return q.w * fmaf(-fa, fmaf(fb, q.z, q.y), q.x);
    });
    gout[i] = make_float4(ri, wi, si, h.w);
    ag = fma(static_cast<double>(ri), static_cast<double>(ui), ag);
    ad = fma(static_cast<double>(wi), static_cast<double>(ui), ad);
    ar = fma(static_cast<double>(ri), static_cast<double>(ri), ar);
  }
  write_partials3(part_out, block_sum3(ag, ad, ar, lds));
  if (lead()) st->iterations = it;
}

// ---------------- multilevel aggregation preconditioner ----------------
//
// M r = D^-1 r + sum_{l=0..2} P_l (P_l^T A P_l)^-diag P_l^T r  (additive):
// P_l is the piecewise-constant prolongation from the aggregates of level l,
// the neighbour-search cells (l = 0) and 2x2 and 4x4 blocks of cells, and
// each aggregate k is scaled by 1 / (1^T A_kk 1), the sum of the matrix block
// of its members.  Dummy and disabled rows belong to no aggregate.  One
// wavefront owns one 4x4 block of cells: lane l takes cell l % 16 and every
// fourth particle of it (from l / 16), and the cell, 2x2 and 4x4 sums are
// butterflies over fixed lane masks, so every sum has a fixed order.
constexpr int NTP = 256;  // threads per block of the aggregate kernels

struct CellGrid {
  const uint32_t* __restrict__ start;   // first slot of each cell in `sorted`
  const uint32_t* __restrict__ count;   // particles per cell
  const uint32_t* __restrict__ sorted;  // particle ids grouped by cell, ascending in a cell
  const uint32_t* __restrict__ key;     // cell of each particle
  const uint8_t* __restrict__ type;
  uint32_t gnx, gnz;                    // cells
  uint32_t g1z;                         // 2x2 blocks along z
  uint32_t bnx, bnz;                    // 4x4 blocks
  size_t off1, off2;                    // level offsets in the aggregate arrays
};

struct BlockLane {
  bool valid;  // the lane's cell lies in the grid
  uint32_t cell, a1, a2, sub;
};
__device__ __forceinline__ BlockLane block_lane(const CellGrid& G, uint32_t b) {
  const uint32_t lane = threadIdx.x & 63, u = lane & 15;
  const uint32_t kx = b / G.bnz, kz = b % G.bnz;
  const uint32_t cx = 4 * kx + (u >> 2), cz = 4 * kz + (u & 3);
  BlockLane L;
  L.valid = cx < G.gnx && cz < G.gnz;
  L.cell = cx * G.gnz + cz;
  L.a1 = (cx >> 1) * G.g1z + (cz >> 1);
  L.a2 = b;
  L.sub = lane >> 4;
  return L;
}
__device__ __forceinline__ bool in_agg(uint8_t t) { return t != kDummy && t != kDisabled; }

// Sums of v over the lanes of one cell, of one 2x2 block and of the 4x4 block.
template <class T>
__device__ __forceinline__ void agg_sums(T v, T& y0, T& y1, T& y2) {
  y0 = v + __shfl_xor(v, 16, 64);
  y0 = y0 + __shfl_xor(y0, 32, 64);
  y1 = y0 + __shfl_xor(y0, 1, 64);
  y1 = y1 + __shfl_xor(y1, 4, 64);
  y2 = y1 + __shfl_xor(y1, 2, 64);
  y2 = y2 + __shfl_xor(y2, 8, 64);
}

// Per row i (dummy and disabled rows: 0), for the aggregates of levels 0..2
// containing i: S_l(i) = A_ii + sum_{j in agg_l(i)} A_ij from the FP64 system.
// Every term of sum_{i in k} S_l(i) = 1^T A_kk 1 is <= 0 (A_ii is minus the
// sum of the row's off-diagonals and the ghost), so FP32 storage loses nothing
// to cancellation.
__global__ void __launch_bounds__(NT, 8) k_ml_rowsum(Sell<double> A, const uint32_t* __restrict__ row_len,
                                                     const uint32_t* __restrict__ key,
                                                     const uint8_t* __restrict__ type, uint32_t gnz,
                                                     float* __restrict__ s0, float* __restrict__ s1,
                                                     float* __restrict__ s2) {
  const Rows rw = block_rows(A.n, A.rpb);
  for (size_t i = rw.lo + threadIdx.x; i < rw.hi; i += NT) {
    double a0 = 0, a1 = 0, a2 = 0;
    const uint32_t ki = key[i];
    if (in_agg(type[i]) && ki != kNoCell) {
      const uint32_t cx = ki / gnz, cz = ki % gnz;
      a0 = a1 = a2 = A.diag[i];
      const size_t base = A.off[i >> 6] + (i & 63);
      const uint32_t m = row_len[i];
      for (uint32_t q = 0; q < m; q++) {
        const size_t o = base + static_cast<size_t>(q) * 64;
        const uint32_t kj = key[sell_col(A, i, o)];
        if (kj == kNoCell) continue;
        const uint32_t jx = kj / gnz, jz = kj % gnz;
        const double a = A.val[o];
        if ((jx >> 2) == (cx >> 2) && (jz >> 2) == (cz >> 2)) {
          a2 += a;
          if ((jx >> 1) == (cx >> 1) && (jz >> 1) == (cz >> 1)) {
            a1 += a;
            if (jx == cx && jz == cz) a0 += a;
          }
        }
      }
    }
    s0[i] = static_cast<float>(a0);
    s1[i] = static_cast<float>(a1);
    s2[i] = static_cast<float>(a2);
  }
}

// 1 / (1^T A_kk 1) for every aggregate (0 where the sum is not negative: no
// correction): the row sums of k_ml_rowsum added over the members, one
// wavefront per 4x4 block of cells.
__global__ void __launch_bounds__(NTP) k_ml_setup(CellGrid G, const float* __restrict__ s0,
                                                  const float* __restrict__ s1, const float* __restrict__ s2,
                                                  float* __restrict__ inv) {
  const uint32_t b = blockIdx.x * (NTP / 64) + (threadIdx.x >> 6);
  if (b >= G.bnx * G.bnz) return;  // wave-uniform
  const BlockLane L = block_lane(G, b);
  const uint32_t cnt = L.valid ? G.count[L.cell] : 0u, st0 = L.valid ? G.start[L.cell] : 0u;
  float t0 = 0, t1 = 0, t2 = 0;
  for (uint32_t k = L.sub; k < cnt; k += 4) {
    const uint32_t i = G.sorted[st0 + k];
    t0 += s0[i];
    t1 += s1[i];
    t2 += s2[i];
  }
  float y0, y1, y2, t;
  agg_sums(t0, y0, t, t);
  agg_sums(t1, t, y1, t);
  agg_sums(t2, t, t, y2);
  const uint32_t lane = threadIdx.x & 63;
  if (L.valid && L.sub == 0) {
    inv[L.cell] = (y0 < 0) ? 1.0f / y0 : 0.0f;
    if ((lane & 5) == 0) inv[G.off1 + L.a1] = (y1 < 0) ? 1.0f / y1 : 0.0f;
    if ((lane & 15) == 0) inv[G.off2 + L.a2] = (y2 < 0) ? 1.0f / y2 : 0.0f;
  }
}

// Each inner iteration `it` >= 1 is two kernels; every block reduces the
// partial sums of the kernel before it itself, in one fixed order, so all
// blocks derive the same scalars without a one-block launch in between:
//   k_ml_spmv(it):  rho_(it-1) and r.r from the partials of k_ml_prec(it-1)
//                   and the stop test (block 0 records them); d += alpha_(it-1)
//                   p_(it-1); unless stopped, beta = rho_(it-1) / rho_(it-2),
//                   p = z + beta p, q = A z + beta q (A p by linearity; z is
//                   the only gathered vector), partial sums of p.q;
//   k_ml_prec(it):  alpha = rho_(it-1) / p.q from the partials of
//                   k_ml_spmv(it) (block 0 records it); r -= alpha q; z = M r;
//                   partial sums of rho_it = r.z and of r.r.
// Iteration 0 is k_ml_prec alone: r = r64 / |r64|, d = 0, z = M r.  The stop
// at iteration k is found by k_ml_spmv(k + 1), which still adds the last
// alpha_k p_k to d; a sweep cut short (iteration limit, stagnation) ends with
// k_ml_finish instead.  (A grid-wide sum finished by the last block to arrive
// would need a device-scope release per block, an L2 write-back on this
// multi-XCD part, far more than the redundant sums.)  k_ml_prec visits the
// particles cell by cell, one wavefront per 4x4 block of cells (every
// particle of the system that is in a cell exactly once; rows outside every
// cell are identity rows whose r, d, z stay 0), so the residual update, the
// aggregate sums and z share one pass.
constexpr int kPerLane = 7;  // particles of a cell held in registers per lane (4 lanes per cell)
constexpr int kPqPerThread = kRedBlocks / NTP;  // p.q partials loaded per thread of k_ml_prec

// Block-uniform: the inner solve stopped at an iteration below k.
__device__ __forceinline__ bool stopped_before(const CgStatus* st, long long k) {
  __shared__ long long s_mark;
  if (threadIdx.x == 0) s_mark = st->stop_mark;
  __syncthreads();
  return s_mark != 0 && s_mark - 1 < k;
}

// 1 for every 4x4 block of cells that holds a particle (one thread per block).
__global__ void __launch_bounds__(256) k_ml_active(CellGrid G, uint32_t* __restrict__ flag) {
  const uint32_t b = blockIdx.x * 256u + threadIdx.x;
  if (b >= G.bnx * G.bnz) return;
  const uint32_t kx = b / G.bnz, kz = b % G.bnz;
  uint32_t any = 0;
  for (uint32_t u = 0; u < 16; u++) {
    const uint32_t cx = 4 * kx + (u >> 2), cz = 4 * kz + (u & 3);
    if (cx < G.gnx && cz < G.gnz) any |= G.count[cx * G.gnz + cz];
  }
  flag[b] = any != 0u;
}

// The flagged blocks in ascending order (off: exclusive scan of the flags).
__global__ void __launch_bounds__(256) k_ml_compact(uint32_t nblk, const uint32_t* __restrict__ flag,
                                                    const uint32_t* __restrict__ off,
                                                    uint32_t* __restrict__ list) {
  const uint32_t b = blockIdx.x * 256u + threadIdx.x;
  if (b < nblk && flag[b]) list[off[b]] = b;
}

// The preconditioner's per-solve constants in the order its wavefronts use
// them, so that an iteration reads them with coalesced loads instead of the
// chain list -> cell table -> sorted ids -> per-particle data: for listed
// block w, slot u < kPerLane and lane l, at (w * kPerLane + u) * 64 + l,
//   pid: the particle id (top bit set: not in an aggregate, a dummy or
//        disabled row; kNoSlot: empty slot), dv: its Jacobi factor 1 / A_ii;
// per (w, l): the lane's cell, 2x2 and 4x4 aggregate scales and its cell's
// particle count (float4; counts above 4 kPerLane take the cell-table path).
constexpr uint32_t kNoSlot = 0xffffffffu;
constexpr uint32_t kOutAgg = 0x80000000u;

// inv == nullptr (the cell-grid multigrid): no aggregate scales, and rows
// without stored off-diagonals (row_len 0: the fixed surface rows) join no
// aggregate either.
__global__ void __launch_bounds__(NTP) k_ml_pack(CellGrid G, const uint32_t* __restrict__ list, uint32_t nact,
                                                 const float* __restrict__ inv, const float* __restrict__ dinv,
                                                 uint32_t* __restrict__ pid, float* __restrict__ dv,
                                                 float4* __restrict__ sc, const uint32_t* __restrict__ row_len) {
  const uint32_t w = blockIdx.x * (NTP / 64) + (threadIdx.x >> 6);
  if (w >= nact) return;  // wave-uniform
  const uint32_t lane = threadIdx.x & 63;
  const BlockLane L = block_lane(G, list[w]);
  const uint32_t cnt = L.valid ? G.count[L.cell] : 0u, st0 = L.valid ? G.start[L.cell] : 0u;
  const bool scales = inv != nullptr && L.valid;
  sc[static_cast<size_t>(w) * 64 + lane] =
      make_float4(scales ? inv[L.cell] : 0.0f, scales ? inv[G.off1 + L.a1] : 0.0f,
                  scales ? inv[G.off2 + L.a2] : 0.0f, static_cast<float>(cnt));
  for (int u = 0; u < kPerLane; u++) {
    const uint32_t k = L.sub + 4u * u;
    const size_t o = (static_cast<size_t>(w) * kPerLane + u) * 64 + lane;
    if (k < cnt) {
      const uint32_t i = G.sorted[st0 + k];
      const bool agg = in_agg(G.type[i]) && (row_len == nullptr || row_len[i] > 0);
      pid[o] = i | (agg ? 0u : kOutAgg);
      dv[o] = dinv[i];
    } else {
      pid[o] = kNoSlot;
      dv[o] = 0.0f;
    }
  }
}

// ---------------- cell-grid aggregation multigrid (MPS_PRECOND_AMG) ----------------
//
// The preconditioner M r = omega_r D^-1 r + P_s V(P_s^T r): aggregates are
// the neighbour-search cells (P piecewise constant over a cell's aggregated
// rows), V is one V-cycle on the cells' Galerkin operator P^T A P, and P_s =
// (I - omega_s D^-1 A) P is P smoothed by one Jacobi step.  Row i of P_s has
// 9 entries, one per cell around i's own (A P's row is the row's stored
// values summed by the direction of the column's cell, which the fill
// computes), so P_s^T r and P_s e are 9-term sums per particle; the coarse
// operator stays P^T A P (a non-Galerkin pair: the Galerkin P_s^T A P would be
// a 25-point stencil).  The
// cell side NL exceeds the Laplacian radius, so a particle couples only to
// its own and the 8 adjacent cells: P^T A P is a 9-point stencil on the cell
// grid, and so is every coarser level (2x2 blocks of the level below,
// Galerkin sums of the stencil).  Each level is a dense grid (empty points:
// zero stencil, zero D^-1).  V at level l: x = omega D^-1 r, residual,
// restriction (sums over 2x2 blocks), the level below, prolongation
// (piecewise constant), post-smoothing x += omega D^-1 (r - S x); the
// coarsest level (at most kAmgDense points) is solved with its inverse.
// The coarse levels whose points fit kAmgLdsPts run inside one workgroup, in LDS.
//
// The weights: the cell levels' stencils are weakly diagonally dominant with
// non-positive off-diagonals (Galerkin sums over aggregates of the HL
// Laplacian's rows, whose off-diagonals are negative and whose row sums are
// not), so lambda_max(D^-1 S) <= 2 (Gershgorin; about 1.8-1.97 on the 1M
// dam break) and Jacobi with omega <= 1 is a convergent smoother: the
// symmetric V-cycle, and M, stay positive definite.  omega = 1 on the cells
// and omega_r = 0.5 on the particles take the PCG on the developed 1M dam
// break from 63 to 54 iterations (FP64, 0.7 on both: 63; 0.7 / 0.9 / 1.0 on
// the cells with 0.5 on the particles: 56 / 55 / 54; early flow 37 -> 35).
// The piecewise-constant prolongation of aggregation underestimates the
// coarse correction: weighting it by 1.1 (below 2, so the V-cycle stays
// positive definite) takes the developed flow to 49 and the early flow to 34
// iterations.  The smoothed P_s (omega_s = 1, corrections unweighted)
// takes them to 36 and 22 (omega_s = 0.67: 37 and 23; P_s with the exact
// coarse solve instead of V: 22 on the developed flow, P: 33); its FP16
// coefficients cost no iterations, and since restriction and prolongation
// use the same rounded values M stays symmetric.
constexpr int kAmgLdsPts = 2624;  // points of the levels the single-workgroup kernel holds in LDS (133 KB + the inverse)
constexpr int kAmgDense = 64;
constexpr int kAmgBlockThreads = 1024;
constexpr float kAmgOmega = 1.0f;      // the cell levels' Jacobi smoother
constexpr float kAmgOmegaRows = 0.5f;  // the particle-level (additive) Jacobi term
constexpr float kAmgCorr = 1.0f;       // the cell levels' coarse-correction weight
constexpr double kAmgOmegaP = 1.0;     // the smoothing of the prolongation
// P_s's rows per packed slot o: coefficients 0..7 as 8 FP16 in q9[o] (16
// bytes), the ninth as one FP16 in a plane after them, q9c[o] (18 bytes a slot)

struct AmgLevel {
  uint32_t nx, nz;
  uint32_t off;  // first point of the level in the per-point arrays
};
struct AmgGrid {
  int levels;
  AmgLevel L[kAmgMaxLevels];
  const float* S;     // 9 stencil coefficients of level l at S + 9 off_l + d npts_l + p (d = 3 (dx + 1) + dz + 1)
  const float* dinv;  // 1 / S_center (0 on empty points)
  float* r;           // right-hand side of every level (level 0: the cells' residual sums)
  float* x;           // pre-smoothed iterate
  float* e;           // result (level 0: the correction per cell)
  const float* inv;   // inverse of the coarsest level (npts x npts, row-major)
  float omega;
  float corr;  // weight of the coarse correction: x = x_pre + corr e(parent)
  // the inner solve whose iteration `it` the cycle belongs to (nullptr: none);
  // a cycle of an iteration past the solve's stop does nothing
  const CgStatus* st = nullptr;
  long long it = 0;
};
__device__ __forceinline__ uint32_t amg_npts(const AmgLevel& L) { return L.nx * L.nz; }

// Block-uniform: the cycle belongs to an iteration launched past the stop
// (every thread of the block calls it, before any divergent exit).
__device__ __forceinline__ bool amg_skip(const AmgGrid& g) { return g.st && stopped_before(g.st, g.it); }

// The per-point vectors of a cycle: r, x, e and 1 / S_center of point k of
// any level at v[k - base] (global memory: base 0; the single-workgroup
// kernel's LDS copy: base = the first point it holds).
struct AmgVec {
  float *r, *x, *e;
  const float* dinv;
  uint32_t base;
};

// Level l -> l + 1: one thread per point of level l + 1; its 2x2 children p
// get x_p = omega r_p / S_pp and their residuals r_p - (S x)_p, summed into
// the parent's right-hand side (x of a neighbour is computed from its r).
template <class SV>
__device__ __forceinline__ void amg_down_point(const AmgGrid& g, const AmgVec& v, const SV& sten, int l, uint32_t P) {
  const AmgLevel F = g.L[l], C = g.L[l + 1];
  const uint32_t X = P / C.nz, Z = P % C.nz;
  const uint32_t fo = F.off - v.base;
  const float* __restrict__ vr = v.r;
  const float* __restrict__ vd = v.dinv;
  // every load first: r and 1 / S_pp on the 4x4 patch around the children,
  // the children's stencils
  float rr[4][4], dd[4][4], st[2][2][9];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int ix = 2 * static_cast<int>(X) - 1 + a, iz = 2 * static_cast<int>(Z) - 1 + b;
      const bool in = ix >= 0 && iz >= 0 && ix < static_cast<int>(F.nx) && iz < static_cast<int>(F.nz);
      const uint32_t q = fo + static_cast<uint32_t>(in ? ix : 0) * F.nz + static_cast<uint32_t>(in ? iz : 0);
      rr[a][b] = in ? vr[q] : 0.0f;
      dd[a][b] = in ? vd[q] : 0.0f;
    }
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int w = 0; w < 2; w++) {
      const uint32_t ix = 2 * X + u, iz = 2 * Z + w;
      const bool in = ix < F.nx && iz < F.nz;
      const uint32_t p = in ? ix * F.nz + iz : 0u;
#pragma unroll
      for (int d = 0; d < 9; d++) st[u][w][d] = in ? sten(l, d, p) : 0.0f;
    }
  float rc = 0.0f;
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int w = 0; w < 2; w++) {
      const uint32_t ix = 2 * X + u, iz = 2 * Z + w;
      if (ix >= F.nx || iz >= F.nz) continue;
      float sum = 0.0f;
#pragma unroll
      for (int dx = -1; dx <= 1; dx++)
#pragma unroll
        for (int dz = -1; dz <= 1; dz++)
          sum = fmaf(st[u][w][3 * (dx + 1) + dz + 1],
                     g.omega * dd[1 + u + dx][1 + w + dz] * rr[1 + u + dx][1 + w + dz], sum);
      v.x[fo + ix * F.nz + iz] = g.omega * dd[1 + u][1 + w] * rr[1 + u][1 + w];
      rc += rr[1 + u][1 + w] - sum;
    }
  v.r[C.off - v.base + P] = rc;
}

// Level l + 1 -> l: one thread per point p of level l: x = x_pre + corr e(parent),
// then e_p = x_p + omega (r_p - (S x)_p) / S_pp.
template <class SV>
__device__ __forceinline__ float amg_up_value(const AmgGrid& g, const AmgVec& v, const SV& sten, int l, uint32_t p) {
  const AmgLevel F = g.L[l], C = g.L[l + 1];
  const uint32_t ix = p / F.nz, iz = p % F.nz;
  const uint32_t fo = F.off - v.base, co = C.off - v.base;
  const float di = v.dinv[fo + p];
  if (di == 0.0f) return 0.0f;  // empty point
  float xq[9], sv[9];
#pragma unroll
  for (int dx = -1; dx <= 1; dx++)
#pragma unroll
    for (int dz = -1; dz <= 1; dz++) {
      const int k = 3 * (dx + 1) + dz + 1;
      const int jx = static_cast<int>(ix) + dx, jz = static_cast<int>(iz) + dz;
      const bool in = jx >= 0 && jz >= 0 && jx < static_cast<int>(F.nx) && jz < static_cast<int>(F.nz);
      const uint32_t ux = static_cast<uint32_t>(in ? jx : 0), uz = static_cast<uint32_t>(in ? jz : 0);
      xq[k] = in ? fmaf(g.corr, v.e[co + (ux >> 1) * C.nz + (uz >> 1)], v.x[fo + ux * F.nz + uz]) : 0.0f;
      sv[k] = in ? sten(l, k, p) : 0.0f;
    }
  float sum = 0.0f;
#pragma unroll
  for (int k = 0; k < 9; k++) sum = fmaf(sv[k], xq[k], sum);
  return fmaf(g.omega * di, v.r[fo + p] - sum, xq[4]);
}

// Level l + 1 -> l: one thread per point p of level l: x = x_pre + e(parent),
// then e_p = x_p + omega (r_p - (S x)_p) / S_pp (0 on empty points).
template <class SV>
__device__ __forceinline__ void amg_up_point(const AmgGrid& g, const AmgVec& v, const SV& sten, int l, uint32_t p) {
  v.e[g.L[l].off - v.base + p] = amg_up_value(g, v, sten, l, p);
}

// Stencil coefficient d of point p of level l, from global memory.
struct AmgStenGlobal {
  const AmgGrid* g;
  __device__ __forceinline__ float operator()(int l, int d, uint32_t p) const {
    const AmgLevel L = g->L[l];
    return g->S[9ull * L.off + static_cast<size_t>(d) * amg_npts(L) + p];
  }
};

__device__ __forceinline__ AmgVec amg_global(const AmgGrid& g) { return AmgVec{g.r, g.x, g.e, g.dinv, 0u}; }

__global__ void __launch_bounds__(256) k_amg_down(AmgGrid g, int l) {
  if (amg_skip(g)) return;
  const uint32_t P = blockIdx.x * 256u + threadIdx.x;
  if (P < amg_npts(g.L[l + 1])) amg_down_point(g, amg_global(g), AmgStenGlobal{&g}, l, P);
}

__global__ void __launch_bounds__(256) k_amg_up(AmgGrid g, int l) {
  if (amg_skip(g)) return;
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p < amg_npts(g.L[l])) amg_up_point(g, amg_global(g), AmgStenGlobal{&g}, l, p);
}

// NLEV levels of the V-cycle in one launch each way (the level kernels
// above cost a launch and a dependent round trip to memory apiece, ~5 us,
// whatever their size).  Workgroup t owns tile t (T x T points) of the
// coarsest of the NLEV + 1 levels l1 .. l1 + NLEV and the 2^k T x 2^k T points
// below it k levels finer; each point is computed by its owner exactly as the
// level kernels compute it.
//   down (k_amg_down_chain): level l1 + k, k = 1 .. NLEV, on region D_k = the
//   owned points and the halo that the coarser levels' residuals need (side
//   s_k = 2 s_(k+1) + 2, origin o_k = 2 o_(k+1) - 1), r and 1 / S_pp in LDS;
//   owned points write r and their children's x to global memory.
//   up (k_amg_up_chain): level l1 + k, k = NLEV - 1 .. 0, on the owned points
//   and a halo of one (their prolongation reads the parents' e), e in LDS;
//   level l1 writes e to global memory.
constexpr int kAmgChainT = 4;
constexpr int kAmgChainLevels = 3;
constexpr int kAmgChainThreads = 1024;

__device__ __forceinline__ bool amg_in(const AmgLevel& L, int ix, int iz) {
  return ix >= 0 && iz >= 0 && ix < static_cast<int>(L.nx) && iz < static_cast<int>(L.nz);
}

template <int T, int NLEV>
struct AmgDown {
  // s_k = 2 s_(k+1) + 2 with s_NLEV = T
  __host__ __device__ static constexpr int side(int k) { return ((T + 2) << (NLEV - k)) - 2; }
  // LDS floats before level k's r (levels 1 .. NLEV - 1; r then 1 / S_pp)
  __host__ __device__ static constexpr int base(int k) {
    int b = 0;
    for (int m = 1; m < k; m++) b += 2 * side(m) * side(m);
    return b;
  }
  static constexpr int floats = base(NLEV);
};

template <int T, int NLEV>
__global__ void __launch_bounds__(kAmgChainThreads) k_amg_down_chain(AmgGrid g, int l1) {
  if (amg_skip(g)) return;
  using D = AmgDown<T, NLEV>;
  static_assert(D::side(NLEV) == T && D::side(NLEV - 1) == 2 * T + 2 && D::floats > 0, "chain regions");
  __shared__ float lds[D::floats > 0 ? D::floats : 1];
  const AmgLevel top = g.L[l1 + NLEV];
  const uint32_t tz_n = (top.nz + T - 1) / T;
  const int tx = static_cast<int>(blockIdx.x / tz_n), tz = static_cast<int>(blockIdx.x % tz_n);
#pragma unroll
  for (int k = 1; k <= NLEV; k++) {
    const AmgLevel F = g.L[l1 + k - 1], C = g.L[l1 + k];
    const int sk = D::side(k), sf = D::side(k - 1);
    const int own = T << (NLEV - k);  // owned side at this level
    const int ox = tx * own - (sk - own) / 2, oz = tz * own - (sk - own) / 2;  // region origin
    const int fox = 2 * ox - 1, foz = 2 * oz - 1;                            // the finer region's origin
    const float* fr = lds + D::base(k - 1);
    const float* fd = fr + sf * sf;
    float* cr = lds + D::base(k);
    float* cd = cr + sk * sk;
    for (int i = threadIdx.x; i < sk * sk; i += kAmgChainThreads) {
      const int qx = ox + i / sk, qz = oz + i % sk;
      if (!amg_in(C, qx, qz)) {
        if (k < NLEV) cr[i] = cd[i] = 0.0f;
        continue;
      }
      float rr[4][4], dd[4][4], st[2][2][9];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
          const int ix = 2 * qx - 1 + a, iz = 2 * qz - 1 + b;
          if (k == 1) {
            const bool in = amg_in(F, ix, iz);
            const uint32_t q = F.off + static_cast<uint32_t>(in ? ix : 0) * F.nz + static_cast<uint32_t>(in ? iz : 0);
            rr[a][b] = in ? g.r[q] : 0.0f;
            dd[a][b] = in ? g.dinv[q] : 0.0f;
          } else {
            const int li = (ix - fox) * sf + (iz - foz);
            rr[a][b] = fr[li];
            dd[a][b] = fd[li];
          }
        }
#pragma unroll
      for (int u = 0; u < 2; u++)
#pragma unroll
        for (int w = 0; w < 2; w++) {
          const int ix = 2 * qx + u, iz = 2 * qz + w;
          const bool in = amg_in(F, ix, iz);
          const uint32_t p = in ? static_cast<uint32_t>(ix) * F.nz + static_cast<uint32_t>(iz) : 0u;
#pragma unroll
          for (int d = 0; d < 9; d++)
            st[u][w][d] = in ? g.S[9ull * F.off + static_cast<size_t>(d) * amg_npts(F) + p] : 0.0f;
        }
      const bool owned = qx >= tx * own && qx < (tx + 1) * own && qz >= tz * own && qz < (tz + 1) * own;
      float rc = 0.0f;
#pragma unroll
      for (int u = 0; u < 2; u++)
#pragma unroll
        for (int w = 0; w < 2; w++) {
          const int ix = 2 * qx + u, iz = 2 * qz + w;
          if (!amg_in(F, ix, iz)) continue;
          float sum = 0.0f;
#pragma unroll
          for (int dx = -1; dx <= 1; dx++)
#pragma unroll
            for (int dz = -1; dz <= 1; dz++)
              sum = fmaf(st[u][w][3 * (dx + 1) + dz + 1],
                         g.omega * dd[1 + u + dx][1 + w + dz] * rr[1 + u + dx][1 + w + dz], sum);
          if (owned) g.x[F.off + static_cast<uint32_t>(ix) * F.nz + static_cast<uint32_t>(iz)] =
              g.omega * dd[1 + u][1 + w] * rr[1 + u][1 + w];
          rc += rr[1 + u][1 + w] - sum;
        }
      const uint32_t qc = C.off + static_cast<uint32_t>(qx) * C.nz + static_cast<uint32_t>(qz);
      if (owned) g.r[qc] = rc;
      if (k < NLEV) {
        cr[i] = rc;
        cd[i] = g.dinv[qc];
      }
    }
    if (k < NLEV) __syncthreads();
  }
}

template <int T, int NLEV>
struct AmgUp {
  // e of level l1 + k (k = 1 .. NLEV) on the owned points and a halo of one
  __host__ __device__ static constexpr int side(int k) { return (T << (NLEV - k)) + 2; }
  __host__ __device__ static constexpr int base(int k) {
    int b = 0;
    for (int m = NLEV; m > k; m--) b += side(m) * side(m);
    return b;
  }
  static constexpr int floats = base(1) + side(1) * side(1);
};

template <int T, int NLEV>
__global__ void __launch_bounds__(kAmgChainThreads) k_amg_up_chain(AmgGrid g, int l1) {
  if (amg_skip(g)) return;
  using U = AmgUp<T, NLEV>;
  __shared__ float lds[U::floats];
  const AmgLevel top = g.L[l1 + NLEV];
  const uint32_t tz_n = (top.nz + T - 1) / T;
  const int tx = static_cast<int>(blockIdx.x / tz_n), tz = static_cast<int>(blockIdx.x % tz_n);
  {
    const int sn = U::side(NLEV), ox = tx * T - 1, oz = tz * T - 1;
    float* e = lds + U::base(NLEV);
    for (int i = threadIdx.x; i < sn * sn; i += kAmgChainThreads) {
      const int qx = ox + i / sn, qz = oz + i % sn;
      e[i] = amg_in(top, qx, qz) ? g.e[top.off + static_cast<uint32_t>(qx) * top.nz + static_cast<uint32_t>(qz)] : 0.0f;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = NLEV - 1; k >= 0; k--) {
    const AmgLevel F = g.L[l1 + k], C = g.L[l1 + k + 1];
    const int own = T << (NLEV - k);
    const int sk = k > 0 ? U::side(k) : own;
    const int ox = tx * own - (k > 0 ? 1 : 0), oz = tz * own - (k > 0 ? 1 : 0);
    const int sc = U::side(k + 1), cox = tx * (own / 2) - 1, coz = tz * (own / 2) - 1;
    const float* ce = lds + U::base(k + 1);
    float* fe = k > 0 ? lds + U::base(k) : nullptr;
    for (int i = threadIdx.x; i < sk * sk; i += kAmgChainThreads) {
      const int px = ox + i / sk, pz = oz + i % sk;
      float val = 0.0f;
      if (amg_in(F, px, pz)) {
        const uint32_t p = static_cast<uint32_t>(px) * F.nz + static_cast<uint32_t>(pz);
        const float di = g.dinv[F.off + p];
        if (di != 0.0f) {
          float xq[9], sv[9];
#pragma unroll
          for (int dx = -1; dx <= 1; dx++)
#pragma unroll
            for (int dz = -1; dz <= 1; dz++) {
              const int kk = 3 * (dx + 1) + dz + 1;
              const int jx = px + dx, jz = pz + dz;
              const bool in = amg_in(F, jx, jz);
              const uint32_t q = F.off + static_cast<uint32_t>(in ? jx : 0) * F.nz + static_cast<uint32_t>(in ? jz : 0);
              const float ep = ce[((jx >> 1) - cox) * sc + ((jz >> 1) - coz)];
              xq[kk] = in ? fmaf(g.corr, ep, g.x[q]) : 0.0f;
              sv[kk] = in ? g.S[9ull * F.off + static_cast<size_t>(kk) * amg_npts(F) + p] : 0.0f;
            }
          float sum = 0.0f;
#pragma unroll
          for (int kk = 0; kk < 9; kk++) sum = fmaf(sv[kk], xq[kk], sum);
          val = fmaf(g.omega * di, g.r[F.off + p] - sum, xq[4]);
        }
        if (k == 0) g.e[F.off + p] = val;
      }
      if (k > 0) fe[i] = val;
    }
    if (k > 0) __syncthreads();
  }
}

// Levels l0 .. levels - 1 in one workgroup (at most kAmgLdsPts points from
// point off_l0 on): their stencils, 1 / S_center and r of level l0 are copied
// to LDS; the way down, the coarsest solve (e = inverse r) and the way up run
// in LDS, ordered by barriers; e of level l0 goes back to global memory.
constexpr int kAmgLdsFloats = 13 * kAmgLdsPts + kAmgDense * kAmgDense;  // 9 stencil + r, x, e, 1 / S_center per point; the inverse
static_assert(kAmgBlockThreads >= 16 * kAmgDense, "16 lanes per row of the coarsest solve");
struct AmgStenLds {
  const AmgGrid* g;
  const float* s;  // the stencils of the levels from l0 on, as in global memory
  uint32_t base;
  __device__ __forceinline__ float operator()(int l, int d, uint32_t p) const {
    const AmgLevel L = g->L[l];
    return s[9u * (L.off - base) + static_cast<uint32_t>(d) * amg_npts(L) + p];
  }
};
static_assert(kAmgLdsPts % 4 == 0 && kAmgLdsFloats % 4 == 0, "float4 regions");
__device__ __forceinline__ void amg_block_body(const AmgGrid& g, int l0) {
  __shared__ float4 lds4[kAmgLdsFloats / 4];
  float* const lds = reinterpret_cast<float*>(lds4);
  const uint32_t base = g.L[l0].off;
  const uint32_t npts = g.L[g.levels - 1].off + amg_npts(g.L[g.levels - 1]) - base;
  float* const ls = lds;
  float* const lv = lds + 9 * kAmgLdsPts;
  const AmgVec v{lv, lv + kAmgLdsPts, lv + 2 * kAmgLdsPts, lv + 3 * kAmgLdsPts, base};
  float* const vd = lv + 3 * kAmgLdsPts;
  const uint32_t n0 = amg_npts(g.L[l0]);
  const AmgLevel B = g.L[g.levels - 1];
  const uint32_t nb = amg_npts(B), bo = B.off - base;
  float* const linv = lds + 13 * kAmgLdsPts;
  // one index space over the stencils, the inverse, 1 / S_center and r of
  // level l0 in float4 units (level offsets are multiples of 4; the arrays
  // and the LDS regions have room for the rounding up), eight loads per
  // thread in flight at a time
  const uint32_t a0 = (9 * npts + 3) / 4, a1 = a0 + (nb * nb + 3) / 4, a2 = a1 + (npts + 3) / 4,
                 a3 = a2 + (n0 + 3) / 4;
  const float4* __restrict__ s4 = reinterpret_cast<const float4*>(g.S + 9ull * base);
  const float4* __restrict__ i4 = reinterpret_cast<const float4*>(g.inv);
  const float4* __restrict__ d4 = reinterpret_cast<const float4*>(g.dinv + base);
  const float4* __restrict__ r4 = reinterpret_cast<const float4*>(g.r + base);
  for (uint32_t k0 = threadIdx.x; k0 < a3; k0 += 8 * kAmgBlockThreads) {
    float4 t[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const uint32_t k = k0 + u * kAmgBlockThreads;
      const float4* src = k < a0 ? s4 + k : (k < a1 ? i4 + (k - a0) : (k < a2 ? d4 + (k - a1) : r4 + (k - a2)));
      t[u] = k < a3 ? *src : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const uint32_t k = k0 + u * kAmgBlockThreads;
      if (k < a3)
        lds4[k < a0 ? k
                    : (k < a1 ? 13 * kAmgLdsPts / 4 + (k - a0)
                              : (k < a2 ? 12 * kAmgLdsPts / 4 + (k - a1) : 9 * kAmgLdsPts / 4 + (k - a2)))] = t[u];
    }
  }
  __syncthreads();
  const AmgStenLds sten{&g, ls, base};
  for (int l = l0; l + 1 < g.levels; l++) {
    const uint32_t n1 = amg_npts(g.L[l + 1]);
    for (uint32_t P = threadIdx.x; P < n1; P += kAmgBlockThreads) amg_down_point(g, v, sten, l, P);
    __syncthreads();
  }
  {
    // e = inverse r: 16 lanes per row (nb <= 64 rows), partial sums over q = part mod 16
    const uint32_t p = threadIdx.x >> 4, part = threadIdx.x & 15;
    float s = 0.0f;
    if (p < nb)
      for (uint32_t q = part; q < nb; q += 16) s = fmaf(linv[p * nb + q], v.r[bo + q], s);
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) s += __shfl_xor(s, m, 16);
    if (p < nb && part == 0) v.e[bo + p] = s;
  }
  __syncthreads();
  for (int l = g.levels - 2; l >= l0; l--) {
    const uint32_t nl = amg_npts(g.L[l]);
    for (uint32_t p = threadIdx.x; p < nl; p += kAmgBlockThreads) amg_up_point(g, v, sten, l, p);
    __syncthreads();
  }
  for (uint32_t p = threadIdx.x; p < n0; p += kAmgBlockThreads) g.e[base + p] = v.e[p];
}
__global__ void __launch_bounds__(kAmgBlockThreads) k_amg_block(AmgGrid g, int l0) {
  if (amg_skip(g)) return;
  amg_block_body(g, l0);
}

// The operands of k_ml_prec.
struct PrecArgs {
  CellGrid G;
  const uint32_t* list;
  uint32_t nact;
  const uint32_t* pid;
  const float* dvp;
  const float4* sc;
  const double* r64;
  double inv_norm;
  float *r, *d;
  const float *q, *dinv;
  float* z;
  unsigned zs;
  const double* part_pq;
  int npq;
  double* part;
  size_t part_stride;
  CgStatus* st;
  // the cell-grid multigrid (modes 1, 2): the residual's cell sums out, the
  // coarse correction per cell in, the Jacobi damping
  float* rc;
  const float* ec;
  float omega;
  const uint32_t* row_len;
  uint32_t bx0, bz0, bnx, bnz;  // the multigrid's level 0: cells [bx0, bx0 + bnx) x [bz0, bz0 + bnz)
  // the smoothed prolongation (modes 1, 2): P_s's rows per packed slot, and
  // mode 1's 9 direction sums per level-0 point, at tq[d * tq_stride + p]
  const uint4* q9;
  const uint16_t* q9c;
  float* tq;
  size_t tq_stride;
};

// P_s's row of packed slot (w, u, lane): 9 FP16 coefficients.
__device__ __forceinline__ void load_q9(const PrecArgs& a, uint32_t w, int u, uint32_t lane, float q[9]) {
  const size_t o = (static_cast<size_t>(w) * kPerLane + u) * 64 + lane;
  const uint4 v = a.q9[o];
  const uint32_t h[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const uint32_t bits = (h[e >> 1] >> (16 * (e & 1))) & 0xffffu;
    q[e] = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(bits)));
  }
  q[8] = static_cast<float>(__builtin_bit_cast(_Float16, a.q9c[o]));
}

// Level-0 point of a cell of the multigrid's box, or -1 outside it.
__device__ __forceinline__ int amg_point(const PrecArgs& a, uint32_t cell, uint32_t gnz) {
  const uint32_t cx = cell / gnz - a.bx0, cz = cell % gnz - a.bz0;  // unsigned: below the box wraps high
  return (cx < a.bnx && cz < a.bnz) ? static_cast<int>(cx * a.bnz + cz) : -1;
}

// One wavefront per listed 4x4 block of cells.  MODE 0: the additive
// multilevel preconditioner (residual update, aggregate sums, z = M r).  The
// cell-grid multigrid splits the pass in two around its coarse cycle: MODE 1
// updates r and writes the cells' residual sums (and the r.r partials); MODE 2
// forms z = omega r / A_ii + e(cell) on aggregated rows, r / A_ii on the
// others (and the r.z partials).
template <int MODE>
__device__ __forceinline__ void ml_prec_body(const PrecArgs& a, long long it, const Frame& f,
                                             const AmgGrid* amg = nullptr) {
  __shared__ double lds[NTP / 64];
  const CellGrid& G = a.G;
  const uint32_t* __restrict__ list = a.list;
  const uint32_t nact = a.nact;
  const uint32_t* __restrict__ pid = a.pid;
  const float* __restrict__ dvp = a.dvp;
  const float4* __restrict__ sc = a.sc;
  const double* __restrict__ r64 = a.r64;
  const double inv_norm = a.inv_norm;
  float* __restrict__ r = a.r;
  float* __restrict__ d = a.d;
  const float* __restrict__ q = a.q;
  const float* __restrict__ dinv = a.dinv;
  float* __restrict__ z = a.z;
  const unsigned zs = a.zs;
  const double* __restrict__ part_pq = a.part_pq;
  const int npq = a.npq;
  double* __restrict__ part = a.part;
  const size_t part_stride = a.part_stride;
  CgStatus* st = a.st;
  if (stopped_before(st, it)) return;
  // z_i is written at z[zs * i] (the value field of the SpMV's records)
  // the p.q partials of k_ml_spmv(it) are loaded first and summed once the
  // particle loads below are in flight (npq <= kRedBlocks)
  double pq_part[kPqPerThread];
#pragma unroll
  for (int u = 0; u < kPqPerThread; u++) {
    const int k = threadIdx.x + u * NTP;
    pq_part[u] = (MODE != 2 && it > 0 && k < npq) ? part_pq[k] : 0.0;
  }
  const uint32_t w = f.bid * (NTP / 64) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  const bool active = w < nact;  // wave-uniform
  float4 scl = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t id[kPerLane];
  float rv[kPerLane], qv[kPerLane], dv[kPerLane];
  if (active) {
    scl = sc[static_cast<size_t>(w) * 64 + lane];
#pragma unroll
    for (int u = 0; u < kPerLane; u++) {
      const size_t o = (static_cast<size_t>(w) * kPerLane + u) * 64 + lane;
      id[u] = pid[o];
      dv[u] = dvp[o];
    }
#pragma unroll
    for (int u = 0; u < kPerLane; u++) {
      rv[u] = qv[u] = 0.0f;
      if (id[u] != kNoSlot) {
        const uint32_t i = id[u] & ~kOutAgg;
        if (MODE == 2) {
          rv[u] = r[i];
        } else if (it == 0) {
          rv[u] = static_cast<float>(r64[i] * inv_norm);
        } else {
          rv[u] = r[i];
          qv[u] = q[i];
        }
      }
    }
  }
  float alpha = 0.0f;
  if (MODE != 2 && it > 0) {
    double s = 0;
#pragma unroll
    for (int u = 0; u < kPqPerThread; u++) s += pq_part[u];
    const double al = st->gamma[(it - 1) & 1] / block_sum<NTP>(s, lds);
    if (lead(f)) st->alpha_ml = al;
    alpha = static_cast<float>(al);
  }
  double rho = 0, rr = 0;
  if (active) {
    const uint32_t cnt = static_cast<uint32_t>(scl.w);
    // a cell beyond 4 kPerLane particles (larger cell capacities): its
    // remaining members through the cell table
    BlockLane L{false, 0, 0, 0, 0};
    uint32_t st0 = 0;
    const bool over = cnt > 4u * kPerLane;
    if (over || MODE != 0) {
      L = block_lane(G, list[w]);
      st0 = L.valid ? G.start[L.cell] : 0u;
    }
    // rows beyond the registers (over): aggregated as k_ml_pack decides
    auto agg_of = [&](uint32_t i) { return in_agg(G.type[i]) && (MODE == 0 || a.row_len[i] > 0); };
    float s = 0.0f, s_over = 0.0f;
    if (MODE != 2) {
#pragma unroll
      for (int u = 0; u < kPerLane; u++) {
        if (id[u] != kNoSlot) {
          const uint32_t i = id[u] & ~kOutAgg;
          const float ri = (it == 0) ? rv[u] : fmaf(-alpha, qv[u], rv[u]);
          rv[u] = ri;
          r[i] = ri;
          if (it == 0) d[i] = 0.0f;
          if (!(id[u] & kOutAgg)) s += ri;
          rr = fma(static_cast<double>(ri), static_cast<double>(ri), rr);
        }
      }
      if (over)
        for (uint32_t k = L.sub + 4u * kPerLane; k < cnt; k += 4) {
          const uint32_t i = G.sorted[st0 + k];
          float ri;
          if (it == 0) {
            ri = static_cast<float>(r64[i] * inv_norm);
            d[i] = 0.0f;
          } else {
            ri = fmaf(-alpha, q[i], r[i]);
          }
          r[i] = ri;
          if (agg_of(i)) s_over += ri;
          rr = fma(static_cast<double>(ri), static_cast<double>(ri), rr);
        }
      s += s_over;
    }
    if (MODE == 1) {
      // the cell's 9 direction sums of P_s^T r: T_d = sum_i q_i[d] r_i over its
      // aggregated rows (rows beyond the registers keep P's row: direction
      // 4 only); lanes 0, 16, 32, 48 hold one cell's members
      float t9[9];
#pragma unroll
      for (int e = 0; e < 9; e++) t9[e] = 0.0f;
#pragma unroll
      for (int u = 0; u < kPerLane; u++) {
        if (id[u] != kNoSlot && !(id[u] & kOutAgg)) {
          float qv9[9];
          load_q9(a, w, u, lane, qv9);
#pragma unroll
          for (int e = 0; e < 9; e++) t9[e] = fmaf(qv9[e], rv[u], t9[e]);
        }
      }
      t9[4] += s_over;
      const int pt = L.valid ? amg_point(a, L.cell, G.gnz) : -1;
#pragma unroll
      for (int e = 0; e < 9; e++) {
        float t = t9[e] + __shfl_xor(t9[e], 16, 64);
        t = t + __shfl_xor(t, 32, 64);
        if (pt >= 0 && L.sub == 0) a.tq[static_cast<size_t>(e) * a.tq_stride + static_cast<uint32_t>(pt)] = t;
      }
    } else if (MODE == 2) {
      // z = omega_r r / A_ii + P_s e on aggregated rows: e of the 9 cells
      // around the lane's cell (level 0 of the cycle; 0 outside the box)
      const int pt = L.valid ? amg_point(a, L.cell, G.gnz) : -1;
      float e9[9];
      if (amg) {
        // the cycle's last up-pass (level 0) folded in: lane l < 36 computes
        // e of cell (l / 6, l % 6) of the 6x6 cells around the wavefront's
        // 4x4 block, and each lane takes its cell's 9 from those lanes
        const uint32_t b = list[w];
        const int cx = static_cast<int>(4 * (b / G.bnz)) - 1 + static_cast<int>(lane / 6) - static_cast<int>(a.bx0);
        const int cz = static_cast<int>(4 * (b % G.bnz)) - 1 + static_cast<int>(lane % 6) - static_cast<int>(a.bz0);
        float val = 0.0f;
        if (lane < 36 && cx >= 0 && cz >= 0 && cx < static_cast<int>(a.bnx) && cz < static_cast<int>(a.bnz))
          val = amg_up_value(*amg, amg_global(*amg), AmgStenGlobal{amg}, 0,
                             static_cast<uint32_t>(cx) * a.bnz + static_cast<uint32_t>(cz));
        const int ux = static_cast<int>((lane & 15u) >> 2), uz = static_cast<int>(lane & 3u);
#pragma unroll
        for (int e = 0; e < 9; e++) e9[e] = __shfl(val, (1 + ux + e / 3 - 1) * 6 + (1 + uz + e % 3 - 1), 64);
      } else {
        const int bx = (pt >= 0) ? pt / static_cast<int>(a.bnz) : 0, bz = (pt >= 0) ? pt % static_cast<int>(a.bnz) : 0;
#pragma unroll
        for (int e = 0; e < 9; e++) {
          const int x = bx + e / 3 - 1, zc = bz + e % 3 - 1;
          const bool in = pt >= 0 && x >= 0 && zc >= 0 && x < static_cast<int>(a.bnx) && zc < static_cast<int>(a.bnz);
          e9[e] = in ? a.ec[static_cast<uint32_t>(x) * a.bnz + static_cast<uint32_t>(zc)] : 0.0f;
        }
      }
      const float om = a.omega;
#pragma unroll
      for (int u = 0; u < kPerLane; u++) {
        if (id[u] != kNoSlot) {
          const uint32_t i = id[u] & ~kOutAgg;
          float zi;
          if (id[u] & kOutAgg) {
            zi = dv[u] * rv[u];
          } else {
            float qv9[9];
            load_q9(a, w, u, lane, qv9);
            float ez = 0.0f;
#pragma unroll
            for (int e = 0; e < 9; e++) ez = fmaf(qv9[e], e9[e], ez);
            zi = fmaf(om * dv[u], rv[u], ez);
          }
          z[static_cast<size_t>(i) * zs] = zi;
          rho = fma(static_cast<double>(rv[u]), static_cast<double>(zi), rho);
        }
      }
      if (over)
        for (uint32_t k = L.sub + 4u * kPerLane; k < cnt; k += 4) {
          const uint32_t i = G.sorted[st0 + k];
          const float ri = r[i];
          const float zi = agg_of(i) ? fmaf(om * dinv[i], ri, e9[4]) : dinv[i] * ri;
          z[static_cast<size_t>(i) * zs] = zi;
          rho = fma(static_cast<double>(ri), static_cast<double>(zi), rho);
        }
    } else {
      float y0, y1, y2;
      agg_sums(s, y0, y1, y2);
      const float corr = y0 * scl.x + y1 * scl.y + y2 * scl.z;
#pragma unroll
      for (int u = 0; u < kPerLane; u++) {
        if (id[u] != kNoSlot) {
          const uint32_t i = id[u] & ~kOutAgg;
          const float zi = (id[u] & kOutAgg) ? dv[u] * rv[u] : fmaf(dv[u], rv[u], corr);
          z[static_cast<size_t>(i) * zs] = zi;
          rho = fma(static_cast<double>(rv[u]), static_cast<double>(zi), rho);
        }
      }
      if (over)
        for (uint32_t k = L.sub + 4u * kPerLane; k < cnt; k += 4) {
          const uint32_t i = G.sorted[st0 + k];
          const float ri = r[i];
          const float zi = in_agg(G.type[i]) ? fmaf(dinv[i], ri, corr) : dinv[i] * ri;
          z[static_cast<size_t>(i) * zs] = zi;
          rho = fma(static_cast<double>(ri), static_cast<double>(zi), rho);
        }
    }
  }
  const double srho = block_sum<NTP>(rho, lds);
  const double srr = block_sum<NTP>(rr, lds);
  if (threadIdx.x == 0) {
    if (MODE != 1) part[f.bid] = srho;
    if (MODE != 2) part[part_stride + f.bid] = srr;
  }
}
template <int MODE>
__global__ void __launch_bounds__(NTP) k_ml_prec(PrecArgs a, long long it) {
  ml_prec_body<MODE>(a, it, whole_grid());
}
// MODE 2 with the cycle's level-0 up-pass folded in (it reads level 1's
// result and level 0's pre-smoothed iterate).
__global__ void __launch_bounds__(NTP) k_ml_prec_up(PrecArgs a, long long it, AmgGrid g) {
  ml_prec_body<2>(a, it, whole_grid(), &g);
}

// The set-up of P_s's rows in packed-slot order (after k_ml_pack): slot o
// of particle i (aggregated) gets q[d] = [d == 4] - omega_s (A P)_i,d / A_ii,
// (A P)_i,d = rows9[9 i + d]; other slots 0.
__global__ void __launch_bounds__(256) k_amg_q(size_t nslots, const uint32_t* __restrict__ pid,
                                               const double* __restrict__ rows9, const double* __restrict__ diag,
                                               double omega_s, uint4* __restrict__ q9,
                                               uint16_t* __restrict__ q9c) {
  const size_t o = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (o >= nslots) return;
  const uint32_t id = pid[o];
  uint32_t h[5] = {};
  if (id != kNoSlot && !(id & kOutAgg)) {
    const double sc = omega_s / diag[id];
#pragma unroll
    for (int e = 0; e < 9; e++) {
      const double v = (e == 4 ? 1.0 : 0.0) - sc * rows9[9ull * id + e];
      const uint32_t bits = __builtin_bit_cast(uint16_t, static_cast<_Float16>(v));
      h[e >> 1] |= bits << (16 * (e & 1));
    }
  }
  q9[o] = make_uint4(h[0], h[1], h[2], h[3]);
  q9c[o] = static_cast<uint16_t>(h[4]);
}

// Level 0's right-hand side P_s^T r from pass 1's direction sums:
// r0(p) = sum_d T_d(p - off(d)), T_d(c) the share of cell c's rows for the
// cell c + off(d) (off(d) = (d / 3 - 1, d % 3 - 1)).
__device__ __forceinline__ void amg_rfix_point(const float* __restrict__ tq, size_t stride, uint32_t nx,
                                               uint32_t nz, float* __restrict__ r0, uint32_t p) {
  const int x = static_cast<int>(p / nz), zc = static_cast<int>(p % nz);
  float s = 0.0f;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    const int sx = x - (e / 3 - 1), sz = zc - (e % 3 - 1);
    if (sx >= 0 && sz >= 0 && sx < static_cast<int>(nx) && sz < static_cast<int>(nz))
      s += tq[static_cast<size_t>(e) * stride + static_cast<uint32_t>(sx) * nz + static_cast<uint32_t>(sz)];
  }
  r0[p] = s;
}
__global__ void __launch_bounds__(256) k_amg_rfix(const float* __restrict__ tq, size_t stride, uint32_t nx,
                                                  uint32_t nz, float* __restrict__ r0, const CgStatus* st,
                                                  long long it) {
  if (stopped_before(st, it)) return;  // an iteration past the stop
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p < nx * nz) amg_rfix_point(tq, stride, nx, nz, r0, p);
}

// ---- set-up (FP64 sums, FP32 levels) ----

// Per aggregated row i (in an aggregate: not dummy / disabled, stored
// off-diagonals, in a cell): its diagonal and off-diagonals to aggregated
// columns, summed by the direction of the column's cell (9 sums, FP64, at
// rows[9 i + d]).  The fill computes the same sums while it stores the row
// when the system is assembled for this solver; this kernel serves a system
// assembled for another one.
__global__ void __launch_bounds__(NT, 8) k_amg_rows(Sell<double> A, const uint32_t* __restrict__ row_len,
                                                    const uint32_t* __restrict__ key,
                                                    const uint8_t* __restrict__ type, uint32_t gnz,
                                                    double* __restrict__ rows) {
  const Rows rw = block_rows(A.n, A.rpb);
  for (size_t i = rw.lo + threadIdx.x; i < rw.hi; i += NT) {
    double acc[9];
#pragma unroll
    for (int d = 0; d < 9; d++) acc[d] = 0;
    const uint32_t ki = key[i];
    const uint32_t m = row_len[i];
    if (in_agg(type[i]) && m > 0 && ki != kNoCell) {
      const int cx = static_cast<int>(ki / gnz), cz = static_cast<int>(ki % gnz);
      acc[4] = A.diag[i];
      const size_t base = A.off[i >> 6] + (i & 63);
      for (uint32_t q = 0; q < m; q++) {
        const size_t o = base + static_cast<size_t>(q) * 64;
        const uint32_t j = sell_col(A, i, o);
        const uint32_t kj = key[j];
        if (kj == kNoCell || !in_agg(type[j]) || row_len[j] == 0) continue;
        const int dx = static_cast<int>(kj / gnz) - cx, dz = static_cast<int>(kj % gnz) - cz;
        const double a = A.val[o];
#pragma unroll
        for (int d = 0; d < 9; d++)
          if (d == 3 * (dx + 1) + dz + 1) acc[d] += a;
      }
    }
#pragma unroll
    for (int d = 0; d < 9; d++) rows[9 * i + d] = acc[d];
  }
}

// Level 0 (the cells of the box x0.. x0 + nx - 1, z0 .. z0 + nz - 1): the
// Galerkin stencil, sums of the member rows' sums.
__global__ void __launch_bounds__(256) k_amg_cells(uint32_t x0, uint32_t z0, uint32_t nx, uint32_t nz, uint32_t gnz,
                                                   const uint32_t* __restrict__ start,
                                                   const uint32_t* __restrict__ count,
                                                   const uint32_t* __restrict__ sorted, const double* __restrict__ rows,
                                                   size_t n, double* __restrict__ S) {
  const size_t pt = blockIdx.x * 256ull + threadIdx.x;
  const size_t npts = static_cast<size_t>(nx) * nz;
  if (pt >= npts) return;
  const size_t c = static_cast<size_t>(x0 + pt / nz) * gnz + (z0 + pt % nz);
  const uint32_t s0 = start[c], cnt = count[c];
  double acc[9];
#pragma unroll
  for (int d = 0; d < 9; d++) acc[d] = 0;
  for (uint32_t k = 0; k < cnt; k++) {
    const uint32_t i = sorted[s0 + k];
#pragma unroll
    for (int d = 0; d < 9; d++) acc[d] += rows[9ull * i + d];
  }
#pragma unroll
  for (int d = 0; d < 9; d++) S[d * npts + pt] = acc[d];
}

// Level l + 1 from level l: the coarse stencil of each 2x2 block (FP64).
__global__ void __launch_bounds__(256) k_amg_coarsen(const double* __restrict__ Sf, uint32_t fnx, uint32_t fnz,
                                                     double* __restrict__ Sc, uint32_t cnx, uint32_t cnz) {
  const uint32_t P = blockIdx.x * 256u + threadIdx.x;
  if (P >= cnx * cnz) return;
  const uint32_t X = P / cnz, Z = P % cnz;
  const size_t nf = static_cast<size_t>(fnx) * fnz, nc = static_cast<size_t>(cnx) * cnz;
  double acc[9];
#pragma unroll
  for (int d = 0; d < 9; d++) acc[d] = 0;
  for (int u = 0; u < 2; u++)
    for (int v = 0; v < 2; v++) {
      const uint32_t ix = 2 * X + u, iz = 2 * Z + v;
      if (ix >= fnx || iz >= fnz) continue;
      const size_t p = static_cast<size_t>(ix) * fnz + iz;
#pragma unroll
      for (int dx = -1; dx <= 1; dx++)
#pragma unroll
        for (int dz = -1; dz <= 1; dz++) {
          const int jx = static_cast<int>(ix) + dx, jz = static_cast<int>(iz) + dz;
          if (jx < 0 || jz < 0 || jx >= static_cast<int>(fnx) || jz >= static_cast<int>(fnz)) continue;
          const int DX = (jx >> 1) - static_cast<int>(X), DZ = (jz >> 1) - static_cast<int>(Z);
          const double a = Sf[(3 * (dx + 1) + dz + 1) * nf + p];
#pragma unroll
          for (int d = 0; d < 9; d++)
            if (d == 3 * (DX + 1) + DZ + 1) acc[d] += a;
        }
    }
#pragma unroll
  for (int d = 0; d < 9; d++) Sc[d * nc + P] = acc[d];
}

// FP32 copies of every level's stencils and 1 / S_pp (one thread per point
// of any level; the padding between levels is skipped).
__global__ void __launch_bounds__(256) k_amg_f32_all(AmgGrid g, const double* __restrict__ S64, size_t pts,
                                                     float* __restrict__ S32, float* __restrict__ dinv) {
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= pts) return;
  int l = 0;
  for (int m = 1; m < g.levels; m++)
    if (g.L[m].off <= k) l = m;
  const AmgLevel L = g.L[l];
  const uint32_t np = L.nx * L.nz, p = k - L.off;
  if (p >= np) return;
#pragma unroll
  for (int d = 0; d < 9; d++) {
    const size_t o = 9ull * L.off + static_cast<size_t>(d) * np + p;
    S32[o] = static_cast<float>(S64[o]);
  }
  const double c = S64[9ull * L.off + 4ull * np + p];
  dinv[L.off + p] = (c != 0.0) ? static_cast<float>(1.0 / c) : 0.0f;
}

// The inverse of the coarsest level's operator (n <= kAmgDense points; empty
// points become identity rows), Gauss-Jordan in FP64 in one workgroup's LDS:
// pivot k turns buffer b into buffer 1 - b (row k divided by the pivot, the
// other rows less their multiple of it), one barrier per pivot.
__global__ void __launch_bounds__(1024) k_amg_dense(const double* __restrict__ S, uint32_t nx, uint32_t nz,
                                                    float* __restrict__ inv) {
  __shared__ double M[2][kAmgDense][2 * kAmgDense];
  const int n = static_cast<int>(nx * nz);
  for (int e = threadIdx.x; e < n * 2 * n; e += blockDim.x) {
    const int p = e / (2 * n), q = e % (2 * n);
    double v = 0;
    if (q >= n) {
      v = (q - n == p) ? 1.0 : 0.0;
    } else {
      const int px = p / static_cast<int>(nz), pz = p % static_cast<int>(nz);
      const int qx = q / static_cast<int>(nz), qz = q % static_cast<int>(nz);
      const double c = S[4 * n + p];
      if (c == 0.0) {
        v = (p == q) ? 1.0 : 0.0;
      } else if (abs(qx - px) <= 1 && abs(qz - pz) <= 1) {
        v = S[(3 * (qx - px + 1) + qz - pz + 1) * n + p];
      }
    }
    M[0][p][q] = v;
  }
  __syncthreads();
  int b = 0;
  for (int k = 0; k < n; k++) {
    const double pk = M[b][k][k];
    for (int e = threadIdx.x; e < n * 2 * n; e += blockDim.x) {
      const int p = e / (2 * n), q = e % (2 * n);
      const double rk = M[b][k][q] / pk;
      M[1 - b][p][q] = (p == k) ? rk : M[b][p][q] - M[b][p][k] * rk;
    }
    b = 1 - b;
    __syncthreads();
  }
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) inv[e] = static_cast<float>(M[b][e / n][n + e % n]);
}

// See above; FIRST is it == 1 (p = z, q = A z, tolerance from r_0.r_0).  The
// partials of k_ml_prec(it - 1) are loaded at the start and summed after the
// gathers of the first kSpmvRows rows per thread.
// (A z)_i and z_i: z from the records (MatFreeRec) or the array z.
template <class Op>
__device__ __forceinline__ float spmv_row(const Op& A, const float* __restrict__ z, size_t i, float& zi) {
  if constexpr (std::is_same<Op, MatFreeRec>::value) {
    return rec_apply<kIterUnroll + 1>(A, i, zi);
  } else {
    zi = z[i];
    return op_apply<kIterUnroll + 1>(A, i, zi, [&](uint32_t j) { return z[j]; });
  }
}

constexpr int kSpmvParts = 3;  // prec partials per thread and array in registers
constexpr int kSpmvRows = 2;   // rows per thread whose A z waits in registers for beta

// The operands of k_ml_spmv.
template <class Op>
struct SpmvArgs {
  Op A;
  const float* z;
  float *p, *q, *d;
  const double* part_ml;
  size_t part_stride;
  int nparts;
  double* part_pq;
  CgStatus* st;
  double tol_rel;
};

template <bool FIRST, class Op>
__device__ __forceinline__ void ml_spmv_body(const SpmvArgs<Op>& args, long long it, const Frame& f) {
  __shared__ double lds[2 * (NT / 64)];
  const Op& A = args.A;
  const float* __restrict__ z = args.z;
  float* __restrict__ p = args.p;
  float* __restrict__ q = args.q;
  float* __restrict__ d = args.d;
  const double* __restrict__ part_ml = args.part_ml;
  const size_t part_stride = args.part_stride;
  const int nparts = args.nparts;
  double* __restrict__ part_pq = args.part_pq;
  CgStatus* st = args.st;
  const double tol_rel = args.tol_rel;
  // a stop at it - 2 or earlier was found (and finished) by an earlier launch
  if (stopped_before(st, it - 1)) return;
  double pa[kSpmvParts], pb[kSpmvParts];
#pragma unroll
  for (int u = 0; u < kSpmvParts; u++) {
    const int k = threadIdx.x + u * NT;
    pa[u] = (k < nparts) ? part_ml[k] : 0.0;
    pb[u] = (k < nparts) ? part_ml[part_stride + k] : 0.0;
  }
  const Rows rw = block_rows(A.n, A.rpb, f);
  float zc[kSpmvRows], sc[kSpmvRows];
#pragma unroll
  for (int u = 0; u < kSpmvRows; u++) {
    const size_t i = rw.lo + threadIdx.x + static_cast<size_t>(u) * NT;
    zc[u] = sc[u] = 0.0f;
    if (i < rw.hi) sc[u] = spmv_row(A, z, i, zc[u]);
  }
  double a = 0, b = 0;
#pragma unroll
  for (int u = 0; u < kSpmvParts; u++) {
    a += pa[u];
    b += pb[u];
  }
  for (int k = threadIdx.x + kSpmvParts * NT; k < nparts; k += NT) {
    a += part_ml[k];
    b += part_ml[part_stride + k];
  }
  double rho, rr;
  block_sum2<NT>(a, b, lds, rho, rr);
  const double tol = FIRST ? rr * tol_rel * tol_rel : st->tol;
  const bool stop = FIRST ? (tol == 0) : (rr < tol);
  if (lead(f)) {
    const int par = static_cast<int>((it - 1) & 1);
    st->gamma[par] = rho;
    st->rr[par] = rr;
    if (FIRST) {
      st->tol = tol;
      st->rr0 = rr;
    }
    if (stop) {
      st->iterations = it - 1;
      st->converged = 1;
      st->stop_mark = it;
    }
  }
  const float alpha = FIRST ? 0.0f : static_cast<float>(st->alpha_ml);
  if (stop) {
    if (!FIRST)
      for (size_t i = rw.lo + threadIdx.x; i < rw.hi; i += NT) d[i] = fmaf(alpha, p[i], d[i]);
    return;
  }
  const float beta = FIRST ? 0.0f : static_cast<float>(rho / st->gamma[it & 1]);
  double acc = 0;
  auto finish = [&](size_t i, float zi, float s) {
    float pi = zi, qi = s;
    if (!FIRST) {
      const float po = p[i];
      d[i] = fmaf(alpha, po, d[i]);
      pi = fmaf(beta, po, zi);
      qi = fmaf(beta, q[i], s);
    }
    p[i] = pi;
    q[i] = qi;
    acc = fma(static_cast<double>(pi), static_cast<double>(qi), acc);
  };
#pragma unroll
  for (int u = 0; u < kSpmvRows; u++) {
    const size_t i = rw.lo + threadIdx.x + static_cast<size_t>(u) * NT;
    if (i < rw.hi) finish(i, zc[u], sc[u]);
  }
  for (size_t i = rw.lo + threadIdx.x + static_cast<size_t>(kSpmvRows) * NT; i < rw.hi; i += NT) {
    float zi;
    const float si = spmv_row(A, z, i, zi);
    finish(i, zi, si);
  }
  const double sp = block_sum<NT>(acc, lds);
  if (threadIdx.x == 0) part_pq[f.bid] = sp;
}
template <bool FIRST, class Op>
__global__ void __launch_bounds__(NT, 8) k_ml_spmv(SpmvArgs<Op> a, long long it) {
  ml_spmv_body<FIRST, Op>(a, it, whole_grid());
}

// d += alpha p after a sweep that ended without a stop (its last k_ml_prec
// computed alpha, no k_ml_spmv followed).
__global__ void __launch_bounds__(256) k_ml_finish(size_t n, const CgStatus* st, const float* __restrict__ p,
                                                   float* __restrict__ d) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) d[i] = fmaf(static_cast<float>(st->alpha_ml), p[i], d[i]);
}

// x += |r| d (FP64 accumulation of the FP32 correction).
__global__ void __launch_bounds__(256) k_ir_correct(size_t n, double* __restrict__ x,
                                                    const float* __restrict__ d, double norm) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) x[i] = fma(static_cast<double>(d[i]), norm, x[i]);
}

__global__ void __launch_bounds__(256) k_jacobi_inv(size_t n, const double* __restrict__ diag,
                                                    float* __restrict__ dinv) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) {
    const double d = diag[i];
    dinv[i] = (d != 0.0) ? static_cast<float>(1.0 / d) : 1.0f;
  }
}

struct Launch {
  int grid;
  size_t rpb;
};
Launch cg_grid(size_t n) {
  size_t blocks = (n + NT - 1) / NT;
  if (blocks > static_cast<size_t>(kRedBlocks)) blocks = kRedBlocks;
  if (blocks == 0) blocks = 1;
  size_t rpb = (n + blocks - 1) / blocks;
  rpb = (rpb + NT - 1) / NT * NT;
  blocks = (n + rpb - 1) / rpb;
  if (blocks == 0) blocks = 1;
  return Launch{static_cast<int>(blocks), rpb};
}

template <class T>
Sell<T> sell(const Context& c, const T* val, const T* diag, size_t rpb) {
  return Sell<T>{c.slice_off,  c.slice_width, val, c.scol, diag, c.ppe_n, rpb,
                 c.ppe_col16 ? c.col16 : nullptr, c.ppe_col16 ? c.slice_wide : nullptr};
}

// Reads the device status; folds the timed launches, dropping those that
// ran after convergence (iteration tags above the converged count).
void read_status(Context& c, CgStatus* dev, CgStatus* host) {
  MPS_HIP(hipMemcpyAsync(host, dev, sizeof(CgStatus), hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
  collect_events(c, host->converged ? host->iterations : 0x7fffffffffffffffLL);
}

// Iterations to enqueue before the host next reads the status: the predicted
// count (the same solve's count in the previous step) plus a margin, then
// short chunks.  Launches past the stop exit at once (a few microseconds);
// every status read costs a host round trip with the GPU idle.
long long next_chunk(long long done, long long predicted) {
  if (done == 0) return predicted > 0 ? predicted + 2 : 32;
  return std::max<long long>(8, std::min<long long>(256, done / 4));
}

// Algorithmic HBM bytes of one SpMV launch: every stored non-zero (value +
// column index: 2 bytes, 4 in wide slices or without 16-bit offsets) once,
// the diagonal once, `extra_per_row` bytes of streamed per-row vectors, and
// the slice table.  The wide slices' count is the last one synchronised.
double spmv_bytes(const Context& c, int vbytes, double extra_per_row) {
  const double nnz = static_cast<double>(c.nnz_offdiag);
  const double cols = c.ppe_col16 ? 2.0 * nnz + 2.0 * std::min(c.mf_nnz_wide, nnz) : 4.0 * nnz;
  return nnz * vbytes + cols + c.ppe_n * (vbytes + extra_per_row) + c.nslices * 13.0;
}

// Plain FP64 CG (the reference algorithm); x = c.sol holds the warm start.
void solve_cg64(Context& c, const StopRule& sr, long long max_iter) {
  const size_t n = c.ppe_n;
  const Launch L = cg_grid(n);
  const Sell<double> A = sell<double>(c, c.sval, c.diag, L.rpb);
  MPS_HIP(hipMemsetAsync(c.st, 0, sizeof(CgStatus), c.stream));
  c.replay_ready = MPS_K_CG_SPMV;
  tracked(c, MPS_K_CG_INIT, spmv_bytes(c, 8, 8.0 + 8 + 8), [&] {
    hipLaunchKernelGGL(k_cg_init<Sell<double>>, dim3(L.grid), dim3(NT), 0, c.stream, A, static_cast<const double*>(c.b),
                       static_cast<const double*>(c.sol), c.r, c.partB, c.partC);
  });
  check_launch();
  long long it = 1;
  // reads r (own row and gathered), reads and writes p and q (the diagonal
  // is in spmv_bytes)
  const double b_spmv = spmv_bytes(c, 8, 8.0 * 5);
  const double b_upd = c.ppe_n * 48.0;
  while (true) {
    const long long chunk = next_chunk(it - 1, c.last_iters);
    for (long long m = 0; m < chunk && it <= max_iter; m++, it++) {
      tracked(c, MPS_K_CG_SPMV, b_spmv, [&] {
        if (it == 1)
          hipLaunchKernelGGL((k_cg_spmv<true, Sell<double>>), dim3(L.grid), dim3(NT), 0, c.stream, A,
                             static_cast<const double*>(c.r), c.pa, c.q,
                             static_cast<const double*>(c.partB), static_cast<const double*>(c.partC),
                             L.grid, c.partA, c.st, sr, it);
        else
          hipLaunchKernelGGL((k_cg_spmv<false, Sell<double>>), dim3(L.grid), dim3(NT), 0, c.stream, A,
                             static_cast<const double*>(c.r), c.pa, c.q,
                             static_cast<const double*>(c.partB), static_cast<const double*>(c.partC),
                             L.grid, c.partA, c.st, sr, it);
      }, it);
      tracked(c, MPS_K_CG_UPDATE, b_upd, [&] {
        hipLaunchKernelGGL(k_cg_update, dim3(L.grid), dim3(NT), 0, c.stream, n, L.rpb, c.sol, c.r,
                           static_cast<const double*>(c.pa), static_cast<const double*>(c.q),
                           static_cast<const double*>(c.partA), L.grid, c.partB, c.st, it);
      }, it);
    }
    check_launch();
    if (it > max_iter) {
      hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(NT), 0, c.stream, static_cast<const double*>(c.partB),
                         L.grid, c.st, sr, it);
      check_launch();
      read_status(c, c.st, c.st_host);
      break;
    }
    read_status(c, c.st, c.st_host);
    if (c.st_host->converged) break;
  }
  const long long launched = std::min(it - 1, max_iter);
  uncount(c, MPS_K_CG_SPMV, launched - c.st_host->iterations, b_spmv);
  uncount(c, MPS_K_CG_UPDATE, launched - c.st_host->iterations, b_upd);
  c.stats.iterations = c.st_host->iterations;
  c.stats.outer_iterations = 1;
  c.stats.rr0 = c.st_host->rr0;
  c.stats.rr = c.st_host->rr[(c.st_host->iterations + 1) & 1];
  c.last_iters = c.st_host->iterations;
  if (!c.st_host->converged)
    throw Failure{MPS_ERR_CG_NOT_CONVERGED,
                  "Conjugate Gradient method couldn't solve Pressure Poison Equation"};
}

// r0 = b - A p with the FP32 matrix-free operator on records holding p (the
// learned warm start's feature input, for which FP32 is ample), written into the feature records
// with the positions and p, and block partials of |r0|^2, |b|^2 and, over
// alive rows, of p^2, diag^2 and their count: sum k at part[k G + block],
// G = gridDim.x.
// {qx, qz, float(p), 0} for every row: the operator's records holding p.
__global__ void __launch_bounds__(256) k_rec_p(size_t n, const uint2* __restrict__ qpos, const double* __restrict__ p,
                                               uint4* __restrict__ rec) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) {
    const uint2 q = qpos[i];
    rec[i] = make_uint4(q.x, q.y, __float_as_uint(static_cast<float>(p[i])), 0u);
  }
}

__global__ void __launch_bounds__(256) k_learn_resid32(MatFreeRec A, const double* __restrict__ b,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ diag,
                                                       const uint8_t* __restrict__ type,
                                                       const double* __restrict__ x, const double* __restrict__ z,
                                                       FeatRec* __restrict__ rec, double* __restrict__ part) {
  __shared__ double lds[4];
  double s[5] = {0, 0, 0, 0, 0};
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < A.n; i += gridDim.x * 256ull) {
    const double pv = p[i], bv = b[i];
    float pf;
    const float ap = rec_apply<4>(A, i, pf);
    const double rv = bv - static_cast<double>(ap);
    const uint8_t t = type[i];
    rec[i] = FeatRec{(t == kDummy) ? __builtin_nan("") : x[i], z[i], pv, rv};
    s[0] += rv * rv;
    s[1] += bv * bv;
    if (t != kDummy && t != kDisabled) {
      s[2] += pv * pv;
      s[3] += diag[i] * diag[i];
      s[4] += 1.0;
    }
  }
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const double tot = block_sum<256>(s[k], lds);
    if (threadIdx.x == 0) part[k * static_cast<size_t>(gridDim.x) + blockIdx.x] = tot;
  }
}

// out[k] = the sum of partials k of k_learn_resid32 (one block; the five
// sums' loads are all issued before any is added), then the feature scales.
__global__ void __launch_bounds__(1024) k_reduce5(const double* __restrict__ parts, int G, double* __restrict__ out) {
  __shared__ double lds[16];
  double v[5] = {0, 0, 0, 0, 0};
  for (int j = threadIdx.x; j < G; j += 1024) {
    double x[5];
#pragma unroll
    for (int k = 0; k < 5; k++) x[k] = parts[static_cast<size_t>(k) * G + j];
#pragma unroll
    for (int k = 0; k < 5; k++) v[k] += x[k];
  }
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const double tot = block_sum<1024>(v[k], lds);
    if (threadIdx.x == 0) out[k] = tot;
  }
  if (threadIdx.x == 0) feat_scales(out);
}

}  // namespace

// Single-block reduction of two partial arrays into c.red_host[0..1] (synchronous).
void reduce2(Context& c, const double* a, const double* b, int nparts) {
  hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(NT), 0, c.stream, a, b, nparts, c.red_out);
  check_launch();
  MPS_HIP(hipMemcpyAsync(c.red_host, c.red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
}

// FP64 residual |b - A x|^2 and |b|^2 (reuses the init kernel on scratch r).
void residual64_launch(Context& c) {
  const Launch L = cg_grid(c.ppe_n);
  const Sell<double> A = sell<double>(c, c.sval, c.diag, L.rpb);
  hipLaunchKernelGGL(k_cg_init<Sell<double>>, dim3(L.grid), dim3(NT), 0, c.stream, A, static_cast<const double*>(c.b),
                     static_cast<const double*>(c.sol), c.r, c.partB, c.partC);
  hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(NT), 0, c.stream, static_cast<const double*>(c.partB),
                     static_cast<const double*>(c.partC), L.grid, c.red_out);
}

void residual64(Context& c, double& rr, double& bb, bool track) {
  auto launch = [&] { residual64_launch(c); };
  if (track) tracked(c, MPS_K_IR_RESIDUAL, spmv_bytes(c, 8, 8.0 + 8 + 8), launch);
  else launch();
  check_launch();
  MPS_HIP(hipMemcpyAsync(c.red_host, c.red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
  rr = c.red_host[0];
  bb = c.red_host[1];
}

namespace {

// Mixed precision: FP32 CG sweeps on A d = r / |r| inside FP64 iterative
// refinement.  `sweep(norm, tol)` runs one FP32 solve to the relative
// tolerance tol, leaves d in c.d32 and returns its iterations; the refinement
// adds |r| d to x in FP64 and recomputes r = b - A x in FP64.
long long max_iter_fallback(const Context& c) { return static_cast<long long>(c.ppe_n); }

template <class Sweep, class Correct>
void refine(Context& c, const StopRule& sr, Sweep sweep, Correct correct) {
  const double inner_tol = c.solver.inner_tol > 0 ? c.solver.inner_tol : 1e-4;
  long long total = 0;
  double rr0 = 0, bb = 0, rr = 0;
  int outer = 0;
  const int max_outer = 64;
  // an explicit iteration limit caps the FP32 iterations of all sweeps; by
  // default each sweep may take up to N (CG's finite-termination bound) and
  // the solve fails when a sweep reduces the FP64 residual norm by less than sqrt(2)
  const long long total_cap = c.solver.max_iter > 0 ? c.solver.max_iter : LLONG_MAX;
  double rr_prev = 0;
  bool done = false;
  for (;; outer++) {
    // the learned warm start may have left r = b - A x and its norms
    if (!(outer == 0 && c.learned_pending && learned_resolve(c, rr, bb))) residual64(c, rr, bb, true);
    if (outer == 0) rr0 = rr;
    const bool stalled = outer > 0 && !(rr < 0.5 * rr_prev);
    rr_prev = rr;
    double target;  // outer r.r to reach
    if (sr.rule == MPS_STOP_RHS_NORM) {
      target = (sr.tol * sr.tol) * bb;
      done = rr <= target;
    } else {
      target = rr0 * sr.tol * sr.tol;
      done = (outer == 0) ? (target == 0) : (rr < target);
    }
    if (stalled && !done && rr == rr && total < total_cap) {
      // the FP32 sweeps no longer make progress: the plain FP64 CG (the
      // reference algorithm) finishes from the current x, to the same target
      StopRule rest = sr;
      if (sr.rule == MPS_STOP_INITIAL_RESIDUAL) rest.tol = sr.tol * std::sqrt(rr0 / rr);
      const long long cap = c.solver.max_iter > 0 ? c.solver.max_iter - total : max_iter_fallback(c);
      solve_cg64(c, rest, cap);
      c.stats.iterations += total;
      c.stats.outer_iterations = outer + 1;
      c.stats.rr0 = rr0;
      c.stats.fallback = 1;
      return;
    }
    if (done || stalled || outer >= max_outer || total >= total_cap || !(rr == rr)) break;
    const double norm = std::sqrt(rr);
    // this sweep reduces the residual by inner_tol, or only as far as the
    // outer target needs (with a factor 2 of margin), whichever is less work
    const double sweep_tol = std::max(inner_tol, 0.5 * std::sqrt(target / rr));
    c.sweep_index = outer;
    const long long inner = sweep(norm, sweep_tol);
    c.pred(outer) = inner;
    total += inner;
    correct(norm);
    check_launch();
  }
  c.stats.iterations = total;
  c.stats.outer_iterations = outer;
  c.stats.rr0 = rr0;
  c.stats.rr = rr;
  c.last_iters = total;
  // rr, bb: the FP64 residual of the returned x, computed above
  c.final_rr = rr;
  c.final_bb = bb;
  if (!done)
    throw Failure{MPS_ERR_CG_NOT_CONVERGED,
                  "Conjugate Gradient method couldn't solve Pressure Poison Equation"};
}

long long inner_limit(const Context& c, long long max_iter) {
  return c.solver.inner_max_iter > 0 ? c.solver.inner_max_iter : max_iter;
}

// Runs iterations 1, 2, ... of an inner solve in chunks until its status
// reports convergence or the limit; `launch(it)` enqueues iteration it.
// Returns the iterations done.
template <class Launch1>
long long run_inner(Context& c, long long limit, Launch1 launch) {
  long long it = 1;
  double rr_last = -1;
  int flat = 0;
  while (true) {
    const long long chunk = next_chunk(it - 1, c.pred(c.sweep_index));
    for (long long m = 0; m < chunk && it <= limit; m++, it++) launch(it);
    check_launch();
    read_status(c, c.st32, c.st32_host);
    if (c.st32_host->converged || it > limit) break;
    // FP32 stagnation: two chunks in a row that lower r.r by less than 19 %
    // (|r| by less than 10 %) end the sweep; the FP64 refinement decides
    const double rr = std::min(c.st32_host->rr[0], c.st32_host->rr[1]);
    flat = (rr_last > 0 && !(rr < 0.81 * rr_last)) ? flat + 1 : 0;
    rr_last = rr;
    if (flat >= 2) return it - 1;
  }
  return c.st32_host->converged ? c.st32_host->iterations : limit;
}

// Jacobi (or no) preconditioner: the one-kernel Chronopoulos-Gear iteration.
//   b_base: bytes of one iteration without the wide slices' extra column
//   bytes (wide: their stored nnz, known after the first synchronisation)
template <class Op>
void solve_mixed_cg(Context& c, const Op& A, double b_base, bool wide, const StopRule& sr, long long max_iter) {
  const size_t n = c.ppe_n;
  const Launch L = cg_grid(n);
  const bool jac = c.solver.precond != MPS_PRECOND_NONE;
  if (jac)
    hipLaunchKernelGGL(k_jacobi_inv, dim3(grid_for(n)), dim3(256), 0, c.stream, n,
                       static_cast<const double*>(c.diag), c.dinv32);
  const float* dinv = jac ? c.dinv32 : nullptr;
  const long long limit = inner_limit(c, max_iter);
  float4* gbuf[2] = {c.g32a, c.g32b};
  double* parts[2] = {c.part3a, c.part3b};
  double b_iter = b_base;
  refine(c, sr, [&](double norm, double sweep_tol) {
    // refine() has synchronised the stream (residual) before every sweep
    if (wide) b_iter = b_base + 2.0 * mf_nnz_wide(c);
    MPS_HIP(hipMemsetAsync(c.st32, 0, sizeof(CgStatus), c.stream));
    tracked(c, MPS_K_CG_INIT32, b_iter + n * 8.0, [&] {
      hipLaunchKernelGGL(k_cgf_init<Op>, dim3(L.grid), dim3(NT), 0, c.stream, A,
                         static_cast<const double*>(c.r), 1.0 / norm, dinv, gbuf[1], c.pa32, c.d32,
                         parts[1]);
    });
    long long launched = 0;
    const long long inner = run_inner(c, limit, [&](long long it) {
      // iteration it reads buffer it % 2 and writes the other
      const int in = static_cast<int>(it & 1);
      tracked(c, MPS_K_CG_ITER32, b_iter, [&] {
        hipLaunchKernelGGL(k_cgf_iter<Op>, dim3(L.grid), dim3(NT), 0, c.stream, A,
                           static_cast<const float4*>(gbuf[in]), gbuf[in ^ 1], c.pa32, c.d32,
                           static_cast<const double*>(parts[in]), L.grid, parts[in ^ 1], c.st32,
                           sweep_tol, it);
      }, it);
      launched = it;
    });
    // launches after convergence did no work: count them out
    uncount(c, MPS_K_CG_ITER32, launched - inner, b_iter);
    return inner;
  }, [&](double norm) {
    hipLaunchKernelGGL(k_ir_correct, dim3(grid_for(n)), dim3(256), 0, c.stream, n, c.sol,
                       static_cast<const float*>(c.d32), norm);
  });
}

// Multilevel preconditioner: two kernels per iteration (see k_ml_prec).
void ensure_zrec(Context& c, size_t n) {
  if (c.zrec_cap < n) {
    dfree(c.zrec);
    c.zrec = dalloc<uint4>(n);
    c.zrec_cap = n;
  }
}

MatFreeRec matfree_rec(const Context& c, size_t rpb) {
  return MatFreeRec{c.slice_off, c.scol, c.slice_off, c.col16, c.slice_wide, c.row_len, c.zrec, c.ghost32,
                    static_cast<float>(3.0 * c.env.r_lap / c.env.n0_lap * c.qpos_inv_h * c.qpos_inv_h *
                                       c.qpos_inv_h),
                    c.ppe_n, rpb};
}

// After ml_setup_ahead: the main stream waits for the auxiliary stream's
// set-up; false if none is pending.
bool join_ml_setup(Context& c) {
  if (!c.ml_setup_pending) return false;
  c.ml_setup_pending = false;
  MPS_HIP(hipStreamWaitEvent(c.stream, c.ev_join, 0));
  return true;
}

// The multilevel preconditioner's grid of aggregates: cells, 2x2 and 4x4
// blocks of cells.
struct MlGeom {
  uint32_t gnx, gnz, g1x, g1z, bnx, bnz;
  size_t off1, off2, total_aggs, nblk, sgrid;
};

MlGeom ml_geom(const Context& c) {
  MlGeom g;
  g.gnx = static_cast<uint32_t>(c.gnx);
  g.gnz = static_cast<uint32_t>(c.gnz);
  g.g1x = (g.gnx + 1) / 2;
  g.g1z = (g.gnz + 1) / 2;
  g.bnx = (g.gnx + 3) / 4;
  g.bnz = (g.gnz + 3) / 4;
  g.off1 = static_cast<size_t>(g.gnx) * g.gnz;
  g.off2 = g.off1 + static_cast<size_t>(g.g1x) * g.g1z;
  g.total_aggs = g.off2 + static_cast<size_t>(g.bnx) * g.bnz;
  g.nblk = static_cast<size_t>(g.bnx) * g.bnz;
  g.sgrid = (g.nblk + NTP / 64 - 1) / (NTP / 64);  // one wavefront per 4x4 block of cells
  return g;
}

CellGrid ml_grid(const Context& c, const MlGeom& g) {
  return CellGrid{c.cell_start, c.cell_count, c.sorted, c.key, c.type, g.gnx, g.gnz, g.g1z, g.bnx, g.bnz,
                  g.off1, g.off2};
}

// The multilevel solver's set-up on c.stream, which does not depend on the
// initial guess: Jacobi factors, the aggregate sums of A from the FP64 system
// (row sums in z, r, d as scratch, unless the fill of this system summed the
// rows already), and the list of the 4x4 cell blocks that hold particles (its
// length reaches the host with the solver's first synchronisation).
void ml_setup(Context& c) {
  const size_t n = c.ppe_n;
  const Launch L = cg_grid(n);
  hipLaunchKernelGGL(k_jacobi_inv, dim3(grid_for(n)), dim3(256), 0, c.stream, n,
                     static_cast<const double*>(c.diag), c.dinv32);
  const MlGeom g = ml_geom(c);
  if (c.agg_cap < g.total_aggs) {
    dfree(c.agg_inv);
    c.agg_inv = dalloc<float>(g.total_aggs);
    c.agg_cap = g.total_aggs;
  }
  if (c.part_ml_cap < g.sgrid) {
    dfree(c.part_ml);
    c.part_ml = dalloc<double>(2 * g.sgrid);
    c.part_ml_cap = g.sgrid;
  }
  if (c.ml_blk_cap < g.nblk) {
    dfree(c.ml_flag);
    dfree(c.ml_off);
    dfree(c.ml_list);
    c.ml_flag = dalloc<uint32_t>(g.nblk);
    c.ml_off = dalloc<uint32_t>(g.nblk);
    c.ml_list = dalloc<uint32_t>(g.nblk);
    c.ml_blk_cap = g.nblk;
  }
  if (!c.ml_active) c.ml_active = dalloc<uint64_t>(1);
  const CellGrid G = ml_grid(c, g);
  const bool sums = c.ml_rowsums_ready;
  c.ml_rowsums_ready = false;
  tracked(c, MPS_K_ML_SETUP, (sums ? 0.0 : c.nnz_offdiag * 12.0 + n * 17.0) + n * 20.0 + g.total_aggs * 4.0 +
                                 g.nblk * 12.0, [&] {
    if (!sums)
      hipLaunchKernelGGL(k_ml_rowsum, dim3(L.grid), dim3(NT), 0, c.stream,
                         sell<double>(c, c.sval, c.diag, L.rpb), static_cast<const uint32_t*>(c.row_len),
                         static_cast<const uint32_t*>(c.key), static_cast<const uint8_t*>(c.type), g.gnz, c.z32,
                         c.r32, c.d32);
    hipLaunchKernelGGL(k_ml_setup, dim3(static_cast<unsigned>(g.sgrid)), dim3(NTP), 0, c.stream, G,
                       static_cast<const float*>(c.z32), static_cast<const float*>(c.r32),
                       static_cast<const float*>(c.d32), c.agg_inv);
    hipLaunchKernelGGL(k_ml_active, dim3(grid_for(g.nblk)), dim3(256), 0, c.stream, G, c.ml_flag);
    scan_counts(c, c.ml_flag, g.nblk, c.ml_off, c.ml_active);
    hipLaunchKernelGGL(k_ml_compact, dim3(grid_for(g.nblk)), dim3(256), 0, c.stream, static_cast<uint32_t>(g.nblk),
                       static_cast<const uint32_t*>(c.ml_flag), static_cast<const uint32_t*>(c.ml_off),
                       c.ml_list);
  });
  MPS_HIP(hipMemcpyAsync(c.ml_active_host, c.ml_active, sizeof(uint64_t), hipMemcpyDeviceToHost, c.stream));
  check_launch();
}

void solve_mixed_ml(Context& c, double b_base, const StopRule& sr, long long max_iter) {
  const size_t n = c.ppe_n;
  const Launch L = cg_grid(n);
  ensure_zrec(c, n);
  const MatFreeRec A = matfree_rec(c, L.rpb);
  float* const zval = reinterpret_cast<float*>(c.zrec) + 2;  // z_i at zval[4 i]
  if (!join_ml_setup(c)) ml_setup(c);  // enqueued ahead on the auxiliary stream, or now
  const MlGeom g = ml_geom(c);
  const size_t stride = c.part_ml_cap;
  const CellGrid G = ml_grid(c, g);
  // rows outside every cell keep r = d = z = 0; the records take the positions
  hipLaunchKernelGGL(k_rec_init, dim3(grid_for(n)), dim3(256), 0, c.stream, n, static_cast<const uint2*>(c.qpos),
                     c.zrec, c.r32, c.d32);
  check_launch();
  double* p_pq = c.part3b;
  const long long limit = inner_limit(c, max_iter);
  // per particle: r (read, written), q, z; per slot of the packed layout: id
  // and Jacobi factor; per lane: the aggregate scales (set once nact is known)
  double b_prec = n * 16.0;
  uint32_t nact = 0;
  size_t pgrid = 1;
  bool packed = false;
  auto prec = [&](double norm, long long it) {
    const PrecArgs a{G,      c.ml_list, nact,   c.ml_pid, c.ml_dv, c.ml_sc,   c.r,       1.0 / norm,
                     c.r32,  c.d32,     c.q32,  c.dinv32, zval,    4u,        p_pq,      L.grid,
                     c.part_ml, stride, c.st32};
    hipLaunchKernelGGL(k_ml_prec<0>, dim3(static_cast<unsigned>(pgrid)), dim3(NTP), 0, c.stream, a, it);
  };
  double b_spmv = b_base;
  c.replay_ready = MPS_K_CG_SPMV32;
  refine(c, sr, [&](double norm, double sweep_tol) {
    // refine() has synchronised the stream (residual) before every sweep: the
    // block count and the wide nnz are on the host
    nact = static_cast<uint32_t>(*c.ml_active_host);
    b_spmv = b_base + 2.0 * mf_nnz_wide(c);
    pgrid = std::max<size_t>(1, (nact + NTP / 64 - 1) / (NTP / 64));
    b_prec = n * 16.0 + static_cast<double>(nact) * 64 * (8.0 * kPerLane + 16);
    if (!packed) {
      const size_t slots = static_cast<size_t>(std::max<uint32_t>(nact, 1)) * 64;
      if (c.ml_pack_cap < slots) {
        dfree(c.ml_pid);
        dfree(c.ml_dv);
        dfree(c.ml_sc);
        c.ml_pid = dalloc<uint32_t>(slots * kPerLane);
        c.ml_dv = dalloc<float>(slots * kPerLane);
        c.ml_sc = dalloc<float4>(slots);
        c.ml_pack_cap = slots;
      }
      tracked(c, MPS_K_ML_SETUP, n * 13.0 + slots * (8.0 * kPerLane + 16), [&] {
        hipLaunchKernelGGL(k_ml_pack, dim3(static_cast<unsigned>(pgrid)), dim3(NTP), 0, c.stream, G,
                           static_cast<const uint32_t*>(c.ml_list), nact, static_cast<const float*>(c.agg_inv),
                           static_cast<const float*>(c.dinv32), c.ml_pid, c.ml_dv, c.ml_sc,
                           static_cast<const uint32_t*>(nullptr));
      });
      packed = true;
    }
    MPS_HIP(hipMemsetAsync(c.st32, 0, sizeof(CgStatus), c.stream));
    tracked(c, MPS_K_CG_INIT32, n * 25.0, [&] { prec(norm, 0); });
    long long launched = 0;
    const long long inner = run_inner(c, limit, [&](long long it) {
      tracked(c, MPS_K_CG_SPMV32, b_spmv, [&] {
        const SpmvArgs<MatFreeRec> a{A,      nullptr, c.pa32,  c.q32, c.d32, c.part_ml, stride,
                                     static_cast<int>(pgrid), p_pq, c.st32, sweep_tol};
        if (it == 1)
          hipLaunchKernelGGL((k_ml_spmv<true, MatFreeRec>), dim3(L.grid), dim3(NT), 0, c.stream, a, it);
        else
          hipLaunchKernelGGL((k_ml_spmv<false, MatFreeRec>), dim3(L.grid), dim3(NT), 0, c.stream, a, it);
      }, it);
      tracked(c, MPS_K_ML_PREC, b_prec, [&] { prec(norm, it); }, it);
      launched = it;
    });
    if (!c.st32_host->converged && launched > 0)
      hipLaunchKernelGGL(k_ml_finish, dim3(grid_for(n)), dim3(256), 0, c.stream, n,
                         static_cast<const CgStatus*>(c.st32), static_cast<const float*>(c.pa32), c.d32);
    // launches after the stop did no work: count them out (k_ml_spmv(inner + 1)
    // only finished d)
    uncount(c, MPS_K_CG_SPMV32, launched - inner, b_spmv);
    uncount(c, MPS_K_ML_PREC, launched - inner, b_prec);
    return inner;
  }, [&](double norm) {
    hipLaunchKernelGGL(k_ir_correct, dim3(grid_for(n)), dim3(256), 0, c.stream, n, c.sol,
                       static_cast<const float*>(c.d32), norm);
  });
}

// ---- cell-grid multigrid (MPS_PRECOND_AMG) ----

// Capacity for P_s's rows of nslots packed slots (q9, then q9c) in c.amg_q9.
void ensure_q9(Context& c, size_t nslots) {
  const size_t m = std::max<size_t>(nslots, 64);
  const size_t words = m + (m + 7) / 8;  // uint4 words: the rows, then the ninth coefficients
  if (c.amg_q9_cap < words) {
    dfree(c.amg_q9);
    c.amg_q9 = dalloc<uint4>(words);
    c.amg_q9_cap = words;
  }
  c.amg_q9c = reinterpret_cast<uint16_t*>(c.amg_q9 + m);
}

AmgGrid amg_grid(const Context& c) {
  AmgGrid g{};
  g.levels = c.amg_levels;
  for (int l = 0; l < c.amg_levels; l++) g.L[l] = AmgLevel{c.amg_nx[l], c.amg_nz[l], c.amg_off[l]};
  g.S = c.amg_S;
  g.dinv = c.amg_dinv;
  g.r = c.amg_r;
  g.x = c.amg_x;
  g.e = c.amg_e;
  g.inv = c.amg_inv;
  g.omega = kAmgOmega;
  g.corr = kAmgCorr;
  return g;
}

// The first level the single-workgroup kernel takes: the points of that
// level and all below fit its LDS (kAmgLdsPts).
int amg_block_level(const Context& c) {
  const size_t end = c.amg_pts;
  int l0 = c.amg_levels - 1;
  while (l0 > 0 && end - c.amg_off[l0 - 1] <= static_cast<size_t>(kAmgLdsPts - 4)) l0--;
  return l0;
}

// Levels and operators from the assembled FP64 system and the current cells
// (on c.stream): Jacobi factors, the occupied 4x4 cell blocks (their count
// reaches the host with the solver's first synchronisation), the rows'
// per-direction sums, the cells' Galerkin stencil, the coarser stencils, their
// FP32 copies and 1 / diagonal, and the coarsest inverse.
void amg_setup(Context& c) {
  const size_t n = c.ppe_n;
  const Launch L = cg_grid(n);
  hipLaunchKernelGGL(k_jacobi_inv, dim3(grid_for(n)), dim3(256), 0, c.stream, n,
                     static_cast<const double*>(c.diag), c.dinv32);
  // the occupied 4x4 blocks of cells, as the multilevel set-up lists them
  const MlGeom mg = ml_geom(c);
  if (c.ml_blk_cap < mg.nblk) {
    dfree(c.ml_flag);
    dfree(c.ml_off);
    dfree(c.ml_list);
    c.ml_flag = dalloc<uint32_t>(mg.nblk);
    c.ml_off = dalloc<uint32_t>(mg.nblk);
    c.ml_list = dalloc<uint32_t>(mg.nblk);
    c.ml_blk_cap = mg.nblk;
  }
  if (!c.ml_active) c.ml_active = dalloc<uint64_t>(1);
  const CellGrid G = ml_grid(c, mg);
  // level shapes
  amg_box(c);
  int lv = 0;
  uint32_t nx = c.amg_box[2], nz = c.amg_box[3];
  size_t off = 0;
  for (;;) {
    if (lv == kAmgMaxLevels) throw Failure{MPS_ERR_BAD_ARG, "cell grid too large for the multigrid levels"};
    c.amg_nx[lv] = nx;
    c.amg_nz[lv] = nz;
    c.amg_off[lv] = static_cast<uint32_t>(off);
    off += static_cast<size_t>(nx) * nz;
    off = (off + 3) & ~static_cast<size_t>(3);  // every level 16-byte aligned (k_amg_block copies float4)
    lv++;
    if (static_cast<size_t>(nx) * nz <= static_cast<size_t>(kAmgDense)) break;
    nx = (nx + 1) / 2;
    nz = (nz + 1) / 2;
  }
  if (off > 0xffffffffull) throw Failure{MPS_ERR_BAD_ARG, "cell grid too large for the multigrid levels"};
  c.amg_levels = lv;
  c.amg_pts = off;
  if (c.amg_cap < off) {
    dfree(c.amg_S64);
    dfree(c.amg_S);
    dfree(c.amg_dinv);
    dfree(c.amg_r);
    dfree(c.amg_x);
    dfree(c.amg_e);
    c.amg_S64 = dalloc<double>(9 * off);
    // (room for k_amg_block's float4 copies rounding the last level up)
    c.amg_S = dalloc<float>(9 * off + 16);
    c.amg_dinv = dalloc<float>(off + 16);
    c.amg_r = dalloc<float>(off + 16);
    c.amg_x = dalloc<float>(off);
    c.amg_e = dalloc<float>(off);
    c.amg_cap = off;
  }
  if (!c.amg_inv) c.amg_inv = dalloc<float>(kAmgDense * kAmgDense);
  ensure_amg_rows(c, n);
  const bool rows_ready = c.amg_rows_ready;  // left by the fill
  const size_t npts0 = static_cast<size_t>(c.amg_box[2]) * c.amg_box[3];
  tracked(c, MPS_K_ML_SETUP, (rows_ready ? 0.0 : c.nnz_offdiag * 16.0 + n * 17.0) + n * 72.0 + npts0 * 8.0 +
                                 off * 9 * 12.0, [&] {
    hipLaunchKernelGGL(k_ml_active, dim3(grid_for(mg.nblk)), dim3(256), 0, c.stream, G, c.ml_flag);
    scan_counts(c, c.ml_flag, mg.nblk, c.ml_off, c.ml_active);
    hipLaunchKernelGGL(k_ml_compact, dim3(grid_for(mg.nblk)), dim3(256), 0, c.stream, static_cast<uint32_t>(mg.nblk),
                       static_cast<const uint32_t*>(c.ml_flag), static_cast<const uint32_t*>(c.ml_off), c.ml_list);
    if (!rows_ready)
      hipLaunchKernelGGL(k_amg_rows, dim3(L.grid), dim3(NT), 0, c.stream, sell<double>(c, c.sval, c.diag, L.rpb),
                         static_cast<const uint32_t*>(c.row_len), static_cast<const uint32_t*>(c.key),
                         static_cast<const uint8_t*>(c.type), static_cast<uint32_t>(c.gnz), c.amg_rows);
    hipLaunchKernelGGL(k_amg_cells, dim3(grid_for(npts0)), dim3(256), 0, c.stream, c.amg_box[0], c.amg_box[1],
                       c.amg_box[2], c.amg_box[3], static_cast<uint32_t>(c.gnz),
                       static_cast<const uint32_t*>(c.cell_start), static_cast<const uint32_t*>(c.cell_count),
                       static_cast<const uint32_t*>(c.sorted), static_cast<const double*>(c.amg_rows), n, c.amg_S64);
    for (int l = 0; l + 1 < lv; l++)
      hipLaunchKernelGGL(k_amg_coarsen, dim3(grid_for(static_cast<size_t>(c.amg_nx[l + 1]) * c.amg_nz[l + 1])),
                         dim3(256), 0, c.stream, static_cast<const double*>(c.amg_S64 + 9ull * c.amg_off[l]),
                         c.amg_nx[l], c.amg_nz[l], c.amg_S64 + 9ull * c.amg_off[l + 1], c.amg_nx[l + 1],
                         c.amg_nz[l + 1]);
    hipLaunchKernelGGL(k_amg_f32_all, dim3(grid_for(off)), dim3(256), 0, c.stream, amg_grid(c),
                       static_cast<const double*>(c.amg_S64), off, c.amg_S, c.amg_dinv);
    hipLaunchKernelGGL(k_amg_dense, dim3(1), dim3(1024), 0, c.stream,
                       static_cast<const double*>(c.amg_S64 + 9ull * c.amg_off[lv - 1]), c.amg_nx[lv - 1],
                       c.amg_nz[lv - 1], c.amg_inv);
    // (level 0's right-hand side needs no clearing: k_amg_rfix writes every
    // level-0 point each cycle, from direction sums cleared per solve)
  });
  MPS_HIP(hipMemcpyAsync(c.ml_active_host, c.ml_active, sizeof(uint64_t), hipMemcpyDeviceToHost, c.stream));
  check_launch();
}

// One V-cycle on the cell-grid levels: level 0's right-hand side in, its
// result (the correction per cell) out.  The level-0 up-pass is left to the
// caller when `fold0` (k_ml_prec_up does it) and the cycle has a level below
// the cells' outside the single-workgroup kernel; returns whether it was left.
// The cycle of inner iteration `it` of the solve with status `st` does
// nothing once that solve has stopped before it (launches past the stop).
bool amg_cycle(Context& c, bool fold0, const CgStatus* st, long long it) {
  AmgGrid g = amg_grid(c);
  g.st = st;
  g.it = it;
  const int l0 = amg_block_level(c);
  // the kAmgChainLevels levels above the single-workgroup ones in one launch
  // each way (level 0 keeps its own kernels)
  const int l1 = (l0 - kAmgChainLevels >= 1) ? l0 - kAmgChainLevels : l0;
  const unsigned tiles = ((c.amg_nx[l0] + kAmgChainT - 1) / kAmgChainT) * ((c.amg_nz[l0] + kAmgChainT - 1) / kAmgChainT);
  for (int l = 0; l < l1; l++)
    hipLaunchKernelGGL(k_amg_down, dim3(grid_for(static_cast<size_t>(c.amg_nx[l + 1]) * c.amg_nz[l + 1])),
                       dim3(256), 0, c.stream, g, l);
  if (l1 < l0)
    hipLaunchKernelGGL((k_amg_down_chain<kAmgChainT, kAmgChainLevels>), dim3(tiles), dim3(kAmgChainThreads), 0,
                       c.stream, g, l1);
  hipLaunchKernelGGL(k_amg_block, dim3(1), dim3(kAmgBlockThreads), 0, c.stream, g, l0);
  if (l1 < l0)
    hipLaunchKernelGGL((k_amg_up_chain<kAmgChainT, kAmgChainLevels>), dim3(tiles), dim3(kAmgChainThreads), 0,
                       c.stream, g, l1);
  const bool left = fold0 && l0 >= 1;
  for (int l = l1 - 1; l >= (left ? 1 : 0); l--)
    hipLaunchKernelGGL(k_amg_up, dim3(grid_for(static_cast<size_t>(c.amg_nx[l]) * c.amg_nz[l])), dim3(256), 0,
                       c.stream, g, l);
  return left;
}

void solve_mixed_amg(Context& c, double b_base, const StopRule& sr, long long max_iter) {
  const size_t n = c.ppe_n;
  const Launch L = cg_grid(n);
  ensure_zrec(c, n);
  const MatFreeRec A = matfree_rec(c, L.rpb);
  float* const zval = reinterpret_cast<float*>(c.zrec) + 2;  // z_i at zval[4 i]
  amg_setup(c);
  const MlGeom g = ml_geom(c);
  if (c.part_ml_cap < g.sgrid) {
    dfree(c.part_ml);
    c.part_ml = dalloc<double>(2 * g.sgrid);
    c.part_ml_cap = g.sgrid;
  }
  const size_t stride = c.part_ml_cap;
  const CellGrid G = ml_grid(c, g);
  hipLaunchKernelGGL(k_rec_init, dim3(grid_for(n)), dim3(256), 0, c.stream, n, static_cast<const uint2*>(c.qpos),
                     c.zrec, c.r32, c.d32);
  check_launch();
  double* p_pq = c.part3b;
  const long long limit = inner_limit(c, max_iter);
  uint32_t nact = 0;
  size_t pgrid = 1;
  bool packed = false;
  // per pass over the particles (k_ml_prec<1>, then k_ml_prec<2>): r, q or z
  // per particle, the packed ids and Jacobi factors per slot, P_s's rows (18
  // bytes per particle); the cycle: level 0's direction sums written and
  // gathered (9 + 9 + 1 floats per point), then r, x, e and the stencils of
  // every level, twice
  double b_pass = n * 30.0;
  const size_t npts0 = static_cast<size_t>(c.amg_box[2]) * c.amg_box[3];
  const double b_cycle = static_cast<double>(c.amg_pts) * 2 * (36 + 16) + npts0 * 19.0 * 4;
  const size_t tq_stride = (npts0 + 63) & ~static_cast<size_t>(63);
  if (c.amg_tq_cap < 9 * tq_stride) {
    dfree(c.amg_tq);
    c.amg_tq = dalloc<float>(9 * tq_stride);
    c.amg_tq_cap = 9 * tq_stride;
  }
  // the direction sums of cells in unlisted (empty) blocks stay zero
  MPS_HIP(hipMemsetAsync(c.amg_tq, 0, 9 * tq_stride * sizeof(float), c.stream));
  // it == 0 (the sweep's start) is accounted as a whole by the caller
  auto prec = [&](double norm, long long it) {
    PrecArgs a{G,         c.ml_list, nact,  c.ml_pid, c.ml_dv,    c.ml_sc, c.r,   1.0 / norm, c.r32,
               c.d32,     c.q32,     c.dinv32, zval, 4u,         p_pq,    L.grid, c.part_ml,  stride,
               c.st32,    c.amg_r,   c.amg_e,   kAmgOmegaRows, c.row_len, c.amg_box[0], c.amg_box[1],
               c.amg_box[2], c.amg_box[3]};
    a.q9 = c.amg_q9;
    a.q9c = c.amg_q9c;
    a.tq = c.amg_tq;
    a.tq_stride = tq_stride;
    auto track = [&](int kid, double bytes, auto&& f) {
      if (it == 0)
        f();
      else
        tracked(c, kid, bytes, f, it);
    };
    track(MPS_K_ML_PREC, b_pass, [&] {
      hipLaunchKernelGGL(k_ml_prec<1>, dim3(static_cast<unsigned>(pgrid)), dim3(NTP), 0, c.stream, a, it);
    });
    bool left = false;
    track(MPS_K_AMG_CYCLE, b_cycle, [&] {
      hipLaunchKernelGGL(k_amg_rfix, dim3(grid_for(npts0)), dim3(256), 0, c.stream,
                         static_cast<const float*>(c.amg_tq), tq_stride, c.amg_box[2], c.amg_box[3], c.amg_r,
                         static_cast<const CgStatus*>(c.st32), it);
      left = amg_cycle(c, true, c.st32, it);
    });
    track(MPS_K_ML_PREC, b_pass, [&] {
      if (left)
        hipLaunchKernelGGL(k_ml_prec_up, dim3(static_cast<unsigned>(pgrid)), dim3(NTP), 0, c.stream, a, it,
                           amg_grid(c));
      else
        hipLaunchKernelGGL(k_ml_prec<2>, dim3(static_cast<unsigned>(pgrid)), dim3(NTP), 0, c.stream, a, it);
    });
  };
  double b_spmv = b_base;
  c.replay_ready = MPS_K_CG_SPMV32;
  refine(c, sr, [&](double norm, double sweep_tol) {
    // refine() has synchronised the stream before every sweep: the block
    // count and the wide nnz are on the host
    nact = static_cast<uint32_t>(*c.ml_active_host);
    b_spmv = b_base + 2.0 * mf_nnz_wide(c);
    pgrid = std::max<size_t>(1, (nact + NTP / 64 - 1) / (NTP / 64));
    b_pass = n * 30.0 + static_cast<double>(nact) * 64 * (8.0 * kPerLane + 8);
    if (!packed) {
      const size_t slots = static_cast<size_t>(std::max<uint32_t>(nact, 1)) * 64;
      if (c.ml_pack_cap < slots) {
        dfree(c.ml_pid);
        dfree(c.ml_dv);
        dfree(c.ml_sc);
        c.ml_pid = dalloc<uint32_t>(slots * kPerLane);
        c.ml_dv = dalloc<float>(slots * kPerLane);
        c.ml_sc = dalloc<float4>(slots);
        c.ml_pack_cap = slots;
      }
      tracked(c, MPS_K_ML_SETUP, n * 17.0 + slots * (8.0 * kPerLane + 16), [&] {
        hipLaunchKernelGGL(k_ml_pack, dim3(static_cast<unsigned>(pgrid)), dim3(NTP), 0, c.stream, G,
                           static_cast<const uint32_t*>(c.ml_list), nact, static_cast<const float*>(nullptr),
                           static_cast<const float*>(c.dinv32), c.ml_pid, c.ml_dv, c.ml_sc,
                           static_cast<const uint32_t*>(c.row_len));
      });
      // P_s's rows in the same slot order
      const size_t nslots = static_cast<size_t>(nact) * kPerLane * 64;
      ensure_q9(c, nslots);
      if (nslots > 0)
        tracked(c, MPS_K_ML_SETUP, nslots * (4.0 + 18.0) + n * 80.0, [&] {
          hipLaunchKernelGGL(k_amg_q, dim3(grid_for(nslots)), dim3(256), 0, c.stream, nslots,
                             static_cast<const uint32_t*>(c.ml_pid), static_cast<const double*>(c.amg_rows),
                             static_cast<const double*>(c.diag), kAmgOmegaP, c.amg_q9, c.amg_q9c);
        });
      packed = true;
    }
    MPS_HIP(hipMemsetAsync(c.st32, 0, sizeof(CgStatus), c.stream));
    tracked(c, MPS_K_CG_INIT32, n * 25.0, [&] { prec(norm, 0); });
    long long launched = 0;
    const long long inner = run_inner(c, limit, [&](long long it) {
      tracked(c, MPS_K_CG_SPMV32, b_spmv, [&] {
        const SpmvArgs<MatFreeRec> a{A,      nullptr, c.pa32,  c.q32, c.d32, c.part_ml, stride,
                                     static_cast<int>(pgrid), p_pq, c.st32, sweep_tol};
        if (it == 1)
          hipLaunchKernelGGL((k_ml_spmv<true, MatFreeRec>), dim3(L.grid), dim3(NT), 0, c.stream, a, it);
        else
          hipLaunchKernelGGL((k_ml_spmv<false, MatFreeRec>), dim3(L.grid), dim3(NT), 0, c.stream, a, it);
      }, it);
      prec(norm, it);
      launched = it;
    });
    if (!c.st32_host->converged && launched > 0)
      hipLaunchKernelGGL(k_ml_finish, dim3(grid_for(n)), dim3(256), 0, c.stream, n,
                         static_cast<const CgStatus*>(c.st32), static_cast<const float*>(c.pa32), c.d32);
    uncount(c, MPS_K_CG_SPMV32, launched - inner, b_spmv);
    uncount(c, MPS_K_ML_PREC, 2 * (launched - inner), b_pass);
    uncount(c, MPS_K_AMG_CYCLE, launched - inner, b_cycle);
    return inner;
  }, [&](double norm) {
    hipLaunchKernelGGL(k_ir_correct, dim3(grid_for(n)), dim3(256), 0, c.stream, n, c.sol,
                       static_cast<const float*>(c.d32), norm);
  });
}

MatFree matfree(const Context& c, size_t rpb) {
  return MatFree{c.slice_off, c.scol, c.slice_off, c.col16, c.slice_wide, c.row_len, c.qpos, c.ghost32,
                 static_cast<float>(3.0 * c.env.r_lap / c.env.n0_lap * c.qpos_inv_h * c.qpos_inv_h *
                                    c.qpos_inv_h),
                 c.ppe_n, rpb};
}

void solve_mixed(Context& c, const StopRule& sr, long long max_iter) {
  const size_t n = c.ppe_n;
  const Launch L = cg_grid(n);
  // per row and iteration of the one-kernel CG: the packed (r, w, s, M) read
  // and written, p and d read and written
  const double vec = 16.0 * 2 + 4.0 * 4;
  if (!c.ppe_from_particles) {
    c.stats.precond_used = c.solver.precond == MPS_PRECOND_NONE ? MPS_PRECOND_NONE : MPS_PRECOND_JACOBI;
    c.pred_kind = static_cast<int>(c.stats.precond_used);
    ppe_make_fp32(c);
    solve_mixed_cg(c, sell<float>(c, c.sval32, c.diag32, L.rpb), spmv_bytes(c, 4, vec), false, sr, max_iter);
    return;
  }
  const MatFree A = matfree(c, L.rpb);
  // int16 columns of narrow slices, u32 of wide ones (2 more bytes each,
  // added by the solvers); row length, ghost and position of every row; slice tables
  const double cols = c.nnz_offdiag * 2.0 + c.nslices * 25.0;
  const double rows = n * (4 + 4 + 8.0);
  const bool cells_match = c.have_neighbors && c.ppe_search_serial == c.search_serial && c.ppe_n == c.n;
  const int pc = effective_precond(c);
  c.stats.precond_used = cells_match && (pc == MPS_PRECOND_MULTILEVEL || pc == MPS_PRECOND_AMG)
                             ? pc
                             : (pc == MPS_PRECOND_NONE ? MPS_PRECOND_NONE : MPS_PRECOND_JACOBI);
  c.pred_kind = static_cast<int>(c.stats.precond_used);
  if (pc == MPS_PRECOND_MULTILEVEL && cells_match)
    // row length and ghost; the own record {position, z}; p, q, d read and written
    solve_mixed_ml(c, cols + n * (4 + 4 + 16 + 24.0), sr, max_iter);
  else if (pc == MPS_PRECOND_AMG && cells_match)
    solve_mixed_amg(c, cols + n * (4 + 4 + 16 + 24.0), sr, max_iter);
  else
    solve_mixed_cg(c, A, cols + rows + n * vec, true, sr, max_iter);
}


// ---------------- batched solves: several scenes per launch ----------------
//
// The mixed multilevel solve of several scenes at once: the scenes' systems
// stay in their own buffers; every kernel of the solve is one grouped launch
// whose grid concatenates the grids the scenes' own launches would use
// (Group), so each scene's blocks compute exactly what its own launch would,
// with its own alpha, beta and stop test (the block partial sums are per
// scene).  A scene that stops, or is not part of the running phase, has no
// blocks in the launch.  The refinement logic of `refine` runs per scene on
// the host between phases.

struct ResidArgs {
  Sell<double> A;
  const double* b;
  const double* x;
  double* r;
  double* part_rr;
  double* part_bb;
  double* out;  // |r|^2, |b|^2
  int nparts;
};

struct VecArgs {
  size_t n;
  double* x;
  float* d;
  const float* p;
  CgStatus* st;
  double norm;
};

__global__ void __launch_bounds__(NT) kg_residual(const ResidArgs* __restrict__ tab, Group g) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  const ResidArgs a = tab[s];
  cg_init_body(a.A, a.b, a.x, a.r, a.part_rr, a.part_bb, f);
}

// One block per scene of the group.
__global__ void __launch_bounds__(NT) kg_reduce2(const ResidArgs* __restrict__ tab, Group g) {
  __shared__ double lds[NT / 64];
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  const ResidArgs a = tab[s];
  const double sa = reduce_partials<NT>(a.part_rr, a.nparts, lds);
  const double sb = reduce_partials<NT>(a.part_bb, a.nparts, lds);
  if (threadIdx.x == 0) {
    a.out[0] = sa;
    a.out[1] = sb;
  }
}

__global__ void __launch_bounds__(NTP) kg_ml_prec(const PrecArgs* __restrict__ tab, Group g, long long it) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  const PrecArgs a = tab[s];
  ml_prec_body<0>(a, it, f);
}

// The cell-grid multigrid of one scene in grouped launches: its levels, its
// single-workgroup level and pass 1's direction sums (k_amg_rfix's input).
// The grouped cycle runs every level above l0 with the per-level kernels
// (the chains compute the same values, one launch per scene per chain).
struct AmgArgs {
  AmgGrid g;
  int l0;
  const float* tq;
  size_t tq_stride;
};
template <int MODE>
__global__ void __launch_bounds__(NTP) kg_ml_prec_m(const PrecArgs* __restrict__ tab, Group g, long long it) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  const PrecArgs a = tab[s];
  ml_prec_body<MODE>(a, it, f);
}
__global__ void __launch_bounds__(NTP) kg_ml_prec_up(const PrecArgs* __restrict__ tab,
                                                     const AmgArgs* __restrict__ am, Group g, long long it) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  const PrecArgs a = tab[s];
  ml_prec_body<2>(a, it, f, &am[s].g);
}
__global__ void __launch_bounds__(256) kg_amg_rfix(const AmgArgs* __restrict__ am, Group g) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  const AmgArgs& A = am[s];
  const uint32_t nx = A.g.L[0].nx, nz = A.g.L[0].nz, p = f.bid * 256u + threadIdx.x;
  if (p < nx * nz) amg_rfix_point(A.tq, A.tq_stride, nx, nz, A.g.r + A.g.L[0].off, p);
}
__global__ void __launch_bounds__(256) kg_amg_down(const AmgArgs* __restrict__ am, Group g, int l) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  const AmgGrid& G = am[s].g;
  const uint32_t P = f.bid * 256u + threadIdx.x;
  if (P < amg_npts(G.L[l + 1])) amg_down_point(G, amg_global(G), AmgStenGlobal{&G}, l, P);
}
__global__ void __launch_bounds__(256) kg_amg_up(const AmgArgs* __restrict__ am, Group g, int l) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  const AmgGrid& G = am[s].g;
  const uint32_t p = f.bid * 256u + threadIdx.x;
  if (p < amg_npts(G.L[l])) amg_up_point(G, amg_global(G), AmgStenGlobal{&G}, l, p);
}
__global__ void __launch_bounds__(kAmgBlockThreads) kg_amg_block(const AmgArgs* __restrict__ am, Group g) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  amg_block_body(am[s].g, am[s].l0);
}

template <bool FIRST>
__global__ void __launch_bounds__(NT, 8) kg_ml_spmv(const SpmvArgs<MatFreeRec>* __restrict__ tab, Group g,
                                                    long long it) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  const SpmvArgs<MatFreeRec> a = tab[s];
  ml_spmv_body<FIRST, MatFreeRec>(a, it, f);
}

// d += alpha p of scenes whose sweep ended without a stop (k_ml_finish).
__global__ void __launch_bounds__(256) kg_ml_finish(const VecArgs* __restrict__ tab, Group g) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  const VecArgs a = tab[s];
  const size_t i = f.bid * 256ull + threadIdx.x;
  if (i < a.n) a.d[i] = fmaf(static_cast<float>(a.st->alpha_ml), a.p[i], a.d[i]);
}

// x += |r| d (k_ir_correct).
__global__ void __launch_bounds__(256) kg_ir_correct(const VecArgs* __restrict__ tab, Group g) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  const VecArgs a = tab[s];
  const size_t i = f.bid * 256ull + threadIdx.x;
  if (i < a.n) a.x[i] = fma(static_cast<double>(a.d[i]), a.norm, a.x[i]);
}

// The inner status of each scene of the group cleared (one block per scene).
__global__ void __launch_bounds__(64) kg_status_clear(const VecArgs* __restrict__ tab, Group g) {
  int s;
  Frame f;
  if (!group_frame(g, s, f)) return;
  if (threadIdx.x == 0) *tab[s].st = CgStatus{};
}

enum BatchKind {
  kResid = 0, kRed2, kPrec, kSpmv, kFinish, kCorrect, kClear,
  // the multigrid scenes: pass 1, pass 2 (with or without level 0's up
  // step), level 0's right-hand side, the single-workgroup levels, and the
  // down and up steps of each level
  kPrec1, kPrec2, kPrecUp, kRfix, kBlock, kDown0, kUp0 = kDown0 + kAmgMaxLevels,
  kKindCount = kUp0 + kAmgMaxLevels
};
static_assert(kKindCount <= BatchWork::kKinds, "block offset regions");
static_assert(sizeof(AmgArgs) <= BatchWork::kArgBytes, "multigrid table entries");
static_assert(sizeof(ResidArgs) <= BatchWork::kArgBytes && sizeof(PrecArgs) <= BatchWork::kArgBytes &&
                  sizeof(SpmvArgs<MatFreeRec>) <= BatchWork::kArgBytes && sizeof(VecArgs) <= BatchWork::kArgBytes,
              "argument tables");

struct SceneRun {
  Context* c;
  Launch L;
  size_t pgrid = 1, vgrid = 1;
  StopRule sr;
  double inner_tol = 1e-4;
  long long limit = 0, total_cap = 0, max_iter = 0;
  long long total = 0;
  int outer = 0;
  double rr0 = 0, bb = 0, rr = 0, rr_prev = 0, target = 0;
  bool refining = true, done = false, fallback = false;
  // the running sweep
  bool sweeping = false, swept = false;
  double norm = 0, sweep_tol = 0;
  long long inner = 0;
  double rr_last = -1;
  int flat = 0;
  // the cell-grid multigrid: its single-workgroup level, the blocks of level
  // 0's right-hand side and of each level's down (level l -> l + 1) and up
  // (level l) step
  bool amg = false;
  int l0 = 0;
  size_t rgrid = 1;
  size_t dgrid[kAmgMaxLevels] = {}, ugrid[kAmgMaxLevels] = {};
};

// Grouped launches of one batched solve: argument tables (4 regions: residual,
// preconditioner, SpMV, vector updates) and block offsets per kind.
struct Batch {
  BatchWork& w;
  std::vector<SceneRun>& sc;
  size_t S;
  template <class T>
  T* tab_host(int region) { return reinterpret_cast<T*>(w.tab_host + region * w.cap * BatchWork::kArgBytes); }
  template <class T>
  const T* tab_dev(int region) const {
    return reinterpret_cast<const T*>(w.tab + region * w.cap * BatchWork::kArgBytes);
  }
  void upload_tab(int region, size_t elem) {
    const size_t off = region * w.cap * BatchWork::kArgBytes;
    MPS_HIP(hipMemcpyAsync(w.tab + off, w.tab_host + off, S * elem, hipMemcpyHostToDevice, w.stream));
  }
  // offsets of `kind` over the scenes `in(s)`, each with blocks(s) blocks;
  // returns the launch's total block count
  template <class In, class Blocks>
  unsigned group(int kind, In in, Blocks blocks) {
    uint32_t* f = w.first_host + kind * (w.cap + 1);
    f[0] = 0;
    for (size_t s = 0; s < S; s++) f[s + 1] = f[s] + (in(s) ? static_cast<uint32_t>(blocks(s)) : 0u);
    MPS_HIP(hipMemcpyAsync(w.first + kind * (w.cap + 1), f, (S + 1) * 4, hipMemcpyHostToDevice, w.stream));
    return f[S];
  }
  Group g(int kind) const { return Group{w.first + kind * (w.cap + 1), nullptr, static_cast<int>(S)}; }
};

void batch_solve_run(std::vector<SceneRun>& sc, BatchWork& w) {
  const size_t S = sc.size();
  Batch B{w, sc, S};
  // static operands
  ResidArgs* rt = B.tab_host<ResidArgs>(0);
  PrecArgs* pt = B.tab_host<PrecArgs>(1);
  SpmvArgs<MatFreeRec>* st = B.tab_host<SpmvArgs<MatFreeRec>>(2);
  VecArgs* vt = B.tab_host<VecArgs>(3);
  double* red_dev = reinterpret_cast<double*>(w.tab + 4 * w.cap * BatchWork::kArgBytes);
  double* red_host = reinterpret_cast<double*>(w.tab_host + 4 * w.cap * BatchWork::kArgBytes);
  AmgArgs* at = B.tab_host<AmgArgs>(5);
  bool any_amg = false;
  for (size_t s = 0; s < S; s++) {
    Context& c = *sc[s].c;
    const Launch L = sc[s].L;
    const MlGeom g = ml_geom(c);
    rt[s] = ResidArgs{sell<double>(c, c.sval, c.diag, L.rpb), c.b, c.sol, c.r, c.partB, c.partC, red_dev + 2 * s,
                      L.grid};
    float* const zval = reinterpret_cast<float*>(c.zrec) + 2;
    pt[s] = PrecArgs{ml_grid(c, g), c.ml_list, static_cast<uint32_t>(*c.ml_active_host), c.ml_pid, c.ml_dv,
                     c.ml_sc, c.r, 1.0, c.r32, c.d32, c.q32, c.dinv32, zval, 4u, c.part3b, L.grid, c.part_ml,
                     c.part_ml_cap, w.st + s};
    st[s] = SpmvArgs<MatFreeRec>{matfree_rec(c, L.rpb), nullptr, c.pa32, c.q32, c.d32, c.part_ml, c.part_ml_cap,
                                 static_cast<int>(sc[s].pgrid), c.part3b, w.st + s, 1e-4};
    vt[s] = VecArgs{c.ppe_n, c.sol, c.d32, c.pa32, w.st + s, 1.0};
    if (sc[s].amg) {
      // pass 1 and 2 of the multigrid: no aggregate scales (k_ml_pack ran
      // with the multigrid's rules), level 0's box, P_s's rows
      pt[s].rc = c.amg_r;
      pt[s].ec = c.amg_e;
      pt[s].omega = kAmgOmegaRows;
      pt[s].row_len = c.row_len;
      pt[s].bx0 = c.amg_box[0];
      pt[s].bz0 = c.amg_box[1];
      pt[s].bnx = c.amg_box[2];
      pt[s].bnz = c.amg_box[3];
      pt[s].q9 = c.amg_q9;
      pt[s].q9c = c.amg_q9c;
      pt[s].tq = c.amg_tq;
      pt[s].tq_stride = c.amg_tq_stride;
      at[s] = AmgArgs{amg_grid(c), sc[s].l0, c.amg_tq, c.amg_tq_stride};
      any_amg = true;
    }
  }
  B.upload_tab(0, sizeof(ResidArgs));
  if (any_amg) B.upload_tab(5, sizeof(AmgArgs));
  auto launch_check = [&] {
    check_launch();
    w.launches++;
  };
  const int max_outer = 64;
  for (;;) {
    // FP64 residual r = b - A x of every scene still refining
    const unsigned nr = B.group(kResid, [&](size_t s) { return sc[s].refining; },
                                [&](size_t s) { return sc[s].L.grid; });
    if (nr == 0) break;
    const unsigned n2 = B.group(kRed2, [&](size_t s) { return sc[s].refining; }, [](size_t) { return 1; });
    hipLaunchKernelGGL(kg_residual, dim3(nr), dim3(NT), 0, w.stream, B.tab_dev<ResidArgs>(0), B.g(kResid));
    launch_check();
    hipLaunchKernelGGL(kg_reduce2, dim3(n2), dim3(NT), 0, w.stream, B.tab_dev<ResidArgs>(0), B.g(kRed2));
    launch_check();
    MPS_HIP(hipMemcpyAsync(red_host, red_dev, 2 * S * sizeof(double), hipMemcpyDeviceToHost, w.stream));
    MPS_HIP(hipStreamSynchronize(w.stream));
    // the refinement decisions of `refine`, per scene
    bool any = false;
    for (size_t s = 0; s < S; s++) {
      SceneRun& r = sc[s];
      r.swept = false;
      if (!r.refining) continue;
      r.rr = red_host[2 * s];
      r.bb = red_host[2 * s + 1];
      if (r.outer == 0) r.rr0 = r.rr;
      const bool stalled = r.outer > 0 && !(r.rr < 0.5 * r.rr_prev);
      r.rr_prev = r.rr;
      if (r.sr.rule == MPS_STOP_RHS_NORM) {
        r.target = (r.sr.tol * r.sr.tol) * r.bb;
        r.done = r.rr <= r.target;
      } else {
        r.target = r.rr0 * r.sr.tol * r.sr.tol;
        r.done = (r.outer == 0) ? (r.target == 0) : (r.rr < r.target);
      }
      if (stalled && !r.done && r.rr == r.rr && r.total < r.total_cap) {
        r.fallback = true;
        r.refining = false;
        continue;
      }
      if (r.done || stalled || r.outer >= max_outer || r.total >= r.total_cap || !(r.rr == r.rr)) {
        r.refining = false;
        continue;
      }
      r.norm = std::sqrt(r.rr);
      r.sweep_tol = std::max(r.inner_tol, 0.5 * std::sqrt(r.target / r.rr));
      r.sweeping = r.swept = true;
      r.inner = 0;
      r.rr_last = -1;
      r.flat = 0;
      pt[s].inv_norm = 1.0 / r.norm;
      st[s].tol_rel = r.sweep_tol;
      vt[s].norm = r.norm;
      r.c->sweep_index = r.outer;
      r.c->pred_kind = r.amg ? MPS_PRECOND_AMG : MPS_PRECOND_MULTILEVEL;
      any = true;
    }
    if (!any) break;
    // one FP32 sweep of every refining scene
    B.upload_tab(1, sizeof(PrecArgs));
    B.upload_tab(2, sizeof(SpmvArgs<MatFreeRec>));
    B.upload_tab(3, sizeof(VecArgs));
    auto sweeping = [&](size_t s) { return sc[s].sweeping; };
    const unsigned ncl = B.group(kClear, sweeping, [](size_t) { return 1; });
    hipLaunchKernelGGL(kg_status_clear, dim3(ncl), dim3(64), 0, w.stream, B.tab_dev<VecArgs>(3), B.g(kClear));
    launch_check();
    // the preconditioner of every sweeping scene: the multilevel pass, or the
    // multigrid's passes around its cycle (grouped per step; a scene joins
    // the level steps it has)
    unsigned np = 0, ns = 0;
    unsigned na[kKindCount] = {};
    int lmax = 0;
    auto regroup = [&] {
      auto ml = [&](size_t s) { return sc[s].sweeping && !sc[s].amg; };
      auto mg = [&](size_t s) { return sc[s].sweeping && sc[s].amg; };
      np = B.group(kPrec, ml, [&](size_t s) { return sc[s].pgrid; });
      ns = B.group(kSpmv, sweeping, [&](size_t s) { return sc[s].L.grid; });
      if (!any_amg) return;
      lmax = 0;
      for (size_t s = 0; s < S; s++)
        if (mg(s)) lmax = std::max(lmax, sc[s].l0);
      na[kPrec1] = B.group(kPrec1, mg, [&](size_t s) { return sc[s].pgrid; });
      na[kRfix] = B.group(kRfix, mg, [&](size_t s) { return sc[s].rgrid; });
      na[kBlock] = B.group(kBlock, mg, [](size_t) { return 1; });
      na[kPrecUp] = B.group(kPrecUp, [&](size_t s) { return mg(s) && sc[s].l0 >= 1; },
                            [&](size_t s) { return sc[s].pgrid; });
      na[kPrec2] = B.group(kPrec2, [&](size_t s) { return mg(s) && sc[s].l0 == 0; },
                           [&](size_t s) { return sc[s].pgrid; });
      for (int l = 0; l < lmax; l++) {
        na[kDown0 + l] = B.group(kDown0 + l, [&](size_t s) { return mg(s) && l < sc[s].l0; },
                                 [&](size_t s) { return sc[s].dgrid[l]; });
        if (l >= 1)
          na[kUp0 + l] = B.group(kUp0 + l, [&](size_t s) { return mg(s) && l < sc[s].l0; },
                                 [&](size_t s) { return sc[s].ugrid[l]; });
      }
    };
    auto precondition = [&](long long it) {
      if (np) {
        hipLaunchKernelGGL(kg_ml_prec, dim3(np), dim3(NTP), 0, w.stream, B.tab_dev<PrecArgs>(1), B.g(kPrec), it);
        launch_check();
      }
      if (!any_amg || na[kPrec1] == 0) return;
      const AmgArgs* am = B.tab_dev<AmgArgs>(5);
      hipLaunchKernelGGL(kg_ml_prec_m<1>, dim3(na[kPrec1]), dim3(NTP), 0, w.stream, B.tab_dev<PrecArgs>(1),
                         B.g(kPrec1), it);
      launch_check();
      hipLaunchKernelGGL(kg_amg_rfix, dim3(na[kRfix]), dim3(256), 0, w.stream, am, B.g(kRfix));
      launch_check();
      for (int l = 0; l < lmax; l++)
        if (na[kDown0 + l]) {
          hipLaunchKernelGGL(kg_amg_down, dim3(na[kDown0 + l]), dim3(256), 0, w.stream, am, B.g(kDown0 + l), l);
          launch_check();
        }
      hipLaunchKernelGGL(kg_amg_block, dim3(na[kBlock]), dim3(kAmgBlockThreads), 0, w.stream, am, B.g(kBlock));
      launch_check();
      for (int l = lmax - 1; l >= 1; l--)
        if (na[kUp0 + l]) {
          hipLaunchKernelGGL(kg_amg_up, dim3(na[kUp0 + l]), dim3(256), 0, w.stream, am, B.g(kUp0 + l), l);
          launch_check();
        }
      if (na[kPrecUp]) {
        hipLaunchKernelGGL(kg_ml_prec_up, dim3(na[kPrecUp]), dim3(NTP), 0, w.stream, B.tab_dev<PrecArgs>(1), am,
                           B.g(kPrecUp), it);
        launch_check();
      }
      if (na[kPrec2]) {
        hipLaunchKernelGGL(kg_ml_prec_m<2>, dim3(na[kPrec2]), dim3(NTP), 0, w.stream, B.tab_dev<PrecArgs>(1),
                           B.g(kPrec2), it);
        launch_check();
      }
    };
    regroup();
    precondition(0);
    long long it = 1, pred = 0, lim = 0;
    for (size_t s = 0; s < S; s++)
      if (sc[s].sweeping) {
        Context& c = *sc[s].c;
        pred = std::max(pred, c.pred(c.sweep_index));
        lim = std::max(lim, sc[s].limit);
      }
    while (true) {
      const long long chunk = next_chunk(it - 1, pred);
      for (long long m = 0; m < chunk && it <= lim; m++, it++) {
        if (it == 1)
          hipLaunchKernelGGL((kg_ml_spmv<true>), dim3(ns), dim3(NT), 0, w.stream,
                             B.tab_dev<SpmvArgs<MatFreeRec>>(2), B.g(kSpmv), it);
        else
          hipLaunchKernelGGL((kg_ml_spmv<false>), dim3(ns), dim3(NT), 0, w.stream,
                             B.tab_dev<SpmvArgs<MatFreeRec>>(2), B.g(kSpmv), it);
        launch_check();
        precondition(it);
      }
      MPS_HIP(hipMemcpyAsync(w.st_host, w.st, S * sizeof(CgStatus), hipMemcpyDeviceToHost, w.stream));
      MPS_HIP(hipStreamSynchronize(w.stream));
      // per scene: stopped (the SpMV after the stop added the last alpha p to
      // d), out of iterations or stagnating (d still needs it: k_ml_finish)
      bool changed = false, finish = false;
      std::vector<char> need_finish(S, 0);
      for (size_t s = 0; s < S; s++) {
        SceneRun& r = sc[s];
        if (!r.sweeping) continue;
        const CgStatus& h = w.st_host[s];
        bool end = false;
        if (h.converged) {
          r.inner = h.iterations;
          end = true;
        } else if (it > r.limit) {
          r.inner = r.limit;
          end = need_finish[s] = 1;
        } else {
          const double q = std::min(h.rr[0], h.rr[1]);
          r.flat = (r.rr_last > 0 && !(q < 0.81 * r.rr_last)) ? r.flat + 1 : 0;
          r.rr_last = q;
          if (r.flat >= 2) {
            r.inner = it - 1;
            end = need_finish[s] = 1;
          }
        }
        if (end) {
          r.sweeping = false;
          changed = true;
          finish = finish || need_finish[s];
        }
      }
      if (finish) {
        const unsigned nf = B.group(kFinish, [&](size_t s) { return need_finish[s] != 0; },
                                    [&](size_t s) { return sc[s].vgrid; });
        hipLaunchKernelGGL(kg_ml_finish, dim3(nf), dim3(256), 0, w.stream, B.tab_dev<VecArgs>(3), B.g(kFinish));
        launch_check();
      }
      if (changed) {
        regroup();
        if (ns == 0) break;
        lim = 0;
        for (size_t s = 0; s < S; s++)
          if (sc[s].sweeping) lim = std::max(lim, sc[s].limit);
      }
    }
    // x += |r| d for the scenes of this sweep
    const unsigned nc = B.group(kCorrect, [&](size_t s) { return sc[s].swept; }, [&](size_t s) { return sc[s].vgrid; });
    hipLaunchKernelGGL(kg_ir_correct, dim3(nc), dim3(256), 0, w.stream, B.tab_dev<VecArgs>(3), B.g(kCorrect));
    launch_check();
    for (size_t s = 0; s < S; s++) {
      SceneRun& r = sc[s];
      if (!r.swept) continue;
      Context& c = *r.c;
      c.pred(r.outer) = r.inner;
      r.total += r.inner;
      r.outer++;
    }
  }
}

}  // namespace

int precond_for_system(Context& c) {
  constexpr int ML = MPS_PRECOND_MULTILEVEL, MG = MPS_PRECOND_AMG;
  settle_solve(c);  // the last solve's time, for the automatic choice
  int pick = c.solver.precond;
  if (pick == MPS_PRECOND_AUTO) {
    if (c.auto_ms[0] < 0) pick = ML;
    else if (c.auto_ms[1] < 0) pick = MG;
    else pick = (c.auto_solves >= c.auto_next) ? (c.auto_best == ML ? MG : ML) : c.auto_best;
  }
  c.precond_now = pick;
  return pick;
}

int effective_precond(const Context& c) {
  if (c.solver.precond != MPS_PRECOND_AUTO) return c.solver.precond;
  return (c.precond_now == MPS_PRECOND_MULTILEVEL || c.precond_now == MPS_PRECOND_AMG) ? c.precond_now
                                                                                         : MPS_PRECOND_MULTILEVEL;
}

void ensure_amg_rows(Context& c, size_t n) {
  if (c.amg_rows_cap < n) {
    dfree(c.amg_rows);
    c.amg_rows = dalloc<double>(9 * n);
    c.amg_rows_cap = n;
    c.amg_rows_ready = false;
  }
}

// Re-launches the last solve's SpMV on its operands with a status block that
// never stops (tolerance -1, unit scalars, alpha 0: the correction d is left
// as it is); only p, q and the partial sums are overwritten.
double replay_spmv(Context& c, int kid, int reps) {
  if (c.replay_ready != kid || c.ppe_n == 0)
    throw Failure{MPS_ERR_STATE, "the last solve did not run this SpMV"};
  if (!c.st_replay) c.st_replay = dalloc<CgStatus>(1);
  CgStatus s{};
  s.rr[0] = s.rr[1] = 1.0;
  s.gamma[0] = s.gamma[1] = 1.0;
  s.tol = -1.0;
  MPS_HIP(hipMemcpyAsync(c.st_replay, &s, sizeof(s), hipMemcpyHostToDevice, c.stream));
  const Launch L = cg_grid(c.ppe_n);
  const size_t pgrid = std::max<size_t>(1, (*c.ml_active_host + NTP / 64 - 1) / (NTP / 64));
  auto launch = [&] {
    if (kid == MPS_K_CG_SPMV32)
      hipLaunchKernelGGL((k_ml_spmv<false, MatFreeRec>), dim3(L.grid), dim3(NT), 0, c.stream,
                         SpmvArgs<MatFreeRec>{matfree_rec(c, L.rpb), nullptr, c.pa32, c.q32, c.d32, c.part_ml,
                                              c.part_ml_cap, static_cast<int>(pgrid), c.part3b, c.st_replay, 1e-8},
                         2LL);
    else
      hipLaunchKernelGGL((k_cg_spmv<false, Sell<double>>), dim3(L.grid), dim3(NT), 0, c.stream,
                         sell<double>(c, c.sval, c.diag, L.rpb), static_cast<const double*>(c.r), c.pa, c.q,
                         static_cast<const double*>(c.partB), static_cast<const double*>(c.partC), L.grid,
                         c.partA, c.st_replay, StopRule{MPS_STOP_RHS_NORM, 1e-8}, 2LL);
  };
  launch();  // warm
  MPS_HIP(hipEventRecord(c.solve_a, c.stream));
  for (int k = 0; k < reps; k++) launch();
  MPS_HIP(hipEventRecord(c.solve_b, c.stream));
  check_launch();
  MPS_HIP(hipEventSynchronize(c.solve_b));
  float ms = 0;
  MPS_HIP(hipEventElapsedTime(&ms, c.solve_a, c.solve_b));
  return ms / reps;
}

// MPS_PRECOND_AUTO: the time of a solve with the multilevel or the multigrid
// preconditioner; the first of each sets the choice, later ones of the other
// (the re-timings) switch it when faster and otherwise double the interval to
// the next re-timing (16 .. 1024 solves).
// A solve on the current choice this much slower than the other's last
// timing, when that timing is at most kAutoFresh solves old, has the other
// timed again on the next solve (right after a switch the flow can change the
// two solvers' costs at different rates; an older timing says nothing about
// the other's cost now, and timing it again is left to the interval).
constexpr double kAutoSlack = 1.05;
constexpr long long kAutoFresh = 16;
// A timing of the other that loses grows the interval 2x, or 8x when it took
// kAutoFar times the current choice's (far behind: not worth timing soon).
constexpr double kAutoFar = 1.5;

void auto_record(Context& c, double ms) {
  constexpr int ML = MPS_PRECOND_MULTILEVEL, MG = MPS_PRECOND_AMG;
  const int used = static_cast<int>(c.stats.precond_used);
  if (c.solver.precond != MPS_PRECOND_AUTO || (used != ML && used != MG)) return;
  const bool first = c.auto_ms[0] < 0 || c.auto_ms[1] < 0;
  const bool retime = !first && used != c.auto_best;
  const double best_ms = c.auto_ms[c.auto_best == ML ? 0 : 1];
  c.auto_ms[used == ML ? 0 : 1] = ms;
  c.auto_at[used == ML ? 0 : 1] = c.auto_solves;
  c.auto_solves++;
  if (first) {
    if (c.auto_ms[0] >= 0 && c.auto_ms[1] >= 0) {
      c.auto_best = c.auto_ms[1] < c.auto_ms[0] ? MG : ML;
      c.auto_interval = 16;
      c.auto_next = c.auto_solves + c.auto_interval;
    }
    return;
  }
  if (!retime) {
    const int other = used == ML ? 1 : 0;
    if (c.auto_solves - c.auto_at[other] <= kAutoFresh && ms > kAutoSlack * c.auto_ms[other])
      c.auto_next = c.auto_solves;
    return;
  }
  if (ms < best_ms) {
    c.auto_best = used;
    c.auto_interval = 16;
  } else {
    c.auto_interval = std::min<long long>(1024, (ms >= kAutoFar * best_ms ? 8 : 2) * c.auto_interval);
  }
  c.auto_next = c.auto_solves + c.auto_interval;
}

// The last solve's time from its events, read on first use: the solve itself
// does not wait for its end event (its last FP64 residual has synchronised
// the stream already); the automatic choice records it here.
void settle_solve(Context& c) {
  if (!c.solve_pending) return;
  c.solve_pending = false;
  MPS_HIP(hipEventSynchronize(c.solve_b));
  float ms = 0;
  MPS_HIP(hipEventElapsedTime(&ms, c.solve_a, c.solve_b));
  c.stats.solve_ms = ms;
  if (c.solve_pending_auto) auto_record(c, ms);
}

void ppe_solve(Context& c) {
  if (!c.have_ppe) throw Failure{MPS_ERR_STATE, "no pressure Poisson system assembled"};
  settle_solve(c);
  const size_t n = c.ppe_n;
  c.stats = mps_solve_stats{};
  c.replay_ready = 0;
  c.final_rr = -1;
  c.learned_pending = false;
  if (n == 0) return;
  StopRule sr;
  sr.rule = c.solver.stop_rule;
  sr.tol = (c.solver.tol > 0) ? c.solver.tol : c.eps;
  const long long max_iter = (c.solver.max_iter > 0) ? c.solver.max_iter : static_cast<long long>(n);
  // snapshot features use x0 = p_prev, so they are taken before any warm start
  const bool on_particles = (c.ppe_n == c.n);
  if (c.dataset && on_particles) collect_before(c);
  MPS_HIP(hipEventRecord(c.solve_a, c.stream));
  int rc = MPS_OK;
  std::string msg;
  try {
    if (c.solver.warm_start == MPS_WARM_LEARNED && on_particles && c.model && c.ppe_from_particles) {
      ml_setup_ahead(c);
      learned_warm_start(c);
    }
    if (c.solver.kind == MPS_SOLVER_MIXED_IR) solve_mixed(c, sr, max_iter);
    else solve_cg64(c, sr, max_iter);
    c.learned_pending = false;
    join_aux(c);
  } catch (Failure& f) {
    c.learned_pending = false;
    join_aux(c);
    if (f.code != MPS_ERR_CG_NOT_CONVERGED) throw;
    rc = f.code;
    msg = f.msg;
  }
  MPS_HIP(hipEventRecord(c.solve_b, c.stream));
  c.solve_pending = true;
  c.solve_pending_auto = false;
  // true FP64 residual of the returned x (reported, not part of the solve;
  // the refinement's last residual is that one already)
  double rr = c.final_rr, bb = c.final_bb;
  if (rr < 0) residual64(c, rr, bb, false);
  c.stats.rel_residual = (bb > 0) ? std::sqrt(rr / bb) : std::sqrt(rr);
  c.stats.initial_rel_residual = (bb > 0) ? std::sqrt(c.stats.rr0 / bb) : std::sqrt(c.stats.rr0);
  if (rc != MPS_OK) {
    settle_solve(c);
    throw Failure{rc, msg};
  }
  c.solve_pending_auto = true;
  if (c.dataset && on_particles) collect_after(c);
}


// The learned warm start's residual r0 of x = p, its feature records and
// scales, from particles' systems only (the matrix-free operator):
// red_out[0..4] = |r0|^2, |b|^2, sum p^2, sum diag^2, alive count and
// red_out[6..9] the scales (feat_scales).  Returns false (nothing enqueued)
// for a system given as a matrix.
bool learned_residual32(Context& c, FeatRec* rec) {
  if (!c.ppe_from_particles || c.ppe_n != c.n) return false;
  const Launch L = cg_grid(c.ppe_n);
  // one row per thread (the gathers want every wavefront the CU can hold);
  // five partial sums per block
  const int G = static_cast<int>((c.ppe_n + 255) / 256);
  if (c.learn_part_cap < static_cast<size_t>(G)) {
    dfree(c.learn_part);
    c.learn_part = dalloc<double>(5 * static_cast<size_t>(G));
    c.learn_part_cap = G;
  }
  ensure_zrec(c, c.ppe_n);
  hipLaunchKernelGGL(k_rec_p, dim3(grid_for(c.ppe_n)), dim3(256), 0, c.stream, c.ppe_n,
                     static_cast<const uint2*>(c.qpos), static_cast<const double*>(c.p), c.zrec);
  hipLaunchKernelGGL(k_learn_resid32, dim3(G), dim3(256), 0, c.stream, matfree_rec(c, L.rpb),
                     static_cast<const double*>(c.b), static_cast<const double*>(c.p),
                     static_cast<const double*>(c.diag), static_cast<const uint8_t*>(c.type),
                     static_cast<const double*>(c.x), static_cast<const double*>(c.z), rec, c.learn_part);
  hipLaunchKernelGGL(k_reduce5, dim3(1), dim3(1024), 0, c.stream, static_cast<const double*>(c.learn_part), G,
                     c.red_out);
  check_launch();
  return true;
}


// With the learned warm start in front of the multilevel mixed solver: the
// solver's set-up, which does not depend on the initial guess, goes to the
// auxiliary stream after the assembly and runs beside the warm start's
// kernels; solve_mixed_ml joins it.
void ml_setup_ahead(Context& c) {
  const bool cells_match = c.have_neighbors && c.ppe_search_serial == c.search_serial && c.ppe_n == c.n;
  if (c.solver.kind != MPS_SOLVER_MIXED_IR || !c.ppe_from_particles || effective_precond(c) != MPS_PRECOND_MULTILEVEL ||
      !cells_match || c.ml_setup_pending)
    return;
  if (!c.stream_aux) {
    // created on first use: an idle stream would still take its turn in the
    // assignment of hardware queues to streams (contexts that never use it,
    // such as many scenes side by side, would share fewer queues)
    MPS_HIP(hipStreamCreateWithFlags(&c.stream_aux, hipStreamNonBlocking));
    MPS_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
    MPS_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
  }
  MPS_HIP(hipEventRecord(c.ev_fork, c.stream));
  MPS_HIP(hipStreamWaitEvent(c.stream_aux, c.ev_fork, 0));
  std::swap(c.stream, c.stream_aux);  // everything below enqueues on c.stream
  try {
    ml_setup(c);
  } catch (...) {
    std::swap(c.stream, c.stream_aux);
    throw;
  }
  std::swap(c.stream, c.stream_aux);
  MPS_HIP(hipEventRecord(c.ev_join, c.stream_aux));
  c.ml_setup_pending = true;
}

// Joins set-up work left on the auxiliary stream (a solve that ended early).
void join_aux(Context& c) { join_ml_setup(c); }

void BatchWork::ensure(size_t scenes) {
  if (scenes <= cap && tab) return;
  const int dev = device;
  hipStream_t s = stream;
  hipEvent_t a = ev_a, b = ev_b;
  stream = nullptr;
  ev_a = ev_b = nullptr;
  release();
  device = dev;
  stream = s;
  ev_a = a;
  ev_b = b;
  cap = std::max<size_t>(scenes, 8);
  tab = dalloc<unsigned char>(kKinds * cap * kArgBytes);
  first = dalloc<uint32_t>(kKinds * (cap + 1));
  st = dalloc<CgStatus>(cap);
  void* h = nullptr;
  MPS_HIP(hipHostMalloc(&h, kKinds * cap * kArgBytes, hipHostMallocDefault));
  tab_host = static_cast<unsigned char*>(h);
  MPS_HIP(hipHostMalloc(&h, kKinds * (cap + 1) * sizeof(uint32_t), hipHostMallocDefault));
  first_host = static_cast<uint32_t*>(h);
  MPS_HIP(hipHostMalloc(&h, cap * sizeof(CgStatus), hipHostMallocDefault));
  st_host = static_cast<CgStatus*>(h);
  std::memset(tab_host, 0, kKinds * cap * kArgBytes);
  if (!stream) MPS_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  if (!ev_a) MPS_HIP(hipEventCreate(&ev_a));
  if (!ev_b) MPS_HIP(hipEventCreate(&ev_b));
}

void BatchWork::release() {
  dfree(tab);
  dfree(first);
  dfree(st);
  if (tab_host) (void)hipHostFree(tab_host);
  if (first_host) (void)hipHostFree(first_host);
  if (st_host) (void)hipHostFree(st_host);
  tab_host = nullptr;
  first_host = nullptr;
  st_host = nullptr;
  if (ev_a) (void)hipEventDestroy(ev_a);
  if (ev_b) (void)hipEventDestroy(ev_b);
  if (stream) (void)hipStreamDestroy(stream);
  ev_a = ev_b = nullptr;
  stream = nullptr;
  cap = 0;
}

bool batch_eligible(const Context& c) {
  const bool cells_match = c.have_neighbors && c.ppe_search_serial == c.search_serial && c.ppe_n == c.n;
  return c.n > 0 && c.have_ppe && c.ppe_from_particles && cells_match && c.solver.kind == MPS_SOLVER_MIXED_IR &&
         (c.solver.precond == MPS_PRECOND_MULTILEVEL || c.solver.precond == MPS_PRECOND_AMG);
}

bool batch_amg(const Context& c) { return c.solver.precond == MPS_PRECOND_AMG; }

void amg_prepare(Context& c) {
  const size_t n = c.ppe_n;
  ensure_zrec(c, n);
  amg_setup(c);
  const MlGeom g = ml_geom(c);
  if (c.part_ml_cap < g.sgrid) {
    dfree(c.part_ml);
    c.part_ml = dalloc<double>(2 * g.sgrid);
    c.part_ml_cap = g.sgrid;
  }
  MPS_HIP(hipStreamSynchronize(c.stream));  // the count of occupied cell blocks
  const CellGrid G = ml_grid(c, g);
  const uint32_t nact = static_cast<uint32_t>(*c.ml_active_host);
  const size_t pgrid = std::max<size_t>(1, (nact + NTP / 64 - 1) / (NTP / 64));
  const size_t slots = static_cast<size_t>(std::max<uint32_t>(nact, 1)) * 64;
  if (c.ml_pack_cap < slots) {
    dfree(c.ml_pid);
    dfree(c.ml_dv);
    dfree(c.ml_sc);
    c.ml_pid = dalloc<uint32_t>(slots * kPerLane);
    c.ml_dv = dalloc<float>(slots * kPerLane);
    c.ml_sc = dalloc<float4>(slots);
    c.ml_pack_cap = slots;
  }
  // the multigrid's rules: no aggregate scales, rows without stored
  // off-diagonals outside every aggregate (as solve_mixed_amg packs)
  hipLaunchKernelGGL(k_ml_pack, dim3(static_cast<unsigned>(pgrid)), dim3(NTP), 0, c.stream, G,
                     static_cast<const uint32_t*>(c.ml_list), nact, static_cast<const float*>(nullptr),
                     static_cast<const float*>(c.dinv32), c.ml_pid, c.ml_dv, c.ml_sc,
                     static_cast<const uint32_t*>(c.row_len));
  const size_t nslots = static_cast<size_t>(nact) * kPerLane * 64;
  ensure_q9(c, nslots);
  if (nslots > 0)
    hipLaunchKernelGGL(k_amg_q, dim3(grid_for(nslots)), dim3(256), 0, c.stream, nslots,
                       static_cast<const uint32_t*>(c.ml_pid), static_cast<const double*>(c.amg_rows),
                       static_cast<const double*>(c.diag), kAmgOmegaP, c.amg_q9, c.amg_q9c);
  const size_t npts0 = static_cast<size_t>(c.amg_box[2]) * c.amg_box[3];
  const size_t tq_stride = (npts0 + 63) & ~static_cast<size_t>(63);
  if (c.amg_tq_cap < 9 * tq_stride) {
    dfree(c.amg_tq);
    c.amg_tq = dalloc<float>(9 * tq_stride);
    c.amg_tq_cap = 9 * tq_stride;
  }
  c.amg_tq_stride = tq_stride;
  MPS_HIP(hipMemsetAsync(c.amg_tq, 0, 9 * tq_stride * sizeof(float), c.stream));
  hipLaunchKernelGGL(k_rec_init, dim3(grid_for(n)), dim3(256), 0, c.stream, n, static_cast<const uint2*>(c.qpos),
                     c.zrec, c.r32, c.d32);
  check_launch();
}

void ml_prepare(Context& c) {
  const size_t n = c.ppe_n;
  ensure_zrec(c, n);
  if (!join_ml_setup(c)) ml_setup(c);
  MPS_HIP(hipStreamSynchronize(c.stream));  // the count of occupied cell blocks
  const MlGeom g = ml_geom(c);
  const CellGrid G = ml_grid(c, g);
  const uint32_t nact = static_cast<uint32_t>(*c.ml_active_host);
  const size_t pgrid = std::max<size_t>(1, (nact + NTP / 64 - 1) / (NTP / 64));
  const size_t slots = static_cast<size_t>(std::max<uint32_t>(nact, 1)) * 64;
  if (c.ml_pack_cap < slots) {
    dfree(c.ml_pid);
    dfree(c.ml_dv);
    dfree(c.ml_sc);
    c.ml_pid = dalloc<uint32_t>(slots * kPerLane);
    c.ml_dv = dalloc<float>(slots * kPerLane);
    c.ml_sc = dalloc<float4>(slots);
    c.ml_pack_cap = slots;
  }
  hipLaunchKernelGGL(k_ml_pack, dim3(static_cast<unsigned>(pgrid)), dim3(NTP), 0, c.stream, G,
                     static_cast<const uint32_t*>(c.ml_list), nact, static_cast<const float*>(c.agg_inv),
                     static_cast<const float*>(c.dinv32), c.ml_pid, c.ml_dv, c.ml_sc,
                     static_cast<const uint32_t*>(nullptr));
  hipLaunchKernelGGL(k_rec_init, dim3(grid_for(n)), dim3(256), 0, c.stream, n, static_cast<const uint2*>(c.qpos),
                     c.zrec, c.r32, c.d32);
  check_launch();
}

void batch_solve(std::vector<Context*>& scenes, BatchWork& w) {
  std::vector<SceneRun> sc;
  sc.reserve(scenes.size());
  for (Context* c : scenes) {
    SceneRun r;
    r.c = c;
    const size_t n = c->ppe_n;
    r.L = cg_grid(n);
    const uint32_t nact = static_cast<uint32_t>(*c->ml_active_host);
    r.pgrid = std::max<size_t>(1, (nact + NTP / 64 - 1) / (NTP / 64));
    r.vgrid = grid_for(n);
    r.amg = batch_amg(*c);
    if (r.amg) {
      r.l0 = amg_block_level(*c);
      r.rgrid = grid_for(static_cast<size_t>(c->amg_nx[0]) * c->amg_nz[0]);
      for (int l = 0; l < r.l0; l++) {
        r.dgrid[l] = grid_for(static_cast<size_t>(c->amg_nx[l + 1]) * c->amg_nz[l + 1]);
        r.ugrid[l] = grid_for(static_cast<size_t>(c->amg_nx[l]) * c->amg_nz[l]);
      }
    }
    r.sr = StopRule{c->solver.stop_rule, (c->solver.tol > 0) ? c->solver.tol : c->eps};
    r.max_iter = (c->solver.max_iter > 0) ? c->solver.max_iter : static_cast<long long>(n);
    r.limit = inner_limit(*c, r.max_iter);
    r.total_cap = c->solver.max_iter > 0 ? c->solver.max_iter : LLONG_MAX;
    r.inner_tol = c->solver.inner_tol > 0 ? c->solver.inner_tol : 1e-4;
    const long long learned = c->stats.learned_used;
    const double prev_rel = c->stats.prev_rel_residual;
    settle_solve(*c);
    c->stats = mps_solve_stats{};
    c->stats.learned_used = learned;
    c->stats.prev_rel_residual = prev_rel;
    c->replay_ready = MPS_K_CG_SPMV32;
    c->final_rr = -1;
    sc.push_back(r);
  }
  if (sc.empty()) return;
  w.ensure(sc.size());
  w.launches = 0;
  MPS_HIP(hipEventRecord(w.ev_a, w.stream));
  batch_solve_run(sc, w);
  MPS_HIP(hipEventRecord(w.ev_b, w.stream));
  MPS_HIP(hipEventSynchronize(w.ev_b));
  float ms = 0;
  MPS_HIP(hipEventElapsedTime(&ms, w.ev_a, w.ev_b));
  w.solve_ms = ms;
  int code = MPS_OK;
  std::string msg;
  for (SceneRun& r : sc) {
    Context& c = *r.c;
    try {
      if (r.fallback) {
        // the FP32 sweeps stalled: the FP64 CG finishes from the current x,
        // to the same target, on the scene's own stream
        StopRule rest = r.sr;
        if (r.sr.rule == MPS_STOP_INITIAL_RESIDUAL) rest.tol = r.sr.tol * std::sqrt(r.rr0 / r.rr);
        const long long cap = c.solver.max_iter > 0 ? c.solver.max_iter - r.total : max_iter_fallback(c);
        solve_cg64(c, rest, cap);
        c.stats.iterations += r.total;
        c.stats.outer_iterations = r.outer + 1;
        c.stats.rr0 = r.rr0;
        c.stats.fallback = 1;
        double rr = 0, bb = 0;
        residual64(c, rr, bb, false);
        c.final_rr = rr;
        c.final_bb = bb;
      } else {
        c.stats.iterations = r.total;
        c.stats.outer_iterations = r.outer;
        c.stats.rr0 = r.rr0;
        c.stats.rr = r.rr;
        c.last_iters = r.total;
        c.final_rr = r.rr;
        c.final_bb = r.bb;
        if (!r.done)
          throw Failure{MPS_ERR_CG_NOT_CONVERGED,
                        "Conjugate Gradient method couldn't solve Pressure Poison Equation"};
      }
    } catch (Failure& f) {
      if (code == MPS_OK) {
        code = f.code;
        msg = f.msg;
      }
    }
    c.stats.solve_ms = w.solve_ms;
    const double bb = c.final_bb;
    c.stats.rel_residual = (bb > 0) ? std::sqrt(c.final_rr / bb) : std::sqrt(c.final_rr);
    c.stats.initial_rel_residual = (bb > 0) ? std::sqrt(c.stats.rr0 / bb) : std::sqrt(c.stats.rr0);
  }
  if (code != MPS_OK) throw Failure{code, msg};
}

}  // namespace mps

// ---------------- allocation-free FP64 CG on a caller's CSR system ----------------
//
// mps_cg_solve_csr_noalloc runs the FP64 CG of solve_cg64 (the same kernels,
// instantiated for a CSR matrix) without allocating: its work vectors r, p, q,
// the partial-sum slots and the status live in `scratch`, a fixed array of
// the code object that every device holding the library keeps until the
// process ends.  Host code never passes `scratch` itself to a kernel (that
// name is a host-side stand-in): the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
namespace mps {
namespace {

constexpr size_t kNoAllocMaxRows = MPS_NOALLOC_MAX_ROWS;
constexpr size_t kStatusDoubles = (sizeof(CgStatus) + 7) / 8;
constexpr size_t kScratchDoubles = 3 * kNoAllocMaxRows + 3 * static_cast<size_t>(kRedBlocks) + kStatusDoubles;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device `dev` (looked up once).
double* scratch_ptr(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Failure{MPS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  return g_scratch_dev[dev];
}

__global__ void k_status_clear(CgStatus* st) {
  if (threadIdx.x == 0) *st = CgStatus{};
}

void csr_solve(size_t n, const uint64_t* rowptr, const uint32_t* col, const double* val, const double* b,
               double* x, double eps, long long max_iter, hipStream_t s, double* work, long long* iters_out) {
  double* r = work;
  double* p = r + kNoAllocMaxRows;
  double* q = p + kNoAllocMaxRows;
  double* partA = q + kNoAllocMaxRows;
  double* partB = partA + kRedBlocks;
  double* partC = partB + kRedBlocks;
  CgStatus* st = reinterpret_cast<CgStatus*>(partC + kRedBlocks);
  const Launch L = cg_grid(n);
  const Csr A{rowptr, col, val, n, L.rpb};
  const StopRule sr{MPS_STOP_INITIAL_RESIDUAL, eps};
  // every value the iterations read is written first: the status by the clear
  // kernel, r and the partial slots by k_cg_init, p and q by the first
  // iteration (FIRST: p = r, q = A r)
  hipLaunchKernelGGL(k_status_clear, dim3(1), dim3(64), 0, s, st);
  hipLaunchKernelGGL(k_cg_init<Csr>, dim3(L.grid), dim3(NT), 0, s, A, b, static_cast<const double*>(x), r, partB,
                     partC);
  check_launch();
  CgStatus hs{};
  long long it = 1;
  while (true) {
    const long long chunk = next_chunk(it - 1, 0);
    for (long long m = 0; m < chunk && it <= max_iter; m++, it++) {
      if (it == 1)
        hipLaunchKernelGGL((k_cg_spmv<true, Csr>), dim3(L.grid), dim3(NT), 0, s, A, static_cast<const double*>(r), p,
                           q, static_cast<const double*>(partB), static_cast<const double*>(partC), L.grid, partA,
                           st, sr, it);
      else
        hipLaunchKernelGGL((k_cg_spmv<false, Csr>), dim3(L.grid), dim3(NT), 0, s, A, static_cast<const double*>(r),
                           p, q, static_cast<const double*>(partB), static_cast<const double*>(partC), L.grid,
                           partA, st, sr, it);
      hipLaunchKernelGGL(k_cg_update, dim3(L.grid), dim3(NT), 0, s, n, L.rpb, x, r, static_cast<const double*>(p),
                         static_cast<const double*>(q), static_cast<const double*>(partA), L.grid, partB, st, it);
    }
    check_launch();
    if (it > max_iter)
      hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(NT), 0, s, static_cast<const double*>(partB), L.grid, st, sr, it);
    MPS_HIP(hipMemcpyAsync(&hs, st, sizeof(CgStatus), hipMemcpyDeviceToHost, s));
    MPS_HIP(hipStreamSynchronize(s));
    if (hs.converged || it > max_iter) break;
  }
  if (iters_out) *iters_out = hs.converged ? hs.iterations : max_iter;
  if (!hs.converged)
    throw Failure{MPS_ERR_CG_NOT_CONVERGED, "Conjugate Gradient method couldn't solve Pressure Poison Equation"};
}

void put_err(char* err, size_t len, const char* msg) {
  if (err && len) std::snprintf(err, len, "%s", msg);
}

}  // namespace
}  // namespace mps

extern "C" int mps_cg_solve_csr_noalloc(int device, size_t n, const uint64_t* rowptr, const uint32_t* col,
                                        const double* val, const double* b, double* x,
                                        double allowable_residual, int64_t max_iter, void* stream,
                                        int64_t* iterations, char* err, size_t err_len) {
  using namespace mps;
  put_err(err, err_len, "");
  if (iterations) *iterations = 0;
  if (n > kNoAllocMaxRows) {
    char msg[160];
    std::snprintf(msg, sizeof msg, "mps_cg_solve_csr_noalloc: %zu rows exceed the fixed work area (at most %zu)", n,
                  kNoAllocMaxRows);
    put_err(err, err_len, msg);
    return MPS_ERR_BAD_ARG;
  }
  if (device < 0 || device >= kMaxDevices || !(allowable_residual > 0) ||
      (n > 0 && (!rowptr || !col || !val || !b || !x))) {
    put_err(err, err_len, "mps_cg_solve_csr_noalloc: bad device, tolerance or pointer");
    return MPS_ERR_BAD_ARG;
  }
  if (n == 0) return MPS_OK;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  int rc = MPS_OK;
  try {
    MPS_HIP(hipSetDevice(device));
    double* work = scratch_ptr(device);
    csr_solve(n, rowptr, col, val, b, x, allowable_residual, max_iter > 0 ? max_iter : static_cast<long long>(n),
              static_cast<hipStream_t>(stream), work, reinterpret_cast<long long*>(iterations));
  } catch (Failure& f) {
    if (f.code == MPS_ERR_HIP) (void)hipGetLastError();
    put_err(err, err_len, f.msg.c_str());
    rc = f.code;
  }
  if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
  return rc;
}
