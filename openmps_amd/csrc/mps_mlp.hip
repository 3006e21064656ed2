// Learned CG warm start: per-particle features, a 132 -> 256 -> 256 -> 1 MLP
// (GELU) with BF16 MFMA GEMMs and FP32 accumulation, its backward pass, Adam,
// snapshot collection from the FP64 solver, and the warm-start application.
//
// Features of particle i (alive rows only; 132 values, padded to K = 160):
//   [0] non-dummy neighbours within r_lap / 32      [1] (n_i - n0) / n0
//   [2] p_prev_i / s_p                               [3] r0_i / s_r
//   then for the 32 nearest non-dummy neighbours j within r_lap (ascending
//   distance, zero padded): (x_j - x_i)/l0, (z_j - z_i)/l0, p_prev_j / s_p, r0_j / s_r
// where r0 = b - A p_prev is the residual of the previous-step guess, s_r its
// RMS over alive rows and s_p the RMS of p_prev.  The network predicts the
// correction p - p_prev in units of s_d = s_r l0^2 / 5 (the size of A^-1 r0
// for HL coefficients ~ 5 / l0^2 on the diagonal).
//
// GEMM: C = op(A) op(B), BF16 operands with arbitrary strides staged through
// LDS in 64 x 32 tiles, v_mfma_f32_16x16x32_bf16 (each wave owns a 32 x 32
// quarter of the 64 x 64 block tile), FP32 accumulation, optional split-K with
// FP32 atomic accumulation for the weight gradients (reduction over the batch).
#include <cmath>
#include <random>

#include "mps_device.hpp"
#include "mps_internal.hpp"

namespace mps {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float kG0 = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kG1 = 0.044715f;

// With u = kG0 (z + kG1 z^3): 0.5 (1 + tanh u) = sigmoid(2u) = 1 / (1 + 2^(-2u log2 e)),
// one hardware exp2 and one reciprocal (v_exp_f32, v_rcp_f32, about 1 ulp
// each) where tanhf costs tens of instructions; the MLP's GELUs outnumber its
// MFMA instructions 2.5 : 1 per row.
constexpr float kGE0 = -2.0f * kG0 * 1.4426950408889634f;  // -2 kG0 log2(e)
constexpr float kGE1 = kGE0 * kG1;
__device__ __forceinline__ float gelu_sig(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * fmaf(kGE1, z * z, kGE0)));
}
__device__ __forceinline__ float gelu_f(float z) { return z * gelu_sig(z); }
// d/dz z s(2u) = s + 2 z s (1 - s) u'(z)
__device__ __forceinline__ float gelu_grad_f(float z) {
  const float s = gelu_sig(z);
  return s + 2.0f * z * s * (1.0f - s) * kG0 * (1.0f + 3.0f * kG1 * z * z);
}

struct GemmArgs {
  int M, N, K;
  const __bf16* A;
  long long sam, sak;  // A(m, k) = A[m * sam + k * sak]
  const __bf16* B;
  long long sbk, sbn;  // B(k, n) = B[k * sbk + n * sbn]
  float* C;
  long long scm, scn;  // C(m, n)
  int atomic;          // 1: C += (split-K partials, FP32 atomics)
  int ksplit;          // K range per blockIdx.z
};

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) __bf16 As[64][40];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[64][40];  // Bs[n][k]
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int kbeg = blockIdx.z * g.ksplit;
  const int kend = (kbeg + g.ksplit < g.K) ? kbeg + g.ksplit : g.K;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int r = t >> 2, kc = (t & 3) * 8;
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    const int m = m0 + r, n = n0 + r;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int k = k0 + kc + e;
      As[r][kc + e] = (m < g.M && k < kend) ? g.A[m * g.sam + k * g.sak] : __bf16(0.0f);
      Bs[r][kc + e] = (n < g.N && k < kend) ? g.B[k * g.sbk + n * g.sbn] : __bf16(0.0f);
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; a++) {
      const bf16x8 af = *reinterpret_cast<const bf16x8*>(&As[wm + a * 16 + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const bf16x8 bfv = *reinterpret_cast<const bf16x8*>(&Bs[wn + b * 16 + (lane & 15)][(lane >> 4) * 8]);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfv, acc[a][b], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = m0 + wm + a * 16 + 4 * (lane >> 4) + v;
        const int n = n0 + wn + b * 16 + (lane & 15);
        if (m < g.M && n < g.N) {
          float* c = g.C + m * g.scm + n * g.scn;
          if (g.atomic) atomicAdd(c, acc[a][b][v]);
          else *c = acc[a][b][v];
        }
      }
}

// z <- z + b (in place, FP32) and h = bf16(gelu(z)) (row-major M x N).
__global__ void __launch_bounds__(256) k_bias_gelu(size_t total, int N, float* __restrict__ z,
                                                   const float* __restrict__ b, __bf16* __restrict__ h) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= total) return;
  const float v = z[i] + b[i % N];
  z[i] = v;
  h[i] = __bf16(gelu_f(v));
}

// y[m] = sum_n gelu(z2[m][n] + b2[n]) w3[n] + b3; one wavefront per row.
// With `a2` set, also stores z2 + b2 there (training).
__global__ void __launch_bounds__(256) k_head(int M, const float* __restrict__ z2,
                                              const float* __restrict__ b2,
                                              const float* __restrict__ w3,
                                              const float* __restrict__ b3, float* __restrict__ y,
                                              float* __restrict__ a2) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= M) return;
  float s = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int n = lane + 64 * q;
    const float v = z2[static_cast<size_t>(row) * kHid + n] + b2[n];
    if (a2) a2[static_cast<size_t>(row) * kHid + n] = v;
    s += gelu_f(v) * w3[n];
  }
  const double ws = wave_sum(static_cast<double>(s));
  if (lane == 0) y[row] = static_cast<float>(ws) + b3[0];
}

// dy = 2 (y - t) / M, loss partial, db3.
__global__ void __launch_bounds__(256) k_loss(int M, const float* __restrict__ y,
                                              const float* __restrict__ t, float* __restrict__ dy,
                                              float* __restrict__ loss, float* __restrict__ db3) {
  __shared__ double lds[4];
  const int m = blockIdx.x * 256 + threadIdx.x;
  double e2 = 0, g = 0;
  if (m < M) {
    const float d = y[m] - t[m];
    const float gd = 2.0f * d / static_cast<float>(M);
    dy[m] = gd;
    e2 = static_cast<double>(d) * d;
    g = gd;
  }
  const double s1 = block_sum<256>(e2, lds);
  const double s2 = block_sum<256>(g, lds);
  if (threadIdx.x == 0) {
    atomicAdd(loss, static_cast<float>(s1 / M));
    atomicAdd(db3, static_cast<float>(s2));
  }
}

// Head backward over 64 rows per block, one thread per hidden column:
// dW3[n] += sum h2 dy, da2 = dy w3 gelu'(a2) (FP32 + BF16), db2 += sum da2.
__global__ void __launch_bounds__(256) k_head_bwd(int M, const float* __restrict__ a2,
                                                  const float* __restrict__ dy,
                                                  const float* __restrict__ w3, float* __restrict__ dw3,
                                                  float* __restrict__ db2, __bf16* __restrict__ da2b) {
  const int n = threadIdx.x;
  const int r0 = blockIdx.x * 64;
  float aw = 0, ab = 0;
  const float wn = w3[n];
  for (int r = r0; r < r0 + 64 && r < M; r++) {
    const float a = a2[static_cast<size_t>(r) * kHid + n];
    const float d = dy[r];
    aw += gelu_f(a) * d;
    const float g = d * wn * gelu_grad_f(a);
    ab += g;
    da2b[static_cast<size_t>(r) * kHid + n] = __bf16(g);
  }
  atomicAdd(&dw3[n], aw);
  atomicAdd(&db2[n], ab);
}

// da1 = dh1 gelu'(a1) (BF16), db1 += column sums; 64 rows per block.
__global__ void __launch_bounds__(256) k_gelu_bwd(int M, const float* __restrict__ a1,
                                                  const float* __restrict__ dh1,
                                                  float* __restrict__ db1, __bf16* __restrict__ da1b) {
  const int n = threadIdx.x;
  const int r0 = blockIdx.x * 64;
  float ab = 0;
  for (int r = r0; r < r0 + 64 && r < M; r++) {
    const size_t o = static_cast<size_t>(r) * kHid + n;
    const float g = dh1[o] * gelu_grad_f(a1[o]);
    ab += g;
    da1b[o] = __bf16(g);
  }
  atomicAdd(&db1[n], ab);
}

__global__ void __launch_bounds__(256) k_adam(size_t n, float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr,
                                              float c1, float c2) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

__global__ void __launch_bounds__(256) k_to_bf16(size_t n, const float* __restrict__ a,
                                                 __bf16* __restrict__ b) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i < n) b[i] = __bf16(a[i]);
}

// Host-side rows (m x 132 FP32) -> BF16 rows padded to 160.
__global__ void __launch_bounds__(256) k_pack_rows(size_t m, const float* __restrict__ x,
                                                   __bf16* __restrict__ xb) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= m * kFeatPad) return;
  const size_t r = i / kFeatPad, k = i % kFeatPad;
  xb[i] = (k < static_cast<size_t>(kFeat)) ? __bf16(x[r * kFeat + k]) : __bf16(0.0f);
}

__global__ void __launch_bounds__(256) k_gather_rows(size_t m, const uint32_t* __restrict__ idx,
                                                     const __bf16* __restrict__ src_x,
                                                     const float* __restrict__ src_t,
                                                     __bf16* __restrict__ xb, float* __restrict__ tb) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= m * (kFeatPad / 8)) return;
  const size_t r = i / (kFeatPad / 8), c = i % (kFeatPad / 8);
  const size_t s = idx[r];
  reinterpret_cast<uint4*>(xb + r * kFeatPad)[c] = reinterpret_cast<const uint4*>(src_x + s * kFeatPad)[c];
  if (c == 0) tb[r] = src_t[s];
}

struct FeatArgs {
  size_t n;
  const double *x, *z, *nd;
  const uint8_t* type;
  const FeatRec* rec;
  const double* scales;  // [7] = 1/s_p, [8] = 1/s_r
  // the assembled system's rows: their off-diagonal columns are the accepted
  // neighbours (non-dummy, r^2 < r_lap^2), ascending (SELL-64 slices: int16
  // offsets from the row, or u32 columns in wide slices)
  const uint64_t* off;
  const uint32_t* col;
  const int16_t* col16;
  const uint8_t* wide;
  const uint32_t* len;
  double inv_l0, n0;
  __bf16* out;
};

// Column k of row i of the system.
__device__ __forceinline__ uint32_t feat_col(const FeatArgs& a, size_t i, bool wide, uint64_t base, uint32_t k) {
  const size_t at = base + static_cast<size_t>(k) * 64 + (i & 63);
  return wide ? a.col[at] : static_cast<uint32_t>(static_cast<int>(i) + a.col16[at]);
}

// Features of every particle (zero rows for dummy / disabled particles).
// One thread per particle.  Its accepted neighbours are the off-diagonal
// columns of its row of the assembled system (the same test, r^2 < r_lap^2
// and not dummy, picked them), read as coalesced SELL slots; their
// positions come from the records.  A row of at most kNbr columns keeps
// them all as (float(r^2), index) pairs in registers, sorted by a bitonic
// network (distance, then ascending storage index of the neighbour).  A longer row goes
// through an insertion sort in LDS (keys and column positions Q, 8 or 16
// bits: 160-192 bytes per thread; stable in column order) that keeps the
// kNbr nearest.
constexpr int kFeatWalk = 8;    // columns whose loads are issued together
constexpr int kFeatGather = 4;  // ranks whose neighbours' state is gathered together
static_assert(kNbr % kFeatGather == 0 && kNbr % kFeatWalk == 0, "whole groups");

template <class Q>
__global__ void __launch_bounds__(256) k_features(FeatArgs a) {
  __shared__ uint32_t skey[kNbr][256];  // float bits: non-negative, so ordered as unsigned
  __shared__ Q sq[kNbr][256];
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const size_t i = blk * 256ull + threadIdx.x;
  const bool live = i < a.n;  // every thread reaches the barriers of the row output
  const int t = threadIdx.x;
  const uint8_t ti = live ? a.type[i] : kDummy;
  const double xi = live ? a.x[i] : 0.0, zi = live ? a.z[i] : 0.0;
  const double inv_sp = a.scales[7], inv_sr = a.scales[8];
  const size_t sl = i >> 6;
  const bool wide = live && a.wide[sl] != 0;
  const uint64_t base = live ? a.off[sl] : 0;
  float f[4] = {0, 0, 0, 0};
  int m = 0;
  bool sorted = false;
  // the kept neighbours as (float bits of r^2) << 32 | index: ascending
  // order is distance order with equal distances by storage index (= column order);
  // unused entries are all ones and sort last
  uint64_t kv[kNbr];
#pragma unroll
  for (int e = 0; e < kNbr; e++) kv[e] = ~0ull;
  if (ti != kDummy && ti != kDisabled) {
    const uint32_t cnt = a.len[i];
    auto key = [&](uint32_t j) {
      const double2 xz = *reinterpret_cast<const double2*>(&a.rec[j]);
      const double dx = xz.x - xi, dz = xz.y - zi;
      return __float_as_uint(static_cast<float>(dx * dx + dz * dz));
    };
    if (cnt <= static_cast<uint32_t>(kNbr)) {  // every column is kept
      m = static_cast<int>(cnt);
#pragma unroll
      for (int k0 = 0; k0 < kNbr; k0 += kFeatWalk) {
        if (k0 < m) {
          uint32_t jj[kFeatWalk];
#pragma unroll
          for (int u = 0; u < kFeatWalk; u++)
            jj[u] = (k0 + u < m) ? feat_col(a, i, wide, base, k0 + u) : static_cast<uint32_t>(i);
#pragma unroll
          for (int u = 0; u < kFeatWalk; u++) {
            const uint64_t kf = key(jj[u]);
            if (k0 + u < m) kv[k0 + u] = (kf << 32) | jj[u];
          }
        }
      }
    } else {  // keep the kNbr nearest: stable insertion sort in column order
      for (uint32_t k = 0; k < cnt; k++) {
        const uint32_t kf = key(feat_col(a, i, wide, base, k));
        int pos;
        if (m < kNbr) pos = m++;
        else if (kf < skey[kNbr - 1][t]) pos = kNbr - 1;  // an equal key with a lower index ranks first
        else continue;
        while (pos > 0 && skey[pos - 1][t] > kf) {
          skey[pos][t] = skey[pos - 1][t];
          sq[pos][t] = sq[pos - 1][t];
          pos--;
        }
        skey[pos][t] = kf;
        sq[pos][t] = static_cast<Q>(k);
      }
#pragma unroll
      for (int e = 0; e < kNbr; e++)
        if (e < m) kv[e] = (static_cast<uint64_t>(skey[e][t]) << 32) | feat_col(a, i, wide, base, sq[e][t]);
      sorted = true;
    }
    f[0] = static_cast<float>(cnt) / 32.0f;
    f[1] = static_cast<float>((a.nd[i] - a.n0) / a.n0);
    f[2] = static_cast<float>(a.rec[i].p * inv_sp);
    f[3] = static_cast<float>(a.rec[i].r0 * inv_sr);
  }
  // bitonic sort of the kNbr pairs in registers (ascending)
  if (!sorted && m > 1) {
#pragma unroll
    for (int k = 2; k <= kNbr; k <<= 1)
#pragma unroll
      for (int d = k >> 1; d > 0; d >>= 1)
#pragma unroll
        for (int e = 0; e < kNbr; e++) {
          const int l = e ^ d;
          if (l > e) {
            const uint64_t lo = kv[e] < kv[l] ? kv[e] : kv[l];
            const uint64_t hi = kv[e] < kv[l] ? kv[l] : kv[e];
            const bool up = (e & k) == 0;
            kv[e] = up ? lo : hi;
            kv[l] = up ? hi : lo;
          }
        }
  }
  // the row in registers (elements 0-3: the particle's own features;
  // 4 + 4s .. 7 + 4s: the neighbour of rank s, zero for s >= m; zeros from
  // kFeat on), gathered four neighbours at a time
  bf16x4 val[kNbr + 1];
  val[0] = bf16x4{__bf16(f[0]), __bf16(f[1]), __bf16(f[2]), __bf16(f[3])};
  const bf16x4 zero4 = bf16x4{__bf16(0.f), __bf16(0.f), __bf16(0.f), __bf16(0.f)};
#pragma unroll
  for (int s0 = 0; s0 < kNbr; s0 += kFeatGather) {
#pragma unroll
    for (int u = 0; u < kFeatGather; u++) val[1 + s0 + u] = zero4;
    if (s0 < m) {
      uint32_t jj[kFeatGather];
#pragma unroll
      for (int u = 0; u < kFeatGather; u++)
        jj[u] = (s0 + u < m) ? static_cast<uint32_t>(kv[s0 + u]) : static_cast<uint32_t>(i);
#pragma unroll
      for (int u = 0; u < kFeatGather; u++) {
        const FeatRec rj = a.rec[jj[u]];
        const bf16x4 v = bf16x4{__bf16(static_cast<float>((rj.x - xi) * a.inv_l0)),
                                __bf16(static_cast<float>((rj.z - zi) * a.inv_l0)),
                                __bf16(static_cast<float>(rj.p * inv_sp)), __bf16(static_cast<float>(rj.r0 * inv_sr))};
        val[1 + s0 + u] = (s0 + u < m) ? v : zero4;
      }
    }
  }
  static_assert(kFeat == 4 + 4 * kNbr && (kNbr + 1) % 2 == 1 && kFeatPad % 8 == 0, "row layout");
  // The rows leave through LDS (the key array, free once every thread is
  // past its sort), 16 rows of a wavefront at a time, as contiguous 16-byte
  // vectors: stored straight from the registers, each wavefront store would
  // touch 64 rows 320 bytes apart.
  constexpr int kStageRows = 16;
  constexpr int kStagePitch = kFeatPad * 2 + 16;  // bytes; the pad spreads a row's vectors over the banks
  constexpr int kRowVec = kFeatPad / 8;
  static_assert(4 * kStageRows * kStagePitch <= static_cast<int>(sizeof(skey)), "staging fits the keys");
  char* const stage = reinterpret_cast<char*>(&skey[0][0]) + (t >> 6) * (kStageRows * kStagePitch);
  const int lane = t & 63;
  const size_t wrow0 = blk * 256ull + (t & ~63);  // the wavefront's first row
  __syncthreads();
#pragma unroll 1
  for (int h = 0; h < 64 / kStageRows; h++) {
    if ((lane / kStageRows) == h) {
      char* const dst = stage + (lane % kStageRows) * kStagePitch;
#pragma unroll
      for (int vv = 0; vv < kRowVec; vv++) {
        const bf16x4 lo = (2 * vv < kNbr + 1) ? val[2 * vv] : zero4;
        const bf16x4 hi = (2 * vv + 1 < kNbr + 1) ? val[2 * vv + 1] : zero4;
        *reinterpret_cast<bf16x8*>(dst + 16 * vv) = bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
      }
    }
    __syncthreads();
    const size_t r0 = wrow0 + h * kStageRows;
#pragma unroll
    for (int k = 0; k < (kStageRows * kRowVec + 63) / 64; k++) {
      const int v = lane + 64 * k, rr = v / kRowVec, cc = v % kRowVec;
      if (v < kStageRows * kRowVec && r0 + rr < a.n)
        *reinterpret_cast<uint4*>(a.out + (r0 + rr) * kFeatPad + 8 * cc) =
            *reinterpret_cast<const uint4*>(stage + rr * kStagePitch + 16 * cc);
    }
    __syncthreads();
  }
}

// sol = p_prev + s_d y on alive rows (s_d = scales[9]).
__global__ void __launch_bounds__(256) k_apply_pred(size_t n, const uint8_t* __restrict__ type,
                                                    const double* __restrict__ p,
                                                    const float* __restrict__ y, const double* __restrict__ scales,
                                                    double* __restrict__ sol) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  const uint8_t t = type[i];
  if (t != kDummy && t != kDisabled) sol[i] = p[i] + scales[9] * static_cast<double>(y[i]);
}

// The prediction is kept when its FP64 residual red[0] = |b - A sol|^2 is
// below kKeep times red[6] = |b - A p_prev|^2, which the feature pass computed
// with the FP32 operator (FP32 rounding moves it by far less than the margin).
constexpr double kKeep = 0.98;
__host__ __device__ __forceinline__ bool keep_prediction(double rr_pred, double rr_prev) {
  return rr_pred < kKeep * rr_prev;
}

// sol = p_prev on alive rows unless the prediction is kept (keep_prediction).
__global__ void __launch_bounds__(256) k_revert_unless_better(size_t n, const uint8_t* __restrict__ type,
                                                              const double* __restrict__ p,
                                                              const double* __restrict__ red,
                                                              double* __restrict__ sol) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n || keep_prediction(red[0], red[6])) return;
  const uint8_t t = type[i];
  if (t != kDummy && t != kDisabled) sol[i] = p[i];
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct CollectArgs {
  size_t n;
  const uint8_t* type;
  const __bf16* feats;
  const double *sol, *p, *x, *z;  // solution, previous pressure, positions
  const uint32_t *ids, *count;    // neighbour lists
  uint32_t kstride;
  double rl2, inv_sd;
  uint64_t seed;
  uint32_t threshold;
  size_t cap;
  unsigned long long* used;
  __bf16* dx;
  float* dt;
};

// Appends (features, target) of a hashed subset of alive rows.  The target is
// the local part of the correction delta = sol - p_prev: delta_i minus the
// mean of delta over i and its Laplacian neighbours (the matrix pattern),
// divided by s_d.  The smooth remainder of delta is not a function of a
// particle's neighbourhood; the local part is, and it carries most of the
// residual A delta.
__global__ void __launch_bounds__(256) k_collect(CollectArgs a) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= a.n) return;
  const uint8_t t = a.type[i];
  if (t == kDummy || t == kDisabled) return;
  if (static_cast<uint32_t>(mix64(a.seed ^ (i * 0x2545F4914F6CDD1Dull)) >> 32) >= a.threshold) return;
  const double di = a.sol[i] - a.p[i];
  double sum = di;
  int cnt = 1;
  const double xi = a.x[i], zi = a.z[i];
  const uint32_t m = a.count[i];
  for (uint32_t q = 0; q < m; q++) {
    const uint32_t j = a.ids[nbr_slot(static_cast<uint32_t>(i), q, a.kstride)];
    if (a.type[j] == kDummy) continue;
    const double ex = a.x[j] - xi, ez = a.z[j] - zi;
    if (!(ex * ex + ez * ez < a.rl2)) continue;
    sum += a.sol[j] - a.p[j];
    cnt++;
  }
  const unsigned long long slot = atomicAdd(a.used, 1ull);
  if (slot >= a.cap) return;
  const uint4* src = reinterpret_cast<const uint4*>(a.feats + i * kFeatPad);
  uint4* dst = reinterpret_cast<uint4*>(a.dx + slot * kFeatPad);
  for (int c = 0; c < kFeatPad / 8; c++) dst[c] = src[c];
  a.dt[slot] = static_cast<float>((di - sum / cnt) * a.inv_sd);
}


// ---- fused inference ---------------------------------------------------------
// y = w3 . gelu(W2 gelu(W1 x + b1) + b2) + b3 in one kernel, nothing but x
// read and y written.  Both layers are computed transposed (hidden units as
// the MFMA M dimension, rows as N) with v_mfma_f32_16x16x32_bf16, whose lane
// layout is: A row l%16 / B column l%16 with K elements 8(l/16)..8(l/16)+7;
// C[4(l/16)+v][l%16], v = 0..3.
//
// Weight-stationary and persistent: one block of kMlpWaves wavefronts per CU;
// wavefront w holds, in registers for the block's lifetime, the A operands of
// hidden units 32w..32w+31 (W1) and of outputs 32w..32w+31 (W2), and the
// block walks batches of kBatch rows.  Per batch the rows go to LDS (loaded
// into registers during the previous batch), every wavefront computes its 32
// hidden units of every row into an LDS tile h1, then its 32 outputs, and the
// head's per-wavefront partial sums meet in LDS.  Weights are read from
// memory once per block; the per-row LDS traffic is the B operands (x, h1).
//
// Layer 1 leaves lane (g, c) of wavefront w with hidden units
// 32w + 16mt + 4g + v (tile mt = 0, 1) of row c: one 16-byte vector whose
// element 4mt + v is layer 2's K slot 8g + 4mt + v of K step w.  That K slot
// stands for hidden unit w2_hidden(w, g, 4mt + v) below, and W2 is stored in
// that order (k_perm_w2) so that h1 is written and read as 16-byte vectors.

__host__ __device__ __forceinline__ int w2_hidden(int ks, int g, int e) {
  return 32 * ks + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4));
}

__global__ void __launch_bounds__(256) k_perm_w2(const float* __restrict__ w2, __bf16* __restrict__ w2p) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kHid * kHid) return;
  const int o = i / kHid, k = i % kHid;
  const int ks = k / 32, g = (k % 32) / 8, e = k % 8;
  w2p[i] = static_cast<__bf16>(w2[o * kHid + w2_hidden(ks, g, e)]);
}

constexpr int kMlpWaves = 8;
constexpr int kMlpThreads = 64 * kMlpWaves;
constexpr int kBatch = 64;             // rows per block iteration: 4 MFMA column tiles
constexpr int kXPitch = kFeatPad + 8;  // LDS row pitches (bf16), padded against bank conflicts
constexpr int kHPitch = kHid + 8;
constexpr int kXVec = kBatch * (kFeatPad / 8);  // 16-byte vectors of one batch of rows
constexpr int kXPer = (kXVec + kMlpThreads - 1) / kMlpThreads;
static_assert(kHid == 32 * kMlpWaves, "one 32-unit slice of each layer per wavefront");

__global__ void __launch_bounds__(kMlpThreads) k_mlp_infer(size_t M, const __bf16* __restrict__ x,
                                                           const __bf16* __restrict__ w1,
                                                           const float* __restrict__ b1,
                                                           const __bf16* __restrict__ w2p,
                                                           const float* __restrict__ b2,
                                                           const float* __restrict__ w3,
                                                           const float* __restrict__ b3, float* __restrict__ y) {
  __shared__ __attribute__((aligned(16))) __bf16 xs[kBatch * kXPitch];
  __shared__ __attribute__((aligned(16))) __bf16 hs[kBatch * kHPitch];
  __shared__ float red[kMlpWaves][kBatch];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, c = lane & 15;
  constexpr int K1 = kFeatPad / 32, K2 = kHid / 32;
  const size_t nbatch = (M + kBatch - 1) / kBatch;
  if (blockIdx.x >= nbatch) return;  // block-uniform, before any barrier
  // this wavefront's weights, biases and head coefficients
  bf16x8 a1[2][K1], a2[2][K2];
  float bias1[2][4], bias2[2][4], head[2][4];
#pragma unroll
  for (int mt = 0; mt < 2; mt++) {
    const int h = 32 * w + 16 * mt + c;
#pragma unroll
    for (int ks = 0; ks < K1; ks++) a1[mt][ks] = *reinterpret_cast<const bf16x8*>(w1 + h * kFeatPad + 32 * ks + 8 * g);
#pragma unroll
    for (int ks = 0; ks < K2; ks++) a2[mt][ks] = *reinterpret_cast<const bf16x8*>(w2p + h * kHid + 32 * ks + 8 * g);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int o = 32 * w + 16 * mt + 4 * g + v;
      bias1[mt][v] = b1[o];
      bias2[mt][v] = b2[o];
      head[mt][v] = w3[o];
    }
  }
  const float bias3 = b3[0];
  // the next batch's rows, global -> registers -> LDS (rows past M read as zero)
  uint4 xr[kXPer];
  auto fetch = [&](size_t bt) {
#pragma unroll
    for (int v = 0; v < kXPer; v++) {
      const int e = tid + v * kMlpThreads;
      const size_t r = bt * kBatch + e / (kFeatPad / 8);
      xr[v] = (e < kXVec && r < M)
                  ? reinterpret_cast<const uint4*>(x + r * kFeatPad)[e % (kFeatPad / 8)]
                  : make_uint4(0u, 0u, 0u, 0u);
    }
  };
  auto put = [&] {
#pragma unroll
    for (int v = 0; v < kXPer; v++) {
      const int e = tid + v * kMlpThreads;
      if (e < kXVec)
        *reinterpret_cast<uint4*>(&xs[(e / (kFeatPad / 8)) * kXPitch + (e % (kFeatPad / 8)) * 8]) = xr[v];
    }
  };
  fetch(blockIdx.x);
  for (size_t bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
    // xs was last read before the previous batch's second barrier, red
    // before this barrier by wavefront 0
    put();
    __syncthreads();
    if (bt + gridDim.x < nbatch) fetch(bt + gridDim.x);
#pragma unroll
    for (int nt = 0; nt < kBatch / 16; nt++) {
      const __bf16* xrow = xs + (16 * nt + c) * kXPitch + 8 * g;
      f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
      for (int ks = 0; ks < K1; ks++) {
        const bf16x8 bfr = *reinterpret_cast<const bf16x8*>(xrow + 32 * ks);
        acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[0][ks], bfr, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[1][ks], bfr, acc[1], 0, 0, 0);
      }
      bf16x8 hv;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        hv[v] = static_cast<__bf16>(gelu_f(acc[0][v] + bias1[0][v]));
        hv[4 + v] = static_cast<__bf16>(gelu_f(acc[1][v] + bias1[1][v]));
      }
      *reinterpret_cast<bf16x8*>(&hs[(16 * nt + c) * kHPitch + 32 * w + 8 * g]) = hv;
    }
    __syncthreads();  // h1 complete
#pragma unroll
    for (int nt = 0; nt < kBatch / 16; nt++) {
      const __bf16* hrow = hs + (16 * nt + c) * kHPitch + 8 * g;
      f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
      for (int ks = 0; ks < K2; ks++) {
        const bf16x8 bfr = *reinterpret_cast<const bf16x8*>(hrow + 32 * ks);
        acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[0][ks], bfr, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[1][ks], bfr, acc[1], 0, 0, 0);
      }
      float part = 0.f;
#pragma unroll
      for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int v = 0; v < 4; v++) part = fmaf(head[mt][v], gelu_f(acc[mt][v] + bias2[mt][v]), part);
      part += __shfl_xor(part, 16, 64);
      part += __shfl_xor(part, 32, 64);
      if (g == 0) red[w][16 * nt + c] = part;
    }
    __syncthreads();  // partial sums complete; hs free
    if (tid < kBatch) {
      const size_t r = bt * kBatch + tid;
      if (r < M) {
        float s = bias3;
#pragma unroll
        for (int k = 0; k < kMlpWaves; k++) s += red[k][tid];
        y[r] = s;
      }
    }
  }
}

unsigned blocks(size_t n) { return static_cast<unsigned>((n + 255) / 256); }

void gemm(hipStream_t s, int M, int N, int K, const __bf16* A, long long sam, long long sak,
          const __bf16* B, long long sbk, long long sbn, float* C, long long scm, long long scn,
          int splits) {
  GemmArgs g{M, N, K, A, sam, sak, B, sbk, sbn, C, scm, scn, splits > 1 ? 1 : 0, 0};
  int ks = (K + splits - 1) / splits;
  ks = (ks + 31) / 32 * 32;
  g.ksplit = ks;
  const int nz = (K + ks - 1) / ks;
  hipLaunchKernelGGL(k_gemm_bf16, dim3((N + 63) / 64, (M + 63) / 64, nz), dim3(256), 0, s, g);
}

}  // namespace

// ------------------------------- Model ----------------------------------------

size_t Model::param_count(int k) {
  static const size_t sz[6] = {static_cast<size_t>(kHid) * kFeatPad, kHid,
                               static_cast<size_t>(kHid) * kHid, kHid, kHid, 1};
  return sz[k];
}

void Model::ensure_batch(size_t m) {
  if (m <= cap) return;
  dfree(xb); dfree(tb); dfree(z1); dfree(z2); dfree(a2); dfree(h1b); dfree(y); dfree(dy);
  dfree(da2b); dfree(dh1); dfree(da1b); dfree(idx);
  xb = dalloc<__bf16>(m * kFeatPad);
  tb = dalloc<float>(m);
  z1 = dalloc<float>(m * kHid);
  z2 = dalloc<float>(m * kHid);
  a2 = dalloc<float>(m * kHid);
  h1b = dalloc<__bf16>(m * kHid);
  y = dalloc<float>(m);
  dy = dalloc<float>(m);
  da2b = dalloc<__bf16>(m * kHid);
  dh1 = dalloc<float>(m * kHid);
  da1b = dalloc<__bf16>(m * kHid);
  idx = dalloc<uint32_t>(m);
  cap = m;
}

void Model::init(int seed) {
  MPS_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  int dev = 0;
  MPS_HIP(hipGetDevice(&dev));
  MPS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  for (int k = 0; k < 6; k++) {
    const size_t n = param_count(k);
    w[k] = dalloc<float>(n);
    g[k] = dalloc<float>(n);
    m1[k] = dalloc<float>(n);
    m2[k] = dalloc<float>(n);
    MPS_HIP(hipMemset(m1[k], 0, n * 4));
    MPS_HIP(hipMemset(m2[k], 0, n * 4));
  }
  w1b = dalloc<__bf16>(param_count(0));
  w2b = dalloc<__bf16>(param_count(2));
  w2p = dalloc<__bf16>(param_count(2));
  loss = dalloc<float>(1);
  // U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases (seeded)
  std::mt19937_64 rng(static_cast<uint64_t>(seed));
  const int fan[6] = {kFeat, kFeat, kHid, kHid, kHid, kHid};
  for (int k = 0; k < 6; k++) {
    std::vector<float> h(param_count(k), 0.0f);
    std::uniform_real_distribution<float> d(-1.0f / std::sqrt(static_cast<float>(fan[k])),
                                            1.0f / std::sqrt(static_cast<float>(fan[k])));
    for (size_t i = 0; i < h.size(); i++) {
      if (k == 0 && (i % kFeatPad) >= static_cast<size_t>(kFeat)) continue;  // K padding stays 0
      h[i] = d(rng);
    }
    MPS_HIP(hipMemcpy(w[k], h.data(), h.size() * 4, hipMemcpyHostToDevice));
  }
  refresh_bf16();
}

void Model::refresh_bf16() {
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(param_count(0))), dim3(256), 0, stream, param_count(0),
                     static_cast<const float*>(w[0]), w1b);
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(param_count(2))), dim3(256), 0, stream, param_count(2),
                     static_cast<const float*>(w[2]), w2b);
  hipLaunchKernelGGL(k_perm_w2, dim3(blocks(param_count(2))), dim3(256), 0, stream,
                     static_cast<const float*>(w[2]), w2p);
  check_launch();
  MPS_HIP(hipStreamSynchronize(stream));
}

void Model::release() {
  for (int k = 0; k < 6; k++) {
    dfree(w[k]); dfree(g[k]); dfree(m1[k]); dfree(m2[k]);
  }
  dfree(w1b); dfree(w2b); dfree(w2p); dfree(loss);
  dfree(xb); dfree(tb); dfree(z1); dfree(z2); dfree(a2); dfree(h1b); dfree(y); dfree(dy);
  dfree(da2b); dfree(dh1); dfree(da1b); dfree(idx);
  if (stream) (void)hipStreamDestroy(stream);
  stream = nullptr;
}

// Forward pass of m rows already in xb (BF16, 160 wide) -> y; with `train`
// the pre-activations stay in z1 (+b1) and a2 for the backward pass.
void Model::forward(hipStream_t s, size_t m, const __bf16* x, float* out, bool train) {
  if (m == 0) return;
  if (!train) {  // inference: the fused kernel, no intermediate in memory
    const size_t nbatch = (m + kBatch - 1) / kBatch;
    const unsigned grid = static_cast<unsigned>(std::min<size_t>(nbatch, static_cast<size_t>(cus > 0 ? cus : 256)));
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(kMlpThreads), 0, s, m, x,
                       static_cast<const __bf16*>(w1b), static_cast<const float*>(w[1]),
                       static_cast<const __bf16*>(w2p), static_cast<const float*>(w[3]),
                       static_cast<const float*>(w[4]), static_cast<const float*>(w[5]), out);
    check_launch();
    return;
  }
  ensure_batch(m);
  const int M = static_cast<int>(m);
  gemm(s, M, kHid, kFeatPad, x, kFeatPad, 1, w1b, 1, kFeatPad, z1, kHid, 1, 1);
  hipLaunchKernelGGL(k_bias_gelu, dim3(blocks(m * kHid)), dim3(256), 0, s, m * kHid, kHid, z1,
                     static_cast<const float*>(w[1]), h1b);
  gemm(s, M, kHid, kHid, h1b, kHid, 1, w2b, 1, kHid, z2, kHid, 1, 1);
  hipLaunchKernelGGL(k_head, dim3((M + 3) / 4), dim3(256), 0, s, M, static_cast<const float*>(z2),
                     static_cast<const float*>(w[3]), static_cast<const float*>(w[4]),
                     static_cast<const float*>(w[5]), out, train ? a2 : static_cast<float*>(nullptr));
  check_launch();
}

// Gradients of the mean squared error for the batch in xb / tb (into g[]).
void Model::backward(hipStream_t s, size_t m) {
  const int M = static_cast<int>(m);
  for (int k = 0; k < 6; k++) MPS_HIP(hipMemsetAsync(g[k], 0, param_count(k) * 4, s));
  MPS_HIP(hipMemsetAsync(loss, 0, 4, s));
  hipLaunchKernelGGL(k_loss, dim3(blocks(m)), dim3(256), 0, s, M, static_cast<const float*>(y),
                     static_cast<const float*>(tb), dy, loss, g[5]);
  hipLaunchKernelGGL(k_head_bwd, dim3((M + 63) / 64), dim3(256), 0, s, M, static_cast<const float*>(a2),
                     static_cast<const float*>(dy), static_cast<const float*>(w[4]), g[4], g[3], da2b);
  const int splits = std::max(1, std::min(64, M / 512));
  // dW2[n][k] = sum_m da2[m][n] h1[m][k]
  gemm(s, kHid, kHid, M, da2b, 1, kHid, h1b, kHid, 1, g[2], kHid, 1, splits);
  // dh1[m][k] = sum_n da2[m][n] W2[n][k]
  gemm(s, M, kHid, kHid, da2b, kHid, 1, w2b, kHid, 1, dh1, kHid, 1, 1);
  hipLaunchKernelGGL(k_gelu_bwd, dim3((M + 63) / 64), dim3(256), 0, s, M, static_cast<const float*>(z1),
                     static_cast<const float*>(dh1), g[1], da1b);
  // dW1[n][k] = sum_m da1[m][n] x[m][k]
  gemm(s, kHid, kFeatPad, M, da1b, 1, kHid, xb, kFeatPad, 1, g[0], kFeatPad, 1, splits);
  check_launch();
}

void Model::adam(hipStream_t s, float lr) {
  adam_t++;
  const float c1 = 1.0f - std::pow(0.9f, static_cast<float>(adam_t));
  const float c2 = 1.0f - std::pow(0.999f, static_cast<float>(adam_t));
  for (int k = 0; k < 6; k++)
    hipLaunchKernelGGL(k_adam, dim3(blocks(param_count(k))), dim3(256), 0, s, param_count(k), w[k],
                       static_cast<const float*>(g[k]), m1[k], m2[k], lr, c1, c2);
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(param_count(0))), dim3(256), 0, s, param_count(0),
                     static_cast<const float*>(w[0]), w1b);
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(param_count(2))), dim3(256), 0, s, param_count(2),
                     static_cast<const float*>(w[2]), w2b);
  hipLaunchKernelGGL(k_perm_w2, dim3(blocks(param_count(2))), dim3(256), 0, s, static_cast<const float*>(w[2]),
                     w2p);
  check_launch();
}

void Model::pack_host_rows(size_t m, const float* x_host) {
  ensure_batch(m);
  float* tmp = dalloc<float>(m * kFeat);
  MPS_HIP(hipMemcpyAsync(tmp, x_host, m * kFeat * 4, hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_pack_rows, dim3(blocks(m * kFeatPad)), dim3(256), 0, stream, m,
                     static_cast<const float*>(tmp), xb);
  check_launch();
  MPS_HIP(hipStreamSynchronize(stream));
  dfree(tmp);
}

// Epochs of shuffled mini-batches over the dataset; returns the mean loss of
// each epoch in losses[].
void Model::train(Dataset& ds, int epochs, int batch, float lr, uint64_t seed, double* losses) {
  const size_t n = ds.size();
  if (n == 0 || batch <= 0) throw Failure{MPS_ERR_BAD_ARG, "empty dataset or batch"};
  const size_t bs = std::min<size_t>(static_cast<size_t>(batch), n);
  ensure_batch(bs);
  std::vector<uint32_t> perm(n);
  for (size_t i = 0; i < n; i++) perm[i] = static_cast<uint32_t>(i);
  std::mt19937_64 rng(seed);
  uint32_t* dperm = dalloc<uint32_t>(n);
  for (int e = 0; e < epochs; e++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    MPS_HIP(hipMemcpyAsync(dperm, perm.data(), n * 4, hipMemcpyHostToDevice, stream));
    double sum = 0;
    int nb = 0;
    for (size_t b0 = 0; b0 + bs <= n; b0 += bs) {
      hipLaunchKernelGGL(k_gather_rows, dim3(blocks(bs * (kFeatPad / 8))), dim3(256), 0, stream, bs,
                         static_cast<const uint32_t*>(dperm + b0), static_cast<const __bf16*>(ds.x),
                         static_cast<const float*>(ds.t), xb, tb);
      forward(stream, bs, xb, y, true);
      backward(stream, bs);
      adam(stream, lr);
      float l = 0;
      MPS_HIP(hipMemcpyAsync(&l, loss, 4, hipMemcpyDeviceToHost, stream));
      MPS_HIP(hipStreamSynchronize(stream));
      sum += l;
      nb++;
    }
    if (losses) losses[e] = nb ? sum / nb : 0.0;
  }
  dfree(dperm);
}

void Dataset::init(size_t capacity) {
  cap = capacity;
  x = dalloc<__bf16>(cap * kFeatPad);
  t = dalloc<float>(cap);
  count = dalloc<unsigned long long>(1);
  MPS_HIP(hipMemset(count, 0, sizeof(unsigned long long)));
}

size_t Dataset::size() const {
  unsigned long long h = 0;
  MPS_HIP(hipMemcpy(&h, count, sizeof(h), hipMemcpyDeviceToHost));
  return std::min<size_t>(h, cap);
}

void Dataset::release() {
  dfree(x);
  dfree(t);
  dfree(count);
}

// --------------------------- use inside a step -------------------------------------

// The residual of p_prev, the scales, and the features of every particle,
// all on the stream without a synchronisation (the scales stay on the
// device, in red_out).
void learned_features(Context& c) {
  const size_t n = c.n;
  if (n == 0) return;
  if (c.feats_borrowed) {  // rows and predictions in a batch's buffers: the records only
    if (!c.feat_rec || c.feats_cap < n) {
      dfree(c.feat_rec);
      c.feat_rec = dalloc<FeatRec>(n);
      c.feats_cap = n;
    }
  } else if (!c.feats || c.feats_cap < n) {
    dfree(c.feats);
    dfree(c.ypred);
    dfree(c.feat_rec);
    c.feats = dalloc<__bf16>(n * kFeatPad);
    c.ypred = dalloc<float>(n);
    c.feat_rec = dalloc<FeatRec>(n);
    c.feats_cap = n;
  }
  if (c.kmax > 65536u) throw Failure{MPS_ERR_BAD_ARG, "row length bound exceeds 16-bit positions"};
  FeatRec* rec = static_cast<FeatRec*>(c.feat_rec);
  // the residual r0 of x = p_prev (a feature input; also the revert test's
  // reference), the records and the scales, in FP32 in one pass
  if (!learned_residual32(c, rec))
    throw Failure{MPS_ERR_STATE, "learned features need the pressure system assembled from the particles"};
  const FeatArgs a{n, c.x, c.z, c.nd, c.type, rec, c.red_out, c.slice_off, c.scol, c.col16, c.slice_wide,
                   c.row_len, 1.0 / c.env.l0, c.env.n0, c.feats};
  // per particle: x, z, n, type, row length, the records (each once) and the
  // row written; the system's column offsets
  tracked(c, MPS_K_FEATURES, n * (8.0 * 3 + 1 + 4 + 32 + kFeatPad * 2) + 2.0 * c.nnz_offdiag, [&] {
    if (c.kmax <= 256u)  // column positions fit 8 bits
      hipLaunchKernelGGL(k_features<uint8_t>, dim3(blocks(n)), dim3(256), 0, c.stream, a);
    else
      hipLaunchKernelGGL(k_features<uint16_t>, dim3(blocks(n)), dim3(256), 0, c.stream, a);
  });
  check_launch();
}

// MPS_WARM_LEARNED: x0 = p_prev + s_d MLP(features), kept only when it lowers
// the FP64 residual (decided on the device).  Nothing waits here: a mixed
// solve reads the outcome at its first synchronisation (learned_resolve),
// which finds r = b - A x0 and its norms already computed; the FP64 CG reads
// it at once.
void learned_warm_start(Context& c) {
  c.stats.learned_used = 0;
  if (!c.model || !c.ppe_from_particles) return;
  const size_t n = c.n;
  learned_features(c);
  const double flops = 2.0 * n * (static_cast<double>(kFeatPad) * kHid + kHid * kHid + kHid);
  tracked(c, MPS_K_MLP, flops, [&] { c.model->forward(c.stream, n, c.feats, c.ypred, false); });
  hipLaunchKernelGGL(k_apply_pred, dim3(blocks(n)), dim3(256), 0, c.stream, n,
                     static_cast<const uint8_t*>(c.type), static_cast<const double*>(c.p),
                     static_cast<const float*>(c.ypred), static_cast<const double*>(c.red_out), c.sol);
  residual64_launch(c);
  hipLaunchKernelGGL(k_revert_unless_better, dim3(blocks(n)), dim3(256), 0, c.stream, n,
                     static_cast<const uint8_t*>(c.type), static_cast<const double*>(c.p),
                     static_cast<const double*>(c.red_out), c.sol);
  check_launch();
  MPS_HIP(hipMemcpyAsync(c.red_host, c.red_out, kRedOut * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  c.learned_pending = true;
  if (c.solver.kind != MPS_SOLVER_MIXED_IR) {
    double rr = 0, bb = 0;
    learned_resolve(c, rr, bb);
  }
}

void batch_learned_start(std::vector<Context*>& scenes, Model& m, hipStream_t s, const __bf16* feats, float* ypred,
                         size_t rows) {
  m.forward(s, rows, feats, ypred, false);
  for (Context* cp : scenes) {
    Context& c = *cp;
    const size_t n = c.n;
    hipStream_t own = c.stream;
    c.stream = s;  // the scene's kernels below run on the batch's stream
    try {
      hipLaunchKernelGGL(k_apply_pred, dim3(blocks(n)), dim3(256), 0, c.stream, n,
                         static_cast<const uint8_t*>(c.type), static_cast<const double*>(c.p),
                         static_cast<const float*>(c.ypred), static_cast<const double*>(c.red_out), c.sol);
      residual64_launch(c);
      hipLaunchKernelGGL(k_revert_unless_better, dim3(blocks(n)), dim3(256), 0, c.stream, n,
                         static_cast<const uint8_t*>(c.type), static_cast<const double*>(c.p),
                         static_cast<const double*>(c.red_out), c.sol);
      check_launch();
      MPS_HIP(hipMemcpyAsync(c.red_host, c.red_out, kRedOut * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    } catch (...) {
      c.stream = own;
      throw;
    }
    c.stream = own;
  }
}

void batch_learned_outcome(Context& c) {
  const double rr_new = c.red_host[0], bbv = c.red_host[1], rr_prev = c.red_host[6];
  c.stats.prev_rel_residual = bbv > 0 ? std::sqrt(rr_prev / bbv) : 0;
  c.stats.learned_used = keep_prediction(rr_new, rr_prev) ? 1 : 0;
}

bool learned_resolve(Context& c, double& rr, double& bb) {
  c.learned_pending = false;
  MPS_HIP(hipStreamSynchronize(c.stream));
  const double rr_new = c.red_host[0], bbv = c.red_host[1], rr_prev = c.red_host[6];
  c.stats.prev_rel_residual = bbv > 0 ? std::sqrt(rr_prev / bbv) : 0;
  if (!keep_prediction(rr_new, rr_prev)) return false;  // x went back to p_prev: r is not its residual
  c.stats.learned_used = 1;
  rr = rr_new;
  bb = bbv;
  return true;
}

void collect_before(Context& c) {
  c.collect_ok = false;
  if (!c.dataset || !c.ppe_from_particles) return;
  learned_features(c);
  MPS_HIP(hipMemcpyAsync(c.red_host, c.red_out, kRedOut * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
  c.collect_sd = c.red_host[9];
  c.collect_ok = c.collect_sd > 0;
}

void collect_after(Context& c) {
  if (!c.dataset || !c.collect_ok) return;
  const double sd = c.collect_sd;
  const uint32_t thr = static_cast<uint32_t>(std::min(1.0, std::max(0.0, c.collect_frac)) * 4294967295.0);
  const CollectArgs a{c.n, c.type, c.feats, c.sol, c.p, c.x, c.z, c.nbr, c.ncount, c.kstride,
                      c.env.r_lap * c.env.r_lap, 1.0 / sd,
                      c.collect_seed ^ (static_cast<uint64_t>(c.collect_step) * 0x9E3779B97F4A7C15ull), thr,
                      c.dataset->cap, c.dataset->count, c.dataset->x, c.dataset->t};
  hipLaunchKernelGGL(k_collect, dim3(blocks(c.n)), dim3(256), 0, c.stream, a);
  check_launch();
  c.collect_step++;
}

void model_forward_host(Model& m, size_t rows, const float* x, float* y) {
  m.pack_host_rows(rows, x);
  m.forward(m.stream, rows, m.xb, m.y, false);
  MPS_HIP(hipMemcpyAsync(y, m.y, rows * 4, hipMemcpyDeviceToHost, m.stream));
  MPS_HIP(hipStreamSynchronize(m.stream));
}

void model_grads_host(Model& m, size_t rows, const float* x, const float* t, float* loss_out) {
  m.pack_host_rows(rows, x);
  MPS_HIP(hipMemcpyAsync(m.tb, t, rows * 4, hipMemcpyHostToDevice, m.stream));
  m.forward(m.stream, rows, m.xb, m.y, true);
  m.backward(m.stream, rows);
  MPS_HIP(hipMemcpyAsync(loss_out, m.loss, 4, hipMemcpyDeviceToHost, m.stream));
  MPS_HIP(hipStreamSynchronize(m.stream));
}

}  // namespace mps
