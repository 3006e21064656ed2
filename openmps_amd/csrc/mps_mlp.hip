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
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float kG0 = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kG1 = 0.044715f;

__device__ __forceinline__ float gelu_f(float z) {
  return 0.5f * z * (1.0f + tanhf(kG0 * (z + kG1 * z * z * z)));
}
__device__ __forceinline__ float gelu_grad_f(float z) {
  const float th = tanhf(kG0 * (z + kG1 * z * z * z));
  return 0.5f * (1.0f + th) + 0.5f * z * (1.0f - th * th) * kG0 * (1.0f + 3.0f * kG1 * z * z);
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
  const double *x, *z, *p, *r0, *nd;
  const uint8_t* type;
  const uint32_t *ids, *count;
  uint32_t kmax;
  double rl2, inv_l0, n0, inv_sp, inv_sr;
  __bf16* out;
};

// Features of every particle (zero rows for dummy / disabled particles).
__global__ void __launch_bounds__(256) k_features(FeatArgs a) {
  __shared__ float skey[kNbr][256];
  __shared__ uint32_t sidx[kNbr][256];
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const size_t i = blk * 256ull + threadIdx.x;
  if (i >= a.n) return;
  const int t = threadIdx.x;
  __bf16* row = a.out + i * kFeatPad;
  const uint8_t ti = a.type[i];
  float f[4] = {0, 0, 0, 0};
  int m = 0;
  if (ti != kDummy && ti != kDisabled) {
    const double xi = a.x[i], zi = a.z[i];
    int cnt = 0;
    const uint32_t c = a.count[i];
    for (uint32_t q = 0; q < c; q++) {
      const uint32_t j = a.ids[nbr_slot(static_cast<uint32_t>(i), q, a.kmax)];
      if (a.type[j] == kDummy) continue;
      const double dx = a.x[j] - xi, dz = a.z[j] - zi;
      const double r2 = dx * dx + dz * dz;
      if (!(r2 < a.rl2)) continue;
      cnt++;
      const float kf = static_cast<float>(r2);
      int pos;
      if (m < kNbr) {
        pos = m++;
      } else if (kf < skey[kNbr - 1][t]) {
        pos = kNbr - 1;
      } else {
        continue;
      }
      while (pos > 0 && skey[pos - 1][t] > kf) {
        skey[pos][t] = skey[pos - 1][t];
        sidx[pos][t] = sidx[pos - 1][t];
        pos--;
      }
      skey[pos][t] = kf;
      sidx[pos][t] = j;
    }
    f[0] = cnt / 32.0f;
    f[1] = static_cast<float>((a.nd[i] - a.n0) / a.n0);
    f[2] = static_cast<float>(a.p[i] * a.inv_sp);
    f[3] = static_cast<float>(a.r0[i] * a.inv_sr);
  }
  // the row as 16-byte stores: elements 0-3 the particle's own features,
  // 4 + 4k .. 4 + 4k + 3 neighbour k's (k < m, else 0), zeros from kFeat on
  const double xi = a.x[i], zi = a.z[i];
  for (int q = 0; q < kFeatPad / 8; q++) {
    bf16x8 v;
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const int e0 = 8 * q + 4 * half;
      float u[4] = {0.f, 0.f, 0.f, 0.f};
      if (e0 == 0) {
        u[0] = f[0];
        u[1] = f[1];
        u[2] = f[2];
        u[3] = f[3];
      } else if (e0 < kFeat && (e0 - 4) / 4 < m) {
        const uint32_t j = sidx[(e0 - 4) / 4][t];
        u[0] = static_cast<float>((a.x[j] - xi) * a.inv_l0);
        u[1] = static_cast<float>((a.z[j] - zi) * a.inv_l0);
        u[2] = static_cast<float>(a.p[j] * a.inv_sp);
        u[3] = static_cast<float>(a.r0[j] * a.inv_sr);
      }
#pragma unroll
      for (int e = 0; e < 4; e++) v[4 * half + e] = __bf16(u[e]);
    }
    *reinterpret_cast<bf16x8*>(row + 8 * q) = v;
  }
}

// sum over alive rows of v^2 and the alive count (block partials).
__global__ void __launch_bounds__(256) k_alive_stats(size_t n, const uint8_t* __restrict__ type,
                                                     const double* __restrict__ v,
                                                     double* __restrict__ part_vv,
                                                     double* __restrict__ part_n) {
  __shared__ double lds[4];
  double s = 0, c = 0;
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) {
    const uint8_t t = type[i];
    if (t != kDummy && t != kDisabled) {
      s += v[i] * v[i];
      c += 1.0;
    }
  }
  const double s1 = block_sum<256>(s, lds);
  const double s2 = block_sum<256>(c, lds);
  if (threadIdx.x == 0) {
    part_vv[blockIdx.x] = s1;
    part_n[blockIdx.x] = s2;
  }
}

// sol = p_prev + s_d y on alive rows.
__global__ void __launch_bounds__(256) k_apply_pred(size_t n, const uint8_t* __restrict__ type,
                                                    const double* __restrict__ p,
                                                    const float* __restrict__ y, double sd,
                                                    double* __restrict__ sol) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
  const uint8_t t = type[i];
  if (t != kDummy && t != kDisabled) sol[i] = p[i] + sd * static_cast<double>(y[i]);
}

__global__ void __launch_bounds__(256) k_revert(size_t n, const uint8_t* __restrict__ type,
                                                const double* __restrict__ p, double* __restrict__ sol) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i >= n) return;
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
  uint32_t kmax;
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
    const uint32_t j = a.ids[nbr_slot(static_cast<uint32_t>(i), q, a.kmax)];
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
// read and y written: one wavefront per 16 rows, both layers transposed
// (hidden units as the MFMA M dimension, rows as N), so layer 1's output
// tiles already sit in the lanes that feed them to layer 2 as its B operand.
// v_mfma_f32_16x16x32_bf16 lane layout: A row l%16 / B column l%16 with K
// elements 8(l/16)..8(l/16)+7; C[4(l/16)+v][l%16], v = 0..3.  Layer 1 lane
// l ends with hidden units 16t + 4(l/16) + v of row l%16, so layer 2's K
// step ks (hidden 32ks..32ks+31) takes tiles 2ks and 2ks+1: its K slot
// 8g + e stands for hidden unit w2_hidden(ks, g, e) below, and W2 is stored
// in that order (k_perm_w2) so that each lane's A operand is one 16-byte load.
constexpr int kRowGroups = 1;  // groups of 16 rows per wavefront (2 needs ~500 registers)
constexpr int kRowsPerWave = 16 * kRowGroups;

__host__ __device__ __forceinline__ int w2_hidden(int ks, int g, int e) {
  return 32 * ks + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4));
}

// w2p[o][32 ks + 8 g + e] = bf16(w2[o][w2_hidden(ks, g, e)]).
__global__ void __launch_bounds__(256) k_perm_w2(const float* __restrict__ w2, __bf16* __restrict__ w2p) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kHid * kHid) return;
  const int o = i / kHid, k = i % kHid;
  const int ks = k / 32, g = (k % 32) / 8, e = k % 8;
  w2p[i] = static_cast<__bf16>(w2[o * kHid + w2_hidden(ks, g, e)]);
}

__global__ void __launch_bounds__(256) k_mlp_infer(size_t M, const __bf16* __restrict__ x,
                                                   const __bf16* __restrict__ w1, const float* __restrict__ b1,
                                                   const __bf16* __restrict__ w2p, const float* __restrict__ b2,
                                                   const float* __restrict__ w3, const float* __restrict__ b3,
                                                   float* __restrict__ y) {
  const size_t wave = blockIdx.x * static_cast<size_t>(blockDim.x / 64) + (threadIdx.x >> 6);
  const size_t r0 = wave * kRowsPerWave;
  if (r0 >= M) return;  // wave-uniform
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  constexpr int K1 = kFeatPad / 32, K2 = kHid / 32;
  bf16x8 xf[kRowGroups][K1];
#pragma unroll
  for (int sg = 0; sg < kRowGroups; sg++) {
    size_t r = r0 + 16 * sg + c;
    r = (r < M) ? r : M - 1;
#pragma unroll
    for (int ks = 0; ks < K1; ks++)
      xf[sg][ks] = *reinterpret_cast<const bf16x8*>(x + r * kFeatPad + 32 * ks + 8 * g);
  }
  bf16x8 h1[kRowGroups][K2];
#pragma unroll
  for (int t2 = 0; t2 < K2; t2++) {
    // keeps the next pair's weight loads from being hoisted (register pressure)
    asm volatile("" ::: "memory");
    f32x4 acc[2][kRowGroups];
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const int h0 = 16 * (2 * t2 + half);
#pragma unroll
      for (int sg = 0; sg < kRowGroups; sg++) acc[half][sg] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < K1; ks++) {
        const bf16x8 af = *reinterpret_cast<const bf16x8*>(w1 + (h0 + c) * kFeatPad + 32 * ks + 8 * g);
#pragma unroll
        for (int sg = 0; sg < kRowGroups; sg++)
          acc[half][sg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, xf[sg][ks], acc[half][sg], 0, 0, 0);
      }
    }
#pragma unroll
    for (int half = 0; half < 2; half++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const float bias = b1[16 * (2 * t2 + half) + 4 * g + v];
#pragma unroll
        for (int sg = 0; sg < kRowGroups; sg++) h1[sg][t2][4 * half + v] = static_cast<__bf16>(gelu_f(acc[half][sg][v] + bias));
      }
  }
  float part[kRowGroups];
#pragma unroll
  for (int sg = 0; sg < kRowGroups; sg++) part[sg] = 0.f;
  for (int ot = 0; ot < kHid / 16; ot++) {
    asm volatile("" ::: "memory");
    const int o0 = 16 * ot;
    f32x4 acc2[kRowGroups];
#pragma unroll
    for (int sg = 0; sg < kRowGroups; sg++) acc2[sg] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < K2; ks++) {
      const bf16x8 af = *reinterpret_cast<const bf16x8*>(w2p + (o0 + c) * kHid + 32 * ks + 8 * g);
#pragma unroll
      for (int sg = 0; sg < kRowGroups; sg++)
        acc2[sg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, h1[sg][ks], acc2[sg], 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int o = o0 + 4 * g + v;
      const float bias = b2[o], wo = w3[o];
#pragma unroll
      for (int sg = 0; sg < kRowGroups; sg++) part[sg] = fmaf(wo, gelu_f(acc2[sg][v] + bias), part[sg]);
    }
  }
#pragma unroll
  for (int sg = 0; sg < kRowGroups; sg++) {
    part[sg] += __shfl_xor(part[sg], 16, 64);
    part[sg] += __shfl_xor(part[sg], 32, 64);
    const size_t r = r0 + 16 * sg + c;
    if (g == 0 && r < M) y[r] = part[sg] + b3[0];
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
    const size_t waves = (m + kRowsPerWave - 1) / kRowsPerWave;
    hipLaunchKernelGGL(k_mlp_infer, dim3(static_cast<unsigned>((waves + 3) / 4)), dim3(256), 0, s, m, x,
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

// Residual of the current x (= p_prev), the scales, and the features of every
// particle into c.feats.  Returns false if there are no alive rows.
bool learned_features(Context& c, double& s_r, double& rr) {
  const size_t n = c.n;
  double bb = 0;
  residual64(c, rr, bb, false);
  hipLaunchKernelGGL(k_alive_stats, dim3(256), dim3(256), 0, c.stream, n, static_cast<const uint8_t*>(c.type),
                     static_cast<const double*>(c.p), c.partB, c.partC);
  reduce2(c, c.partB, c.partC, 256);
  const double pp = c.red_host[0], nalive = c.red_host[1];
  if (nalive <= 0) return false;
  hipLaunchKernelGGL(k_alive_stats, dim3(256), dim3(256), 0, c.stream, n, static_cast<const uint8_t*>(c.type),
                     static_cast<const double*>(c.diag), c.partB, c.partC);
  reduce2(c, c.partB, c.partC, 256);
  c.learned_drms = std::sqrt(c.red_host[0] / nalive);
  s_r = std::sqrt(rr / nalive);
  const double s_p = std::sqrt(pp / nalive);
  if (!c.feats || c.feats_cap < n) {
    dfree(c.feats);
    dfree(c.ypred);
    c.feats = dalloc<__bf16>(n * kFeatPad);
    c.ypred = dalloc<float>(n);
    c.feats_cap = n;
  }
  FeatArgs a{n, c.x, c.z, c.p, c.r, c.nd, c.type, c.nbr, c.ncount, c.kstride, c.env.r_lap * c.env.r_lap,
             1.0 / c.env.l0, c.env.n0, s_p > 0 ? 1.0 / s_p : 0.0, s_r > 0 ? 1.0 / s_r : 0.0, c.feats};
  tracked(c, MPS_K_FEATURES, n * (8.0 * 5 + 1 + 4 + kFeatPad * 2) + 4.0 * c.nbr_total, [&] {
    hipLaunchKernelGGL(k_features, dim3(blocks(n)), dim3(256), 0, c.stream, a);
  });
  check_launch();
  return true;
}

// Scale of the model output: the residual's RMS over the diagonal's RMS, the
// size of a Jacobi correction.
double learned_scale(const Context& c, double s_r) {
  return c.learned_drms > 0 ? s_r / c.learned_drms : 0.0;
}

// MPS_WARM_LEARNED: x0 = p_prev + s_d MLP(features); keeps p_prev when the
// prediction does not lower the FP64 residual.
void learned_warm_start(Context& c) {
  c.stats.learned_used = 0;
  if (!c.model) return;
  double s_r = 0, rr_prev = 0;
  if (!learned_features(c, s_r, rr_prev) || s_r == 0) return;
  const size_t n = c.n;
  const double flops = 2.0 * n * (static_cast<double>(kFeatPad) * kHid + kHid * kHid + kHid);
  tracked(c, MPS_K_MLP, flops, [&] { c.model->forward(c.stream, n, c.feats, c.ypred, false); });
  hipLaunchKernelGGL(k_apply_pred, dim3(blocks(n)), dim3(256), 0, c.stream, n,
                     static_cast<const uint8_t*>(c.type), static_cast<const double*>(c.p),
                     static_cast<const float*>(c.ypred), learned_scale(c, s_r), c.sol);
  check_launch();
  double rr_new = 0, bb = 0;
  residual64(c, rr_new, bb, false);
  c.stats.prev_rel_residual = bb > 0 ? std::sqrt(rr_prev / bb) : 0;
  if (rr_new < rr_prev) {
    c.stats.learned_used = 1;
  } else {
    hipLaunchKernelGGL(k_revert, dim3(blocks(n)), dim3(256), 0, c.stream, n,
                       static_cast<const uint8_t*>(c.type), static_cast<const double*>(c.p), c.sol);
    check_launch();
  }
}

void collect_before(Context& c) {
  if (!c.dataset) return;
  double rr = 0;
  c.collect_ok = learned_features(c, c.collect_sr, rr) && c.collect_sr > 0;
}

void collect_after(Context& c) {
  if (!c.dataset || !c.collect_ok) return;
  const double sd = learned_scale(c, c.collect_sr);
  if (!(sd > 0)) return;
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
