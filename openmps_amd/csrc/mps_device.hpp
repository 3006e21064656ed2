// Device-side helpers shared by the particle, PPE and CG kernels.
//
// All FP64 arithmetic here must round exactly as the reference's host code:
// the library is compiled with -ffp-contract=off, and every expression keeps
// the association order of the reference source it cites.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace mps {

// Particle::W, Particle.hpp:71-75
__device__ __forceinline__ double weight(double r, double re) {
  return ((0 < r) && (r < re)) ? (re / r - 1) : 0;
}

// Slot k of particle i in the neighbour layout: per slice of 64 particles,
// groups of four consecutive slots of one particle (16 bytes), the 64
// particles' groups side by side; kstride (slots per particle) is a
// multiple of 4.  The search stores and the walks load a whole group at once.
__host__ __device__ __forceinline__ size_t nbr_slot(uint32_t i, uint32_t k, uint32_t kstride) {
  return ((static_cast<size_t>(i >> 6) * (kstride >> 2) + (k >> 2)) * 64 + (i & 63u)) * 4 + (k & 3u);
}

// SPP ghost particle, Computer.hpp:657-693.  (sx, sz) is sum W(|dx|) dx over
// the r_e-filtered neighbours; the ghost exists when nWithoutSpp_i < n0 and
// the centroid offset exceeds DBL_EPSILON.
__device__ __forceinline__ bool ghost_pos(double xi, double zi, double sx, double sz, double nws_i,
                                          double n0, double re, double& gx, double& gz) {
  sx /= n0;
  sz /= n0;
  if (!(nws_i < n0)) return false;
  const double rg = sqrt(sx * sx + sz * sz);
  if (!(rg > DBL_EPSILON)) return false;
  const double wspp = n0 - nws_i;
  const double rspp = re / (wspp + 1);
  gx = xi - (rspp / rg) * sx;
  gz = zi - (rspp / rg) * sz;
  return true;
}

// Walks particle i's m-entry neighbour list in list order, U neighbours per
// round trip to memory: the U ids, then load(j) for all U of them (the
// gathered fields of neighbour j, returned as a T), then use(j, T) in list
// order.  A walk that issues each neighbour's loads only after the previous
// one's arithmetic pays a full memory latency per neighbour.
template <int U, class T, class Load, class Use>
__device__ __forceinline__ void walk_list(uint32_t i, uint32_t m, const uint32_t* __restrict__ ids, uint32_t kmax,
                                          Load load, Use use) {
  static_assert(U % 4 == 0, "whole slot groups");
  uint32_t q = 0;
  for (; q + U <= m; q += U) {
    uint32_t jj[U];
#pragma unroll
    for (int v = 0; v < U; v += 4) {
      const uint4 g = *reinterpret_cast<const uint4*>(ids + nbr_slot(i, q + v, kmax));
      jj[v] = g.x;
      jj[v + 1] = g.y;
      jj[v + 2] = g.z;
      jj[v + 3] = g.w;
    }
    T f[U];
#pragma unroll
    for (int v = 0; v < U; v++) f[v] = load(jj[v]);
#pragma unroll
    for (int v = 0; v < U; v++) use(jj[v], f[v]);
  }
  for (; q < m; q++) {
    const uint32_t j = ids[nbr_slot(i, q, kmax)];
    use(j, load(j));
  }
}

// XCD-aware block order: hardware hands consecutive workgroups to the 8 XCDs
// round-robin; remapping makes each XCD walk one contiguous range of
// particles, so gathered neighbour data stays in that XCD's L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned per = nb / 8u;
  if (per == 0u || b >= per * 8u) return b;
  return (b % 8u) * per + b / 8u;
}

// Wavefront (64-lane) sum / max.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned wave_min_u32(unsigned v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned t = __shfl_xor(v, o, 64);
    v = (t < v) ? t : v;
  }
  return v;
}
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned t = __shfl_xor(v, o, 64);
    v = (t > v) ? t : v;
  }
  return v;
}

// Deterministic block sum of one double per thread (blockDim.x == NT):
// fixed wavefront butterflies, then the NT/64 wavefront sums in order.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) lds[wv] = v;
  __syncthreads();
  double s = lds[0];
#pragma unroll
  for (int k = 1; k < NT / 64; k++) s += lds[k];
  __syncthreads();
  return s;
}

// Sum of `count` partials in a fixed order that every block reproduces
// identically (thread t takes t, t+NT, ...; then block_sum).
template <int NT>
__device__ __forceinline__ double reduce_partials(const double* part, int count, double* lds) {
  double s = 0;
  for (int k = threadIdx.x; k < count; k += NT) s += part[k];
  return block_sum<NT>(s, lds);
}

// Block sums of two values per thread at once (lds: 2 * NT / 64 entries),
// in the order of block_sum.
template <int NT>
__device__ __forceinline__ void block_sum2(double a, double b, double* lds, double& sa, double& sb) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    lds[wv] = a;
    lds[NT / 64 + wv] = b;
  }
  __syncthreads();
  sa = lds[0];
  sb = lds[NT / 64];
#pragma unroll
  for (int k = 1; k < NT / 64; k++) {
    sa += lds[k];
    sb += lds[NT / 64 + k];
  }
  __syncthreads();
}

// What the learned warm start's feature pass gathers of a neighbour, one
// 32-byte record (one cache line access where four arrays cost four): the
// position (x = NaN for dummy particles, which no r^2 < r_lap^2 test
// accepts), the pressure p_prev and the residual r0 = b - A p_prev.
struct __attribute__((aligned(32))) FeatRec {
  double x, z, p, r0;
};

// The feature scales from the reduced sums out[0] = |r0|^2, out[2] = sum p^2,
// out[3] = sum diag^2, out[4] = alive count (one thread): out[6..9] =
// |r0|^2, 1/s_p, 1/s_r and s_d = s_r / RMS(diag) over the alive rows, with
// s_v = RMS(v); zero where a scale is zero or no row is alive.
__device__ __forceinline__ void feat_scales(double* out) {
  const double rr = out[0], pp = out[2], dd = out[3], na = out[4];
  double inv_sp = 0, inv_sr = 0, sd = 0;
  if (na > 0) {
    const double s_r = sqrt(rr / na), s_p = sqrt(pp / na), drms = sqrt(dd / na);
    inv_sp = s_p > 0 ? 1.0 / s_p : 0.0;
    inv_sr = s_r > 0 ? 1.0 / s_r : 0.0;
    sd = (drms > 0 && s_r > 0) ? s_r / drms : 0.0;
  }
  out[6] = rr;
  out[7] = inv_sp;
  out[8] = inv_sr;
  out[9] = sd;
}

}  // namespace mps
