// Per-particle stages of Computer::ForwardTime on the GPU, one thread per
// particle, FP64 throughout.
//
// Every neighbour sum runs sequentially over the particle's own list, in list
// order, exactly as Computer::AccumulateNeighbor (Computer.hpp:618-695) does,
// so the results are bitwise equal to the reference's.  Wavefront reductions
// are used only where the reference reduces globally (DetermineDt).
//
// The neighbour loop fuses what the reference does in separate passes where
// the inputs allow it: density + HS + ECS (first density pass), density +
// SaveX (second pass), and the SPP centroid of every stage.
#include <algorithm>
#include <utility>

#include "mps_device.hpp"
#include "mps_internal.hpp"

namespace mps {
namespace {

struct NV {
  const uint32_t* __restrict__ ids;
  const uint32_t* __restrict__ count;
  uint32_t kmax;
};

struct Consts {
  double re, re2, rl, rl2, n0, n0l, nu, rho, dt, gx, gz, d2;
};

constexpr int kWalk = 4;  // neighbours whose loads are issued together (walk_list)

Consts consts(const Context& c) {
  Consts k;
  k.re = c.env.r_e;
  k.re2 = k.re * k.re;
  k.rl = c.env.r_lap;
  k.rl2 = k.rl * k.rl;
  k.n0 = c.env.n0;
  k.n0l = c.env.n0_lap;
  k.nu = c.env.nu;
  k.rho = c.env.rho;
  k.dt = c.dt;
  k.gx = c.env.gx;
  k.gz = c.env.gz;
  const double d = c.env.l0 - c.env.max_dx;  // DynamicStabilize, Computer.hpp:1570
  k.d2 = d * d;
  return k;
}

__device__ __forceinline__ size_t thread_particle() {
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  return blk * static_cast<size_t>(blockDim.x) + threadIdx.x;
}

// ComputeNeighborDensities (Computer.hpp:780-833), fused with
// NeighborDensityVariationSpeed + ComputeErrorCorrection (:838-910) and with
// SaveX (:1025-1039) as the template flags ask.
template <bool DENS, bool ECS, bool SAVEX>
__global__ void __launch_bounds__(kBlock) k_density(size_t n, const double* __restrict__ x,
                                                    const double* __restrict__ z,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ w,
                                                    const uint8_t* __restrict__ type, NV nb, Consts k,
                                                    double* __restrict__ nd, double* __restrict__ nws,
                                                    double* __restrict__ ecs, double* __restrict__ ox,
                                                    double* __restrict__ oz) {
  const size_t i = thread_particle();
  if (i >= n) return;
  const uint8_t ti = type[i];
  const bool alive = (ti != kDummy) && (ti != kDisabled);
  const double xi = x[i], zi = z[i];
  if (SAVEX && alive) {
    ox[i] = xi;
    oz[i] = zi;
  }
  if (!alive) {
    if (DENS) nws[i] = k.n0;
    return;
  }
  double ui = 0, wi = 0;
  if (ECS) {
    ui = u[i];
    wi = w[i];
  }
  double s = 0.0, h = 0.0;
  struct F {
    double x, z, u, w;
  };
  walk_list<kWalk, F>(static_cast<uint32_t>(i), nb.count[i], nb.ids, nb.kmax,
                      [&](uint32_t j) {
                        F f{x[j], z[j], 0.0, 0.0};
                        if (ECS) {
                          f.u = u[j];
                          f.w = w[j];
                        }
                        return f;
                      },
                      [&](uint32_t, const F& f) {
                        const double dx = f.x - xi, dz = f.z - zi;
                        const double r2 = dx * dx + dz * dz;
                        if (r2 < k.re2) {
                          const double r = sqrt(r2);
                          if (DENS) s += weight(r, k.re);
                          if (ECS) {
                            const double du = f.u - ui, dw = f.w - wi;
                            h += (dx * du + dz * dw) / (r * r * r);
                          }
                        }
                      });
  double thisNws, thisN;
  if (DENS) {
    thisNws = s;
    thisN = (s < k.n0) ? k.n0 : s;  // std::max(thisN, n0)
    nws[i] = thisNws;
    nd[i] = thisN;
  } else {
    thisNws = nws[i];
    thisN = nd[i];
  }
  if (ECS) {
    const double speed = (thisNws < k.n0) ? 0.0 : -k.re * h;
    const double e = (thisN - k.n0) / k.n0;
    ecs[i] = fabs(e) * speed + fabs(speed) * e;
  }
}

// NeighborDensityVariationSpeed(i) for every particle (HS, Computer.hpp:838-872).
__global__ void __launch_bounds__(kBlock) k_hs(size_t n, const double* __restrict__ x,
                                               const double* __restrict__ z,
                                               const double* __restrict__ u,
                                               const double* __restrict__ w, NV nb, Consts k,
                                               const double* __restrict__ nws, double* __restrict__ out) {
  const size_t i = thread_particle();
  if (i >= n) return;
  const double xi = x[i], zi = z[i], ui = u[i], wi = w[i];
  double h = 0.0;
  const uint32_t cnt = nb.count[i];
  for (uint32_t q = 0; q < cnt; q++) {
    const uint32_t j = nb.ids[nbr_slot(static_cast<uint32_t>(i), q, nb.kmax)];
    if (j == i) continue;
    const double dx = x[j] - xi, dz = z[j] - zi;
    const double r2 = dx * dx + dz * dz;
    if (r2 < k.re2) {
      const double r = sqrt(r2);
      const double du = u[j] - ui, dw = w[j] - wi;
      h += (dx * du + dz * dw) / (r * r * r);
    }
  }
  out[i] = (nws[i] < k.n0) ? 0.0 : -k.re * h;
}

// ComputeExplicitForces first pass (Computer.hpp:928-987): HL viscosity over
// non-dummy neighbours, the SPP ghost term, plus gravity.
__global__ void __launch_bounds__(kBlock) k_explicit(size_t n, const double* __restrict__ x,
                                                     const double* __restrict__ z,
                                                     const double* __restrict__ u,
                                                     const double* __restrict__ w,
                                                     const uint8_t* __restrict__ type, NV nb, Consts k,
                                                     const double* __restrict__ nws,
                                                     double* __restrict__ ax, double* __restrict__ az) {
  const size_t i = thread_particle();
  if (i >= n) return;
  if (type[i] != kFluid) return;
  const double xi = x[i], zi = z[i], ui = u[i], wi = w[i];
  double vx = 0, vz = 0, sx = 0, sz = 0;
  struct F {
    double x, z, u, w;
    uint8_t t;
  };
  walk_list<kWalk, F>(static_cast<uint32_t>(i), nb.count[i], nb.ids, nb.kmax,
                      [&](uint32_t j) { return F{x[j], z[j], u[j], w[j], type[j]}; },
                      [&](uint32_t, const F& f) {
                        const double dx = f.x - xi, dz = f.z - zi;
                        const double r2 = dx * dx + dz * dz;
                        if (r2 < k.rl2 && f.t != kDummy) {
                          const double r = sqrt(r2);
                          const double c = k.nu * 3.0 * k.rl / k.n0l / (r * r * r);
                          vx += c * (f.u - ui);
                          vz += c * (f.w - wi);
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
    const double c = k.nu * 3.0 * k.rl / k.n0l / (r * r * r);
    vx += c * (ui - ui);
    vz += c * (wi - wi);
  }
  ax[i] = vx + k.gx;
  az[i] = vz + k.gz;
}

// The first ComputeNeighborDensities pass with HS + ECS (Computer.hpp:780-910)
// and the acceleration pass of ComputeExplicitForces (:914-990) in one walk of
// the neighbour list: the accelerations read the unmoved positions and
// velocities and only the particle's own nWithoutSpp (for the SPP ghost),
// which this walk has just summed.  Every value is computed as k_density and
// k_explicit compute it (same expressions, same list order).
__global__ void __launch_bounds__(kBlock) k_pack_state(size_t n, const double* __restrict__ x,
                                                       const double* __restrict__ z, const double* __restrict__ u,
                                                       const double* __restrict__ w, double2* __restrict__ xz,
                                                       double2* __restrict__ uw) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  xz[i] = make_double2(x[i], z[i]);
  uw[i] = make_double2(u[i], w[i]);
}

__global__ void __launch_bounds__(kBlock) k_density_explicit(size_t n, const double* __restrict__ x,
                                                             const double* __restrict__ z,
                                                             const double* __restrict__ u,
                                                             const double* __restrict__ w,
                                                             const uint8_t* __restrict__ type, NV nb,
                                                             Consts k, double* __restrict__ nd,
                                                             double* __restrict__ nws, double* __restrict__ ecs,
                                                             double* __restrict__ ax, double* __restrict__ az,
                                                             const double2* __restrict__ xz,
                                                             const double2* __restrict__ uw) {
  const size_t i = thread_particle();
  if (i >= n) return;
  const uint8_t ti = type[i];
  if (ti == kDummy || ti == kDisabled) {
    nws[i] = k.n0;
    return;
  }
  const bool fluid = (ti == kFluid);
  const double xi = x[i], zi = z[i], ui = u[i], wi = w[i];
  double s = 0.0, h = 0.0;                     // density, HS
  double vx = 0, vz = 0, sx = 0, sz = 0;       // viscosity, ghost direction (fluid)
  struct F {
    double x, z, u, w;
    uint8_t t;
  };
  walk_list<kWalk, F>(static_cast<uint32_t>(i), nb.count[i], nb.ids, nb.kmax,
                      [&](uint32_t j) {
                        const double2 pj = xz[j], vj = uw[j];
                        return F{pj.x, pj.y, vj.x, vj.y, type[j]};
                      },
                      [&](uint32_t, const F& f) {
                        const double dx = f.x - xi, dz = f.z - zi;
                        const double r2 = dx * dx + dz * dz;
                        if (r2 < k.re2) {
                          const double r = sqrt(r2);
                          s += weight(r, k.re);
                          const double du = f.u - ui, dw = f.w - wi;
                          h += (dx * du + dz * dw) / (r * r * r);
                        }
                        if (fluid) {
                          if (r2 < k.rl2 && f.t != kDummy) {
                            const double r = sqrt(r2);
                            const double c = k.nu * 3.0 * k.rl / k.n0l / (r * r * r);
                            vx += c * (f.u - ui);
                            vz += c * (f.w - wi);
                          }
                          if (r2 < k.re2) {
                            const double wg = weight(sqrt(r2), k.re);
                            sx += wg * dx;
                            sz += wg * dz;
                          }
                        }
                      });
  const double thisN = (s < k.n0) ? k.n0 : s;
  nws[i] = s;
  nd[i] = thisN;
  const double speed = (s < k.n0) ? 0.0 : -k.re * h;
  const double e = (thisN - k.n0) / k.n0;
  ecs[i] = fabs(e) * speed + fabs(speed) * e;
  if (!fluid) return;
  double gxp, gzp;
  if (ghost_pos(xi, zi, sx, sz, s, k.n0, k.re, gxp, gzp)) {
    const double rx = xi - gxp, rz = zi - gzp;
    const double r = sqrt(rx * rx + rz * rz);
    const double c = k.nu * 3.0 * k.rl / k.n0l / (r * r * r);
    vx += c * (ui - ui);
    vz += c * (wi - wi);
  }
  ax[i] = vx + k.gx;
  az[i] = vz + k.gz;
}

// ComputeExplicitForces second pass (Computer.hpp:996-1020).
// Non-fluid targets: wall_x / wall_z (positions at AddParticles, or the
// callbacks' targets), displaced by the group's prescribed motion at time t
// when device kinematics are on (group != nullptr).
__global__ void __launch_bounds__(kBlock) k_explicit_update(size_t n, double* __restrict__ x,
                                                            double* __restrict__ z, double* __restrict__ u,
                                                            double* __restrict__ w,
                                                            const uint8_t* __restrict__ type,
                                                            const double* __restrict__ ax,
                                                            const double* __restrict__ az,
                                                            const double* __restrict__ wall_x,
                                                            const double* __restrict__ wall_z, double dt,
                                                            const uint8_t* __restrict__ group,
                                                            const mps_wall_motion* __restrict__ motion,
                                                            unsigned n_motions, double t,
                                                            double2* __restrict__ xz, double2* __restrict__ uw) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (type[i] == kFluid) {
    const double un = u[i] + ax[i] * dt, wn = w[i] + az[i] * dt;
    const double xn = x[i] + un * dt, zn = z[i] + wn * dt;
    u[i] = un;
    w[i] = wn;
    x[i] = xn;
    z[i] = zn;
    xz[i] = make_double2(xn, zn);
    uw[i] = make_double2(un, wn);
  } else {
    double nx = wall_x[i], nz = wall_z[i];
    const unsigned g = group ? group[i] : 0u;
    if (g >= 1 && g <= n_motions) {
      const mps_wall_motion m = motion[g - 1];
      const double sn = sin(m.omega * t + m.phase);
      nx = nx + (m.vx * t + m.ax * sn);
      nz = nz + (m.vz * t + m.az * sn);
    }
    const double un = (nx - x[i]) / dt, wn = (nz - z[i]) / dt;
    u[i] = un;
    w[i] = wn;
    x[i] = nx;
    z[i] = nz;
    xz[i] = make_double2(nx, nz);
    uw[i] = make_double2(un, wn);
  }
}

// Pressure write-back of ComputeImplicitForces, Computer.hpp:1076-1097.
__global__ void __launch_bounds__(kBlock) k_writeback(size_t n, const uint8_t* __restrict__ type,
                                                      const double* __restrict__ sol,
                                                      double* __restrict__ p) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const uint8_t t = type[i];
  if (t != kDummy && t != kDisabled) {
    const double v = sol[i];
    p[i] = (v < 0) ? 0 : v;
  }
}

// ModifyByPressureGradient, midpoint form (Computer.hpp:1501-1521).
__global__ void __launch_bounds__(kBlock) k_gradient(size_t n, const double* __restrict__ x,
                                                     const double* __restrict__ z,
                                                     const double* __restrict__ p,
                                                     const uint8_t* __restrict__ type, NV nb, Consts k,
                                                     const double* __restrict__ nws,
                                                     double* __restrict__ dux, double* __restrict__ duz) {
  const size_t i = thread_particle();
  if (i >= n) return;
  if (type[i] != kFluid) return;
  const double xi = x[i], zi = z[i], pi = p[i];
  double fx = 0, fz = 0, sx = 0, sz = 0;
  struct F {
    double x, z, p;
    uint8_t t;
  };
  walk_list<kWalk, F>(static_cast<uint32_t>(i), nb.count[i], nb.ids, nb.kmax,
                      [&](uint32_t j) { return F{x[j], z[j], p[j], type[j]}; },
                      [&](uint32_t, const F& f) {
                        const double dx = f.x - xi, dz = f.z - zi;
                        const double r2 = dx * dx + dz * dz;
                        if (r2 < k.re2) {
                          const double wg = weight(sqrt(r2), k.re);
                          sx += wg * dx;
                          sz += wg * dz;
                          if (f.t != kDummy) {
                            const double s = (f.p + pi) / r2 * wg;
                            fx += s * dx;
                            fz += s * dz;
                          }
                        }
                      });
  double gxp, gzp;
  if (ghost_pos(xi, zi, sx, sz, nws[i], k.n0, k.re, gxp, gzp)) {
    const double dx = gxp - xi, dz = gzp - zi;
    const double r2 = dx * dx + dz * dz;
    const double s = (0.0 + pi) / r2 * weight(sqrt(r2), k.re);
    fx += s * dx;
    fz += s * dz;
  }
  const double coef = -k.dt / k.rho * 2 / k.n0;
  dux[i] = coef * fx;
  duz[i] = coef * fz;
}

// DynamicStabilize first pass (Computer.hpp:1576-1635).
__global__ void __launch_bounds__(kBlock) k_ds(size_t n, const double* __restrict__ x,
                                               const double* __restrict__ z,
                                               const double* __restrict__ ox,
                                               const double* __restrict__ oz,
                                               const uint8_t* __restrict__ type, NV nb, Consts k,
                                               double* __restrict__ dux, double* __restrict__ duz) {
  const size_t i = thread_particle();
  if (i >= n) return;
  if (type[i] != kFluid) return;
  const double xi = x[i], zi = z[i], oxi = ox[i], ozi = oz[i];
  double fx = 0, fz = 0;
  // X0 is read only for the few neighbours closer than d = (1 - C) l0
  struct F {
    double x, z;
    uint8_t t;
  };
  walk_list<kWalk, F>(static_cast<uint32_t>(i), nb.count[i], nb.ids, nb.kmax,
                      [&](uint32_t j) { return F{x[j], z[j], type[j]}; },
                      [&](uint32_t j, const F& f) {
    const double dx = f.x - xi, dz = f.z - zi;
    const double r2 = dx * dx + dz * dz;
    if (r2 < k.re2 && f.t != kDummy && r2 < k.d2) {
      const double e0x = ox[j] - oxi, e0z = oz[j] - ozi;
      const double nrm = sqrt(e0x * e0x + e0z * e0z);
      const double ex = e0x / nrm, ez = e0z / nrm;
      const double rpar = dx * ex + dz * ez;
      const double px = dx - rpar * ex, pz = dz - rpar * ez;
      const double rperp2 = px * px + pz * pz;
      const double m = sqrt(k.d2 - rperp2) - rpar;
      fx += m * ex;
      fz += m * ez;
    }
                      });
  const double coef = -1.0 / (2 * k.dt * k.n0);
  dux[i] = coef * fx;
  duz[i] = coef * fz;
}

// Second passes of ModifyByPressureGradient / DynamicStabilize
// (Computer.hpp:1544-1563, :1638-1655).
__global__ void __launch_bounds__(kBlock) k_apply_du(size_t n, double* __restrict__ x,
                                                     double* __restrict__ z, double* __restrict__ u,
                                                     double* __restrict__ w,
                                                     const uint8_t* __restrict__ type,
                                                     const double* __restrict__ dux,
                                                     const double* __restrict__ duz, double dt) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (type[i] != kFluid) return;
  const double ax = dux[i], az = duz[i];
  u[i] = u[i] + ax;
  w[i] = w[i] + az;
  x[i] = x[i] + ax * dt;
  z[i] = z[i] + az * dt;
}

// max over particles of u.u ignoring NaN (DetermineDt, Computer.hpp:764-771).
__global__ void __launch_bounds__(256) k_max_speed2(size_t n, const double* __restrict__ u,
                                                    const double* __restrict__ w,
                                                    double* __restrict__ part) {
  __shared__ double s[4];
  double m = 0;
  bool any = false;
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) {
    const double v = u[i] * u[i] + w[i] * w[i];
    if (!any || v > m) {
      if (v == v) {
        m = v;
        any = true;
      }
    }
  }
  if (!any) m = -1.0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(m, o, 64);
    m = (t > m) ? t : m;
  }
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = s[0];
    for (int q = 1; q < 4; q++) b = (s[q] > b) ? s[q] : b;
    part[blockIdx.x] = b;
  }
}

__global__ void __launch_bounds__(256) k_copy_probe(size_t n4, const float4* __restrict__ a,
                                                    float4* __restrict__ b) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i < n4) b[i] = a[i];
}

__global__ void k_math_check(size_t n, const double* a, const double* b, double* s, double* d) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  s[i] = sqrt(a[i]);
  d[i] = a[i] / b[i];
}

NV nview(const Context& c) { return NV{c.nbr, c.ncount, c.kstride}; }

}  // namespace

void density(Context& c, bool ecs, bool savex, bool dens) {
  const size_t n = c.n;
  if (n == 0) return;
  const Consts k = consts(c);
  const NV nb = nview(c);
  const double bytes = n * (16.0 + 1 + 4) + 4.0 * c.nbr_total + (ecs ? 16.0 * n + 8.0 * n : 0) +
                       (dens ? 16.0 * n : 16.0 * n) + (savex ? 16.0 * n : 0);
  const int kid = savex ? MPS_K_DENSITY_SAVE : MPS_K_DENSITY_ECS;
  tracked(c, kid, bytes, [&] {
    const dim3 g(grid_for(n)), b(kBlock);
#define MPS_DENS(D, E, S)                                                                         \
  hipLaunchKernelGGL((k_density<D, E, S>), g, b, 0, c.stream, n, c.x, c.z, c.u, c.w,              \
                     static_cast<const uint8_t*>(c.type), nb, k, c.nd, c.nws, c.ecs, c.ox, c.oz)
    if (dens && ecs && !savex) MPS_DENS(true, true, false);
    else if (dens && !ecs && savex) MPS_DENS(true, false, true);
    else if (dens && !ecs && !savex) MPS_DENS(true, false, false);
    else if (!dens && ecs && !savex) MPS_DENS(false, true, false);
    else if (!dens && !ecs && savex) MPS_DENS(false, false, true);
    else MPS_DENS(true, true, true);
#undef MPS_DENS
  });
  check_launch();
}

void density_variation(Context& c, double* out_dev) {
  const size_t n = c.n;
  if (n == 0) return;
  hipLaunchKernelGGL(k_hs, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z, c.u, c.w,
                     nview(c), consts(c), static_cast<const double*>(c.nws), out_dev);
  check_launch();
}

// Wall targets and the integration of ComputeExplicitForces (accelerations in dux, duz).
void explicit_update(Context& c) {
  const size_t n = c.n;
  if (n == 0) return;
  // positionWallPre(t, dt) then positionWall(i, t, dt) for every non-fluid i
  // (Computer.hpp:993, :1015), after the accelerations and before the update.
  if (c.wall_fn || c.wall_pre) wall_targets_from_callback(c);
  // x, z, u, w, type, accelerations and wall targets read; x, z, u, w and
  // the packed pairs written
  tracked(c, MPS_K_UPDATE, n * (32.0 + 1 + 16 + 16 + 32 + 32), [&] {
    hipLaunchKernelGGL(k_explicit_update, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z,
                       c.u, c.w, static_cast<const uint8_t*>(c.type),
                       static_cast<const double*>(c.dux), static_cast<const double*>(c.duz),
                       static_cast<const double*>(c.wall_x), static_cast<const double*>(c.wall_z),
                       c.dt, static_cast<const uint8_t*>(c.wall_motions ? c.wall_group : nullptr),
                       static_cast<const mps_wall_motion*>(c.wall_motion), static_cast<unsigned>(c.wall_motions),
                       c.t, c.xz, c.uw);
  });
  check_launch();
}

void explicit_forces(Context& c) {
  const size_t n = c.n;
  if (n == 0) return;
  const Consts k = consts(c);
  tracked(c, MPS_K_EXPLICIT, n * (32.0 + 1 + 4 + 8 + 16) + 4.0 * c.nbr_total, [&] {
    hipLaunchKernelGGL(k_explicit, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z, c.u, c.w,
                       static_cast<const uint8_t*>(c.type), nview(c), k,
                       static_cast<const double*>(c.nws), c.dux, c.duz);
  });
  check_launch();
  explicit_update(c);
}

void pack_state(Context& c) {
  const size_t n = c.n;
  if (n == 0) return;
  hipLaunchKernelGGL(k_pack_state, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n,
                     static_cast<const double*>(c.x), static_cast<const double*>(c.z),
                     static_cast<const double*>(c.u), static_cast<const double*>(c.w), c.xz, c.uw);
  check_launch();
}

void density_and_explicit(Context& c) {
  const size_t n = c.n;
  if (n == 0) return;
  pack_state(c);
  // x, z, u, w, type read; n, nWithoutSpp, ECS, accelerations written
  tracked(c, MPS_K_DENSITY_ECS, n * (32.0 + 1 + 4 + 24 + 16) + 4.0 * c.nbr_total, [&] {
    hipLaunchKernelGGL(k_density_explicit, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z, c.u,
                       c.w, static_cast<const uint8_t*>(c.type), nview(c), consts(c), c.nd, c.nws, c.ecs,
                       c.dux, c.duz, static_cast<const double2*>(c.xz), static_cast<const double2*>(c.uw));
  });
  check_launch();
  explicit_update(c);
}

void wall_targets_from_callback(Context& c) {
  const size_t n = c.n;
  std::vector<uint8_t> t(n);
  MPS_HIP(hipMemcpyAsync(t.data(), c.type, n, hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
  if (c.wall_pre) c.wall_pre(c.wall_user, c.t, c.dt);
  if (!c.wall_fn) return;
  // the callback sees the caller's indices, ascending (the reference's loop
  // order); the targets go to the storage slots of those particles
  const std::vector<uint32_t>& orig = storage_order(c);
  std::vector<std::pair<uint32_t, uint32_t>> walls;  // (caller index, storage slot)
  for (size_t k = 0; k < n; k++)
    if (t[k] != kFluid) walls.emplace_back(orig[k], static_cast<uint32_t>(k));
  if (walls.empty()) return;
  std::sort(walls.begin(), walls.end());
  std::vector<uint32_t> idx(walls.size());
  for (size_t m = 0; m < walls.size(); m++) idx[m] = walls[m].first;
  std::vector<double> hx(idx.size()), hz(idx.size());
  c.wall_fn(c.wall_user, c.t, c.dt, idx.size(), idx.data(), hx.data(), hz.data());
  // scatter through full host mirrors of the target arrays
  std::vector<double> wx(n), wz(n);
  MPS_HIP(hipMemcpyAsync(wx.data(), c.wall_x, n * 8, hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipMemcpyAsync(wz.data(), c.wall_z, n * 8, hipMemcpyDeviceToHost, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
  for (size_t m = 0; m < walls.size(); m++) {
    wx[walls[m].second] = hx[m];
    wz[walls[m].second] = hz[m];
  }
  MPS_HIP(hipMemcpyAsync(c.wall_x, wx.data(), n * 8, hipMemcpyHostToDevice, c.stream));
  MPS_HIP(hipMemcpyAsync(c.wall_z, wz.data(), n * 8, hipMemcpyHostToDevice, c.stream));
  MPS_HIP(hipStreamSynchronize(c.stream));
}

void pressure_writeback(Context& c) {
  const size_t n = c.n;
  if (n == 0) return;
  tracked(c, MPS_K_OTHER, n * 17.0, [&] {
    hipLaunchKernelGGL(k_writeback, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n,
                       static_cast<const uint8_t*>(c.type), static_cast<const double*>(c.sol), c.p);
  });
  check_launch();
}

void pressure_gradient(Context& c) {
  const size_t n = c.n;
  if (n == 0) return;
  const Consts k = consts(c);
  tracked(c, MPS_K_GRADIENT, n * (24.0 + 1 + 4 + 8 + 16) + 4.0 * c.nbr_total, [&] {
    hipLaunchKernelGGL(k_gradient, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z, c.p,
                       static_cast<const uint8_t*>(c.type), nview(c), k,
                       static_cast<const double*>(c.nws), c.dux, c.duz);
  });
  tracked(c, MPS_K_UPDATE, n * (32.0 + 1 + 16 + 32), [&] {
    hipLaunchKernelGGL(k_apply_du, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z, c.u, c.w,
                       static_cast<const uint8_t*>(c.type), static_cast<const double*>(c.dux),
                       static_cast<const double*>(c.duz), c.dt);
  });
  check_launch();
}

void dynamic_stabilize(Context& c) {
  const size_t n = c.n;
  if (n == 0) return;
  const Consts k = consts(c);
  tracked(c, MPS_K_DS, n * (32.0 + 1 + 4 + 16) + 4.0 * c.nbr_total, [&] {
    hipLaunchKernelGGL(k_ds, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z,
                       static_cast<const double*>(c.ox), static_cast<const double*>(c.oz),
                       static_cast<const uint8_t*>(c.type), nview(c), k, c.dux, c.duz);
  });
  tracked(c, MPS_K_UPDATE, n * (32.0 + 1 + 16 + 32), [&] {
    hipLaunchKernelGGL(k_apply_du, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, n, c.x, c.z, c.u, c.w,
                       static_cast<const uint8_t*>(c.type), static_cast<const double*>(c.dux),
                       static_cast<const double*>(c.duz), c.dt);
  });
  check_launch();
}

// Computer::DetermineDt, Computer.hpp:759-777.
double max_speed_dt(Context& c) {
  const size_t n = c.n;
  double maxU = 0;
  if (n > 0) {
    const int nb = 256;
    double* part = c.partA;  // scratch: CG partial slots are free outside a solve
    hipLaunchKernelGGL(k_max_speed2, dim3(nb), dim3(256), 0, c.stream, n, static_cast<const double*>(c.u),
                       static_cast<const double*>(c.w), part);
    check_launch();
    std::vector<double> h(nb);
    double first[2];
    MPS_HIP(hipMemcpyAsync(h.data(), part, nb * 8, hipMemcpyDeviceToHost, c.stream));
    MPS_HIP(hipMemcpyAsync(&first[0], c.u, 8, hipMemcpyDeviceToHost, c.stream));
    MPS_HIP(hipMemcpyAsync(&first[1], c.w, 8, hipMemcpyDeviceToHost, c.stream));
    MPS_HIP(hipStreamSynchronize(c.stream));
    const double v0 = first[0] * first[0] + first[1] * first[1];
    double m = -1.0;
    for (int q = 0; q < nb; q++) m = (h[q] > m) ? h[q] : m;
    // std::max_element keeps the first element when it is NaN
    const double best = (v0 != v0) ? v0 : m;
    maxU = sqrt(best);
  }
  const double dtmax = c.env.max_dt, dxmax = c.env.max_dx;
  if (maxU == 0) return dtmax;
  const double a = dxmax / maxU;
  return (dtmax < a) ? dtmax : a;
}

double device_copy_probe(size_t bytes) {
  const size_t n4 = bytes / 16;
  if (n4 == 0 || n4 > (1ull << 36)) throw Failure{MPS_ERR_BAD_ARG, "copy probe size out of range"};
  float4* a = dalloc<float4>(n4);
  float4* b = dalloc<float4>(n4);
  MPS_HIP(hipMemset(a, 0, n4 * 16));
  hipEvent_t e0, e1;
  MPS_HIP(hipEventCreate(&e0));
  MPS_HIP(hipEventCreate(&e1));
  // one untimed launch, then the mean of five
  const unsigned blocks = grid_for(n4, 256);
  hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, 0, n4, static_cast<const float4*>(a), b);
  MPS_HIP(hipEventRecord(e0, 0));
  for (int k = 0; k < 5; k++)
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, 0, n4, static_cast<const float4*>(a), b);
  MPS_HIP(hipEventRecord(e1, 0));
  MPS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MPS_HIP(hipEventElapsedTime(&ms, e0, e1));
  ms /= 5;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dfree(a);
  dfree(b);
  return ms;
}

void device_math_check(size_t n, const double* a, const double* b, double* s, double* d) {
  double *da = dalloc<double>(n), *db = dalloc<double>(n), *ds = dalloc<double>(n), *dd = dalloc<double>(n);
  MPS_HIP(hipMemcpy(da, a, n * 8, hipMemcpyHostToDevice));
  MPS_HIP(hipMemcpy(db, b, n * 8, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_math_check, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, da, db, ds, dd);
  check_launch();
  MPS_HIP(hipDeviceSynchronize());
  MPS_HIP(hipMemcpy(s, ds, n * 8, hipMemcpyDeviceToHost));
  MPS_HIP(hipMemcpy(d, dd, n * 8, hipMemcpyDeviceToHost));
  dfree(da);
  dfree(db);
  dfree(ds);
  dfree(dd);
}

}  // namespace mps
