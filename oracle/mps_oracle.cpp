// CPU oracle for the per-timestep MPS path of OpenMps (Computer::ForwardTime).
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++17 + OpenMP and
// without Boost or ViennaCL, the arithmetic of the reference solver so that the
// HIP implementation can be checked against it bit for bit.  Only tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it; the
// product library never links or calls it.
//
// Each function names the reference lines it follows (paths relative to the
// OpenMps source tree).  Default method set of src/OpenMps/defines.hpp:
// PRESSURE_GRADIENT_MIDPOINT, MPS_HS, MPS_HL, MPS_ECS, MPS_DS, MPS_SPP, DIM=2.
//
// Floating-point rules that make the restatement bit-exact:
//  * compile with -ffp-contract=off (no FMA contraction);
//  * every per-particle sum runs sequentially in neighbour-list order;
//  * expressions keep the association order of the reference source.
//
// Extension (not in the reference): an optional Laplacian radius r_lap used by
// the viscosity term and the PPE matrix.  With r_lap == r_e every result is the
// reference's.  The SPP ghost is always built from the r_e neighbourhood.

#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {

enum : uint8_t { kFluid = 0, kWall = 1, kDummy = 2, kDisabled = 3 };  // Particle.hpp:16-29

// Particle::W, Particle.hpp:71-75
inline double weight(double r, double re) { return ((0 < r) && (r < re)) ? (re / r - 1) : 0; }

// std::min / std::max as the reference calls them (first argument wins ties / NaN).
inline double ref_min(double a, double b) { return (b < a) ? b : a; }
inline double ref_max(double a, double b) { return (a < b) ? b : a; }

// Lattice sum of Environment.hpp:164-209 (loop order i, j in [-R, R), R = ceil(r/l0)).
double lattice_n0(double r_by_l0, double l0, double r) {
  const int range = static_cast<int>(std::ceil(r_by_l0));
  double n0 = 0;
  for (int i = -range; i < range; i++)
    for (int j = -range; j < range; j++) {
      if (i == 0 && j == 0) continue;
      const double px = i * l0, pz = j * l0;
      const double d = std::sqrt(std::fabs(px) * std::fabs(px) + std::fabs(pz) * std::fabs(pz));
      if (d < r) n0 += weight(d, r);
    }
  return n0;
}

struct Env {
  double max_dt, max_dx, l0, r_e, r_lap, gx, gz, rho, nu, min_x, min_z, max_x, max_z;
  double neighbor_length, n0, n0_lap, courant;
};

// Environment ctor, Environment.hpp:101-216 (MaxDt :138, MaxDx :143, R_e :145,
// G :149, NeighborLength :161, n0 :164-209).
Env make_env(double maxDt, double courant, double g, double rho, double nu, double reByl0,
             double l0, double minX, double minZ, double maxX, double maxZ, double rlapByl0) {
  Env e{};
  e.max_dt = ref_min(maxDt, std::sqrt(2 * (courant * l0) / g));
  e.max_dx = courant * l0;
  e.l0 = l0;
  e.r_e = reByl0 * l0;
  e.gx = 0;
  e.gz = -g;
  e.rho = rho;
  e.nu = nu;
  e.min_x = minX;
  e.min_z = minZ;
  e.max_x = maxX;
  e.max_z = maxZ;
  e.courant = courant;
  const double rmax_by_l0 = (rlapByl0 > reByl0) ? rlapByl0 : reByl0;
  e.neighbor_length = rmax_by_l0 * l0 * (1 + courant * 2);
  e.n0 = lattice_n0(reByl0, l0, e.r_e);
  if (rlapByl0 == reByl0) {
    e.r_lap = e.r_e;
    e.n0_lap = e.n0;
  } else {
    e.r_lap = rlapByl0 * l0;
    e.n0_lap = lattice_n0(rlapByl0, l0, e.r_lap);
  }
  return e;
}

// Deterministic FP64 dot product: fixed 4096-element chunks summed in order,
// chunk sums added in order.  Independent of the OpenMP thread count.
double det_dot(const std::vector<double>& a, const std::vector<double>& b) {
  const size_t n = a.size();
  const size_t chunk = 4096;
  const size_t nc = (n + chunk - 1) / chunk;
  std::vector<double> part(nc, 0.0);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < static_cast<long>(nc); c++) {
    double s = 0;
    const size_t lo = c * chunk, hi = std::min(n, lo + chunk);
    for (size_t i = lo; i < hi; i++) s += a[i] * b[i];
    part[c] = s;
  }
  double s = 0;
  for (size_t c = 0; c < nc; c++) s += part[c];
  return s;
}

class Oracle {
 public:
  Env env;
  double eps;
  double t = 0, dt = 0;
  std::vector<double> x, z, u, w, p, n;
  std::vector<uint8_t> type;
  std::vector<double> wall_x, wall_z;  // static wall positions, Main.cpp:305-318
  // grid, Grid.hpp:137-150
  long gnx = 0, gnz = 0;
  size_t cap = 0;
  std::vector<uint32_t> cells;  // [gnx][gnz][1+cap], slot 0 = count
  // neighbour table, Computer.hpp:473 / :1767-1770
  size_t kmax = 0;
  std::vector<uint32_t> nbr, ncount;
  std::vector<double> nws, ecs, ox, oz, dux, duz;
  // PPE, Computer.hpp:462-536
  std::vector<uint64_t> rowptr;
  std::vector<uint32_t> col;
  std::vector<double> val, b, sol;
  long cg_iters = -1;
  double cg_rr0 = 0, cg_rr = 0;
  int verbose = 0;        // 1: progress lines on stderr from the CG loop
  int stop_rule = 0;      // 0: r.r < eps^2 r0.r0 (reference); 1: |r| <= tol |b|
  double stop_tol = 0;    // tolerance of rule 1
  std::string err;

  Oracle(const Env& e, double allowable) : env(e), eps(allowable) {
    const double NL = env.neighbor_length;
    gnx = static_cast<long>(std::ceil((env.max_x - env.min_x) / NL)) + 2;
    gnz = static_cast<long>(std::ceil((env.max_z - env.min_z) / NL)) + 2;
    const long c1 = static_cast<long>(std::ceil(NL / env.l0)) + 1;
    cap = static_cast<size_t>(c1 * c1);
    kmax = cap * 9;
    cells.assign(static_cast<size_t>(gnx * gnz) * (1 + cap), 0);
  }

  size_t size() const { return x.size(); }

  // Computer::AddParticles, Computer.hpp:1754-1777 (initial positions double as
  // the static wall positions of Main.cpp:305-318).
  void add(size_t cnt, const double* px, const double* pz, const double* pu, const double* pw,
           const double* pp, const double* pn, const uint8_t* pt) {
    for (size_t k = 0; k < cnt; k++) {
      x.push_back(px[k]);
      z.push_back(pz[k]);
      u.push_back(pu[k]);
      w.push_back(pw[k]);
      p.push_back(pp[k]);
      n.push_back(pn[k]);
      type.push_back(pt[k]);
      wall_x.push_back(px[k]);
      wall_z.push_back(pz[k]);
    }
    const size_t N = size();
    nbr.assign(N * kmax, 0);
    ncount.assign(N, 0);
    nws.resize(N, 0.0);
    ecs.resize(N, 0.0);
    ox.resize(N, 0.0);
    oz.resize(N, 0.0);
    dux.resize(N, 0.0);
    duz.resize(N, 0.0);
  }

  // Grid::Block, Grid.hpp:250-254, with the range test of Grid::Store :278-290
  // done on the floored double (NaN / inf / huge fall outside, as on x86).
  bool block_of(double px, double pz, long& bx, long& bz) const {
    const double fx = std::floor((px - env.min_x) / env.neighbor_length);
    const double fz = std::floor((pz - env.min_z) / env.neighbor_length);
    if (!(fx >= 0 && fx < static_cast<double>(gnx) && fz >= 0 && fz < static_cast<double>(gnz)))
      return false;
    bx = static_cast<long>(fx);
    bz = static_cast<long>(fz);
    return true;
  }
  uint32_t* cell(long bx, long bz) { return &cells[static_cast<size_t>(bx * gnz + bz) * (1 + cap)]; }
  const uint32_t* cell(long bx, long bz) const {
    return &cells[static_cast<size_t>(bx * gnz + bz) * (1 + cap)];
  }

  // Computer::SearchNeighbor, Computer.hpp:698-756, with Grid::Store
  // (Grid.hpp:276-331) and the 9-block iterator order of Grid.hpp:393-401 and
  // :433-518 (x-major, blocks outside the grid or empty are skipped).
  int search() {
    const size_t N = size();
    for (long c = 0; c < gnx * gnz; c++) cells[static_cast<size_t>(c) * (1 + cap)] = 0;
    for (size_t i = 0; i < N; i++) {
      if (type[i] == kDisabled) continue;
      long bx = 0, bz = 0;
      if (!block_of(x[i], z[i], bx, bz)) {
        type[i] = kDisabled;
        continue;
      }
      uint32_t* c = cell(bx, bz);
      const uint32_t l = c[0]++;
      if (l >= cap) {
        err = "Too many particle in a block";
        return 2;
      }
      c[1 + l] = static_cast<uint32_t>(i);
    }
    const double NL = env.neighbor_length;
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(N); ii++) {
      const size_t i = static_cast<size_t>(ii);
      if (type[i] == kDisabled) continue;
      long bx = 0, bz = 0;
      block_of(x[i], z[i], bx, bz);
      uint32_t cnt = 0;
      for (int ox_ = -1; ox_ <= 1; ox_++)
        for (int oz_ = -1; oz_ <= 1; oz_++) {
          const long cx = bx + ox_, cz = bz + oz_;
          if (cx < 0 || cx >= gnx || cz < 0 || cz >= gnz) continue;
          const uint32_t* c = cell(cx, cz);
          for (uint32_t l = 0; l < c[0]; l++) {
            const uint32_t j = c[1 + l];
            if (j == i || type[j] == kDisabled) continue;
            const double rx = x[i] - x[j], rz = z[i] - z[j];
            const double r = std::sqrt(rx * rx + rz * rz);  // Computer::R, :568-572
            if (r < NL) nbr[i * kmax + cnt++] = j;
          }
        }
      ncount[i] = cnt;
    }
    return 0;
  }

  // Neighbour walk of Computer::AccumulateNeighbor, Computer.hpp:618-662:
  // f(j, dx, dz, r2) for list entries with |x_j - x_i|^2 < rf^2, dx = x_j - x_i.
  template <class F>
  void pairs(size_t i, double rf, F&& f) const {
    const double rf2 = rf * rf;
    const uint32_t* lst = &nbr[i * kmax];
    for (uint32_t k = 0; k < ncount[i]; k++) {
      const uint32_t j = lst[k];
      if (j == i) continue;
      const double dx = x[j] - x[i], dz = z[j] - z[i];
      const double r2 = dx * dx + dz * dz;
      if (r2 < rf2) f(j, dx, dz, r2);
    }
  }

  // SPP ghost particle, Computer.hpp:657-693: exists when nWithoutSpp_i < n0
  // and the weighted centroid offset is above DBL_EPSILON.
  bool ghost(size_t i, double& gx, double& gz) const {
    const double n0 = env.n0, re = env.r_e;
    double sx = 0, sz = 0;
    pairs(i, re, [&](uint32_t, double dx, double dz, double r2) {
      const double wgt = weight(std::sqrt(r2), re);
      sx += wgt * dx;
      sz += wgt * dz;
    });
    sx /= n0;
    sz /= n0;
    const double thisN = nws[i];
    if (!(thisN < n0)) return false;
    const double rg = std::sqrt(sx * sx + sz * sz);
    if (!(rg > DBL_EPSILON)) return false;
    const double wspp = n0 - thisN;
    const double rspp = re / (wspp + 1);
    gx = x[i] - (rspp / rg) * sx;
    gz = z[i] - (rspp / rg) * sz;
    return true;
  }

  // Computer::ComputeNeighborDensities, Computer.hpp:780-833.
  void density() {
    const double n0 = env.n0, re = env.r_e;
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(size()); ii++) {
      const size_t i = static_cast<size_t>(ii);
      nws[i] = n0;
      if (type[i] == kDummy || type[i] == kDisabled) continue;
      double s = 0.0;
      pairs(i, re, [&](uint32_t, double, double, double r2) { s += weight(std::sqrt(r2), re); });
      nws[i] = s;
      n[i] = ref_max(s, n0);
    }
  }

  // Computer::NeighborDensityVariationSpeed (HS), Computer.hpp:838-872.
  double hs(size_t i) const {
    const double n0 = env.n0, re = env.r_e;
    if (nws[i] < n0) return 0.0;
    double s = 0.0;
    pairs(i, re, [&](uint32_t j, double dx, double dz, double r2) {
      (void)r2;
      const double du = u[j] - u[i], dw = w[j] - w[i];
      const double r = std::sqrt(std::fabs(dx) * std::fabs(dx) + std::fabs(dz) * std::fabs(dz));
      s += (dx * du + dz * dw) / (r * r * r);
    });
    return -re * s;
  }

  // Computer::ComputeErrorCorrection (ECS), Computer.hpp:877-910.
  void error_correction() {
    const double n0 = env.n0;
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(size()); ii++) {
      const size_t i = static_cast<size_t>(ii);
      if (type[i] == kDummy || type[i] == kDisabled) continue;
      const double speed = hs(i);
      const double e = (n[i] - n0) / n0;
      ecs[i] = std::fabs(e) * speed + std::fabs(speed) * e;
    }
  }

  // Computer::ComputeExplicitForces, Computer.hpp:914-1021 (HL viscosity :963-965,
  // gravity :985, integration :1006-1010, walls :1013-1019).
  void explicit_forces() {
    const double n0l = env.n0_lap, rl = env.r_lap, nu = env.nu;
    const size_t N = size();
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(N); ii++) {
      const size_t i = static_cast<size_t>(ii);
      if (type[i] != kFluid) continue;
      double vx = 0, vz = 0;
      pairs(i, rl, [&](uint32_t j, double, double, double r2) {
        if (type[j] == kDummy) return;
        const double r = std::sqrt(r2);
        const double c = nu * 3 * rl / n0l / (r * r * r);
        vx += c * (u[j] - u[i]);
        vz += c * (w[j] - w[i]);
      });
      double gxp, gzp;
      if (ghost(i, gxp, gzp)) {
        const double rx = x[i] - gxp, rz = z[i] - gzp;
        const double r = std::sqrt(rx * rx + rz * rz);
        const double c = nu * 3 * rl / n0l / (r * r * r);
        vx += c * (u[i] - u[i]);
        vz += c * (w[i] - w[i]);
      }
      dux[i] = vx + env.gx;
      duz[i] = vz + env.gz;
    }
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(N); ii++) {
      const size_t i = static_cast<size_t>(ii);
      if (type[i] == kFluid) {
        u[i] += dux[i] * dt;
        w[i] += duz[i] * dt;
        x[i] += u[i] * dt;
        z[i] += w[i] * dt;
      } else {
        const double nx_ = wall_x[i], nz_ = wall_z[i];
        const double nu_ = (nx_ - x[i]) / dt, nw_ = (nz_ - z[i]) / dt;
        x[i] = nx_;
        z[i] = nz_;
        u[i] = nu_;
        w[i] = nw_;
      }
    }
  }

  // Computer::SaveX, Computer.hpp:1025-1039.
  void save_x() {
    for (size_t i = 0; i < size(); i++)
      if (type[i] != kDummy && type[i] != kDisabled) {
        ox[i] = x[i];
        oz[i] = z[i];
      }
  }

  // Computer::SetPressurePoissonEquation, Computer.hpp:1145-1356: RHS with HS
  // and ECS (:1205-1219), HL matrix rows with the SPP ghost on the diagonal
  // (:1232-1326), stored as CSR with ascending columns (the layout that
  // viennacl::copy produces from the uBLAS compressed matrix, :1354).
  void ppe_assemble() {
    const size_t N = size();
    const double n0 = env.n0, n0l = env.n0_lap, rl = env.r_lap, rho = env.rho;
    b.assign(N, 0.0);
    sol.assign(N, 0.0);
    std::vector<std::vector<std::pair<uint32_t, double>>> rows(N);
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(N); ii++) {
      const size_t i = static_cast<size_t>(ii);
      auto& row = rows[i];
      if (type[i] == kDummy || type[i] == kDisabled) {
        b[i] = 0;
        sol[i] = 0;
        row.emplace_back(static_cast<uint32_t>(i), 1.0);
        continue;
      }
      const double speed = hs(i);
      b[i] = -rho / (n0 * dt) * (speed + ecs[i]);
      sol[i] = p[i];
      double diag = 0.0;
      pairs(i, rl, [&](uint32_t j, double, double, double r2) {
        if (type[j] == kDummy) return;
        const double r = std::sqrt(r2);
        const double a = 3 * rl / n0l / (r * r * r);
        row.emplace_back(j, a);
        diag += -a;
      });
      double gxp, gzp;
      if (ghost(i, gxp, gzp)) {
        const double rx = x[i] - gxp, rz = z[i] - gzp;
        const double r = std::sqrt(rx * rx + rz * rz);
        const double a = 3 * rl / n0l / (r * r * r);
        diag += -a;
      }
      row.emplace_back(static_cast<uint32_t>(i), diag);
    }
    rowptr.assign(N + 1, 0);
    for (size_t i = 0; i < N; i++) rowptr[i + 1] = rowptr[i] + rows[i].size();
    col.resize(rowptr[N]);
    val.resize(rowptr[N]);
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(N); ii++) {
      auto& row = rows[static_cast<size_t>(ii)];
      std::sort(row.begin(), row.end(),
                [](const std::pair<uint32_t, double>& a, const std::pair<uint32_t, double>& c) {
                  return a.first < c.first;
                });
      size_t o = rowptr[static_cast<size_t>(ii)];
      for (auto& e : row) {
        col[o] = e.first;
        val[o] = e.second;
        o++;
      }
    }
  }

  void spmv(const std::vector<double>& v, std::vector<double>& out) const {
    const size_t N = b.size();
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(N); ii++) {
      double s = 0;
      for (uint64_t k = rowptr[ii]; k < rowptr[ii + 1]; k++) s += val[k] * v[col[k]];
      out[ii] = s;
    }
  }

  // Computer::SolvePressurePoissonEquation, Computer.hpp:1359-1429:
  // unpreconditioned CG from x0, stopping when r.r < eps^2 (r0.r0), at most N
  // iterations, otherwise the reference throws Computer::Exception.
  int ppe_solve() {
    const size_t N = b.size();
    std::vector<double> r(N), pp(N), Ap(N);
    spmv(sol, Ap);
    for (size_t i = 0; i < N; i++) r[i] = b[i] - Ap[i];
    pp = r;
    double rr = det_dot(r, r);
    double residual0 = rr * eps * eps;
    cg_rr0 = rr;
    bool conv = (residual0 == 0);
    if (stop_rule == 1) {
      residual0 = stop_tol * stop_tol * det_dot(b, b);
      conv = (rr <= residual0);
    }
    long it = 0;
    for (size_t k = 0; k < N && !conv; k++) {
      spmv(pp, Ap);
      const double alpha = rr / det_dot(pp, Ap);
#pragma omp parallel for schedule(static)
      for (long i = 0; i < static_cast<long>(N); i++) {
        sol[i] += alpha * pp[i];
        r[i] -= alpha * Ap[i];
      }
      const double rrNew = det_dot(r, r);
      it++;
      conv = (stop_rule == 1) ? (rrNew <= residual0) : (rrNew < residual0);
      if (verbose && it % 500 == 0) {
        std::fprintf(stderr, "    oracle CG iteration %ld, r.r / target %.3e\n", it, rrNew / residual0);
        std::fflush(stderr);
      }
      if (!conv) {
        const double beta = rrNew / rr;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < static_cast<long>(N); i++) pp[i] = r[i] + beta * pp[i];
        rr = rrNew;
      } else {
        rr = rrNew;
      }
    }
    cg_iters = it;
    cg_rr = rr;
    if (!conv) {
      err = "Conjugate Gradient method couldn't solve Pressure Poison Equation";
      return 1;
    }
    return 0;
  }

  // Pressure write-back of Computer::ComputeImplicitForces, Computer.hpp:1076-1097.
  void pressure_writeback() {
    for (size_t i = 0; i < size(); i++)
      if (type[i] != kDummy && type[i] != kDisabled) {
        const double v = sol[i];
        p[i] = (v < 0) ? 0 : v;
      }
  }

  // Computer::ModifyByPressureGradient (midpoint form), Computer.hpp:1433-1564
  // (:1501-1521, update :1544-1563).
  void pressure_gradient() {
    const double re = env.r_e, n0 = env.n0;
    const size_t N = size();
    const double coef = -dt / env.rho * 2 / n0;
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(N); ii++) {
      const size_t i = static_cast<size_t>(ii);
      if (type[i] != kFluid) continue;
      double sx = 0, sz = 0;
      const double pi = p[i];
      pairs(i, re, [&](uint32_t j, double dx, double dz, double r2) {
        if (type[j] == kDummy) return;
        const double s = (p[j] + pi) / r2 * weight(std::sqrt(r2), re);
        sx += s * dx;
        sz += s * dz;
      });
      double gxp, gzp;
      if (ghost(i, gxp, gzp)) {
        const double dx = gxp - x[i], dz = gzp - z[i];
        const double r2 = dx * dx + dz * dz;
        const double s = (0.0 + pi) / r2 * weight(std::sqrt(r2), re);
        sx += s * dx;
        sz += s * dz;
      }
      dux[i] = coef * sx;
      duz[i] = coef * sz;
    }
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(N); ii++) {
      const size_t i = static_cast<size_t>(ii);
      if (type[i] != kFluid) continue;
      u[i] += dux[i];
      w[i] += duz[i];
      x[i] += dux[i] * dt;
      z[i] += duz[i] * dt;
    }
  }

  // Computer::DynamicStabilize (DS), Computer.hpp:1566-1657.
  void dynamic_stabilize() {
    const double re = env.r_e, n0 = env.n0;
    const double d = env.l0 - env.max_dx;
    const double d2 = d * d;
    const double coef = -1.0 / (2 * dt * n0);
    const size_t N = size();
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(N); ii++) {
      const size_t i = static_cast<size_t>(ii);
      if (type[i] != kFluid) continue;
      double sx = 0, sz = 0;
      pairs(i, re, [&](uint32_t j, double dx, double dz, double r2) {
        if (type[j] == kDummy) return;
        if (!(r2 < d2)) return;
        const double e0x = ox[j] - ox[i], e0z = oz[j] - oz[i];
        const double nrm = std::sqrt(std::fabs(e0x) * std::fabs(e0x) + std::fabs(e0z) * std::fabs(e0z));
        const double ex = e0x / nrm, ez = e0z / nrm;
        const double rpar = dx * ex + dz * ez;
        const double px_ = dx - rpar * ex, pz_ = dz - rpar * ez;
        const double rperp2 = px_ * px_ + pz_ * pz_;
        const double m = std::sqrt(d2 - rperp2) - rpar;
        sx += m * ex;
        sz += m * ez;
      });
      dux[i] = coef * sx;
      duz[i] = coef * sz;
    }
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < static_cast<long>(N); ii++) {
      const size_t i = static_cast<size_t>(ii);
      if (type[i] != kFluid) continue;
      u[i] += dux[i];
      w[i] += duz[i];
      x[i] += dux[i] * dt;
      z[i] += duz[i] * dt;
    }
  }

  // Computer::DetermineDt, Computer.hpp:759-777 (first maximum of |u|^2).
  double determine_dt() const {
    size_t best = 0;
    for (size_t i = 1; i < size(); i++)
      if ((u[best] * u[best] + w[best] * w[best]) < (u[i] * u[i] + w[i] * w[i])) best = i;
    double maxU = 0;
    if (size() > 0)
      maxU = std::sqrt(std::fabs(u[best]) * std::fabs(u[best]) + std::fabs(w[best]) * std::fabs(w[best]));
    return (maxU == 0) ? env.max_dt : ref_min(env.max_dx / maxU, env.max_dt);
  }

  // Computer::ForwardTime(dt), Computer.hpp:1700-1742.
  int forward(double step) {
    dt = step;
    t += dt;
    int rc = search();
    if (rc) return rc;
    density();
    error_correction();
    explicit_forces();
    density();
    save_x();
    ppe_assemble();
    rc = ppe_solve();
    if (rc) return rc;
    pressure_writeback();
    pressure_gradient();
    dynamic_stabilize();
    return 0;
  }
};

}  // namespace

extern "C" {

void* oracle_create(double maxDt, double courant, double g, double rho, double nu, double reByl0,
                    double l0, double minX, double minZ, double maxX, double maxZ, double rlapByl0,
                    double eps) {
  if (!(rlapByl0 > 0)) rlapByl0 = reByl0;
  return new Oracle(make_env(maxDt, courant, g, rho, nu, reByl0, l0, minX, minZ, maxX, maxZ, rlapByl0),
                    eps);
}
void oracle_destroy(void* h) { delete static_cast<Oracle*>(h); }
const char* oracle_last_error(void* h) { return static_cast<Oracle*>(h)->err.c_str(); }

// out[0..] = max_dt, max_dx, l0, r_e, r_lap, gx, gz, rho, nu, min_x, min_z,
//            max_x, max_z, neighbor_length, n0, n0_lap, courant
void oracle_env(void* h, double* out) {
  const Env& e = static_cast<Oracle*>(h)->env;
  const double v[] = {e.max_dt, e.max_dx, e.l0,    e.r_e,   e.r_lap, e.gx, e.gz,   e.rho, e.nu,
                      e.min_x,  e.min_z,  e.max_x, e.max_z, e.neighbor_length, e.n0, e.n0_lap, e.courant};
  std::memcpy(out, v, sizeof(v));
}
// out = nx, nz, cap, kmax
void oracle_grid(void* h, long* out) {
  Oracle* o = static_cast<Oracle*>(h);
  out[0] = o->gnx;
  out[1] = o->gnz;
  out[2] = static_cast<long>(o->cap);
  out[3] = static_cast<long>(o->kmax);
}

void oracle_add_particles(void* h, size_t cnt, const double* x, const double* z, const double* u,
                          const double* w, const double* p, const double* n, const uint8_t* t) {
  static_cast<Oracle*>(h)->add(cnt, x, z, u, w, p, n, t);
}
size_t oracle_size(void* h) { return static_cast<Oracle*>(h)->size(); }

void oracle_get_particles(void* h, double* x, double* z, double* u, double* w, double* p, double* n,
                          uint8_t* t) {
  Oracle* o = static_cast<Oracle*>(h);
  const size_t N = o->size();
  if (x) std::memcpy(x, o->x.data(), N * 8);
  if (z) std::memcpy(z, o->z.data(), N * 8);
  if (u) std::memcpy(u, o->u.data(), N * 8);
  if (w) std::memcpy(w, o->w.data(), N * 8);
  if (p) std::memcpy(p, o->p.data(), N * 8);
  if (n) std::memcpy(n, o->n.data(), N * 8);
  if (t) std::memcpy(t, o->type.data(), N);
}
void oracle_set_particles(void* h, const double* x, const double* z, const double* u, const double* w,
                          const double* p, const double* n, const uint8_t* t) {
  Oracle* o = static_cast<Oracle*>(h);
  const size_t N = o->size();
  if (x) std::memcpy(o->x.data(), x, N * 8);
  if (z) std::memcpy(o->z.data(), z, N * 8);
  if (u) std::memcpy(o->u.data(), u, N * 8);
  if (w) std::memcpy(o->w.data(), w, N * 8);
  if (p) std::memcpy(o->p.data(), p, N * 8);
  if (n) std::memcpy(o->n.data(), n, N * 8);
  if (t) std::memcpy(o->type.data(), t, N);
}
// which: 0 nWithoutSpp, 1 ecs, 2 saved x, 3 saved z, 4 du x, 5 du z, 6 ppe b, 7 ppe x
void oracle_get_aux(void* h, int which, double* out) {
  Oracle* o = static_cast<Oracle*>(h);
  const std::vector<double>* v[] = {&o->nws, &o->ecs, &o->ox, &o->oz, &o->dux, &o->duz, &o->b, &o->sol};
  if (which < 0 || which > 7) return;
  std::memcpy(out, v[which]->data(), v[which]->size() * 8);
}
void oracle_set_aux(void* h, int which, const double* in) {
  Oracle* o = static_cast<Oracle*>(h);
  std::vector<double>* v[] = {&o->nws, &o->ecs, &o->ox, &o->oz, &o->dux, &o->duz, &o->b, &o->sol};
  if (which < 0 || which > 7) return;
  std::memcpy(v[which]->data(), in, v[which]->size() * 8);
}
void oracle_set_time(void* h, double t, double dt) {
  static_cast<Oracle*>(h)->t = t;
  static_cast<Oracle*>(h)->dt = dt;
}
void oracle_get_time(void* h, double* out) {
  out[0] = static_cast<Oracle*>(h)->t;
  out[1] = static_cast<Oracle*>(h)->dt;
}
void oracle_set_wall_positions(void* h, const double* x, const double* z) {
  Oracle* o = static_cast<Oracle*>(h);
  std::memcpy(o->wall_x.data(), x, o->size() * 8);
  std::memcpy(o->wall_z.data(), z, o->size() * 8);
}

int oracle_stage_search(void* h) { return static_cast<Oracle*>(h)->search(); }
void oracle_stage_density(void* h) { static_cast<Oracle*>(h)->density(); }
void oracle_stage_error_correction(void* h) { static_cast<Oracle*>(h)->error_correction(); }
void oracle_stage_explicit(void* h) { static_cast<Oracle*>(h)->explicit_forces(); }
void oracle_stage_save_x(void* h) { static_cast<Oracle*>(h)->save_x(); }
void oracle_stage_ppe_assemble(void* h) { static_cast<Oracle*>(h)->ppe_assemble(); }
int oracle_stage_ppe_solve(void* h) { return static_cast<Oracle*>(h)->ppe_solve(); }
void oracle_stage_pressure_writeback(void* h) { static_cast<Oracle*>(h)->pressure_writeback(); }
void oracle_stage_pressure_gradient(void* h) { static_cast<Oracle*>(h)->pressure_gradient(); }
void oracle_stage_dynamic_stabilize(void* h) { static_cast<Oracle*>(h)->dynamic_stabilize(); }
void oracle_density_variation(void* h, double* out) {
  Oracle* o = static_cast<Oracle*>(h);
  for (size_t i = 0; i < o->size(); i++) out[i] = o->hs(i);
}
double oracle_determine_dt(void* h) { return static_cast<Oracle*>(h)->determine_dt(); }
int oracle_forward_time(void* h, double dt) { return static_cast<Oracle*>(h)->forward(dt); }
int oracle_forward_time_cfl(void* h, double* dt_out) {
  Oracle* o = static_cast<Oracle*>(h);
  const double dt = o->determine_dt();
  if (dt_out) *dt_out = dt;
  return o->forward(dt);
}

void oracle_neighbor_counts(void* h, uint32_t* out) {
  Oracle* o = static_cast<Oracle*>(h);
  std::memcpy(out, o->ncount.data(), o->size() * 4);
}
// ids laid out as CSR by the exclusive scan of the counts.
void oracle_neighbors(void* h, uint32_t* ids) {
  Oracle* o = static_cast<Oracle*>(h);
  size_t off = 0;
  for (size_t i = 0; i < o->size(); i++) {
    std::memcpy(ids + off, &o->nbr[i * o->kmax], o->ncount[i] * 4);
    off += o->ncount[i];
  }
}

size_t oracle_ppe_dim(void* h) { return static_cast<Oracle*>(h)->b.size(); }
size_t oracle_ppe_nnz(void* h) {
  Oracle* o = static_cast<Oracle*>(h);
  return o->rowptr.empty() ? 0 : o->rowptr.back();
}
void oracle_get_ppe(void* h, uint64_t* rowptr, uint32_t* col, double* val) {
  Oracle* o = static_cast<Oracle*>(h);
  std::memcpy(rowptr, o->rowptr.data(), o->rowptr.size() * 8);
  std::memcpy(col, o->col.data(), o->col.size() * 4);
  std::memcpy(val, o->val.data(), o->val.size() * 8);
}
// Installs an arbitrary system (columns ascending per row) for the CG stage.
void oracle_set_ppe(void* h, size_t n, const uint64_t* rowptr, const uint32_t* col, const double* val,
                    const double* b, const double* x0) {
  Oracle* o = static_cast<Oracle*>(h);
  o->rowptr.assign(rowptr, rowptr + n + 1);
  o->col.assign(col, col + rowptr[n]);
  o->val.assign(val, val + rowptr[n]);
  o->b.assign(b, b + n);
  o->sol.assign(x0, x0 + n);
}
// Extension used by the CPU baseline: rule 1 stops at |b - A x| <= tol |b|.
void oracle_set_verbose(void* h, int on) { static_cast<Oracle*>(h)->verbose = on; }
void oracle_set_stop_rule(void* h, int rule, double tol) {
  static_cast<Oracle*>(h)->stop_rule = rule;
  static_cast<Oracle*>(h)->stop_tol = tol;
}
void oracle_cg_stats(void* h, double* out) {
  Oracle* o = static_cast<Oracle*>(h);
  out[0] = static_cast<double>(o->cg_iters);
  out[1] = o->cg_rr0;
  out[2] = o->cg_rr;
}

}  // extern "C"
