// mps_computer.hpp - C++17 facade with the interface of OpenMps's Computer,
// implemented over the C ABI of mps.h (libmps_hip.so, MI355X / gfx950).
//
// Drop-in for the reference's public API (src/OpenMps/):
//   OpenMps::Particle            Particle.hpp:12-126
//   OpenMps::Environment         Environment.hpp:9-256 (ctor :101-216)
//   OpenMps::Computer<PW, PWP>   Computer.hpp:434-1790
//     ctor(allowableResidual, env, posWall, posWallPre)   :1674-1691
//     AddParticles(src)                                  :1754-1777
//     ForwardTime(dt) / ForwardTime()                    :1700-1751
//     Particles(), GetEnvironment()                      :1780-1789
//     Computer::Exception                                :1792-1798
//   OpenMps::CreateComputer(...)  Computer.hpp:1800-1815
//   OpenMps::Grid::Exception      Grid.hpp:122-129
//   OpenMps::OutputToCsv          Main.cpp:31-67 (same header and format)
// Particles live on the GPU; Particles() returns a host mirror refreshed
// lazily after each step.  positionWall(i, t, dt) / positionWallPre(t, dt)
// are called exactly where the reference calls them (once per step, for every
// non-fluid particle, between the explicit forces and the position update).
// Pass OpenMps::StaticWalls to keep walls at their initial positions on the
// device without host callbacks (what Main.cpp:305-318 does).
#ifndef MPS_COMPUTER_HPP_
#define MPS_COMPUTER_HPP_

#include <cstddef>
#include <cstdint>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "mps.h"

namespace OpenMps {

constexpr std::size_t DIM = 2;
constexpr std::size_t AXIS_X = 0;
constexpr std::size_t AXIS_Z = 1;

// Two-component vector standing in for ublas::c_vector<double, 2> (Vector.hpp:18).
struct Vector {
  double v[2] = {0.0, 0.0};
  double& operator[](std::size_t i) { return v[i]; }
  double operator[](std::size_t i) const { return v[i]; }
};
inline Vector CreateVector(double x, double z) { return Vector{{x, z}}; }

class Particle final {
 public:
  enum class Type { IncompressibleNewton, Wall, Dummy, Disabled };
  explicit Particle(Type t) : type_(t) {}
  void Disable() { type_ = Type::Disabled; }
  // Particle::W, Particle.hpp:71-75
  static double W(double r, double r_e) { return ((0 < r) && (r < r_e)) ? (r_e / r - 1) : 0; }
  const Vector& X() const { return x_; }
  Vector& X() { return x_; }
  const Vector& U() const { return u_; }
  Vector& U() { return u_; }
  const double& P() const { return p_; }
  double& P() { return p_; }
  const double& N() const { return n_; }
  double& N() { return n_; }
  const Type& TYPE() const { return type_; }

 private:
  Vector x_, u_;
  double p_ = 0, n_ = 0;
  Type type_;
};

class Environment final {
  mps_env e_;  // declared first: the public constants below are initialised from it
  double t_ = 0, dt_ = 0;

 public:
  Environment(double maxDt, double courant, double g, double rho, double nu, double r_eByl_0,
              double l_0, double minX, double minZ, double maxX, double maxZ)
      : e_(make(maxDt, courant, g, rho, nu, r_eByl_0, l_0, minX, minZ, maxX, maxZ)),
        MaxDt(e_.max_dt), MaxDx(e_.max_dx), L_0(e_.l0), R_e(e_.r_e), G(CreateVector(e_.gx, e_.gz)),
        Rho(e_.rho), Nu(e_.nu), MinX(CreateVector(e_.min_x, e_.min_z)),
        MaxX(CreateVector(e_.max_x, e_.max_z)), NeighborLength(e_.neighbor_length) {}
  explicit Environment(const mps_env& e)
      : e_(e), MaxDt(e.max_dt), MaxDx(e.max_dx), L_0(e.l0), R_e(e.r_e), G(CreateVector(e.gx, e.gz)),
        Rho(e.rho), Nu(e.nu), MinX(CreateVector(e.min_x, e.min_z)), MaxX(CreateVector(e.max_x, e.max_z)),
        NeighborLength(e.neighbor_length) {}
  Environment(const Environment&) = default;

  // Extension: a larger Laplacian radius for viscosity and the PPE matrix.
  Environment WithLaplacianRadius(double r_lapByl_0) const {
    mps_env e = e_;
    if (mps_env_set_laplacian_radius(&e, r_lapByl_0) != MPS_OK)
      throw std::invalid_argument("invalid Laplacian radius");
    return Environment(e, t_, dt_);
  }

  double N0() const { return e_.n0; }
  double T() const { return t_; }
  double Dt() const { return dt_; }
  const mps_env& Native() const { return e_; }

  const double MaxDt, MaxDx, L_0, R_e;
  const Vector G;
  const double Rho, Nu;
  const Vector MinX, MaxX;
  const double NeighborLength;

 private:
  template <class, class>
  friend class Computer;
  Environment(const mps_env& e, double t, double dt) : Environment(e) {
    t_ = t;
    dt_ = dt;
  }
  static mps_env make(double maxDt, double courant, double g, double rho, double nu, double r,
                      double l0, double minX, double minZ, double maxX, double maxZ) {
    mps_env e{};
    if (mps_env_init(&e, maxDt, courant, g, rho, nu, r, l0, minX, minZ, maxX, maxZ) != MPS_OK)
      throw std::invalid_argument("invalid environment parameters");
    return e;
  }
};

class Grid final {
 public:
  struct Exception : public std::runtime_error {
    explicit Exception(const std::string& msg) : std::runtime_error(msg) {}
  };
};

// Tag: walls stay where mps_add_particles put them (device resident).
struct StaticWalls {};

template <class POSITION_WALL, class POSITION_WALL_PRE>
class Computer final {
 public:
  struct Exception : public std::runtime_error {
    explicit Exception(const std::string& msg) : std::runtime_error(msg) {}
  };

  Computer(double allowableResidual, const Environment& env, const POSITION_WALL& posWall,
           const POSITION_WALL_PRE& posWallPre, int device = 0)
      : s_(new State{nullptr, env, posWall, posWallPre, {}, true}) {
    check(mps_create(&env.Native(), allowableResidual, device, &s_->ctx), "mps_create");
    if (!std::is_same<typename std::decay<POSITION_WALL>::type, StaticWalls>::value)
      check(mps_set_wall_callbacks(s_->ctx, &Computer::WallThunk, &Computer::PreThunk, s_.get()),
            "mps_set_wall_callbacks");
  }
  // Move-only, as the reference (Computer.hpp:1693-1697); the callbacks point
  // at the heap state, which does not move.
  Computer(Computer&&) noexcept = default;
  Computer(const Computer&) = delete;
  Computer& operator=(Computer&&) noexcept = delete;
  Computer& operator=(const Computer&) = delete;

  void ForwardTime(double dt) {
    check(mps_forward_time(s_->ctx, dt), "ForwardTime");
    s_->dirty = true;
  }
  void ForwardTime() {
    double dt = 0;
    check(mps_forward_time_cfl(s_->ctx, &dt), "ForwardTime");
    s_->dirty = true;
  }

  template <class PARTICLES>
  void AddParticles(PARTICLES&& src) {
    const std::size_t n = src.size();
    std::vector<double> x(n), z(n), u(n), w(n), p(n), nd(n);
    std::vector<uint8_t> t(n);
    std::size_t k = 0;
    for (const auto& q : src) {
      x[k] = q.X()[AXIS_X];
      z[k] = q.X()[AXIS_Z];
      u[k] = q.U()[AXIS_X];
      w[k] = q.U()[AXIS_Z];
      p[k] = q.P();
      nd[k] = q.N();
      t[k] = static_cast<uint8_t>(q.TYPE());
      k++;
    }
    check(mps_add_particles(s_->ctx, n, x.data(), z.data(), u.data(), w.data(), p.data(), nd.data(),
                            t.data()),
          "AddParticles");
    s_->dirty = true;
  }

  const std::vector<Particle>& Particles() const {
    if (s_->dirty) Refresh();
    return s_->mirror;
  }

  Environment GetEnvironment() const {
    double t = 0, dt = 0;
    mps_get_time(s_->ctx, &t, &dt);
    return Environment(s_->env.Native(), t, dt);
  }

  // Extensions: pressure solver selection and statistics of the last solve.
  void SetSolver(const mps_solver_opts& o) { check(mps_set_solver(s_->ctx, &o), "SetSolver"); }
  // Device-side wall kinematics in place of the positionWall callbacks
  // (mps_set_wall_motion): group has one entry per particle.
  void SetWallMotion(const std::vector<mps_wall_motion>& motions, const std::vector<uint8_t>& group) {
    check(mps_set_wall_motion(s_->ctx, motions.size(), motions.data(), group.data()), "SetWallMotion");
  }
  mps_solve_stats SolveStats() const {
    mps_solve_stats st{};
    mps_get_solve_stats(s_->ctx, &st);
    return st;
  }
  mps_ctx* Native() const { return s_->ctx; }

 private:
  struct State {
    mps_ctx* ctx;
    Environment env;
    POSITION_WALL posWall;
    POSITION_WALL_PRE posWallPre;
    std::vector<Particle> mirror;
    bool dirty;
    ~State() { mps_destroy(ctx); }
  };

  void check(int rc, const char* what) const {
    if (rc == MPS_OK) return;
    const std::string msg = s_->ctx ? mps_last_error(s_->ctx) : what;
    if (rc == MPS_ERR_CG_NOT_CONVERGED) throw Exception(msg);
    if (rc == MPS_ERR_CELL_OVERFLOW) throw Grid::Exception(msg);
    throw std::runtime_error(std::string(what) + ": " + msg);
  }
  void Refresh() const {
    const std::size_t n = mps_size(s_->ctx);
    std::vector<double> x(n), z(n), u(n), w(n), p(n), nd(n);
    std::vector<uint8_t> t(n);
    check(mps_get_particles(s_->ctx, x.data(), z.data(), u.data(), w.data(), p.data(), nd.data(),
                            t.data()),
          "Particles");
    s_->mirror.clear();
    s_->mirror.reserve(n);
    for (std::size_t k = 0; k < n; k++) {
      Particle q(static_cast<Particle::Type>(t[k]));
      q.X() = CreateVector(x[k], z[k]);
      q.U() = CreateVector(u[k], w[k]);
      q.P() = p[k];
      q.N() = nd[k];
      s_->mirror.push_back(q);
    }
    s_->dirty = false;
  }
  template <class PW>
  static void CallWall(PW& pw, std::size_t i, double t, double dt, double& x, double& z) {
    const auto v = pw(i, t, dt);
    x = v[AXIS_X];
    z = v[AXIS_Z];
  }
  static void CallWall(StaticWalls&, std::size_t, double, double, double&, double&) {}
  template <class PWP>
  static void CallPre(PWP& pwp, double t, double dt) { pwp(t, dt); }
  static void CallPre(StaticWalls&, double, double) {}
  static void WallThunk(void* user, double t, double dt, std::size_t count, const uint32_t* idx,
                        double* x, double* z) {
    auto* st = static_cast<State*>(user);
    for (std::size_t k = 0; k < count; k++) CallWall(st->posWall, idx[k], t, dt, x[k], z[k]);
  }
  static void PreThunk(void* user, double t, double dt) {
    auto* st = static_cast<State*>(user);
    CallPre(st->posWallPre, t, dt);
  }

  std::unique_ptr<State> s_;
};

// CreateComputer, Computer.hpp:1800-1815.
template <class PW, class PWP>
inline Computer<PW, PWP> CreateComputer(double allowableResidual, const Environment& env,
                                        const PW& posWall, const PWP& posWallPre) {
  return Computer<PW, PWP>(allowableResidual, env, posWall, posWallPre);
}

// OutputToCsv body, Main.cpp:31-67: header "Type, x, z, u, w, p, n", the
// stream's default 6 significant digits; returns the non-disabled count.
template <class COMPUTER>
inline std::size_t OutputToCsv(const COMPUTER& computer, std::ostream& out) {
  out << "Type, x, z, u, w, p, n" << std::endl;
  std::size_t count = 0;
  for (const auto& q : computer.Particles()) {
    out << static_cast<int>(q.TYPE()) << ", " << q.X()[0] << ", " << q.X()[1] << ", " << q.U()[0]
        << ", " << q.U()[1] << ", " << q.P() << ", " << q.N() << std::endl;
    if (q.TYPE() != Particle::Type::Disabled) count++;
  }
  return count;
}

}  // namespace OpenMps

#endif  // MPS_COMPUTER_HPP_
