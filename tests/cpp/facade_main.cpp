// Drives the C++ facade (include/mps_computer.hpp) the way the reference's
// Main.cpp:288-400 drives OpenMps::Computer: environment, particles from a CSV
// body, static walls through the positionWall / positionWallPre callbacks,
// ForwardTime loop, CSV output.  Usage: facade_main <in.csv> <out.csv> <steps>
// <dt> <eps> <maxDt> <courant> <g> <rho> <nu> <r_eByl_0> <l_0> <minX> <minZ>
// <maxX> <maxZ> [static]
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "mps_computer.hpp"

int main(int argc, char** argv) {
  if (argc < 17) {
    std::fprintf(stderr, "usage: facade_main in.csv out.csv steps dt eps maxDt courant g rho nu "
                         "r_eByl_0 l_0 minX minZ maxX maxZ [static]\n");
    return 2;
  }
  const int steps = std::atoi(argv[3]);
  const double dt = std::atof(argv[4]), eps = std::atof(argv[5]);
  double a[11];
  for (int k = 0; k < 11; k++) a[k] = std::atof(argv[6 + k]);
  const bool static_walls = argc > 17 && std::string(argv[17]) == "static";
  OpenMps::Environment env(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10]);

  std::vector<OpenMps::Particle> particles;
  std::ifstream in(argv[1]);
  std::string line;
  std::getline(in, line);  // header "Type, x, z, u, w, p, n"
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    for (char& ch : line)
      if (ch == ',') ch = ' ';
    std::istringstream ss(line);
    int t;
    double x, z, u, w, p, n;
    ss >> t >> x >> z >> u >> w >> p >> n;
    OpenMps::Particle q(static_cast<OpenMps::Particle::Type>(t));
    q.X() = OpenMps::CreateVector(x, z);
    q.U() = OpenMps::CreateVector(u, w);
    q.P() = p;
    q.N() = n;
    particles.push_back(q);
  }
  std::vector<OpenMps::Vector> initial;
  for (const auto& q : particles) initial.push_back(q.X());
  int pre_calls = 0;
  try {
    if (static_walls) {
      auto computer = OpenMps::CreateComputer(eps, env, OpenMps::StaticWalls{}, OpenMps::StaticWalls{});
      computer.AddParticles(particles);
      for (int s = 0; s < steps; s++) computer.ForwardTime(dt);
      std::ofstream out(argv[2]);
      out << std::setprecision(17);
      OpenMps::OutputToCsv(computer, out);
    } else {
      const auto positionWall = [&initial](std::size_t i, double, double) { return initial[i]; };
      const auto positionWallPre = [&pre_calls](double, double) { pre_calls++; };
      auto computer = OpenMps::CreateComputer(eps, env, positionWall, positionWallPre);
      computer.AddParticles(particles);
      for (int s = 0; s < steps; s++) computer.ForwardTime(dt);
      std::ofstream out(argv[2]);
      out << std::setprecision(17);
      OpenMps::OutputToCsv(computer, out);
      std::printf("positionWallPre calls: %d\n", pre_calls);
    }
  } catch (decltype(OpenMps::CreateComputer(eps, env, OpenMps::StaticWalls{}, OpenMps::StaticWalls{}))::Exception& e) {
    std::printf("!!!!ERROR!!!!\n%s\n", e.what());
    return 1;
  }
  std::printf("done %d steps\n", steps);
  return 0;
}
