// C++ host driver of the projection drop-in (p2d::PROJECTION_2D over the C ABI),
// used by tests/test_gpu_cpp_host.py.
//
//   projection_demo <dir> <n> <precision 0|1> <rho_minus> <rho_plus> <sigma>
//
// reads <dir>/phi.bin (n*n), u.bin ((n+1)*n), v.bin (n*(n+1)) as raw FP64,
// runs DetermineDivergence(dt = 1) and DeterminePressure() (checking that they
// leave u, v alone), then Solve(dt = 1) (checking the same pressure), and writes
// <dir>/p_out.bin (the reference's gauge, p(0,0) = 0), u_out.bin, v_out.bin and
// one line of stats to stdout.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "p2d_projection.hpp"

static std::vector<double> read_all(const std::string& path, size_t n) {
  std::vector<double> a(n);
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f || std::fread(a.data(), sizeof(double), n, f) != n) {
    std::fprintf(stderr, "cannot read %s\n", path.c_str());
    std::exit(2);
  }
  std::fclose(f);
  return a;
}

static void write_all(const std::string& path, const std::vector<double>& a) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(a.data(), sizeof(double), a.size(), f) != a.size()) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(2);
  }
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s dir n precision rho_minus rho_plus sigma\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  const int n = std::atoi(argv[2]);
  const size_t N = (size_t)n * n, NU = (size_t)(n + 1) * n;
  std::vector<double> phi = read_all(dir + "/phi.bin", N);
  std::vector<double> u = read_all(dir + "/u.bin", NU);
  std::vector<double> v = read_all(dir + "/v.bin", NU);
  try {
    p2d::PROJECTION_2D projection(n, n, 1.0 / n);
    projection.precision = std::atoi(argv[3]);
    projection.air_bubble_rising = true;  // minus side (phi <= 0) holds air_density
    projection.air_density = std::atof(argv[4]);
    projection.water_density = std::atof(argv[5]);
    projection.surface_tension = std::atof(argv[6]);
    projection.tolerance = 0;  // the reference's absolute rule off: the fixture's relative one instead
    projection.relative_tolerance = 1e-12;
    projection.div_tolerance = 1e-10;
    projection.max_iteration = 1000;
    // DeterminePressure first: it must leave u, v untouched and give Solve's pressure
    const std::vector<double> u0 = u, v0 = v;
    projection.DetermineDivergence(1.0, u.data(), v.data());
    projection.DeterminePressure(phi.data());
    const std::vector<double> p_only = projection.pressure_field;
    if (u != u0 || v != v0) {
      std::fprintf(stderr, "DeterminePressure changed the velocity\n");
      return 4;
    }
    projection.Solve(1.0, phi.data(), u.data(), v.data());
    if (p_only != projection.pressure_field) {
      std::fprintf(stderr, "DeterminePressure and Solve disagree on the pressure\n");
      return 5;
    }
    write_all(dir + "/p_out.bin", projection.pressure_field);
    write_all(dir + "/u_out.bin", u);
    write_all(dir + "/v_out.bin", v);
    const p2d_stats& s = projection.last_stats;
    std::printf("status %d iterations %d outer %d max_div %.3e max_u %.6f max_v %.6f\n", s.status, s.iterations,
                s.outer, s.max_div, projection.max_velocity_x, projection.max_velocity_y);
    return projection.converged() ? 0 : 1;
  } catch (const p2d::Error& e) {
    std::fprintf(stderr, "p2d error: %s\n", e.what());
    return 3;
  }
}
