/* p2d_projection.hpp -- header-only C++ host side of the projection drop-in.
 *
 * `p2d::PROJECTION_2D` keeps the names and meaning of the reference class's
 * projection interface (PROJECTION_2D.h:107-257) over plain arrays, and runs
 * every projection through the C ABI of p2d.h (HIP kernels for gfx950):
 *
 *   reference                                      here
 *   PROJECTION_2D(levelsets, u_mac_x, u_mac_y,...)  PROJECTION_2D(nx, ny, dx)  (fields are passed to Solve)
 *   InitializeFromBlock(block) :132-232             the public controls below, set by the caller
 *   tolerance, max_iteration                        relative residual target, inner iteration cap
 *   air_bubble_rising / water_drop :259-321         which density lies on the minus side (phi <= 0)
 *   water_density, air_density, surface_tension     the same
 *   Solve(dt) :235-242, Solve(dt, thread_id) :244-257  Solve(dt, phi, u, v)
 *   DeterminePressure() :480-509 + the update      inside Solve (one C ABI call, p2d_project);
 *                                                   DeterminePressure alone on a copy of u, v
 *   pressure_field, max_velocity_x/y                the same (pressure_field is mean-zero)
 *
 * Arrays are row-major, i fastest, without ghost cells: phi and pressure
 * nx*ny, u (nx+1)*ny, v nx*(ny+1); the reference's ghosted FIELD_STRUCTURE_2D
 * arrays copy in and out row by row, skipping ghost_width cells on each side
 * (ARRAY_2D::Index1D, ARRAY_2D.h:140-146).  Errors are thrown as
 * p2d::Error carrying the P2D_* status (the reference exits or stays silent).
 * Not thread-safe: call from one thread (inside BEGIN_HEAD_THREAD_WORK in the
 * reference's threaded Solve).
 */
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "p2d.h"

namespace p2d {

struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string& what)
      : std::runtime_error(what + ": " + p2d_status_string(st)), status(st) {}
};

class PROJECTION_2D {
 public:
  // controls (PROJECTION_2D.h:132-232 reads them from the script)
  double tolerance = 1e-10;   // relative residual ||r||_2 / ||r_0||_2 of the pressure solve
  double div_tolerance = 1e-10;  // max |D u_proj| after the projection (0: off)
  int max_iteration = 1000;
  bool air_bubble_rising = true;  // minus side (phi <= 0) holds air_density
  bool water_drop = false;        // minus side holds water_density
  double water_density = 1000.0, air_density = 1.0;
  double surface_tension = 0.0;   // jump sigma * kappa on the interface (delta formulation)
  int precision = P2D_FP64;       // P2D_FP64 or P2D_MIXED (FP32 inner solve in FP64 refinement)

  // outputs
  std::vector<double> pressure_field;  // nx*ny, mean zero
  double max_velocity_x = 0.0, max_velocity_y = 0.0;
  p2d_stats last_stats{};

  PROJECTION_2D(int nx, int ny, double dx) : nx_(nx), ny_(ny), dx_(dx) {}
  ~PROJECTION_2D() {
    if (ctx_) p2d_destroy(ctx_);
  }
  PROJECTION_2D(const PROJECTION_2D&) = delete;
  PROJECTION_2D& operator=(const PROJECTION_2D&) = delete;

  // Solve(dt) (PROJECTION_2D.h:235-242): u, v projected in place from the level
  // set phi; mem = P2D_MEM_HOST or P2D_MEM_DEVICE (device pointers stay on the GPU).
  void Solve(double dt, const double* phi, double* u, double* v, int mem = P2D_MEM_HOST) {
    ensure_context();
    const double rho_minus = water_drop ? water_density : air_density;
    const double rho_plus = water_drop ? air_density : water_density;
    check(p2d_set_fluid(ctx_, phi, rho_minus, rho_plus, surface_tension, mem), "p2d_set_fluid");
    const int st = p2d_project(ctx_, u, v, dt, &last_stats, mem);
    if (st != P2D_OK && st != P2D_NOT_CONVERGED) throw Error(st, "p2d_project");
    max_velocity_x = last_stats.max_u;
    max_velocity_y = last_stats.max_v;
    pressure_field.resize((size_t)nx_ * ny_);
    check(p2d_get_pressure(ctx_, pressure_field.data(), 1, P2D_MEM_HOST), "p2d_get_pressure");
  }

  // DeterminePressure() (PROJECTION_2D.h:480-509, also called directly by the solver,
  // EULERIAN_FLUID_SOLVER_2D.h:1513, 1681): the pressure of the same solve, with u and v
  // left untouched (the projection runs on a copy).
  void DeterminePressure(double dt, const double* phi, const double* u, const double* v) {
    std::vector<double> uc(u, u + (size_t)(nx_ + 1) * ny_), vc(v, v + (size_t)nx_ * (ny_ + 1));
    Solve(dt, phi, uc.data(), vc.data(), P2D_MEM_HOST);
  }

  bool converged() const { return last_stats.status == P2D_OK; }
  p2d_ctx* context() { return ctx_; }

 private:
  int nx_, ny_;
  double dx_;
  p2d_ctx* ctx_ = nullptr;
  p2d_config cfg_{};

  static void check(int st, const char* what) {
    if (st != P2D_OK) throw Error(st, what);
  }
  // (Re)creates the context when a control it depends on changed.
  void ensure_context() {
    if (ctx_ && cfg_.precision == precision && cfg_.rtol == tolerance && cfg_.div_tol == div_tolerance &&
        cfg_.max_iterations == max_iteration)
      return;
    if (ctx_) p2d_destroy(ctx_);
    ctx_ = nullptr;
    p2d_config_default(&cfg_, nx_, ny_, dx_);
    cfg_.precision = precision;
    cfg_.rtol = tolerance;
    cfg_.div_tol = div_tolerance;
    cfg_.max_iterations = max_iteration;
    check(p2d_create(&cfg_, &ctx_), "p2d_create");
  }
};

}  // namespace p2d
