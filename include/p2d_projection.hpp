/* p2d_projection.hpp -- header-only C++ host side of the projection drop-in.
 *
 * `p2d::PROJECTION_2D` keeps the names and meaning of the reference class's
 * projection interface (PROJECTION_2D.h:107-257) over plain arrays, and runs
 * every projection through the C ABI of p2d.h (HIP kernels for gfx950):
 *
 *   reference                                      here
 *   PROJECTION_2D(levelsets, u_mac_x, u_mac_y,...)  PROJECTION_2D(nx, ny, dx)  (fields are passed to Solve)
 *   InitializeFromBlock(block) :132-232             the public controls below, set by the caller
 *   tolerance, max_iteration (1e-4, 30)             the same, with the reference's absolute rule
 *                                                   ||r||_2 < tolerance (CG_METHOD.cpp:52)
 *   air_bubble_rising / water_drop :259-321         which density lies on the minus side (phi <= 0)
 *   water_density, air_density, surface_tension     the same
 *   Solve(dt) :235-242, Solve(dt, thread_id) :244-257  Solve(dt, phi, u, v)  (one C ABI call, p2d_project)
 *   DetermineDivergence(dt) :395-478                DetermineDivergence(dt, u, v) -> divergence_field
 *   DeterminePressure() :480-509                    DeterminePressure(phi): the solve for whatever
 *                                                   divergence_field holds (p2d_solve_pressure)
 *   pressure_field, max_velocity_x/y                the same (pressure_field in the reference's gauge,
 *                                                   p(0,0) = 0)
 *
 * Additional controls, off (0) by default: relative_tolerance (||r||_2 <= it * ||b||_2) and
 * div_tolerance (max|D u_proj| after the projection); every rule that is on must hold.
 *
 * Arrays are row-major, i fastest, without ghost cells: phi, divergence and
 * pressure nx*ny, u (nx+1)*ny, v nx*(ny+1); the reference's ghosted FIELD_STRUCTURE_2D
 * arrays copy in and out row by row, skipping ghost_width cells on each side
 * (ARRAY_2D::Index1D, ARRAY_2D.h:140-146).  Errors are thrown as
 * p2d::Error carrying the P2D_* status (the reference exits or stays silent); a
 * solve that reaches max_iteration returns, as the reference's does, with
 * converged() false.  Not thread-safe: call from one thread (inside
 * BEGIN_HEAD_THREAD_WORK in the reference's threaded Solve).
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
  // controls (PROJECTION_2D.h:132-232 reads them from the script; defaults are the reference's)
  double tolerance = 1e-4;        // stop when ||b - A p||_2 < tolerance (0: off)
  int max_iteration = 30;
  double relative_tolerance = 0;  // stop when ||b - A p||_2 <= relative_tolerance ||b||_2 (0: off)
  double div_tolerance = 0;       // stop when max|D u_proj| <= div_tolerance (0: off)
  bool air_bubble_rising = true;  // minus side (phi <= 0) holds air_density
  bool water_drop = false;        // minus side holds water_density
  double water_density = 1000.0, air_density = 1.0;
  double surface_tension = 0.0;   // jump sigma * kappa on the interface (delta formulation)
  int precision = P2D_FP64;       // P2D_FP64 or P2D_MIXED (FP32 inner solve in FP64 refinement)

  // fields
  std::vector<double> divergence_field;  // nx*ny, (D u)/dt: written by DetermineDivergence or the caller
  std::vector<double> pressure_field;    // nx*ny, p(0,0) = 0
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
    set_fluid(phi, mem);
    const int st = p2d_project(ctx_, u, v, dt, &last_stats, mem);
    if (st != P2D_OK && st != P2D_NOT_CONVERGED) throw Error(st, "p2d_project");
    max_velocity_x = last_stats.max_u;
    max_velocity_y = last_stats.max_v;
    pressure_field.resize((size_t)nx_ * ny_);
    check(p2d_get_pressure(ctx_, pressure_field.data(), 0, P2D_MEM_HOST), "p2d_get_pressure");
  }

  // DetermineDivergence(dt) (PROJECTION_2D.h:395-414): divergence_field = (D u) / dt from
  // host arrays (divergence_field is a host field, as the reference's member is).
  void DetermineDivergence(double dt, const double* u, const double* v) {
    ensure_context();
    divergence_field.resize((size_t)nx_ * ny_);
    check(p2d_divergence(ctx_, u, v, dt, divergence_field.data(), P2D_MEM_HOST), "p2d_divergence");
  }

  // DeterminePressure() (PROJECTION_2D.h:480-509, also called directly by the solver,
  // EULERIAN_FLUID_SOLVER_2D.h:1513, 1681): the pressure for whatever divergence_field holds,
  // with the jump of the level set phi; velocities are not involved.
  void DeterminePressure(const double* phi) {
    if (divergence_field.size() != (size_t)nx_ * ny_) throw Error(P2D_BAD_ARG, "divergence_field size");
    set_fluid(phi, P2D_MEM_HOST);
    pressure_field.resize((size_t)nx_ * ny_);
    const int st = p2d_solve_pressure(ctx_, divergence_field.data(), pressure_field.data(), &last_stats,
                                      P2D_MEM_HOST);
    if (st != P2D_OK && st != P2D_NOT_CONVERGED) throw Error(st, "p2d_solve_pressure");
  }

  // The two steps Solve takes before its update: DetermineDivergence(dt), DeterminePressure().
  void DeterminePressure(double dt, const double* phi, const double* u, const double* v) {
    DetermineDivergence(dt, u, v);
    DeterminePressure(phi);
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
  void set_fluid(const double* phi, int mem) {
    ensure_context();
    const double rho_minus = water_drop ? water_density : air_density;
    const double rho_plus = water_drop ? air_density : water_density;
    check(p2d_set_fluid(ctx_, phi, rho_minus, rho_plus, surface_tension, mem), "p2d_set_fluid");
  }
  // (Re)creates the context when a control it depends on changed.
  void ensure_context() {
    if (ctx_ && cfg_.precision == precision && cfg_.abs_tol == tolerance && cfg_.rtol == relative_tolerance &&
        cfg_.div_tol == div_tolerance && cfg_.max_iterations == max_iteration)
      return;
    if (ctx_) p2d_destroy(ctx_);
    ctx_ = nullptr;
    p2d_config_default(&cfg_, nx_, ny_, dx_);
    cfg_.precision = precision;
    cfg_.abs_tol = tolerance;
    cfg_.rtol = relative_tolerance;
    cfg_.div_tol = div_tolerance;
    cfg_.max_iterations = max_iteration;
    check(p2d_create(&cfg_, &ctx_), "p2d_create");
  }
};

}  // namespace p2d
