// Semi-implicit viscosity solve of one MAC velocity component, FP64, matrix-free:
//
//   POISSON_SOLVER::SolveForViscosity (POISSON_SOLVER.cpp:175-184) builds, per node
//   (i, j) of the component grid (BuildLinearSystemNodeForSemiImplicitViscosity,
//   POISSON_SOLVER.cpp:620-712),
//     (1 + (c1 + c2 + c3 + c4)/dx^2 + c5) x_ij - c2/dx^2 x_W - c1/dx^2 x_E - c4/dy^2 x_S - c3/dy^2 x_N
//        = explicit_ij
//   (the diagonal scales all four coefficients by 1/dx^2, the off-diagonals the y pair by
//   1/dy^2), and solves it with CG.  The boundary-condition field's one-wide ghost ring is
//   never written (SetupBoundaryConditionsForVelocity, VISCOSITY_2D.h:647-663, visits only
//   the grid's own nodes), so it keeps 0 = BC_FULL and a ghost neighbour is read as unknown
//   0 once the interior is renumbered: the coupling goes to the first node.  VISC_LITERAL
//   reproduces that; VISC_DIRICHLET takes the ghost neighbours as zero velocity (the
//   intent of SetupBoundaryConditionsForVelocity: bc = BC_DIR, velocity = 0).
//
// The system is not symmetric (the rows carry 1/(Re rho r)), so the GPU solves it with
// Jacobi-preconditioned BiCGStab (van der Vorst 1992) and stops on the reference's rule
// ||r||_2 < tolerance.  Scalars stay on the device (VScal); every kernel returns at once
// after convergence, so the host can enqueue iterations in chunks.
//
// The oil-water coefficients (VISCOSITY_2D::ApplyViscosity, oil_water_simulation,
// semi_implicit_approach, dimensionless_form, VISCOSITY_2D.h:546-640) are computed on the
// device from the level set: k_ow_coef_x / k_ow_coef_y.
#pragma once

#include "p2d_common.hpp"

namespace p2d {

constexpr int VISC_LITERAL = 0;
constexpr int VISC_DIRICHLET = 1;

struct VScal {
  double rho, alpha, omega, beta, r2, tol2, b2, rinf;
  int iter, done, max_iter, flag;  // flag: 1 = breakdown (rho, omega or (r^, v) zero, or not finite)
};

// VISCOSITY_2D::HeavisideFunction (VISCOSITY_2D.h:667-686), PI = 3.141592.
__device__ __forceinline__ double heaviside_ref(double phi, double eps) {
#pragma clang fp contract(off)
  const double pi = 3.141592;
  const double oe = 1.0 / eps, op = 1.0 / pi;
  if (phi < -eps) return 0.0;
  if (phi > eps) return 1.0;
  return 0.5 + phi * oe * 0.5 + 0.5 * op * sin(pi * phi * oe);
}
// Dimensionless oil/water property: oil/water + (1 - oil/water) H (VISCOSITY_2D.h:776-781, 817-822).
__device__ __forceinline__ double dimless(double phi, double eps, double oil, double water) {
#pragma clang fp contract(off)
  return oil / water + (1.0 - oil / water) * heaviside_ref(phi, eps);
}

struct OilWater {
  double dx, dy, x_min, dt, Re, rho_oil, rho_water, mu_oil, mu_water, eps;
};

// x-face coefficients and explicit term (VISCOSITY_2D.h:553-591); nodes (i, j), i <= nx.
// The x-face grid starts at x_min - dx/2; v ghosts are zero; the level set's ghosts are
// linear extrapolations (phi_at).
__global__ void __launch_bounds__(BLOCK) k_ow_coef_x(int nx, int ny, OilWater o, const double* __restrict__ phi,
                                                     const double* __restrict__ u, const double* __restrict__ v,
                                                     double* c1, double* c2, double* c3, double* c4, double* c5,
                                                     double* ex) {
#pragma clang fp contract(off)
  const int n = (nx + 1) * ny;
  for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) {
    const int j = k / (nx + 1), i = k - j * (nx + 1);
    const double pw = phi_at(phi, nx, ny, i - 1, j), pe = phi_at(phi, nx, ny, i, j);
    const double pwn = phi_at(phi, nx, ny, i - 1, j + 1), pen = phi_at(phi, nx, ny, i, j + 1);
    const double pws = phi_at(phi, nx, ny, i - 1, j - 1), pes = phi_at(phi, nx, ny, i, j - 1);
    const double half = 0.5 * (pw + pe);
    const double cu = 0.25 * (pw + pe + pwn + pen);
    const double cd = 0.25 * (pw + pe + pws + pes);
    const double rho = dimless(half, o.eps, o.rho_oil, o.rho_water);
    const double mu_h = dimless(half, o.eps, o.mu_oil, o.mu_water);
    const double mu_cu = dimless(cu, o.eps, o.mu_oil, o.mu_water), mu_cd = dimless(cd, o.eps, o.mu_oil, o.mu_water);
    const double mu_r = dimless(pe, o.eps, o.mu_oil, o.mu_water), mu_c = dimless(pw, o.eps, o.mu_oil, o.mu_water);
    const double gx = o.x_min - 0.5 * o.dx;
    const double one_over_r = 1.0 / (gx + (i + 0.5) * o.dx);
    const double one_over_r2 = one_over_r * one_over_r;
    const double rre = o.Re * rho;
    c1[k] = one_over_r * (gx + (i + 1) * o.dx) * 2.0 * o.dt / rre * mu_r;
    c2[k] = one_over_r * (gx + i * o.dx) * 2.0 * o.dt / rre * mu_c;
    c3[k] = o.dt / rre * mu_cu;
    c4[k] = o.dt / rre * mu_cd;
    c5[k] = 2.0 * o.dt / rre * mu_h * one_over_r2;
    auto V = [&](int ii, int jj) { return (ii < 0 || ii >= nx || jj < 0 || jj > ny) ? 0.0 : v[(size_t)jj * nx + ii]; };
    const double odx = 1.0 / o.dx, ody = 1.0 / o.dy;
    ex[k] = u[k] + o.dt / rre *
                       (ody * (mu_cu * odx * (V(i, j + 1) - V(i - 1, j + 1)) - mu_cd * odx * (V(i, j) - V(i - 1, j))));
  }
}

// y-face coefficients and explicit term (VISCOSITY_2D.h:593-640); nodes (i, j), j <= ny.
// The y-face grid starts at x_min; the explicit term reads the x velocity before (uo) and
// after (un) the x solve; u ghosts are zero.
__global__ void __launch_bounds__(BLOCK) k_ow_coef_y(int nx, int ny, OilWater o, const double* __restrict__ phi,
                                                     const double* __restrict__ uo, const double* __restrict__ un,
                                                     const double* __restrict__ v, double* c1, double* c2, double* c3,
                                                     double* c4, double* c5, double* ex) {
#pragma clang fp contract(off)
  const int n = nx * (ny + 1);
  for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) {
    const int j = k / nx, i = k - j * nx;
    const double ps = phi_at(phi, nx, ny, i, j - 1), pn = phi_at(phi, nx, ny, i, j);
    const double pse = phi_at(phi, nx, ny, i + 1, j - 1), pne = phi_at(phi, nx, ny, i + 1, j);
    const double psw = phi_at(phi, nx, ny, i - 1, j - 1), pnw = phi_at(phi, nx, ny, i - 1, j);
    const double half = 0.5 * (ps + pn);
    const double cu = 0.25 * (ps + pse + pn + pne);
    const double cl = 0.25 * (psw + ps + pnw + pn);
    const double rho = dimless(half, o.eps, o.rho_oil, o.rho_water);
    const double mu_h = dimless(half, o.eps, o.mu_oil, o.mu_water);
    const double mu_cu = dimless(cu, o.eps, o.mu_oil, o.mu_water), mu_cl = dimless(cl, o.eps, o.mu_oil, o.mu_water);
    const double mu_u = dimless(pn, o.eps, o.mu_oil, o.mu_water), mu_c = dimless(ps, o.eps, o.mu_oil, o.mu_water);
    const double r = o.x_min + i * o.dx;
    const double one_over_r = 1.0 / r;
    const double rre = o.Re * rho;
    c1[k] = one_over_r * o.dt / rre * mu_cu * (r + 0.5 * o.dx);
    c2[k] = one_over_r * o.dt / rre * mu_cl * (r - 0.5 * o.dx);
    c3[k] = 2.0 * o.dt / rre * mu_u;
    c4[k] = 2.0 * o.dt / rre * mu_c;
    c5[k] = 0.0;
    auto U = [&](const double* a, int ii, int jj) {
      return (ii < 0 || ii > nx || jj < 0 || jj >= ny) ? 0.0 : a[(size_t)jj * (nx + 1) + ii];
    };
    const double odx = 1.0 / o.dx, ody = 1.0 / o.dy;
    ex[k] = v[k] + o.dt / rre *
                       ((odx * (mu_cu * ody * (U(uo, i + 1, j) - U(uo, i + 1, j - 1)) -
                                mu_cl * ody * (U(uo, i, j) - U(uo, i, j - 1)))) +
                        one_over_r * mu_h * ody *
                            (0.5 * (U(un, i, j) + U(un, i + 1, j)) - 0.5 * (U(un, i, j - 1) + U(un, i + 1, j - 1))));
  }
}

// The row of node (i, j): diagonal and the four off-diagonal magnitudes
// (A = dg x - cw x_W - ce x_E - cs x_S - cn x_N), the diagonal's sum in the reference's
// order (full neighbours W, E, S, N, then Dirichlet ones W, E, S, N).
__global__ void __launch_bounds__(BLOCK) k_visc_rows(int nxc, int nyc, double inv_dxdx, double inv_dydy, int mode,
                                                     const double* __restrict__ c1, const double* __restrict__ c2,
                                                     const double* __restrict__ c3, const double* __restrict__ c4,
                                                     const double* __restrict__ c5, double* dg, double* cw,
                                                     double* ce, double* cs, double* cn) {
#pragma clang fp contract(off)
  const int n = nxc * nyc;
  for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) {
    const int j = k / nxc, i = k - j * nxc;
    const bool in_w = i > 0, in_e = i < nxc - 1, in_s = j > 0, in_n = j < nyc - 1;
    const double a2 = c2[k], a1 = c1[k], a4 = c4[k], a3 = c3[k];
    double coef = 0.0;
    if (mode == VISC_LITERAL) {
      coef += a2;
      coef += a1;
      coef += a4;
      coef += a3;
    } else {
      if (in_w) coef += a2;
      if (in_e) coef += a1;
      if (in_s) coef += a4;
      if (in_n) coef += a3;
      if (!in_w) coef += a2;
      if (!in_e) coef += a1;
      if (!in_s) coef += a4;
      if (!in_n) coef += a3;
    }
    if (coef == 0.0) coef = 1.0;
    dg[k] = 1.0 + inv_dxdx * coef + c5[k];
    cw[k] = a2 * inv_dxdx;
    ce[k] = a1 * inv_dxdx;
    cs[k] = a4 * inv_dydy;
    cn[k] = a3 * inv_dydy;
  }
}

// (A x)_k with the ghost neighbours per mode (literal: x[0]; Dirichlet: 0).
__device__ __forceinline__ double visc_apply_at(int k, int i, int j, int nxc, int nyc, int mode,
                                                const double* __restrict__ x, const double* __restrict__ dg,
                                                const double* __restrict__ cw, const double* __restrict__ ce,
                                                const double* __restrict__ cs, const double* __restrict__ cn) {
  const double g = mode == VISC_LITERAL ? x[0] : 0.0;
  const double xw = i > 0 ? x[k - 1] : g;
  const double xe = i < nxc - 1 ? x[k + 1] : g;
  const double xs = j > 0 ? x[k - nxc] : g;
  const double xn = j < nyc - 1 ? x[k + nxc] : g;
  return dg[k] * x[k] - cw[k] * xw - ce[k] * xe - cs[k] * xs - cn[k] * xn;
}

struct VStencil {
  const double *dg, *cw, *ce, *cs, *cn;
  int nxc, nyc, mode;
};

// r = b - A x, rhat = r, p = v = 0; partials (r.r, b.b).
__global__ void __launch_bounds__(BLOCK) k_bs_init(VStencil A, const double* __restrict__ b, const double* __restrict__ x,
                                                   double* r, double* rhat, double* p, double* v, double* partial) {
  const int n = A.nxc * A.nyc;
  double acc[2] = {0.0, 0.0};
  for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) {
    const int j = k / A.nxc, i = k - j * A.nxc;
    const double rk = b[k] - visc_apply_at(k, i, j, A.nxc, A.nyc, A.mode, x, A.dg, A.cw, A.ce, A.cs, A.cn);
    r[k] = rk;
    rhat[k] = rk;
    p[k] = 0.0;
    v[k] = 0.0;
    acc[0] += rk * rk;
    acc[1] += b[k] * b[k];
  }
  block_partial<2>(acc, 0u, partial);
}
__global__ void __launch_bounds__(FINISH_THREADS) k_bs_fin_init(const double* __restrict__ partial, int G, VScal* s) {
  double a[2];
  finish_reduce<2>(a, partial, G, 0u);
  if (threadIdx.x == 0) {
    s->r2 = a[0];
    s->b2 = a[1];
    s->rho = a[0];  // (rhat, r) with rhat = r
    s->alpha = 1.0;
    s->omega = 1.0;
    s->beta = 0.0;
    s->iter = 0;
    s->flag = 0;
    s->done = !(a[0] >= s->tol2) ? 1 : 0;  // converged already (or NaN: flagged below)
    if (!isfinite(a[0])) {
      s->flag = 1;
      s->done = 1;
    }
  }
}

// p = r + beta (p - omega v); y = p / dg.
__global__ void __launch_bounds__(BLOCK) k_bs_dir(int n, const double* __restrict__ r, double* p,
                                                  const double* __restrict__ v, const double* __restrict__ dg,
                                                  double* y, const VScal* s) {
  if (s->done) return;
  const double beta = s->beta, omega = s->omega;
  for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) {
    const double pk = r[k] + beta * (p[k] - omega * v[k]);
    p[k] = pk;
    y[k] = pk / dg[k];
  }
}
// out = A in; partial (rhat . out)  [v = A y]  or (out . s, out . out)  [t = A z].
template <bool TS>
__global__ void __launch_bounds__(BLOCK) k_bs_apply(VStencil A, const double* __restrict__ in, double* out,
                                                    const double* __restrict__ w, double* partial, const VScal* s) {
  if (s->done) return;
  const int n = A.nxc * A.nyc;
  double acc[2] = {0.0, 0.0};
  for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) {
    const int j = k / A.nxc, i = k - j * A.nxc;
    const double o = visc_apply_at(k, i, j, A.nxc, A.nyc, A.mode, in, A.dg, A.cw, A.ce, A.cs, A.cn);
    out[k] = o;
    acc[0] += o * w[k];
    if (TS) acc[1] += o * o;
  }
  block_partial<2>(acc, 0u, partial);
}
// alpha = rho / (rhat . v)
__global__ void __launch_bounds__(FINISH_THREADS) k_bs_fin_alpha(const double* __restrict__ partial, int G, VScal* s) {
  if (s->done) return;
  double a[2];
  finish_reduce<2>(a, partial, G, 0u);
  if (threadIdx.x == 0) {
    if (a[0] == 0.0 || !isfinite(a[0])) {
      s->flag = 1;
      s->done = 1;
      return;
    }
    s->alpha = s->rho / a[0];
  }
}
// s = r - alpha v; z = s / dg.
__global__ void __launch_bounds__(BLOCK) k_bs_half(int n, const double* __restrict__ r, const double* __restrict__ v,
                                                   const double* __restrict__ dg, double* sv, double* z,
                                                   const VScal* s) {
  if (s->done) return;
  const double alpha = s->alpha;
  for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) {
    const double sk = r[k] - alpha * v[k];
    sv[k] = sk;
    z[k] = sk / dg[k];
  }
}
// omega = (t . s) / (t . t); a zero t means s = 0: the half step solved the system.
__global__ void __launch_bounds__(FINISH_THREADS) k_bs_fin_omega(const double* __restrict__ partial, int G, VScal* s) {
  if (s->done) return;
  double a[2];
  finish_reduce<2>(a, partial, G, 0u);
  if (threadIdx.x == 0) s->omega = a[1] > 0.0 ? a[0] / a[1] : 0.0;
}
// x += alpha y + omega z; r = s - omega t; partials (r.r, rhat.r, max|r|).
__global__ void __launch_bounds__(BLOCK) k_bs_update(int n, double* x, const double* __restrict__ y,
                                                     const double* __restrict__ z, const double* __restrict__ sv,
                                                     const double* __restrict__ t, double* r,
                                                     const double* __restrict__ rhat, double* partial,
                                                     const VScal* s) {
  if (s->done) return;
  const double alpha = s->alpha, omega = s->omega;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) {
    x[k] += alpha * y[k] + omega * z[k];
    const double rk = sv[k] - omega * t[k];
    r[k] = rk;
    acc[0] += rk * rk;
    acc[1] += rhat[k] * rk;
    acc[2] = fmax(acc[2], fabs(rk));
  }
  block_partial<3>(acc, 4u, partial);
}
// r2, rho, beta; stop on ||r||_2 < tolerance, on the iteration limit or on a breakdown.
__global__ void __launch_bounds__(FINISH_THREADS) k_bs_fin_update(const double* __restrict__ partial, int G,
                                                                  VScal* s) {
  if (s->done) return;
  double a[3];
  finish_reduce<3>(a, partial, G, 4u);
  if (threadIdx.x == 0) {
    const double rho_old = s->rho;
    s->r2 = a[0];
    s->rinf = a[2];
    s->rho = a[1];
    s->iter += 1;
    if (!isfinite(a[0])) {
      s->flag = 1;
      s->done = 1;
      return;
    }
    if (a[0] < s->tol2 || s->iter >= s->max_iter) {
      s->done = 1;
      return;
    }
    if (a[1] == 0.0 || s->omega == 0.0) {  // restart-free breakdown: report it
      s->flag = 1;
      s->done = 1;
      return;
    }
    s->beta = (a[1] / rho_old) * (s->alpha / s->omega);
  }
}

}  // namespace p2d
