// p2d level-set kernels: the Sussman re-initialisation (WENO5 one-sided
// derivatives + Godunov Hamiltonian), the WENO5 level-set and ENO3 velocity
// advection stages, the explicit air-water viscosity update and unit normals.  FP64 with
// contraction off: the arithmetic keeps the reference's operation order
// (ADVECTION_METHOD_2D.h, LEVELSET_2D.h), so the results are reproducible
// operation by operation on the CPU.
#pragma once
#include "p2d_common.hpp"

namespace p2d {

constexpr int RI_W = 64, RI_H = 16, RI_G = 3;  // re-initialisation tile: 64 x 16 cells, 3-cell ghost ring

// Sign function of the re-initialisation, s = phi0 / sqrt(phi0^2 + dx^2)
// (EULERIAN_FLUID_SOLVER_2D.h:2380-2383), from the level set at entry.
__global__ void __launch_bounds__(BLOCK) k_reinit_sign(size_t N, double h, const double* __restrict__ phi,
                                                       double* __restrict__ s) {
#pragma clang fp contract(off)
  for (size_t k = blockIdx.x * (size_t)BLOCK + threadIdx.x; k < N; k += (size_t)gridDim.x * BLOCK)
    s[k] = phi[k] / sqrt(phi[k] * phi[k] + h * h);
}

// WENO5 derivative from the five one-sided differences v1..v5: smoothness
// indicators, nonlinear weights with epsilon, the three candidate stencils
// (ADVECTION_METHOD_2D.h:844-929, in its expression forms).
// x / y given z = RN(1/y), correctly rounded without the general division sequence:
// q = x z, then one FMA correction q + (x - q y) z (Markstein 1990: with z the correctly
// rounded reciprocal and q within one ulp of x / y, the remainder x - q y is exact and the
// corrected sum rounds to x / y).  Valid for |x / y| >= 2^-1000, where the product cannot
// lose bits to underflow; the WENO inputs are scaled differences of level-set values, zero
// or far above that.  The remainder is formed as -(q y - x): it is the same number (round
// to nearest is symmetric) except that an exact zero comes out as -0, so the correction
// adds -0 to q and an exact quotient keeps q with the sign of its zero.  This equals the
// form that tests the remainder for zero and selects q, bit for bit, on every finite x
// (140 M random and edge-case inputs on the CPU), without the compare and select (~4 VALU
// instructions per quotient, ~40 quotients per re-initialised cell).  An infinite x gives
// NaN instead of an infinity; the level set is finite.
__device__ __forceinline__ double div_known(double x, double y, double z) {
#pragma clang fp contract(off)
  const double q = x * z;
  const double r = -fma(q, y, -x);
  return fma(r, z, q);
}
// a / b correctly rounded for positive normal a, b whose quotient and 1 / b are normal and
// whose exponents differ by less than 768: the reciprocal estimate, two Newton steps
// r += r (1 - b r) and one Markstein correction q + (a - b q) r.  These are the steps of the
// general FP64 division sequence, whose operand scaling (v_div_scale), scaled FMA
// (v_div_fmas) and special-case fixup (v_div_fixup) leave every value unchanged in that range,
// so the result is the same correctly rounded quotient, three instructions shorter.  The
// WENO weights use it: numerators 0.1, 0.3, 0.6 and 1, denominators (eps + s)^2 >= 1e-12
// and their sums, far inside the range for any finite level set.
__device__ __forceinline__ double div_pos(double a, double b) {
#pragma clang fp contract(off)
  const double r0 = __builtin_amdgcn_rcp(b);
  const double e0 = fma(-b, r0, 1.0);
  const double r1 = fma(r0, e0, r0);
  const double e1 = fma(-b, r1, 1.0);
  const double r2 = fma(r1, e1, r1);
  const double q = a * r2;
  const double rem = fma(-b, q, a);
  return fma(rem, r2, q);
}
__device__ __forceinline__ double div3(double x) { return div_known(x, 3.0, 1.0 / 3.0); }
__device__ __forceinline__ double div6(double x) { return div_known(x, 6.0, 1.0 / 6.0); }

__device__ __forceinline__ double weno5_d(double v1, double v2, double v3, double v4, double v5, double eps) {
#pragma clang fp contract(off)
  const double s1 = (double)13 / 12 * (v1 - 2 * v2 + v3) * (v1 - 2 * v2 + v3) +
                    (double)1 / 4 * (v1 - 4 * v2 + 3 * v3) * (v1 - 4 * v2 + 3 * v3);
  const double s2 = (double)13 / 12 * (v2 - 2 * v3 + v4) * (v2 - 2 * v3 + v4) + (double)1 / 4 * (v2 - v4) * (v2 - v4);
  const double s3 = (double)13 / 12 * (v3 - 2 * v4 + v5) * (v3 - 2 * v4 + v5) +
                    (double)1 / 4 * (3 * v3 - 4 * v4 + v5) * (3 * v3 - 4 * v4 + v5);
  // (T)1/10*(T)1/x is (0.1 * 1) / x: one rounded quotient of the constant
  const double a1 = div_pos((double)1 / 10 * (double)1, (eps + s1) * (eps + s1));
  const double a2 = div_pos((double)6 / 10 * (double)1, (eps + s2) * (eps + s2));
  const double a3 = div_pos((double)3 / 10 * (double)1, (eps + s3) * (eps + s3));
  const double sum = a1 + a2 + a3;
  // the three quotients by sum share one correctly rounded reciprocal (Markstein's
  // correction holds for any divisor whose reciprocal is correctly rounded)
  const double zs = div_pos(1.0, sum);
  const double w1 = div_known(a1, sum, zs), w2 = div_known(a2, sum, zs), w3 = div_known(a3, sum, zs);
  // the reference's v*(T)7/6 is (v * 7) / 6, a rounded division (div3, div6 give the same
  // bits); -(T)1/6 is a constant
  return w1 * (div3(v1 * (double)1) - div6(v2 * (double)7) + div6(v3 * (double)11)) +
         w2 * (v2 * (-(double)1 / 6) + div6(v3 * (double)5) + div3(v4 * (double)1)) +
         w3 * (div3(v3 * (double)1) + div6(v4 * (double)5) - div6(v5 * (double)1));
}

// |grad phi| by the reference's selection among the one-sided derivatives
// (ADVECTION_METHOD_2D.h:936-1060): upwind by the sign of s times each
// derivative, the larger magnitude at a shock, none at an expansion.  NaN when
// no case applies (NaN derivatives): the cell is then left unchanged.
// Each direction contributes the square of its chosen derivative (+0 at an
// expansion) and one square root takes their sum: the reference's cases take the
// root of the same sums, since adding +0 to a square changes no bit, and the
// lanes of a wavefront no longer diverge into separate square roots.
__device__ __forceinline__ double godunov_part(double s, double sm, double sp, double m, double p) {
#pragma clang fp contract(off)
  if ((sp >= 0) && (sm >= 0)) return m * m;
  if ((sp <= 0) && (sm <= 0)) return p * p;
  if ((sp > 0) && (sm < 0)) return 0.0;
  if ((sp < 0) && (sm > 0)) {
    const double ss = s * (fabs(p) - fabs(m)) / (p - m);
    return ss > 0 ? m * m : p * p;
  }
  return __longlong_as_double(0x7ff8000000000000LL);
}
__device__ __forceinline__ double godunov_h(double s, double mx, double px, double my, double py) {
#pragma clang fp contract(off)
  const double x2 = godunov_part(s, s * mx, s * px, mx, px);
  const double y2 = godunov_part(s, s * my, s * py, my, py);
  return sqrt(x2 + y2);
}

// Stages a 64 x 16 tile of phi plus a 3-cell ring in LDS with the ghost rule
// of FIELD_STRUCTURE_2D::FillGhostCellsFrom (FIELD_STRUCTURE_2D.h:528-582):
// every ghost of a face row takes the second interior value; corners, never
// read by the dimension-by-dimension stencils, are 0.
__device__ __forceinline__ void stage_weno_tile(double (*g)[RI_W + 2 * RI_G], const double* __restrict__ phi, int nx,
                                                int ny, int i0, int j0) {
  for (int t = threadIdx.x; t < (RI_H + 2 * RI_G) * (RI_W + 2 * RI_G); t += BLOCK) {
    const int a = t % (RI_W + 2 * RI_G), b = t / (RI_W + 2 * RI_G);
    const int i = i0 - RI_G + a, j = j0 - RI_G + b;
    const bool xin = i >= 0 && i < nx, yin = j >= 0 && j < ny;
    double v = 0.0;
    if (xin && yin) v = phi[ci(i, j, nx)];
    else if (yin) v = phi[ci(i < 0 ? 1 : nx - 2, j, nx)];
    else if (xin) v = phi[ci(i, j < 0 ? 1 : ny - 2, nx)];
    g[b][a] = v;
  }
}

// One re-initialisation step, ADVECTION_METHOD_2D::WENO5thReinitialization
// (ADVECTION_METHOD_2D.h:805-1076): out = phi - dt s (H - 1), a sign flip
// clamped to +-eps.  The tile is staged by stage_weno_tile.
// Grid dim3(ceil(nx/64), ceil(ny/16)), BLOCK threads.
__global__ void __launch_bounds__(BLOCK) k_reinit_step(int nx, int ny, double h, double dt, double eps,
                                                       const double* __restrict__ phi, const double* __restrict__ sg,
                                                       double* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double g[RI_H + 2 * RI_G][RI_W + 2 * RI_G];
  int bx, by;
  xcd_tile(bx, by);
  const int i0 = bx * RI_W, j0 = by * RI_H;
  stage_weno_tile(g, phi, nx, ny, i0, j0);
  __syncthreads();
  const double ih = 1.0 / h;
  const int c = threadIdx.x & 63, w = wave_id();
#pragma unroll
  for (int m = 0; m < RI_H / (BLOCK / 64); ++m) {
    const int r = w + (BLOCK / 64) * m, i = i0 + c, j = j0 + r;
    if (i >= nx || j >= ny) continue;
    const int a = c + RI_G, b = r + RI_G;
    auto X = [&](int d) { return g[b][a + d]; };
    auto Y = [&](int d) { return g[b + d][a]; };
    const double mx = weno5_d(ih * (X(-2) - X(-3)), ih * (X(-1) - X(-2)), ih * (X(0) - X(-1)), ih * (X(1) - X(0)),
                              ih * (X(2) - X(1)), eps);
    const double px = weno5_d(ih * (X(3) - X(2)), ih * (X(2) - X(1)), ih * (X(1) - X(0)), ih * (X(0) - X(-1)),
                              ih * (X(-1) - X(-2)), eps);
    const double my = weno5_d(ih * (Y(-2) - Y(-3)), ih * (Y(-1) - Y(-2)), ih * (Y(0) - Y(-1)), ih * (Y(1) - Y(0)),
                              ih * (Y(2) - Y(1)), eps);
    const double py = weno5_d(ih * (Y(3) - Y(2)), ih * (Y(2) - Y(1)), ih * (Y(1) - Y(0)), ih * (Y(0) - Y(-1)),
                              ih * (Y(-1) - Y(-2)), eps);
    const size_t k = ci(i, j, nx);
    const double s = sg[k];
    const double H = godunov_h(s, mx, px, my, py);
    double v = g[b][a];
    if (H == H) v = v - dt * s * (H - 1);
    if (s * v < 0) {
      if (s < 0) v = -eps;
      if (s > 0) v = eps;
    }
    out[k] = v;
  }
}

// One forward-Euler stage of the level-set advection, ADVECTION_2D::Solve_Levelset
// with use_mac_grid and use_5th_weno (ADVECTION_2D.cpp:165-200) =
// ADVECTION_METHOD_2D::WENO5th on a scalar field (ADVECTION_METHOD_2D.h:508-618):
// the cell velocity is the mean of its two faces, each component picks the
// upwind one-sided WENO5 derivative (u > 0: phi_x^-), out = phi - dt (u phi_x
// + v phi_y).  Only the two derivatives used are formed.  u: (nx+1) x ny
// x-faces, v: nx x (ny+1) y-faces.  Grid dim3(ceil(nx/64), ceil(ny/16)).
__global__ void __launch_bounds__(BLOCK) k_advect_ls(int nx, int ny, double h, double dt, double eps,
                                                     const double* __restrict__ phi, const double* __restrict__ u,
                                                     const double* __restrict__ v, double* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double g[RI_H + 2 * RI_G][RI_W + 2 * RI_G];
  int bx, by;
  xcd_tile(bx, by);
  const int i0 = bx * RI_W, j0 = by * RI_H;
  stage_weno_tile(g, phi, nx, ny, i0, j0);
  __syncthreads();
  const double ih = 1.0 / h;
  const int c = threadIdx.x & 63, w = wave_id();
#pragma unroll
  for (int m = 0; m < RI_H / (BLOCK / 64); ++m) {
    const int r = w + (BLOCK / 64) * m, i = i0 + c, j = j0 + r;
    if (i >= nx || j >= ny) continue;
    const int a = c + RI_G, b = r + RI_G;
    auto X = [&](int d) { return g[b][a + d]; };
    auto Y = [&](int d) { return g[b + d][a]; };
    const double uv = (u[fxi(i + 1, j, nx)] + u[fxi(i, j, nx)]) * 0.5;
    const double vv = (v[fyi(i, j + 1, nx)] + v[fyi(i, j, nx)]) * 0.5;
    const double dx = uv > 0 ? weno5_d(ih * (X(-2) - X(-3)), ih * (X(-1) - X(-2)), ih * (X(0) - X(-1)),
                                       ih * (X(1) - X(0)), ih * (X(2) - X(1)), eps)
                             : weno5_d(ih * (X(3) - X(2)), ih * (X(2) - X(1)), ih * (X(1) - X(0)),
                                       ih * (X(0) - X(-1)), ih * (X(-1) - X(-2)), eps);
    const double dy = vv > 0 ? weno5_d(ih * (Y(-2) - Y(-3)), ih * (Y(-1) - Y(-2)), ih * (Y(0) - Y(-1)),
                                       ih * (Y(1) - Y(0)), ih * (Y(2) - Y(1)), eps)
                             : weno5_d(ih * (Y(3) - Y(2)), ih * (Y(2) - Y(1)), ih * (Y(1) - Y(0)),
                                       ih * (Y(0) - Y(-1)), ih * (Y(-1) - Y(-2)), eps);
    out[ci(i, j, nx)] = g[b][a] - dt * (uv * dx + vv * dy);
  }
}

// ENO3 one-sided derivatives (minus, plus) along a line of seven values
// f[0..6] = f(-3..3) with spacing dx: divided differences 1/dx, 1/(2dx),
// (1/dx)/3 and the choice by the smaller |difference|
// (ADVECTION_METHOD_2D.h:2054-2126, in its expression forms).
__device__ __forceinline__ void eno3_d(const double* f, double dx, double* m, double* p) {
#pragma clang fp contract(off)
  const double o1 = 1.0 / dx, o2 = 1.0 / (2 * dx), o3 = o1 * 1 / 3;
  const double d1n3 = o1 * (f[1] - f[0]), d1n2 = o1 * (f[2] - f[1]), d1n1 = o1 * (f[3] - f[2]);
  const double d10 = o1 * (f[4] - f[3]), d1p1 = o1 * (f[5] - f[4]), d1p2 = o1 * (f[6] - f[5]);
  const double d2n2 = o2 * (d1n2 - d1n3), d2n1 = o2 * (d1n1 - d1n2), d20 = o2 * (d10 - d1n1);
  const double d2p1 = o2 * (d1p1 - d10), d2p2 = o2 * (d1p2 - d1p1);
  const double d3n2 = o3 * (d2n1 - d2n2), d3n1 = o3 * (d20 - d2n1), d30 = o3 * (d2p1 - d20), d3p1 = o3 * (d2p2 - d2p1);
  if (fabs(d2n1) <= fabs(d20)) {
    *m = fabs(d3n2) <= fabs(d3n1) ? d1n1 + d2n1 * dx + 2 * d3n2 * dx * dx : d1n1 + d2n1 * dx + 2 * d3n1 * dx * dx;
  } else {
    *m = fabs(d3n1) <= fabs(d30) ? d1n1 + d20 * dx - d3n1 * dx * dx : d1n1 + d20 * dx - d30 * dx * dx;
  }
  if (fabs(d20) <= fabs(d2p1)) {
    *p = fabs(d3n1) <= fabs(d30) ? d10 - d20 * dx - d3n1 * dx * dx : d10 - d20 * dx - d30 * dx * dx;
  } else {
    *p = fabs(d30) <= fabs(d3p1) ? d10 - d2p1 * dx + 2 * d30 * dx * dx : d10 - d2p1 * dx + 2 * d3p1 * dx * dx;
  }
}

// A face field of fnx x fny values read through the ghost rule of
// FillGhostCellsFrom (second interior value beyond each face; corners unused).
__device__ __forceinline__ double face_ghost_at(const double* f, int fnx, int fny, int i, int j) {
  i = i < 0 ? 1 : (i >= fnx ? fnx - 2 : i);
  j = j < 0 ? 1 : (j >= fny ? fny - 2 : j);
  return f[(size_t)i + (size_t)j * fnx];
}

// One forward-Euler ENO3 stage of one MAC velocity component (XCOMP: u on the
// (nx+1) x ny x-faces, else v on the nx x (ny+1) y-faces) by the old velocity,
// ADVECTION_2D::Solve_Velocity with use_3rd_eno_v (ADVECTION_2D.cpp:95-160) =
// ADVECTION_METHOD_2D::ENO3rd (ADVECTION_METHOD_2D.h:2038-2232): the advecting
// velocity at an x-face is (u, mean of the four y-faces around it), at a
// y-face (mean of the four x-faces, v); upwind ENO3 derivative per component;
// out = f - dt (u f_x + v f_y).  Grid dim3(ceil(fnx/64), ceil(fny/16)).
template <bool XCOMP>
__global__ void __launch_bounds__(BLOCK) k_advect_vel(int nx, int ny, double h, double dt, const double* __restrict__ u,
                                                      const double* __restrict__ v, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int fnx = XCOMP ? nx + 1 : nx, fny = XCOMP ? ny : ny + 1;
  const double* f = XCOMP ? u : v;
  __shared__ double g[RI_H + 2 * RI_G][RI_W + 2 * RI_G];
  int bx, by;
  xcd_tile(bx, by);
  const int i0 = bx * RI_W, j0 = by * RI_H;
  stage_weno_tile(g, f, fnx, fny, i0, j0);
  __syncthreads();
  const int c = threadIdx.x & 63, w = wave_id();
#pragma unroll
  for (int m = 0; m < RI_H / (BLOCK / 64); ++m) {
    const int r = w + (BLOCK / 64) * m, i = i0 + c, j = j0 + r;
    if (i >= fnx || j >= fny) continue;
    const int a = c + RI_G, b = r + RI_G;
    double lx[7], ly[7], mx, px, my, py;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      lx[k] = g[b][a - 3 + k];
      ly[k] = g[b - 3 + k][a];
    }
    eno3_d(lx, h, &mx, &px);
    eno3_d(ly, h, &my, &py);
    double uv, vv;
    if (XCOMP) {
      uv = g[b][a];
      vv = (face_ghost_at(v, nx, ny + 1, i, j) + face_ghost_at(v, nx, ny + 1, i, j + 1) +
            face_ghost_at(v, nx, ny + 1, i - 1, j) + face_ghost_at(v, nx, ny + 1, i - 1, j + 1)) *
           0.25;
    } else {
      uv = (face_ghost_at(u, nx + 1, ny, i, j - 1) + face_ghost_at(u, nx + 1, ny, i + 1, j - 1) +
            face_ghost_at(u, nx + 1, ny, i, j) + face_ghost_at(u, nx + 1, ny, i + 1, j)) *
           0.25;
      vv = g[b][a];
    }
    const double dxv = uv > 0 ? mx : px, dyv = vv > 0 ? my : py;
    out[(size_t)i + (size_t)j * fnx] = g[b][a] - dt * (uv * dxv + vv * dyv);
  }
}

// Mollified Heaviside of the viscosity step, VISCOSITY_2D::HeavisideFunction
// (VISCOSITY_2D.h:688-707, the reference's PI = 3.141592).
__device__ __forceinline__ double heaviside_mollified(double phi, double eps) {
#pragma clang fp contract(off)
  const double pi = 3.141592, oe = (double)1 / eps, op = (double)1 / pi;
  if (phi < -eps) return 0;
  if (phi > eps) return 1;
  return (double)0.5 + phi * oe * (double)0.5 + (double)0.5 * op * sin(pi * phi * oe);
}

// Explicit air-water viscosity update of one MAC component from the old
// velocity (u0, v0, 0 outside the grid), VISCOSITY_2D::ApplyViscosity
// (VISCOSITY_2D.h:460-548): density and viscosities from the mollified
// Heaviside of level-set averages (the face, its two corners across the
// other direction, the two cells), value = minus + (plus - minus) H;
// x-face: u += 2 coef/dx (mu_r (u_E - u) - mu_c (u - u_W))/dx
//            + coef/dy (mu_cu (du/dy + dv/dx)_N - mu_cd (du/dy + dv/dx)_S), coef = dt/rho_half;
// y-face the mirror image.  Every face, walls too.  One thread per face,
// grid dim3(ceil(fnx/BLOCK), ceil(fny/BAND)).
template <bool XCOMP>
__global__ void __launch_bounds__(BLOCK) k_viscosity(int nx, int ny, double h, double dt, const double* __restrict__ phi,
                                                     double rm, double rp, double mm, double mp, double eps,
                                                     const double* __restrict__ u0, const double* __restrict__ v0,
                                                     double* __restrict__ out) {
#pragma clang fp contract(off)
  const int fnx = XCOMP ? nx + 1 : nx, fny = XCOMP ? ny : ny + 1;
  const int i = (int)(blockIdx.x * BLOCK + threadIdx.x);
  if (i >= fnx) return;
  auto P = [&](int a, int b) { return phi_at(phi, nx, ny, a, b); };
  auto U = [&](int a, int b) { return (a < 0 || a > nx || b < 0 || b >= ny) ? 0.0 : u0[fxi(a, b, nx)]; };
  auto V = [&](int a, int b) { return (a < 0 || a >= nx || b < 0 || b > ny) ? 0.0 : v0[fyi(a, b, nx)]; };
  auto rho = [&](double p) { return rm + (rp - rm) * heaviside_mollified(p, eps); };
  auto mu = [&](double p) { return mm + (mp - mm) * heaviside_mollified(p, eps); };
  const double ox = 1.0 / h, oy = 1.0 / h;
  const int j1 = min(fny, (int)(blockIdx.y + 1) * BAND);
  for (int j = (int)blockIdx.y * BAND; j < j1; ++j) {
    if (XCOMP) {
      const double ph = (double)0.5 * (P(i - 1, j) + P(i, j));
      const double pu = (double)0.25 * (P(i - 1, j) + P(i, j) + P(i - 1, j + 1) + P(i, j + 1));
      const double pd = (double)0.25 * (P(i - 1, j - 1) + P(i, j - 1) + P(i - 1, j) + P(i, j));
      const double mr = mu(P(i, j)), mc = mu(P(i - 1, j)), mcu = mu(pu), mcd = mu(pd);
      const double coef = dt * ((double)1 / rho(ph));
      const double first =
          (double)2 * coef * ox * (mr * ox * (U(i + 1, j) - U(i, j)) - mc * ox * (U(i, j) - U(i - 1, j)));
      const double second = coef * oy * (mcu * (oy * (U(i, j + 1) - U(i, j)) + ox * (V(i, j + 1) - V(i - 1, j + 1))) -
                                         mcd * (oy * (U(i, j) - U(i, j - 1)) + ox * (V(i, j) - V(i - 1, j))));
      out[fxi(i, j, nx)] = U(i, j) + (first + second);
    } else {
      const double ph = (double)0.5 * (P(i, j - 1) + P(i, j));
      const double pu = (double)0.25 * (P(i, j - 1) + P(i + 1, j - 1) + P(i, j) + P(i + 1, j));
      const double pl = (double)0.25 * (P(i - 1, j - 1) + P(i, j - 1) + P(i - 1, j) + P(i, j));
      const double mu_u = mu(P(i, j)), mc = mu(P(i, j - 1)), mcu = mu(pu), mcl = mu(pl);
      const double coef = dt * ((double)1 / rho(ph));
      const double first =
          (double)2 * coef * oy * (mu_u * oy * (V(i, j + 1) - V(i, j)) - mc * oy * (V(i, j) - V(i, j - 1)));
      const double second = coef * ox * (mcu * (oy * (U(i + 1, j) - U(i + 1, j - 1)) + ox * (V(i + 1, j) - V(i, j))) -
                                         mcl * (oy * (U(i, j) - U(i, j - 1)) + ox * (V(i, j) - V(i - 1, j))));
      out[fyi(i, j, nx)] = V(i, j) + (first + second);
    }
  }
}

// Time-step helpers.  out = a x + b y (the TVD-RK3 stage combinations,
// EULERIAN_FLUID_SOLVER_2D.h:1700-1795: 0.75 / 0.25 and 1/3 / 2/3).
// out may be y (element k is read before it is written).
__global__ void __launch_bounds__(BLOCK) k_lincomb(size_t n, double a, const double* __restrict__ x, double b,
                                                   const double* y, double* out) {
#pragma clang fp contract(off)
  for (size_t k = blockIdx.x * (size_t)BLOCK + threadIdx.x; k < n; k += (size_t)gridDim.x * BLOCK)
    out[k] = a * x[k] + b * y[k];
}
// Gravity, EULERIAN_FLUID_SOLVER_2D::ApplyGraivity (EULERIAN_FLUID_SOLVER_2D.h:3249-3269):
// every x-face += dt gx, every y-face += dt gy.
__global__ void __launch_bounds__(BLOCK) k_gravity(size_t nu, size_t nv, double gxdt, double gydt, double* __restrict__ u,
                                                   double* __restrict__ v) {
  for (size_t k = blockIdx.x * (size_t)BLOCK + threadIdx.x; k < nu + nv; k += (size_t)gridDim.x * BLOCK) {
    if (k < nu) u[k] += gxdt;
    else v[k - nu] += gydt;
  }
}
// Wall velocities to zero, SetupBoundaryConditionForVelocity (air_bubble_rising
// branch, EULERIAN_FLUID_SOLVER_2D.h:2227-2270): u at i = 0, nx and v at j = 0, ny.
__global__ void __launch_bounds__(BLOCK) k_wall_zero(int nx, int ny, double* __restrict__ u, double* __restrict__ v) {
  const int t = (int)(blockIdx.x * BLOCK + threadIdx.x);
  if (t < ny) {
    u[fxi(0, t, nx)] = 0.0;
    u[fxi(nx, t, nx)] = 0.0;
  }
  if (t < nx) {
    v[fyi(t, 0, nx)] = 0.0;
    v[fyi(t, ny, nx)] = 0.0;
  }
}

// Unit normals, LEVELSET_2D::ComputeNormals (LEVELSET_2D.h:593-620): central
// differences over the linear-extrapolation ghosts, one-sided forward
// differences where the central gradient vanishes, scaled by 1/|n|
// (VECTOR_2D::MakeThisUnit), (0, 0) when |n| = 0.  Grid dim3(ceil(nx/BLOCK),
// ceil(ny/BAND)): one column of BAND rows per thread.
__global__ void __launch_bounds__(BLOCK) k_normals(int nx, int ny, double h, const double* __restrict__ phi,
                                                   double* __restrict__ onx, double* __restrict__ ony) {
#pragma clang fp contract(off)
  const int i = (int)(blockIdx.x * BLOCK + threadIdx.x);
  if (i >= nx) return;
  const double one_over_2dx = 1.0 / (2.0 * h), one_over_dx = 1.0 / h;
  const int j1 = min(ny, (int)(blockIdx.y + 1) * BAND);
  for (int j = (int)blockIdx.y * BAND; j < j1; ++j) {
    const double c = phi_at(phi, nx, ny, i, j), r = phi_at(phi, nx, ny, i + 1, j), t = phi_at(phi, nx, ny, i, j + 1);
    double a = (r - phi_at(phi, nx, ny, i - 1, j)) * one_over_2dx;
    double b = (t - phi_at(phi, nx, ny, i, j - 1)) * one_over_2dx;
    if (sqrt(a * a + b * b) == 0) {
      a = (r - c) * one_over_dx;
      b = (t - c) * one_over_dx;
    }
    const double mag = sqrt(a * a + b * b), inv = 1.0 / mag;
    const size_t k = ci(i, j, nx);
    onx[k] = mag != 0 ? a * inv : 0.0;
    ony[k] = mag != 0 ? b * inv : 0.0;
  }
}

}  // namespace p2d
