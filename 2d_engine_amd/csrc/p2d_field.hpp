// p2d field kernels: coefficient assembly, right-hand side, the FP64
// refinement step, the face update, diagnostics and the adjoint's face kernels.
//
// All of them are band kernels: block (bx, by) owns columns bx*BLOCK .. +BLOCK
// and rows by*BAND .. +BAND, so the rows above and below a cell are read by the
// same block (same XCD, same L2) and each element is fetched from HBM about once.
// Launch geometry: dim3(ceil(nx/BLOCK), ceil(ny/BAND)) (rows ny+1 for face sweeps).
#pragma once
#include "p2d_common.hpp"

namespace p2d {

#define P2D_BAND(NROWS)                                  \
  const int i = (int)(blockIdx.x * BLOCK + threadIdx.x); \
  const int j0 = (int)(blockIdx.y * BAND);               \
  const int j1 = min((int)(NROWS), j0 + BAND)

// --------------------------------------------------------------------------- assembly
// Per cell: jc = sigma * kappa (PROJECTION_2D.h:830-834), beta on its west and
// south faces (0 on the walls), v = 1/rho per side (A2).
__global__ void __launch_bounds__(BLOCK) k_coeff(int nx, int ny, double h, const double* __restrict__ phi,
                                                 double vm, double vp, double sigma, double* __restrict__ bw,
                                                 double* __restrict__ bs, double* __restrict__ jc, Scal* s) {
  P2D_BAND(ny);
  if (i >= nx) return;
  for (int j = j0; j < j1; ++j) {
    const size_t k = ci(i, j, nx);
    const double pc = phi[k];
    const double vc = minus_side(pc) ? vm : vp;
    double w = 0.0, so = 0.0;
    if (i > 0) {
      const double pl = phi[k - 1];
      w = gfm_beta(minus_side(pl) ? vm : vp, vc, pl, pc);
    }
    if (j > 0) {
      const double pb = phi[k - nx];
      so = gfm_beta(minus_side(pb) ? vm : vp, vc, pb, pc);
    }
    bw[k] = w;
    bs[k] = so;
    double jv = 0.0;
    if (sigma != 0.0) {
      bool bad = false;
      const double kap = curvature_at(phi, nx, ny, h, i, j, &bad);
      if (bad) s->err = 1;
      jv = sigma * kap;
    }
    jc[k] = jv;
  }
}

// C ABI face layout <-> cell-pitched layout.
__global__ void __launch_bounds__(BLOCK) k_faces_to_cells(int nx, int ny, const double* __restrict__ fx,
                                                          const double* __restrict__ fy, double* __restrict__ bw,
                                                          double* __restrict__ bs) {
  P2D_BAND(ny);
  if (i >= nx) return;
  for (int j = j0; j < j1; ++j) {
    bw[ci(i, j, nx)] = i > 0 ? fx[fxi(i, j, nx)] : 0.0;
    bs[ci(i, j, nx)] = j > 0 ? fy[fyi(i, j, nx)] : 0.0;
  }
}
__global__ void __launch_bounds__(BLOCK) k_cells_to_faces(int nx, int ny, const double* __restrict__ bw,
                                                          const double* __restrict__ bs, double* __restrict__ fx,
                                                          double* __restrict__ fy) {
  P2D_BAND(ny + 1);
  if (i > nx) return;
  for (int j = j0; j < j1; ++j) {
    if (j < ny) fx[fxi(i, j, nx)] = (i == 0 || i == nx) ? 0.0 : bw[ci(i, j, nx)];
    if (i < nx) fy[fyi(i, j, nx)] = (j == 0 || j == ny) ? 0.0 : bs[ci(i, j, nx)];
  }
}

// --------------------------------------------------------------------------- right-hand side
// Right-hand-side statistics: mean (removed by the residual kernel, A0 being
// singular on constants) and the norms of b - mean into s->red[0..1].
__device__ __forceinline__ void rhs_finalize(const double (&acc)[3], size_t N, Scal* s) {
  const double m = acc[2] / (double)N;
  s->bmean = m;
  s->red[0] = fmax(0.0, acc[0] - (double)N * m * m);
  s->red[1] = acc[1] + fabs(m);
}

// b' = h^2 (-D u/dt + D(B J)), the reference's b = -div - F (POISSON_SOLVER.cpp:951)
// with J_f = s_f a_f / h.  u: (nx+1) x ny x-faces, v: nx x (ny+1) y-faces.
__global__ void __launch_bounds__(BLOCK) k_rhs(int nx, int ny, double h, double dt, const double* __restrict__ u,
                                               const double* __restrict__ v, const double* __restrict__ phi,
                                               const double* __restrict__ jc, const double* __restrict__ bw,
                                               const double* __restrict__ bs, double* __restrict__ b,
                                               double* partial, unsigned* counter, Scal* s) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h, idt = 1.0 / dt, h2 = h * h;
  const Coef<double> C{bw, bs, nx, ny};
  double acc[3] = {0.0, 0.0, 0.0};
  P2D_BAND(ny);
  for (int j = j0; j < j1 && i < nx; ++j) {
    const size_t k = ci(i, j, nx);
    const double d = (ih * (u[fxi(i + 1, j, nx)] - u[fxi(i, j, nx)]) + ih * (v[fyi(i, j + 1, nx)] - v[fyi(i, j, nx)])) * idt;
    double Jl = 0, Jr = 0, Jb = 0, Jt = 0;
    if (jc) {
      if (i > 0) Jl = face_sa(phi, jc, k - 1, k) / h;
      if (i + 1 < nx) Jr = face_sa(phi, jc, k, k + 1) / h;
      if (j > 0) Jb = face_sa(phi, jc, k - nx, k) / h;
      if (j + 1 < ny) Jt = face_sa(phi, jc, k, k + nx) / h;
    }
    const double f = (C.E(i, j) * Jr - C.W(i, j) * Jl + C.N(i, j) * Jt - C.S(i, j) * Jb) * ih;
    const double bv = h2 * (-d + f);
    b[k] = bv;
    acc[0] += bv * bv;
    acc[1] = fmax(acc[1], fabs(bv));
    acc[2] += bv;
  }
  if (grid_reduce<3>(acc, 2u, partial, counter) && threadIdx.x == 0) rhs_finalize(acc, (size_t)nx * ny, s);
}

// --------------------------------------------------------------------------- FP64 refinement step
// On the double-double pressure: p_new = p_old + e (e may be NULL),
// r = (b' - mean b') - A0' p_new with accurate face differences, stored as R;
// ||r||_2^2, ||r||_inf into s->red[0..1].  The block marches down its band and
// keeps the rows above/below in registers.
template <class TE, class R>
__global__ void __launch_bounds__(BLOCK) k_correct_resid(int nx, int ny, const double* __restrict__ bw,
                                                         const double* __restrict__ bs, const double* __restrict__ b,
                                                         const double* __restrict__ ph, const double* __restrict__ pl,
                                                         const TE* __restrict__ e, double* __restrict__ qh,
                                                         double* __restrict__ ql, R* __restrict__ r,
                                                         double* partial, unsigned* counter, Scal* s) {
#pragma clang fp contract(off)
  double acc[3] = {0.0, 0.0, 0.0};
  const double bm = s->bmean;
  const Coef<double> C{bw, bs, nx, ny};
  P2D_BAND(ny);
  auto P = [&](size_t q, double& hh, double& ll) {
    if (e) dd_add(ph[q], pl[q], (double)e[q], hh, ll);
    else {
      hh = ph[q];
      ll = pl[q];
    }
  };
  if (i < nx && j0 < j1) {
    double sh = 0, sl = 0, ch, cl, nh = 0, nl = 0;  // south, centre, north
    if (j0 > 0) P(ci(i, j0 - 1, nx), sh, sl);
    P(ci(i, j0, nx), ch, cl);
    for (int j = j0; j < j1; ++j) {
      const size_t k = ci(i, j, nx);
      if (j + 1 < ny) P(k + nx, nh, nl);
      double a = 0;
      if (i > 0) {
        double wh, wl;
        P(k - 1, wh, wl);
        a += C.W(i, j) * ((ch - wh) + (cl - wl));
      }
      if (i + 1 < nx) {
        double eh, el;
        P(k + 1, eh, el);
        a += C.E(i, j) * ((ch - eh) + (cl - el));
      }
      if (j > 0) a += C.S(i, j) * ((ch - sh) + (cl - sl));
      if (j + 1 < ny) a += C.N(i, j) * ((ch - nh) + (cl - nl));
      const double rv = (b[k] - bm) - a;
      if (qh) {
        qh[k] = ch;
        ql[k] = cl;
      }
      r[k] = (R)rv;
      acc[0] += rv * rv;
      acc[1] = fmax(acc[1], fabs(rv));
      acc[2] += rv;
      sh = ch;
      sl = cl;
      ch = nh;
      cl = nl;
    }
  }
  if (grid_reduce<3>(acc, 2u, partial, counter) && threadIdx.x == 0) {
    s->red[0] = acc[0];
    s->red[1] = acc[1];
    s->red[2] = acc[2];
  }
}

// --------------------------------------------------------------------------- face update / diagnostics
// Consistent A14: u_f -= dt beta_f (p_R - p_L + s a)/h on interior faces (walls untouched).
__global__ void __launch_bounds__(BLOCK) k_face_update(int nx, int ny, double h, double dt,
                                                       const double* __restrict__ ph, const double* __restrict__ pl,
                                                       const double* __restrict__ phi, const double* __restrict__ jc,
                                                       const double* __restrict__ bw, const double* __restrict__ bs,
                                                       double* __restrict__ u, double* __restrict__ v) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h;
  P2D_BAND(ny);
  if (i >= nx) return;
  for (int j = j0; j < j1; ++j) {
    const size_t k = ci(i, j, nx);
    if (i > 0) {
      double gap = dd_diff(ph, pl, k, k - 1);
      gap += face_sa(phi, jc, k - 1, k);
      u[fxi(i, j, nx)] -= dt * gap * bw[k] * ih;
    }
    if (j > 0) {
      double gap = dd_diff(ph, pl, k, k - nx);
      gap += face_sa(phi, jc, k - nx, k);
      v[fyi(i, j, nx)] -= dt * gap * bs[k] * ih;
    }
  }
}

// max |D u| (cell divergence, not divided by dt), max |u|, max |v| into
// s->red[0..2].  Grid covers (nx+1) columns and ny+1 rows.
__global__ void __launch_bounds__(BLOCK) k_diagnostics(int nx, int ny, double h, const double* __restrict__ u,
                                                       const double* __restrict__ v, double* partial,
                                                       unsigned* counter, Scal* s) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h;
  double acc[3] = {0.0, 0.0, 0.0};
  P2D_BAND(ny + 1);
  for (int j = j0; j < j1 && i <= nx; ++j) {
    if (j < ny) acc[1] = fmax(acc[1], fabs(u[fxi(i, j, nx)]));
    if (i < nx) acc[2] = fmax(acc[2], fabs(v[fyi(i, j, nx)]));
    if (i < nx && j < ny) {
      const double d = ih * (u[fxi(i + 1, j, nx)] - u[fxi(i, j, nx)]) + ih * (v[fyi(i, j + 1, nx)] - v[fyi(i, j, nx)]);
      acc[0] = fmax(acc[0], fabs(d));
    }
  }
  if (grid_reduce<3>(acc, 7u, partial, counter) && threadIdx.x == 0) {
    for (int q = 0; q < 3; ++q) s->red[q] = acc[q];
  }
}

// Sum of p = hi + lo into s->red[3] (for the mean); 1-D grid-stride launch.
__global__ void __launch_bounds__(BLOCK) k_psum(size_t N, const double* __restrict__ ph, const double* __restrict__ pl,
                                                double* partial, unsigned* counter, Scal* s) {
  double acc[1] = {0.0};
  for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < N; k += (size_t)gridDim.x * BLOCK) acc[0] += ph[k] + pl[k];
  if (grid_reduce<1>(acc, 0u, partial, counter) && threadIdx.x == 0) s->red[3] = acc[0];
}

// out = p - mean (mean_zero) or p - p(0,0) (the reference's pin, POISSON_SOLVER.cpp:789-807).
__global__ void __launch_bounds__(BLOCK) k_shift(size_t N, const double* __restrict__ ph, const double* __restrict__ pl,
                                                 double* __restrict__ out, int mean_zero, const Scal* s) {
#pragma clang fp contract(off)
  const size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k >= N) return;
  if (mean_zero) out[k] = (ph[k] - s->red[3] / (double)N) + pl[k];
  else out[k] = (ph[k] - ph[0]) + (pl[k] - pl[0]);
}

// --------------------------------------------------------------------------- adjoint
// b' = -h^2 D(B g) with g taken on interior faces only; mean and norms as k_rhs.
__global__ void __launch_bounds__(BLOCK) k_adj_rhs(int nx, int ny, double h, const double* __restrict__ gu,
                                                   const double* __restrict__ gv, const double* __restrict__ bw,
                                                   const double* __restrict__ bs, double* __restrict__ b,
                                                   double* partial, unsigned* counter, Scal* s) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h, h2 = h * h;
  const Coef<double> C{bw, bs, nx, ny};
  double acc[3] = {0.0, 0.0, 0.0};
  P2D_BAND(ny);
  for (int j = j0; j < j1 && i < nx; ++j) {
    const size_t k = ci(i, j, nx);
    const double gl = i > 0 ? C.W(i, j) * gu[fxi(i, j, nx)] : 0.0;
    const double gr = i + 1 < nx ? C.E(i, j) * gu[fxi(i + 1, j, nx)] : 0.0;
    const double gb = j > 0 ? C.S(i, j) * gv[fyi(i, j, nx)] : 0.0;
    const double gt = j + 1 < ny ? C.N(i, j) * gv[fyi(i, j + 1, nx)] : 0.0;
    const double bv = h2 * (-((gr - gl) * ih + (gt - gb) * ih));
    b[k] = bv;
    acc[0] += bv * bv;
    acc[1] = fmax(acc[1], fabs(bv));
    acc[2] += bv;
  }
  if (grid_reduce<3>(acc, 2u, partial, counter) && threadIdx.x == 0) rhs_finalize(acc, (size_t)nx * ny, s);
}

// du = g - G mu, dbeta = -dt du (G p + s a / h) on interior faces; walls 0.
__global__ void __launch_bounds__(BLOCK) k_adj_faces(int nx, int ny, double h, double dt,
                                                     const double* __restrict__ gu, const double* __restrict__ gv,
                                                     const double* __restrict__ muh, const double* __restrict__ mul,
                                                     const double* __restrict__ ph, const double* __restrict__ pl,
                                                     const double* __restrict__ phi, const double* __restrict__ jc,
                                                     double* du, double* dv, double* dbx, double* dby) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h;
  P2D_BAND(ny);
  if (i >= nx) return;
  for (int j = j0; j < j1; ++j) {
    const size_t k = ci(i, j, nx);
    {
      const size_t f = fxi(i, j, nx);
      double t = 0, g = 0;
      if (i > 0) {
        t = gu[f] - dd_diff(muh, mul, k, k - 1) * ih;
        const double w = dd_diff(ph, pl, k, k - 1) * ih + face_sa(phi, jc, k - 1, k) / h;
        g = -dt * t * w;
      }
      if (du) du[f] = t;
      if (dbx) dbx[f] = g;
      if (i == nx - 1) {
        if (du) du[fxi(nx, j, nx)] = 0.0;
        if (dbx) dbx[fxi(nx, j, nx)] = 0.0;
      }
    }
    {
      const size_t f = fyi(i, j, nx);
      double t = 0, g = 0;
      if (j > 0) {
        t = gv[f] - dd_diff(muh, mul, k, k - nx) * ih;
        const double w = dd_diff(ph, pl, k, k - nx) * ih + face_sa(phi, jc, k - nx, k) / h;
        g = -dt * t * w;
      }
      if (dv) dv[f] = t;
      if (dby) dby[f] = g;
      if (j == ny - 1) {
        if (dv) dv[fyi(i, ny, nx)] = 0.0;
        if (dby) dby[fyi(i, ny, nx)] = 0.0;
      }
    }
  }
}

// --------------------------------------------------------------------------- test hook
// y = scale * A0' x (scale = 1/h^2 gives the unscaled operator).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_apply(int nx, int ny, const T* __restrict__ bw, const T* __restrict__ bs,
                                                 const T* __restrict__ x, T* __restrict__ y, double scale) {
  const Coef<T> C{bw, bs, nx, ny};
  P2D_BAND(ny);
  if (i >= nx) return;
  for (int j = j0; j < j1; ++j) {
    const size_t c = ci(i, j, nx);
    const T xc = x[c];
    T a = 0;
    if (i > 0) a += C.W(i, j) * (xc - x[c - 1]);
    if (i + 1 < nx) a += C.E(i, j) * (xc - x[c + 1]);
    if (j > 0) a += C.S(i, j) * (xc - x[c - nx]);
    if (j + 1 < ny) a += C.N(i, j) * (xc - x[c + nx]);
    y[c] = (T)((double)a * scale);
  }
}

template <class A, class B>
__global__ void __launch_bounds__(BLOCK) k_cast(size_t n, const A* __restrict__ a, B* __restrict__ b) {
  const size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k < n) b[k] = (B)a[k];
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_fill(size_t n, T* a, T v) {
  const size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k < n) a[k] = v;
}

}  // namespace p2d
