// p2d field kernels: coefficient assembly, right-hand side, the FP64
// refinement step, the face update, diagnostics and the adjoint's face kernels.
//
// Most are band kernels: block (bx, by) owns columns bx*BLOCK .. +BLOCK and rows
// by*BAND .. +BAND (launch dim3(ceil(nx/BLOCK), ceil(ny/BAND)), rows ny+1 for
// face sweeps).  The refinement step, which runs several times per projection,
// is a tile kernel: the double-double pressure of a 64 x DTY tile and its ring
// is staged in LDS with every load issued at once.
//
// Interface faces.  At a face across the interface the pressure difference
// p_R - p_L nearly cancels the ghost-fluid jump s a (both ~ sigma kappa while
// their sum ~ h |u|).  The face flux G_f = beta_f (p_R - p_L + s_f a_f) is
// therefore formed in one expression (fdiff) with the high words differenced
// exactly, in the residual and in the face update alike; the right-hand side
// keeps only the velocity part -h^2 D u/dt.  Rounding each term separately
// would floor max |D u_proj| at ~eps sigma kappa dt / h^2.
#pragma once
#include "p2d_common.hpp"

namespace p2d {

#define P2D_BAND(NROWS)                                  \
  const int i = (int)(blockIdx.x * BLOCK + threadIdx.x); \
  const int j0 = (int)(blockIdx.y * BAND);               \
  const int j1 = min((int)(NROWS), j0 + BAND)

constexpr int FTY = 16;  // rows per block of the refinement tile kernel

// p_R - p_L + sa for double-double p = (h, l): the high words are differenced
// exactly and the jump added exactly (two_sum), so only the (small) result is
// rounded.
__device__ __forceinline__ double fdiff(double hR, double lR, double hL, double lL, double sa) {
#pragma clang fp contract(off)
  double d, de, t, te;
  two_sum(hR, -hL, d, de);
  two_sum(d, sa, t, te);
  return t + (te + (de + (lR - lL)));
}

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
    // sigma = 0: no jump term, jc is not read (p2d_get_coefficients reports zeros)
    if (sigma != 0.0) {
      bool bad = false;
      const double kap = curvature_at(phi, nx, ny, h, i, j, &bad);
      if (bad) s->err = 1;
      jc[k] = sigma * kap;
    }
  }
}

// Ghost-fluid jump term s a on the west and south face of every cell
// (POISSON_SOLVER.cpp:874-890 restated per face: a interpolated from jc by the
// level set, s = +1 when the cell on the right/top is on the minus side).
__global__ void __launch_bounds__(BLOCK) k_jump(int nx, int ny, const double* __restrict__ phi,
                                                const double* __restrict__ jc, double* __restrict__ saw,
                                                double* __restrict__ sas) {
  P2D_BAND(ny);
  if (i >= nx) return;
  for (int j = j0; j < j1; ++j) {
    const size_t k = ci(i, j, nx);
    saw[k] = i > 0 ? face_sa(phi, jc, k - 1, k) : 0.0;
    sas[k] = j > 0 ? face_sa(phi, jc, k - nx, k) : 0.0;
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
// Right-hand-side statistics (block partials sum r0^2, max |r0|, sum r0, sum b):
// the initial residual's norms into s->red[0..2] and the mean of b' (removed by
// the residual kernel, A0 being singular on constants) into s->bmean.
__global__ void __launch_bounds__(FINISH_THREADS) k_fin_rhs(const double* __restrict__ partial, int G, double N,
                                                             Scal* s) {
  double a[4];
  finish_reduce<4>(a, partial, G, 2u);
  if (threadIdx.x == 0) {
    s->red[0] = a[0];
    s->red[1] = a[1];
    s->red[2] = a[2];
    s->bmean = a[3] / N;
  }
}

// Velocity part of b' = h^2 (-D u/dt + D(B J)), the reference's b = -div - F
// (POISSON_SOLVER.cpp:951) with J_f = s_f a_f / h: b = -h^2 D u/dt; the jump
// part enters the residual face by face (fdiff).  Also the initial residual
// (p = 0) r0 = b - sum_f +-beta_f s_f a_f, stored as R for the first solve
// pass (its mean is removed there).  u: (nx+1) x ny x-faces, v: nx x (ny+1)
// y-faces; saw/sas may be NULL (no jump).  With div given (nx x ny, the
// reference's divergence_field, already divided by dt) u, v and dt are unused
// and b = -h^2 div.
template <class R>
__global__ void __launch_bounds__(BLOCK) k_rhs(int nx, int ny, double h, double dt, const double* __restrict__ u,
                                               const double* __restrict__ v, const double* __restrict__ div,
                                               const double* __restrict__ bw, const double* __restrict__ bs,
                                               const double* __restrict__ saw, const double* __restrict__ sas,
                                               double* __restrict__ b, R* __restrict__ r0, double* partial) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h, idt = 1.0 / dt, h2 = h * h;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  P2D_BAND(ny);
  for (int j = j0; j < j1 && i < nx; ++j) {
    const size_t k = ci(i, j, nx);
    const double d =
        div ? div[k]
            : (ih * (u[fxi(i + 1, j, nx)] - u[fxi(i, j, nx)]) + ih * (v[fyi(i, j + 1, nx)] - v[fyi(i, j, nx)])) * idt;
    const double bv = h2 * (-d);
    double g = 0.0;
    if (saw) {
      if (i > 0) g += bw[k] * saw[k];
      if (i + 1 < nx) g -= bw[k + 1] * saw[k + 1];
      if (j > 0) g += bs[k] * sas[k];
      if (j + 1 < ny) g -= bs[k + nx] * sas[k + nx];
    }
    const double rv = bv - g;
    b[k] = bv;
    r0[k] = (R)rv;
    acc[0] += rv * rv;
    acc[1] = fmax(acc[1], fabs(rv));
    acc[2] += rv;
    acc[3] += bv;
  }
  block_partial<4>(acc, 2u, partial);
}

// PROJECTION_2D::DetermineDivergence (air-water, PROJECTION_2D.h:395-414):
// div = (1/dx (u_{i+1} - u_i) + 1/dy (v_{j+1} - v_j)) then div *= 1/dt, in that
// operation order (dx = dy = h).
__global__ void __launch_bounds__(BLOCK) k_divergence(int nx, int ny, double h, double dt,
                                                      const double* __restrict__ u, const double* __restrict__ v,
                                                      double* __restrict__ div) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h, idt = 1.0 / dt;
  P2D_BAND(ny);
  for (int j = j0; j < j1 && i < nx; ++j) {
    double d = ih * (u[fxi(i + 1, j, nx)] - u[fxi(i, j, nx)]) + ih * (v[fyi(i, j + 1, nx)] - v[fyi(i, j, nx)]);
    d *= idt;
    div[ci(i, j, nx)] = d;
  }
}

// --------------------------------------------------------------------------- FP64 refinement step
// On the double-double pressure (high part FP64, low part FP32: 77 significant bits, the
// low part rounded once per update): p_new = p_old + e (e may be NULL; ph/pl NULL: p_old = 0;
// pd: an inner update left pending, e + alpha pd formed in TE first),
// r = (b' - mean b') - sum_f +-G_f with G_f = beta_f fdiff(p, s a) (west and
// south faces +, east and north -), stored as R; block partials sum r^2,
// max |r|, sum r.  saw/sas may be NULL (no jump).  Tile kernel over FTX x FTY
// cells: lane c of a wavefront holds the column pair (2c, 2c+1) of a row and
// each wavefront FTY / 4 consecutive rows, so every array moves with one
// two-element load or store per lane, a row's north faces are the next row's
// south ones and the east face of a pair's second cell comes from the next lane.
// p_new of the tile plus a one-cell ring is staged in LDS.  The finest level has
// an even nx.  Grid dim3(ceil(nx/FTX), ceil(ny/FTY)), BLOCK threads; p_new of the
// tile interior into (qh, ql).
constexpr int FTX = 128;
template <class TE, class R>
__global__ void __launch_bounds__(BLOCK) k_correct_resid(int nx, int ny, const double* __restrict__ bw,
                                                         const double* __restrict__ bs,
                                                         const double* __restrict__ saw,
                                                         const double* __restrict__ sas, const double* __restrict__ b,
                                                         const double* __restrict__ ph, const float* __restrict__ pl,
                                                         const TE* __restrict__ e, const TE* __restrict__ pd,
                                                         double* __restrict__ qh, float* __restrict__ ql,
                                                         R* __restrict__ r, double* partial, const Scal* s) {
#pragma clang fp contract(off)
  // row layout: [1] west ring, [2 + 2c, 3 + 2c] pair c, [FTX + 2] east ring
  __shared__ Pair<double> sh2[FTY + 2][FTX / 2 + 2], sl2[FTY + 2][FTX / 2 + 2];
  double(*sh)[FTX + 4] = reinterpret_cast<double(*)[FTX + 4]>(&sh2[0][0]);
  double(*sl)[FTX + 4] = reinterpret_cast<double(*)[FTX + 4]>(&sl2[0][0]);
  const double bm = s->bmean;
  const TE ta = pd ? (TE)s->alpha : (TE)0;
  const int c = threadIdx.x & 63, w = wave_id();
  int bx, by;
  xcd_tile(bx, by);
  const int i0 = bx * FTX, j0 = by * FTY;
  const int iA = i0 + 2 * c;
  const bool col_in = iA < nx;
  // p_new of one cell from p_old (hi, lo) and the correction (e, the pending update pd, both
  // or neither)
  const bool corr = e != nullptr || pd != nullptr;
  auto pnew = [&](double oh, double ol, double ev, double& hh, double& ll) {
    if (!ph) {
      hh = ev;
      ll = 0.0;
    } else if (corr) {
      dd_add(oh, ol, ev, hh, ll);
      ll = (double)(float)ll;  // the stored low part is FP32: use what is stored
    } else {
      hh = oh;
      ll = ol;
    }
  };
  constexpr int NW = BLOCK / 64, RR = (FTY + 2 + NW - 1) / NW;
#pragma unroll
  for (int m = 0; m < RR; ++m) {
    const int rr = w + NW * m, jj = j0 - 1 + rr;
    if (rr >= FTY + 2) continue;
    Pair<double> h2{0.0, 0.0}, l2{0.0, 0.0};
    if (col_in && jj >= 0 && jj < ny) {
      const unsigned k = cell_u(iA, jj, nx);
      const Pair<double> oh = ph ? ld_at<Pair<double>>(ph, k) : Pair<double>{0.0, 0.0};
      const Pair<float> ol = ph ? ld_at<Pair<float>>(pl, k) : Pair<float>{0.0f, 0.0f};
      Pair<TE> e2 = e ? ld_at<Pair<TE>>(e, k) : Pair<TE>{(TE)0, (TE)0};
      if (pd) {  // the inner pass's pending update, in TE as the head applies it
        const Pair<TE> d2 = ld_at<Pair<TE>>(pd, k);
        e2.a += ta * d2.a;
        e2.b += ta * d2.b;
      }
      pnew(oh.a, (double)ol.a, (double)e2.a, h2.a, l2.a);
      pnew(oh.b, (double)ol.b, (double)e2.b, h2.b, l2.b);
    }
    sh2[rr][1 + c] = h2;
    sl2[rr][1 + c] = l2;
  }
  if (threadIdx.x < 2 * (FTY + 2)) {
    const int side = (int)threadIdx.x / (FTY + 2), rr = (int)threadIdx.x % (FTY + 2);
    const int ii = side ? i0 + FTX : i0 - 1, jj = j0 - 1 + rr;
    double hh = 0.0, ll = 0.0;
    if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) {
      const size_t q = ci(ii, jj, nx);
      const TE e0 = e ? e[q] : (TE)0;
      const TE ev = pd ? e0 + ta * pd[q] : e0;
      pnew(ph ? ph[q] : 0.0, ph ? (double)pl[q] : 0.0, (double)ev, hh, ll);
    }
    sh[rr][side ? FTX + 2 : 1] = hh;
    sl[rr][side ? FTX + 2 : 1] = ll;
  }
  // faces and jumps of the wave's rows (issued before the barrier): south of rows
  // rb .. rb + RPW (the last is the north of the wave's top row), west of A and B
  // (B's west = A's east), east of B from the next lane
  constexpr int RPW = FTY / NW;
  const int rb = w * RPW;
  double sA[RPW + 1], sB[RPW + 1], jsA[RPW + 1], jsB[RPW + 1], wA[RPW], wB[RPW], eB[RPW], jwA[RPW], jwB[RPW],
      jeB[RPW];
#pragma unroll
  for (int m = 0; m <= RPW; ++m) {
    const int j = j0 + rb + m;
    const bool in = col_in && j < ny;
    const unsigned k = cell_u(iA, j, nx);
    const Pair<double> s2 = in ? ld_at<Pair<double>>(bs, k) : Pair<double>{0.0, 0.0};
    const Pair<double> j2 = in && sas ? ld_at<Pair<double>>(sas, k) : Pair<double>{0.0, 0.0};
    sA[m] = s2.a;
    sB[m] = s2.b;
    jsA[m] = j2.a;
    jsB[m] = j2.b;
  }
#pragma unroll
  for (int m = 0; m < RPW; ++m) {
    const int j = j0 + rb + m;
    const bool in = col_in && j < ny;
    const unsigned k = cell_u(iA, j, nx);
    const Pair<double> w2 = in ? ld_at<Pair<double>>(bw, k) : Pair<double>{0.0, 0.0};
    const Pair<double> j2 = in && saw ? ld_at<Pair<double>>(saw, k) : Pair<double>{0.0, 0.0};
    wA[m] = w2.a;
    wB[m] = w2.b;
    jwA[m] = j2.a;
    jwB[m] = j2.b;
  }
#pragma unroll
  for (int m = 0; m < RPW; ++m) {
    const int j = j0 + rb + m;
    double ev = lane_next(wA[m]), jv = lane_next(jwA[m]);
    if (c == 63 && iA + 2 < nx && j < ny) {
      const unsigned k = cell_u(iA + 2, j, nx);
      ev = ld_at<double>(bw, k);
      jv = saw ? ld_at<double>(saw, k) : 0.0;
    }
    eB[m] = ev;
    jeB[m] = jv;
  }
  __syncthreads();
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int m = 0; m < RPW; ++m) {
    const int rr = rb + m, j = j0 + rr;
    if (!col_in || j >= ny) continue;
    const unsigned k = cell_u(iA, j, nx);
    const Pair<double> ch2 = sh2[rr + 1][1 + c], cl2 = sl2[rr + 1][1 + c];
    const Pair<double> dh2 = sh2[rr][1 + c], dl2 = sl2[rr][1 + c];
    const Pair<double> uh2 = sh2[rr + 2][1 + c], ul2 = sl2[rr + 2][1 + c];
    const Pair<double> b2 = ld_at<Pair<double>>(b, k);
    // cell A (column iA): west, east (= B), south, north
    double gA = 0.0;
    if (iA > 0) gA += wA[m] * fdiff(ch2.a, cl2.a, sh[rr + 1][1 + 2 * c], sl[rr + 1][1 + 2 * c], jwA[m]);
    gA -= wB[m] * fdiff(ch2.b, cl2.b, ch2.a, cl2.a, jwB[m]);
    if (j > 0) gA += sA[m] * fdiff(ch2.a, cl2.a, dh2.a, dl2.a, jsA[m]);
    if (j + 1 < ny) gA -= sA[m + 1] * fdiff(uh2.a, ul2.a, ch2.a, cl2.a, jsA[m + 1]);
    // cell B (column iA + 1): west (= A), east, south, north
    double gB = 0.0;
    gB += wB[m] * fdiff(ch2.b, cl2.b, ch2.a, cl2.a, jwB[m]);
    if (iA + 2 < nx) gB -= eB[m] * fdiff(sh[rr + 1][4 + 2 * c], sl[rr + 1][4 + 2 * c], ch2.b, cl2.b, jeB[m]);
    if (j > 0) gB += sB[m] * fdiff(ch2.b, cl2.b, dh2.b, dl2.b, jsB[m]);
    if (j + 1 < ny) gB -= sB[m + 1] * fdiff(uh2.b, ul2.b, ch2.b, cl2.b, jsB[m + 1]);
    const double ra = (b2.a - bm) - gA, rbv = (b2.b - bm) - gB;
    if (qh) {
      st_at(qh, k, ch2);
      st_at(ql, k, Pair<float>{(float)cl2.a, (float)cl2.b});
    }
    st_at(r, k, Pair<R>{(R)ra, (R)rbv});
    acc[0] += ra * ra;
    acc[1] = fmax(acc[1], fabs(ra));
    acc[2] += ra;
    acc[0] += rbv * rbv;
    acc[1] = fmax(acc[1], fabs(rbv));
    acc[2] += rbv;
  }
  block_partial<3>(acc, 2u, partial);
}

// --------------------------------------------------------------------------- face update / diagnostics
// Consistent A14: u_f -= dt G_f / h, G_f = beta_f fdiff(p, s a), on interior
// faces (walls untouched): the same G_f as the residual, so D u_proj = -dt r / h^2
// up to the rounding of u.
__global__ void __launch_bounds__(BLOCK) k_face_update(int nx, int ny, double h, double dt,
                                                       const double* __restrict__ ph, const float* __restrict__ pl,
                                                       const double* __restrict__ saw, const double* __restrict__ sas,
                                                       const double* __restrict__ bw, const double* __restrict__ bs,
                                                       double* __restrict__ u, double* __restrict__ v) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h;
  P2D_BAND(ny);
  if (i >= nx) return;
  for (int j = j0; j < j1; ++j) {
    const size_t k = ci(i, j, nx);
    const double ch = ph[k], cl = (double)pl[k];
    if (i > 0) {
      const double G = bw[k] * fdiff(ch, cl, ph[k - 1], pl[k - 1], saw ? saw[k] : 0.0);
      u[fxi(i, j, nx)] -= dt * G * ih;
    }
    if (j > 0) {
      const double G = bs[k] * fdiff(ch, cl, ph[k - nx], pl[k - nx], sas ? sas[k] : 0.0);
      v[fyi(i, j, nx)] -= dt * G * ih;
    }
  }
}

// max |D u| (cell divergence, not divided by dt), max |u|, max |v| into
// s->red[0..2].  Grid covers (nx+1) columns and ny+1 rows.
__global__ void __launch_bounds__(BLOCK) k_diagnostics(int nx, int ny, double h, const double* __restrict__ u,
                                                       const double* __restrict__ v, double* partial, Scal* s) {
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
  block_partial<3>(acc, 7u, partial);
}

// The maxima the time-step restriction needs, EULERIAN_FLUID_SOLVER_2D::CFLOneTimeStep
// (EULERIAN_FLUID_SOLVER_2D.h:3517-3637): max|u| over every x-face, max|v| over every
// y-face and, when phi is given, max|curvature| over the cells; block partials for
// k_fin_store<3>.  A vanishing level-set gradient sets s->err, as in k_coeff.
__global__ void __launch_bounds__(BLOCK) k_cfl_max(int nx, int ny, double h, const double* __restrict__ phi,
                                                   const double* __restrict__ u, const double* __restrict__ v,
                                                   double* partial, Scal* s) {
  double acc[3] = {0.0, 0.0, 0.0};
  bool bad = false;
  P2D_BAND(ny + 1);
  for (int j = j0; j < j1 && i <= nx; ++j) {
    if (j < ny) acc[0] = fmax(acc[0], fabs(u[fxi(i, j, nx)]));
    if (i < nx) acc[1] = fmax(acc[1], fabs(v[fyi(i, j, nx)]));
    if (phi && i < nx && j < ny) acc[2] = fmax(acc[2], fabs(curvature_at(phi, nx, ny, h, i, j, &bad)));
  }
  if (bad) s->err = 1;
  block_partial<3>(acc, 7u, partial);
}

// EULERIAN_FLUID_SOLVER_2D::CFLOneTimeStep's formula (EULERIAN_FLUID_SOLVER_2D.h:3517-3637)
// from the maxima: advection c = max|u|/h + max|v|/h, gravity g = sqrt(9.8/h), viscosity
// w = max(mu+/rho+, mu-/rho-)(2/h^2 + 2/h^2), surface tension s = sqrt(sigma max|kappa| /
// (min rho h^2)); dt = cfl / (0.5 (c + w + sqrt((c + w)^2 + 4 g^2 + 4 s^2))), 0 when that
// denominator is 0.  One expression order on host and device (no contraction).
__host__ __device__ inline double cfl_step(double umax, double vmax, double kmax, double h, double rho_m,
                                           double rho_p, double mu_m, double mu_p, double sigma, double cfl_number) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h, ih2 = 1.0 / (h * h);
  const double cf = umax * ih + vmax * ih;
  const double gf = sqrt(9.8 * ih);
  const double vf = fmax(mu_p / rho_p, mu_m / rho_m) * (2 * ih2 + 2 * ih2);
  const double sf = sqrt(sigma * kmax / (fmin(rho_p, rho_m) * (h * h)));
  const double cfl = 0.5 * (cf + vf + sqrt((cf + vf) * (cf + vf) + 4 * (gf * gf) + 4 * (sf * sf)));
  return cfl == 0 ? 0.0 : cfl_number / cfl;
}

// The CFL step from the maxima k_cfl_max and k_fin_store<3> left in s->red[0..2], into *dt
// (device memory); NaN when the curvature was undefined (s->err).
__global__ void k_cfl_final(const Scal* s, double h, double rho_m, double rho_p, double mu_m, double mu_p,
                            double sigma, double cfl_number, double* dt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  *dt = s->err ? __longlong_as_double(0x7ff8000000000000LL)
               : cfl_step(s->red[0], s->red[1], s->red[2], h, rho_m, rho_p, mu_m, mu_p, sigma, cfl_number);
}

// Clears the error flag of s (a fill kernel, so that graph capture records no memset node).
__global__ void k_clear_err(Scal* s) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s->err = 0;
}

// Sum of p = hi + lo into s->red[3] (for the mean); 1-D grid-stride launch.
__global__ void __launch_bounds__(BLOCK) k_psum(size_t N, const double* __restrict__ ph, const float* __restrict__ pl,
                                                double* partial, Scal* s) {
  double acc[1] = {0.0};
  for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < N; k += (size_t)gridDim.x * BLOCK) acc[0] += ph[k] + (double)pl[k];
  block_partial<1>(acc, 0u, partial);
}

// out = p - mean (mean_zero) or p - p(0,0) (the reference's pin, POISSON_SOLVER.cpp:789-807).
__global__ void __launch_bounds__(BLOCK) k_shift(size_t N, const double* __restrict__ ph, const float* __restrict__ pl,
                                                 double* __restrict__ out, int mean_zero, const Scal* s) {
#pragma clang fp contract(off)
  const size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k >= N) return;
  if (mean_zero) out[k] = (ph[k] - s->red[3] / (double)N) + (double)pl[k];
  else out[k] = (ph[k] - ph[0]) + ((double)pl[k] - (double)pl[0]);
}

// --------------------------------------------------------------------------- adjoint
// b' = -h^2 D(B g) with g taken on interior faces only; r0 = b' (no jump) as R;
// block partials as k_rhs.
template <class R>
__global__ void __launch_bounds__(BLOCK) k_adj_rhs(int nx, int ny, double h, const double* __restrict__ gu,
                                                   const double* __restrict__ gv, const double* __restrict__ bw,
                                                   const double* __restrict__ bs, double* __restrict__ b,
                                                   R* __restrict__ r0, double* partial) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h, h2 = h * h;
  const Coef<double> C{bw, bs, nx, ny};
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  P2D_BAND(ny);
  for (int j = j0; j < j1 && i < nx; ++j) {
    const size_t k = ci(i, j, nx);
    const double gl = i > 0 ? C.W(i, j) * gu[fxi(i, j, nx)] : 0.0;
    const double gr = i + 1 < nx ? C.E(i, j) * gu[fxi(i + 1, j, nx)] : 0.0;
    const double gb = j > 0 ? C.S(i, j) * gv[fyi(i, j, nx)] : 0.0;
    const double gt = j + 1 < ny ? C.N(i, j) * gv[fyi(i, j + 1, nx)] : 0.0;
    const double bv = h2 * (-((gr - gl) * ih + (gt - gb) * ih));
    b[k] = bv;
    r0[k] = (R)bv;
    acc[0] += bv * bv;
    acc[1] = fmax(acc[1], fabs(bv));
    acc[2] += bv;
    acc[3] += bv;
  }
  block_partial<4>(acc, 2u, partial);
}

// du = g - G mu, dbeta = -dt du (G p + s a / h) on interior faces; walls 0.
__global__ void __launch_bounds__(BLOCK) k_adj_faces(int nx, int ny, double h, double dt,
                                                     const double* __restrict__ gu, const double* __restrict__ gv,
                                                     const double* __restrict__ muh, const float* __restrict__ mul,
                                                     const double* __restrict__ ph, const float* __restrict__ pl,
                                                     const double* __restrict__ saw, const double* __restrict__ sas,
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
        t = gu[f] - fdiff(muh[k], mul[k], muh[k - 1], mul[k - 1], 0.0) * ih;
        const double w = fdiff(ph[k], pl[k], ph[k - 1], pl[k - 1], saw ? saw[k] : 0.0) * ih;
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
        t = gv[f] - fdiff(muh[k], mul[k], muh[k - nx], mul[k - nx], 0.0) * ih;
        const double w = fdiff(ph[k], pl[k], ph[k - nx], pl[k - nx], sas ? sas[k] : 0.0) * ih;
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
