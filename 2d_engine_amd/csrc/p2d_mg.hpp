// p2d multigrid preconditioner and Krylov kernels (templated on the storage
// type T: float for the mixed path, double for the FP64 path; reductions and
// the Krylov scalars are always FP64).
//
// One preconditioned CG iteration is
//   pre(level 0)       [x += a p, r -= a q of the previous iteration, ||r||]
//                      + nu red-black sweeps from 0 + residual + restriction
//   pre(level 1..l0-1)  the same, without the update
//   small              levels l0..L-1 in one workgroup, in LDS, bottom pseudo-inverse
//   post(l0-1..1)      prolongation + nu black-red sweeps
//   post(level 0)      the same, writes z and reduces r.z, sum z, sum r
//   dir                p = (z - mean z) + beta p_old, q = A0' p, p.q -> alpha
// The pre/post kernels are temporally blocked: a 64-column region (one
// wavefront wide) of TY output rows plus a halo of H cells is staged once, the
// sweeps run on LDS, and only the tile interior is written back.  Halo widths:
// pre H = 2 nu + 1 (first red sweep from zero needs no neighbours; each later
// half-sweep and the residual lose one exact ring; restriction reads one ring
// outside the tile), post H = 2 nu.
#pragma once
#include "p2d_common.hpp"

namespace p2d {

constexpr int TY = 32;                     // output rows per pre/post tile
constexpr int MAX_NU = 3;                  // sweeps per side supported
constexpr int RYMAX = TY + 2 * (2 * MAX_NU + 1);  // region rows at most
constexpr int MAXR = (RYMAX + 3) / 4;      // region rows per wavefront at most
constexpr int SMALL_MAX_CELLS = 1024;      // top level of the single-workgroup V-cycle (32^2)

template <class T>
struct Level {
  int nx, ny;
  const T* bw;  // west-face coefficient per cell
  const T* bs;  // south-face coefficient per cell
  const T* wx;  // parent interpolation weight per cell (x, y); unused on the bottom level
  const T* wy;
  T* xa;  // pre-smoothed correction
  T* xb;  // post-smoothed correction (the coarse correction read by the finer level)
  T* r;   // right-hand side (restricted residual)
};

// 1-D weight of coarse index I for fine index i with parent weight w (n = coarse size).
template <class T>
__device__ __forceinline__ T w1(int i, int I, int n, T w) {
  const int P = i >> 1;
  int Q = P + ((i & 1) ? 1 : -1);
  Q = Q < 0 ? 0 : (Q > n - 1 ? n - 1 : Q);
  T s = 0;
  if (P == I) s += w;
  if (Q == I) s += (T)1 - w;
  return s;
}

// Operator-dependent parent weight of fine index i along a row/column whose face
// coefficients are face(k), k = 0..nf (walls 0): the resistance ratio
// R_far / (R_near + R_far) between the parent and the neighbouring coarse
// centre; 3/4 for constant coefficients, 1 at a wall.
template <class T, class F>
__device__ __forceinline__ T od_weight(F face, int i, int nf) {
  const int P = i >> 1;
  const T near_ = (T)0.5 / face(2 * P + 1);
  T far_;
  if ((i & 1) == 0) {
    if (P == 0) return (T)1;
    far_ = (T)1 / face(2 * P) + (T)0.5 / face(2 * P - 1);
  } else {
    if (2 * P + 2 >= nf) return (T)1;
    far_ = (T)1 / face(2 * P + 2) + (T)0.5 / face(2 * P + 3);
  }
  const T big = (T)1e30;
  if (!(near_ < big) || !(far_ < big)) return far_ >= near_ ? (T)1 : (T)0;
  return far_ / (near_ + far_);
}

// --------------------------------------------------------------------------- multigrid setup (FP64 masters)
// Parent weights of every cell of a fine level, computed in FP64, stored as T.
template <class T>
__global__ void __launch_bounds__(BLOCK) k_mg_weights(int nx, int ny, const double* __restrict__ bw,
                                                      const double* __restrict__ bs, T* __restrict__ wx,
                                                      T* __restrict__ wy) {
  const int i = (int)(blockIdx.x * BLOCK + threadIdx.x), j = (int)blockIdx.y;
  if (i >= nx) return;
  auto fx = [&](int k) { return k < nx ? bw[ci(k, j, nx)] : 0.0; };
  auto fy = [&](int k) { return k < ny ? bs[ci(i, k, nx)] : 0.0; };
  double a = od_weight<double>(fx, i, nx);
  double b = od_weight<double>(fy, j, ny);
  const int Q = (i >> 1) + ((i & 1) ? 1 : -1), Qj = (j >> 1) + ((j & 1) ? 1 : -1);
  if (Q < 0 || Q > (nx >> 1) - 1) a = 1.0;
  if (Qj < 0 || Qj > (ny >> 1) - 1) b = 1.0;
  wx[ci(i, j, nx)] = (T)a;
  wy[ci(i, j, nx)] = (T)b;
}

// Coarse face = average over the two fine rows of the series conductance
// 2 / (0.5/b_{k-1} + 1/b_k + 0.5/b_{k+1}) between the two coarse centres.
template <class F>
__device__ __forceinline__ double coarse_face(F face, int k, int nf) {
  if (k <= 0 || k >= nf) return 0.0;
  const double r = 0.5 / face(k - 1) + 1.0 / face(k) + 0.5 / face(k + 1);
  return 2.0 / r;
}
// Grid: dim3(ceil(cnx/BLOCK), cny).
__global__ void __launch_bounds__(BLOCK) k_mg_coarsen(int fnx, int fny, const double* __restrict__ fbw,
                                                      const double* __restrict__ fbs, double* __restrict__ cbw,
                                                      double* __restrict__ cbs) {
  const int cnx = fnx >> 1;  // the grid covers the fny / 2 coarse rows
  const int I = (int)(blockIdx.x * BLOCK + threadIdx.x), J = (int)blockIdx.y;
  if (I >= cnx) return;
  auto row = [&](int j) { return [=](int k) { return k < fnx ? fbw[ci(k, j, fnx)] : 0.0; }; };
  auto col = [&](int i) { return [=](int k) { return k < fny ? fbs[ci(i, k, fnx)] : 0.0; }; };
  cbw[ci(I, J, cnx)] = 0.5 * (coarse_face(row(2 * J), 2 * I, fnx) + coarse_face(row(2 * J + 1), 2 * I, fnx));
  cbs[ci(I, J, cnx)] = 0.5 * (coarse_face(col(2 * I), 2 * J, fny) + coarse_face(col(2 * I + 1), 2 * J, fny));
}

// Pseudo-inverse of the singular bottom operator A (null space = constants):
// A+ = (A + a 1 1^T)^-1 - 1 1^T / (a n^2) with a n = mean diagonal; the shifted
// matrix is SPD, inverted by Gauss-Jordan without pivoting in one workgroup
// (work is n x 2n in global memory).
__global__ void __launch_bounds__(SMALL_THREADS) k_mg_bottom_inverse(int nx, int ny, const double* __restrict__ bw,
                                                                     const double* __restrict__ bs, double* work,
                                                                     double* inv64, float* inv32) {
  const int n = nx * ny, w = 2 * n, tid = threadIdx.x;
  const Coef<double> C{bw, bs, nx, ny};
  __shared__ double alpha_sh;
  if (tid == 0) {
    double dsum = 0;
    for (int c = 0; c < n; ++c) {
      const int i = c % nx, j = c / nx;
      dsum += C.W(i, j) + C.E(i, j) + C.S(i, j) + C.N(i, j);
    }
    alpha_sh = dsum / n / n;
  }
  __syncthreads();
  const double alpha = alpha_sh;
  for (int q = tid; q < n * w; q += SMALL_THREADS) work[q] = (q % w) < n ? alpha : 0.0;
  __syncthreads();
  for (int c = tid; c < n; c += SMALL_THREADS) {
    const int i = c % nx, j = c / nx;
    double* row = work + (size_t)c * w;
    row[c] += C.W(i, j) + C.E(i, j) + C.S(i, j) + C.N(i, j);
    if (i > 0) row[c - 1] -= C.W(i, j);
    if (i + 1 < nx) row[c + 1] -= C.E(i, j);
    if (j > 0) row[c - nx] -= C.S(i, j);
    if (j + 1 < ny) row[c + nx] -= C.N(i, j);
    row[n + c] = 1.0;
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const double d = 1.0 / work[(size_t)k * w + k];
    __syncthreads();
    for (int c = tid; c < w; c += SMALL_THREADS) work[(size_t)k * w + c] *= d;
    __syncthreads();
    for (int q = tid; q < n * w; q += SMALL_THREADS) {
      const int r = q / w, c = q % w;
      if (r == k || c == k) continue;
      work[q] -= work[(size_t)r * w + k] * work[(size_t)k * w + c];
    }
    __syncthreads();
    for (int r = tid; r < n; r += SMALL_THREADS)
      if (r != k) work[(size_t)r * w + k] = 0.0;
    __syncthreads();
  }
  const double corr = 1.0 / (alpha * n * (double)n);
  for (int q = tid; q < n * n; q += SMALL_THREADS) {
    const double v = work[(size_t)(q / n) * w + n + (q % n)] - corr;
    inv64[q] = v;
    inv32[q] = (float)v;
  }
}

// --------------------------------------------------------------------------- solver state kernels
// Start an inner solve from the residual norms left in s->red by the residual kernel.
// The replaced residual (red[0] = sum r^2, red[1] = max |r|, red[2] = sum r) is
// taken without its mean m: rshift = m is subtracted by the next fused update
// and the segment starts from sum r^2 - N m^2.
__device__ __forceinline__ void segment_start(Scal* s, double ncells) {
  const double m = s->red[2] / ncells;
  s->rshift = m;
  s->r0_2 = fmax(s->red[0] - ncells * m * m, 0.0);
  s->r2 = s->r0_2;
  s->rinf = s->red[1];
}
__global__ void k_inner_init(Scal* s, double ncells, double rel2, double abs2, double ainf, int max_iter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  segment_start(s, ncells);
  s->rel2 = rel2;
  s->abs2 = abs2;
  s->ainf = ainf;
  s->max_iter = max_iter;
  s->iter = 0;
  s->seg0 = 0;
  s->flag = 0;
  s->pending = 0;
  s->rpar = 0;
  s->rz = s->rz_old = s->pq = s->alpha = s->beta = s->zmean = 0.0;
  const bool met = (s->r2 <= abs2 && s->rinf <= ainf) || s->r0_2 == 0.0;
  s->done = (met || max_iter <= 0) ? 1 : 0;
  if (!(s->r2 == s->r2)) {
    s->done = 1;
    s->flag = 1;
  }
}
// Continue after a residual replacement: new segment start, keep the iteration
// count; keep the direction and rz_old (CG with residual replacement) or, with
// restart, begin a new direction sequence (beta = 0).  FP32 segments restart:
// their recursive residual drifts from the FP64 one, and the stale rz_old would
// give a beta that lets the old direction dominate.
__global__ void k_inner_continue(Scal* s, double ncells, double rel2, double abs2, double ainf, int restart) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (restart) s->seg0 = s->iter;
  segment_start(s, ncells);
  s->rel2 = rel2;
  s->abs2 = abs2;
  s->ainf = ainf;
  s->pending = 0;
  const bool met = (s->r2 <= abs2 && s->rinf <= ainf) || s->r0_2 == 0.0;
  s->done = (met || s->iter >= s->max_iter) ? 1 : 0;
}

// Shared staging of a pre/post tile: every thread owns column c of the 64-wide
// region and rows w, w+4, ... ; it keeps r, the four face coefficients and
// 1/diag of its cells in registers.
template <class T>
struct TileCells {
  T rr[MAXR], bW[MAXR], bE[MAXR], bS[MAXR], bN[MAXR], inv[MAXR];
  bool val[MAXR];
};

template <class T>
__device__ __forceinline__ void tile_sweep(TileCells<T>& tc, T (*xs)[66], int c, int w, int RY, int gi, int gy0,
                                           int color, bool zero) {
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + 4 * m;
    if (r >= RY || !tc.val[m] || ((gi + gy0 + r) & 1) != color) continue;
    const T off = zero ? (T)0
                       : tc.bW[m] * xs[r + 1][c] + tc.bE[m] * xs[r + 1][c + 2] + tc.bS[m] * xs[r][c + 1] +
                             tc.bN[m] * xs[r + 2][c + 1];
    xs[r + 1][c + 1] = (tc.rr[m] + off) * tc.inv[m];
  }
}

// --------------------------------------------------------------------------- pre-smoothing tile kernel
// Grid dim3(ceil(nx/TX), ceil(ny/TY)), TX = 64 - 2H, H = 2 nu + 1.
// FUSE (finest level): r_in = rbuf[s->rpar], the pending update r -= a q,
// x += a p is applied first (tile interiors write r to rbuf[1 - rpar] and x),
// ||r|| is reduced and the convergence decision taken.
template <class T, bool FUSE>
__global__ void __launch_bounds__(BLOCK) k_mg_pre(Level<T> F, Level<T> Cc, const T* r_lvl, T* x_lvl, int nu,
                                                  T* rA, T* rB, T* xsol, const T* pA, const T* pB,
                                                  const T* __restrict__ qv, double* partial, unsigned* counter,
                                                  Scal* s) {
  if (s->done) return;
  __shared__ T xs[RYMAX + 2][66];
  __shared__ T rs[RYMAX][64];
  const int H = 2 * nu + 1, TX = 64 - 2 * H, RY = TY + 2 * H;
  const int tx0 = (int)blockIdx.x * TX, ty0 = (int)blockIdx.y * TY;
  const int gx0 = tx0 - H, gy0 = ty0 - H;
  const int c = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int gi = gx0 + c;
  const int nx = F.nx, ny = F.ny;
  const T* rin = r_lvl;
  T* rout = nullptr;
  const T* pdir = nullptr;
  double alpha = 0.0, shift = 0.0;
  bool pend = false;
  if (FUSE) {
    const int rp = s->rpar;
    rin = rp ? rB : rA;
    rout = rp ? rA : rB;
    pend = s->pending != 0;
    alpha = s->alpha;
    shift = s->rshift;
    pdir = ((s->iter + 1) & 1) ? pB : pA;
  }
  double acc[2] = {0.0, 0.0};
  TileCells<T> tc;
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + 4 * m, gj = gy0 + r;
    const bool v = r < RY && gi >= 0 && gi < nx && gj >= 0 && gj < ny;
    tc.val[m] = v;
    T rv = 0, bw = 0, be = 0, bsv = 0, bn = 0;
    if (v) {
      const size_t k = ci(gi, gj, nx);
      rv = rin[k];
      if (FUSE) rv = (T)(((double)rv - (pend ? alpha * (double)qv[k] : 0.0)) - shift);
      if (FUSE && c >= H && c < H + TX && r >= H && r < H + TY) {
        rout[k] = rv;
        if (pend) xsol[k] = (T)((double)xsol[k] + alpha * (double)pdir[k]);
        acc[0] += (double)rv * (double)rv;
        acc[1] = fmax(acc[1], fabs((double)rv));
      }
      bw = F.bw[k];
      be = gi + 1 < nx ? F.bw[k + 1] : (T)0;
      bsv = F.bs[k];
      bn = gj + 1 < ny ? F.bs[k + nx] : (T)0;
    }
    tc.rr[m] = rv;
    tc.bW[m] = bw;
    tc.bE[m] = be;
    tc.bS[m] = bsv;
    tc.bN[m] = bn;
    tc.inv[m] = v ? (T)1 / (bw + be + bsv + bn) : (T)0;
    if (r < RY) xs[r + 1][c + 1] = 0;
  }
  for (int t = threadIdx.x; t < 66; t += BLOCK) {
    xs[0][t] = 0;
    xs[RY + 1][t] = 0;
  }
  for (int t = threadIdx.x; t < RY; t += BLOCK) {
    xs[t + 1][0] = 0;
    xs[t + 1][65] = 0;
  }
  __syncthreads();
  for (int sw = 0; sw < nu; ++sw)
    for (int color = 0; color < 2; ++color) {
      tile_sweep<T>(tc, xs, c, w, RY, gi, gy0, color, sw == 0 && color == 0);
      __syncthreads();
    }
  // residual (reference order: r - (diag x - offsum)) and the interior write of x
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + 4 * m;
    if (r >= RY) continue;
    T res = 0;
    if (tc.val[m]) {
      const T x = xs[r + 1][c + 1];
      const T off =
          tc.bW[m] * xs[r + 1][c] + tc.bE[m] * xs[r + 1][c + 2] + tc.bS[m] * xs[r][c + 1] + tc.bN[m] * xs[r + 2][c + 1];
      res = tc.rr[m] - ((tc.bW[m] + tc.bE[m] + tc.bS[m] + tc.bN[m]) * x - off);
      if (c >= H && c < H + TX && r >= H && r < H + TY) x_lvl[ci(gi, gy0 + r, nx)] = x;
    }
    rs[r][c] = res;
  }
  __syncthreads();
  // restriction r_H = P^T res over the coarse cells of this tile
  const int NCX = TX / 2, NCY = TY / 2, cx0 = tx0 / 2, cy0 = ty0 / 2;
  for (int t = threadIdx.x; t < NCX * NCY; t += BLOCK) {
    const int I = cx0 + t % NCX, J = cy0 + t / NCX;
    if (I >= Cc.nx || J >= Cc.ny) continue;
    T sum = 0;
    for (int j = 2 * J - 1; j <= 2 * J + 2; ++j) {
      if (j < 0 || j >= ny) continue;
      for (int i = 2 * I - 1; i <= 2 * I + 2; ++i) {
        if (i < 0 || i >= nx) continue;
        const size_t kk = ci(i, j, nx);
        const T wy = w1<T>(j, J, Cc.ny, F.wy[kk]);
        const T wx = w1<T>(i, I, Cc.nx, F.wx[kk]);
        if (wx == (T)0 || wy == (T)0) continue;
        sum += wx * wy * rs[j - gy0][i - gx0];
      }
    }
    Cc.r[ci(I, J, Cc.nx)] = sum;
  }
  if (FUSE) {
    if (grid_reduce<2>(acc, 2u, partial, counter) && threadIdx.x == 0) {
      s->rpar = 1 - s->rpar;
      s->rshift = 0.0;
      if (pend) {
        s->r2 = acc[0];
        s->rinf = acc[1];
        s->rz_old = s->rz;
        s->iter += 1;
        const bool met = acc[0] <= s->rel2 * s->r0_2 || (acc[0] <= s->abs2 && acc[1] <= s->ainf);
        if (!(acc[0] == acc[0])) s->flag = 1;
        if (met || s->iter >= s->max_iter || s->flag) s->done = 1;
        s->pending = 0;
      }
    }
  }
}

// --------------------------------------------------------------------------- post-smoothing tile kernel
// Grid dim3(ceil(nx/TX), ceil(ny/TY)), TX = 64 - 2H, H = 2 nu.
// x = x_pre + P x_coarse, then nu (black, red) sweeps; the tile interior is
// written to x_out.  LEVEL0: r = rbuf[s->rpar], x_out = z, and r.z, sum z,
// sum r are reduced: zmean, rz = r.(z - zmean), beta.
template <class T, bool LEVEL0>
__global__ void __launch_bounds__(BLOCK) k_mg_post(Level<T> F, Level<T> Cc, const T* r_lvl, const T* rA, const T* rB,
                                                   const T* x_in, T* x_out, int nu, double* partial,
                                                   unsigned* counter, Scal* s) {
  if (s->done) return;
  __shared__ T xs[RYMAX + 2][66];
  const int H = 2 * nu, TX = 64 - 2 * H, RY = TY + 2 * H;
  const int tx0 = (int)blockIdx.x * TX, ty0 = (int)blockIdx.y * TY;
  const int gx0 = tx0 - H, gy0 = ty0 - H;
  const int c = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int gi = gx0 + c;
  const int nx = F.nx, ny = F.ny;
  const T* rin = LEVEL0 ? (s->rpar ? rB : rA) : r_lvl;
  TileCells<T> tc;
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + 4 * m, gj = gy0 + r;
    const bool v = r < RY && gi >= 0 && gi < nx && gj >= 0 && gj < ny;
    tc.val[m] = v;
    T rv = 0, bw = 0, be = 0, bsv = 0, bn = 0, x0 = 0;
    if (v) {
      const size_t k = ci(gi, gj, nx);
      rv = rin[k];
      bw = F.bw[k];
      be = gi + 1 < nx ? F.bw[k + 1] : (T)0;
      bsv = F.bs[k];
      bn = gj + 1 < ny ? F.bs[k + nx] : (T)0;
      // prolongation of the coarse correction
      const int I = gi >> 1, J = gj >> 1;
      int I2 = I + ((gi & 1) ? 1 : -1), J2 = J + ((gj & 1) ? 1 : -1);
      I2 = I2 < 0 ? 0 : (I2 > Cc.nx - 1 ? Cc.nx - 1 : I2);
      J2 = J2 < 0 ? 0 : (J2 > Cc.ny - 1 ? Cc.ny - 1 : J2);
      const T wx = F.wx[k], wy = F.wy[k];
      const T* xc = Cc.xb;
      const T e = wx * wy * xc[ci(I, J, Cc.nx)] + ((T)1 - wx) * wy * xc[ci(I2, J, Cc.nx)] +
                  wx * ((T)1 - wy) * xc[ci(I, J2, Cc.nx)] + ((T)1 - wx) * ((T)1 - wy) * xc[ci(I2, J2, Cc.nx)];
      x0 = x_in[k] + e;
    }
    tc.rr[m] = rv;
    tc.bW[m] = bw;
    tc.bE[m] = be;
    tc.bS[m] = bsv;
    tc.bN[m] = bn;
    tc.inv[m] = v ? (T)1 / (bw + be + bsv + bn) : (T)0;
    if (r < RY) xs[r + 1][c + 1] = x0;
  }
  for (int t = threadIdx.x; t < 66; t += BLOCK) {
    xs[0][t] = 0;
    xs[RY + 1][t] = 0;
  }
  for (int t = threadIdx.x; t < RY; t += BLOCK) {
    xs[t + 1][0] = 0;
    xs[t + 1][65] = 0;
  }
  __syncthreads();
  for (int sw = 0; sw < nu; ++sw)
    for (int cc = 0; cc < 2; ++cc) {
      tile_sweep<T>(tc, xs, c, w, RY, gi, gy0, 1 - cc, false);
      __syncthreads();
    }
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + 4 * m;
    if (r >= RY || !tc.val[m] || c < H || c >= H + TX || r < H || r >= H + TY) continue;
    const T z = xs[r + 1][c + 1];
    x_out[ci(gi, gy0 + r, nx)] = z;
    if (LEVEL0) {
      acc[0] += (double)tc.rr[m] * (double)z;
      acc[1] += (double)z;
      acc[2] += (double)tc.rr[m];
    }
  }
  if (LEVEL0) {
    if (grid_reduce<3>(acc, 0u, partial, counter) && threadIdx.x == 0) {
      const double N = (double)nx * (double)ny;
      const double zm = acc[1] / N;
      const double rz = acc[0] - zm * acc[2];
      s->zmean = zm;
      s->rz = rz;
      s->beta = s->iter == s->seg0 ? 0.0 : rz / s->rz_old;
    }
  }
}

// --------------------------------------------------------------------------- single-workgroup coarse V-cycle
// Levels l0..L-1 (l0 >= 1; at most SMALL_MAX_CELLS cells on l0) in one
// workgroup: r of level l0 and the coefficients of every level are staged in
// LDS, the weights are recomputed from them, the bottom uses the dense
// pseudo-inverse binv (global), and the result is written to Ls[l0].xb.
template <class T>
__global__ void __launch_bounds__(SMALL_THREADS) k_mg_small(const Level<T>* __restrict__ Ls, int l0, int L, int nu_pre,
                                                            int nu_post, const T* __restrict__ binv, const Scal* s) {
  if (s->done) return;
  extern __shared__ unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x;
  // LDS layout per level: x, r, bw, bs (n_l each)
  T* X[16];
  T* R[16];
  T* BW[16];
  T* BS[16];
  {
    size_t off = 0;
    for (int l = l0; l < L; ++l) {
      const size_t n = (size_t)Ls[l].nx * Ls[l].ny;
      X[l] = smem + off;
      R[l] = smem + off + n;
      BW[l] = smem + off + 2 * n;
      BS[l] = smem + off + 3 * n;
      off += 4 * n;
    }
  }
  for (int l = l0; l < L; ++l) {
    const int n = Ls[l].nx * Ls[l].ny;
    for (int t = tid; t < n; t += SMALL_THREADS) {
      BW[l][t] = Ls[l].bw[t];
      BS[l][t] = Ls[l].bs[t];
      if (l == l0) R[l][t] = Ls[l].r[t];
    }
  }
  __syncthreads();
  auto diag = [&](int l, int i, int j) {
    const int nx = Ls[l].nx, ny = Ls[l].ny;
    const T e = i + 1 < nx ? BW[l][i + 1 + j * nx] : (T)0;
    const T nn = j + 1 < ny ? BS[l][i + (j + 1) * nx] : (T)0;
    return BW[l][i + j * nx] + e + BS[l][i + j * nx] + nn;
  };
  auto offs = [&](int l, int i, int j) {
    const int nx = Ls[l].nx, ny = Ls[l].ny, k = i + j * nx;
    const T* x = X[l];
    T o = 0;
    if (i > 0) o += BW[l][k] * x[k - 1];
    if (i + 1 < nx) o += BW[l][k + 1] * x[k + 1];
    if (j > 0) o += BS[l][k] * x[k - nx];
    if (j + 1 < ny) o += BS[l][k + nx] * x[k + nx];
    return o;
  };
  auto sweep = [&](int l, int color, bool zero) {
    const int nx = Ls[l].nx, ny = Ls[l].ny, half = (nx + 1) >> 1;
    for (int t = tid; t < half * ny; t += SMALL_THREADS) {
      const int j = t / half, i = 2 * (t % half) + ((j + color) & 1);
      if (i >= nx) continue;
      const int k = i + j * nx;
      const T o = zero ? (T)0 : offs(l, i, j);
      X[l][k] = (R[l][k] + o) / diag(l, i, j);
    }
    __syncthreads();
  };
  auto wts = [&](int l, int i, int j, T& wx, T& wy) {
    const int nx = Ls[l].nx, ny = Ls[l].ny;
    auto fx = [&](int k) { return k < nx ? BW[l][k + j * nx] : (T)0; };
    auto fy = [&](int k) { return k < ny ? BS[l][i + k * nx] : (T)0; };
    wx = od_weight<T>(fx, i, nx);
    wy = od_weight<T>(fy, j, ny);
    const int Q = (i >> 1) + ((i & 1) ? 1 : -1), Qj = (j >> 1) + ((j & 1) ? 1 : -1);
    if (Q < 0 || Q > (nx >> 1) - 1) wx = (T)1;
    if (Qj < 0 || Qj > (ny >> 1) - 1) wy = (T)1;
  };
  for (int l = l0; l < L - 1; ++l) {
    for (int sw = 0; sw < nu_pre; ++sw) {
      sweep(l, 0, sw == 0);
      sweep(l, 1, false);
    }
    const int nx = Ls[l].nx, ny = Ls[l].ny, cnx = Ls[l + 1].nx, cny = Ls[l + 1].ny;
    for (int t = tid; t < cnx * cny; t += SMALL_THREADS) {
      const int I = t % cnx, J = t / cnx;
      T sum = 0;
      for (int j = 2 * J - 1; j <= 2 * J + 2; ++j) {
        if (j < 0 || j >= ny) continue;
        for (int i = 2 * I - 1; i <= 2 * I + 2; ++i) {
          if (i < 0 || i >= nx) continue;
          T wx, wy;
          wts(l, i, j, wx, wy);
          const T a = w1<T>(j, J, cny, wy), b = w1<T>(i, I, cnx, wx);
          if (a == (T)0 || b == (T)0) continue;
          const int k = i + j * nx;
          const T res = R[l][k] - (diag(l, i, j) * X[l][k] - offs(l, i, j));
          sum += b * a * res;
        }
      }
      R[l + 1][t] = sum;
    }
    __syncthreads();
  }
  {
    const int n = Ls[L - 1].nx * Ls[L - 1].ny;
    for (int a = tid; a < n; a += SMALL_THREADS) {
      double acc = 0;
      for (int q = 0; q < n; ++q) acc += (double)binv[(size_t)a * n + q] * (double)R[L - 1][q];
      X[L - 1][a] = (T)acc;
    }
    __syncthreads();
  }
  for (int l = L - 2; l >= l0; --l) {
    const int nx = Ls[l].nx, ny = Ls[l].ny, cnx = Ls[l + 1].nx, cny = Ls[l + 1].ny;
    for (int t = tid; t < nx * ny; t += SMALL_THREADS) {
      const int i = t % nx, j = t / nx;
      T wx, wy;
      wts(l, i, j, wx, wy);
      const int I = i >> 1, J = j >> 1;
      int I2 = I + ((i & 1) ? 1 : -1), J2 = J + ((j & 1) ? 1 : -1);
      I2 = I2 < 0 ? 0 : (I2 > cnx - 1 ? cnx - 1 : I2);
      J2 = J2 < 0 ? 0 : (J2 > cny - 1 ? cny - 1 : J2);
      const T* xc = X[l + 1];
      X[l][t] += wx * wy * xc[I + J * cnx] + ((T)1 - wx) * wy * xc[I2 + J * cnx] + wx * ((T)1 - wy) * xc[I + J2 * cnx] +
                 ((T)1 - wx) * ((T)1 - wy) * xc[I2 + J2 * cnx];
    }
    __syncthreads();
    for (int sw = 0; sw < nu_post; ++sw) {
      sweep(l, 1, false);
      sweep(l, 0, false);
    }
  }
  const int n0 = Ls[l0].nx * Ls[l0].ny;
  for (int t = tid; t < n0; t += SMALL_THREADS) Ls[l0].xb[t] = X[l0][t];
}

// --------------------------------------------------------------------------- Krylov direction kernel
// p = (z - mean z) + beta p_old (FP64 arithmetic, stored as T; p_old/p_new =
// pbuf[iter & 1] / pbuf[(iter+1) & 1]), q = A0' p, alpha = rz / p.q; marks the
// update pending.  Band kernel marching down rows with the rows above/below in
// registers.  Grid dim3(ceil(nx/BLOCK), ceil(ny/BAND)).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_dir(int nx, int ny, const T* __restrict__ bw, const T* __restrict__ bs,
                                               const T* __restrict__ z, T* pA, T* pB, T* __restrict__ q,
                                               double* partial, unsigned* counter, Scal* s) {
  if (s->done) return;
  const double beta = s->beta, zm = s->zmean;
  const bool first = s->iter == s->seg0;
  const T* pold = (s->iter & 1) ? pB : pA;
  T* pnew = (s->iter & 1) ? pA : pB;
  const Coef<T> C{bw, bs, nx, ny};
  const int i = (int)(blockIdx.x * BLOCK + threadIdx.x);
  const int j0 = (int)(blockIdx.y * BAND), j1 = min(ny, j0 + BAND);
  auto P = [&](size_t k) {
    return first ? (T)((double)z[k] - zm) : (T)(((double)z[k] - zm) + beta * (double)pold[k]);
  };
  double acc[1] = {0.0};
  if (i < nx && j0 < j1) {
    T ps = j0 > 0 ? P(ci(i, j0 - 1, nx)) : (T)0;
    T pc = P(ci(i, j0, nx));
    T pn = 0;
    for (int j = j0; j < j1; ++j) {
      const size_t k = ci(i, j, nx);
      if (j + 1 < ny) pn = P(k + nx);
      T a = 0;
      if (i > 0) a += C.W(i, j) * (pc - P(k - 1));
      if (i + 1 < nx) a += C.E(i, j) * (pc - P(k + 1));
      if (j > 0) a += C.S(i, j) * (pc - ps);
      if (j + 1 < ny) a += C.N(i, j) * (pc - pn);
      pnew[k] = pc;
      q[k] = a;
      acc[0] += (double)pc * (double)a;
      ps = pc;
      pc = pn;
    }
  }
  if (grid_reduce<1>(acc, 0u, partial, counter) && threadIdx.x == 0) {
    s->pq = acc[0];
    if (!(acc[0] > 0.0)) {
      s->done = 1;
      s->flag = 1;
      s->alpha = 0.0;
    } else {
      s->alpha = s->rz / acc[0];
      s->pending = 1;
    }
  }
}

}  // namespace p2d
