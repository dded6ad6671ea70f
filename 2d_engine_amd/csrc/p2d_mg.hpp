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
// The pre/post kernels are temporally blocked: a 128-column region (one
// wavefront wide, two columns per lane) of TY output rows plus a halo of H
// cells is staged once, the sweeps run on LDS, and only the tile interior is
// written back.  Every global
// load of a tile is independent of the others, so a tile costs about one
// memory round trip; coarse levels (small grids) are latency-bound.  Halo widths:
// pre H = 2 nu + 1 (first red sweep from zero needs no neighbours; each later
// half-sweep and the residual lose one exact ring; restriction reads one ring
// outside the tile), post H = 2 nu.
#pragma once
#include "p2d_common.hpp"

namespace p2d {

constexpr int TBLOCK = 1024;               // pre/post tile kernels: 16 wavefronts
constexpr int TWAVES = TBLOCK / 64;
constexpr int TY = 32;                     // output rows per pre/post tile
constexpr int MAX_NU = 3;                  // sweeps per side supported
constexpr int RYMAX = TY + 2 * (2 * MAX_NU + 1);  // region rows at most
constexpr int MAXR = (RYMAX + TWAVES - 1) / TWAVES;  // region rows per thread at most
constexpr int NCXMAX = 64;                 // coarse columns of a tile at most (TX / 2 < 64)

// Tiles are 128 columns wide; the x halo HX is the smallest even width >= H so
// that the region starts at an even column; TX = 128 - 2 HX interior columns.
__host__ __device__ constexpr int tile_hx(int H) { return H + (H & 1); }
__host__ __device__ constexpr int tile_tx(int H) { return 128 - 2 * tile_hx(H); }
// FP32 tiles keep two blocks resident per CU (<= 64 VGPRs) so that one block's
// loads overlap the other's LDS sweeps; FP64 tiles need the registers.
template <class T>
__host__ __device__ constexpr int tile_min_blocks() { return sizeof(T) == 4 ? 2 : 1; }
constexpr int DTY = 16;                    // rows per block of the direction kernel
constexpr int SMALL_MAX_CELLS = 1024;      // top level of the single-workgroup V-cycle at most (32^2)
constexpr int MAXL = 24;                   // multigrid levels at most
constexpr int BOTTOM_LDS_MAX = 64;         // bottom sizes inverted in LDS

template <class T>
struct Level {
  int nx, ny;
  const T* bw;  // west-face coefficient per cell
  const T* bs;  // south-face coefficient per cell
  const T* wx;  // parent interpolation weight per cell (x, y); none on the bottom level
  const T* wy;
  T* xa;  // pre-smoothed correction
  T* xb;  // post-smoothed correction (the coarse correction read by the finer level)
  T* r;   // right-hand side (restricted residual)
};

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
// Coarse face = average over the two fine rows of the series conductance
// 2 / (0.5/b_{k-1} + 1/b_k + 0.5/b_{k+1}) between the two coarse centres.
template <class F>
__device__ __forceinline__ double coarse_face(F face, int k, int nf) {
  if (k <= 0 || k >= nf) return 0.0;
  const double r = 0.5 / face(k - 1) + 1.0 / face(k) + 0.5 / face(k + 1);
  return 2.0 / r;
}

// One setup sweep over a fine level (FP64 coefficients bw, bs), every output
// optional (NULL): the parent weights of each cell, computed in FP64 and
// stored as T; the FP32 copy of the coefficients; and, from the thread of the
// lower-left child of each coarse cell, the coarse FP64 coefficients.
// Grid dim3(ceil(nx/BLOCK), ny).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_mg_setup(int nx, int ny, const double* __restrict__ bw,
                                                    const double* __restrict__ bs, T* __restrict__ wx,
                                                    T* __restrict__ wy, float* __restrict__ f32bw,
                                                    float* __restrict__ f32bs, double* __restrict__ cbw,
                                                    double* __restrict__ cbs) {
  const int i = (int)(blockIdx.x * BLOCK + threadIdx.x), j = (int)blockIdx.y;
  if (i >= nx) return;
  const size_t k = ci(i, j, nx);
  auto row = [&](int jj) { return [=](int q) { return q < nx ? bw[ci(q, jj, nx)] : 0.0; }; };
  auto col = [&](int ii) { return [=](int q) { return q < ny ? bs[ci(ii, q, nx)] : 0.0; }; };
  if (wx) {
    double a = od_weight<double>(row(j), i, nx);
    double b = od_weight<double>(col(i), j, ny);
    const int Q = (i >> 1) + ((i & 1) ? 1 : -1), Qj = (j >> 1) + ((j & 1) ? 1 : -1);
    if (Q < 0 || Q > (nx >> 1) - 1) a = 1.0;
    if (Qj < 0 || Qj > (ny >> 1) - 1) b = 1.0;
    wx[k] = (T)a;
    wy[k] = (T)b;
  }
  if (f32bw) {
    f32bw[k] = (float)bw[k];
    f32bs[k] = (float)bs[k];
  }
  if (cbw && (i & 1) == 0 && (j & 1) == 0) {
    const int I = i >> 1, J = j >> 1, cnx = nx >> 1;
    cbw[ci(I, J, cnx)] = 0.5 * (coarse_face(row(2 * J), 2 * I, nx) + coarse_face(row(2 * J + 1), 2 * I, nx));
    cbs[ci(I, J, cnx)] = 0.5 * (coarse_face(col(2 * I), 2 * J, ny) + coarse_face(col(2 * I + 1), 2 * J, ny));
  }
}

// Pseudo-inverse of the singular bottom operator A (null space = constants):
// A+ = (A + a 1 1^T)^-1 - 1 1^T / (a n^2) with a n = mean diagonal; the shifted
// matrix is SPD and is inverted by Gauss-Jordan without pivoting in one
// workgroup: in place in LDS for n <= BOTTOM_LDS_MAX, else on an n x 2n
// augmented matrix in global memory (work).
__global__ void __launch_bounds__(SMALL_THREADS) k_mg_bottom_inverse(int nx, int ny, const double* __restrict__ bw,
                                                                     const double* __restrict__ bs, double* work,
                                                                     double* inv64, float* inv32) {
  const int n = nx * ny, tid = threadIdx.x;
  const Coef<double> C{bw, bs, nx, ny};
  __shared__ double alpha_sh;
  __shared__ double a[BOTTOM_LDS_MAX * BOTTOM_LDS_MAX];
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
  const double corr = 1.0 / (alpha * n * (double)n);
  if (n <= BOTTOM_LDS_MAX) {
    for (int q = tid; q < n * n; q += SMALL_THREADS) a[q] = alpha;
    __syncthreads();
    for (int c = tid; c < n; c += SMALL_THREADS) {
      const int i = c % nx, j = c / nx;
      double* row = a + c * n;
      row[c] += C.W(i, j) + C.E(i, j) + C.S(i, j) + C.N(i, j);
      if (i > 0) row[c - 1] -= C.W(i, j);
      if (i + 1 < nx) row[c + 1] -= C.E(i, j);
      if (j > 0) row[c - nx] -= C.S(i, j);
      if (j + 1 < ny) row[c + nx] -= C.N(i, j);
    }
    __syncthreads();
    // in-place inversion: pivot row scaled, rank-1 update off row/column k,
    // column k = -a_ik / a_kk, a_kk = 1 / a_kk
    for (int k = 0; k < n; ++k) {
      const double d = 1.0 / a[k * n + k];
      for (int j = tid; j < n; j += SMALL_THREADS)
        if (j != k) a[k * n + j] *= d;
      __syncthreads();
      for (int q = tid; q < n * n; q += SMALL_THREADS) {
        const int r = q / n, c = q % n;
        if (r != k && c != k) a[q] -= a[r * n + k] * a[k * n + c];
      }
      __syncthreads();
      for (int r = tid; r < n; r += SMALL_THREADS)
        if (r != k) a[r * n + k] *= -d;
      if (tid == 0) a[k * n + k] = d;
      __syncthreads();
    }
    for (int q = tid; q < n * n; q += SMALL_THREADS) {
      const double v = a[q] - corr;
      inv64[q] = v;
      inv32[q] = (float)v;
    }
    return;
  }
  const int w = 2 * n;
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
// Start of a solve from the initial residual's norms in red: the reference
// norm b2 (mean-free), the outer targets (rtol on ||r||_2 relative to b2, atol
// on max |r|; <= 0 disables) and the first segment's inner targets: relative
// rel2 and, when inner_abs, absolute ones with a margin for the recursive
// residual (0.25 out2, 0.5 outi).  Sets outer_done when r0 already meets the
// outer targets.
__global__ void k_inner_init(Scal* s, double ncells, double rel2, int inner_abs, double rtol, double atol,
                             int max_iter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double m = s->red[2] / ncells;
  const double b2 = fmax(s->red[0] - ncells * m * m, 0.0);
  s->b2 = b2;
  s->out2 = rtol > 0.0 ? rtol * rtol * b2 : -1.0;
  s->outi = atol > 0.0 ? atol : -1.0;
  segment_start(s, ncells);
  s->rel2 = rel2;
  s->abs2 = inner_abs && s->out2 >= 0.0 ? 0.25 * s->out2 : (inner_abs ? 1e300 : -1.0);
  s->ainf = inner_abs && s->outi >= 0.0 ? 0.5 * s->outi : (inner_abs ? 1e300 : -1.0);
  s->max_iter = max_iter;
  s->iter = 0;
  s->seg0 = 0;
  s->outer_done = 0;
  s->flag = 0;
  s->pending = 0;
  s->rpar = 0;
  s->rz = s->rz_old = s->pq = s->alpha = s->beta = s->zmean = 0.0;
  if (!(s->red[0] == s->red[0])) {
    s->done = 1;
    s->flag = 1;
    return;
  }
  if (((s->out2 < 0.0 || s->red[0] <= s->out2) && (s->outi < 0.0 || s->red[1] <= s->outi)) || b2 == 0.0) {
    s->outer_done = 1;
    s->done = 1;
    return;
  }
  const bool met = (s->r2 <= s->abs2 && s->rinf <= s->ainf) || s->r0_2 == 0.0;
  s->done = (met || max_iter <= 0) ? 1 : 0;
}
// Continue after a residual replacement (red = norms of the true residual):
// stop when the outer targets hold, else start a new segment (same inner
// targets as the first) that keeps the iteration count and
// begins a new direction sequence (beta = 0).  The recursive residual of the
// finished segment may have drifted far from the true one (FP32, or FP64 with
// large pressure levels); with the old direction and rz_old kept, beta would
// be rz_new / rz_old >> 1 and the stale direction would dominate.
__global__ void k_inner_continue(Scal* s, double ncells) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (!(s->red[0] == s->red[0])) {
    s->flag = 1;
    s->done = 1;
    return;
  }
  if ((s->out2 < 0.0 || s->red[0] <= s->out2) && (s->outi < 0.0 || s->red[1] <= s->outi)) {
    s->outer_done = 1;
    s->done = 1;
    return;
  }
  s->seg0 = s->iter;
  segment_start(s, ncells);
  s->pending = 0;
  const bool met = (s->r2 <= s->abs2 && s->rinf <= s->ainf) || s->r0_2 == 0.0;
  s->done = (met || s->iter >= s->max_iter) ? 1 : 0;
}

// --------------------------------------------------------------------------- pair-layout tiles
// A tile region is 128 columns by RY = TY + 2H rows.  Lane c of every
// wavefront holds the column pair (2c, 2c+1): both children of one coarse
// column, since the region starts at an even column.  Wave w holds rows w,
// w + TWAVES, ...  The two cells of a pair have opposite colours, so a
// red-black half-sweep updates exactly one cell per pair and row, picked per
// row (the same for the whole wavefront): no lane idles on the colour test.
template <class T>
struct alignas(2 * sizeof(T)) Pair {
  T a, b;
};

// Per row m of this thread: the pair's right-hand side, face coefficients
// (wA west of A, wB west of B = east of A, eB east of B, s/n south/north),
// 1/diag and validity.
template <class T>
struct PairRows {
  T rA[MAXR], rB[MAXR], wA[MAXR], wB[MAXR], eB[MAXR], sA[MAXR], sB[MAXR], nA[MAXR], nB[MAXR], iA[MAXR], iB[MAXR];
  bool vA[MAXR], vB[MAXR];
};

// Stages row m (region row r, global row gj) of the pair at global columns
// (giA, giA + 1), k = ci(giA, gj, nx) (even: nx is even on every tile level).
template <class T>
__device__ __forceinline__ void stage_coeffs(PairRows<T>& pr, int m, const Level<T>& F, int giA, int gj, bool row_in,
                                             size_t k) {
  const int nx = F.nx, ny = F.ny;
  const bool vA = row_in && giA >= 0 && giA < nx, vB = row_in && giA + 1 >= 0 && giA + 1 < nx;
  pr.vA[m] = vA;
  pr.vB[m] = vB;
  T wA = 0, wB = 0, eB = 0, sA = 0, sB = 0, nA = 0, nB = 0;
  if (vA) {
    const Pair<T> w2 = reinterpret_cast<const Pair<T>*>(F.bw)[k >> 1];
    const Pair<T> s2 = reinterpret_cast<const Pair<T>*>(F.bs)[k >> 1];
    wA = w2.a;
    wB = w2.b;
    sA = s2.a;
    sB = s2.b;
    if (giA + 2 < nx) eB = F.bw[k + 2];
    if (gj + 1 < ny) {
      const Pair<T> n2 = reinterpret_cast<const Pair<T>*>(F.bs)[(k + nx) >> 1];
      nA = n2.a;
      nB = n2.b;
    }
  }
  pr.wA[m] = wA;
  pr.wB[m] = wB;
  pr.eB[m] = eB;
  pr.sA[m] = sA;
  pr.sB[m] = sB;
  pr.nA[m] = nA;
  pr.nB[m] = nB;
  pr.iA[m] = vA ? (T)1 / (wA + wB + sA + nA) : (T)0;
  pr.iB[m] = vB ? (T)1 / (wB + eB + sB + nB) : (T)0;
}

template <class T>
__device__ __forceinline__ void zero_ring(Pair<T> (*xs)[66], int RY) {
  for (int t = threadIdx.x; t < 66; t += TBLOCK) {
    xs[0][t] = Pair<T>{0, 0};
    xs[RY + 1][t] = Pair<T>{0, 0};
  }
  for (int t = threadIdx.x; t < RY; t += TBLOCK) {
    xs[t + 1][0] = Pair<T>{0, 0};
    xs[t + 1][65] = Pair<T>{0, 0};
  }
}

// One red-black half-sweep: cells of one colour read only cells of the other,
// so all neighbour reads are gathered first, then the new values written.
// Cell (i,j) has colour (i + j) & 1; A (even column) has colour (gy0 + r) & 1.
template <class T>
__device__ __forceinline__ void pair_sweep(PairRows<T>& pr, Pair<T> (*xs)[66], int c, int w, int RY, int gy0,
                                           int color, bool zero) {
  T off[MAXR];
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + TWAVES * m;
    off[m] = 0;
    if (zero || r >= RY) continue;
    const Pair<T> own = xs[r + 1][c + 1], up = xs[r + 2][c + 1], dn = xs[r][c + 1];
    if (((gy0 + r) & 1) == color)
      off[m] = pr.wA[m] * xs[r + 1][c].b + pr.wB[m] * own.b + pr.sA[m] * dn.a + pr.nA[m] * up.a;
    else
      off[m] = pr.wB[m] * own.a + pr.eB[m] * xs[r + 1][c + 2].a + pr.sB[m] * dn.b + pr.nB[m] * up.b;
  }
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + TWAVES * m;
    if (r >= RY) continue;
    if (((gy0 + r) & 1) == color) {
      if (pr.vA[m]) xs[r + 1][c + 1].a = (pr.rA[m] + off[m]) * pr.iA[m];
    } else {
      if (pr.vB[m]) xs[r + 1][c + 1].b = (pr.rB[m] + off[m]) * pr.iB[m];
    }
  }
}

// --------------------------------------------------------------------------- pre-smoothing tile kernel
// Grid dim3(ceil(nx/TX), ceil(ny/TY)), TBLOCK threads, H = 2 nu + 1, x halo
// HX = tile_hx(H) (even), TX = 128 - 2 HX.
// FUSE (finest level): r_in = rbuf[s->rpar], the pending update r -= a q,
// x += a p and the residual shift are applied first (tile interiors write r to
// rbuf[1 - rpar] and x), and the block partials of ||r|| are stored.
// Restriction r_H = P^T res: each cell splits its residual into the four
// weight products; the x sums run across lanes, the y sums through LDS.
template <class T, bool FUSE>
__global__ void __launch_bounds__(TBLOCK) __attribute__((amdgpu_waves_per_eu(4 * tile_min_blocks<T>()))) k_mg_pre(Level<T> F, Level<T> Cc, const T* r_lvl, T* x_lvl, int nu,
                                                   T* rA, T* rB, T* xsol, const T* pA, const T* pB,
                                                   const T* __restrict__ qv, double* partial, Scal* s) {
  if (s->done) return;
  __shared__ Pair<T> xs[RYMAX + 2][66];
  const int H = 2 * nu + 1, HX = tile_hx(H), TX = 128 - 2 * HX, RY = TY + 2 * H;
  int bx, by;
  xcd_tile(bx, by);
  const int tx0 = bx * TX, ty0 = by * TY;
  const int gx0 = tx0 - HX, gy0 = ty0 - H;
  const int c = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int giA = gx0 + 2 * c;
  const int nx = F.nx, ny = F.ny;
  const int hc = HX / 2, NCX = TX / 2;
  const bool pair_int = c >= hc && c < hc + NCX;       // both cells in the tile interior
  const bool pair_fp = c >= hc - 1 && c <= hc + NCX;   // restriction footprint
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
  const T ta = (T)alpha, tshift = (T)shift;
  PairRows<T> pr;
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + TWAVES * m, gj = gy0 + r;
    const bool row_in = r < RY && gj >= 0 && gj < ny;
    const size_t k = row_in && giA >= 0 && giA < nx ? ci(giA, gj, nx) : 0;
    stage_coeffs<T>(pr, m, F, giA, gj, row_in, k);
    T ra = 0, rb = 0;
    if (pr.vA[m]) {
      const Pair<T> r2 = reinterpret_cast<const Pair<T>*>(rin)[k >> 1];
      ra = r2.a;
      rb = r2.b;
      if (FUSE) {
        // the inner recursion runs in T: r -= a q, x += a p (a, shift rounded to T)
        if (pend) {
          const Pair<T> q2 = reinterpret_cast<const Pair<T>*>(qv)[k >> 1];
          ra -= ta * q2.a;
          rb -= ta * q2.b;
        }
        ra -= tshift;
        rb -= tshift;
        if (pair_int && r >= H && r < H + TY) {
          reinterpret_cast<Pair<T>*>(rout)[k >> 1] = Pair<T>{ra, rb};
          if (pend) {
            Pair<T> x2 = reinterpret_cast<Pair<T>*>(xsol)[k >> 1];
            const Pair<T> p2 = reinterpret_cast<const Pair<T>*>(pdir)[k >> 1];
            x2.a += ta * p2.a;
            x2.b += ta * p2.b;
            reinterpret_cast<Pair<T>*>(xsol)[k >> 1] = x2;
          }
        }
      }
    }
    pr.rA[m] = ra;
    pr.rB[m] = rb;
    if (r < RY) xs[r + 1][c + 1] = Pair<T>{0, 0};
  }
  zero_ring<T>(xs, RY);
  __syncthreads();
  for (int sw = 0; sw < nu; ++sw)
    for (int color = 0; color < 2; ++color) {
      pair_sweep<T>(pr, xs, c, w, RY, gy0, color, sw == 0 && color == 0);
      __syncthreads();
    }
  // interpolation weights of the restriction footprint (loaded only now: not live in the sweeps)
  T wxA[MAXR], wxB[MAXR], wyA[MAXR], wyB[MAXR];
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + TWAVES * m, gj = gy0 + r;
    wxA[m] = wxB[m] = wyA[m] = wyB[m] = 0;
    if (pr.vA[m] && pair_fp && r >= H - 1 && r <= H + TY) {
      const size_t k = ci(giA, gj, nx);
      const Pair<T> wx2 = reinterpret_cast<const Pair<T>*>(F.wx)[k >> 1];
      const Pair<T> wy2 = reinterpret_cast<const Pair<T>*>(F.wy)[k >> 1];
      wxA[m] = wx2.a;
      wxB[m] = wx2.b;
      wyA[m] = wy2.a;
      wyB[m] = wy2.b;
    }
  }
  // residual (reference order: r - (diag x - offsum)) and the interior write of x
  T resA[MAXR], resB[MAXR];
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + TWAVES * m;
    resA[m] = 0;
    resB[m] = 0;
    if (r >= RY) continue;
    const Pair<T> own = xs[r + 1][c + 1], up = xs[r + 2][c + 1], dn = xs[r][c + 1];
    const T xw = xs[r + 1][c].b, xe = xs[r + 1][c + 2].a;
    if (pr.vA[m])
      resA[m] = pr.rA[m] - ((pr.wA[m] + pr.wB[m] + pr.sA[m] + pr.nA[m]) * own.a -
                            (pr.wA[m] * xw + pr.wB[m] * own.b + pr.sA[m] * dn.a + pr.nA[m] * up.a));
    if (pr.vB[m])
      resB[m] = pr.rB[m] - ((pr.wB[m] + pr.eB[m] + pr.sB[m] + pr.nB[m]) * own.b -
                            (pr.wB[m] * own.a + pr.eB[m] * xe + pr.sB[m] * dn.b + pr.nB[m] * up.b));
    if (pr.vA[m] && pair_int && r >= H && r < H + TY)
      reinterpret_cast<Pair<T>*>(x_lvl)[ci(giA, gy0 + r, nx) >> 1] = own;
  }
  __syncthreads();  // xs is reused below for the row sums us[2][RYMAX][NCXMAX]
  T* us = &xs[0][0].a;
  const int NCY = TY / 2, cx0 = tx0 / 2, cy0 = ty0 / 2;
  const int ic = c - hc;
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + TWAVES * m;
    if (r >= RY) continue;  // uniform across the wavefront
    // a*: part of res for the parent row (weight wy), b*: for the neighbouring row (1 - wy);
    // p*: part for the parent column (wx), q*: for the neighbouring column (1 - wx).
    const T aA = resA[m] * wyA[m], bA = resA[m] * ((T)1 - wyA[m]);
    const T aB = resB[m] * wyB[m], bB = resB[m] * ((T)1 - wyB[m]);
    const T paA = wxA[m] * aA, qaA = ((T)1 - wxA[m]) * aA, pbA = wxA[m] * bA, qbA = ((T)1 - wxA[m]) * bA;
    const T paB = wxB[m] * aB, qaB = ((T)1 - wxB[m]) * aB, pbB = wxB[m] * bB, qbB = ((T)1 - wxB[m]) * bB;
    // coarse column I (this lane) gathers q(2I-1) + p(2I) + p(2I+1) + q(2I+2)
    const T sa = (paA + paB) + (__shfl_up(qaB, 1, 64) + __shfl_down(qaA, 1, 64));
    const T sb = (pbA + pbB) + (__shfl_up(qbB, 1, 64) + __shfl_down(qbA, 1, 64));
    if (pair_int && r >= H - 1 && r <= H + TY) {
      us[r * NCXMAX + ic] = sa;
      us[(RYMAX + r) * NCXMAX + ic] = sb;
    }
  }
  __syncthreads();
  // coarse row J gathers B(2J-1) + A(2J) + A(2J+1) + B(2J+2)
  for (int t = threadIdx.x; t < NCX * NCY; t += TBLOCK) {
    const int a = t % NCX, b = t / NCX, I = cx0 + a, J = cy0 + b;
    if (I >= Cc.nx || J >= Cc.ny) continue;
    const int jr = 2 * b + H;  // region row of fine row 2J
    Cc.r[ci(I, J, Cc.nx)] = (us[(RYMAX + jr - 1) * NCXMAX + a] + us[jr * NCXMAX + a]) +
                            (us[(jr + 1) * NCXMAX + a] + us[(RYMAX + jr + 2) * NCXMAX + a]);
  }
  if (FUSE) {
    // norms of the updated residual over the tile interior (pr.r holds it)
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < MAXR; ++m) {
      const int r = w + TWAVES * m;
      if (!pr.vA[m] || !pair_int || r < H || r >= H + TY) continue;
      const double ra = (double)pr.rA[m], rb = (double)pr.rB[m];
      acc[0] += ra * ra + rb * rb;
      acc[1] = fmax(acc[1], fmax(fabs(ra), fabs(rb)));
      acc[2] += ra + rb;
    }
    block_partial<3, TBLOCK>(acc, 2u, partial);
  }
}

// Finish of the fused finest pre kernel: residual buffers swap, the shift is
// spent, and a pending update counts as an iteration with its convergence test
// on the mean-free norm sum r^2 - (sum r)^2 / N (the recursive residual's
// constant component is roundoff that no iteration reduces).
__global__ void __launch_bounds__(FINISH_THREADS) k_fin_pre0(const double* __restrict__ partial, int G,
                                                             double ncells, Scal* s) {
  if (s->done) return;
  double a[3];
  finish_reduce<3>(a, partial, G, 2u);
  if (threadIdx.x == 0) {
    s->rpar = 1 - s->rpar;
    s->rshift = 0.0;
    if (s->pending) {
      a[0] = fmax(a[0] - a[2] * a[2] / ncells, 0.0);
      s->r2 = a[0];
      s->rinf = a[1];
      s->rz_old = s->rz;
      s->iter += 1;
      const bool met = a[0] <= s->rel2 * s->r0_2 || (a[0] <= s->abs2 && a[1] <= s->ainf);
      if (!(a[0] == a[0])) s->flag = 1;
      if (met || s->iter >= s->max_iter || s->flag) s->done = 1;
      s->pending = 0;
    }
  }
}

// --------------------------------------------------------------------------- post-smoothing tile kernel
// Grid dim3(ceil(nx/TX), ceil(ny/TY)), TBLOCK threads, H = 2 nu (even): HX = H,
// TX = 128 - 2H.  x = x_pre + P x_coarse (the coarse values around the tile are
// staged in LDS), then nu (black, red) sweeps; the tile interior is written to
// x_out.  LEVEL0: r = rbuf[s->rpar], x_out = z, and the block partials of r.z,
// sum z, sum r are stored.
constexpr int PCW = 66;                           // staged coarse columns
constexpr int PCH = (TY + 4 * MAX_NU) / 2 + 3;    // staged coarse rows
template <class T, bool LEVEL0>
__global__ void __launch_bounds__(TBLOCK) __attribute__((amdgpu_waves_per_eu(4 * tile_min_blocks<T>()))) k_mg_post(Level<T> F, Level<T> Cc, const T* r_lvl, const T* rA,
                                                    const T* rB, const T* x_in, T* x_out, int nu, double* partial,
                                                    Scal* s) {
  if (s->done) return;
  __shared__ Pair<T> xs[RYMAX + 2][66];
  __shared__ T xcs[PCH][PCW];
  const int H = 2 * nu, HX = tile_hx(H), TX = 128 - 2 * HX, RY = TY + 2 * H;
  int bx, by;
  xcd_tile(bx, by);
  const int tx0 = bx * TX, ty0 = by * TY;
  const int gx0 = tx0 - HX, gy0 = ty0 - H;
  const int c = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int giA = gx0 + 2 * c;
  const int nx = F.nx, ny = F.ny, cnx = Cc.nx, cny = Cc.ny;
  const int hc = HX / 2, NCX = TX / 2;
  const bool pair_int = c >= hc && c < hc + NCX;
  const int cxa = gx0 / 2 - 1, cya = (gy0 >> 1) - 1;  // first staged coarse column / row (gx0 is even)
  const T* rin = LEVEL0 ? (s->rpar ? rB : rA) : r_lvl;
  // the coarse correction around the tile into LDS first, then each row is
  // staged and prolongated at once (the row's x_pre and weights die there)
  for (int t = threadIdx.x; t < PCH * PCW; t += TBLOCK) {
    const int I = min(max(cxa + t % PCW, 0), cnx - 1), J = min(max(cya + t / PCW, 0), cny - 1);
    xcs[t / PCW][t % PCW] = Cc.xb[ci(I, J, cnx)];
  }
  zero_ring<T>(xs, RY);
  __syncthreads();
  PairRows<T> pr;
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + TWAVES * m, gj = gy0 + r;
    const bool row_in = r < RY && gj >= 0 && gj < ny;
    const size_t k = row_in && giA >= 0 && giA < nx ? ci(giA, gj, nx) : 0;
    stage_coeffs<T>(pr, m, F, giA, gj, row_in, k);
    Pair<T> rr{0, 0}, x0{0, 0};
    if (pr.vA[m]) {
      rr = reinterpret_cast<const Pair<T>*>(rin)[k >> 1];
      x0 = reinterpret_cast<const Pair<T>*>(x_in)[k >> 1];
      const Pair<T> wx2 = reinterpret_cast<const Pair<T>*>(F.wx)[k >> 1];
      const Pair<T> wy2 = reinterpret_cast<const Pair<T>*>(F.wy)[k >> 1];
      // prolongation: A (column 2I) takes coarse I and I-1, B takes I and I+1
      const int I = giA >> 1, J = gj >> 1;
      int J2 = J + ((gj & 1) ? 1 : -1);
      J2 = J2 < 0 ? 0 : (J2 > cny - 1 ? cny - 1 : J2);
      const int IA2 = I - 1 < 0 ? 0 : I - 1, IB2 = I + 1 > cnx - 1 ? cnx - 1 : I + 1;
      const int a = I - cxa, aA = IA2 - cxa, aB = IB2 - cxa, b = J - cya, b2 = J2 - cya;
      const T wxa = wx2.a, wya = wy2.a, wxb = wx2.b, wyb = wy2.b;
      x0.a += wxa * wya * xcs[b][a] + ((T)1 - wxa) * wya * xcs[b][aA] + wxa * ((T)1 - wya) * xcs[b2][a] +
              ((T)1 - wxa) * ((T)1 - wya) * xcs[b2][aA];
      x0.b += wxb * wyb * xcs[b][a] + ((T)1 - wxb) * wyb * xcs[b][aB] + wxb * ((T)1 - wyb) * xcs[b2][a] +
              ((T)1 - wxb) * ((T)1 - wyb) * xcs[b2][aB];
    }
    pr.rA[m] = rr.a;
    pr.rB[m] = rr.b;
    if (r < RY) xs[r + 1][c + 1] = x0;
  }
  __syncthreads();
  for (int sw = 0; sw < nu; ++sw)
    for (int cc = 0; cc < 2; ++cc) {
      pair_sweep<T>(pr, xs, c, w, RY, gy0, 1 - cc, false);
      __syncthreads();
    }
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int m = 0; m < MAXR; ++m) {
    const int r = w + TWAVES * m;
    if (r >= RY || !pr.vA[m] || !pair_int || r < H || r >= H + TY) continue;
    const Pair<T> z = xs[r + 1][c + 1];
    reinterpret_cast<Pair<T>*>(x_out)[ci(giA, gy0 + r, nx) >> 1] = z;
    if (LEVEL0) {
      acc[0] += (double)pr.rA[m] * (double)z.a + (double)pr.rB[m] * (double)z.b;
      acc[1] += (double)z.a + (double)z.b;
      acc[2] += (double)pr.rA[m] + (double)pr.rB[m];
    }
  }
  if (LEVEL0) block_partial<3, TBLOCK>(acc, 0u, partial);
}

// Finish of the finest post kernel: zmean, rz = r.(z - zmean), beta.
__global__ void __launch_bounds__(FINISH_THREADS) k_fin_post0(const double* __restrict__ partial, int G, double N,
                                                              Scal* s) {
  if (s->done) return;
  double a[3];
  finish_reduce<3>(a, partial, G, 0u);
  if (threadIdx.x == 0) {
    const double zm = a[1] / N;
    const double rz = a[0] - zm * a[2];
    s->zmean = zm;
    s->rz = rz;
    s->beta = s->iter == s->seg0 ? 0.0 : rz / s->rz_old;
  }
}

// --------------------------------------------------------------------------- single-workgroup coarse V-cycle
// Levels l0..L-1 (at most SMALL_LV, each of at most SMALL_THREADS cells) in
// one workgroup.  Thread t owns cell t of every level and keeps that cell's
// coefficients, 1/diag and right-hand side in registers; LDS holds x and r of
// every level, the interpolation weights above the bottom, the residual and
// the two restriction row-sum arrays, and the bottom pseudo-inverse when the
// bottom has at most BOTTOM_LDS_MAX cells.  Every global load is issued in the
// first phase.  The bottom applies the dense pseudo-inverse; the cycle's
// result goes to level l0's xb.
constexpr int SMALL_LV = 4;

template <class T>
struct SmallArgs {
  Level<T> lv[SMALL_LV];  // levels l0.. in order; the last used one is the bottom
  int nl;                 // levels in the cycle
  const T* binv;          // bottom pseudo-inverse, nb x nb, row-major
};

// LDS layout of k_mg_small in elements of T (offsets), shared by host and device.
struct SmallLayout {
  int x[SMALL_LV], r[SMALL_LV], wx[SMALL_LV], wy[SMALL_LV];
  int xb, rb, tmp, tp, tq, binv, total;
};
template <class T>
__host__ __device__ inline SmallLayout small_layout(const Level<T>* lv, int nl) {
  SmallLayout o{};
  int off = 0;
#pragma unroll
  for (int q = 0; q < SMALL_LV; ++q) {
    if (q >= nl) break;
    const int n = lv[q].nx * lv[q].ny;
    o.x[q] = off;
    off += n;
    o.r[q] = off;
    off += n;
    if (q + 1 < nl) {
      o.wx[q] = off;
      off += n;
      o.wy[q] = off;
      off += n;
    } else {
      o.xb = o.x[q];
      o.rb = o.r[q];
    }
  }
  const int n0 = lv[0].nx * lv[0].ny, nh = lv[0].ny * (lv[0].nx / 2);
  o.tmp = off;
  off += n0;
  o.tp = off;
  off += nh;
  o.tq = off;
  off += nh;
  const int nb = lv[nl - 1].nx * lv[nl - 1].ny;
  o.binv = off;
  if (nb <= BOTTOM_LDS_MAX) off += nb * nb;
  o.total = off;
  return o;
}

// One red-black half-sweep of cell t (coordinates i, j) of a small level.
template <class T>
__device__ __forceinline__ void small_half(T* X, int t, int n, int nx, int ny, int i, int j, T cw, T ce, T cs, T cn,
                                           T id, T rr, int color, bool zero) {
  if (t < n && ((i + j) & 1) == color) {
    T o = 0;
    if (!zero)
      o = cw * X[i > 0 ? t - 1 : t] + ce * X[i + 1 < nx ? t + 1 : t] + cs * X[j > 0 ? t - nx : t] +
          cn * X[j + 1 < ny ? t + nx : t];
    X[t] = (rr + o) * id;
  }
  __syncthreads();
}

template <class T>
__global__ void __launch_bounds__(SMALL_THREADS) k_mg_small(SmallArgs<T> A, int nu_pre, int nu_post, const Scal* s) {
  if (s->done) return;
  extern __shared__ unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  const int nl = A.nl, t = threadIdx.x;
  const SmallLayout L = small_layout<T>(A.lv, nl);
  int nb = 0;
  int ci_[SMALL_LV], cj_[SMALL_LV];
  T cw[SMALL_LV], ce[SMALL_LV], cs[SMALL_LV], cn[SMALL_LV], cd[SMALL_LV], rr[SMALL_LV];
  // every global load of the cycle: coefficients (registers), r of level l0,
  // weights and the bottom pseudo-inverse (LDS); x of every level cleared
#pragma unroll
  for (int q = 0; q < SMALL_LV; ++q) {
    ci_[q] = cj_[q] = 0;
    cw[q] = ce[q] = cs[q] = cn[q] = cd[q] = rr[q] = (T)0;
    if (q >= nl) continue;
    const Level<T> F = A.lv[q];
    const int nx = F.nx, ny = F.ny;
    if (q == nl - 1) nb = nx * ny;
    if (t < nx * ny) {
      const int j = t / nx, i = t - j * nx;
      ci_[q] = i;
      cj_[q] = j;
      const T w = F.bw[t], e = i + 1 < nx ? F.bw[t + 1] : (T)0;
      const T so = F.bs[t], no = j + 1 < ny ? F.bs[t + nx] : (T)0;
      cd[q] = w + e + so + no;
      cw[q] = i > 0 ? w : (T)0;
      ce[q] = e;
      cs[q] = j > 0 ? so : (T)0;
      cn[q] = no;
      if (q == 0) {
        rr[0] = F.r[t];
        sm[L.r[0] + t] = rr[0];
      }
      if (q + 1 < nl) {
        sm[L.wx[q] + t] = F.wx[t];
        sm[L.wy[q] + t] = F.wy[t];
      }
      sm[L.x[q] + t] = (T)0;
    }
  }
  if (nb <= BOTTOM_LDS_MAX)
    for (int e = t; e < nb * nb; e += SMALL_THREADS) sm[L.binv + e] = A.binv[e];
  __syncthreads();
  // pre-smoothing, residual and restriction, level by level
#pragma unroll
  for (int q = 0; q + 1 < SMALL_LV; ++q) {
    if (q + 1 >= nl) break;
    const int nx = A.lv[q].nx, ny = A.lv[q].ny, n = nx * ny, cnx = nx / 2, cny = ny / 2;
    T* X = sm + L.x[q];
    const T id = cd[q] != (T)0 ? (T)1 / cd[q] : (T)0;
    if (q > 0 && t < n) rr[q] = sm[L.r[q] + t];
    for (int sw = 0; sw < nu_pre; ++sw) {
      small_half<T>(X, t, n, nx, ny, ci_[q], cj_[q], cw[q], ce[q], cs[q], cn[q], id, rr[q], 0, sw == 0);
      small_half<T>(X, t, n, nx, ny, ci_[q], cj_[q], cw[q], ce[q], cs[q], cn[q], id, rr[q], 1, false);
    }
    T* R = sm + L.tmp;
    if (t < n) {
      const int i = ci_[q], j = cj_[q];
      const T o = cw[q] * X[i > 0 ? t - 1 : t] + ce[q] * X[i + 1 < nx ? t + 1 : t] +
                  cs[q] * X[j > 0 ? t - nx : t] + cn[q] * X[j + 1 < ny ? t + nx : t];
      R[t] = rr[q] - (cd[q] * X[t] - o);
    }
    __syncthreads();
    // restriction r_H = P^T res in two passes: per fine row and coarse column
    // the x-weighted sums of res * wy (tp) and res * (1 - wy) (tq), then per
    // coarse cell the rows 2J-1 (tq), 2J, 2J+1 (tp), 2J+2 (tq)
    const T* WX = sm + L.wx[q];
    const T* WY = sm + L.wy[q];
    if (t < ny * cnx) {
      const int jj = t / cnx, I = t - jj * cnx;
      T pa = 0, pb = 0, qa = 0, qb = 0;
#pragma unroll
      for (int d = -1; d <= 2; ++d) {
        const int i = 2 * I + d;
        if (i < 0 || i >= nx) continue;
        const int k = i + jj * nx;
        const T res = R[k], wy = WY[k], wx = WX[k];
        const T a = res * wy, b = res * ((T)1 - wy);
        const T fx = (d == 0 || d == 1) ? wx : (T)1 - wx;
        if (d == 0 || d == 1) {
          pa += fx * a;
          pb += fx * b;
        } else {
          qa += fx * a;
          qb += fx * b;
        }
      }
      sm[L.tp + t] = pa + qa;
      sm[L.tq + t] = pb + qb;
    }
    __syncthreads();
    if (t < cnx * cny) {
      const int J = t / cnx, I = t - J * cnx;
      const T* TP = sm + L.tp;
      const T* TQ = sm + L.tq;
      const T lo = 2 * J - 1 >= 0 ? TQ[(2 * J - 1) * cnx + I] : (T)0;
      const T hi = 2 * J + 2 < ny ? TQ[(2 * J + 2) * cnx + I] : (T)0;
      sm[L.r[q + 1] + t] = (lo + TP[2 * J * cnx + I]) + (TP[(2 * J + 1) * cnx + I] + hi);
    }
    __syncthreads();
  }
  {
    // bottom: x = binv r, one wavefront per row, FP64 accumulation
    const T* Rb = sm + L.rb;
    T* Xb = sm + L.xb;
    const T* Bi = nb <= BOTTOM_LDS_MAX ? sm + L.binv : A.binv;
    const int lane = t & 63;
    for (int a = t >> 6; a < nb; a += SMALL_THREADS / 64) {
      double acc = 0;
      for (int qq = lane; qq < nb; qq += 64) acc += (double)Bi[(size_t)a * nb + qq] * (double)Rb[qq];
      acc = wave_sum(acc);
      if (lane == 0) Xb[a] = (T)acc;
    }
    __syncthreads();
  }
  // prolongation and post-smoothing, coarse to fine
#pragma unroll
  for (int q = SMALL_LV - 2; q >= 0; --q) {
    if (q + 1 >= nl) continue;
    const int nx = A.lv[q].nx, ny = A.lv[q].ny, n = nx * ny, cnx = nx / 2, cny = ny / 2;
    T* X = sm + L.x[q];
    const T* xc = sm + L.x[q + 1];
    if (t < n) {
      const int i = ci_[q], j = cj_[q];
      const T wx = sm[L.wx[q] + t], wy = sm[L.wy[q] + t];
      const int I = i >> 1, J = j >> 1;
      int I2 = I + ((i & 1) ? 1 : -1), J2 = J + ((j & 1) ? 1 : -1);
      I2 = I2 < 0 ? 0 : (I2 > cnx - 1 ? cnx - 1 : I2);
      J2 = J2 < 0 ? 0 : (J2 > cny - 1 ? cny - 1 : J2);
      X[t] += wx * wy * xc[I + J * cnx] + ((T)1 - wx) * wy * xc[I2 + J * cnx] + wx * ((T)1 - wy) * xc[I + J2 * cnx] +
              ((T)1 - wx) * ((T)1 - wy) * xc[I2 + J2 * cnx];
    }
    __syncthreads();
    const T id = cd[q] != (T)0 ? (T)1 / cd[q] : (T)0;
    for (int sw = 0; sw < nu_post; ++sw) {
      small_half<T>(X, t, n, nx, ny, ci_[q], cj_[q], cw[q], ce[q], cs[q], cn[q], id, rr[q], 1, false);
      small_half<T>(X, t, n, nx, ny, ci_[q], cj_[q], cw[q], ce[q], cs[q], cn[q], id, rr[q], 0, false);
    }
  }
  const int n0 = A.lv[0].nx * A.lv[0].ny;
  if (t < n0) A.lv[0].xb[t] = sm[L.x[0] + t];
}

// --------------------------------------------------------------------------- Krylov direction kernel
// p = (z - mean z) + beta p_old (FP64 arithmetic, stored as T; p_old/p_new =
// pbuf[iter & 1] / pbuf[(iter+1) & 1]), q = A0' p, alpha = rz / p.q; marks the
// update pending.  Tile kernel: p of a 64 x DTY tile plus a one-cell ring is
// formed in LDS (all loads issued at once), then the stencil reads LDS.
// Grid dim3(ceil(nx/64), ceil(ny/DTY)), BLOCK threads.
template <class T>
__global__ void __launch_bounds__(BLOCK) k_dir(int nx, int ny, const T* __restrict__ bw, const T* __restrict__ bs,
                                               const T* __restrict__ z, T* pA, T* pB, T* __restrict__ q,
                                               double* partial, Scal* s) {
  if (s->done) return;
  __shared__ T ps[DTY + 2][66];
  const double beta = s->beta, zm = s->zmean;
  const bool first = s->iter == s->seg0;
  const T* pold = (s->iter & 1) ? pB : pA;
  T* pnew = (s->iter & 1) ? pA : pB;
  const int c = threadIdx.x & 63, w = threadIdx.x >> 6;
  int bx, by;
  xcd_tile(bx, by);
  const int i0 = bx * 64, j0 = by * DTY;
  const int i = i0 + c;
  auto P = [&](int ii, int jj) -> T {
    if (ii < 0 || ii >= nx || jj < 0 || jj >= ny) return (T)0;
    const size_t k = ci(ii, jj, nx);
    return first ? (T)((double)z[k] - zm) : (T)(((double)z[k] - zm) + beta * (double)pold[k]);
  };
  constexpr int NW = BLOCK / 64, RR = (DTY + 2 + NW - 1) / NW;
#pragma unroll
  for (int m = 0; m < RR; ++m) {
    const int r = w + NW * m;
    if (r < DTY + 2) ps[r][c + 1] = P(i, j0 - 1 + r);
  }
  if (threadIdx.x < 2 * (DTY + 2)) {
    const int side = (int)threadIdx.x / (DTY + 2), r = (int)threadIdx.x % (DTY + 2);
    ps[r][side ? 65 : 0] = P(side ? i0 + 64 : i0 - 1, j0 - 1 + r);
  }
  __syncthreads();
  double acc[1] = {0.0};
#pragma unroll
  for (int m = 0; m < DTY / NW; ++m) {
    const int r = w + NW * m, j = j0 + r;
    if (i < nx && j < ny) {
      const size_t k = ci(i, j, nx);
      const T pc = ps[r + 1][c + 1];
      T a = 0;
      if (i > 0) a += bw[k] * (pc - ps[r + 1][c]);
      if (i + 1 < nx) a += bw[k + 1] * (pc - ps[r + 1][c + 2]);
      if (j > 0) a += bs[k] * (pc - ps[r][c + 1]);
      if (j + 1 < ny) a += bs[k + nx] * (pc - ps[r + 2][c + 1]);
      pnew[k] = pc;
      q[k] = a;
      acc[0] += (double)pc * (double)a;
    }
  }
  block_partial<1>(acc, 0u, partial);
}

// Finish of the direction kernel: alpha = rz / p.q and the update pending
// (breakdown when p.q is not positive).
__global__ void __launch_bounds__(FINISH_THREADS) k_fin_dir(const double* __restrict__ partial, int G, Scal* s) {
  if (s->done) return;
  double a[1];
  finish_reduce<1>(a, partial, G, 0u);
  if (threadIdx.x == 0) {
    s->pq = a[0];
    if (!(a[0] > 0.0)) {
      s->done = 1;
      s->flag = 1;
      s->alpha = 0.0;
    } else {
      s->alpha = s->rz / a[0];
      s->pending = 1;
    }
  }
}

}  // namespace p2d
