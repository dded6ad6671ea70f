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
// cells is staged once into registers, the sweeps run there (strips of
// consecutive rows per thread, lane shifts across columns, LDS only between
// strips), and only the tile interior is written back.  The global loads of
// a tile are independent of each other, so a tile costs about one memory
// round trip; coarse levels (small grids) are latency-bound.  Halo widths:
// pre H = 2 nu + 1 (first red sweep from zero needs no neighbours; each later
// half-sweep and the residual lose one exact ring; restriction reads one ring
// outside the tile), post H = 2 nu.
#pragma once
#include "p2d_common.hpp"

namespace p2d {

constexpr int TBLOCK = 1024;               // pre/post tile kernels of the finest level: 16 wavefronts
constexpr int MAX_NU = 3;                  // sweeps per side supported
constexpr int NCXMAX = 64;                 // coarse columns of a tile at most (TX / 2 < 64)
constexpr int KR = 3;                      // consecutive region rows per thread of a tile

// Tile shapes of the pre/post kernels: NW wavefronts of KR region rows each and ty
// output rows per tile, so that a region (ty + 2H rows, H <= 2 MAX_NU + 1) fits the
// NW KR rows: 16 wavefronts and 34 rows, 12 and 22, 8 and 10.  The half-sweeps are
// VALU issue bound per SIMD, so on the coarse levels, where CUs would otherwise
// idle, tiles of fewer wavefronts shorten each one; the host picks the shape per
// level (the finest level keeps the 16-wavefront tiles, whose halos cost least).
template <int NW>
struct Tile {
  static constexpr int waves = NW, threads = 64 * NW;
  static constexpr int ty = NW == 16 ? 34 : (NW == 12 ? 22 : 10);
  static_assert(NW == 16 || NW == 12 || NW == 8, "tile shapes");
  static constexpr int rymax = ty + 2 * (2 * MAX_NU + 1);  // region rows at most
  static constexpr int bnd = NW * 2 * 64;                  // one boundary buffer: [wave][bottom, top][lane]
  static constexpr int lds_pairs = 2 * bnd;                // two buffers (also the restriction's row sums
                                                           // and the block reduction)
  // The post-smoothing (halo H = 2 nu <= 2 MAX_NU) fills the strips with ty_post output
  // rows (36, 24, 12).
  static constexpr int ty_post = KR * NW - 4 * MAX_NU;
  static constexpr int pch = (ty_post + 4 * MAX_NU) / 2 + 3;  // staged coarse rows of the post kernel
  static_assert(KR * NW >= rymax, "a tile region must fit the strips");
  static_assert(ty % 2 == 0 && ty_post % 2 == 0 && ty_post >= ty, "tiles cover whole coarse rows");
  static_assert(2 * rymax * NCXMAX <= 2 * lds_pairs, "restriction row sums must fit the boundary buffers");
  static_assert(3 * threads * sizeof(double) <= lds_pairs * 2 * sizeof(float), "block reduction must fit them");
};
constexpr int TY = Tile<16>::ty;           // output rows of the finest level's tiles

// Tiles are 128 columns wide; the x halo HX is the smallest even width >= H so
// that the region starts at an even column; TX = 128 - 2 HX interior columns.
__host__ __device__ constexpr int tile_hx(int H) { return H + (H & 1); }
__host__ __device__ constexpr int tile_tx(int H) { return 128 - 2 * tile_hx(H); }
// Blocks per CU a tile kernel is compiled for: the finest level's FP32 tiles
// keep two resident (<= 64 VGPRs) so that one block's loads overlap the
// other's sweeps; FP64 tiles, and the coarse levels (a few blocks per CU at
// most), take the registers of one.
template <class T, bool FINEST>
__host__ __device__ constexpr int tile_blocks() { return sizeof(T) == 4 && FINEST ? 2 : 1; }
constexpr int DTY = 16;                    // rows per block of the direction kernel
constexpr int SMALL_MAX_CELLS = 4096;      // top level of the single-workgroup V-cycle at most (64^2)
constexpr int MAXL = 24;                   // multigrid levels at most
constexpr int BOTTOM_LDS_MAX = 64;         // bottom sizes inverted in LDS

template <class T>
struct Level {
  int nx, ny;
  const T* bw;  // west-face coefficient per cell
  const T* bs;  // south-face coefficient per cell
  const T* wx;  // parent interpolation weight per cell (x, y), FP64 levels; none on the bottom level
  const T* wy;
  const unsigned* wh;  // FP32 levels: both weights of a cell in one word (see WCell); none on the bottom
  T* xa;  // pre-smoothed correction
  T* xb;  // post-smoothed correction (the coarse correction read by the finer level)
  T* r;   // right-hand side (restricted residual)
  T om, omc;  // the level's smoother over-relaxation omega and 1 - omega (SOR; omega = 1: Gauss-Seidel)
};

// Interpolation weights as stored.  FP64 levels keep wx and wy in two FP64
// arrays.  FP32 levels keep both in one 32-bit word per cell (wh: wx in the low
// half, wy in the high half), each rounded to binary16 through binary32 (round
// to nearest even): the weights lie in [0, 1], the V-cycle stays symmetric (R =
// P^T with the same stored weights), and the finest level's two reads of them
// per cycle take 4 bytes per cell instead of 8.
__device__ __forceinline__ unsigned w_half_bits(double w) {
  return (unsigned)__builtin_bit_cast(unsigned short, (_Float16)(float)w);
}
__device__ __forceinline__ unsigned w_pack(double wx, double wy) { return w_half_bits(wx) | (w_half_bits(wy) << 16); }
__device__ __forceinline__ float w_lo(unsigned u) { return (float)__builtin_bit_cast(_Float16, (unsigned short)(u & 0xffffu)); }
__device__ __forceinline__ float w_hi(unsigned u) { return (float)__builtin_bit_cast(_Float16, (unsigned short)(u >> 16)); }
template <class T>
struct WCell;
template <>
struct WCell<float> {
  unsigned u = 0;
  __device__ __forceinline__ float x() const { return w_lo(u); }
  __device__ __forceinline__ float y() const { return w_hi(u); }
};
template <>
struct WCell<double> {
  double wx = 0, wy = 0;
  __device__ __forceinline__ double x() const { return wx; }
  __device__ __forceinline__ double y() const { return wy; }
};
// The weights of the cell pair (k, k + 1).
template <class T>
__device__ __forceinline__ void ld_wpair(const Level<T>& F, unsigned k, WCell<T>& a, WCell<T>& b) {
  if constexpr (sizeof(T) == 4) {
    const Pair<unsigned> u = ld_at<Pair<unsigned>>(F.wh, k);
    a.u = u.a;
    b.u = u.b;
  } else {
    const Pair<double> x = ld_at<Pair<double>>(F.wx, k), y = ld_at<Pair<double>>(F.wy, k);
    a.wx = x.a;
    a.wy = y.a;
    b.wx = x.b;
    b.wy = y.b;
  }
}
template <class T>
__device__ __forceinline__ WCell<T> ld_wcell(const Level<T>& F, int e) {
  WCell<T> w;
  if constexpr (sizeof(T) == 4) {
    w.u = F.wh[e];
  } else {
    w.wx = F.wx[e];
    w.wy = F.wy[e];
  }
  return w;
}
// Stores the weights of cell k (FP32 levels: packed into wh; FP64: wx, wy).
template <class T>
__device__ __forceinline__ void st_weights(T* wx, T* wy, unsigned* wh, size_t k, double a, double b) {
  if (wh) {
    wh[k] = w_pack(a, b);
  } else {
    wx[k] = (T)a;
    wy[k] = (T)b;
  }
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
// stored as T (wx, wy) or packed (wh, FP32 levels); the FP32 copy of the coefficients; and, from the thread of the
// lower-left child of each coarse cell, the coarse FP64 coefficients.
// Grid dim3(ceil(nx/BLOCK), ny).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_mg_setup(int nx, int ny, const double* __restrict__ bw,
                                                    const double* __restrict__ bs, T* __restrict__ wx,
                                                    T* __restrict__ wy, unsigned* __restrict__ wh,
                                                    float* __restrict__ f32bw,
                                                    float* __restrict__ f32bs, double* __restrict__ cbw,
                                                    double* __restrict__ cbs) {
  const int i = (int)(blockIdx.x * BLOCK + threadIdx.x), j = (int)blockIdx.y;
  if (i >= nx) return;
  const size_t k = ci(i, j, nx);
  auto row = [&](int jj) { return [=](int q) { return q < nx ? bw[ci(q, jj, nx)] : 0.0; }; };
  auto col = [&](int ii) { return [=](int q) { return q < ny ? bs[ci(ii, q, nx)] : 0.0; }; };
  if (wx || wh) {
    double a = od_weight<double>(row(j), i, nx);
    double b = od_weight<double>(col(i), j, ny);
    const int Q = (i >> 1) + ((i & 1) ? 1 : -1), Qj = (j >> 1) + ((j & 1) ? 1 : -1);
    if (Q < 0 || Q > (nx >> 1) - 1) a = 1.0;
    if (Qj < 0 || Qj > (ny >> 1) - 1) b = 1.0;
    st_weights<T>(wx, wy, wh, k, a, b);
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

// od_weight from the faces' reciprocals (k_mg_setup_tiled): the parent weight of
// fine index i along a row/column whose face coefficients are face(k), k = 0..nf
// (walls 0), given their reciprocals
// rface(k) = 1 / face(k) (inf at a wall): the resistance ratio R_far / (R_near +
// R_far) between the parent and the neighbouring coarse centre; 3/4 for
// constant coefficients, 1 at a wall.  0.5 / f is formed as 0.5 (1 / f), which
// is the same number.
template <class T, class F>
__device__ __forceinline__ T od_weight_rcp(F rface, int i, int nf) {
  const int P = i >> 1;
  const T near_ = (T)0.5 * rface(2 * P + 1);
  T far_;
  if ((i & 1) == 0) {
    if (P == 0) return (T)1;
    far_ = rface(2 * P) + (T)0.5 * rface(2 * P - 1);
  } else {
    if (2 * P + 2 >= nf) return (T)1;
    far_ = rface(2 * P + 2) + (T)0.5 * rface(2 * P + 3);
  }
  const T big = (T)1e30;
  if (!(near_ < big) || !(far_ < big)) return far_ >= near_ ? (T)1 : (T)0;
  return far_ / (near_ + far_);
}

// Coarse face = average over the two fine rows of the series conductance
// 2 / (0.5/b_{k-1} + 1/b_k + 0.5/b_{k+1}) between the two coarse centres
// (reciprocals rface as above).
template <class F>
__device__ __forceinline__ double coarse_face_rcp(F rface, int k, int nf) {
  if (k <= 0 || k >= nf) return 0.0;
  const double r = 0.5 * rface(k - 1) + rface(k) + 0.5 * rface(k + 1);
  return 2.0 / r;
}

// k_mg_setup for the large levels, with the same outputs bit for bit.  Tile
// kernel over STX x STY cells: the reciprocals of the faces the tile's weights
// and coarse faces read (x faces of rows j0 .. j0 + STY, columns i0 - 1 ..
// i0 + STX + 1; y faces of columns i0 .. i0 + STX, rows j0 - 1 .. j0 + STY + 1)
// are formed once into LDS.  Faces beyond the grid are 0 (reciprocal inf), as at
// the walls.  Grid dim3(ceil(nx/STX), ceil(ny/STY)), BLOCK threads.
constexpr int STX = 64, STY = 16;
template <class T>
__global__ void __launch_bounds__(BLOCK) k_mg_setup_tiled(int nx, int ny, const double* __restrict__ bw,
                                                    const double* __restrict__ bs, T* __restrict__ wx,
                                                    T* __restrict__ wy, unsigned* __restrict__ wh,
                                                    float* __restrict__ f32bw,
                                                    float* __restrict__ f32bs, double* __restrict__ cbw,
                                                    double* __restrict__ cbs) {
  __shared__ double rx[STY + 1][STX + 3], ry[STY + 3][STX + 1];
  int bx, by;
  xcd_tile(bx, by);
  const int i0 = bx * STX, j0 = by * STY;
  const bool need = wx || wh || cbw;
  if (need) {
    for (int t = threadIdx.x; t < (STY + 1) * (STX + 3); t += BLOCK) {
      const int a = t % (STX + 3), r = t / (STX + 3), q = i0 - 1 + a, jj = j0 + r;
      const double f = (q >= 0 && q < nx && jj < ny) ? bw[ci(q, jj, nx)] : 0.0;
      rx[r][a] = 1.0 / f;
    }
    for (int t = threadIdx.x; t < (STY + 3) * (STX + 1); t += BLOCK) {
      const int a = t % (STX + 1), r = t / (STX + 1), ii = i0 + a, q = j0 - 1 + r;
      const double f = (ii < nx && q >= 0 && q < ny) ? bs[ci(ii, q, nx)] : 0.0;
      ry[r][a] = 1.0 / f;
    }
    __syncthreads();
  }
  const int c = threadIdx.x & 63, w = wave_id();
  constexpr int NW = BLOCK / 64;
#pragma unroll
  for (int m = 0; m < STY / NW; ++m) {
    const int i = i0 + c, j = j0 + w + NW * m;
    if (i >= nx || j >= ny) continue;
    const size_t k = ci(i, j, nx);
    auto row = [&](int jj) { return [=](int q) { return rx[jj - j0][q - (i0 - 1)]; }; };
    auto col = [&](int ii) { return [=](int q) { return ry[q - (j0 - 1)][ii - i0]; }; };
    if (wx || wh) {
      double a = od_weight_rcp<double>(row(j), i, nx);
      double b = od_weight_rcp<double>(col(i), j, ny);
      const int Q = (i >> 1) + ((i & 1) ? 1 : -1), Qj = (j >> 1) + ((j & 1) ? 1 : -1);
      if (Q < 0 || Q > (nx >> 1) - 1) a = 1.0;
      if (Qj < 0 || Qj > (ny >> 1) - 1) b = 1.0;
      st_weights<T>(wx, wy, wh, k, a, b);
    }
    if (f32bw) {
      f32bw[k] = (float)bw[k];
      f32bs[k] = (float)bs[k];
    }
    if (cbw && (i & 1) == 0 && (j & 1) == 0) {
      const int I = i >> 1, J = j >> 1, cnx = nx >> 1;
      cbw[ci(I, J, cnx)] = 0.5 * (coarse_face_rcp(row(2 * J), 2 * I, nx) + coarse_face_rcp(row(2 * J + 1), 2 * I, nx));
      cbs[ci(I, J, cnx)] = 0.5 * (coarse_face_rcp(col(2 * I), 2 * J, ny) + coarse_face_rcp(col(2 * I + 1), 2 * J, ny));
    }
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
  __shared__ double diag[SMALL_THREADS];  // n <= SMALL_THREADS (the host's bottom limit)
  __shared__ double a[BOTTOM_LDS_MAX * BOTTOM_LDS_MAX];
  // diagonal entries loaded in parallel, summed in cell order by one thread
  for (int c = tid; c < n; c += SMALL_THREADS) {
    const int i = c % nx, j = c / nx;
    diag[c] = C.W(i, j) + C.E(i, j) + C.S(i, j) + C.N(i, j);
  }
  __syncthreads();
  if (tid == 0) {
    double dsum = 0;
    for (int c = 0; c < n; ++c) dsum += diag[c];
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
      row[c] += diag[c];
      if (i > 0) row[c - 1] -= C.W(i, j);
      if (i + 1 < nx) row[c + 1] -= C.E(i, j);
      if (j > 0) row[c - nx] -= C.S(i, j);
      if (j + 1 < ny) row[c + nx] -= C.N(i, j);
    }
    __syncthreads();
    // Gauss-Jordan with the matrix in the registers of the first GJ_THREADS threads
    // (thread t owns entries t + m GJ_THREADS; four wavefronts, one per SIMD, each
    // with OWN independent entries): step k needs the old pivot row and column,
    // which their owners publish in LDS (double-buffered: one barrier per step).
    // New values: a_kk = 1/a_kk, a_kc = a_kc / a_kk, a_rk = -a_rk / a_kk,
    // a_rc -= a_rk (a_kc / a_kk), computed branch-free with the pivot row, column
    // and element selected.  The other threads only pass the barriers.
    constexpr int GJ_THREADS = SMALL_THREADS;
    constexpr int OWN = BOTTOM_LDS_MAX * BOTTOM_LDS_MAX / GJ_THREADS;
    static_assert(OWN * GJ_THREADS == BOTTOM_LDS_MAX * BOTTOM_LDS_MAX && GJ_THREADS <= SMALL_THREADS,
                  "entries per thread");
    __shared__ double prow[2][BOTTOM_LDS_MAX], pcol[2][BOTTOM_LDS_MAX];
    const bool gj = tid < GJ_THREADS;
    double e[OWN];
    int er[OWN], ec[OWN];  // entry coordinates; -2: no entry (never a pivot row or the next one)
#pragma unroll
    for (int m = 0; m < OWN; ++m) {
      const int q = tid + m * GJ_THREADS;
      const bool valid = gj && q < n * n;
      er[m] = valid ? q / n : -2;
      ec[m] = valid ? q % n : -2;
      e[m] = valid ? a[q] : 0.0;
      if (er[m] == 0) prow[0][ec[m]] = e[m];
      if (ec[m] == 0) pcol[0][er[m]] = e[m];
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
      const int b = k & 1, nb = b ^ 1;
      if (gj) {
        const double d = 1.0 / prow[b][k];
#pragma unroll
        for (int m = 0; m < OWN; ++m) {
          const int r = er[m], c = ec[m];
          const double pc = pcol[b][r < 0 ? 0 : r], pr = prow[b][c < 0 ? 0 : c];
          const double gen = e[m] - pc * (pr * d), rowv = e[m] * d, colv = e[m] * -d;
          const double v = r == k ? (c == k ? d : rowv) : (c == k ? colv : gen);
          e[m] = v;
          if (r == k + 1) prow[nb][c] = v;
          if (c == k + 1) pcol[nb][r] = v;
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < OWN; ++m)
      if (er[m] >= 0) a[er[m] * n + ec[m]] = e[m];
    __syncthreads();
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
    row[c] += diag[c];
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
  s->near = 0;
}
// Start of a solve from the initial residual's norms in red: the reference
// norm b2 (mean-free), the outer targets (rtol on ||r||_2 relative to b2, abs2
// on ||r||_2^2, atol on max |r|; <= 0 disables; the sum-of-squares targets
// combine into the smaller one) and the first segment's inner targets: relative
// rel2 and, when inner_abs, absolute ones with a margin for the recursive
// residual (0.25 out2, 0.5 outi).  Sets outer_done when r0 already meets the
// outer targets.
__global__ void k_inner_init(Scal* s, double ncells, double rel2, int inner_abs, double rtol, double atol,
                             double abs2, int max_iter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double m = s->red[2] / ncells;
  const double b2 = fmax(s->red[0] - ncells * m * m, 0.0);
  s->b2 = b2;
  s->out2 = rtol > 0.0 ? rtol * rtol * b2 : -1.0;
  if (abs2 > 0.0) s->out2 = s->out2 < 0.0 ? abs2 : fmin(s->out2, abs2);
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
  s->rz = s->rz_old = s->pq = s->alpha = s->beta = s->zmean = s->r2raw = s->rsum = 0.0;
  s->near = 0;
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

// --------------------------------------------------------------------------- strip-layout tiles
// A tile region is 128 columns by RY = TY + 2H rows.  Lane c of every
// wavefront holds the column pair (2c, 2c+1): both children of one coarse
// column, since the region starts at an even column.  Wave w holds the KR
// consecutive region rows KR w .. KR w + KR - 1 (a strip).  The iterate stays
// in registers: horizontal neighbours come from the adjacent lanes (DPP
// wavefront shifts), vertical ones from the thread's own rows, and only the
// strip's bottom and top rows pass through LDS to the waves below and above
// (two buffers, one barrier per half-sweep).  The two cells of a pair have
// opposite colours, so a red-black half-sweep updates exactly one cell per
// pair and row, the same one for the whole wavefront: no lane idles on the
// colour test.  Cells outside the grid or the region hold 0 and are never
// updated.


// Per row m of the strip: right-hand side, faces (wA west of A, wB west of B
// = east of A, eB east of B, s south; the north faces are the next row's
// south faces, nt* for the top row), 1/diag, iterate and validity.
template <class T>
struct Strip {
  T rA[KR], rB[KR], wA[KR], wB[KR], eB[KR], sA[KR], sB[KR], iA[KR], iB[KR], xA[KR], xB[KR];
  T ntA, ntB;
  bool v[KR];
  __device__ __forceinline__ T nA(int m) const { return m + 1 < KR ? sA[m + 1 < KR ? m + 1 : m] : ntA; }
  __device__ __forceinline__ T nB(int m) const { return m + 1 < KR ? sB[m + 1 < KR ? m + 1 : m] : ntB; }
};

// Loads the strip's faces (rows up to RY: the row above the region supplies
// the top row's north faces); the iterate starts at 0.
template <class T>
__device__ __forceinline__ void strip_stage(Strip<T>& S, const Level<T>& F, int giA, int gy0, int w, int RY) {
  const int nx = F.nx, ny = F.ny;
  const bool col_in = giA >= 0 && giA < nx;  // nx is even on every tile level: giA + 1 < nx too
#pragma unroll
  for (int m = 0; m < KR; ++m) {
    const int r = KR * w + m, gj = gy0 + r;
    const bool ld = col_in && r <= RY && gj >= 0 && gj < ny;
    S.v[m] = ld && r < RY;
    T wA = 0, wB = 0, eB = 0, sA = 0, sB = 0;
    if (ld) {
      const unsigned k = cell_u(giA, gj, nx);
      const Pair<T> w2 = ld_at<Pair<T>>(F.bw, k);
      const Pair<T> s2 = ld_at<Pair<T>>(F.bs, k);
      wA = w2.a;
      wB = w2.b;
      sA = s2.a;
      sB = s2.b;
      if (giA + 2 < nx) eB = ld_at<T>(F.bw, k + 2);
    }
    S.wA[m] = wA;
    S.wB[m] = wB;
    S.eB[m] = eB;
    S.sA[m] = sA;
    S.sB[m] = sB;
    S.xA[m] = 0;
    S.xB[m] = 0;
  }
  {
    const int r = KR * w + KR - 1, gj = gy0 + r + 1;
    T a = 0, b = 0;
    if (col_in && r + 1 <= RY && gj >= 0 && gj < ny) {
      const Pair<T> n2 = ld_at<Pair<T>>(F.bs, cell_u(giA, gj, nx));
      a = n2.a;
      b = n2.b;
    }
    S.ntA = a;
    S.ntB = b;
  }
}
// 1/diag of the strip's cells (after strip_stage; kept out of it so that the
// post kernel forms them only once the prolongation's registers are free).
// FP32 uses the hardware reciprocal (1 ulp): the smoother only needs a close
// 1/diag, and the IEEE division is ten instructions per cell.
__device__ __forceinline__ float smoother_inv(float d) { return __builtin_amdgcn_rcpf(d); }
// FP64: the hardware reciprocal estimate refined by two Newton-Raphson steps
// r <- r + r (1 - d r) (each squares the relative error): within an ulp or two of
// 1/d in five instructions, where the IEEE division takes about twenty.
__device__ __forceinline__ double smoother_inv(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  return fma(r, fma(-d, r, 1.0), r);
}
// The smoother keeps omega / diag (SOR: x <- (1 - omega) x + omega x_GS = omc x + (r + offsum) omega / diag).
template <class T>
__device__ __forceinline__ void strip_diag(Strip<T>& S, T om) {
#pragma unroll
  for (int m = 0; m < KR; ++m) {
    S.iA[m] = S.v[m] ? om * smoother_inv(S.wA[m] + S.wB[m] + S.sA[m] + S.nA(m)) : (T)0;
    S.iB[m] = S.v[m] ? om * smoother_inv(S.wB[m] + S.eB[m] + S.sB[m] + S.nB(m)) : (T)0;
  }
}

// Boundary rows of the neighbouring strips from buffer rd: the top row of
// wave w - 1 (below this strip) and the bottom row of wave w + 1 (above).
template <class T, int NW>
__device__ __forceinline__ void strip_halo(const Pair<T>* rd, int c, int w, Pair<T>& dn, Pair<T>& up) {
  dn = w > 0 ? rd[((w - 1) * 2 + 1) * 64 + c] : Pair<T>{0, 0};
  up = w + 1 < NW ? rd[(w + 1) * 2 * 64 + c] : Pair<T>{0, 0};
}

// Half-sweep k of colour `color` (cells (i + j) & 1 == color; A, the even
// column, has colour (gy0 + r) & 1): reads the boundary buffer (k + 1) & 1,
// publishes the strip's boundary rows to buffer k & 1.  A cell of the swept
// colour reads only cells of the other colour, which this half-sweep leaves
// unchanged, so the rows update in place in any order.  zero: the iterate is
// 0 (first pre-smoothing half-sweep), no neighbour is read.  omc = 1 - omega
// (the strip's 1/diag carries omega).
template <class T, int NW>
__device__ __forceinline__ void strip_half(Strip<T>& S, Pair<T>* bnd, int c, int w, int gy0, int color, bool zero,
                                           int k, T omc) {
  Pair<T> dn{0, 0}, up{0, 0};
  if (!zero) strip_halo<T, NW>(bnd + ((k + 1) & 1) * Tile<NW>::bnd, c, w, dn, up);
#pragma unroll
  for (int m = 0; m < KR; ++m) {
    const bool doA = ((gy0 + KR * w + m) & 1) == color;  // the same for the whole wavefront
    const T xl = lane_prev(S.xB[m]), xr = lane_next(S.xA[m]);
    const T dA = m > 0 ? S.xA[m > 0 ? m - 1 : 0] : dn.a, dB = m > 0 ? S.xB[m > 0 ? m - 1 : 0] : dn.b;
    const T uA = m + 1 < KR ? S.xA[m + 1 < KR ? m + 1 : m] : up.a;
    const T uB = m + 1 < KR ? S.xB[m + 1 < KR ? m + 1 : m] : up.b;
    // cells outside the grid or the region have 1/diag = 0 and right-hand side 0, and
    // their neighbour sums are finite: they stay 0 without a per-lane test
    if (doA) {
      if (zero) {
        S.xA[m] = S.rA[m] * S.iA[m];
      } else {
        const T o = S.wA[m] * xl + S.wB[m] * S.xB[m] + S.sA[m] * dA + S.nA(m) * uA;
        S.xA[m] = (S.rA[m] + o) * S.iA[m] + omc * S.xA[m];
      }
    } else {
      if (zero) {
        S.xB[m] = S.rB[m] * S.iB[m];
      } else {
        const T o = S.wB[m] * S.xA[m] + S.eB[m] * xr + S.sB[m] * dB + S.nB(m) * uB;
        S.xB[m] = (S.rB[m] + o) * S.iB[m] + omc * S.xB[m];
      }
    }
  }
  Pair<T>* wr = bnd + (k & 1) * Tile<NW>::bnd;
  wr[(w * 2) * 64 + c] = Pair<T>{S.xA[0], S.xB[0]};
  wr[(w * 2 + 1) * 64 + c] = Pair<T>{S.xA[KR - 1], S.xB[KR - 1]};
  __syncthreads();
}

// --------------------------------------------------------------------------- pre-smoothing tile kernel
// Grid dim3(ceil(nx/TX), ceil(ny/ty)), 64 NW threads (tile shape Tile<NW>), H = 2 nu + 1, x halo
// HX = tile_hx(H) (even), TX = 128 - 2 HX.
// FUSE (finest level): r_lvl is the current residual buffer; the update r -= a q,
// x += a p (pend: an update is pending, p = pdir; xfirst: it is the first of its
// pass, x = a p, so x needs no clearing between passes) and the residual shift are
// applied first (tile interiors write r to rout and x), and the block partials of
// ||r|| are stored.  The host picks the buffers from its replay of the solver
// state, so their loads do not wait for the state; a guarded head (the last of an
// enqueued chunk, which an early end at the direction finish usually makes
// unnecessary) reads the state first and issues no load once the pass is done.
// Restriction r_H = P^T res: each cell splits its residual into the four
// weight products; the x sums run across lanes, the y sums through LDS.
template <class T, bool FUSE, bool BUSY = false, int NW = 16>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4 * tile_blocks<T, FUSE || BUSY>())))
k_mg_pre(Level<T> F, Level<T> Cc, const T* r_lvl, T* x_lvl, int nu, T* rout, T* xsol, const T* pdir,
         const T* __restrict__ qv, int pend, int xfirst, int guard, double* partial, Scal* s) {
  using Sh = Tile<NW>;
  constexpr int TY = Sh::ty, RYMAX = Sh::rymax;
  __shared__ Pair<T> lds[Sh::lds_pairs];
  if (FUSE && guard && s->done) return;
  const int H = 2 * nu + 1, HX = tile_hx(H), TX = 128 - 2 * HX, RY = TY + 2 * H;
  int bx, by;
  xcd_tile(bx, by);
  const int tx0 = bx * TX, ty0 = by * TY;
  const int gx0 = tx0 - HX, gy0 = ty0 - H;
  const int c = threadIdx.x & 63, w = wave_id();
  const int giA = gx0 + 2 * c;
  const int nx = F.nx;
  const int hc = HX / 2, NCX = TX / 2;
  const bool pair_int = c >= hc && c < hc + NCX;       // both cells in the tile interior
  const bool pair_fp = c >= hc - 1 && c <= hc + NCX;   // restriction footprint
  // The coefficients (and, below the finest level, the right-hand side and the weights)
  // do not depend on the solver state: their loads are issued before `done` is read, so
  // that their round trip overlaps that of the state.
  Strip<T> S;
  strip_stage<T>(S, F, giA, gy0, w, RY);
  WCell<T> wcA[KR], wcB[KR];
  auto load_weights = [&]() {
#pragma unroll
    for (int m = 0; m < KR; ++m) {
      const int r = KR * w + m;
      wcA[m] = WCell<T>{};
      wcB[m] = WCell<T>{};
      if (S.v[m] && pair_fp && r >= H - 1 && r <= H + TY) ld_wpair<T>(F, cell_u(giA, gy0 + r, nx), wcA[m], wcB[m]);
    }
  };
  // interpolation weights of the restriction footprint: loaded with the rest on
  // coarse levels, late in the sweeps where two blocks share a CU (the finest
  // level, and BUSY coarse levels with more tiles than CUs): registers are short there
  if (!(FUSE || BUSY)) load_weights();
  if (!FUSE) {
#pragma unroll
    for (int m = 0; m < KR; ++m) {
      Pair<T> r2{0, 0};
      if (S.v[m]) r2 = ld_at<Pair<T>>(r_lvl, cell_u(giA, gy0 + KR * w + m, nx));
      S.rA[m] = r2.a;
      S.rB[m] = r2.b;
    }
  }
  // finest level: the update's operands, loaded before the state is read (the
  // buffers come from the host); alpha and the shift are read with `done`
  Pair<T> q2[KR], x2[KR], p2[KR];
#pragma unroll
  for (int m = 0; m < KR && FUSE; ++m) {
    const int r = KR * w + m;
    Pair<T> r2{0, 0};
    q2[m] = x2[m] = p2[m] = Pair<T>{0, 0};
    if (S.v[m]) {
      const unsigned k = cell_u(giA, gy0 + r, nx);
      r2 = ld_at<Pair<T>>(r_lvl, k);
      if (pend) {
        q2[m] = ld_at<Pair<T>>(qv, k);
        if (pair_int && r >= H && r < H + TY) {
          if (!xfirst) x2[m] = ld_at<Pair<T>>(xsol, k);
          p2[m] = ld_at<Pair<T>>(pdir, k);
        }
      }
    }
    S.rA[m] = r2.a;
    S.rB[m] = r2.b;
  }
  double alpha = 0.0, shift = 0.0;
  if (FUSE) {
    alpha = s->alpha;
    shift = s->rshift;
  }
  if (s->done) return;
  const T ta = (T)alpha, tshift = (T)shift;
  strip_diag<T>(S, F.om);
#pragma unroll
  for (int m = 0; m < KR && FUSE; ++m) {
    const int r = KR * w + m;
    if (!S.v[m]) continue;
    const unsigned k = cell_u(giA, gy0 + r, nx);
    // the inner recursion runs in T: r -= a q, x += a p (a, shift rounded to T)
    T ra = S.rA[m], rb = S.rB[m];
    if (pend) {
      ra -= ta * q2[m].a;
      rb -= ta * q2[m].b;
    }
    ra -= tshift;
    rb -= tshift;
    if (pair_int && r >= H && r < H + TY) {
      st_at(rout, k, Pair<T>{ra, rb});
      if (pend) st_at(xsol, k, Pair<T>{x2[m].a + ta * p2[m].a, x2[m].b + ta * p2[m].b});
    }
    S.rA[m] = ra;
    S.rB[m] = rb;
  }
  // on the finest level the weights' loads are issued before the last
  // half-sweep, whose work covers part of their latency
  for (int k = 0; k < 2 * nu - 1; ++k) strip_half<T, NW>(S, lds, c, w, gy0, k & 1, k == 0, k, F.omc);
  if (FUSE || BUSY) load_weights();
  strip_half<T, NW>(S, lds, c, w, gy0, 1, false, 2 * nu - 1, F.omc);
  // residual (reference order: r - (diag x - offsum)) with the last published
  // boundary rows (buffer (2 nu - 1) & 1), and the interior write of x
  T resA[KR], resB[KR];
  {
    Pair<T> dn, up;
    strip_halo<T, NW>(lds + Sh::bnd, c, w, dn, up);
#pragma unroll
    for (int m = 0; m < KR; ++m) {
      const int r = KR * w + m;
      const T xl = lane_prev(S.xB[m]), xr = lane_next(S.xA[m]);
      const T dA = m > 0 ? S.xA[m > 0 ? m - 1 : 0] : dn.a, dB = m > 0 ? S.xB[m > 0 ? m - 1 : 0] : dn.b;
      const T uA = m + 1 < KR ? S.xA[m + 1 < KR ? m + 1 : m] : up.a;
      const T uB = m + 1 < KR ? S.xB[m + 1 < KR ? m + 1 : m] : up.b;
      resA[m] = 0;
      resB[m] = 0;
      if (!S.v[m]) continue;
      resA[m] = S.rA[m] - ((S.wA[m] + S.wB[m] + S.sA[m] + S.nA(m)) * S.xA[m] -
                           (S.wA[m] * xl + S.wB[m] * S.xB[m] + S.sA[m] * dA + S.nA(m) * uA));
      resB[m] = S.rB[m] - ((S.wB[m] + S.eB[m] + S.sB[m] + S.nB(m)) * S.xB[m] -
                           (S.wB[m] * S.xA[m] + S.eB[m] * xr + S.sB[m] * dB + S.nB(m) * uB));
      if (pair_int && r >= H && r < H + TY)
        st_at(x_lvl, cell_u(giA, gy0 + r, nx), Pair<T>{S.xA[m], S.xB[m]});
    }
  }
  __syncthreads();  // the boundary buffers are reused below for the row sums us[2][RYMAX][NCXMAX]
  T* us = &lds[0].a;
  const int NCY = TY / 2, cx0 = tx0 / 2, cy0 = ty0 / 2;
  const int ic = c - hc;
#pragma unroll
  for (int m = 0; m < KR; ++m) {
    const int r = KR * w + m;
    // a*: part of res for the parent row (weight wy), b*: for the neighbouring row (1 - wy);
    // p*: part for the parent column (wx), q*: for the neighbouring column (1 - wx).
    // (the (1 - w) shares as differences: res = a + b, a = pa + qa, b = pb + qb)
    const T wxA = wcA[m].x(), wyA = wcA[m].y(), wxB = wcB[m].x(), wyB = wcB[m].y();
    const T aA = resA[m] * wyA, bA = resA[m] - aA;
    const T aB = resB[m] * wyB, bB = resB[m] - aB;
    const T paA = wxA * aA, qaA = aA - paA, pbA = wxA * bA, qbA = bA - pbA;
    const T paB = wxB * aB, qaB = aB - paB, pbB = wxB * bB, qbB = bB - pbB;
    // coarse column I (this lane) gathers q(2I-1) + p(2I) + p(2I+1) + q(2I+2)
    const T sa = (paA + paB) + (lane_prev(qaB) + lane_next(qaA));
    const T sb = (pbA + pbB) + (lane_prev(qbB) + lane_next(qbA));
    if (pair_int && r >= H - 1 && r <= H + TY) {
      us[r * NCXMAX + ic] = sa;
      us[(RYMAX + r) * NCXMAX + ic] = sb;
    }
  }
  __syncthreads();
  // coarse row J gathers B(2J-1) + A(2J) + A(2J+1) + B(2J+2)
  for (int t = threadIdx.x; t < NCX * NCY; t += Sh::threads) {
    const int a = t % NCX, b = t / NCX, I = cx0 + a, J = cy0 + b;
    if (I >= Cc.nx || J >= Cc.ny) continue;
    const int jr = 2 * b + H;  // region row of fine row 2J
    st_at(Cc.r, cell_u(I, J, Cc.nx), (us[(RYMAX + jr - 1) * NCXMAX + a] + us[jr * NCXMAX + a]) +
                            (us[(jr + 1) * NCXMAX + a] + us[(RYMAX + jr + 2) * NCXMAX + a]));
  }
  if (FUSE) {
    // norms of the updated residual over the tile interior (S.r holds it)
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < KR; ++m) {
      const int r = KR * w + m;
      if (!S.v[m] || !pair_int || r < H || r >= H + TY) continue;
      const double ra = (double)S.rA[m], rb = (double)S.rB[m];
      acc[0] += ra * ra + rb * rb;
      acc[1] = fmax(acc[1], fmax(fabs(ra), fabs(rb)));
      acc[2] += ra + rb;
    }
    block_partial_lds<3, Sh::threads>(acc, 2u, partial, reinterpret_cast<double*>(lds));
  }
}

// The next direction kernel prepares the early convergence test (k_dir)
// when the residual is within NEAR_END (in sum r^2) of its pass's relative target,
// or when one more iteration at the rate of the last one would bring it within
// NEAR_RATE of the target: the over-relaxed multigrid cuts sum r^2 by ~100x per
// iteration, so a fixed window alone would miss the last iteration of most passes.
constexpr double NEAR_END = 64.0;
constexpr double NEAR_RATE = 16.0;

// Finish of the fused finest pre kernel: residual buffers swap, the shift is
// spent, and a pending update counts as an iteration with its convergence test
// on the mean-free norm sum r^2 - (sum r)^2 / N (the recursive residual's
// constant component is roundoff that no iteration reduces).
// The state is read together with the partials (one memory round trip); a
// finished solve leaves it unchanged.
__global__ void __launch_bounds__(FINISH_THREADS) k_fin_pre0(const double* __restrict__ partial, int G,
                                                             double ncells, Scal* s) {
  const int done = s->done, pending = s->pending, rpar = s->rpar, iter = s->iter, max_iter = s->max_iter;
  const int flag = s->flag;
  const double rz = s->rz, rel2 = s->rel2, r0_2 = s->r0_2, abs2 = s->abs2, ainf = s->ainf, r2_prev = s->r2;
  double a[3];
  finish_reduce<3>(a, partial, G, 2u);
  if (threadIdx.x != 0 || done) return;
  s->rpar = 1 - rpar;
  s->rshift = 0.0;
  s->r2raw = a[0];
  s->rsum = a[2];
  s->near = 0;
  if (pending) {
    a[0] = fmax(a[0] - a[2] * a[2] / ncells, 0.0);
    const double next = r2_prev > 0.0 ? a[0] * fmin(1.0, a[0] / r2_prev) : a[0];  // one more iteration at this rate
    s->near = a[0] <= NEAR_END * rel2 * r0_2 || next <= NEAR_RATE * rel2 * r0_2 ? 1 : 0;
    s->r2 = a[0];
    s->rinf = a[1];
    s->rz_old = rz;
    s->iter = iter + 1;
    const bool met = a[0] <= rel2 * r0_2 || (a[0] <= abs2 && a[1] <= ainf);
    const bool bad = flag || !(a[0] == a[0]);
    if (bad) s->flag = 1;
    if (met || iter + 1 >= max_iter || bad) s->done = 1;
    s->pending = 0;
  }
}

// --------------------------------------------------------------------------- post-smoothing tile kernel
// Grid dim3(ceil(nx/TX), ceil(ny/ty)), 64 NW threads (tile shape Tile<NW>), H = 2 nu (even): HX = H,
// TX = 128 - 2H.  x = x_pre + P x_coarse (the coarse values around the tile are
// staged in LDS while the strip's own loads are in flight), then nu (black,
// red) sweeps; the tile interior is written to x_out.  LEVEL0: r_lvl is the
// current residual buffer (picked by the host), x_out = z, and the block
// partials of r.z, sum z, sum r are stored.
constexpr int PCW = 66;                           // staged coarse columns
template <class T, bool LEVEL0, bool BUSY = false, int NW = 16>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4 * tile_blocks<T, LEVEL0 || BUSY>())))
k_mg_post(Level<T> F, Level<T> Cc, const T* r_lvl, const T* x_in, T* x_out, int nu, double* partial, Scal* s) {
  using Sh = Tile<NW>;
  constexpr int TY = Sh::ty_post, PCH = Sh::pch, TBLK = Sh::threads;
  __shared__ Pair<T> lds[Sh::lds_pairs];
  __shared__ T xcs[PCH][PCW];
  const int H = 2 * nu, HX = tile_hx(H), TX = 128 - 2 * HX, RY = TY + 2 * H;
  int bx, by;
  xcd_tile(bx, by);
  const int tx0 = bx * TX, ty0 = by * TY;
  const int gx0 = tx0 - HX, gy0 = ty0 - H;
  const int c = threadIdx.x & 63, w = wave_id();
  const int giA = gx0 + 2 * c;
  const int nx = F.nx, cnx = Cc.nx, cny = Cc.ny;
  const int hc = HX / 2, NCX = TX / 2;
  const bool pair_int = c >= hc && c < hc + NCX;
  const int cxa = gx0 / 2 - 1, cya = (gy0 >> 1) - 1;  // first staged coarse column / row (gx0 is even)
  // Every load is issued before `done` is read (the coarse values go to LDS after the test).
  constexpr int XCS_PER_THREAD = (PCH * PCW + TBLK - 1) / TBLK;
  T xcv[XCS_PER_THREAD];
#pragma unroll
  for (int q = 0; q < XCS_PER_THREAD; ++q) {
    const int t = (int)threadIdx.x + q * TBLK;
    const int I = min(max(cxa + t % PCW, 0), cnx - 1), J = min(max(cya + t / PCW, 0), cny - 1);
    xcv[q] = t < PCH * PCW ? ld_at<T>(Cc.xb, cell_u(I, J, cnx)) : (T)0;
  }
  Strip<T> S;
  strip_stage<T>(S, F, giA, gy0, w, RY);
  WCell<T> wcA[KR], wcB[KR];
#pragma unroll
  for (int m = 0; m < KR; ++m) {
    Pair<T> rr{0, 0}, x0{0, 0};
    wcA[m] = WCell<T>{};
    wcB[m] = WCell<T>{};
    if (S.v[m]) {
      const unsigned k = cell_u(giA, gy0 + KR * w + m, nx);
      rr = ld_at<Pair<T>>(r_lvl, k);
      x0 = ld_at<Pair<T>>(x_in, k);
      ld_wpair<T>(F, k, wcA[m], wcB[m]);
    }
    S.rA[m] = rr.a;
    S.rB[m] = rr.b;
    S.xA[m] = x0.a;
    S.xB[m] = x0.b;
  }
  if (s->done) return;
#pragma unroll
  for (int q = 0; q < XCS_PER_THREAD; ++q) {
    const int t = (int)threadIdx.x + q * TBLK;
    if (t < PCH * PCW) xcs[t / PCW][t % PCW] = xcv[q];
  }
  __syncthreads();  // xcs
  // prolongation: A (column 2I) takes coarse I and I-1, B takes I and I+1
#pragma unroll
  for (int m = 0; m < KR; ++m) {
    if (!S.v[m]) continue;
    const int gj = gy0 + KR * w + m;
    const int I = giA >> 1, J = gj >> 1;
    int J2 = J + ((gj & 1) ? 1 : -1);
    J2 = J2 < 0 ? 0 : (J2 > cny - 1 ? cny - 1 : J2);
    const int IA2 = I - 1 < 0 ? 0 : I - 1, IB2 = I + 1 > cnx - 1 ? cnx - 1 : I + 1;
    const int a = I - cxa, aA = IA2 - cxa, aB = IB2 - cxa, b = J - cya, b2 = J2 - cya;
    const T wxa = wcA[m].x(), wya = wcA[m].y(), wxb = wcB[m].x(), wyb = wcB[m].y();
    S.xA[m] += wxa * wya * xcs[b][a] + ((T)1 - wxa) * wya * xcs[b][aA] + wxa * ((T)1 - wya) * xcs[b2][a] +
               ((T)1 - wxa) * ((T)1 - wya) * xcs[b2][aA];
    S.xB[m] += wxb * wyb * xcs[b][a] + ((T)1 - wxb) * wyb * xcs[b][aB] + wxb * ((T)1 - wyb) * xcs[b2][a] +
               ((T)1 - wxb) * ((T)1 - wyb) * xcs[b2][aB];
  }
  strip_diag<T>(S, F.om);
  // publish the boundary rows for the first half-sweep (it reads buffer 1)
  lds[Sh::bnd + (w * 2) * 64 + c] = Pair<T>{S.xA[0], S.xB[0]};
  lds[Sh::bnd + (w * 2 + 1) * 64 + c] = Pair<T>{S.xA[KR - 1], S.xB[KR - 1]};
  __syncthreads();
  for (int k = 0; k < 2 * nu; ++k) strip_half<T, NW>(S, lds, c, w, gy0, 1 - (k & 1), false, k, F.omc);
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int m = 0; m < KR; ++m) {
    const int r = KR * w + m;
    if (!S.v[m] || !pair_int || r < H || r >= H + TY) continue;
    st_at(x_out, cell_u(giA, gy0 + r, nx), Pair<T>{S.xA[m], S.xB[m]});
    if (LEVEL0) {
      acc[0] += (double)S.rA[m] * (double)S.xA[m] + (double)S.rB[m] * (double)S.xB[m];
      acc[1] += (double)S.xA[m] + (double)S.xB[m];
      acc[2] += (double)S.rA[m] + (double)S.rB[m];
    }
  }
  if (LEVEL0) block_partial_lds<3, TBLK>(acc, 0u, partial, reinterpret_cast<double*>(lds));
}

// Finish of the finest post kernel: zmean, rz = r.(z - zmean), beta.
__global__ void __launch_bounds__(FINISH_THREADS) k_fin_post0(const double* __restrict__ partial, int G, double N,
                                                              Scal* s) {
  const int done = s->done, first = s->iter == s->seg0;
  const double rz_old = s->rz_old;
  double a[3];
  finish_reduce<3>(a, partial, G, 0u);
  if (threadIdx.x != 0 || done) return;
  const double zm = a[1] / N;
  const double rz = a[0] - zm * a[2];
  s->zmean = zm;
  s->rz = rz;
  s->beta = first ? 0.0 : rz / rz_old;
}

// --------------------------------------------------------------------------- single-workgroup coarse V-cycle
// Levels l0..L-1 (at most SMALL_LV; the top one of at most SMALL_MAX_CELLS
// cells, the others of at most SMALL_THREADS) in one workgroup.  Thread t owns
// the cell pairs (2p, 2p + 1), p = t + SMALL_THREADS v (v < SMALL_CPT / 2), of the
// top level (horizontal neighbours, opposite colours: every red-black half-sweep
// updates one cell of each pair, with no lane idle; slot 2v holds the pair's red
// cell, 2v + 1 its black one) and cell t of every level below, and keeps each
// owned cell's coefficients, diagonal and
// right-hand side in registers ("slots"); LDS holds x of every level, r of
// the levels below the top (filled by the restriction), the interpolation
// weights above the bottom, the residual and the two restriction row-sum
// arrays, and the bottom pseudo-inverse when the bottom has at most
// BOTTOM_LDS_MAX cells.  Every global load is issued in the first phase.  The
// bottom applies the dense pseudo-inverse; the cycle's result goes to level
// l0's xb.
constexpr int SMALL_LV = 4;
constexpr int SMALL_CPT = SMALL_MAX_CELLS / SMALL_THREADS;  // top-level cells per thread
constexpr int SMALL_SLOTS = SMALL_CPT + SMALL_LV - 1;
static_assert(SMALL_CPT * SMALL_THREADS == SMALL_MAX_CELLS && SMALL_MAX_CELLS <= 4 * SMALL_THREADS,
              "levels below the top must have at most one cell per thread");
__host__ __device__ constexpr int small_cpt(int q) { return q == 0 ? SMALL_CPT : 1; }
__host__ __device__ constexpr int small_slot(int q, int u) { return q == 0 ? u : SMALL_CPT - 1 + q; }
// The cell of thread t's slot u on level q (width nx) of the cycle: the top level's slots 2v
// and 2v + 1 are the red and the black cell of the pair (2p, 2p + 1), p = t + SMALL_THREADS v
// (one row: every level above the bottom has an even width; a bottom-only cycle has no sweeps,
// so any one-to-one map serves it).
__host__ __device__ inline int small_cell(int q, int t, int u, int nx) {
  if (q != 0) return t;
  const int e0 = 2 * (t + (u >> 1) * SMALL_THREADS), j = e0 / nx;
  return e0 + ((u & 1) ^ (j & 1));
}
static_assert(SMALL_CPT % 2 == 0, "the top level's slots come in pairs");

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
    o.r[q] = off;  // the top level's r lives in registers unless it is the bottom
    if (q > 0 || nl == 1) off += n;
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

// NL = A.nl (a template parameter: the level loops unroll into straight-line
// code, so the arguments and the first loads of every level are issued at once).
template <class T, int NL>
__global__ void __launch_bounds__(SMALL_THREADS) k_mg_small(SmallArgs<T> A, int nu_pre, int nu_post, const Scal* s) {
  const int done = s->done;  // tested once the loads are issued: their round trips overlap
  extern __shared__ unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  constexpr int nl = NL;
  const int t = threadIdx.x;
  const SmallLayout L = small_layout<T>(A.lv, nl);
  int nb = 0;
#pragma unroll
  for (int q = 0; q < SMALL_LV; ++q)
    if (q == nl - 1) nb = A.lv[q].nx * A.lv[q].ny;
  // slots of levels below the bottom stay zero (cd = 0: their 1/diag is 0)
  int ci_[SMALL_SLOTS] = {}, cj_[SMALL_SLOTS] = {};
  T cw[SMALL_SLOTS] = {}, ce[SMALL_SLOTS] = {}, cs[SMALL_SLOTS] = {}, cn[SMALL_SLOTS] = {}, cd[SMALL_SLOTS] = {},
    id[SMALL_SLOTS] = {}, rr[SMALL_SLOTS] = {};
  // Every global load of the cycle is issued before any is used (one round trip):
  // the bottom pseudo-inverse and the weights first (they go to LDS), then the
  // coefficients and the top level's r (registers); x is cleared.
  constexpr int BINV_PT = BOTTOM_LDS_MAX * BOTTOM_LDS_MAX / SMALL_THREADS;
  static_assert(BINV_PT * SMALL_THREADS == BOTTOM_LDS_MAX * BOTTOM_LDS_MAX, "pseudo-inverse staging");
  const bool binv_lds = nb <= BOTTOM_LDS_MAX;
  T bv[BINV_PT];
#pragma unroll
  for (int k = 0; k < BINV_PT; ++k) {
    const int e = t + k * SMALL_THREADS;
    bv[k] = binv_lds && e < nb * nb ? A.binv[e] : (T)0;
  }
  WCell<T> wc[SMALL_SLOTS];
#pragma unroll
  for (int q = 0; q + 1 < SMALL_LV; ++q) {
    if (q + 1 >= nl) break;
    const Level<T> F = A.lv[q];
    const int n = F.nx * F.ny;
#pragma unroll
    for (int u = 0; u < small_cpt(q); ++u) {
      const int sl = small_slot(q, u), e = small_cell(q, t, u, A.lv[q].nx);
      wc[sl] = e < n ? ld_wcell<T>(F, e) : WCell<T>{};
    }
  }
#pragma unroll
  for (int q = 0; q < SMALL_LV; ++q) {
    if (q >= nl) break;
    const Level<T> F = A.lv[q];
    const int nx = F.nx, ny = F.ny, n = nx * ny;
#pragma unroll
    for (int u = 0; u < small_cpt(q); ++u) {
      const int sl = small_slot(q, u), e = small_cell(q, t, u, A.lv[q].nx);
      ci_[sl] = cj_[sl] = 0;
      cw[sl] = ce[sl] = cs[sl] = cn[sl] = rr[sl] = (T)0;
      if (e >= n) continue;
      const int j = e / nx, i = e - j * nx;
      ci_[sl] = i;
      cj_[sl] = j;
      cw[sl] = F.bw[e];
      ce[sl] = i + 1 < nx ? F.bw[e + 1] : (T)0;
      cs[sl] = F.bs[e];
      cn[sl] = j + 1 < ny ? F.bs[e + nx] : (T)0;
      if (q == 0) rr[sl] = F.r[e];
    }
  }
  if (binv_lds) {
#pragma unroll
    for (int k = 0; k < BINV_PT; ++k) {
      const int e = t + k * SMALL_THREADS;
      if (e < nb * nb) sm[L.binv + e] = bv[k];
    }
  }
#pragma unroll
  for (int q = 0; q < SMALL_LV; ++q) {
    if (q >= nl) break;
    const int n = A.lv[q].nx * A.lv[q].ny;
#pragma unroll
    for (int u = 0; u < small_cpt(q); ++u) {
      const int sl = small_slot(q, u), e = small_cell(q, t, u, A.lv[q].nx);
      if (e >= n) continue;
      if (q + 1 < nl) {
        sm[L.wx[q] + e] = wc[sl].x();
        sm[L.wy[q] + e] = wc[sl].y();
      }
      sm[L.x[q] + e] = (T)0;
    }
  }
#pragma unroll
  for (int sl = 0; sl < SMALL_SLOTS; ++sl) {
    // the west / south faces at the walls (i = 0, j = 0) are stored zeros
    cd[sl] = cw[sl] + ce[sl] + cs[sl] + cn[sl];
    if (ci_[sl] == 0) cw[sl] = (T)0;
    if (cj_[sl] == 0) cs[sl] = (T)0;
  }
  if (nl == 1) {
#pragma unroll
    for (int u = 0; u < SMALL_CPT; ++u)
      if (small_cell(0, t, u, A.lv[0].nx) < nb) sm[L.r[0] + small_cell(0, t, u, A.lv[0].nx)] = rr[u];
  }
  if (done) return;
  // omega / diag (the levels below the finest share omega)
  const T om = A.lv[0].om, omc = A.lv[0].omc;
#pragma unroll
  for (int sl = 0; sl < SMALL_SLOTS; ++sl) id[sl] = cd[sl] != (T)0 ? om / cd[sl] : (T)0;
  __syncthreads();
  // one red-black half-sweep of level q (cells (i + j) & 1 == color).  On the top level
  // the cell of each pair with the colour is picked by selects: one update per pair.
  auto half = [&](int q, int color, bool zero) __attribute__((always_inline)) {
    const int nx = A.lv[q].nx, ny = A.lv[q].ny, n = nx * ny;
    T* X = sm + L.x[q];
    auto update = [&](int e, int i, int j, T w_, T e_, T s_, T n_, T r_, T id_) __attribute__((always_inline)) {
      if (zero) {
        X[e] = r_ * id_;
      } else {
        const T o = w_ * X[i > 0 ? e - 1 : e] + e_ * X[i + 1 < nx ? e + 1 : e] + s_ * X[j > 0 ? e - nx : e] +
                    n_ * X[j + 1 < ny ? e + nx : e];
        X[e] = (r_ + o) * id_ + omc * X[e];
      }
    };
    if (q == 0) {
#pragma unroll
      for (int v = 0; v < SMALL_CPT / 2; ++v) {
        const int sl = small_slot(0, 2 * v + color);  // the pair's cell of this colour
        if (2 * (t + v * SMALL_THREADS) >= n) continue;
        const int i = ci_[sl], j = cj_[sl];
        update(i + j * nx, i, j, cw[sl], ce[sl], cs[sl], cn[sl], rr[sl], id[sl]);
      }
    } else {
      const int sl = small_slot(q, 0), e = t, i = ci_[sl], j = cj_[sl];
      if (e < n && ((i + j) & 1) == color) update(e, i, j, cw[sl], ce[sl], cs[sl], cn[sl], rr[sl], id[sl]);
    }
    __syncthreads();
  };
  // pre-smoothing, residual and restriction, level by level
#pragma unroll
  for (int q = 0; q + 1 < SMALL_LV; ++q) {
    if (q + 1 >= nl) break;
    const int nx = A.lv[q].nx, ny = A.lv[q].ny, n = nx * ny, cnx = nx / 2, cny = ny / 2;
    T* X = sm + L.x[q];
    if (q > 0 && t < n) rr[small_slot(q, 0)] = sm[L.r[q] + t];
    for (int sw = 0; sw < nu_pre; ++sw) {
      half(q, 0, sw == 0);
      half(q, 1, false);
    }
    T* R = sm + L.tmp;
#pragma unroll
    for (int u = 0; u < small_cpt(q); ++u) {
      const int sl = small_slot(q, u), e = small_cell(q, t, u, A.lv[q].nx), i = ci_[sl], j = cj_[sl];
      if (e >= n) continue;
      const T o = cw[sl] * X[i > 0 ? e - 1 : e] + ce[sl] * X[i + 1 < nx ? e + 1 : e] +
                  cs[sl] * X[j > 0 ? e - nx : e] + cn[sl] * X[j + 1 < ny ? e + nx : e];
      R[e] = rr[sl] - (cd[sl] * X[e] - o);
    }
    __syncthreads();
    // restriction r_H = P^T res in two passes: per fine row and coarse column
    // the x-weighted sums of res * wy (tp) and res * (1 - wy) (tq), then per
    // coarse cell the rows 2J-1 (tq), 2J, 2J+1 (tp), 2J+2 (tq)
    const T* WX = sm + L.wx[q];
    const T* WY = sm + L.wy[q];
    for (int e = t; e < ny * cnx; e += SMALL_THREADS) {
      const int jj = e / cnx, I = e - jj * cnx;
      T pa = 0, pb = 0, qa = 0, qb = 0;
#pragma unroll
      for (int d = -1; d <= 2; ++d) {
        const int i = 2 * I + d;
        if (i < 0 || i >= nx) continue;
        const int k = i + jj * nx;
        const T res = R[k], wy = WY[k], wx = WX[k];
        const T a = res * wy, b = res * ((T)1 - wy);
        if (d == 0 || d == 1) {
          pa += wx * a;
          pb += wx * b;
        } else {
          qa += ((T)1 - wx) * a;
          qb += ((T)1 - wx) * b;
        }
      }
      sm[L.tp + e] = pa + qa;
      sm[L.tq + e] = pb + qb;
    }
    __syncthreads();
    for (int e = t; e < cnx * cny; e += SMALL_THREADS) {
      const int J = e / cnx, I = e - J * cnx;
      const T* TP = sm + L.tp;
      const T* TQ = sm + L.tq;
      const T lo = 2 * J - 1 >= 0 ? TQ[(2 * J - 1) * cnx + I] : (T)0;
      const T hi = 2 * J + 2 < ny ? TQ[(2 * J + 2) * cnx + I] : (T)0;
      sm[L.r[q + 1] + e] = (lo + TP[2 * J * cnx + I]) + (TP[(2 * J + 1) * cnx + I] + hi);
    }
    __syncthreads();
  }
  {
    // bottom: x = binv r, FP64 accumulation.  A pseudo-inverse in LDS (nb <= 64):
    // 16 lanes per row (all rows at once), each summing every 16th column, then
    // the row's 16 lanes by DPP (quad permutations, row shifts by 4 and 8: the
    // total lands in the row's last lane).  Larger bottoms: one wavefront per row.
    const T* Rb = sm + L.rb;
    T* Xb = sm + L.xb;
    static_assert(BOTTOM_LDS_MAX * 16 == SMALL_THREADS, "16 lanes per pseudo-inverse row");
    if (binv_lds) {
      const int a = t >> 4, k = t & 15;
      double acc = 0;
#pragma unroll
      for (int m = 0; m < BOTTOM_LDS_MAX / 16; ++m) {
        const int qq = k + 16 * m;
        if (a < nb && qq < nb) acc += (double)sm[L.binv + a * nb + qq] * (double)Rb[qq];
      }
      acc += dpp_move<0xb1>(acc);   // quad_perm [1,0,3,2]
      acc += dpp_move<0x4e>(acc);   // quad_perm [2,3,0,1]
      acc += dpp_move<0x114>(acc);  // row_shr:4
      acc += dpp_move<0x118>(acc);  // row_shr:8
      if (k == 15 && a < nb) Xb[a] = (T)acc;
    } else {
      const int lane = t & 63;
      for (int a = t >> 6; a < nb; a += SMALL_THREADS / 64) {
        double acc = 0;
        for (int qq = lane; qq < nb; qq += 64) acc += (double)A.binv[(size_t)a * nb + qq] * (double)Rb[qq];
        acc = wave_sum(acc);
        if (lane == 0) Xb[a] = (T)acc;
      }
    }
    __syncthreads();
  }
  // prolongation and post-smoothing, coarse to fine
#pragma unroll
  for (int q = SMALL_LV - 2; q >= 0; --q) {
    if (q + 1 >= nl) continue;
    const int nx = A.lv[q].nx, n = nx * A.lv[q].ny, cnx = nx / 2, cny = A.lv[q].ny / 2;
    T* X = sm + L.x[q];
    const T* xc = sm + L.x[q + 1];
#pragma unroll
    for (int u = 0; u < small_cpt(q); ++u) {
      const int sl = small_slot(q, u), e = small_cell(q, t, u, A.lv[q].nx), i = ci_[sl], j = cj_[sl];
      if (e >= n) continue;
      const T wx = sm[L.wx[q] + e], wy = sm[L.wy[q] + e];
      const int I = i >> 1, J = j >> 1;
      int I2 = I + ((i & 1) ? 1 : -1), J2 = J + ((j & 1) ? 1 : -1);
      I2 = I2 < 0 ? 0 : (I2 > cnx - 1 ? cnx - 1 : I2);
      J2 = J2 < 0 ? 0 : (J2 > cny - 1 ? cny - 1 : J2);
      X[e] += wx * wy * xc[I + J * cnx] + ((T)1 - wx) * wy * xc[I2 + J * cnx] + wx * ((T)1 - wy) * xc[I + J2 * cnx] +
              ((T)1 - wx) * ((T)1 - wy) * xc[I2 + J2 * cnx];
    }
    __syncthreads();
    for (int sw = 0; sw < nu_post; ++sw) {
      half(q, 1, false);
      half(q, 0, false);
    }
  }
  const int n0 = A.lv[0].nx * A.lv[0].ny;
  for (int e = t; e < n0; e += SMALL_THREADS) A.lv[0].xb[e] = sm[L.x[0] + e];
}

// --------------------------------------------------------------------------- Krylov direction kernel
// p = (z - mean z) + beta p_old (FP64 arithmetic, stored as T; first: the first
// direction of a sequence, p_old unused; the host picks p_old and p_new from its
// replay of the solver state), q = A0' p, the block partials of p.q.  When the
// state says the pass is near its end: also r.q, q.q and
// sum q with the current residual rcur, from which k_fin_dir knows the next
// residual's norm before the update is applied.  The choice depends on the solve's
// own state only, so a solve ends its passes in the same places whatever ran before.
// Tile kernel over 128 x DTY cells: lane c of a wavefront holds the column pair
// (2c, 2c+1) of a row, so every array moves with one two-element
// load or store per lane.  p of the tile plus a one-cell ring is formed in LDS
// (all loads issued at once), then the stencil reads LDS.  The finest level has
// an even nx (it coarsens at least once), so a pair is inside the grid or
// outside it as a whole.  Grid dim3(ceil(nx/128), ceil(ny/DTY)), BLOCK threads.
constexpr int DTX = 128;  // columns per block of the direction kernel
template <class T>
__global__ void __launch_bounds__(BLOCK) k_dir(int nx, int ny, const T* __restrict__ bw, const T* __restrict__ bs,
                                               const T* __restrict__ z, const T* pold, T* pnew, T* __restrict__ q,
                                               const T* __restrict__ rcur, int first, double* partial, Scal* s) {
  // row layout: [1] west ring, [2 + 2c, 3 + 2c] pair c, [DTX + 2] east ring (pairs 2-element aligned)
  __shared__ Pair<T> ps2[DTY + 2][DTX / 2 + 2];
  T(*ps)[DTX + 4] = reinterpret_cast<T(*)[DTX + 4]>(&ps2[0][0]);
  const int c = threadIdx.x & 63, w = wave_id();
  int bx, by;
  xcd_tile(bx, by);
  const int i0 = bx * DTX, j0 = by * DTY;
  const int iA = i0 + 2 * c;
  const bool col_in = iA < nx;
  // Each wave owns RPW consecutive rows, so a row's north faces are the next row's
  // south ones, loaded once; the east face of B is the next lane's west face of A
  // (DPP), only the last lane loads its own.  These loads depend on no solver state
  // and are issued before it is read.
  constexpr int NW = BLOCK / 64, RR = (DTY + 2 + NW - 1) / NW;
  constexpr int RPW = DTY / NW;
  const int rb = w * RPW;
  T sA[RPW + 1], sB[RPW + 1], wA[RPW], wB[RPW], eB[RPW], eL[RPW];
#pragma unroll
  for (int m = 0; m <= RPW; ++m) {
    const int j = j0 + rb + m;
    const Pair<T> s2 = (col_in && j < ny) ? ld_at<Pair<T>>(bs, cell_u(iA, j, nx)) : Pair<T>{(T)0, (T)0};
    sA[m] = s2.a;
    sB[m] = s2.b;
  }
#pragma unroll
  for (int m = 0; m < RPW; ++m) {
    const int j = j0 + rb + m;
    const Pair<T> w2 = (col_in && j < ny) ? ld_at<Pair<T>>(bw, cell_u(iA, j, nx)) : Pair<T>{(T)0, (T)0};
    wA[m] = w2.a;
    wB[m] = w2.b;
    eL[m] = (c == 63 && iA + 2 < nx && j < ny) ? ld_at<T>(bw, cell_u(iA + 2, j, nx)) : (T)0;
  }
  // z and p_old of the tile and its ring, loaded before the state is read
  Pair<T> z2[RR], o2[RR];
#pragma unroll
  for (int m = 0; m < RR; ++m) {
    const int r = w + NW * m, jj = j0 - 1 + r;
    z2[m] = o2[m] = Pair<T>{(T)0, (T)0};
    if (r < DTY + 2 && col_in && jj >= 0 && jj < ny) {
      const unsigned k = cell_u(iA, jj, nx);
      z2[m] = ld_at<Pair<T>>(z, k);
      if (!first) o2[m] = ld_at<Pair<T>>(pold, k);
    }
  }
  const int near = s->near;
  Pair<T> rc[RPW];
#pragma unroll
  for (int m = 0; m < RPW; ++m) {
    const int j = j0 + rb + m;
    rc[m] = (near && col_in && j < ny) ? ld_at<Pair<T>>(rcur, cell_u(iA, j, nx)) : Pair<T>{(T)0, (T)0};
  }
  T zr = 0, orr = 0;
  const bool ring = threadIdx.x < 2 * (DTY + 2);
  const int side = (int)threadIdx.x / (DTY + 2), rr_ = (int)threadIdx.x % (DTY + 2);
  if (ring) {
    const int ii = side ? i0 + DTX : i0 - 1, jj = j0 - 1 + rr_;
    if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) {
      const unsigned k = cell_u(ii, jj, nx);
      zr = z[k];
      if (!first) orr = pold[k];
    }
  }
  const double beta = s->beta, zm = s->zmean;
  if (s->done) return;
  auto pform = [&](T zz, T po) -> T {
    return first ? (T)((double)zz - zm) : (T)(((double)zz - zm) + beta * (double)po);
  };
#pragma unroll
  for (int m = 0; m < RR; ++m) {
    const int r = w + NW * m, jj = j0 - 1 + r;
    if (r >= DTY + 2) continue;
    Pair<T> v{(T)0, (T)0};
    if (col_in && jj >= 0 && jj < ny) {
      v.a = pform(z2[m].a, o2[m].a);
      v.b = pform(z2[m].b, o2[m].b);
    }
    ps2[r][1 + c] = v;
  }
  if (ring) {
    const int ii = side ? i0 + DTX : i0 - 1, jj = j0 - 1 + rr_;
    ps[rr_][side ? DTX + 2 : 1] = (ii >= 0 && ii < nx && jj >= 0 && jj < ny) ? pform(zr, orr) : (T)0;
  }
#pragma unroll
  for (int m = 0; m < RPW; ++m) {
    // the lane move runs on every lane (a lane outside the active mask would read as 0)
    const T e = lane_next(wA[m]);
    eB[m] = c == 63 ? eL[m] : e;
  }
  __syncthreads();
  double acc[4] = {};
#pragma unroll
  for (int m = 0; m < RPW; ++m) {
    const int r = rb + m, j = j0 + r;
    if (!col_in || j >= ny) continue;
    const unsigned k = cell_u(iA, j, nx);
    const Pair<T> pc = ps2[r + 1][1 + c], ps_ = ps2[r][1 + c], pn = ps2[r + 2][1 + c];
    const T pa = pc.a, pb = pc.b;
    // cell A (column iA): west, east (= B), south, north
    T qa = 0;
    if (iA > 0) qa += wA[m] * (pa - ps[r + 1][1 + 2 * c]);
    qa += wB[m] * (pa - pb);
    if (j > 0) qa += sA[m] * (pa - ps_.a);
    if (j + 1 < ny) qa += sA[m + 1] * (pa - pn.a);
    // cell B (column iA + 1): west (= A), east, south, north
    T qb = 0;
    qb += wB[m] * (pb - pa);
    if (iA + 2 < nx) qb += eB[m] * (pb - ps[r + 1][4 + 2 * c]);
    if (j > 0) qb += sB[m] * (pb - ps_.b);
    if (j + 1 < ny) qb += sB[m + 1] * (pb - pn.b);
    st_at(pnew, k, Pair<T>{pa, pb});
    st_at(q, k, Pair<T>{qa, qb});
    acc[0] += (double)pa * (double)qa;
    acc[0] += (double)pb * (double)qb;
    if (near) {
      acc[1] += (double)rc[m].a * (double)qa + (double)rc[m].b * (double)qb;
      acc[2] += (double)qa * (double)qa + (double)qb * (double)qb;
      acc[3] += (double)qa + (double)qb;
    }
  }
  if (near) {
    block_partial<4>(acc, 0u, partial);
    return;
  }
  double a1[1] = {acc[0]};
  block_partial<1>(a1, 0u, partial);
}

// Finish of the direction kernel: alpha = rz / p.q and the update pending
// (breakdown when p.q is not positive).  Near a pass's end, the residual the update will
// leave, r' = r - alpha q, has sum r'^2 = sum r^2 - 2 alpha r.q + alpha^2 q.q and
// sum r' = sum r - alpha sum q (FP64 sums of the FP32 vectors); when its
// mean-free norm already meets the pass's relative target, the pass ends here, one
// iteration on and with the update still pending (the refinement applies it), and
// the next head is not run.
__global__ void __launch_bounds__(FINISH_THREADS) k_fin_dir(const double* __restrict__ partial, int G, double ncells,
                                                            Scal* s) {
  const int done = s->done, iter = s->iter, near = s->near;
  const double rz = s->rz, r2raw = s->r2raw, rsum = s->rsum, rel2 = s->rel2, r0_2 = s->r0_2;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  if (near) {
    finish_reduce<4>(a, partial, G, 0u);
  } else {
    double b[1];
    finish_reduce<1>(b, partial, G, 0u);
    a[0] = b[0];
  }
  if (threadIdx.x != 0 || done) return;
  s->pq = a[0];
  if (!(a[0] > 0.0)) {
    s->done = 1;
    s->flag = 1;
    s->alpha = 0.0;
    return;
  }
  const double alpha = rz / a[0];
  s->alpha = alpha;
  s->pending = 1;
  if (near) {
    const double r2n = r2raw - 2.0 * alpha * a[1] + alpha * alpha * a[2];
    const double sn = rsum - alpha * a[3];
    const double mf = fmax(r2n - sn * sn / ncells, 0.0);
    if (mf == mf && mf <= rel2 * r0_2) {
      s->iter = iter + 1;
      s->r2 = mf;
      s->rinf = sqrt(fmax(r2n, 0.0));  // an upper bound of max |r'|
      s->done = 1;
    }
  }
}

}  // namespace p2d
