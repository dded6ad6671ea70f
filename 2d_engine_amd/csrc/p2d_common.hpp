// p2d device code for gfx950 (CDNA4), shared pieces: layout, solver state,
// deterministic FP64 reductions, ghost-fluid and double-double helpers.
//
// Layout.  Every field the kernels touch is cell-pitched: nx x ny, i fastest,
// cell (i,j) at i + j*nx.  Face coefficients too: bw(i,j) is beta on the WEST
// face of cell (i,j) (the x-face between (i-1,j) and (i,j)), bs(i,j) on its
// SOUTH face; the wall faces are implicit zeros (bw(0,j) = bs(i,0) = 0 are
// stored, the east/north walls are not).  So a stencil row reads aligned,
// equal-pitch arrays only.  The C ABI's face layout ((nx+1) x ny and
// nx x (ny+1)) is converted at the boundary.
//
// Solver operators are h^2-scaled: (A' x)_c = sum_f beta_f (x_c - x_nb),
// A' = h^2 A with A = -D B G of the reference's variable-coefficient matrix
// (POISSON_SOLVER.cpp:749-872) without its pin: A' is singular on constants on
// every level, the right-hand side is taken mean-free and the coarsest level is
// solved with a pseudo-inverse.
//
// Reductions accumulate in FP64 and are deterministic: 64-lane wave shuffles,
// a fixed-order block tree, per-block partials, then a one-block finish kernel
// sums the partials in block order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p2d {

constexpr int BLOCK = 256;          // 4 wavefronts of 64
constexpr int BAND = 4;             // rows per block of the band (row-marching) kernels
constexpr int SMALL_THREADS = 1024; // single-workgroup coarse V-cycle
constexpr int RED_K = 4;            // most quantities any kernel reduces

__device__ __forceinline__ size_t ci(int i, int j, int nx) { return (size_t)i + (size_t)j * nx; }
__device__ __forceinline__ size_t fxi(int i, int j, int nx) { return (size_t)i + (size_t)j * (nx + 1); }
__device__ __forceinline__ size_t fyi(int i, int j, int nx) { return (size_t)i + (size_t)j * nx; }

// Face coefficients of cell (i,j) in the cell-pitched layout.
template <class T>
struct Coef {
  const T* bw;
  const T* bs;
  int nx, ny;
  __device__ __forceinline__ T W(int i, int j) const { return bw[ci(i, j, nx)]; }
  __device__ __forceinline__ T E(int i, int j) const { return i + 1 < nx ? bw[ci(i + 1, j, nx)] : (T)0; }
  __device__ __forceinline__ T S(int i, int j) const { return bs[ci(i, j, nx)]; }
  __device__ __forceinline__ T N(int i, int j) const { return j + 1 < ny ? bs[ci(i, j + 1, nx)] : (T)0; }
};

__device__ __forceinline__ int nblocks() { return (int)(gridDim.x * gridDim.y); }
__device__ __forceinline__ int block_id() { return (int)(blockIdx.y * gridDim.x + blockIdx.x); }

// XCD-aware tile order.  The dispatcher deals workgroups to the 8 XCDs
// round-robin by linear id; tiles are renumbered so that every XCD owns one
// contiguous run of row-major tiles, and the halo rows/columns that
// neighbouring tiles share are fetched once into that XCD's L2.  Bijective for
// any grid size.  Returns the (x, y) tile of this workgroup.
constexpr int NXCD = 8;
__device__ __forceinline__ void xcd_tile(int& tx, int& ty) {
  const int G = nblocks(), w = block_id();
  const int x = w % NXCD, k = w / NXCD, q = G / NXCD, rem = G % NXCD;
  const int t = x * q + min(x, rem) + k;
  tx = t % (int)gridDim.x;
  ty = t / (int)gridDim.x;
}

// Device-resident solver state.  Kernels of an inner iteration read `done` on
// entry and return at once when it is set, so the host can enqueue iterations
// ahead without waiting for the convergence test.
struct Scal {
  double rz, rz_old, pq, alpha, beta;
  double r2, rinf, r0_2;
  double rel2;    // stop when r2 <= rel2 * r0_2
  double abs2;    // ... or when r2 <= abs2 and rinf <= ainf
  double ainf;
  double bmean;   // mean of the right-hand side (removed: A0 is singular on constants)
  double zmean;   // mean of the preconditioned residual (removed: keeps iterates mean-free)
  double b2;      // mean-free ||r0||^2 of the solve (the relative target's reference)
  double out2;    // outer targets: sum r^2 <= out2, max |r| <= outi (< 0: disabled)
  double outi;
  double rshift;  // mean of a replaced residual, removed by the next fused update (its constant
                  // component is roundoff that no iterate can reduce: A0 p has zero sum)
  double r2raw, rsum;  // sum r^2 and sum r of the current recursive residual (the last head's)
  double red[8];  // results of stand-alone reductions
  int iter, done, max_iter, flag;  // flag: 1 = breakdown / NaN
  int pending;                     // 1: a direction and alpha await the update x += a p, r -= a q
  int rpar;                        // which of the two residual buffers holds the current residual
  int err;                         // coefficient assembly error (zero level-set gradient)
  int seg0;                        // iteration at which the current direction sequence started (beta = 0)
  int outer_done;                  // the outer targets hold after a refinement
  int near;                        // the last head's residual is near the pass's target (k_fin_pre0)
};

// Index of the calling thread's wavefront in its block, as a wave-uniform (scalar) value:
// tests on it become scalar branches instead of per-lane masks.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// --------------------------------------------------------------------------- reductions
// Cross-lane moves of a 64-bit value (two 32-bit DPP moves); lanes without a
// source receive 0.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xf, 0xf, true);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Wavefront sum / max (of non-negative values: the empty lanes enter as 0),
// returned to every lane: quad permutations, row shifts by 4 and 8, then the
// row broadcasts of lanes 15 and 31 leave the total in lane 63.
template <bool MAX>
__device__ __forceinline__ double wave_reduce(double v) {
  auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
  v = op(v, dpp_move<0xb1>(v));   // quad_perm [1,0,3,2]
  v = op(v, dpp_move<0x4e>(v));   // quad_perm [2,3,0,1]
  v = op(v, dpp_move<0x114>(v));  // row_shr:4
  v = op(v, dpp_move<0x118>(v));  // row_shr:8
  v = op(v, dpp_move<0x142>(v));  // row_bcast:15
  v = op(v, dpp_move<0x143>(v));  // row_bcast:31
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double wave_sum(double v) { return wave_reduce<false>(v); }
__device__ __forceinline__ double wave_max(double v) { return wave_reduce<true>(v); }

// Block reduction; bit k of maxmask selects max (else sum).  Result in thread 0.
// Called by every thread of the block (the cross-lane moves need whole wavefronts).
template <int K, int NT>
__device__ __forceinline__ void block_reduce(double (&v)[K], unsigned maxmask) {
  __shared__ double sh[K][NT / 64];
  const int lane = threadIdx.x & 63, w = wave_id();
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = ((maxmask >> k) & 1) ? wave_max(v[k]) : wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][w] = v[k];
  }
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double a = lane < NT / 64 ? sh[k][lane] : 0.0;
      v[k] = ((maxmask >> k) & 1) ? wave_max(a) : wave_sum(a);
    }
  }
}

// Grid reductions run in two launches.  The producing kernel stores one
// partial per block (block_partial: partial[k * G + block], G blocks); a
// one-block finish kernel launched next on the stream combines them in block
// order (finish_reduce) and applies the solver logic.  There is no
// cross-block synchronisation inside a kernel: the L2 caches of the XCDs are
// not coherent with each other, so a device-scope fence per block would write
// back L2 every time; the kernel boundary makes the partials visible once.
constexpr int FINISH_THREADS = 1024;

template <int K, int NT = BLOCK>
__device__ __forceinline__ void block_partial(double (&v)[K], unsigned maxmask, double* partial) {
  block_reduce<K, NT>(v, maxmask);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partial[(size_t)k * nblocks() + block_id()] = v[k];
  }
}

// The same for large blocks with LDS to spare: every thread stores its values
// in sh (K * NT doubles, free for the caller's use again after the call), the
// first wavefront sums the NT / 64 values of each lane in thread order and
// reduces across lanes.  One cross-lane reduction per block instead of one
// per wavefront.
template <int K, int NT>
__device__ __forceinline__ void block_partial_lds(double (&v)[K], unsigned maxmask, double* partial, double* sh) {
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) sh[k * NT + threadIdx.x] = v[k];
  __syncthreads();
  if (threadIdx.x < 64) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const bool mx = (maxmask >> k) & 1;
      double a = sh[k * NT + threadIdx.x];
      for (int q = 1; q < NT / 64; ++q) {
        const double b = sh[k * NT + q * 64 + threadIdx.x];
        a = mx ? fmax(a, b) : a + b;
      }
      a = mx ? wave_max(a) : wave_sum(a);
      if (threadIdx.x == 0) partial[(size_t)k * nblocks() + block_id()] = a;
    }
  }
}

// In a FINISH_THREADS-thread block: totals of G partials into thread 0's a.
// Maxima are of non-negative quantities (start at 0).  A thread loads all its
// partials (blocks t, t + FINISH_THREADS, ...; up to FIN_U of them) before it
// adds any, so G <= FIN_U * FINISH_THREADS partials take one memory round trip;
// the additions keep block order.
constexpr int FIN_U = 4;
template <int K>
__device__ __forceinline__ void finish_reduce(double (&a)[K], const double* __restrict__ partial, int G,
                                              unsigned maxmask) {
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = 0.0;
  for (int b0 = threadIdx.x; b0 < G; b0 += FIN_U * FINISH_THREADS) {
    double x[FIN_U][K];
#pragma unroll
    for (int u = 0; u < FIN_U; ++u) {
      const int b = b0 + u * FINISH_THREADS;
#pragma unroll
      for (int k = 0; k < K; ++k) x[u][k] = b < G ? partial[(size_t)k * G + b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < FIN_U; ++u) {
      if (b0 + u * FINISH_THREADS >= G) break;
#pragma unroll
      for (int k = 0; k < K; ++k) a[k] = ((maxmask >> k) & 1) ? fmax(a[k], x[u][k]) : a[k] + x[u][k];
    }
  }
  block_reduce<K, FINISH_THREADS>(a, maxmask);
}

// Totals into s->red[dst + k].
template <int K>
__global__ void __launch_bounds__(FINISH_THREADS) k_fin_store(const double* __restrict__ partial, int G,
                                                              unsigned maxmask, int dst, Scal* s) {
  double a[K];
  finish_reduce<K>(a, partial, G, maxmask);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) s->red[dst + k] = a[k];
  }
}

// --------------------------------------------------------------------------- pairs, lane moves, 32-bit offsets
// Two adjacent cells of a row, moved with one load or store.
template <class T>
struct alignas(2 * sizeof(T)) Pair {
  T a, b;
};

// The value held by the previous / next lane of the wavefront (0 for lane 0 /
// lane 63): DPP wave_shr:1 / wave_shl:1 with bound_ctrl.
__device__ __forceinline__ int dpp_prev(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int dpp_next(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ float lane_prev(float v) { return __int_as_float(dpp_prev(__float_as_int(v))); }
__device__ __forceinline__ float lane_next(float v) { return __int_as_float(dpp_next(__float_as_int(v))); }
__device__ __forceinline__ double lane_prev(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_prev((int)(unsigned)b), hi = (unsigned)dpp_prev((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double lane_next(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_next((int)(unsigned)b), hi = (unsigned)dpp_next((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Tile kernels address their level's arrays with 32-bit byte offsets, so a load
// or store takes the array's base in scalar registers and one VGPR offset.  This
// holds because p2d_create rejects grids whose FP64 cell array (the largest array
// of any level) reaches 2^32 bytes (P2D_MAX_ARRAY_BYTES).
__device__ __forceinline__ unsigned cell_u(int i, int j, int nx) { return (unsigned)i + (unsigned)j * (unsigned)nx; }
template <class V, class T>
__device__ __forceinline__ V ld_at(const T* base, unsigned cell) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + cell * (unsigned)sizeof(T));
}
template <class V, class T>
__device__ __forceinline__ void st_at(T* base, unsigned cell, V v) {
  *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + cell * (unsigned)sizeof(T)) = v;
}

// --------------------------------------------------------------------------- ghost-fluid pieces
__device__ __forceinline__ bool minus_side(double phi) { return phi <= 0.0; }

// POISSON_SOLVER.cpp:743-746 with (L,R) = (left/bottom, right/top) cell:
// beta = vR vL (|phiL|+|phiR|) / (vR |phiL| + vL |phiR|); 0/0 -> vR.
__device__ __forceinline__ double gfm_beta(double vL, double vR, double phiL, double phiR) {
#pragma clang fp contract(off)
  const double aL = fabs(phiL), aR = fabs(phiR);
  const double den = vR * aL + vL * aR;
  if (den == 0.0) return vR;
  return vR * vL * (aL + aR) / den;
}

// POISSON_SOLVER.cpp:879: a = (jc_c |phi_n| + jc_n |phi_c|)/(|phi_c| + |phi_n|).
__device__ __forceinline__ double face_jump(double jcL, double jcR, double phiL, double phiR) {
#pragma clang fp contract(off)
  const double aL = fabs(phiL), aR = fabs(phiR);
  return (jcR * aL + jcL * aR) / (aR + aL);
}

// Jump part of the face gradient, s a (s = +1 if the R cell is minus side),
// 0 when both cells lie on one side or jc is absent.
__device__ __forceinline__ double face_sa(const double* phi, const double* jc, size_t kL, size_t kR) {
#pragma clang fp contract(off)
  if (!jc) return 0.0;
  const double pL = phi[kL], pR = phi[kR];
  if (minus_side(pL) == minus_side(pR)) return 0.0;
  const double a = face_jump(jc[kL], jc[kR], pL, pR);
  return minus_side(pR) ? a : -a;
}

// Double-double pressure (hi + lo).  The refinement loop keeps p in two FP64
// words so that face differences stay accurate when p carries large levels
// (pressure of light drops embedded in a heavy fluid, jumps of sigma kappa);
// a single FP64 word would round them at ~eps |p| and floor max|D u|.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
  s = a + b;
  const double bp = s - a;
  const double ap = s - bp;
  e = (a - ap) + (b - bp);
}
__device__ __forceinline__ void dd_add(double hi, double lo, double e, double& rh, double& rl) {
#pragma clang fp contract(off)
  double s, t;
  two_sum(hi, e, s, t);
  t += lo;
  rh = s + t;
  rl = t - (rh - s);
}

// Level set with linear-extrapolation ghosts (FIELD_STRUCTURE_2D.h:585-639);
// corners extrapolated from the y ghosts along x.
__device__ __forceinline__ double phi_y(const double* phi, int nx, int ny, int i, int j) {
#pragma clang fp contract(off)
  if (j < 0) return 2.0 * phi[ci(i, 0, nx)] - phi[ci(i, 1, nx)];
  if (j >= ny) return 2.0 * phi[ci(i, ny - 1, nx)] - phi[ci(i, ny - 2, nx)];
  return phi[ci(i, j, nx)];
}
__device__ __forceinline__ double phi_at(const double* phi, int nx, int ny, int i, int j) {
#pragma clang fp contract(off)
  if (i < 0) return 2.0 * phi_y(phi, nx, ny, 0, j) - phi_y(phi, nx, ny, 1, j);
  if (i >= nx) return 2.0 * phi_y(phi, nx, ny, nx - 1, j) - phi_y(phi, nx, ny, nx - 2, j);
  return phi_y(phi, nx, ny, i, j);
}

// Curvature, LEVELSET_2D.h:929-984: central differences, one-sided fallback,
// clamp to +-1/h.  Sets *bad when both gradients vanish (reference exit(0)).
__device__ double curvature_at(const double* phi, int nx, int ny, double h, int i, int j, bool* bad) {
#pragma clang fp contract(off)
  const double one_over_2dx = 1.0 / (2.0 * h), one_over_dx2 = 1.0 / (h * h), one_over_dx = 1.0 / h;
  const double c = phi_at(phi, nx, ny, i, j);
  const double l = phi_at(phi, nx, ny, i - 1, j), r = phi_at(phi, nx, ny, i + 1, j);
  const double b = phi_at(phi, nx, ny, i, j - 1), t = phi_at(phi, nx, ny, i, j + 1);
  const double rt = phi_at(phi, nx, ny, i + 1, j + 1), rb = phi_at(phi, nx, ny, i + 1, j - 1);
  const double lt = phi_at(phi, nx, ny, i - 1, j + 1), lb = phi_at(phi, nx, ny, i - 1, j - 1);
  double px = (r - l) * one_over_2dx;
  double py = (t - b) * one_over_2dx;
  const double pxx = (r - 2 * c + l) * one_over_dx2;
  const double pyy = (t - 2 * c + b) * one_over_dx2;
  const double pxy = (rt - rb - lt + lb) * one_over_2dx * one_over_2dx;
  double mag = sqrt(px * px + py * py);
  double deno = mag * mag * mag;
  double curv;
  if (deno != 0) {
    curv = -(px * px * pyy - 2.0 * px * py * pxy + py * py * pxx) * (1.0 / deno);
  } else {
    px = (r - c) * one_over_dx;
    py = (t - c) * one_over_dx;
    mag = sqrt(px * px + py * py);
    deno = mag * mag * mag;
    if (deno == 0) {
      *bad = true;
      return 0.0;
    }
    curv = -(px * px * pyy - 2.0 * px * py * pxy + py * py * pxx) * (1.0 / deno);
  }
  const double tol = 1.0 / h;
  if (fabs(curv) <= tol) return curv;
  return curv > tol ? tol : -tol;
}

}  // namespace p2d
