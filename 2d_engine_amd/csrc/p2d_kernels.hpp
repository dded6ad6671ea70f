// p2d device code for gfx950 (CDNA4): coefficient assembly, right-hand side,
// matrix-free GFM stencil, fused Krylov kernels, geometric multigrid, face
// update and the adjoint kernels.  Included once, by p2d.hip.
//
// Conventions (shared with include/p2d.h): nx x ny cells, i fastest; x-faces
// (nx+1) x ny, y-faces nx x (ny+1).  Solver operators are h^2-scaled:
// (A' x)_c = sum_f beta_f (x_c - x_nb), so A' = h^2 A with A = -D B G of the
// reference's variable-coefficient matrix (POISSON_SOLVER.cpp:749-872) without
// its pin: A' is singular on constants on every level, the right-hand side is
// taken mean-free and the coarsest level is solved with a pseudo-inverse.
//
// Reductions accumulate in FP64 and are deterministic: 64-lane wave shuffles,
// a fixed-order block tree, then the last block to finish (atomic ticket)
// sums the per-block partials in block order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p2d {

constexpr int BLOCK = 256;            // 4 wavefronts of 64
constexpr int MAX_RED_BLOCKS = 2048;  // partial buffer rows per reduced quantity
constexpr int SMALL_THREADS = 1024;   // single-workgroup coarse V-cycle

__device__ __forceinline__ size_t ci(int i, int j, int nx) { return (size_t)i + (size_t)j * nx; }
__device__ __forceinline__ size_t fxi(int i, int j, int nx) { return (size_t)i + (size_t)j * (nx + 1); }
__device__ __forceinline__ size_t fyi(int i, int j, int nx) { return (size_t)i + (size_t)j * nx; }

// Launch geometry.  Cell kernels: grid (ceil(nx/BLOCK), ny), i = blockIdx.x*BLOCK + tid,
// j = blockIdx.y (no integer division in the hot loops).  Reduction kernels:
// grid (ceil(nx/BLOCK), gy <= ny) looping rows j = blockIdx.y, blockIdx.y+gy, ...
__device__ __forceinline__ int nblocks() { return (int)(gridDim.x * gridDim.y); }
__device__ __forceinline__ int block_id() { return (int)(blockIdx.y * gridDim.x + blockIdx.x); }
__device__ __forceinline__ int cell_i() { return (int)(blockIdx.x * BLOCK + threadIdx.x); }

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
  double red[8];  // results of stand-alone reductions
  int iter, done, max_iter, flag;  // flag: 1 = breakdown / NaN
  int err;                         // coefficient assembly error (zero level-set gradient)
  int pad[3];
};

// --------------------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Block reduction; bit k of maxmask selects max (else sum).  Result in thread 0.
template <int K, int NT>
__device__ __forceinline__ void block_reduce(double (&v)[K], unsigned maxmask) {
  __shared__ double sh[K][NT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = ((maxmask >> k) & 1) ? wave_max(v[k]) : wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double a = sh[k][0];
      for (int q = 1; q < NT / 64; ++q) a = ((maxmask >> k) & 1) ? fmax(a, sh[k][q]) : a + sh[k][q];
      v[k] = a;
    }
  }
}

// Grid reduction.  Returns true in the last block; thread 0 of that block holds
// the totals in v.  partial must hold K * gridDim.x doubles; counter starts at 0
// and is left at 0.
template <int K>
__device__ bool grid_reduce(double (&v)[K], unsigned maxmask, double* partial, unsigned* counter) {
  block_reduce<K, BLOCK>(v, maxmask);
  __shared__ int last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partial[(size_t)k * nblocks() + block_id()] = v[k];
    __threadfence();
    const unsigned t = atomicAdd(counter, 1u);
    last = (t == (unsigned)nblocks() - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double a[K];
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = 0.0;
  const volatile double* vp = partial;
  const int G = nblocks();
  for (int b = threadIdx.x; b < G; b += BLOCK) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double x = vp[(size_t)k * G + b];
      a[k] = ((maxmask >> k) & 1) ? fmax(a[k], x) : a[k] + x;
    }
  }
  block_reduce<K, BLOCK>(a, maxmask);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = a[k];
    *counter = 0;
  }
  return true;
}

// --------------------------------------------------------------------------- GFM pieces
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
// words so that face differences (hi_R - hi_L) + (lo_R - lo_L) stay accurate
// when p carries large levels (pressure of light drops embedded in a heavy
// fluid); a single FP64 word would round them at ~eps |p| and floor max|D u|.
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
__device__ __forceinline__ double dd_diff(const double* hi, const double* lo, size_t kR, size_t kL) {
#pragma clang fp contract(off)
  return (hi[kR] - hi[kL]) + (lo[kR] - lo[kL]);
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
    if (deno == 0) { *bad = true; return 0.0; }
    curv = -(px * px * pyy - 2.0 * px * py * pxy + py * py * pxx) * (1.0 / deno);
  }
  const double tol = 1.0 / h;
  if (fabs(curv) <= tol) return curv;
  return curv > tol ? tol : -tol;
}

// --------------------------------------------------------------------------- assembly kernels
// Per cell: jc = sigma * kappa (PROJECTION_2D.h:830-834), beta on its left and
// bottom faces (and the right/top wall faces, = 0), v = 1/rho per side (A2).
__global__ void __launch_bounds__(BLOCK) k_coeff(int nx, int ny, double h, const double* __restrict__ phi,
                                                 double vm, double vp, double sigma, double* __restrict__ bx,
                                                 double* __restrict__ by, double* __restrict__ jc, Scal* s) {
  const int i = cell_i(), j = blockIdx.y;
  if (i >= nx) return;
  const size_t k = ci(i, j, nx);
  const double pc = phi[k];
  const double vc = minus_side(pc) ? vm : vp;
  if (i == 0) bx[fxi(0, j, nx)] = 0.0;
  else {
    const double pl = phi[k - 1];
    bx[fxi(i, j, nx)] = gfm_beta(minus_side(pl) ? vm : vp, vc, pl, pc);
  }
  if (i == nx - 1) bx[fxi(nx, j, nx)] = 0.0;
  if (j == 0) by[fyi(i, 0, nx)] = 0.0;
  else {
    const double pb = phi[k - nx];
    by[fyi(i, j, nx)] = gfm_beta(minus_side(pb) ? vm : vp, vc, pb, pc);
  }
  if (j == ny - 1) by[fyi(i, ny, nx)] = 0.0;
  double jv = 0.0;
  if (sigma != 0.0) {
    bool bad = false;
    const double kap = curvature_at(phi, nx, ny, h, i, j, &bad);
    if (bad) s->err = 1;
    jv = sigma * kap;
  }
  jc[k] = jv;
}

// Right-hand-side statistics: mean (removed by the residual kernel, A0 being
// singular on constants) and the norms of b - mean into s->red[0..1].
__device__ __forceinline__ void rhs_finalize(const double (&acc)[3], size_t N, Scal* s) {
  const double m = acc[2] / (double)N;
  s->bmean = m;
  s->red[0] = fmax(0.0, acc[0] - (double)N * m * m);
  s->red[1] = acc[1] + fabs(m);
}

// b' = h^2 (-D u/dt + D(B J)), the reference's b = -div - F (POISSON_SOLVER.cpp:951)
// with J_f = s_f a_f / h; also ||b'||_2^2 and ||b'||_inf into s->red[0..1].
__global__ void __launch_bounds__(BLOCK) k_rhs(int nx, int ny, double h, double dt, const double* __restrict__ u,
                                               const double* __restrict__ v, const double* __restrict__ phi,
                                               const double* __restrict__ jc, const double* __restrict__ bx,
                                               const double* __restrict__ by, double* __restrict__ b,
                                               double* partial, unsigned* counter, Scal* s) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h, idt = 1.0 / dt, h2 = h * h;
  double acc[3] = {0.0, 0.0, 0.0};
  const size_t N = (size_t)nx * ny;
  const int i = cell_i();
  for (int j = blockIdx.y; j < ny && i < nx; j += gridDim.y) {
    const size_t k = ci(i, j, nx);
    const double d = (ih * (u[fxi(i + 1, j, nx)] - u[fxi(i, j, nx)]) + ih * (v[fyi(i, j + 1, nx)] - v[fyi(i, j, nx)])) * idt;
    double Jl = 0, Jr = 0, Jb = 0, Jt = 0;
    if (jc) {
      if (i > 0) Jl = face_sa(phi, jc, k - 1, k) / h;
      if (i + 1 < nx) Jr = face_sa(phi, jc, k, k + 1) / h;
      if (j > 0) Jb = face_sa(phi, jc, k - nx, k) / h;
      if (j + 1 < ny) Jt = face_sa(phi, jc, k, k + nx) / h;
    }
    const double f = (bx[fxi(i + 1, j, nx)] * Jr - bx[fxi(i, j, nx)] * Jl + by[fyi(i, j + 1, nx)] * Jt -
                      by[fyi(i, j, nx)] * Jb) * ih;
    const double bv = h2 * (-d + f);
    b[k] = bv;
    acc[0] += bv * bv;
    acc[1] = fmax(acc[1], fabs(bv));
    acc[2] += bv;
  }
  if (grid_reduce<3>(acc, 2u, partial, counter) && threadIdx.x == 0) rhs_finalize(acc, N, s);
}

// --------------------------------------------------------------------------- consistent operator A0'
template <class T>
__device__ __forceinline__ T a0_apply(const T* __restrict__ bx, const T* __restrict__ by, const T* __restrict__ x,
                                      int i, int j, int nx, int ny) {
  const size_t c = ci(i, j, nx);
  const T xc = x[c];
  T s = 0;
  if (i > 0) s += bx[fxi(i, j, nx)] * (xc - x[c - 1]);
  if (i + 1 < nx) s += bx[fxi(i + 1, j, nx)] * (xc - x[c + 1]);
  if (j > 0) s += by[fyi(i, j, nx)] * (xc - x[c - nx]);
  if (j + 1 < ny) s += by[fyi(i, j + 1, nx)] * (xc - x[c + nx]);
  return s;
}

// y = scale * A0' x (test hook; scale = 1/h^2 gives the unscaled operator).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_apply(int nx, int ny, const T* __restrict__ bx, const T* __restrict__ by,
                                                 const T* __restrict__ x, T* __restrict__ y, double scale) {
  const int i = cell_i(), j = blockIdx.y;
  if (i >= nx) return;
  y[ci(i, j, nx)] = (T)((double)a0_apply<T>(bx, by, x, i, j, nx, ny) * scale);
}

// FP64 refinement step on the double-double pressure: p_new = p_old + e (e may
// be NULL), r = (b' - mean b') - A0' p_new with accurate face differences,
// stored as R; ||r||_2^2, ||r||_inf into s->red[0..1].
template <class TE, class R>
__global__ void __launch_bounds__(BLOCK) k_correct_resid(int nx, int ny, const double* __restrict__ bx,
                                                         const double* __restrict__ by, const double* __restrict__ b,
                                                         const double* __restrict__ ph, const double* __restrict__ pl,
                                                         const TE* __restrict__ e, double* __restrict__ qh,
                                                         double* __restrict__ ql, R* __restrict__ r,
                                                         double* partial, unsigned* counter, Scal* s) {
#pragma clang fp contract(off)
  double acc[2] = {0.0, 0.0};
  const double bm = s->bmean;
  const int i = cell_i();
  for (int j = blockIdx.y; j < ny && i < nx; j += gridDim.y) {
    const size_t k = ci(i, j, nx);
    auto P = [&](size_t q, double& h, double& l) {
      if (e) dd_add(ph[q], pl[q], (double)e[q], h, l);
      else { h = ph[q]; l = pl[q]; }
    };
    double ch, cl;
    P(k, ch, cl);
    double a = 0;
    auto arm = [&](double beta, size_t q) {
      double nh, nl;
      P(q, nh, nl);
      a += beta * ((ch - nh) + (cl - nl));
    };
    if (i > 0) arm(bx[fxi(i, j, nx)], k - 1);
    if (i + 1 < nx) arm(bx[fxi(i + 1, j, nx)], k + 1);
    if (j > 0) arm(by[fyi(i, j, nx)], k - nx);
    if (j + 1 < ny) arm(by[fyi(i, j + 1, nx)], k + nx);
    const double rv = (b[k] - bm) - a;
    if (qh) {
      qh[k] = ch;
      ql[k] = cl;
    }
    r[k] = (R)rv;
    acc[0] += rv * rv;
    acc[1] = fmax(acc[1], fabs(rv));
  }
  if (grid_reduce<2>(acc, 2u, partial, counter) && threadIdx.x == 0) {
    s->red[0] = acc[0];
    s->red[1] = acc[1];
  }
}

// --------------------------------------------------------------------------- Krylov (PCG) kernels
// Initialise the inner solve from the norms left in s->red by the residual kernel.
__global__ void k_inner_init(Scal* s, double rel2, double abs2, double ainf, int max_iter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  s->r0_2 = s->red[0];
  s->r2 = s->red[0];
  s->rinf = s->red[1];
  s->rel2 = rel2;
  s->abs2 = abs2;
  s->ainf = ainf;
  s->max_iter = max_iter;
  s->iter = 0;
  s->flag = 0;
  s->rz = s->rz_old = s->pq = s->alpha = s->beta = 0.0;
  const bool met = (s->r2 <= abs2 && s->rinf <= ainf) || s->r0_2 == 0.0;
  s->done = (met || max_iter <= 0) ? 1 : 0;
  if (!(s->r2 == s->r2)) { s->done = 1; s->flag = 1; }
}

// z is taken mean-free (z - mean z): the preconditioner's output on the constant
// null space of A0 is dropped, so search directions and the solution stay
// mean-free and the FP64 residual keeps its full precision.
// rz = r . (z - mean z) ; beta = rz / rz_old (0 on the first iteration).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_dot_rz(size_t N, const T* __restrict__ r, const T* __restrict__ z,
                                                  double* partial, unsigned* counter, Scal* s) {
  if (s->done) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < N; k += (size_t)gridDim.x * BLOCK) {
    const double rv = (double)r[k], zv = (double)z[k];
    acc[0] += rv * zv;
    acc[1] += zv;
    acc[2] += rv;
  }
  if (grid_reduce<3>(acc, 0u, partial, counter) && threadIdx.x == 0) {
    const double zm = acc[1] / (double)N;
    const double rz = acc[0] - zm * acc[2];
    s->zmean = zm;
    s->rz = rz;
    s->beta = s->iter == 0 ? 0.0 : rz / s->rz_old;
  }
}

// p_new = z + beta p_old (FP64 arithmetic, stored as T), q = A0' p_new, and
// alpha = rz / (p_new . q).  The neighbours' p_new are recomputed from z and
// p_old, so p_old and p_new must be distinct buffers.
template <class T>
__global__ void __launch_bounds__(BLOCK) k_dir_apply(int nx, int ny, const T* __restrict__ bx,
                                                     const T* __restrict__ by, const T* __restrict__ z,
                                                     const T* __restrict__ p_old, T* __restrict__ p_new,
                                                     T* __restrict__ q, double* partial, unsigned* counter, Scal* s) {
  if (s->done) return;
  const double beta = s->beta, zm = s->zmean;
  const bool first = s->iter == 0;
  double acc[1] = {0.0};
  const int i = cell_i();
  for (int j = blockIdx.y; j < ny && i < nx; j += gridDim.y) {
    const size_t k = ci(i, j, nx);
    auto P = [&](size_t c) {
      return first ? (T)((double)z[c] - zm) : (T)(((double)z[c] - zm) + beta * (double)p_old[c]);
    };
    const T pc = P(k);
    T a = 0;
    if (i > 0) a += bx[fxi(i, j, nx)] * (pc - P(k - 1));
    if (i + 1 < nx) a += bx[fxi(i + 1, j, nx)] * (pc - P(k + 1));
    if (j > 0) a += by[fyi(i, j, nx)] * (pc - P(k - nx));
    if (j + 1 < ny) a += by[fyi(i, j + 1, nx)] * (pc - P(k + nx));
    p_new[k] = pc;
    q[k] = a;
    acc[0] += (double)pc * (double)a;
  }
  if (grid_reduce<1>(acc, 0u, partial, counter) && threadIdx.x == 0) {
    s->pq = acc[0];
    if (!(acc[0] > 0.0)) { s->done = 1; s->flag = 1; s->alpha = 0.0; }
    else s->alpha = s->rz / acc[0];
  }
}

// x += alpha p, r -= alpha q; norms; convergence decision.
template <class T>
__global__ void __launch_bounds__(BLOCK) k_update(size_t N, T* __restrict__ x, T* __restrict__ r,
                                                  const T* __restrict__ p, const T* __restrict__ q, double* partial,
                                                  unsigned* counter, Scal* s) {
  if (s->done) return;
  const double alpha = s->alpha;
  double acc[2] = {0.0, 0.0};
  for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < N; k += (size_t)gridDim.x * BLOCK) {
    const T xn = (T)((double)x[k] + alpha * (double)p[k]);
    const T rn = (T)((double)r[k] - alpha * (double)q[k]);
    x[k] = xn;
    r[k] = rn;
    acc[0] += (double)rn * (double)rn;
    acc[1] = fmax(acc[1], fabs((double)rn));
  }
  if (grid_reduce<2>(acc, 2u, partial, counter) && threadIdx.x == 0) {
    s->r2 = acc[0];
    s->rinf = acc[1];
    s->rz_old = s->rz;
    s->iter += 1;
    const bool met = acc[0] <= s->rel2 * s->r0_2 || (acc[0] <= s->abs2 && acc[1] <= s->ainf);
    if (!(acc[0] == acc[0])) s->flag = 1;
    if (met || s->iter >= s->max_iter || s->flag) s->done = 1;
  }
}

// --------------------------------------------------------------------------- multigrid
template <class T>
struct LevelDesc {
  int nx, ny;
  const T* bx;
  const T* by;
  const T* wx;  // interpolation weight on the parent coarse cell, per cell of this (fine) level
  const T* wy;
  T* x;
  T* r;
  T* res;
};

template <class T>
__device__ __forceinline__ T lvl_diag(const LevelDesc<T>& L, int i, int j) {
  const int nx = L.nx;
  return L.bx[fxi(i, j, nx)] + L.bx[fxi(i + 1, j, nx)] + L.by[fyi(i, j, nx)] + L.by[fyi(i, j + 1, nx)];
}
template <class T>
__device__ __forceinline__ T lvl_off(const LevelDesc<T>& L, const T* __restrict__ x, int i, int j) {
  const int nx = L.nx, ny = L.ny;
  const size_t c = ci(i, j, nx);
  T s = 0;
  if (i > 0) s += L.bx[fxi(i, j, nx)] * x[c - 1];
  if (i + 1 < nx) s += L.bx[fxi(i + 1, j, nx)] * x[c + 1];
  if (j > 0) s += L.by[fyi(i, j, nx)] * x[c - nx];
  if (j + 1 < ny) s += L.by[fyi(i, j + 1, nx)] * x[c + nx];
  return s;
}
// Red-black Gauss-Seidel update of one cell; zero = treat x as 0 (first sweep from x = 0).
template <class T>
__device__ __forceinline__ void gs_cell(const LevelDesc<T>& L, const T* r, T* x, int i, int j, bool zero) {
  const size_t c = ci(i, j, L.nx);
  const T off = zero ? (T)0 : lvl_off(L, x, i, j);
  x[c] = (r[c] + off) / lvl_diag(L, i, j);
}
template <class T>
__device__ __forceinline__ void resid_cell(const LevelDesc<T>& L, const T* r, const T* x, int i, int j) {
  const size_t c = ci(i, j, L.nx);
  L.res[c] = r[c] - (lvl_diag(L, i, j) * x[c] - lvl_off(L, x, i, j));
}
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
// Coarse right-hand side = P^T res (exact transpose of prolongation).
template <class T>
__device__ __forceinline__ void restrict_cell(const LevelDesc<T>& F, const LevelDesc<T>& Cc, int I, int J) {
  T s = 0;
  for (int j = 2 * J - 1; j <= 2 * J + 2; ++j) {
    if (j < 0 || j >= F.ny) continue;
    for (int i = 2 * I - 1; i <= 2 * I + 2; ++i) {
      if (i < 0 || i >= F.nx) continue;
      const size_t c = ci(i, j, F.nx);
      const T wy = w1<T>(j, J, Cc.ny, F.wy[c]);
      const T wx = w1<T>(i, I, Cc.nx, F.wx[c]);
      if (wx == (T)0 || wy == (T)0) continue;
      s += wx * wy * F.res[c];
    }
  }
  Cc.r[ci(I, J, Cc.nx)] = s;
}
template <class T>
__device__ __forceinline__ void prolong_cell(const LevelDesc<T>& F, const LevelDesc<T>& Cc, T* x, int i, int j) {
  const int I = i >> 1, J = j >> 1;
  int I2 = I + ((i & 1) ? 1 : -1), J2 = J + ((j & 1) ? 1 : -1);
  I2 = I2 < 0 ? 0 : (I2 > Cc.nx - 1 ? Cc.nx - 1 : I2);
  J2 = J2 < 0 ? 0 : (J2 > Cc.ny - 1 ? Cc.ny - 1 : J2);
  const size_t c = ci(i, j, F.nx);
  const T wx = F.wx[c], wy = F.wy[c];
  const T* xc = Cc.x;
  const T e = wx * wy * xc[ci(I, J, Cc.nx)] + ((T)1 - wx) * wy * xc[ci(I2, J, Cc.nx)] +
              wx * ((T)1 - wy) * xc[ci(I, J2, Cc.nx)] + ((T)1 - wx) * ((T)1 - wy) * xc[ci(I2, J2, Cc.nx)];
  x[c] += e;
}

// Grid-wide level kernels.  Colour c: cells with (i+j)&1 == c; one thread per such cell.
template <class T>
__global__ void __launch_bounds__(BLOCK) k_gs(LevelDesc<T> L, const T* __restrict__ r, T* __restrict__ x, int color,
                                              int zero, const Scal* s) {
  if (s->done) return;
  const int j = blockIdx.y, m = cell_i();
  const int i = 2 * m + ((j + color) & 1);
  if (i >= L.nx) return;
  gs_cell<T>(L, r, x, i, j, zero != 0);
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_resid(LevelDesc<T> L, const T* __restrict__ r, const T* __restrict__ x,
                                                 const Scal* s) {
  if (s->done) return;
  const int i = cell_i();
  if (i >= L.nx) return;
  resid_cell<T>(L, r, x, i, blockIdx.y);
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_restrict(LevelDesc<T> F, LevelDesc<T> Cc, const Scal* s) {
  if (s->done) return;
  const int I = cell_i();
  if (I >= Cc.nx) return;
  restrict_cell<T>(F, Cc, I, blockIdx.y);
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_prolong(LevelDesc<T> F, LevelDesc<T> Cc, T* __restrict__ x, const Scal* s) {
  if (s->done) return;
  const int i = cell_i();
  if (i >= F.nx) return;
  prolong_cell<T>(F, Cc, x, i, blockIdx.y);
}

// Coarse levels l0..nl-1 of the V-cycle in one workgroup (levels small enough
// that a grid launch would be pure launch latency).  r_in / x_out replace the
// r / x buffers of level l0.
template <class T>
__global__ void __launch_bounds__(SMALL_THREADS) k_vcycle_small(const LevelDesc<T>* __restrict__ Ls, int l0, int nl,
                                                                const T* r_in, T* x_out, int nu_pre, int nu_post,
                                                                const T* __restrict__ binv, const Scal* s) {
  if (s->done) return;
  const int tid = threadIdx.x;
  for (int l = l0; l < nl - 1; ++l) {
    const LevelDesc<T> L = Ls[l];
    const T* r = (l == l0) ? r_in : L.r;
    T* x = (l == l0) ? x_out : L.x;
    const int half = (L.nx + 1) >> 1;
    for (int sw = 0; sw < nu_pre; ++sw)
      for (int color = 0; color < 2; ++color) {
        for (int t = tid; t < half * L.ny; t += SMALL_THREADS) {
          const int j = t / half, i = 2 * (t % half) + ((j + color) & 1);
          if (i < L.nx) gs_cell<T>(L, r, x, i, j, sw == 0 && color == 0);
        }
        __syncthreads();
      }
    for (int t = tid; t < L.nx * L.ny; t += SMALL_THREADS) resid_cell<T>(L, r, x, t % L.nx, t / L.nx);
    __syncthreads();
    const LevelDesc<T> Cc = Ls[l + 1];
    for (int t = tid; t < Cc.nx * Cc.ny; t += SMALL_THREADS) restrict_cell<T>(L, Cc, t % Cc.nx, t / Cc.nx);
    __syncthreads();
  }
  {  // bottom: dense pseudo-inverse of the coarsest operator
    const LevelDesc<T> B = Ls[nl - 1];
    const T* r = (nl - 1 == l0) ? r_in : B.r;
    T* x = (nl - 1 == l0) ? x_out : B.x;
    const int n = B.nx * B.ny;
    for (int a = tid; a < n; a += SMALL_THREADS) {
      double acc = 0;
      for (int c = 0; c < n; ++c) acc += (double)binv[(size_t)a * n + c] * (double)r[c];
      x[a] = (T)acc;
    }
    __syncthreads();
  }
  for (int l = nl - 2; l >= l0; --l) {
    const LevelDesc<T> L = Ls[l];
    const LevelDesc<T> Cc = Ls[l + 1];
    const T* r = (l == l0) ? r_in : L.r;
    T* x = (l == l0) ? x_out : L.x;
    for (int t = tid; t < L.nx * L.ny; t += SMALL_THREADS) prolong_cell<T>(L, Cc, x, t % L.nx, t / L.nx);
    __syncthreads();
    const int half = (L.nx + 1) >> 1;
    for (int sw = 0; sw < nu_post; ++sw)
      for (int cc = 0; cc < 2; ++cc) {
        const int color = 1 - cc;
        for (int t = tid; t < half * L.ny; t += SMALL_THREADS) {
          const int j = t / half, i = 2 * (t % half) + ((j + color) & 1);
          if (i < L.nx) gs_cell<T>(L, r, x, i, j, false);
        }
        __syncthreads();
      }
  }
}

// --------------------------------------------------------------------------- multigrid setup (FP64)
// Operator-dependent parent weight (resistance ratio along the fine row/column),
// 3/4 for constant coefficients, 1 at a wall.  b(k) = fine face k (0..nf).
__device__ double od_weight(const double* bface, size_t stride, int i, int nf) {
  auto b = [&](int k) { return bface[(size_t)k * stride]; };
  const int P = i >> 1;
  const double near_ = 0.5 / b(2 * P + 1);
  double far_;
  if ((i & 1) == 0) {
    if (P == 0) return 1.0;
    far_ = 1.0 / b(2 * P) + 0.5 / b(2 * P - 1);
  } else {
    if (2 * P + 2 >= nf) return 1.0;
    far_ = 1.0 / b(2 * P + 2) + 0.5 / b(2 * P + 3);
  }
  if (!(near_ < 1e300) || !(far_ < 1e300)) return far_ >= near_ ? 1.0 : 0.0;
  return far_ / (near_ + far_);
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_mg_weights(int nx, int ny, const double* __restrict__ bx,
                                                      const double* __restrict__ by, T* __restrict__ wx,
                                                      T* __restrict__ wy) {
  const int i = cell_i(), j = blockIdx.y;
  if (i >= nx) return;
  const size_t k = ci(i, j, nx);
  const int Q = (i >> 1) + ((i & 1) ? 1 : -1), Qj = (j >> 1) + ((j & 1) ? 1 : -1);
  double a = od_weight(bx + fxi(0, j, nx), 1, i, nx);
  double b = od_weight(by + fyi(i, 0, nx), (size_t)nx, j, ny);
  if (Q < 0 || Q > (nx >> 1) - 1) a = 1.0;
  if (Qj < 0 || Qj > (ny >> 1) - 1) b = 1.0;
  wx[k] = (T)a;
  wy[k] = (T)b;
}
// Coarse face = average over the two fine rows of the series conductance
// 2 / (0.5/b_{k-1} + 1/b_k + 0.5/b_{k+1}) between the two coarse centres.
__device__ __forceinline__ double coarse_face(const double* b, size_t stride, int k, int nf) {
  if (k <= 0 || k >= nf) return 0.0;
  const double r = 0.5 / b[(size_t)(k - 1) * stride] + 1.0 / b[(size_t)k * stride] + 0.5 / b[(size_t)(k + 1) * stride];
  return 2.0 / r;
}
__global__ void __launch_bounds__(BLOCK) k_mg_coarsen(int fnx, int fny, const double* __restrict__ fbx,
                                                      const double* __restrict__ fby, double* __restrict__ cbx,
                                                      double* __restrict__ cby) {
  const int cnx = fnx >> 1, cny = fny >> 1;
  const size_t nfx = (size_t)(cnx + 1) * cny, nfy = (size_t)cnx * (cny + 1);
  const size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k < nfx) {
    const int kk = (int)k, I = kk % (cnx + 1), J = kk / (cnx + 1);
    cbx[k] = 0.5 * (coarse_face(fbx + fxi(0, 2 * J, fnx), 1, 2 * I, fnx) +
                    coarse_face(fbx + fxi(0, 2 * J + 1, fnx), 1, 2 * I, fnx));
  } else if (k < nfx + nfy) {
    const size_t q = k - nfx;
    const int qq = (int)q, I = qq % cnx, J = qq / cnx;
    cby[q] = 0.5 * (coarse_face(fby + fyi(2 * I, 0, fnx), (size_t)fnx, 2 * J, fny) +
                    coarse_face(fby + fyi(2 * I + 1, 0, fnx), (size_t)fnx, 2 * J, fny));
  }
}
// Pseudo-inverse of the singular bottom operator A (null space = constants):
// A+ = (A + a 1 1^T)^-1 - 1 1^T / (a n^2) with a n = mean diagonal; the shifted
// matrix is SPD, inverted by Gauss-Jordan without pivoting in one workgroup
// (work is n x 2n in global memory).
__global__ void __launch_bounds__(SMALL_THREADS) k_mg_bottom_inverse(int nx, int ny, const double* __restrict__ bx,
                                                                     const double* __restrict__ by, double* work,
                                                                     double* inv64, float* inv32, Scal* s) {
  const int n = nx * ny, w = 2 * n, tid = threadIdx.x;
  __shared__ double alpha_sh;
  if (tid == 0) {
    double dsum = 0;
    for (int c = 0; c < n; ++c) {
      const int i = c % nx, j = c / nx;
      dsum += bx[fxi(i, j, nx)] + bx[fxi(i + 1, j, nx)] + by[fyi(i, j, nx)] + by[fyi(i, j + 1, nx)];
    }
    alpha_sh = dsum / n / n;
  }
  __syncthreads();
  const double alpha = alpha_sh;
  for (int q = tid; q < n * w; q += SMALL_THREADS) work[q] = (q % w) < n ? alpha : 0.0;
  __syncthreads();
  for (int c = tid; c < n; c += SMALL_THREADS) {
    const int i = c % nx, j = c / nx;
    const double d = bx[fxi(i, j, nx)] + bx[fxi(i + 1, j, nx)] + by[fyi(i, j, nx)] + by[fyi(i, j + 1, nx)];
    double* row = work + (size_t)c * w;
    row[c] += d;
    if (i > 0) row[c - 1] -= bx[fxi(i, j, nx)];
    if (i + 1 < nx) row[c + 1] -= bx[fxi(i + 1, j, nx)];
    if (j > 0) row[c - nx] -= by[fyi(i, j, nx)];
    if (j + 1 < ny) row[c + nx] -= by[fyi(i, j + 1, nx)];
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
      if (r == k) continue;
      const double f = work[(size_t)r * w + k];
      if (c == k) continue;
      work[q] -= f * work[(size_t)k * w + c];
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

template <class A, class B>
__global__ void __launch_bounds__(BLOCK) k_cast(size_t n, const A* __restrict__ a, B* __restrict__ b) {
  const size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k < n) b[k] = (B)a[k];
}

// --------------------------------------------------------------------------- face update / diagnostics
// Consistent A14: u_f -= dt beta_f (p_R - p_L + s a)/h on interior faces (walls untouched).
__global__ void __launch_bounds__(BLOCK) k_face_update(int nx, int ny, double h, double dt,
                                                       const double* __restrict__ ph, const double* __restrict__ pl,
                                                       const double* __restrict__ phi,
                                                       const double* __restrict__ jc, const double* __restrict__ bx,
                                                       const double* __restrict__ by, double* __restrict__ u,
                                                       double* __restrict__ v) {
#pragma clang fp contract(off)
  const int i = cell_i(), j = blockIdx.y;
  if (i >= nx) return;
  const size_t k = ci(i, j, nx);
  const double ih = 1.0 / h;
  if (i > 0) {
    double gap = dd_diff(ph, pl, k, k - 1);
    gap += face_sa(phi, jc, k - 1, k);
    const size_t f = fxi(i, j, nx);
    u[f] -= dt * gap * bx[f] * ih;
  }
  if (j > 0) {
    double gap = dd_diff(ph, pl, k, k - nx);
    gap += face_sa(phi, jc, k - nx, k);
    const size_t f = fyi(i, j, nx);
    v[f] -= dt * gap * by[f] * ih;
  }
}

// max |D u| (cell divergence, not divided by dt), max |u|, max |v|, and the sum of
// p (for the mean) into s->red[0..3].
__global__ void __launch_bounds__(BLOCK) k_diagnostics(int nx, int ny, double h, const double* __restrict__ u,
                                                       const double* __restrict__ v, const double* __restrict__ p,
                                                       double* partial, unsigned* counter, Scal* s) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int i = cell_i();
  for (int j = blockIdx.y; j <= ny && i <= nx; j += gridDim.y) {
    if (j < ny) acc[1] = fmax(acc[1], fabs(u[fxi(i, j, nx)]));
    if (i < nx) acc[2] = fmax(acc[2], fabs(v[fyi(i, j, nx)]));
    if (i < nx && j < ny) {
      const double d = ih * (u[fxi(i + 1, j, nx)] - u[fxi(i, j, nx)]) + ih * (v[fyi(i, j + 1, nx)] - v[fyi(i, j, nx)]);
      acc[0] = fmax(acc[0], fabs(d));
      if (p) acc[3] += p[ci(i, j, nx)];
    }
  }
  if (grid_reduce<4>(acc, 7u, partial, counter) && threadIdx.x == 0) {
    for (int q = 0; q < 4; ++q) s->red[q] = acc[q];
  }
}

// Sum of p into s->red[3] (for the mean).
__global__ void __launch_bounds__(BLOCK) k_psum(size_t N, const double* __restrict__ ph, const double* __restrict__ pl,
                                                double* partial, unsigned* counter, Scal* s) {
  double acc[1] = {0.0};
  for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < N; k += (size_t)gridDim.x * BLOCK) acc[0] += ph[k] + pl[k];
  if (grid_reduce<1>(acc, 0u, partial, counter) && threadIdx.x == 0) s->red[3] = acc[0];
}

// Wall-normal faces carry no flux: beta = 0 there.
__global__ void __launch_bounds__(BLOCK) k_zero_walls(int nx, int ny, double* bx, double* by) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t < ny) {
    bx[fxi(0, t, nx)] = 0.0;
    bx[fxi(nx, t, nx)] = 0.0;
  }
  if (t < nx) {
    by[fyi(t, 0, nx)] = 0.0;
    by[fyi(t, ny, nx)] = 0.0;
  }
}

// out = p - shift, where shift = s->red[3] / N (mean) or p[0] (pin).
__global__ void __launch_bounds__(BLOCK) k_shift(size_t N, const double* __restrict__ ph, const double* __restrict__ pl,
                                                 double* __restrict__ out, int mean_zero, const Scal* s) {
#pragma clang fp contract(off)
  const size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k >= N) return;
  if (mean_zero) out[k] = (ph[k] - s->red[3] / (double)N) + pl[k];
  else out[k] = (ph[k] - ph[0]) + (pl[k] - pl[0]);
}

// --------------------------------------------------------------------------- adjoint
// b' = -h^2 D(B g) with g taken on interior faces only; norms into s->red[0..1].
__global__ void __launch_bounds__(BLOCK) k_adj_rhs(int nx, int ny, double h, const double* __restrict__ gu,
                                                   const double* __restrict__ gv, const double* __restrict__ bx,
                                                   const double* __restrict__ by, double* __restrict__ b,
                                                   double* partial, unsigned* counter, Scal* s) {
#pragma clang fp contract(off)
  const double ih = 1.0 / h, h2 = h * h;
  double acc[3] = {0.0, 0.0, 0.0};
  const size_t N = (size_t)nx * ny;
  const int i = cell_i();
  for (int j = blockIdx.y; j < ny && i < nx; j += gridDim.y) {
    const size_t k = ci(i, j, nx);
    const double gl = i > 0 ? bx[fxi(i, j, nx)] * gu[fxi(i, j, nx)] : 0.0;
    const double gr = i + 1 < nx ? bx[fxi(i + 1, j, nx)] * gu[fxi(i + 1, j, nx)] : 0.0;
    const double gb = j > 0 ? by[fyi(i, j, nx)] * gv[fyi(i, j, nx)] : 0.0;
    const double gt = j + 1 < ny ? by[fyi(i, j + 1, nx)] * gv[fyi(i, j + 1, nx)] : 0.0;
    const double bv = h2 * (-((gr - gl) * ih + (gt - gb) * ih));
    b[k] = bv;
    acc[0] += bv * bv;
    acc[1] = fmax(acc[1], fabs(bv));
    acc[2] += bv;
  }
  if (grid_reduce<3>(acc, 2u, partial, counter) && threadIdx.x == 0) rhs_finalize(acc, N, s);
}

// du = g - G mu, dbeta = -dt du (G p + s a / h) on interior faces; walls 0.
__global__ void __launch_bounds__(BLOCK) k_adj_faces(int nx, int ny, double h, double dt,
                                                     const double* __restrict__ gu, const double* __restrict__ gv,
                                                     const double* __restrict__ muh, const double* __restrict__ mul,
                                                     const double* __restrict__ ph, const double* __restrict__ pl,
                                                     const double* __restrict__ phi, const double* __restrict__ jc,
                                                     double* du, double* dv, double* dbx, double* dby) {
#pragma clang fp contract(off)
  const int i = cell_i(), j = blockIdx.y;
  if (i >= nx) return;
  const size_t k = ci(i, j, nx);
  const double ih = 1.0 / h;
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

}  // namespace p2d
