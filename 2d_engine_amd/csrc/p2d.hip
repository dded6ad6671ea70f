// p2d host library: context, solver orchestration and the C ABI of include/p2d.h.
//
// One projection (PROJECTION_2D::Solve(dt), PROJECTION_2D.h:235-257) is
//   k_coeff          face beta + sigma*kappa from phi        (A2, A4, A5, A8 beta)
//   multigrid setup  coarse beta, interpolation weights, bottom inverse
//   k_rhs            b' = h^2 (-D u/dt + D(B J))             (A3, A8 rhs)
//   solve            FP64 refinement around an inner MG-PCG in FP64 or FP32
//                    (replaces CSR + CG/PCG: A9-A13, A15)
//   k_face_update    consistent pressure-gradient face update (A14)
//   k_diagnostics    FP64 max|D u|, max|u|, max|v|
// all on the context's stream.  The inner Krylov loop keeps alpha/beta and the
// convergence flag on the device; the host enqueues iterations in chunks and
// reads the flag once per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "p2d.h"
#include "p2d_kernels.hpp"

using namespace p2d;

namespace {

constexpr int SMALL_CELLS = 16384;  // levels with <= 128^2 cells run in the single-workgroup kernel
constexpr int BOTTOM_MAX = 64;      // coarsen until the grid has <= 64 cells (or a side is odd)
constexpr int BOTTOM_LIMIT = 1024;  // largest bottom grid accepted (dense inverse)

inline unsigned blocks_for(size_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }
inline unsigned red_blocks(size_t n) { return (unsigned)std::min<size_t>(blocks_for(n), MAX_RED_BLOCKS); }
// Cell kernels: one row of BLOCK cells per block, one block row per grid row.
inline dim3 cgrid(int nx, int ny) { return dim3(blocks_for((size_t)nx), (unsigned)ny); }
// Reduction kernels: at most MAX_RED_BLOCKS blocks, rows looped.
inline dim3 rgrid(int nx, int ny) {
  const unsigned gx = blocks_for((size_t)nx);
  const unsigned gy = std::min<unsigned>((unsigned)ny, std::max(1u, (unsigned)MAX_RED_BLOCKS / gx));
  return dim3(gx, gy);
}
// Red-black sweeps: one thread per same-colour cell of a row.
inline dim3 gsgrid(int nx, int ny) { return dim3(blocks_for((size_t)((nx + 1) >> 1)), (unsigned)ny); }

template <class T>
struct MGData {
  int L = 0;
  int small_l0 = 0;
  bool valid = false;
  std::vector<LevelDesc<T>> lv;
  LevelDesc<T>* d_lv = nullptr;
  T* binv = nullptr;
};

template <class T>
struct KVec {  // inner Krylov vectors on the finest level
  T *x = nullptr, *r = nullptr, *z = nullptr, *q = nullptr;
  T* p[2] = {nullptr, nullptr};
};

}  // namespace

struct p2d_ctx {
  p2d_config cfg;
  int nx = 0, ny = 0;
  size_t N = 0, NU = 0, NV = 0;
  double h = 0;
  hipStream_t own = nullptr, stream = nullptr;
  std::vector<void*> allocs;

  // coefficients (FP64 masters) and their FP32 copies
  double *phi = nullptr, *jc = nullptr, *bx = nullptr, *by = nullptr;
  float *bx32 = nullptr, *by32 = nullptr;
  bool have_fluid = false, explicit_coeff = false, has_jc = true, coeff_ready = false, f32_ready = false;
  double rho_m = 1, rho_p = 1, sigma = 0;

  // right-hand side, solutions, staging
  double* b = nullptr;
  double* ph[2] = {nullptr, nullptr};  // pressure, double-double (hi, lo), double-buffered
  double* pl[2] = {nullptr, nullptr};
  int pcur = 0;
  double* muh[2] = {nullptr, nullptr};  // adjoint solution, same layout
  double* mul[2] = {nullptr, nullptr};
  double *su = nullptr, *sv = nullptr, *sg1 = nullptr, *sg2 = nullptr, *sp = nullptr;
  double* sa[4] = {nullptr, nullptr, nullptr, nullptr};  // adjoint output staging (host calls)
  bool have_pressure = false;
  double last_dt = 1.0;

  KVec<double> v64;
  KVec<float> v32;
  MGData<double> mg64;
  MGData<float> mg32;
  std::vector<double*> cbx, cby;  // FP64 master coefficients per level (level 0 = bx/by)
  double* bottom_work = nullptr;
  double* binv64 = nullptr;
  float* binv32 = nullptr;

  double* partial = nullptr;
  unsigned* counter = nullptr;
  Scal* s = nullptr;
  Scal* hs = nullptr;  // pinned host mirror
  int last_iters[2] = {0, 0};
};

namespace {

#define HIPCHK(call)                                                                                       \
  do {                                                                                                     \
    hipError_t e_ = (call);                                                                                \
    if (e_ != hipSuccess) {                                                                                \
      fprintf(stderr, "p2d: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return P2D_HIP_ERROR;                                                                                \
    }                                                                                                      \
  } while (0)
#define CHK(call)                    \
  do {                               \
    int rc_ = (call);                \
    if (rc_ != P2D_OK) return rc_;   \
  } while (0)

int launched_ok() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "p2d: kernel launch failed: %s\n", hipGetErrorString(e));
    return P2D_HIP_ERROR;
  }
  return P2D_OK;
}

template <class T>
int dalloc(p2d_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
  HIPCHK(hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T)));
  c->allocs.push_back(q);
  *p = (T*)q;
  return P2D_OK;
}

int num_levels(int nx, int ny, int max_levels) {
  int n = 1;
  while ((nx % 2 == 0) && (ny % 2 == 0) && (long)nx * ny > BOTTOM_MAX && (max_levels <= 0 || n < max_levels)) {
    nx /= 2;
    ny /= 2;
    ++n;
  }
  return n;
}

int sync_scal(p2d_ctx* c) {
  HIPCHK(hipMemcpyAsync(c->hs, c->s, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

template <class T>
int alloc_mg(p2d_ctx* c, MGData<T>& mg, const T* bx0, const T* by0) {
  const int L = num_levels(c->nx, c->ny, c->cfg.mg_max_levels);
  mg.L = L;
  mg.lv.resize(L);
  mg.small_l0 = L - 1;
  for (int l = 0; l < L; ++l) {
    LevelDesc<T>& d = mg.lv[l];
    d.nx = c->nx >> l;
    d.ny = c->ny >> l;
    const size_t n = (size_t)d.nx * d.ny;
    if (l == 0) {
      d.bx = bx0;
      d.by = by0;
    } else {
      T *a, *b;
      CHK(dalloc(c, &a, (size_t)(d.nx + 1) * d.ny));
      CHK(dalloc(c, &b, (size_t)d.nx * (d.ny + 1)));
      d.bx = a;
      d.by = b;
    }
    T *wx, *wy, *x, *r, *res;
    CHK(dalloc(c, &wx, n));
    CHK(dalloc(c, &wy, n));
    CHK(dalloc(c, &x, n));
    CHK(dalloc(c, &r, n));
    CHK(dalloc(c, &res, n));
    d.wx = wx;
    d.wy = wy;
    d.x = x;
    d.r = r;
    d.res = res;
    if ((long)n <= SMALL_CELLS && l < mg.small_l0) mg.small_l0 = l;
  }
  CHK(dalloc(c, &mg.d_lv, (size_t)L));
  HIPCHK(hipMemcpy(mg.d_lv, mg.lv.data(), sizeof(LevelDesc<T>) * L, hipMemcpyHostToDevice));
  return P2D_OK;
}

// Coarse FP64 coefficients and the bottom pseudo-inverse (both precisions).
int mg_setup_common(p2d_ctx* c) {
  const int L = c->mg64.L;
  for (int l = 0; l + 1 < L; ++l) {
    const int fnx = c->nx >> l, fny = c->ny >> l;
    const size_t nf = (size_t)((fnx >> 1) + 1) * (fny >> 1) + (size_t)(fnx >> 1) * ((fny >> 1) + 1);
    k_mg_coarsen<<<blocks_for(nf), BLOCK, 0, c->stream>>>(fnx, fny, c->cbx[l], c->cby[l], c->cbx[l + 1],
                                                           c->cby[l + 1]);
  }
  const LevelDesc<double>& B = c->mg64.lv[L - 1];
  k_mg_bottom_inverse<<<1, SMALL_THREADS, 0, c->stream>>>(B.nx, B.ny, c->cbx[L - 1], c->cby[L - 1], c->bottom_work,
                                                          c->binv64, c->binv32, c->s);
  return launched_ok();
}

template <class T>
int mg_setup_prec(p2d_ctx* c, MGData<T>& mg) {
  const int L = mg.L;
  for (int l = 0; l + 1 < L; ++l) {
    const LevelDesc<T>& d = mg.lv[l];
    k_mg_weights<T><<<cgrid(d.nx, d.ny), BLOCK, 0, c->stream>>>(d.nx, d.ny, c->cbx[l], c->cby[l],
                                                                              const_cast<T*>(d.wx),
                                                                              const_cast<T*>(d.wy));
  }
  return launched_ok();
}

// FP32 copies of every level's coefficients.
int mg_cast32(p2d_ctx* c) {
  for (int l = 0; l < c->mg32.L; ++l) {
    const LevelDesc<float>& d = c->mg32.lv[l];
    const size_t a = (size_t)(d.nx + 1) * d.ny, b = (size_t)d.nx * (d.ny + 1);
    k_cast<double, float><<<blocks_for(a), BLOCK, 0, c->stream>>>(a, c->cbx[l], const_cast<float*>(d.bx));
    k_cast<double, float><<<blocks_for(b), BLOCK, 0, c->stream>>>(b, c->cby[l], const_cast<float*>(d.by));
  }
  return launched_ok();
}

int ensure_mg(p2d_ctx* c, int prec) {
  if (!c->coeff_ready) return P2D_BAD_ARG;
  if (!c->mg64.valid && !c->mg32.valid) CHK(mg_setup_common(c));
  if (prec == P2D_PREC_F64 && !c->mg64.valid) {
    CHK(mg_setup_prec(c, c->mg64));
    c->mg64.valid = true;
  }
  if (prec == P2D_PREC_F32 && !c->mg32.valid) {
    if (!c->mg64.valid) {
      // common setup already ran above (or with mg64); keep the master coefficients
    }
    CHK(mg_cast32(c));
    CHK(mg_setup_prec(c, c->mg32));
    c->mg32.valid = true;
    c->f32_ready = true;
  }
  return P2D_OK;
}

void invalidate(p2d_ctx* c) {
  c->mg64.valid = false;
  c->mg32.valid = false;
  c->f32_ready = false;
}

template <class T>
int vcycle(p2d_ctx* c, MGData<T>& mg, const T* r0, T* z0) {
  const int nu1 = c->cfg.mg_pre_sweeps, nu2 = c->cfg.mg_post_sweeps;
  const int l0 = mg.small_l0;
  for (int l = 0; l < l0; ++l) {
    const LevelDesc<T>& L = mg.lv[l];
    const T* r = l == 0 ? r0 : L.r;
    T* x = l == 0 ? z0 : L.x;
    const dim3 gb = gsgrid(L.nx, L.ny);
    for (int sw = 0; sw < nu1; ++sw) {
      k_gs<T><<<gb, BLOCK, 0, c->stream>>>(L, r, x, 0, sw == 0, c->s);
      k_gs<T><<<gb, BLOCK, 0, c->stream>>>(L, r, x, 1, 0, c->s);
    }
    k_resid<T><<<cgrid(L.nx, L.ny), BLOCK, 0, c->stream>>>(L, r, x, c->s);
    const LevelDesc<T>& Cc = mg.lv[l + 1];
    k_restrict<T><<<cgrid(Cc.nx, Cc.ny), BLOCK, 0, c->stream>>>(L, Cc, c->s);
  }
  {
    const T* r = l0 == 0 ? r0 : mg.lv[l0].r;
    T* x = l0 == 0 ? z0 : mg.lv[l0].x;
    k_vcycle_small<T><<<1, SMALL_THREADS, 0, c->stream>>>(mg.d_lv, l0, mg.L, r, x, nu1, nu2, mg.binv, c->s);
  }
  for (int l = l0 - 1; l >= 0; --l) {
    const LevelDesc<T>& L = mg.lv[l];
    const LevelDesc<T>& Cc = mg.lv[l + 1];
    const T* r = l == 0 ? r0 : L.r;
    T* x = l == 0 ? z0 : L.x;
    k_prolong<T><<<cgrid(L.nx, L.ny), BLOCK, 0, c->stream>>>(L, Cc, x, c->s);
    const dim3 gb = gsgrid(L.nx, L.ny);
    for (int sw = 0; sw < nu2; ++sw) {
      k_gs<T><<<gb, BLOCK, 0, c->stream>>>(L, r, x, 1, 0, c->s);
      k_gs<T><<<gb, BLOCK, 0, c->stream>>>(L, r, x, 0, 0, c->s);
    }
  }
  return P2D_OK;
}

template <class T>
__global__ void k_fill(size_t n, T* a, T v) {
  const size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k < n) a[k] = v;
}
template <class T>
__global__ void k_pattern(size_t n, T* a) {
  const size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k < n) a[k] = (T)(((k * 2654435761u) % 1000u) * 1e-3 - 0.5);
}
__global__ void k_scal_arm(Scal* s, int iter, double beta) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    s->done = 0;
    s->iter = iter;
    s->beta = beta;
    s->rz = 1.0;
    s->flag = 0;
  }
}

// Inner MG-PCG on A0' e = r (r in v.r, mutated), e in v.x.
template <class T>
int inner_solve(p2d_ctx* c, MGData<T>& mg, KVec<T>& v, const T* bx, const T* by, double rel2, double abs2,
                double ainf, int max_iter, int slot, int* iters) {
  const size_t N = c->N;
  k_inner_init<<<1, 1, 0, c->stream>>>(c->s, rel2, abs2, ainf, max_iter);
  k_fill<T><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, v.x, (T)0);
  const unsigned rb = red_blocks(N);
  const dim3 rg = rgrid(c->nx, c->ny);
  int launched = 0;
  int chunk = std::max(2, c->last_iters[slot] > 0 ? c->last_iters[slot] : 8);
  while (true) {
    for (int k = 0; k < chunk; ++k) {
      const int it = launched + k;
      CHK(vcycle<T>(c, mg, v.r, v.z));
      k_dot_rz<T><<<rb, BLOCK, 0, c->stream>>>(N, v.r, v.z, c->partial, c->counter, c->s);
      k_dir_apply<T><<<rg, BLOCK, 0, c->stream>>>(c->nx, c->ny, bx, by, v.z, v.p[it & 1], v.p[(it + 1) & 1], v.q,
                                                 c->partial, c->counter, c->s);
      k_update<T><<<rb, BLOCK, 0, c->stream>>>(N, v.x, v.r, v.p[(it + 1) & 1], v.q, c->partial, c->counter, c->s);
    }
    CHK(launched_ok());
    launched += chunk;
    CHK(sync_scal(c));
    if (c->hs->done) break;
    chunk = std::max(2, std::min(8, chunk / 2));
  }
  *iters = c->hs->iter;
  c->last_iters[slot] = c->hs->iter;
  if (c->hs->flag) return P2D_NAN;
  return P2D_OK;
}

struct SolveResult {
  int status = P2D_OK, iterations = 0, outer = 0;
  double r2 = 0, rinf = 0;
};

// FP64 iterative refinement: x_out = sol[*cur] solves A0' x = b' (b' scaled).
// rtol: ||r||_2 <= rtol ||b||_2 ; atol: ||r'||_inf <= atol (scaled units); <= 0 disables.
int solve(p2d_ctx* c, const double* b, double b2, double* solh[2], double* soll[2], int* cur, double rtol,
          double atol, SolveResult* out) {
  const size_t N = c->N;
  const bool mixed = c->cfg.precision == P2D_MIXED;
  const dim3 rb = rgrid(c->nx, c->ny);
  CHK(ensure_mg(c, mixed ? P2D_PREC_F32 : P2D_PREC_F64));
  k_fill<double><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, solh[0], 0.0);
  k_fill<double><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, soll[0], 0.0);
  *cur = 0;
  // r = b - A0' * 0
  if (mixed)
    k_correct_resid<float, float><<<rb, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bx, c->by, b, solh[0], soll[0],
                                                              nullptr, nullptr, nullptr, c->v32.r, c->partial,
                                                              c->counter, c->s);
  else
    k_correct_resid<double, double><<<rb, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bx, c->by, b, solh[0], soll[0],
                                                                nullptr, nullptr, nullptr, c->v64.r, c->partial,
                                                                c->counter, c->s);
  CHK(launched_ok());
  CHK(sync_scal(c));
  double r2 = c->hs->red[0], rinf = c->hs->red[1];
  const double big = 1e300;
  auto converged = [&](double a2, double ainf) {
    return (rtol <= 0 || a2 <= rtol * rtol * b2) && (atol <= 0 || ainf <= atol);
  };
  const double abs2 = rtol > 0 ? 0.25 * rtol * rtol * b2 : big;
  const double ainf = atol > 0 ? 0.5 * atol : big;
  int total = 0, outer = 0, stalled = 0;
  double best_r2 = r2, best_rinf = rinf;
  int status = P2D_OK;
  const bool verbose = std::getenv("P2D_VERBOSE") != nullptr;
  if (verbose)
    fprintf(stderr, "p2d solve: %s b2=%.3e rtol=%.1e atol=%.3e r0: r2=%.3e rinf=%.3e\n", mixed ? "mixed" : "fp64", b2,
            rtol, atol, r2, rinf);
  while (!converged(r2, rinf) && outer < c->cfg.max_outer) {
    const int budget = c->cfg.max_iterations - total;
    if (budget <= 0) break;
    int it = 0;
    if (mixed) {
      // The FP32 recursive residual is not trusted against the absolute targets: the
      // inner solve stops on relative reduction only, no deeper than this pass needs.
      double need = 1.0;
      if (rtol > 0) need = std::min(need, std::sqrt(abs2 / std::max(r2, 1e-300)));
      if (atol > 0) need = std::min(need, ainf / std::max(rinf, 1e-300));
      const double rel = std::min(0.5, std::max(c->cfg.inner_rtol, 0.3 * need));
      status = inner_solve<float>(c, c->mg32, c->v32, c->bx32, c->by32, rel * rel, -1.0, -1.0, budget, 1, &it);
      k_correct_resid<float, float><<<rb, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bx, c->by, b, solh[*cur],
                                                                soll[*cur], c->v32.x, solh[1 - *cur],
                                                                soll[1 - *cur], c->v32.r, c->partial, c->counter,
                                                                c->s);
    } else {
      status = inner_solve<double>(c, c->mg64, c->v64, c->bx, c->by, 1e-28, abs2, ainf, budget, 0, &it);
      k_correct_resid<double, double><<<rb, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bx, c->by, b, solh[*cur],
                                                                  soll[*cur], c->v64.x, solh[1 - *cur],
                                                                  soll[1 - *cur], c->v64.r, c->partial,
                                                                  c->counter, c->s);
    }
    CHK(launched_ok());
    *cur = 1 - *cur;
    total += it;
    ++outer;
    CHK(sync_scal(c));
    r2 = c->hs->red[0];
    rinf = c->hs->red[1];
    if (verbose)
      fprintf(stderr, "  pass %d: inner %d its (status %d) -> true r2=%.3e (rel %.3e) rinf=%.3e\n", outer, it, status,
              r2, std::sqrt(r2 / std::max(b2, 1e-300)), rinf);
    if (status != P2D_OK) break;
    if (!(r2 == r2)) {
      status = P2D_NAN;
      break;
    }
    // stagnation: two passes without halving the true residual (FP64 floor reached)
    if (r2 > 0.25 * best_r2 && rinf > 0.5 * best_rinf) {
      if (++stalled >= 2) break;
    } else {
      stalled = 0;
    }
    best_r2 = std::min(best_r2, r2);
    best_rinf = std::min(best_rinf, rinf);
  }
  if (status == P2D_OK && !converged(r2, rinf)) status = P2D_NOT_CONVERGED;
  out->status = status;
  out->iterations = total;
  out->outer = outer;
  out->r2 = r2;
  out->rinf = rinf;
  return P2D_OK;
}

int stage_in(p2d_ctx* c, double* dst, const double* src, size_t n, int mem) {
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), mem == P2D_MEM_HOST ? hipMemcpyHostToDevice
                                                                          : hipMemcpyDeviceToDevice,
                        c->stream));
  return P2D_OK;
}
int stage_out(p2d_ctx* c, double* dst, const double* src, size_t n, int mem) {
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), mem == P2D_MEM_HOST ? hipMemcpyDeviceToHost
                                                                          : hipMemcpyDeviceToDevice,
                        c->stream));
  return P2D_OK;
}

int make_f32_level0(p2d_ctx* c) {
  if (c->f32_ready) return P2D_OK;
  k_cast<double, float><<<blocks_for(c->NU), BLOCK, 0, c->stream>>>(c->NU, c->bx, c->bx32);
  k_cast<double, float><<<blocks_for(c->NV), BLOCK, 0, c->stream>>>(c->NV, c->by, c->by32);
  c->f32_ready = true;
  return launched_ok();
}

}  // namespace

// =========================================================================== C ABI
extern "C" {

void p2d_config_default(p2d_config* cfg, int nx, int ny, double h) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->nx = nx;
  cfg->ny = ny;
  cfg->h = h;
  cfg->precision = P2D_FP64;
  cfg->mg_max_levels = 0;
  cfg->mg_pre_sweeps = 2;
  cfg->mg_post_sweeps = 2;
  cfg->rtol = 1e-10;
  cfg->div_tol = 1e-10;
  cfg->max_iterations = 1000;
  cfg->max_outer = 12;
  cfg->inner_rtol = 1e-4;
}

const char* p2d_status_string(int st) {
  switch (st) {
    case P2D_OK: return "ok";
    case P2D_NOT_CONVERGED: return "not converged";
    case P2D_NAN: return "NaN or breakdown";
    case P2D_BAD_ARG: return "bad argument";
    case P2D_HIP_ERROR: return "HIP error";
    case P2D_NO_DEVICE: return "no device";
    default: return "unknown";
  }
}

int p2d_create(const p2d_config* cfg, p2d_ctx** out) {
  if (!cfg || !out) return P2D_BAD_ARG;
  *out = nullptr;
  if (cfg->nx < 2 || cfg->ny < 2 || cfg->nx > 65535 || cfg->ny > 65535 || !(cfg->h > 0) || cfg->mg_pre_sweeps < 1 || cfg->mg_post_sweeps < 1 ||
      cfg->max_iterations < 1 || cfg->max_outer < 1 || (cfg->precision != P2D_FP64 && cfg->precision != P2D_MIXED) ||
      (!(cfg->rtol > 0) && !(cfg->div_tol > 0)) || !(cfg->inner_rtol > 0 && cfg->inner_rtol < 1))
    return P2D_BAD_ARG;
  {
    const int L = num_levels(cfg->nx, cfg->ny, cfg->mg_max_levels);
    const long nb = (long)(cfg->nx >> (L - 1)) * (cfg->ny >> (L - 1));
    if (nb > BOTTOM_LIMIT) return P2D_BAD_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return P2D_NO_DEVICE;
  p2d_ctx* c = new p2d_ctx;
  c->cfg = *cfg;
  c->nx = cfg->nx;
  c->ny = cfg->ny;
  c->h = cfg->h;
  c->N = (size_t)c->nx * c->ny;
  c->NU = (size_t)(c->nx + 1) * c->ny;
  c->NV = (size_t)c->nx * (c->ny + 1);
  auto fail = [&](int rc) {
    p2d_destroy(c);
    return rc;
  };
  if (hipStreamCreate(&c->own) != hipSuccess) return fail(P2D_HIP_ERROR);
  c->stream = c->own;
  const size_t N = c->N;
  int rc = P2D_OK;
#define A(p, n) \
  if ((rc = dalloc(c, &(p), (n))) != P2D_OK) return fail(rc)
  A(c->phi, N);
  A(c->jc, N);
  A(c->bx, c->NU);
  A(c->by, c->NV);
  A(c->bx32, c->NU);
  A(c->by32, c->NV);
  A(c->b, N);
  for (int q = 0; q < 2; ++q) {
    A(c->ph[q], N);
    A(c->pl[q], N);
    A(c->muh[q], N);
    A(c->mul[q], N);
  }
  A(c->su, c->NU);
  A(c->sv, c->NV);
  A(c->sg1, c->NU);
  A(c->sg2, c->NV);
  A(c->sp, std::max(c->NU, c->NV));
  A(c->sa[0], c->NU);
  A(c->sa[1], c->NV);
  A(c->sa[2], c->NU);
  A(c->sa[3], c->NV);
  A(c->v64.x, N);
  A(c->v64.r, N);
  A(c->v64.z, N);
  A(c->v64.q, N);
  A(c->v64.p[0], N);
  A(c->v64.p[1], N);
  A(c->v32.x, N);
  A(c->v32.r, N);
  A(c->v32.z, N);
  A(c->v32.q, N);
  A(c->v32.p[0], N);
  A(c->v32.p[1], N);
  A(c->partial, (size_t)8 * MAX_RED_BLOCKS);
  A(c->counter, 4);
  A(c->s, 1);
#undef A
  if ((rc = alloc_mg<double>(c, c->mg64, c->bx, c->by)) != P2D_OK) return fail(rc);
  if ((rc = alloc_mg<float>(c, c->mg32, c->bx32, c->by32)) != P2D_OK) return fail(rc);
  c->cbx.resize(c->mg64.L);
  c->cby.resize(c->mg64.L);
  for (int l = 0; l < c->mg64.L; ++l) {
    c->cbx[l] = const_cast<double*>(c->mg64.lv[l].bx);
    c->cby[l] = const_cast<double*>(c->mg64.lv[l].by);
  }
  {
    const LevelDesc<double>& B = c->mg64.lv[c->mg64.L - 1];
    const size_t nb = (size_t)B.nx * B.ny;
    if ((rc = dalloc(c, &c->bottom_work, 2 * nb * nb)) != P2D_OK) return fail(rc);
    if ((rc = dalloc(c, &c->binv64, nb * nb)) != P2D_OK) return fail(rc);
    if ((rc = dalloc(c, &c->binv32, nb * nb)) != P2D_OK) return fail(rc);
    c->mg64.binv = c->binv64;
    c->mg32.binv = c->binv32;
  }
  if (hipHostMalloc((void**)&c->hs, sizeof(Scal), hipHostMallocDefault) != hipSuccess) return fail(P2D_HIP_ERROR);
  std::memset(c->hs, 0, sizeof(Scal));
  *out = c;
  return P2D_OK;
}

int p2d_destroy(p2d_ctx* c) {
  if (!c) return P2D_BAD_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->hs) (void)hipHostFree(c->hs);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
  return P2D_OK;
}

int p2d_set_stream(p2d_ctx* c, void* stream) {
  if (!c) return P2D_BAD_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->stream = stream ? (hipStream_t)stream : c->own;
  return P2D_OK;
}

int p2d_set_fluid(p2d_ctx* c, const double* phi, double rho_minus, double rho_plus, double sigma, int mem) {
  if (!c || !phi || !(rho_minus > 0) || !(rho_plus > 0) || !std::isfinite(sigma)) return P2D_BAD_ARG;
  CHK(stage_in(c, c->phi, phi, c->N, mem));
  c->rho_m = rho_minus;
  c->rho_p = rho_plus;
  c->sigma = sigma;
  c->have_fluid = true;
  c->explicit_coeff = false;
  c->has_jc = true;
  c->coeff_ready = false;
  invalidate(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_set_coefficients(p2d_ctx* c, const double* beta_x, const double* beta_y, const double* phi,
                         const double* jc, int mem) {
  if (!c || !beta_x || !beta_y || !phi) return P2D_BAD_ARG;
  CHK(stage_in(c, c->bx, beta_x, c->NU, mem));
  CHK(stage_in(c, c->by, beta_y, c->NV, mem));
  CHK(stage_in(c, c->phi, phi, c->N, mem));
  if (jc) CHK(stage_in(c, c->jc, jc, c->N, mem));
  c->has_jc = jc != nullptr;
  // wall faces carry no flux
  k_zero_walls<<<blocks_for((size_t)std::max(c->nx, c->ny)), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bx, c->by);
  CHK(launched_ok());
  c->explicit_coeff = true;
  c->have_fluid = true;
  c->coeff_ready = true;
  invalidate(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_project(p2d_ctx* c, double* u, double* v, double dt, p2d_stats* st, int mem) {
  if (!c || !u || !v || !(dt > 0) || !c->have_fluid) return P2D_BAD_ARG;
  double* du = u;
  double* dv = v;
  if (mem == P2D_MEM_HOST) {
    CHK(stage_in(c, c->su, u, c->NU, mem));
    CHK(stage_in(c, c->sv, v, c->NV, mem));
    du = c->su;
    dv = c->sv;
  }
  c->hs->err = 0;
  if (!c->explicit_coeff) {
    HIPCHK(hipMemsetAsync(&c->s->err, 0, sizeof(int), c->stream));
    k_coeff<<<cgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, c->phi, 1.0 / c->rho_m, 1.0 / c->rho_p,
                                                   c->sigma, c->bx, c->by, c->jc, c->s);
    CHK(launched_ok());
    c->coeff_ready = true;
    invalidate(c);
  }
  if (c->cfg.precision == P2D_MIXED) CHK(make_f32_level0(c));
  const double* jc = c->has_jc ? c->jc : nullptr;
  k_rhs<<<rgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, du, dv, c->phi, jc, c->bx, c->by, c->b,
                                                c->partial, c->counter, c->s);
  CHK(launched_ok());
  CHK(sync_scal(c));
  const bool bad_geom = c->hs->err != 0;
  const double b2 = c->hs->red[0];
  const double atol = c->cfg.div_tol > 0 ? 0.5 * c->cfg.div_tol * c->h * c->h / dt : 0.0;
  SolveResult sr;
  CHK(solve(c, c->b, b2, c->ph, c->pl, &c->pcur, c->cfg.rtol, atol, &sr));
  k_face_update<<<cgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, c->ph[c->pcur], c->pl[c->pcur],
                                                              c->phi, jc, c->bx, c->by, du, dv);
  k_diagnostics<<<rgrid(c->nx + 1, c->ny + 1), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, du, dv, nullptr,
                                                                      c->partial, c->counter, c->s);
  CHK(launched_ok());
  if (mem == P2D_MEM_HOST) {
    CHK(stage_out(c, u, c->su, c->NU, mem));
    CHK(stage_out(c, v, c->sv, c->NV, mem));
  }
  CHK(sync_scal(c));
  c->have_pressure = true;
  c->last_dt = dt;
  if (st) {
    st->status = bad_geom ? P2D_NAN : sr.status;
    st->iterations = sr.iterations;
    st->outer = sr.outer;
    st->rel_residual = b2 > 0 ? std::sqrt(sr.r2 / b2) : 0.0;
    st->max_abs_residual = sr.rinf / (c->h * c->h);
    st->max_div = c->hs->red[0];
    st->max_u = c->hs->red[1];
    st->max_v = c->hs->red[2];
  }
  if (bad_geom) return P2D_NAN;
  return sr.status;
}

int p2d_get_pressure(p2d_ctx* c, double* p, int mean_zero, int mem) {
  if (!c || !p || !c->have_pressure) return P2D_BAD_ARG;
  const size_t N = c->N;
  const double* sh = c->ph[c->pcur];
  const double* sl = c->pl[c->pcur];
  k_psum<<<red_blocks(N), BLOCK, 0, c->stream>>>(N, sh, sl, c->partial, c->counter, c->s);
  double* dst = mem == P2D_MEM_HOST ? c->sp : p;
  k_shift<<<blocks_for(N), BLOCK, 0, c->stream>>>(N, sh, sl, dst, mean_zero, c->s);
  CHK(launched_ok());
  if (mem == P2D_MEM_HOST) CHK(stage_out(c, p, c->sp, N, mem));
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_get_coefficients(p2d_ctx* c, double* bx, double* by, double* jc, int mem) {
  if (!c || !c->coeff_ready) return P2D_BAD_ARG;
  if (bx) CHK(stage_out(c, bx, c->bx, c->NU, mem));
  if (by) CHK(stage_out(c, by, c->by, c->NV, mem));
  if (jc) CHK(stage_out(c, jc, c->jc, c->N, mem));
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_adjoint(p2d_ctx* c, const double* gu, const double* gv, double* du, double* dv, double* dbx, double* dby,
                p2d_stats* st, int mem) {
  if (!c || !gu || !gv || !c->have_pressure) return P2D_BAD_ARG;
  const double* g1 = gu;
  const double* g2 = gv;
  if (mem == P2D_MEM_HOST) {
    CHK(stage_in(c, c->sg1, gu, c->NU, mem));
    CHK(stage_in(c, c->sg2, gv, c->NV, mem));
    g1 = c->sg1;
    g2 = c->sg2;
  }
  if (c->cfg.precision == P2D_MIXED) CHK(make_f32_level0(c));
  k_adj_rhs<<<rgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, g1, g2, c->bx, c->by, c->b, c->partial,
                                                    c->counter, c->s);
  CHK(launched_ok());
  CHK(sync_scal(c));
  const double b2 = c->hs->red[0];
  SolveResult sr;
  int mcur = 0;
  const double rt = c->cfg.rtol > 0 ? c->cfg.rtol : 1e-10;
  CHK(solve(c, c->b, b2, c->muh, c->mul, &mcur, rt, 0.0, &sr));
  const double* jc = c->has_jc ? c->jc : nullptr;
  double *o1 = du, *o2 = dv, *o3 = dbx, *o4 = dby;
  if (mem == P2D_MEM_HOST) {
    o1 = du ? c->sa[0] : nullptr;
    o2 = dv ? c->sa[1] : nullptr;
    o3 = dbx ? c->sa[2] : nullptr;
    o4 = dby ? c->sa[3] : nullptr;
  }
  k_adj_faces<<<cgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, c->last_dt, g1, g2, c->muh[mcur],
                                                           c->mul[mcur], c->ph[c->pcur], c->pl[c->pcur], c->phi,
                                                           jc, o1, o2, o3, o4);
  CHK(launched_ok());
  if (mem == P2D_MEM_HOST) {
    if (du) CHK(stage_out(c, du, o1, c->NU, mem));
    if (dv) CHK(stage_out(c, dv, o2, c->NV, mem));
    if (dbx) CHK(stage_out(c, dbx, o3, c->NU, mem));
    if (dby) CHK(stage_out(c, dby, o4, c->NV, mem));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  if (st) {
    st->status = sr.status;
    st->iterations = sr.iterations;
    st->outer = sr.outer;
    st->rel_residual = b2 > 0 ? std::sqrt(sr.r2 / b2) : 0.0;
    st->max_abs_residual = sr.rinf / (c->h * c->h);
    st->max_div = 0;
    st->max_u = st->max_v = 0;
  }
  return sr.status;
}

int p2d_apply_stencil(p2d_ctx* c, const void* x, void* y, int prec, int mem) {
  if (!c || !x || !y || !c->coeff_ready || (prec != P2D_PREC_F64 && prec != P2D_PREC_F32)) return P2D_BAD_ARG;
  const size_t N = c->N;
  const double scale = 1.0 / (c->h * c->h);
  const auto kind = mem == P2D_MEM_HOST;
  if (prec == P2D_PREC_F64) {
    const double* xi = (const double*)x;
    double* yo = (double*)y;
    if (kind) {
      HIPCHK(hipMemcpyAsync(c->v64.z, x, N * 8, hipMemcpyHostToDevice, c->stream));
      xi = c->v64.z;
      yo = c->v64.q;
    }
    k_apply<double><<<cgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bx, c->by, xi, yo, scale);
    CHK(launched_ok());
    if (kind) HIPCHK(hipMemcpyAsync(y, c->v64.q, N * 8, hipMemcpyDeviceToHost, c->stream));
  } else {
    CHK(make_f32_level0(c));
    const float* xi = (const float*)x;
    float* yo = (float*)y;
    if (kind) {
      HIPCHK(hipMemcpyAsync(c->v32.z, x, N * 4, hipMemcpyHostToDevice, c->stream));
      xi = c->v32.z;
      yo = c->v32.q;
    }
    k_apply<float><<<cgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bx32, c->by32, xi, yo, scale);
    CHK(launched_ok());
    if (kind) HIPCHK(hipMemcpyAsync(y, c->v32.q, N * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_vcycle(p2d_ctx* c, const void* r, void* z, int prec, int mem) {
  if (!c || !r || !z || !c->coeff_ready || (prec != P2D_PREC_F64 && prec != P2D_PREC_F32)) return P2D_BAD_ARG;
  const size_t N = c->N;
  const bool host = mem == P2D_MEM_HOST;
  if (prec == P2D_PREC_F32) CHK(make_f32_level0(c));
  CHK(ensure_mg(c, prec));
  k_scal_arm<<<1, 1, 0, c->stream>>>(c->s, 0, 0.0);
  if (prec == P2D_PREC_F64) {
    const double* ri = (const double*)r;
    double* zo = (double*)z;
    if (host) {
      HIPCHK(hipMemcpyAsync(c->v64.r, r, N * 8, hipMemcpyHostToDevice, c->stream));
      ri = c->v64.r;
      zo = c->v64.z;
    }
    CHK(vcycle<double>(c, c->mg64, ri, zo));
    CHK(launched_ok());
    if (host) HIPCHK(hipMemcpyAsync(z, c->v64.z, N * 8, hipMemcpyDeviceToHost, c->stream));
  } else {
    const float* ri = (const float*)r;
    float* zo = (float*)z;
    if (host) {
      HIPCHK(hipMemcpyAsync(c->v32.r, r, N * 4, hipMemcpyHostToDevice, c->stream));
      ri = c->v32.r;
      zo = c->v32.z;
    }
    CHK(vcycle<float>(c, c->mg32, ri, zo));
    CHK(launched_ok());
    if (host) HIPCHK(hipMemcpyAsync(z, c->v32.z, N * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_mg_levels(p2d_ctx* c) { return c ? c->mg64.L : -1; }

int p2d_time_kernel(p2d_ctx* c, int which, int reps, double* avg_ms) {
  if (!c || !avg_ms || reps < 1 || which < 0 || which > 3 || !c->coeff_ready) return P2D_BAD_ARG;
  const size_t N = c->N;
  const dim3 rb = rgrid(c->nx, c->ny);
  CHK(make_f32_level0(c));
  // deterministic non-constant inputs so p.Ap > 0 and no kernel takes its early exit
  k_pattern<float><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v32.z);
  k_pattern<float><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v32.p[0]);
  k_pattern<double><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v64.z);
  k_pattern<double><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v64.p[0]);
  k_pattern<float><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v32.x);
  k_pattern<double><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->ph[0]);
  k_fill<double><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->pl[0], 0.0);
  k_scal_arm<<<1, 1, 0, c->stream>>>(c->s, 1, 0.25);
  CHK(ensure_mg(c, P2D_PREC_F32));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const LevelDesc<float>& L0 = c->mg32.lv[0];
  auto launch = [&]() {
    switch (which) {
      case 0:
        k_dir_apply<float><<<rb, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bx32, c->by32, c->v32.z, c->v32.p[0],
                                                         c->v32.p[1], c->v32.q, c->partial, c->counter, c->s);
        break;
      case 1:
        k_dir_apply<double><<<rb, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bx, c->by, c->v64.z, c->v64.p[0],
                                                          c->v64.p[1], c->v64.q, c->partial, c->counter, c->s);
        break;
      case 2:
        k_gs<float><<<gsgrid(L0.nx, L0.ny), BLOCK, 0, c->stream>>>(L0, c->v32.p[0],
                                                                                            c->v32.z, 0, 0, c->s);
        break;
      default:
        k_correct_resid<float, float><<<rb, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bx, c->by, c->b, c->ph[0],
                                                                  c->pl[0], c->v32.x, c->ph[1], c->pl[1], c->v32.r,
                                                                  c->partial, c->counter, c->s);
    }
  };
  k_scal_arm<<<1, 1, 0, c->stream>>>(c->s, 1, 0.25);
  launch();  // warm
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int k = 0; k < reps; ++k) launch();
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  CHK(launched_ok());
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / reps;
  // the pattern overwrote solver state
  c->have_pressure = false;
  return P2D_OK;
}

}  // extern "C"
