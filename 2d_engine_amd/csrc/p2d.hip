// p2d host library: context, solver orchestration and the C ABI of include/p2d.h.
//
// One projection (PROJECTION_2D::Solve(dt), PROJECTION_2D.h:235-257) is
//   k_coeff          face beta + sigma*kappa from phi        (A2, A4, A5, A8 beta)
//   multigrid setup  coarse beta, interpolation weights, bottom pseudo-inverse
//   k_rhs            b' = h^2 (-D u/dt + D(B J)), its mean and norms (A3, A8 rhs)
//   solve            FP64 iterative refinement around a preconditioned CG in FP64
//                    or FP32 (replaces the CSR matrix and CG/PCG: A9-A13, A15)
//   k_face_update    consistent pressure-gradient face update (A14)
//   k_diagnostics    FP64 max|D u|, max|u|, max|v|
// all on the context's stream.  The CG keeps alpha/beta, the convergence flag
// and its buffer parities on the device; the host enqueues iterations in chunks
// and reads the state once per chunk.  Refinement passes recompute the true
// residual in FP64 (double-double pressure) and hand it back to the CG, which
// continues with its search direction (residual replacement).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "p2d.h"
#include "p2d_common.hpp"
#include "p2d_field.hpp"
#include "p2d_mg.hpp"
#include "p2d_levelset.hpp"
#include "p2d_visc.hpp"

using namespace p2d;

namespace {

constexpr int BOTTOM_MAX = 64;      // coarsen until the grid has <= 64 cells (or a side is odd)
constexpr int BOTTOM_LIMIT = 1024;  // largest bottom grid accepted (dense pseudo-inverse)
static_assert(BOTTOM_LIMIT <= SMALL_THREADS, "k_mg_bottom_inverse keeps one diagonal entry per thread");
constexpr int GRIDSTRIDE_BLOCKS = 2048;
constexpr long SETUP_TILED_MIN_CELLS = 512L * 512;  // levels set up by k_mg_setup_tiled

inline unsigned blocks_for(size_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }
inline unsigned gs_blocks(size_t n) { return (unsigned)std::min<size_t>(blocks_for(n), GRIDSTRIDE_BLOCKS); }
// Band kernels: BLOCK columns x BAND rows per block.
inline dim3 bgrid(int ncols, int nrows) {
  return dim3(blocks_for((size_t)ncols), (unsigned)((nrows + BAND - 1) / BAND));
}
// Pre/post tiles: tile_tx(H) columns x ty rows of output per block.
inline dim3 tgrid(int nx, int ny, int H, int ty = TY) {
  const int TX = tile_tx(H);
  return dim3((unsigned)((nx + TX - 1) / TX), (unsigned)((ny + ty - 1) / ty));
}
// Direction kernel: DTX x DTY cells per block.
inline dim3 dirgrid(int nx, int ny) {
  return dim3((unsigned)((nx + DTX - 1) / DTX), (unsigned)((ny + DTY - 1) / DTY));
}
// Level-set tile kernels (re-initialisation, advection): RI_W x RI_H cells per block.
inline dim3 dgrid(int nx, int ny) {
  return dim3((unsigned)((nx + RI_W - 1) / RI_W), (unsigned)((ny + RI_H - 1) / RI_H));
}
// Refinement tile kernel: FTX x FTY cells per block.
inline dim3 fgrid(int nx, int ny) {
  return dim3((unsigned)((nx + FTX - 1) / FTX), (unsigned)((ny + FTY - 1) / FTY));
}
constexpr int MAX_PASSES = 32;
// Kernels p2d_time_kernel / p2d_kernel_timing name (finest level).
enum {
  KT_DIR_F32 = 0, KT_DIR_F64 = 1, KT_PRE0_F32 = 2, KT_POST0_F32 = 3, KT_RESID_F32 = 4,
  KT_RESID_F64 = 5, KT_PRE0_F64 = 6, KT_POST0_F64 = 7, KT_COUNT = 8
};
inline size_t count(dim3 g) { return (size_t)g.x * g.y * g.z; }
inline double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}
inline double ms_since(std::chrono::steady_clock::time_point a) {
  return ms_between(a, std::chrono::steady_clock::now());
}

template <class T>
struct MG {
  int L = 0, l0 = 0;
  bool valid = false;
  std::vector<Level<T>> lv;
  T* binv = nullptr;
  size_t small_lds = 0;
};

template <class T>
struct KVec {  // Krylov vectors on the finest level
  T *x = nullptr, *z = nullptr, *q = nullptr;
  T* r[2] = {nullptr, nullptr};
  T* p[2] = {nullptr, nullptr};
};

}  // namespace

struct p2d_ctx {
  p2d_config cfg;
  int nx = 0, ny = 0;
  size_t N = 0, NU = 0, NV = 0;
  double h = 0;
  hipStream_t own = nullptr, stream = nullptr;
  // The bottom pseudo-inverse runs on a side stream, forked from the solve
  // stream after the multigrid setup and joined before its first use (the
  // single-workgroup cycle), so that it overlaps the first V-cycle's descent.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool join_pending = false;
  std::vector<void*> allocs;
  size_t lds_limit = 65536;  // dynamic LDS available to the single-workgroup V-cycle
  // bench hook: HIP-event timing of every launch of selected finest-level kernels
  struct KTimer {
    std::vector<hipEvent_t> ev;  // start/end pairs recorded since the last harvest
    size_t used = 0;
    double total_ms = 0;
    long launches = 0;
  };
  unsigned timing_mask = 0;
  KTimer ktimer[KT_COUNT];

  // fluid state and coefficients (FP64 masters, cell-pitched)
  double *phi = nullptr, *jc = nullptr, *bw = nullptr, *bs = nullptr;
  double *saw = nullptr, *sas = nullptr;  // ghost-fluid jump s a on the west / south face of each cell
  bool have_fluid = false, explicit_coeff = false, has_jc = true, coeff_ready = false;
  // set by p2d_advance for its second and third projection of a step: the level set, and
  // so the coefficients and the multigrid hierarchy, are those of the first
  bool reuse_coeff = false;
  int num_cu = 256;  // compute units of the device (tile kernels with more tiles run two blocks per CU)
  double rho_m = 1, rho_p = 1, sigma = 0;

  // right-hand side, solutions (double-double with an FP32 low part, double-buffered), staging
  double* b = nullptr;
  double* ph[2] = {nullptr, nullptr};
  float* pl[2] = {nullptr, nullptr};
  int pcur = 0;
  double* muh[2] = {nullptr, nullptr};
  float* mul[2] = {nullptr, nullptr};
  double *su = nullptr, *sv = nullptr, *sg1 = nullptr, *sg2 = nullptr, *sp = nullptr;
  // time-step state (allocated by the first p2d_advance): phi, u, v and their stage-start copies
  double *st_phi = nullptr, *st_phin = nullptr, *st_u = nullptr, *st_v = nullptr, *st_un = nullptr, *st_vn = nullptr;
  double* sa[4] = {nullptr, nullptr, nullptr, nullptr};
  bool have_pressure = false;
  bool pressure_of_projection = false;  // the kept pressure is a p2d_project's (the adjoint needs one)
  double last_dt = 1.0;

  KVec<double> v64;
  KVec<float> v32;
  MG<double> mg64;  // its level coefficients are the FP64 masters of every level
  MG<float> mg32;
  bool mg_common = false;
  double* bottom_work = nullptr;
  double* binv64 = nullptr;
  float* binv32 = nullptr;

  double* partial = nullptr;
  Scal* s = nullptr;
  Scal* hs = nullptr;  // pinned host mirror
  // The host's replay of the Krylov bookkeeping of the inner iterations it has
  // enqueued (the device's rpar, iter, pending, seg0 as each enqueued kernel will
  // find them, while the solve runs): the finest-level kernels get their residual
  // and direction buffers from it, so their loads need not wait for the state.
  // Reloaded from hs at every synchronisation; once the device has set done, the
  // kernels return without touching the buffers and a stale replay is harmless.
  struct Track {
    int rpar = 0, iter = 0, pending = 0, seg0 = 0;
  } trk;
  int last_seg[2][MAX_PASSES] = {};  // iterations of each pass of the last solve (FP64, mixed)
  int last_passes[2] = {0, 0};       // refinement passes of the last solve
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
#define CHK(call)                  \
  do {                             \
    int rc_ = (call);              \
    if (rc_ != P2D_OK) return rc_; \
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

// Kernel timing: an event pair around each launch of a selected kernel on the
// context stream; durations are harvested at the host synchronisations.
void kt_begin(p2d_ctx* c, int which) {
  if (!((c->timing_mask >> which) & 1)) return;
  auto& t = c->ktimer[which];
  while (t.ev.size() < t.used + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) {
      c->timing_mask &= ~(1u << which);
      return;
    }
    t.ev.push_back(e);
  }
  (void)hipEventRecord(t.ev[t.used], c->stream);
}
void kt_end(p2d_ctx* c, int which) {
  if (!((c->timing_mask >> which) & 1)) return;
  auto& t = c->ktimer[which];
  (void)hipEventRecord(t.ev[t.used + 1], c->stream);
  t.used += 2;
}
void kt_harvest(p2d_ctx* c) {
  for (auto& t : c->ktimer) {
    for (size_t k = 0; k + 1 < t.used; k += 2) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, t.ev[k], t.ev[k + 1]) == hipSuccess) {
        t.total_ms += ms;
        ++t.launches;
      }
    }
    t.used = 0;
  }
}

int sync_scal(p2d_ctx* c) {
  HIPCHK(hipMemcpyAsync(c->hs, c->s, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  kt_harvest(c);
  c->trk.rpar = c->hs->rpar;
  c->trk.iter = c->hs->iter;
  c->trk.pending = c->hs->pending;
  c->trk.seg0 = c->hs->seg0;
  return P2D_OK;
}

// Orders the solve stream after the pending bottom pseudo-inverse, if any.
int join_side(p2d_ctx* c) {
  if (!c->join_pending) return P2D_OK;
  c->join_pending = false;
  HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  return P2D_OK;
}

// The single-workgroup cycle for nl levels (1 .. SMALL_LV).
template <class T>
const void* small_kernel(int nl) {
  static_assert(SMALL_LV == 4, "one instantiation per level count");
  switch (nl) {
    case 1: return reinterpret_cast<const void*>(&k_mg_small<T, 1>);
    case 2: return reinterpret_cast<const void*>(&k_mg_small<T, 2>);
    case 3: return reinterpret_cast<const void*>(&k_mg_small<T, 3>);
    default: return reinterpret_cast<const void*>(&k_mg_small<T, 4>);
  }
}

// Levels: l = 0 uses the context's coefficients (bw0/bs0) when given; every
// level owns xa, xb, r; levels above l0 own their interpolation weights.
template <class T>
int alloc_mg(p2d_ctx* c, MG<T>& mg, const T* bw0, const T* bs0) {
  const int L = num_levels(c->nx, c->ny, c->cfg.mg_max_levels);
  mg.L = L;
  mg.lv.resize(L);
  for (int l = 0; l < L; ++l) {
    mg.lv[l].nx = c->nx >> l;
    mg.lv[l].ny = c->ny >> l;
  }
  // l0: the first level of at most SMALL_MAX_CELLS cells whose single-workgroup
  // V-cycle (at most SMALL_LV levels down to the bottom) fits the LDS limit
  auto small_bytes = [&](int a) { return (size_t)small_layout<T>(mg.lv.data() + a, L - a).total * sizeof(T); };
  mg.l0 = -1;
  for (int l = 1; l < L; ++l)
    if ((long)(c->nx >> l) * (c->ny >> l) <= SMALL_MAX_CELLS && L - l <= SMALL_LV && small_bytes(l) <= c->lds_limit) {
      mg.l0 = l;
      break;
    }
  if (mg.l0 < 0) return P2D_BAD_ARG;
  mg.small_lds = small_bytes(mg.l0);
  if (mg.small_lds > 65536 &&
      hipFuncSetAttribute(small_kernel<T>(L - mg.l0), hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)mg.small_lds) != hipSuccess)
    return P2D_HIP_ERROR;
  for (int l = 0; l < L; ++l) {
    Level<T>& d = mg.lv[l];
    d.nx = c->nx >> l;
    d.ny = c->ny >> l;
    const size_t n = (size_t)d.nx * d.ny;
    if (l == 0 && bw0) {
      d.bw = bw0;
      d.bs = bs0;
    } else {
      T *a, *b;
      CHK(dalloc(c, &a, n));
      CHK(dalloc(c, &b, n));
      d.bw = a;
      d.bs = b;
    }
    T *wx = nullptr, *wy = nullptr, *xa, *xb, *r;
    unsigned* wh = nullptr;
    if (l + 1 < L) {
      if (sizeof(T) == 4) {
        CHK(dalloc(c, &wh, n));
      } else {
        CHK(dalloc(c, &wx, n));
        CHK(dalloc(c, &wy, n));
      }
    }
    CHK(dalloc(c, &xa, n));
    CHK(dalloc(c, &xb, n));
    CHK(dalloc(c, &r, n));
    d.wx = wx;
    d.wy = wy;
    d.wh = wh;
    d.xa = xa;
    d.xb = xb;
    d.r = r;
    const double om = l == 0 ? c->cfg.mg_omega : c->cfg.mg_omega_coarse;  // SOR over-relaxation of the level
    d.om = (T)om;
    d.omc = (T)(1.0 - om);
  }
  return P2D_OK;
}

// Multigrid setup for precision prec: per fine level one k_mg_setup sweep that
// writes what is missing (the coarse FP64 coefficients shared by both
// hierarchies, the FP32 coefficient copies, the interpolation weights), then
// the bottom pseudo-inverse.
template <class T>
int mg_setup(p2d_ctx* c, MG<T>& mg, bool cast32) {
  MG<double>& m = c->mg64;
  const bool common = c->mg_common;
  CHK(join_side(c));  // the setup may rewrite what a pending inverse reads
  for (int l = 0; l < m.L; ++l) {
    const Level<double>& f = m.lv[l];
    T* wx = const_cast<T*>(mg.lv[l].wx);  // null on the bottom level and on FP32 levels
    T* wy = const_cast<T*>(mg.lv[l].wy);
    unsigned* wh = const_cast<unsigned*>(mg.lv[l].wh);  // FP32 levels but the bottom
    float* f32bw = cast32 ? const_cast<float*>(c->mg32.lv[l].bw) : nullptr;
    float* f32bs = cast32 ? const_cast<float*>(c->mg32.lv[l].bs) : nullptr;
    double* cbw = !common && l + 1 < m.L ? const_cast<double*>(m.lv[l + 1].bw) : nullptr;
    double* cbs = !common && l + 1 < m.L ? const_cast<double*>(m.lv[l + 1].bs) : nullptr;
    if (!wx && !wh && !f32bw && !cbw) continue;
    // large levels: the tiled sweep (each face reciprocal formed once); small ones
    // are latency-bound, where the per-cell sweep's shorter chain wins
    if ((long)f.nx * f.ny >= SETUP_TILED_MIN_CELLS)
      k_mg_setup_tiled<T><<<dim3((unsigned)((f.nx + STX - 1) / STX), (unsigned)((f.ny + STY - 1) / STY)), BLOCK, 0,
                            c->stream>>>(f.nx, f.ny, f.bw, f.bs, wx, wy, wh, f32bw, f32bs, cbw, cbs);
    else
      k_mg_setup<T><<<dim3(blocks_for((size_t)f.nx), (unsigned)f.ny), BLOCK, 0, c->stream>>>(
          f.nx, f.ny, f.bw, f.bs, wx, wy, wh, f32bw, f32bs, cbw, cbs);
  }
  if (!common) {
    const Level<double>& B = m.lv[m.L - 1];
    HIPCHK(hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    k_mg_bottom_inverse<<<1, SMALL_THREADS, 0, c->side>>>(B.nx, B.ny, B.bw, B.bs, c->bottom_work, c->binv64,
                                                          c->binv32);
    CHK(launched_ok());
    HIPCHK(hipEventRecord(c->ev_join, c->side));
    c->join_pending = true;
    c->mg_common = true;
  }
  mg.valid = true;
  return launched_ok();
}

int ensure_mg(p2d_ctx* c, int prec) {
  if (!c->coeff_ready) return P2D_BAD_ARG;
  if (prec == P2D_PREC_F64 && !c->mg64.valid) CHK(mg_setup<double>(c, c->mg64, false));
  if (prec == P2D_PREC_F32 && !c->mg32.valid) CHK(mg_setup<float>(c, c->mg32, true));
  return P2D_OK;
}

void invalidate(p2d_ctx* c) {
  c->mg_common = false;
  c->mg64.valid = false;
  c->mg32.valid = false;
}

// One V-cycle on the finest level's residual rbuf[s->rpar] (after the pending
// Krylov update, if any), result in v.z, in two parts: the head (finest
// pre-smoothing fused with the update, and its finish, which applies the
// convergence test) and the tail (the rest; every tail kernel returns at once
// once the test has set s->done).
template <class T>
void vcycle_head(p2d_ctx* c, MG<T>& mg, KVec<T>& v, int guard = 0) {
  const int nu1 = c->cfg.mg_pre_sweeps;
  const Level<T>* lv = mg.lv.data();
  const dim3 gpre = tgrid(lv[0].nx, lv[0].ny, 2 * nu1 + 1);
  const int kpre = sizeof(T) == 4 ? KT_PRE0_F32 : KT_PRE0_F64;
  p2d_ctx::Track& t = c->trk;
  kt_begin(c, kpre);
  // the pending update is the pass's first when no update has been applied since its start
  const int xfirst = t.pending && t.iter == t.seg0 ? 1 : 0;
  k_mg_pre<T, true><<<gpre, TBLOCK, 0, c->stream>>>(lv[0], lv[1], v.r[t.rpar], lv[0].xa, nu1, v.r[1 - t.rpar], v.x,
                                                    v.p[(t.iter + 1) & 1], v.q, t.pending, xfirst, guard, c->partial,
                                                    c->s);
  kt_end(c, kpre);
  k_fin_pre0<<<1, FINISH_THREADS, 0, c->stream>>>(c->partial, (int)count(gpre), (double)c->N, c->s);
  // k_fin_pre0: the residual buffers swap; a pending update counts as an iteration
  t.rpar = 1 - t.rpar;
  if (t.pending) {
    t.iter += 1;
    t.pending = 0;
  }
}

template <class T>
void vcycle_tail(p2d_ctx* c, MG<T>& mg, KVec<T>& v) {
  const int nu1 = c->cfg.mg_pre_sweeps, nu2 = c->cfg.mg_post_sweeps;
  const int Hpre = 2 * nu1 + 1, Hpost = 2 * nu2;
  const Level<T>* lv = mg.lv.data();
  // The finest post-smoothing: FP32 in 16-wavefront tiles two blocks per CU; FP64, whose tiles
  // take the registers of one 16-wavefront block per CU, in 8-wavefront tiles, two per CU, so
  // that one block's loads overlap the other's sweeps (FP64 2048^2: 64 -> 57 us per launch).
  // Its halo is 2 nu, so its tiles take ty_post rows (36 and 12), one pre-smoothing halo
  // row more on each side than the pre kernel's (FP64 RT 2048^2 4.91 -> 4.82 ms).
  constexpr int NW0 = sizeof(T) == 4 ? 16 : 8;
  const dim3 gpost = tgrid(lv[0].nx, lv[0].ny, Hpost, Tile<NW0>::ty_post);
  const int kpost = sizeof(T) == 4 ? KT_POST0_F32 : KT_POST0_F64;
  // Tile shape of a coarse level: 8-wavefront tiles (10 rows) when at most two of
  // them per CU cover it; else, when 16-wavefront tiles (34 rows) would need more
  // than one per CU, 12-wavefront tiles (22 rows) if at most two per CU cover it
  // (fewer wavefronts per SIMD on the busiest CUs); else 16-wavefront tiles, FP32
  // ones two blocks per CU when they outnumber the CUs (BUSY variants).
  const size_t ncu = (size_t)c->num_cu;
  auto shape = [&](int l, int H, bool post) {
    const int t8 = post ? Tile<8>::ty_post : Tile<8>::ty, t12 = post ? Tile<12>::ty_post : Tile<12>::ty;
    const int t16 = post ? Tile<16>::ty_post : Tile<16>::ty;
    if (count(tgrid(lv[l].nx, lv[l].ny, H, t8)) <= 2 * ncu) return 8;
    if (count(tgrid(lv[l].nx, lv[l].ny, H, t16)) > ncu && count(tgrid(lv[l].nx, lv[l].ny, H, t12)) <= 2 * ncu)
      return 12;
    return 16;
  };
  auto busy = [&](int l, int H, bool post) {
    return sizeof(T) == 4 && count(tgrid(lv[l].nx, lv[l].ny, H, post ? Tile<16>::ty_post : Tile<16>::ty)) > ncu;
  };
  for (int l = 1; l < mg.l0; ++l) {
    const int sh = shape(l, Hpre, false);
    const Level<T>& F = lv[l];
    if (sh == 8)
      k_mg_pre<T, false, false, 8><<<tgrid(F.nx, F.ny, Hpre, Tile<8>::ty), Tile<8>::threads, 0, c->stream>>>(
          F, lv[l + 1], F.r, F.xa, nu1, nullptr, nullptr, nullptr, nullptr, 0, 0, 0, nullptr, c->s);
    else if (sh == 12)
      k_mg_pre<T, false, false, 12><<<tgrid(F.nx, F.ny, Hpre, Tile<12>::ty), Tile<12>::threads, 0, c->stream>>>(
          F, lv[l + 1], F.r, F.xa, nu1, nullptr, nullptr, nullptr, nullptr, 0, 0, 0, nullptr, c->s);
    else if (busy(l, Hpre, false))
      k_mg_pre<T, false, true><<<tgrid(F.nx, F.ny, Hpre), TBLOCK, 0, c->stream>>>(
          F, lv[l + 1], F.r, F.xa, nu1, nullptr, nullptr, nullptr, nullptr, 0, 0, 0, nullptr, c->s);
    else
      k_mg_pre<T, false><<<tgrid(F.nx, F.ny, Hpre), TBLOCK, 0, c->stream>>>(
          F, lv[l + 1], F.r, F.xa, nu1, nullptr, nullptr, nullptr, nullptr, 0, 0, 0, nullptr, c->s);
  }
  (void)join_side(c);  // enqueue-only; a failure shows at the next launch check
  SmallArgs<T> sa{};
  sa.nl = mg.L - mg.l0;
  for (int q = 0; q < sa.nl; ++q) sa.lv[q] = lv[mg.l0 + q];
  sa.binv = mg.binv;
  switch (sa.nl) {
    case 1: k_mg_small<T, 1><<<1, SMALL_THREADS, mg.small_lds, c->stream>>>(sa, nu1, nu2, c->s); break;
    case 2: k_mg_small<T, 2><<<1, SMALL_THREADS, mg.small_lds, c->stream>>>(sa, nu1, nu2, c->s); break;
    case 3: k_mg_small<T, 3><<<1, SMALL_THREADS, mg.small_lds, c->stream>>>(sa, nu1, nu2, c->s); break;
    default: k_mg_small<T, 4><<<1, SMALL_THREADS, mg.small_lds, c->stream>>>(sa, nu1, nu2, c->s); break;
  }
  for (int l = mg.l0 - 1; l >= 1; --l) {
    const int sh = shape(l, Hpost, true);
    const Level<T>& F = lv[l];
    if (sh == 8)
      k_mg_post<T, false, false, 8><<<tgrid(F.nx, F.ny, Hpost, Tile<8>::ty_post), Tile<8>::threads, 0, c->stream>>>(
          F, lv[l + 1], F.r, F.xa, F.xb, nu2, nullptr, c->s);
    else if (sh == 12)
      k_mg_post<T, false, false, 12><<<tgrid(F.nx, F.ny, Hpost, Tile<12>::ty_post), Tile<12>::threads, 0,
                                        c->stream>>>(
          F, lv[l + 1], F.r, F.xa, F.xb, nu2, nullptr, c->s);
    else if (busy(l, Hpost, true))
      k_mg_post<T, false, true><<<tgrid(F.nx, F.ny, Hpost, Tile<16>::ty_post), TBLOCK, 0, c->stream>>>(
          F, lv[l + 1], F.r, F.xa, F.xb, nu2, nullptr, c->s);
    else
      k_mg_post<T, false><<<tgrid(F.nx, F.ny, Hpost, Tile<16>::ty_post), TBLOCK, 0, c->stream>>>(
          F, lv[l + 1], F.r, F.xa, F.xb, nu2, nullptr, c->s);
  }
  kt_begin(c, kpost);
  k_mg_post<T, true, false, NW0><<<gpost, Tile<NW0>::threads, 0, c->stream>>>(lv[0], lv[1], v.r[c->trk.rpar],
                                                                              lv[0].xa, v.z, nu2, c->partial, c->s);
  kt_end(c, kpost);
  k_fin_post0<<<1, FINISH_THREADS, 0, c->stream>>>(c->partial, (int)count(gpost), (double)c->N, c->s);
}

template <class T>
void vcycle(p2d_ctx* c, MG<T>& mg, KVec<T>& v) {
  vcycle_head<T>(c, mg, v);
  vcycle_tail<T>(c, mg, v);
}

// The part of a CG iteration after the V-cycle's head: the rest of the V-cycle,
// the new direction and alpha, with the direction kernel and finish that can end
// a pass one head early (both precisions: FP64 RT 2048^2 4.95 -> 4.88 ms, 4096^2
// 17.28 -> 17.02 ms, the same iterations).
template <class T>
void cg_tail(p2d_ctx* c, MG<T>& mg, KVec<T>& v) {
  vcycle_tail<T>(c, mg, v);
  const Level<T>& l0 = mg.lv[0];
  const dim3 g = dirgrid(c->nx, c->ny);
  const int kd = sizeof(T) == 4 ? KT_DIR_F32 : KT_DIR_F64;
  p2d_ctx::Track& t = c->trk;
  const T* pold = v.p[t.iter & 1];
  T* pnew = v.p[1 - (t.iter & 1)];
  const int first = t.iter == t.seg0 ? 1 : 0;
  kt_begin(c, kd);
  k_dir<T><<<g, BLOCK, 0, c->stream>>>(c->nx, c->ny, l0.bw, l0.bs, v.z, pold, pnew, v.q, v.r[t.rpar], first,
                                       c->partial, c->s);
  kt_end(c, kd);
  k_fin_dir<<<1, FINISH_THREADS, 0, c->stream>>>(c->partial, (int)count(g), (double)c->N, c->s);
  t.pending = 1;  // k_fin_dir
}

// Residual launch plus its finish (sum r^2, max |r|, sum r into s->red[0..2]).
template <class T>
void correct_resid(p2d_ctx* c, const double* b, const double* saw, const double* sas, const double* ph,
                   const float* pl, const T* e, const T* pd, double* qh, float* ql, T* r) {
  const dim3 g = fgrid(c->nx, c->ny);
  const int kr = sizeof(T) == 4 ? KT_RESID_F32 : KT_RESID_F64;
  kt_begin(c, kr);
  k_correct_resid<T, T><<<g, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bw, c->bs, saw, sas, b, ph, pl, e, pd, qh, ql,
                                                    r, c->partial, c->s);
  kt_end(c, kr);
  k_fin_store<3><<<1, FINISH_THREADS, 0, c->stream>>>(c->partial, (int)count(g), 2u, 0, c->s);
}

struct SolveResult {
  int status = P2D_OK, iterations = 0, outer = 0;
  double r2 = 0, rinf = 0, b2 = 0;
};

// Right-hand side of one solve: b' (velocity part, mean in s->bmean) and the
// per-face jump terms (NULL: none).  The initial residual r0 (p = 0) is already
// in the solve precision's rbuf[0] and its norms in s->red (k_rhs, k_adj_rhs).
struct Rhs {
  const double* b;
  const double* saw;
  const double* sas;
};

// Solves A0' x = b' - jumps (h^2-scaled, mean removed) for x = sol[*cur], the
// pressure in double-double.  rtol: ||r||_2 <= rtol ||r_0||_2; atol:
// ||r||_inf <= atol (scaled); abs2: ||r||_2^2 <= abs2 (scaled); <= 0 disables.
//
// The host enqueues inner iterations in chunks sized from the previous solve's
// pass lengths; the kernels skip work once s->done is set.  When a pass (inner
// segment) has converged, the refinement (true residual in FP64, correction
// added to the double-double pressure) and the continue kernel are enqueued
// without waiting: the continue kernel applies the outer convergence test on
// the device, so a pass costs one host synchronisation.
template <class T>
int solve_t(p2d_ctx* c, MG<T>& mg, KVec<T>& v, int slot, const Rhs& rhs, double* solh[2], float* soll[2],
            int* cur, double rtol, double atol, double abs2, SolveResult* out) {
  const size_t N = c->N;
  const bool mixed = sizeof(T) == 4;
  *cur = 0;
  // Inner targets: FP64 runs to the outer targets (with a margin for the
  // recursive residual) or 1e-14 relative; FP32 runs passes of inner_rtol
  // relative reduction between FP64 residual replacements.
  const double rel = mixed ? c->cfg.inner_rtol : 1e-14;
  const bool verbose = std::getenv("P2D_VERBOSE") != nullptr;
  const auto t_solve = std::chrono::steady_clock::now();
  k_inner_init<<<1, 1, 0, c->stream>>>(c->s, (double)N, rel * rel, mixed ? 0 : 1, rtol, atol, abs2,
                                       c->cfg.max_iterations);
  c->trk = p2d_ctx::Track{};  // k_inner_init: rpar = iter = pending = seg0 = 0
  CHK(launched_ok());
  int status = P2D_OK, outer = 0, stalled = 0, iters = 0, seg_start = 0;
  double r2 = 0, rinf = 0, b2 = 0, best_r2 = 1e300, best_rinf = 1e300;
  int* last = c->last_seg[slot];
  const int last_passes = c->last_passes[slot];
  auto predict = [&](int pass) { return std::max(2, (pass < MAX_PASSES && last[pass] > 0 ? last[pass] : 6) + 1); };
  int chunk = predict(0);
  bool report = false;  // a refinement's norms wait in red (the iterations do not touch red)
  bool first = true;
  // A chunk ends after the head of its last iteration (the convergence test):
  // a pass that ends there skips the tail's launches, and one that goes on
  // enqueues the tail first.
  bool tail_pending = false;
  while (true) {
    const auto t_launch0 = std::chrono::steady_clock::now();
    for (int k = 0; k < chunk; ++k) {
      if (tail_pending) cg_tail<T>(c, mg, v);
      vcycle_head<T>(c, mg, v, k + 1 == chunk ? 1 : 0);
      tail_pending = true;
    }
    CHK(launched_ok());
    const auto t_launch1 = std::chrono::steady_clock::now();
    CHK(sync_scal(c));
    iters = c->hs->iter;
    if (first) {
      first = false;
      b2 = c->hs->b2;
      best_r2 = r2 = c->hs->r0_2;
      best_rinf = rinf = c->hs->red[1];
      if (verbose)
        fprintf(stderr, "p2d solve: %s b2=%.3e rtol=%.1e atol=%.3e r0 rinf=%.3e\n", mixed ? "mixed" : "fp64", b2, rtol,
                atol, rinf);
    }
    if (verbose)
      fprintf(stderr, "    sync at %.3f ms (launch %.3f ms, wait %.3f ms): iter %d done %d recursive r2=%.3e rinf=%.3e\n",
              ms_since(t_solve), ms_between(t_launch0, t_launch1), ms_since(t_launch1), c->hs->iter, c->hs->done,
              c->hs->r2, c->hs->rinf);
    if (report) {
      report = false;
      r2 = c->hs->red[0];
      rinf = c->hs->red[1];
      if (verbose)
        fprintf(stderr, "  pass %d: %d its -> true r2=%.3e (rel %.3e) rinf=%.3e\n", outer, seg_start, r2,
                std::sqrt(r2 / std::max(b2, 1e-300)), rinf);
      if (!(r2 == r2)) {
        status = P2D_NAN;
        break;
      }
      // a pass that improves neither norm by 2x has hit the rounding floor of the
      // FP64 residual (~eps max |beta grad p| per cell): stop after the next refinement
      if (r2 > 0.25 * best_r2 && rinf > 0.5 * best_rinf) {
        ++stalled;
      } else {
        stalled = 0;
      }
      best_r2 = std::min(best_r2, r2);
      best_rinf = std::min(best_rinf, rinf);
    }
    if (c->hs->flag) {
      status = P2D_NAN;
      break;
    }
    if (c->hs->outer_done) break;
    if (!c->hs->done) {
      chunk = chunk == 0 ? predict(outer) : 2;  // after a look at the outer test: the predicted pass length
      continue;
    }
    // end of a pass: refinement p += x (double-double), the true residual into the
    // buffer the next pass reads; the last head's tail is not needed
    tail_pending = false;
    if (outer < MAX_PASSES) last[outer] = iters - seg_start;
    T* rdst = c->hs->rpar ? v.r[1] : v.r[0];
    const double* oh = outer > 0 ? solh[*cur] : nullptr;  // the first pass starts from p = 0
    const float* ol = outer > 0 ? soll[*cur] : nullptr;
    // a pass that ended at its direction's finish leaves that update pending: the
    // refinement applies it (direction p[iter & 1], alpha in the state); x holds the
    // pass's correction only if a head applied an update (the first writes x = a p)
    const T* pd = c->hs->pending ? v.p[c->hs->iter & 1] : nullptr;
    const bool x_written = iters - seg_start - (c->hs->pending ? 1 : 0) > 0;
    correct_resid<T>(c, rhs.b, rhs.saw, rhs.sas, oh, ol, x_written ? v.x : nullptr, pd, solh[1 - *cur],
                     soll[1 - *cur], rdst);
    *cur = 1 - *cur;
    ++outer;
    if (iters >= c->cfg.max_iterations || outer >= c->cfg.max_outer || stalled >= 1) {
      CHK(launched_ok());
      CHK(sync_scal(c));
      r2 = c->hs->red[0];
      rinf = c->hs->red[1];
      if (verbose)
        fprintf(stderr, "  pass %d: %d its -> true r2=%.3e (rel %.3e) rinf=%.3e (last)\n", outer, iters, r2,
                std::sqrt(r2 / std::max(b2, 1e-300)), rinf);
      if (!(r2 == r2)) status = P2D_NAN;
      break;
    }
    k_inner_continue<<<1, 1, 0, c->stream>>>(c->s, (double)N);
    c->trk.seg0 = c->trk.iter;  // k_inner_continue: a new direction sequence, nothing pending
    c->trk.pending = 0;
    CHK(launched_ok());
    report = true;
    seg_start = iters;
    // when the previous solve ended after this many passes, look at the outer test
    // before enqueueing more iterations
    chunk = outer >= last_passes ? 0 : predict(outer);
  }
  CHK(join_side(c));  // a solve that ended before any single-workgroup cycle
  c->last_passes[slot] = outer;
  if (outer == 0) {  // no refinement ran: p = 0
    k_fill<double><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, solh[*cur], 0.0);
    k_fill<float><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, soll[*cur], 0.0f);
    CHK(launched_ok());
    r2 = c->hs->red[0];
    rinf = c->hs->red[1];
  }
  auto converged = [&](double a2, double ai) {
    return (rtol <= 0 || a2 <= rtol * rtol * b2) && (atol <= 0 || ai <= atol) && (abs2 <= 0 || a2 <= abs2);
  };
  if (status == P2D_OK && !converged(r2, rinf)) status = P2D_NOT_CONVERGED;
  out->status = status;
  out->iterations = iters;
  out->outer = outer;
  out->r2 = r2;
  out->rinf = rinf;
  out->b2 = b2;
  return P2D_OK;
}

int solve(p2d_ctx* c, const Rhs& rhs, double* solh[2], float* soll[2], int* cur, double rtol, double atol,
          double abs2, SolveResult* out) {
  if (c->cfg.precision == P2D_MIXED) {
    CHK(ensure_mg(c, P2D_PREC_F32));
    return solve_t<float>(c, c->mg32, c->v32, 1, rhs, solh, soll, cur, rtol, atol, abs2, out);
  }
  CHK(ensure_mg(c, P2D_PREC_F64));
  return solve_t<double>(c, c->mg64, c->v64, 0, rhs, solh, soll, cur, rtol, atol, abs2, out);
}

int stage_in(p2d_ctx* c, double* dst, const double* src, size_t n, int mem) {
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double),
                        mem == P2D_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c->stream));
  return P2D_OK;
}
int stage_out(p2d_ctx* c, double* dst, const double* src, size_t n, int mem) {
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double),
                        mem == P2D_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, c->stream));
  return P2D_OK;
}

__global__ void k_scal_arm(Scal* s, int iter, double beta, double alpha, int pending) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    s->done = 0;
    s->iter = iter;
    s->seg0 = 0;
    s->beta = beta;
    s->alpha = alpha;
    s->pending = pending;
    s->rpar = 0;
    s->rz = 1.0;
    s->rz_old = 1.0;
    s->zmean = 0.0;
    s->rshift = 0.0;
    s->r2raw = 0.0;
    s->rsum = 0.0;
    s->near = 0;
    s->flag = 0;
    s->outer_done = 0;
    s->rel2 = 0.0;
    s->abs2 = -1.0;
    s->ainf = -1.0;
    s->r0_2 = 1.0;
    s->max_iter = 1 << 30;
  }
}
template <class T>
__global__ void k_pattern(size_t n, T* a) {
  const size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (k < n) a[k] = (T)(((k * 2654435761u) % 1000u) * 1e-3 - 0.5);
}

// Coefficients of the current fluid: face beta and sigma kappa (k_coeff) unless
// they are explicit or reused, then the per-face jump terms s a (k_jump).
int assemble(p2d_ctx* c) {
  if (!c->explicit_coeff && !(c->reuse_coeff && c->coeff_ready)) {
    HIPCHK(hipMemsetAsync(&c->s->err, 0, sizeof(int), c->stream));
    k_coeff<<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, c->phi, 1.0 / c->rho_m,
                                                          1.0 / c->rho_p, c->sigma, c->bw, c->bs, c->jc, c->s);
    CHK(launched_ok());
    c->coeff_ready = true;
    invalidate(c);
  }
  if (c->has_jc)
    k_jump<<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->phi, c->jc, c->saw, c->sas);
  return launched_ok();
}

// b' and the first residual in the solve precision, from the velocity (du, dv, dt)
// or from a given divergence field (div); norms into s->red, mean into s->bmean.
int build_rhs(p2d_ctx* c, const double* du, const double* dv, const double* div, double dt) {
  const double* saw = c->has_jc ? c->saw : nullptr;
  const double* sas = c->has_jc ? c->sas : nullptr;
  if (c->cfg.precision == P2D_MIXED)
    k_rhs<float><<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, du, dv, div, c->bw, c->bs,
                                                               saw, sas, c->b, c->v32.r[0], c->partial);
  else
    k_rhs<double><<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, du, dv, div, c->bw, c->bs,
                                                                saw, sas, c->b, c->v64.r[0], c->partial);
  k_fin_rhs<<<1, FINISH_THREADS, 0, c->stream>>>(c->partial, (int)count(bgrid(c->nx, c->ny)), (double)c->N, c->s);
  return launched_ok();
}

// The configuration's absolute rule ||r||_2 <= abs_tol on the unscaled system, as a
// target on the sum of squares of the h^2-scaled residual (0: off).
double abs2_target(const p2d_ctx* c) {
  if (!(c->cfg.abs_tol > 0)) return 0.0;
  const double t = c->cfg.abs_tol * c->h * c->h;
  return t * t;
}

}  // namespace

// --------------------------------------------------------------------------- static scratch
// p2d_cfl_dt_noalloc's device memory: the block partials of k_cfl_max (3 per block, at most
// P2D_NOALLOC_MAX_BLOCKS blocks) and, after them, the Scal its kernels reduce into.  One copy
// per device, for the life of the process (1.7 MB).  Host code never passes `scratch` itself
// to a kernel: it is a host-side stand-in; each device's address comes from
// hipGetSymbolAddress (noalloc_scratch) and is kept in g_scratch.
constexpr size_t NOALLOC_SCAL_DOUBLES = (sizeof(Scal) + 255) / 256 * 32;
static_assert(sizeof(Scal) <= NOALLOC_SCAL_DOUBLES * sizeof(double), "Scal fits its slot");
constexpr size_t NOALLOC_DOUBLES = 3 * (size_t)P2D_NOALLOC_MAX_BLOCKS + NOALLOC_SCAL_DOUBLES;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double* g_scratch[NOALLOC_MAX_DEVICES] = {};

// The current device's copy of `scratch` (its address looked up once per device).
int noalloc_scratch(double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return P2D_NO_DEVICE;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return P2D_HIP_ERROR;
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return P2D_OK;
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
  cfg->mg_pre_sweeps = 3;
  cfg->mg_post_sweeps = 3;
  cfg->rtol = 1e-10;
  cfg->div_tol = 1e-10;
  cfg->max_iterations = 1000;
  cfg->max_outer = 12;
  cfg->inner_rtol = 1e-4;
  cfg->abs_tol = 0.0;
  cfg->mg_omega = 1.4;
  cfg->mg_omega_coarse = 1.5;
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
  if (cfg->nx < 2 || cfg->ny < 2 || cfg->nx > 65535 || cfg->ny > 65535 || !(cfg->h > 0) || cfg->mg_pre_sweeps < 1 ||
      cfg->mg_pre_sweeps > MAX_NU || cfg->mg_post_sweeps < 1 || cfg->mg_post_sweeps > MAX_NU ||
      cfg->max_iterations < 1 || cfg->max_outer < 1 || (cfg->precision != P2D_FP64 && cfg->precision != P2D_MIXED) ||
      (!(cfg->rtol > 0) && !(cfg->div_tol > 0) && !(cfg->abs_tol > 0)) || !(cfg->abs_tol >= 0) ||
      !(cfg->inner_rtol > 0 && cfg->inner_rtol < 1) || !(cfg->mg_omega > 0 && cfg->mg_omega < 2) ||
      !(cfg->mg_omega_coarse > 0 && cfg->mg_omega_coarse < 2))
    return P2D_BAD_ARG;
  // The tile kernels address a level's arrays with 32-bit byte offsets (ld_at / st_at):
  // the largest array, nx*ny doubles, must stay below 2^32 bytes.
  if ((unsigned long long)cfg->nx * (unsigned long long)cfg->ny * sizeof(double) > (unsigned long long)P2D_MAX_ARRAY_BYTES)
    return P2D_BAD_ARG;
  {
    // the V-cycle needs at least one coarsening and a bottom that fits the dense solve
    const int L = num_levels(cfg->nx, cfg->ny, cfg->mg_max_levels);
    const long nb = (long)(cfg->nx >> (L - 1)) * (cfg->ny >> (L - 1));
    if (L < 2 || nb > BOTTOM_LIMIT) return P2D_BAD_ARG;
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
  if (hipStreamCreate(&c->own) != hipSuccess || hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)
    return fail(P2D_HIP_ERROR);
  {
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
      c->lds_limit = (size_t)lds - 1024;  // less the kernel's static LDS
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      c->num_cu = cus;
  }
  c->stream = c->own;
  const size_t N = c->N;
  int rc = P2D_OK;
#define A(p, n) \
  if ((rc = dalloc(c, &(p), (n))) != P2D_OK) return fail(rc)
  A(c->phi, N);
  A(c->jc, N);
  A(c->saw, N);
  A(c->sas, N);
  A(c->bw, N);
  A(c->bs, N);
  A(c->b, N);
  for (int q = 0; q < 2; ++q) {
    A(c->ph[q], N);
    A(c->pl[q], N);
    A(c->muh[q], N);
    A(c->mul[q], N);
    A(c->v64.r[q], N);
    A(c->v64.p[q], N);
    A(c->v32.r[q], N);
    A(c->v32.p[q], N);
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
  A(c->v64.z, N);
  A(c->v64.q, N);
  A(c->v32.x, N);
  A(c->v32.z, N);
  A(c->v32.q, N);
  {
    // partial sums: the largest block count of any reducing launch
    size_t mb = std::max<size_t>(count(bgrid(c->nx + 1, c->ny + 1)), GRIDSTRIDE_BLOCKS);
    for (int H = 2; H <= 2 * MAX_NU + 1; ++H) mb = std::max(mb, count(tgrid(c->nx, c->ny, H)));
    mb = std::max(mb, count(dirgrid(c->nx, c->ny)));
    mb = std::max(mb, count(fgrid(c->nx, c->ny)));
    A(c->partial, (size_t)RED_K * mb);
  }
  A(c->s, 1);
#undef A
  if ((rc = alloc_mg<double>(c, c->mg64, c->bw, c->bs)) != P2D_OK) return fail(rc);
  if ((rc = alloc_mg<float>(c, c->mg32, nullptr, nullptr)) != P2D_OK) return fail(rc);
  {
    const Level<double>& B = c->mg64.lv[c->mg64.L - 1];
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
  if (c->side) (void)hipStreamSynchronize(c->side);
  for (void* p : c->allocs) (void)hipFree(p);
  for (auto& t : c->ktimer)
    for (hipEvent_t e : t.ev) (void)hipEventDestroy(e);
  if (c->hs) (void)hipHostFree(c->hs);
  if (c->own) (void)hipStreamDestroy(c->own);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
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
  c->has_jc = sigma != 0.0;
  c->coeff_ready = false;
  invalidate(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_set_coefficients(p2d_ctx* c, const double* beta_x, const double* beta_y, const double* phi,
                         const double* jc, int mem) {
  if (!c || !beta_x || !beta_y || !phi) return P2D_BAD_ARG;
  CHK(stage_in(c, c->sa[0], beta_x, c->NU, mem));
  CHK(stage_in(c, c->sa[1], beta_y, c->NV, mem));
  CHK(stage_in(c, c->phi, phi, c->N, mem));
  if (jc) CHK(stage_in(c, c->jc, jc, c->N, mem));
  // face layout -> cell-pitched layout; wall faces carry no flux
  k_faces_to_cells<<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->sa[0], c->sa[1], c->bw, c->bs);
  CHK(launched_ok());
  c->has_jc = jc != nullptr;
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
  CHK(assemble(c));
  const double* saw = c->has_jc ? c->saw : nullptr;
  const double* sas = c->has_jc ? c->sas : nullptr;
  CHK(build_rhs(c, du, dv, nullptr, dt));
  const double atol = c->cfg.div_tol > 0 ? 0.5 * c->cfg.div_tol * c->h * c->h / dt : 0.0;
  SolveResult sr;
  CHK(solve(c, Rhs{c->b, saw, sas}, c->ph, c->pl, &c->pcur, c->cfg.rtol, atol, abs2_target(c), &sr));
  c->pressure_of_projection = true;
  const bool bad_geom = c->hs->err != 0;
  const double b2 = sr.b2;
  k_face_update<<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, c->ph[c->pcur],
                                                              c->pl[c->pcur], saw, sas, c->bw, c->bs, du, dv);
  k_diagnostics<<<bgrid(c->nx + 1, c->ny + 1), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, du, dv, c->partial, c->s);
  k_fin_store<3><<<1, FINISH_THREADS, 0, c->stream>>>(c->partial, (int)count(bgrid(c->nx + 1, c->ny + 1)), 7u, 0,
                                                      c->s);
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
  const float* sl = c->pl[c->pcur];
  k_psum<<<gs_blocks(N), BLOCK, 0, c->stream>>>(N, sh, sl, c->partial, c->s);
  k_fin_store<1><<<1, FINISH_THREADS, 0, c->stream>>>(c->partial, (int)gs_blocks(N), 0u, 3, c->s);
  double* dst = mem == P2D_MEM_HOST ? c->sp : p;
  k_shift<<<blocks_for(N), BLOCK, 0, c->stream>>>(N, sh, sl, dst, mean_zero, c->s);
  CHK(launched_ok());
  if (mem == P2D_MEM_HOST) CHK(stage_out(c, p, c->sp, N, mem));
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_divergence(p2d_ctx* c, const double* u, const double* v, double dt, double* div, int mem) {
  if (!c || !u || !v || !div || !(dt > 0)) return P2D_BAD_ARG;
  const double* du = u;
  const double* dv = v;
  double* out = div;
  if (mem == P2D_MEM_HOST) {
    CHK(stage_in(c, c->su, u, c->NU, mem));
    CHK(stage_in(c, c->sv, v, c->NV, mem));
    du = c->su;
    dv = c->sv;
    out = c->sp;
  }
  k_divergence<<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, du, dv, out);
  CHK(launched_ok());
  if (mem == P2D_MEM_HOST) CHK(stage_out(c, div, c->sp, c->N, mem));
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_solve_pressure(p2d_ctx* c, const double* div, double* p, p2d_stats* st, int mem) {
  if (!c || !div || !c->have_fluid) return P2D_BAD_ARG;
  const double* dd = div;
  if (mem == P2D_MEM_HOST) {
    CHK(stage_in(c, c->sg1, div, c->N, mem));
    dd = c->sg1;
  }
  c->hs->err = 0;
  CHK(assemble(c));
  const double* saw = c->has_jc ? c->saw : nullptr;
  const double* sas = c->has_jc ? c->sas : nullptr;
  CHK(build_rhs(c, nullptr, nullptr, dd, 1.0));
  const double atol = c->cfg.div_tol > 0 ? 0.5 * c->cfg.div_tol * c->h * c->h : 0.0;
  SolveResult sr;
  CHK(solve(c, Rhs{c->b, saw, sas}, c->ph, c->pl, &c->pcur, c->cfg.rtol, atol, abs2_target(c), &sr));
  const bool bad_geom = c->hs->err != 0;
  c->have_pressure = true;
  c->pressure_of_projection = false;
  if (p) CHK(p2d_get_pressure(c, p, 0, mem));
  if (st) {
    st->status = bad_geom ? P2D_NAN : sr.status;
    st->iterations = sr.iterations;
    st->outer = sr.outer;
    st->rel_residual = sr.b2 > 0 ? std::sqrt(sr.r2 / sr.b2) : 0.0;
    st->max_abs_residual = sr.rinf / (c->h * c->h);
    st->max_div = 0;
    st->max_u = st->max_v = 0;
  }
  if (bad_geom) return P2D_NAN;
  return sr.status;
}

int p2d_get_coefficients(p2d_ctx* c, double* bx, double* by, double* jc, int mem) {
  if (!c || !c->coeff_ready) return P2D_BAD_ARG;
  k_cells_to_faces<<<bgrid(c->nx + 1, c->ny + 1), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bw, c->bs, c->sa[2],
                                                                         c->sa[3]);
  CHK(launched_ok());
  if (bx) CHK(stage_out(c, bx, c->sa[2], c->NU, mem));
  if (by) CHK(stage_out(c, by, c->sa[3], c->NV, mem));
  if (jc && c->has_jc) CHK(stage_out(c, jc, c->jc, c->N, mem));
  if (jc && !c->has_jc && mem != P2D_MEM_HOST)  // no surface tension: jc = 0 (k_coeff leaves it unwritten)
    k_fill<double><<<blocks_for(c->N), BLOCK, 0, c->stream>>>(c->N, jc, 0.0);
  CHK(launched_ok());
  HIPCHK(hipStreamSynchronize(c->stream));
  if (jc && !c->has_jc && mem == P2D_MEM_HOST) std::fill(jc, jc + c->N, 0.0);
  return P2D_OK;
}

int p2d_adjoint(p2d_ctx* c, const double* gu, const double* gv, double* du, double* dv, double* dbx, double* dby,
                p2d_stats* st, int mem) {
  if (!c || !gu || !gv || !c->have_pressure || !c->pressure_of_projection) return P2D_BAD_ARG;
  const double* g1 = gu;
  const double* g2 = gv;
  if (mem == P2D_MEM_HOST) {
    CHK(stage_in(c, c->sg1, gu, c->NU, mem));
    CHK(stage_in(c, c->sg2, gv, c->NV, mem));
    g1 = c->sg1;
    g2 = c->sg2;
  }
  if (c->cfg.precision == P2D_MIXED)
    k_adj_rhs<float><<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, g1, g2, c->bw, c->bs, c->b,
                                                                   c->v32.r[0], c->partial);
  else
    k_adj_rhs<double><<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, g1, g2, c->bw, c->bs, c->b,
                                                                    c->v64.r[0], c->partial);
  k_fin_rhs<<<1, FINISH_THREADS, 0, c->stream>>>(c->partial, (int)count(bgrid(c->nx, c->ny)), (double)c->N, c->s);
  CHK(launched_ok());
  SolveResult sr;
  int mcur = 0;
  const double rt = c->cfg.rtol > 0 ? c->cfg.rtol : 1e-10;
  CHK(solve(c, Rhs{c->b, nullptr, nullptr}, c->muh, c->mul, &mcur, rt, 0.0, 0.0, &sr));
  const double b2 = sr.b2;
  const double* saw = c->has_jc ? c->saw : nullptr;
  const double* sas = c->has_jc ? c->sas : nullptr;
  double *o1 = du, *o2 = dv, *o3 = dbx, *o4 = dby;
  if (mem == P2D_MEM_HOST) {
    o1 = du ? c->sa[0] : nullptr;
    o2 = dv ? c->sa[1] : nullptr;
    o3 = dbx ? c->sa[2] : nullptr;
    o4 = dby ? c->sa[3] : nullptr;
  }
  k_adj_faces<<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, c->last_dt, g1, g2, c->muh[mcur],
                                                            c->mul[mcur], c->ph[c->pcur], c->pl[c->pcur], saw, sas,
                                                            o1, o2, o3, o4);
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
  const bool host = mem == P2D_MEM_HOST;
  if (prec == P2D_PREC_F64) {
    const double* xi = (const double*)x;
    double* yo = (double*)y;
    if (host) {
      HIPCHK(hipMemcpyAsync(c->v64.z, x, N * 8, hipMemcpyHostToDevice, c->stream));
      xi = c->v64.z;
      yo = c->v64.q;
    }
    k_apply<double><<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->bw, c->bs, xi, yo, scale);
    CHK(launched_ok());
    if (host) HIPCHK(hipMemcpyAsync(y, c->v64.q, N * 8, hipMemcpyDeviceToHost, c->stream));
  } else {
    CHK(ensure_mg(c, P2D_PREC_F32));
    const Level<float>& l0 = c->mg32.lv[0];
    const float* xi = (const float*)x;
    float* yo = (float*)y;
    if (host) {
      HIPCHK(hipMemcpyAsync(c->v32.z, x, N * 4, hipMemcpyHostToDevice, c->stream));
      xi = c->v32.z;
      yo = c->v32.q;
    }
    k_apply<float><<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, l0.bw, l0.bs, xi, yo, scale);
    CHK(launched_ok());
    if (host) HIPCHK(hipMemcpyAsync(y, c->v32.q, N * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_vcycle(p2d_ctx* c, const void* r, void* z, int prec, int mem) {
  if (!c || !r || !z || !c->coeff_ready || (prec != P2D_PREC_F64 && prec != P2D_PREC_F32)) return P2D_BAD_ARG;
  const size_t N = c->N;
  const bool host = mem == P2D_MEM_HOST;
  CHK(ensure_mg(c, prec));
  // armed with no pending update: the fused finest-level kernel only copies r
  k_scal_arm<<<1, 1, 0, c->stream>>>(c->s, 0, 0.0, 0.0, 0);
  c->trk = p2d_ctx::Track{};
  if (prec == P2D_PREC_F64) {
    HIPCHK(hipMemcpyAsync(c->v64.r[0], r, N * 8, host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c->stream));
    vcycle<double>(c, c->mg64, c->v64);
    CHK(launched_ok());
    HIPCHK(hipMemcpyAsync(z, c->v64.z, N * 8, host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, c->stream));
  } else {
    HIPCHK(hipMemcpyAsync(c->v32.r[0], r, N * 4, host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c->stream));
    vcycle<float>(c, c->mg32, c->v32);
    CHK(launched_ok());
    HIPCHK(hipMemcpyAsync(z, c->v32.z, N * 4, host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_mg_levels(p2d_ctx* c) { return c ? c->mg64.L : -1; }

// Level-set re-initialisation and normals on the context's grid.  They use the
// FP64 Krylov vectors x, z, q of the context as scratch (free between solves:
// the pressure and the coefficients are kept elsewhere).
// --- device-pointer building blocks of the level-set, advection and viscosity entry points
// and of p2d_advance (enqueued on the context stream, no synchronisation)
namespace {
void ls_advect_stage(p2d_ctx* c, const double* phi, const double* u, const double* v, double dt, double eps,
                     double* out) {
  k_advect_ls<<<dgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, eps, phi, u, v, out);
}
// `iterations` Sussman steps from phi, ping-ponging between phi and tmp (the sign array in
// v64.q); returns the buffer that holds the result
double* reinit_run(p2d_ctx* c, double* phi, double* tmp, int iterations, double scaling, double eps) {
  k_reinit_sign<<<blocks_for(c->N), BLOCK, 0, c->stream>>>(c->N, c->h, phi, c->v64.q);
  const double dt = c->h / scaling;
  for (int it = 0; it < iterations; ++it) {
    k_reinit_step<<<dgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, eps, phi, c->v64.q, tmp);
    std::swap(phi, tmp);
  }
  return phi;
}
// one ENO3 velocity advection stage (u, v) -> (uo, vo)
void vel_advect_stage(p2d_ctx* c, const double* u, const double* v, double dt, double* uo, double* vo) {
  const dim3 gu((unsigned)((c->nx + 1 + RI_W - 1) / RI_W), (unsigned)((c->ny + RI_H - 1) / RI_H));
  const dim3 gv((unsigned)((c->nx + RI_W - 1) / RI_W), (unsigned)((c->ny + 1 + RI_H - 1) / RI_H));
  k_advect_vel<true><<<gu, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, u, v, uo);
  k_advect_vel<false><<<gv, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, u, v, vo);
}
// one explicit viscosity step (u, v) -> (uo, vo) with the context's fluid
void viscosity_step(p2d_ctx* c, const double* u, const double* v, double dt, double mm, double mp, double eps,
                    double* uo, double* vo) {
  const dim3 gu(blocks_for((size_t)c->nx + 1), (unsigned)((c->ny + BAND - 1) / BAND));
  const dim3 gv(blocks_for((size_t)c->nx), (unsigned)((c->ny + 1 + BAND - 1) / BAND));
  k_viscosity<true><<<gu, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, c->phi, c->rho_m, c->rho_p, mm, mp, eps, u, v,
                                                 uo);
  k_viscosity<false><<<gv, BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dt, c->phi, c->rho_m, c->rho_p, mm, mp, eps, u,
                                                  v, vo);
}
void lincomb(p2d_ctx* c, size_t n, double a, const double* x, double b, const double* y, double* out) {
  k_lincomb<<<blocks_for(n), BLOCK, 0, c->stream>>>(n, a, x, b, y, out);
}
}  // namespace

int p2d_reinitialize(p2d_ctx* c, double* phi, int iterations, double scaling, double epsilon, int mem) {
  if (!c || !phi || iterations < 0 || !(scaling > 0) || !(epsilon >= 0) || c->nx < 3 || c->ny < 3)
    return P2D_BAD_ARG;
  CHK(stage_in(c, c->v64.x, phi, c->N, mem));
  const double* r = reinit_run(c, c->v64.x, c->v64.z, iterations, scaling, epsilon);
  CHK(launched_ok());
  CHK(stage_out(c, phi, r, c->N, mem));
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

// EULERIAN_FLUID_SOLVER_2D::CFLOneTimeStep for the air-water solver
// (EULERIAN_FLUID_SOLVER_2D.h:3517-3637): advection c = max|u|/h + max|v|/h, gravity
// g = sqrt(9.8/h), viscosity w = max(mu-/rho-, mu+/rho+)(2/h^2 + 2/h^2), surface tension
// s = sqrt(sigma max|kappa| / (min(rho) h^2)); dt = cfl / (0.5 (c + w + sqrt((c + w)^2 + 4 g^2
// + 4 s^2))), 0 when the denominator is 0.  The curvature is that of phi (the level set the
// last projection used); phi may be NULL when sigma is 0.
int p2d_cfl_dt(p2d_ctx* c, const double* phi, const double* u, const double* v, const p2d_step_params* prm,
               double cfl_number, double* dt, int mem) {
  if (!c || !u || !v || !prm || !dt || (!phi && prm->sigma != 0.0) || !(prm->rho_minus > 0) ||
      !(prm->rho_plus > 0) || c->nx < 2 || c->ny < 2)
    return P2D_BAD_ARG;
  const double* du = u;
  const double* dv = v;
  const double* dp = prm->sigma != 0.0 ? phi : nullptr;
  if (mem == P2D_MEM_HOST) {
    CHK(stage_in(c, c->su, u, c->NU, mem));
    CHK(stage_in(c, c->sv, v, c->NV, mem));
    du = c->su;
    dv = c->sv;
    if (dp) {
      CHK(stage_in(c, c->v64.x, phi, c->N, mem));
      dp = c->v64.x;
    }
  }
  c->hs->err = 0;
  HIPCHK(hipMemsetAsync(&c->s->err, 0, sizeof(int), c->stream));
  k_cfl_max<<<bgrid(c->nx + 1, c->ny + 1), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, dp, du, dv, c->partial, c->s);
  k_fin_store<3><<<1, FINISH_THREADS, 0, c->stream>>>(c->partial, (int)count(bgrid(c->nx + 1, c->ny + 1)), 7u, 0,
                                                      c->s);
  CHK(launched_ok());
  CHK(sync_scal(c));
  if (c->hs->err) return P2D_NAN;
  *dt = cfl_step(c->hs->red[0], c->hs->red[1], c->hs->red[2], c->h, prm->rho_minus, prm->rho_plus, prm->mu_minus,
                 prm->mu_plus, prm->sigma, cfl_number);
  return P2D_OK;
}

void p2d_step_params_default(p2d_step_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->gravity_x = 0.0;
  p->gravity_y = -9.8;
  p->rho_minus = 1.0;
  p->rho_plus = 1000.0;
  p->sigma = 0.0;
  p->mu_minus = 0.0;
  p->mu_plus = 0.0;
  p->mollification = 0.0;
  p->weno_epsilon = 1e-6;
  p->reinit_iterations = 25;
  p->reinit_scaling = 5.0;
  p->rk_order = 3;
}

int p2d_cfl_dt_noalloc(int nx, int ny, double h, const double* phi, const double* u, const double* v,
                       const p2d_step_params* prm, double cfl_number, double* dt, void* hip_stream) {
  if (nx < 2 || ny < 2 || !(h > 0) || !u || !v || !prm || !dt || (!phi && prm->sigma != 0.0) ||
      !(prm->rho_minus > 0) || !(prm->rho_plus > 0))
    return P2D_BAD_ARG;
  const dim3 g = bgrid(nx + 1, ny + 1);
  if (count(g) > (size_t)P2D_NOALLOC_MAX_BLOCKS) {
    fprintf(stderr,
            "p2d_cfl_dt_noalloc: a %d x %d grid needs %zu reduction blocks; the static scratch holds %d "
            "(P2D_NOALLOC_MAX_BLOCKS)\n",
            nx, ny, count(g), P2D_NOALLOC_MAX_BLOCKS);
    return P2D_BAD_ARG;
  }
  double* base = nullptr;
  CHK(noalloc_scratch(&base));
  double* partial = base;
  Scal* s = reinterpret_cast<Scal*>(base + 3 * (size_t)P2D_NOALLOC_MAX_BLOCKS);
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  const double* dp = prm->sigma != 0.0 ? phi : nullptr;
  k_clear_err<<<1, 1, 0, st>>>(s);
  k_cfl_max<<<g, BLOCK, 0, st>>>(nx, ny, h, dp, u, v, partial, s);
  k_fin_store<3><<<1, FINISH_THREADS, 0, st>>>(partial, (int)count(g), 7u, 0, s);
  k_cfl_final<<<1, 1, 0, st>>>(s, h, prm->rho_minus, prm->rho_plus, prm->mu_minus, prm->mu_plus, prm->sigma,
                               cfl_number, dt);
  return launched_ok();
}

// One time step of the air-water solver with third-order TVD Runge-Kutta,
// EULERIAN_FLUID_SOLVER_2D::AdvanceOneTimeStepThread, order_for_time_advancing == 3,
// no vortex sheet (EULERIAN_FLUID_SOLVER_2D.h:1365-1797):
//   level set:  phi <- L(L(phi)); 0.75 phi_n + 0.25 phi; L; 1/3 phi_n + 2/3 phi
//               (L = one WENO5 advection stage by the current velocity)
//   re-init:    the same combination of three Reinitialization calls
//   velocity:   the same combination of three SolveForRK stages, each
//               ENO3 advection, gravity, viscosity (Sourcing), wall velocities 0,
//               projection with the new level set's density and jump
// Returns the last projection's statistics; P2D_NOT_CONVERGED if any did not converge.
int p2d_advance(p2d_ctx* c, double* phi, double* u, double* v, double dt, const p2d_step_params* prm, p2d_stats* out,
                int mem) {
  if (!c || !phi || !u || !v || !prm || !(dt > 0) || c->nx < 3 || c->ny < 3 || !(prm->rho_minus > 0) ||
      !(prm->rho_plus > 0) || prm->reinit_iterations < 0 || !(prm->reinit_scaling > 0) || !(prm->weno_epsilon >= 0) ||
      prm->rk_order < 1 || prm->rk_order > 3)
    return P2D_BAD_ARG;
  if (!c->st_phi) {
    CHK(dalloc(c, &c->st_phi, c->N));
    CHK(dalloc(c, &c->st_phin, c->N));
    CHK(dalloc(c, &c->st_u, c->NU));
    CHK(dalloc(c, &c->st_v, c->NV));
    CHK(dalloc(c, &c->st_un, c->NU));
    CHK(dalloc(c, &c->st_vn, c->NV));
  }
  double *P = c->st_phi, *Pn = c->st_phin, *U = c->st_u, *V = c->st_v, *Un = c->st_un, *Vn = c->st_vn;
  CHK(stage_in(c, P, phi, c->N, mem));
  CHK(stage_in(c, U, u, c->NU, mem));
  CHK(stage_in(c, V, v, c->NV, mem));
  const double eps = prm->weno_epsilon;
  const double third = 1 / (double)3, two_thirds = 2 / (double)3;
  const int order = prm->rk_order;
  const size_t pb = c->N * sizeof(double);
  // level-set advection by the step's velocity: one stage (order 1,
  // EULERIAN_FLUID_SOLVER_2D.h:1191), two averaged 1/2-1/2 (order 2, :1308-1313), or
  // TVD-RK3 (:1688-1716)
  double* T1 = c->v64.x;
  if (order == 1) {
    ls_advect_stage(c, P, U, V, dt, eps, T1);
    HIPCHK(hipMemcpyAsync(P, T1, pb, hipMemcpyDeviceToDevice, c->stream));
  } else {
    HIPCHK(hipMemcpyAsync(Pn, P, pb, hipMemcpyDeviceToDevice, c->stream));
    ls_advect_stage(c, P, U, V, dt, eps, T1);
    ls_advect_stage(c, T1, U, V, dt, eps, P);
    if (order == 2) {
      lincomb(c, c->N, 0.5, Pn, 0.5, P, P);
    } else {
      lincomb(c, c->N, 0.75, Pn, 0.25, P, P);
      ls_advect_stage(c, P, U, V, dt, eps, T1);
      lincomb(c, c->N, third, Pn, two_thirds, T1, P);
    }
  }
  // re-initialisation, the same combination (order 1: :1193; order 2: :1324-1329; order 3:
  // :1718-1741)
  if (prm->reinit_iterations > 0) {
    const int it = prm->reinit_iterations;
    const double sc = prm->reinit_scaling;
    if (order == 1) {
      double* r = reinit_run(c, P, T1, it, sc, eps);
      if (r != P) HIPCHK(hipMemcpyAsync(P, r, pb, hipMemcpyDeviceToDevice, c->stream));
    } else {
      HIPCHK(hipMemcpyAsync(Pn, P, pb, hipMemcpyDeviceToDevice, c->stream));
      double* r = reinit_run(c, P, T1, it, sc, eps);
      r = reinit_run(c, r, r == P ? T1 : P, it, sc, eps);
      if (order == 2) {
        lincomb(c, c->N, 0.5, Pn, 0.5, r, P);
      } else {
        lincomb(c, c->N, 0.75, Pn, 0.25, r, P);
        r = reinit_run(c, P, T1, it, sc, eps);
        lincomb(c, c->N, third, Pn, two_thirds, r, P);
      }
    }
  }
  CHK(launched_ok());
  // the projection's fluid: the new level set
  CHK(p2d_set_fluid(c, P, prm->rho_minus, prm->rho_plus, prm->sigma, P2D_MEM_DEVICE));
  const bool visc = prm->mu_minus != 0.0 || prm->mu_plus != 0.0;
  const double moll = prm->mollification > 0 ? prm->mollification : 1.5 * c->h;
  int status = P2D_OK;
  p2d_stats st{};
  // SolveForRK (:1857-1878) from the velocity in (cu, cv), alternating with (sa[0], sa[1]);
  // (cu, cv) is left pointing at the result
  double *cu = U, *cv = V;
  auto other_u = [&](const double* p) { return p == U ? c->sa[0] : U; };
  auto other_v = [&](const double* p) { return p == V ? c->sa[1] : V; };
  auto stage = [&]() -> int {
    double *nu = other_u(cu), *nv = other_v(cv);
    vel_advect_stage(c, cu, cv, dt, nu, nv);
    cu = nu;
    cv = nv;
    k_gravity<<<blocks_for(c->NU + c->NV), BLOCK, 0, c->stream>>>(c->NU, c->NV, dt * prm->gravity_x,
                                                                   dt * prm->gravity_y, cu, cv);
    if (visc) {
      nu = other_u(cu);
      nv = other_v(cv);
      viscosity_step(c, cu, cv, dt, prm->mu_minus, prm->mu_plus, moll, nu, nv);
      cu = nu;
      cv = nv;
    }
    k_wall_zero<<<blocks_for((size_t)std::max(c->nx, c->ny)), BLOCK, 0, c->stream>>>(c->nx, c->ny, cu, cv);
    CHK(launched_ok());
    const int rc = p2d_project(c, cu, cv, dt, &st, P2D_MEM_DEVICE);
    c->reuse_coeff = true;  // the next stages project with the same level set
    if (rc == P2D_NOT_CONVERGED) status = P2D_NOT_CONVERGED;
    else if (rc != P2D_OK) return rc;
    return P2D_OK;
  };
  // velocity: SolveForRK once (order 1, :1197), twice averaged 1/2-1/2 (order 2,
  // :1334-1362), or TVD-RK3 (:1743-1797)
  if (order >= 2) {
    HIPCHK(hipMemcpyAsync(Un, U, c->NU * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(Vn, V, c->NV * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  }
  c->reuse_coeff = false;
  int src = stage();
  if (src == P2D_OK && order >= 2) src = stage();
  if (src != P2D_OK) {
    c->reuse_coeff = false;
    return src;
  }
  if (order == 1) {
    c->reuse_coeff = false;
    if (cu != U) HIPCHK(hipMemcpyAsync(U, cu, c->NU * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (cv != V) HIPCHK(hipMemcpyAsync(V, cv, c->NV * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  } else if (order == 2) {
    c->reuse_coeff = false;
    lincomb(c, c->NU, 0.5, Un, 0.5, cu, U);
    lincomb(c, c->NV, 0.5, Vn, 0.5, cv, V);
  } else {
    lincomb(c, c->NU, 0.75, Un, 0.25, cu, U);
    lincomb(c, c->NV, 0.75, Vn, 0.25, cv, V);
    cu = U;
    cv = V;
    src = stage();
    c->reuse_coeff = false;
    if (src != P2D_OK) return src;
    lincomb(c, c->NU, third, Un, two_thirds, cu, U);
    lincomb(c, c->NV, third, Vn, two_thirds, cv, V);
  }
  CHK(launched_ok());
  CHK(stage_out(c, phi, P, c->N, mem));
  CHK(stage_out(c, u, U, c->NU, mem));
  CHK(stage_out(c, v, V, c->NV, mem));
  HIPCHK(hipStreamSynchronize(c->stream));
  st.status = status;
  if (out) *out = st;
  return status;
}

int p2d_advect_levelset(p2d_ctx* c, double* phi, const double* u, const double* v, double dt, double epsilon,
                        int mem) {
  if (!c || !phi || !u || !v || !(epsilon >= 0) || !(dt == dt) || c->nx < 3 || c->ny < 3) return P2D_BAD_ARG;
  const double* du = u;
  const double* dv = v;
  if (mem == P2D_MEM_HOST) {
    CHK(stage_in(c, c->su, u, c->NU, mem));
    CHK(stage_in(c, c->sv, v, c->NV, mem));
    du = c->su;
    dv = c->sv;
  }
  CHK(stage_in(c, c->v64.x, phi, c->N, mem));
  ls_advect_stage(c, c->v64.x, du, dv, dt, epsilon, c->v64.z);
  CHK(launched_ok());
  CHK(stage_out(c, phi, c->v64.z, c->N, mem));
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_advect_velocity(p2d_ctx* c, double* u, double* v, double dt, int mem) {
  if (!c || !u || !v || !(dt == dt) || c->nx < 3 || c->ny < 3) return P2D_BAD_ARG;
  // old velocity in su/sv, the new components in sa[0]/sa[1], then back
  CHK(stage_in(c, c->su, u, c->NU, mem));
  CHK(stage_in(c, c->sv, v, c->NV, mem));
  vel_advect_stage(c, c->su, c->sv, dt, c->sa[0], c->sa[1]);
  CHK(launched_ok());
  CHK(stage_out(c, u, c->sa[0], c->NU, mem));
  CHK(stage_out(c, v, c->sa[1], c->NV, mem));
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_apply_viscosity(p2d_ctx* c, double* u, double* v, double dt, double mu_minus, double mu_plus,
                        double epsilon, int mem) {
  // the densities come from p2d_set_fluid (explicit coefficients carry none)
  if (!c || !u || !v || !c->have_fluid || c->explicit_coeff || !(epsilon > 0) || !(dt == dt) || c->nx < 2 ||
      c->ny < 2)
    return P2D_BAD_ARG;
  CHK(stage_in(c, c->su, u, c->NU, mem));
  CHK(stage_in(c, c->sv, v, c->NV, mem));
  viscosity_step(c, c->su, c->sv, dt, mu_minus, mu_plus, epsilon, c->sa[0], c->sa[1]);
  CHK(launched_ok());
  CHK(stage_out(c, u, c->sa[0], c->NU, mem));
  CHK(stage_out(c, v, c->sa[1], c->NV, mem));
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_levelset_normals(p2d_ctx* c, const double* phi, double* normal_x, double* normal_y, int mem) {
  if (!c || !phi || !normal_x || !normal_y || c->nx < 2 || c->ny < 2) return P2D_BAD_ARG;
  CHK(stage_in(c, c->v64.x, phi, c->N, mem));
  k_normals<<<bgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, c->v64.x, c->v64.z, c->v64.q);
  CHK(launched_ok());
  CHK(stage_out(c, normal_x, c->v64.z, c->N, mem));
  CHK(stage_out(c, normal_y, c->v64.q, c->N, mem));
  HIPCHK(hipStreamSynchronize(c->stream));
  return P2D_OK;
}

int p2d_kernel_timing(p2d_ctx* c, int which, int enable) {
  if (!c || which < 0 || which >= KT_COUNT) return P2D_BAD_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  kt_harvest(c);
  auto& t = c->ktimer[which];
  t.total_ms = 0;
  t.launches = 0;
  if (enable) c->timing_mask |= 1u << which;
  else c->timing_mask &= ~(1u << which);
  return P2D_OK;
}

int p2d_kernel_time(p2d_ctx* c, int which, double* total_ms, int* launches) {
  if (!c || which < 0 || which >= KT_COUNT || !total_ms || !launches) return P2D_BAD_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  kt_harvest(c);
  *total_ms = c->ktimer[which].total_ms;
  *launches = (int)c->ktimer[which].launches;
  return P2D_OK;
}

int p2d_time_kernel(p2d_ctx* c, int which, int reps, double* avg_ms) {
  if (!c || !avg_ms || reps < 1 || which < 0 || which > 5 || !c->coeff_ready) return P2D_BAD_ARG;
  const size_t N = c->N;
  CHK(ensure_mg(c, P2D_PREC_F32));
  CHK(ensure_mg(c, P2D_PREC_F64));
  // deterministic non-constant inputs so p.Ap > 0 and no kernel takes its early exit
  k_pattern<float><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v32.z);
  k_pattern<float><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v32.p[0]);
  k_pattern<float><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v32.p[1]);
  k_pattern<float><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v32.r[0]);
  k_pattern<float><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v32.q);
  k_pattern<double><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v64.z);
  k_pattern<double><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v64.p[0]);
  k_pattern<double><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v64.p[1]);
  k_pattern<float><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->v32.x);
  k_pattern<double><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->ph[0]);
  k_fill<float><<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->pl[0], 0.0f);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const Level<float>* L32 = c->mg32.lv.data();
  const int nu1 = c->cfg.mg_pre_sweeps, nu2 = c->cfg.mg_post_sweeps;
  auto launch = [&]() {
    switch (which) {
      case 0:
        k_dir<float><<<dirgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(
            c->nx, c->ny, L32[0].bw, L32[0].bs, c->v32.z, c->v32.p[1], c->v32.p[0], c->v32.q, c->v32.r[0], 0, c->partial,
            c->s);
        break;
      case 1:
        k_dir<double><<<dirgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(
            c->nx, c->ny, c->bw, c->bs, c->v64.z, c->v64.p[1], c->v64.p[0], c->v64.q, c->v64.r[0], 0, c->partial,
            c->s);
        break;
      case 2:
        k_mg_pre<float, true><<<tgrid(c->nx, c->ny, 2 * nu1 + 1), TBLOCK, 0, c->stream>>>(
            L32[0], L32[1], c->v32.r[0], L32[0].xa, nu1, c->v32.r[1], c->v32.x, c->v32.p[0], c->v32.q, 1, 0, 0,
            c->partial, c->s);
        break;
      case 3:
        k_mg_post<float, true><<<tgrid(c->nx, c->ny, 2 * nu2, Tile<16>::ty_post), TBLOCK, 0, c->stream>>>(
            L32[0], L32[1], c->v32.r[0], L32[0].xa, c->v32.z, nu2, c->partial, c->s);
        break;
      case 5:  // one re-initialisation step of the fluid's level set
        k_reinit_step<<<dgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(c->nx, c->ny, c->h, c->h / 5.0, 1e-6, c->phi,
                                                                    c->v64.q, c->v64.z);
        break;
      default:
        k_correct_resid<float, float><<<fgrid(c->nx, c->ny), BLOCK, 0, c->stream>>>(
            c->nx, c->ny, c->bw, c->bs, nullptr, nullptr, c->b, c->ph[0], c->pl[0], c->v32.x, nullptr, c->ph[1],
            c->pl[1], c->v32.r[0], c->partial, c->s);
    }
  };
  k_scal_arm<<<1, 1, 0, c->stream>>>(c->s, 1, 0.25, 1e-3, 1);
  if (which == 5) k_reinit_sign<<<blocks_for(N), BLOCK, 0, c->stream>>>(N, c->h, c->phi, c->v64.q);
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
  c->have_pressure = false;  // the patterns overwrote solver state
  return P2D_OK;
}

}  // extern "C"

// --------------------------------------------------------------------------- semi-implicit viscosity
// One grid's solver state for p2d_visc_*: coefficient, stencil and BiCGStab vectors sized for
// the larger MAC component, staging for host arrays, the device scalars and their pinned mirror.
struct p2d_visc {
  int nx = 0, ny = 0;
  double dx = 0, dy = 0;
  size_t nmax = 0;
  hipStream_t stream = nullptr, own = nullptr;
  std::vector<void*> allocs;
  double* c[5] = {};
  double *ex = nullptr, *dg = nullptr, *cw = nullptr, *ce = nullptr, *cs = nullptr, *cn = nullptr;
  double *x = nullptr, *r = nullptr, *rhat = nullptr, *p = nullptr, *v = nullptr, *y = nullptr, *sv = nullptr,
         *z = nullptr, *t = nullptr;
  double *uo = nullptr, *hphi = nullptr, *hu = nullptr, *hv = nullptr, *partial = nullptr;
  VScal* s = nullptr;
  VScal* hs = nullptr;
};

namespace {
int valloc(p2d_visc* V, double** p, size_t n) {
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(double)));
  HIPCHK(hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(double)));
  V->allocs.push_back(q);
  *p = (double*)q;
  return P2D_OK;
}
int vcopy(p2d_visc* V, double* dst, const double* src, size_t n, hipMemcpyKind kind) {
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), kind, V->stream));
  return P2D_OK;
}
// Solves the component system for V->x (initial guess in V->x) with the coefficients in
// V->c[0..4] and the right-hand side in V->ex.
int visc_solve_device(p2d_visc* V, int comp, int mode, double tolerance, int max_iterations, p2d_stats* st) {
  const int nxc = comp == 0 ? V->nx + 1 : V->nx, nyc = comp == 0 ? V->ny : V->ny + 1;
  const int n = nxc * nyc;
  const unsigned G = gs_blocks((size_t)n);
  const double inv_dxdx = 1.0 / (V->dx * V->dx), inv_dydy = 1.0 / (V->dy * V->dy);
  hipStream_t q = V->stream;
  k_visc_rows<<<G, BLOCK, 0, q>>>(nxc, nyc, inv_dxdx, inv_dydy, mode, V->c[0], V->c[1], V->c[2], V->c[3], V->c[4],
                                  V->dg, V->cw, V->ce, V->cs, V->cn);
  std::memset(V->hs, 0, sizeof(VScal));
  V->hs->tol2 = tolerance * tolerance;
  V->hs->max_iter = max_iterations;
  HIPCHK(hipMemcpyAsync(V->s, V->hs, sizeof(VScal), hipMemcpyHostToDevice, q));
  const VStencil A{V->dg, V->cw, V->ce, V->cs, V->cn, nxc, nyc, mode};
  k_bs_init<<<G, BLOCK, 0, q>>>(A, V->ex, V->x, V->r, V->rhat, V->p, V->v, V->partial);
  k_bs_fin_init<<<1, FINISH_THREADS, 0, q>>>(V->partial, (int)G, V->s);
  CHK(launched_ok());
  int chunk = 4;
  for (;;) {
    HIPCHK(hipMemcpyAsync(V->hs, V->s, sizeof(VScal), hipMemcpyDeviceToHost, q));
    HIPCHK(hipStreamSynchronize(q));
    if (V->hs->done) break;
    for (int k = 0; k < chunk; ++k) {
      k_bs_dir<<<G, BLOCK, 0, q>>>(n, V->r, V->p, V->v, V->dg, V->y, V->s);
      k_bs_apply<false><<<G, BLOCK, 0, q>>>(A, V->y, V->v, V->rhat, V->partial, V->s);
      k_bs_fin_alpha<<<1, FINISH_THREADS, 0, q>>>(V->partial, (int)G, V->s);
      k_bs_half<<<G, BLOCK, 0, q>>>(n, V->r, V->v, V->dg, V->sv, V->z, V->s);
      k_bs_apply<true><<<G, BLOCK, 0, q>>>(A, V->z, V->t, V->sv, V->partial, V->s);
      k_bs_fin_omega<<<1, FINISH_THREADS, 0, q>>>(V->partial, (int)G, V->s);
      k_bs_update<<<G, BLOCK, 0, q>>>(n, V->x, V->y, V->z, V->sv, V->t, V->r, V->rhat, V->partial, V->s);
      k_bs_fin_update<<<1, FINISH_THREADS, 0, q>>>(V->partial, (int)G, V->s);
    }
    CHK(launched_ok());
    chunk = std::min(chunk * 2, 64);
  }
  const VScal& h = *V->hs;
  int status = P2D_OK;
  if (h.flag) status = P2D_NAN;
  else if (!(h.r2 < h.tol2)) status = P2D_NOT_CONVERGED;
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->status = status;
    st->iterations = h.iter;
    st->outer = 1;
    st->rel_residual = h.b2 > 0 ? std::sqrt(h.r2 / h.b2) : 0.0;
    st->max_abs_residual = std::sqrt(h.r2);
  }
  return status;
}
}  // namespace

int p2d_visc_create(int nx, int ny, double dx, double dy, p2d_visc** out) {
  if (!out) return P2D_BAD_ARG;
  *out = nullptr;
  if (nx < 2 || ny < 2 || !(dx > 0) || !(dy > 0) || (long)(nx + 1) * (ny + 1) > (1L << 30)) return P2D_BAD_ARG;
  p2d_visc* V = new p2d_visc();
  V->nx = nx;
  V->ny = ny;
  V->dx = dx;
  V->dy = dy;
  V->nmax = std::max((size_t)(nx + 1) * ny, (size_t)nx * (ny + 1));
  auto fail = [&](int rc) {
    p2d_visc_destroy(V);
    return rc;
  };
  if (hipStreamCreate(&V->own) != hipSuccess) return fail(P2D_HIP_ERROR);
  V->stream = V->own;
  const size_t M = V->nmax;
  double** vecs[] = {&V->c[0], &V->c[1], &V->c[2], &V->c[3], &V->c[4], &V->ex, &V->dg, &V->cw, &V->ce, &V->cs,
                     &V->cn, &V->x, &V->r, &V->rhat, &V->p, &V->v, &V->y, &V->sv, &V->z, &V->t, &V->uo, &V->hu,
                     &V->hv};
  for (double** a : vecs) {
    const int rc = valloc(V, a, M);
    if (rc != P2D_OK) return fail(rc);
  }
  int rc = valloc(V, &V->hphi, (size_t)nx * ny);
  if (rc == P2D_OK) rc = valloc(V, &V->partial, 3 * (size_t)GRIDSTRIDE_BLOCKS);
  if (rc != P2D_OK) return fail(rc);
  if (hipMalloc((void**)&V->s, sizeof(VScal)) != hipSuccess) return fail(P2D_HIP_ERROR);
  V->allocs.push_back(V->s);
  if (hipHostMalloc((void**)&V->hs, sizeof(VScal), hipHostMallocDefault) != hipSuccess) return fail(P2D_HIP_ERROR);
  *out = V;
  return P2D_OK;
}

int p2d_visc_destroy(p2d_visc* V) {
  if (!V) return P2D_BAD_ARG;
  if (V->stream) (void)hipStreamSynchronize(V->stream);
  for (void* p : V->allocs) (void)hipFree(p);
  if (V->hs) (void)hipHostFree(V->hs);
  if (V->own) (void)hipStreamDestroy(V->own);
  delete V;
  return P2D_OK;
}

int p2d_visc_set_stream(p2d_visc* V, void* stream) {
  if (!V) return P2D_BAD_ARG;
  HIPCHK(hipStreamSynchronize(V->stream));
  V->stream = stream ? (hipStream_t)stream : V->own;
  return P2D_OK;
}

int p2d_visc_solve(p2d_visc* V, int component, double* velocity, const double* c1, const double* c2,
                   const double* c3, const double* c4, const double* c5, const double* explicit_term, int boundary,
                   double tolerance, int max_iterations, p2d_stats* stats, int mem) {
  if (!V || !velocity || !c1 || !c2 || !c3 || !c4 || !c5 || !explicit_term || (component != 0 && component != 1) ||
      (boundary != P2D_VISC_REFERENCE && boundary != P2D_VISC_DIRICHLET) || !(tolerance >= 0) || max_iterations < 0)
    return P2D_BAD_ARG;
  const size_t n = component == 0 ? (size_t)(V->nx + 1) * V->ny : (size_t)V->nx * (V->ny + 1);
  const hipMemcpyKind in = mem == P2D_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  const double* cs[5] = {c1, c2, c3, c4, c5};
  for (int k = 0; k < 5; ++k) CHK(vcopy(V, V->c[k], cs[k], n, in));
  CHK(vcopy(V, V->ex, explicit_term, n, in));
  CHK(vcopy(V, V->x, velocity, n, in));
  const int rc = visc_solve_device(V, component, boundary == P2D_VISC_REFERENCE ? VISC_LITERAL : VISC_DIRICHLET,
                                   tolerance, max_iterations, stats);
  if (rc != P2D_OK && rc != P2D_NOT_CONVERGED) return rc;
  CHK(vcopy(V, velocity, V->x, n, mem == P2D_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice));
  HIPCHK(hipStreamSynchronize(V->stream));
  return rc;
}

void p2d_oil_water_params_default(p2d_oil_water_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->x_min = 0.0;
  p->Re = 0.94983;
  p->rho_oil = 0.905;
  p->rho_water = 0.995;
  p->mu_oil = 6.01;
  p->mu_water = 0.01;
  p->mollification = 0.0;
  p->tolerance = 1e-6;
  p->max_iterations = 10000;
  p->boundary = P2D_VISC_REFERENCE;
}

// VISCOSITY_2D::ApplyViscosity for the oil-water solver with semi_implicit_approach and
// dimensionless_form (VISCOSITY_2D.h:546-640): the x component's coefficients and explicit
// term from the level set and the current (u, v), its solve, then the y component's from
// the level set, v and u before and after the x solve, and its solve.
int p2d_visc_oil_water(p2d_visc* V, const double* phi, double* u, double* v, double dt,
                       const p2d_oil_water_params* prm, p2d_stats* stats_x, p2d_stats* stats_y, int mem) {
  if (!V || !phi || !u || !v || !prm || !(dt > 0) || !(prm->Re > 0) || !(prm->rho_water > 0) ||
      !(prm->mu_water > 0) || !(prm->tolerance >= 0) || prm->max_iterations < 0 ||
      (prm->boundary != P2D_VISC_REFERENCE && prm->boundary != P2D_VISC_DIRICHLET))
    return P2D_BAD_ARG;
  const int nx = V->nx, ny = V->ny;
  const size_t NU = (size_t)(nx + 1) * ny, NV = (size_t)nx * (ny + 1), N = (size_t)nx * ny;
  const hipMemcpyKind in = mem == P2D_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  const hipMemcpyKind back = mem == P2D_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  CHK(vcopy(V, V->hphi, phi, N, in));
  CHK(vcopy(V, V->hu, u, NU, in));
  CHK(vcopy(V, V->hv, v, NV, in));
  CHK(vcopy(V, V->uo, V->hu, NU, hipMemcpyDeviceToDevice));
  const OilWater o{V->dx, V->dy, prm->x_min, dt, prm->Re, prm->rho_oil, prm->rho_water, prm->mu_oil, prm->mu_water,
                   prm->mollification > 0 ? prm->mollification : 1.5 * V->dx};
  const int mode = prm->boundary == P2D_VISC_REFERENCE ? VISC_LITERAL : VISC_DIRICHLET;
  k_ow_coef_x<<<gs_blocks(NU), BLOCK, 0, V->stream>>>(nx, ny, o, V->hphi, V->hu, V->hv, V->c[0], V->c[1], V->c[2],
                                                       V->c[3], V->c[4], V->ex);
  CHK(vcopy(V, V->x, V->hu, NU, hipMemcpyDeviceToDevice));
  CHK(launched_ok());
  int rx = visc_solve_device(V, 0, mode, prm->tolerance, prm->max_iterations, stats_x);
  if (rx != P2D_OK && rx != P2D_NOT_CONVERGED) return rx;
  CHK(vcopy(V, V->hu, V->x, NU, hipMemcpyDeviceToDevice));
  k_ow_coef_y<<<gs_blocks(NV), BLOCK, 0, V->stream>>>(nx, ny, o, V->hphi, V->uo, V->hu, V->hv, V->c[0], V->c[1],
                                                       V->c[2], V->c[3], V->c[4], V->ex);
  CHK(vcopy(V, V->x, V->hv, NV, hipMemcpyDeviceToDevice));
  CHK(launched_ok());
  int ry = visc_solve_device(V, 1, mode, prm->tolerance, prm->max_iterations, stats_y);
  if (ry != P2D_OK && ry != P2D_NOT_CONVERGED) return ry;
  CHK(vcopy(V, u, V->hu, NU, back));
  CHK(vcopy(V, v, V->x, NV, back));
  HIPCHK(hipStreamSynchronize(V->stream));
  return (rx == P2D_OK && ry == P2D_OK) ? P2D_OK : P2D_NOT_CONVERGED;
}
