// p2d oracle -- TEST INFRASTRUCTURE ONLY.
//
// A clean-room FP64 CPU restatement of 2D_Engine's pressure projection
// (PROJECTION_2D + POISSON_SOLVER jump-condition variable-coefficient assembly +
// CG / PCG(IC0) / PCG(Jacobi)), plus the CPU mirror of this library's own
// geometric multigrid preconditioner and hand-written adjoint.
//
// Only tests, the smoke check and the CPU-baseline leg of the benchmark may
// load this library, and only as the checker / baseline.  The product path
// (the HIP library) never calls it.
//
// Layout (identical to the product's C ABI): nx x ny cells, spacing h, arrays
// i-fastest with no ghosts.  Cell (i,j) -> i + j*nx.  x-faces u: (nx+1) x ny,
// face (i,j) -> i + j*(nx+1) sits between cells (i-1,j) and (i,j).  y-faces v:
// nx x (ny+1), face (i,j) -> i + j*nx sits between cells (i,j-1) and (i,j).
// "minus" side = phi <= 0 (reference: POISSON_SOLVER.cpp:875-949).
//
// Every function cites the reference lines it follows.  Where this file
// deviates from the reference on purpose, the comment says "DEVIATION".
//
// How each part is pinned (tests/test_oracle_*.py; the reference itself does
// not build here: Windows.h and Boost):
//   face beta, jump terms, assembly  closed forms, matrix invariants, the
//                                    Liu-Fedkiw-Kang known answer (test 3) under
//                                    refinement, balanced-force drop
//   curvature                        circle and RT closed forms at O(h^2), clamp,
//                                    a linear level set (corner cells included)
//   IC(0) factor and solve           L L^T = A on A's pattern; equal to a dense
//                                    textbook IC(0) and its triangular solves
//   CG / IC(0)-PCG iterates          iterate by iterate against dense textbook
//                                    restatements; iteration counts under the
//                                    absolute rule and the off-by-one cap
//   CG / IC(0)-PCG / Jacobi-PCG /    agreement on the mean-zero pressure (the
//   MG-PCG                           system fixes it)
//   multigrid V-cycle                symmetry, positivity, linearity, spectrum; the
//                                    smoother (SOR omega included) against a dense
//                                    textbook red-black sweep around the nu = 0 cycle
//   consistent update, adjoint       D u_proj = -dt r identity, central differences,
//                                    idempotence and self-adjointness
//   reference-literal update         closed forms: linear pressure (wall faces against
//                                    the ghost pressure 0), face-aligned interface,
//                                    exactly zero level set, mollification band
// Not reproducible by design: the curvature of the four corner cells, where the
// reference reads corner ghosts that it never fills (LEVELSET_2D.h:935); here they
// are extrapolated (DEVIATION in levelset_with_ghosts) and pinned by the linear case.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Boundary-condition codes (reference MACROS.h:3-12).
constexpr int BC_FULL = 0;
constexpr int BC_DIR = -1;
constexpr int BC_NEUM = -4;
constexpr double REF_PI = 3.141592;  // MACROS.h:5 (the reference's truncated pi)

// A ghost-padded field with one ghost layer: (nx+2) x (ny+2), index (i,j) with
// i in [-1, nx], j in [-1, ny].  Mirrors the reference's ARRAY_2D ghost box
// (ARRAY_2D.h:140-146) for the one layer the 5-point path touches.
struct Pad {
  int nx, ny;
  std::vector<double> a;
  Pad(int nx_, int ny_, double fill = 0.0) : nx(nx_), ny(ny_), a((size_t)(nx_ + 2) * (ny_ + 2), fill) {}
  double& operator()(int i, int j) { return a[(size_t)(i + 1) + (size_t)(j + 1) * (nx + 2)]; }
  double operator()(int i, int j) const { return a[(size_t)(i + 1) + (size_t)(j + 1) * (nx + 2)]; }
};
struct PadI {
  int nx, ny;
  std::vector<int> a;
  PadI(int nx_, int ny_, int fill = 0) : nx(nx_), ny(ny_), a((size_t)(nx_ + 2) * (ny_ + 2), fill) {}
  int& operator()(int i, int j) { return a[(size_t)(i + 1) + (size_t)(j + 1) * (nx + 2)]; }
  int operator()(int i, int j) const { return a[(size_t)(i + 1) + (size_t)(j + 1) * (nx + 2)]; }
};

inline size_t C(int i, int j, int nx) { return (size_t)i + (size_t)j * nx; }
inline size_t FX(int i, int j, int nx) { return (size_t)i + (size_t)j * (nx + 1); }
inline size_t FY(int i, int j, int nx) { return (size_t)i + (size_t)j * nx; }

// Interior copy + linear-extrapolation face ghosts.
// FIELD_STRUCTURE_2D::FillGhostCellsContinuousDerivativesFrom, FIELD_STRUCTURE_2D.h:585-639:
//   left ghost = 2*phi(i+1) - phi(i+2), right = 2*phi(i-1) - phi(i-2), same in y.
// DEVIATION: the reference leaves the four corner ghosts unfilled; here they are
// extrapolated from the filled face ghosts along x (equal to extrapolating along y).
Pad levelset_with_ghosts(int nx, int ny, const double* phi) {
  Pad p(nx, ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) p(i, j) = phi[C(i, j, nx)];
  for (int j = 0; j < ny; ++j) {
    p(-1, j) = 2.0 * p(0, j) - p(1, j);
    p(nx, j) = 2.0 * p(nx - 1, j) - p(nx - 2, j);
  }
  for (int i = 0; i < nx; ++i) {
    p(i, ny) = 2.0 * p(i, ny - 1) - p(i, ny - 2);
    p(i, -1) = 2.0 * p(i, 0) - p(i, 1);
  }
  p(-1, -1) = 2.0 * p(0, -1) - p(1, -1);
  p(nx, -1) = 2.0 * p(nx - 1, -1) - p(nx - 2, -1);
  p(-1, ny) = 2.0 * p(0, ny) - p(1, ny);
  p(nx, ny) = 2.0 * p(nx - 1, ny) - p(nx - 2, ny);
  return p;
}

// Mirror ghosts: FIELD_STRUCTURE_2D::FillGhostCellsFrom(.., copy_real_data),
// FIELD_STRUCTURE_2D.h:528-582: left ghost = interior(i_start+1), right = interior(i_end-1).
Pad mirror_with_ghosts(int nx, int ny, const double* f) {
  Pad p(nx, ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) p(i, j) = f[C(i, j, nx)];
  for (int j = 0; j < ny; ++j) {
    p(-1, j) = p(1, j);
    p(nx, j) = p(nx - 2, j);
  }
  for (int i = 0; i < nx; ++i) {
    p(i, ny) = p(i, ny - 2);
    p(i, -1) = p(i, 1);
  }
  return p;
}

// GFM face coefficient, POISSON_SOLVER.cpp:743-746:
//   beta = v_c v_n (|phi_n| + |phi_c|) / (v_c |phi_n| + v_n |phi_c|),  v = 1/rho.
// Written with (L,R) = (left/bottom cell, right/top cell); the reference's
// evaluation from either cell gives the same bits (products and sums commute).
// DEVIATION: the reference returns 0/0 = NaN when phi_L = phi_R = 0; both
// cells are then on the minus side and this returns v_R.
inline double gfm_beta(double vL, double vR, double phiL, double phiR) {
  const double aL = std::fabs(phiL), aR = std::fabs(phiR);
  const double den = vR * aL + vL * aR;
  if (den == 0.0) return vR;
  return vR * vL * (aL + aR) / den;
}

// Jump value interpolated to the face, POISSON_SOLVER.cpp:879 / 897 (a_l, a_r):
//   a = (jc_c |phi_n| + jc_n |phi_c|) / (|phi_c| + |phi_n|).
inline double face_jump(double jcL, double jcR, double phiL, double phiR) {
  const double aL = std::fabs(phiL), aR = std::fabs(phiR);
  return (jcR * aL + jcL * aR) / (aR + aL);
}

inline bool minus_side(double phi) { return phi <= 0.0; }

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// A2: projection density, PROJECTION_2D.h:323-393 (air_bubble_rising branch):
//   rho = rho_minus (air) where phi <= 0, rho_plus (water) where phi > 0;
//   threaded variant also stores 1/rho (388-392).
void orc_density(int nx, int ny, const double* phi, double rho_minus, double rho_plus,
                 double* rho, double* inv_rho) {
  for (size_t k = 0; k < (size_t)nx * ny; ++k) {
    const double r = phi[k] <= 0.0 ? rho_minus : rho_plus;
    if (rho) rho[k] = r;
    if (inv_rho) inv_rho[k] = 1.0 / r;
  }
}

// A3: divergence scaled by 1/dt, PROJECTION_2D.h:435-478:
//   div = (1/dx)(u(i+1,j)-u(i,j)) + (1/dy)(v(i,j+1)-v(i,j)); div *= 1/dt.
void orc_divergence(int nx, int ny, double h, double dt, const double* u, const double* v,
                    double* div) {
  const double ih = 1.0 / h, idt = 1.0 / dt;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double d = ih * (u[FX(i + 1, j, nx)] - u[FX(i, j, nx)]) + ih * (v[FY(i, j + 1, nx)] - v[FY(i, j, nx)]);
      div[C(i, j, nx)] = d * idt;
    }
}

// A4: curvature from the level set, LEVELSET_2D::ComputeCurvaturesThread,
// LEVELSET_2D.h:898-990 (ghosts from FIELD_STRUCTURE_2D.h:641-716):
//   central phi_x, phi_y, phi_xx, phi_yy, phi_xy; kappa = -(px^2 pyy - 2 px py pxy + py^2 pxx)/|grad|^3;
//   one-sided px, py if |grad| == 0; clamp to +-1/min(dx,dy).
// Returns 0, or -1 where the reference would exit(0) (zero one-sided gradient).
int orc_curvature(int nx, int ny, double h, const double* phi_in, double* kappa) {
  const Pad phi = levelset_with_ghosts(nx, ny, phi_in);
  const double one_over_2dx = 1.0 / (2.0 * h), one_over_dx2 = 1.0 / (h * h), one_over_dx = 1.0 / h;
  const double tol = 1.0 / h;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double px = (phi(i + 1, j) - phi(i - 1, j)) * one_over_2dx;
      double py = (phi(i, j + 1) - phi(i, j - 1)) * one_over_2dx;
      const double pxx = (phi(i + 1, j) - 2 * phi(i, j) + phi(i - 1, j)) * one_over_dx2;
      const double pyy = (phi(i, j + 1) - 2 * phi(i, j) + phi(i, j - 1)) * one_over_dx2;
      const double pxy = (phi(i + 1, j + 1) - phi(i + 1, j - 1) - phi(i - 1, j + 1) + phi(i - 1, j - 1)) *
                         one_over_2dx * one_over_2dx;
      double mag = std::sqrt(px * px + py * py);
      double deno = mag * mag * mag;
      double curv;
      if (deno != 0) {
        curv = -(px * px * pyy - 2.0 * px * py * pxy + py * py * pxx) * (1.0 / deno);
      } else {
        px = (phi(i + 1, j) - phi(i, j)) * one_over_dx;
        py = (phi(i, j + 1) - phi(i, j)) * one_over_dx;
        mag = std::sqrt(px * px + py * py);
        deno = mag * mag * mag;
        if (deno == 0) return -1;
        curv = -(px * px * pyy - 2.0 * px * py * pxy + py * py * pxx) * (1.0 / deno);
      }
      if (std::fabs(curv) <= tol) kappa[C(i, j, nx)] = curv;
      else if (curv > tol) kappa[C(i, j, nx)] = tol;
      else kappa[C(i, j, nx)] = -tol;
    }
  return 0;
}

// Unit normals, LEVELSET_2D::ComputeNormals, LEVELSET_2D.h:593-620 (central
// differences over the linear-extrapolation ghosts; one-sided if zero).
void orc_normals(int nx, int ny, double h, const double* phi_in, double* nxo, double* nyo) {
  const Pad phi = levelset_with_ghosts(nx, ny, phi_in);
  const double one_over_2dx = 1.0 / (2.0 * h), one_over_dx = 1.0 / h;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double a = (phi(i + 1, j) - phi(i - 1, j)) * one_over_2dx, b = (phi(i, j + 1) - phi(i, j - 1)) * one_over_2dx;
      if (std::sqrt(a * a + b * b) == 0) {
        a = (phi(i + 1, j) - phi(i, j)) * one_over_dx;
        b = (phi(i, j + 1) - phi(i, j)) * one_over_dx;
      }
      // VECTOR_2D::MakeThisUnit (VECTOR_2D.h:136-150): scaled by 1/|n|, (0,0) when |n| = 0
      const double m = std::sqrt(a * a + b * b), inv = 1.0 / m;
      nxo[C(i, j, nx)] = m != 0 ? a * inv : 0.0;
      nyo[C(i, j, nx)] = m != 0 ? b * inv : 0.0;
    }
}

// GFM face coefficients beta = 1/rho_f on interior faces (POISSON_SOLVER.cpp:743-746
// with v = 1/rho from A2); wall faces are 0 (no-flux walls).
void orc_face_beta(int nx, int ny, const double* phi, double rho_minus, double rho_plus, double* bx,
                   double* by) {
  const double vm = 1.0 / rho_minus, vp = 1.0 / rho_plus;
  auto v = [&](int i, int j) { return phi[C(i, j, nx)] <= 0.0 ? vm : vp; };
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      double b = 0.0;
      if (i > 0 && i < nx) b = gfm_beta(v(i - 1, j), v(i, j), phi[C(i - 1, j, nx)], phi[C(i, j, nx)]);
      bx[FX(i, j, nx)] = b;
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double b = 0.0;
      if (j > 0 && j < ny) b = gfm_beta(v(i, j - 1), v(i, j), phi[C(i, j - 1, nx)], phi[C(i, j, nx)]);
      by[FY(i, j, nx)] = b;
    }
}

// ---------------------------------------------------------------------------
// A6-A9: reference-literal CSR assembly,
// POISSON_SOLVER::BuildLinearSystemNodeJumpConditionVaribleCoefficient, POISSON_SOLVER.cpp:714-976,
// with AssignSequentialindicesToFullCells (1392-1418), CountNonZeroElements (1488-1535)
// and the cursor-style CSR_MATRIX::AssignValue (CSR_MATRIX.h:199-212).
//
// bc_pad: (nx+2)x(ny+2) codes (BC_FULL=0 / BC_DIR=-1 / BC_NEUM=-4) over the ghost box;
//         NULL -> Neumann ghost ring, full interior (PROJECTION_2D.h:1010-1047).
// p_pad:  (nx+2)x(ny+2) pressure incl. Dirichlet values (x0 + lifts); NULL -> 0.
// variable: nx*ny, v (= 1/rho, or the LFK "beta" test coefficient); ghosts mirrored
//           (POISSON_SOLVER.cpp:131-138).
// div: nx*ny right-hand side ("div"), b = -div - F (951).
// jc, jcd: nx*ny jumps on solution / derivative, ghost ring 0 (PROJECTION_2D.h:205-208);
//          either may be NULL (zero).  nrm_x/nrm_y: unit normals (needed only if jcd).
// Outputs: row_ptr[N+1], col[nnz], val[nnz], b[N], x0[N], row_of_cell[nx*ny] (-1 if not full).
// Returns the number of rows N (nnz is row_ptr[N]).
int orc_assemble_reference(int nx, int ny, double h, const int* bc_pad, const double* p_pad,
                           const double* phi_in, const double* variable_in, const double* div,
                           const double* jc_in, const double* jcd_in, const double* nrm_x,
                           const double* nrm_y, int* row_ptr, int* col, double* val, double* b,
                           double* x0, int* row_of_cell) {
  PadI bc(nx, ny, BC_NEUM);
  if (bc_pad) std::memcpy(bc.a.data(), bc_pad, sizeof(int) * bc.a.size());
  else
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) bc(i, j) = BC_FULL;
  Pad pressure(nx, ny);
  if (p_pad) std::memcpy(pressure.a.data(), p_pad, sizeof(double) * pressure.a.size());
  const Pad levelset = levelset_with_ghosts(nx, ny, phi_in);
  const Pad variable = mirror_with_ghosts(nx, ny, variable_in);
  Pad jc(nx, ny), jcd(nx, ny), nx_(nx, ny), ny_(nx, ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      jc(i, j) = jc_in ? jc_in[C(i, j, nx)] : 0.0;
      jcd(i, j) = jcd_in ? jcd_in[C(i, j, nx)] : 0.0;
      nx_(i, j) = nrm_x ? nrm_x[C(i, j, nx)] : 0.0;
      ny_(i, j) = nrm_y ? nrm_y[C(i, j, nx)] : 0.0;
    }
  // AssignSequentialindicesToFullCells: bc >= 0 becomes the row id (i-fastest).
  int N = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      if (bc(i, j) > -1) bc(i, j) = N++;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) row_of_cell[C(i, j, nx)] = bc(i, j) > -1 ? bc(i, j) : -1;

  const double dxdx = h * h, inv_dxdx = 1.0 / dxdx, one_over_dx = 1.0 / h, one_over_dx2 = 1.0 / (h * h);
  int cursor = 0, prev_row = -1;
  auto assign = [&](int row, int column, double value) {  // CSR_MATRIX.h:199-212
    val[cursor] = value;
    if (row != prev_row) {
      row_ptr[row] = cursor;
      prev_row = row;
    }
    col[cursor] = column;
    ++cursor;
  };
  const int i_start = 0, j_start = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      if (bc(i, j) < 0) continue;
      const int row = bc(i, j);
      double coef = 0;
      const double lc = levelset(i, j), ll = levelset(i - 1, j), lr = levelset(i + 1, j), lb = levelset(i, j - 1),
                   lt = levelset(i, j + 1);
      const double beta_l = variable(i, j) * variable(i - 1, j) * (std::fabs(ll) + std::fabs(lc)) /
                            (variable(i, j) * std::fabs(ll) + variable(i - 1, j) * std::fabs(lc));
      const double beta_r = variable(i, j) * variable(i + 1, j) * (std::fabs(lr) + std::fabs(lc)) /
                            (variable(i + 1, j) * std::fabs(lc) + variable(i, j) * std::fabs(lr));
      const double beta_b = variable(i, j) * variable(i, j - 1) * (std::fabs(lb) + std::fabs(lc)) /
                            (variable(i, j) * std::fabs(lb) + variable(i, j - 1) * std::fabs(lc));
      const double beta_t = variable(i, j) * variable(i, j + 1) * (std::fabs(lt) + std::fabs(lc)) /
                            (variable(i, j + 1) * std::fabs(lc) + variable(i, j) * std::fabs(lt));
      // 749-768: full neighbours.
      if (bc(i - 1, j) > -1) { coef += beta_l; assign(row, bc(i - 1, j), -beta_l * inv_dxdx); }
      if (bc(i + 1, j) > -1) { coef += beta_r; assign(row, bc(i + 1, j), -beta_r * inv_dxdx); }
      if (bc(i, j - 1) > -1) { coef += beta_b; assign(row, bc(i, j - 1), -beta_b * inv_dxdx); }
      if (bc(i, j + 1) > -1) { coef += beta_t; assign(row, bc(i, j + 1), -beta_t * inv_dxdx); }
      // 770-786: Dirichlet neighbours.
      if (bc(i - 1, j) == BC_DIR) coef += beta_l;
      if (bc(i + 1, j) == BC_DIR) coef += beta_r;
      if (bc(i, j - 1) == BC_DIR) coef += beta_b;
      if (bc(i, j + 1) == BC_DIR) coef += beta_t;
      // 789-807: Neumann faces count only at (i_start, j_start) -- the pin.
      if (i == i_start && j == j_start) {
        if (bc(i - 1, j) == BC_NEUM) coef += beta_l;
        if (bc(i + 1, j) == BC_NEUM) coef += beta_r;
        if (bc(i, j - 1) == BC_NEUM) coef += beta_b;
        if (bc(i, j + 1) == BC_NEUM) coef += beta_t;
      }
      if (coef == 0) coef = 1;  // 867-870
      assign(row, row, inv_dxdx * coef);  // 872
      // 875-949: jump terms per arm (a on the solution, b on the derivative).
      auto arm = [&](double lcen, double lnb, double jc_c, double jc_n, double jcd_c, double nrm_c, double jcd_n,
                     double nrm_n, double beta, double var_n, double dir_sign) {
        // dir_sign: -1 for left/bottom arms, +1 for right/top arms (sign of the b-term).
        const double subcell = std::fabs(lnb) / (std::fabs(lcen) + std::fabs(lnb));
        const double a = (jc_c * std::fabs(lnb) + jc_n * std::fabs(lcen)) / (std::fabs(lcen) + std::fabs(lnb));
        const double bb = (jcd_c * nrm_c * std::fabs(lnb) + jcd_n * nrm_n * std::fabs(lcen)) /
                          (std::fabs(lcen) + std::fabs(lnb));
        if ((lcen <= 0 && lnb <= 0) || (lcen > 0 && lnb > 0)) return 0.0;
        if (lcen <= 0 && lnb > 0) return one_over_dx2 * a * beta + dir_sign * one_over_dx * beta * bb * subcell / var_n;
        return -one_over_dx2 * a * beta - dir_sign * one_over_dx * beta * bb * subcell / var_n;
      };
      const double F_L = arm(lc, ll, jc(i, j), jc(i - 1, j), jcd(i, j), nx_(i, j), jcd(i - 1, j), nx_(i - 1, j),
                             beta_l, variable(i - 1, j), -1.0);
      const double F_R = arm(lc, lr, jc(i, j), jc(i + 1, j), jcd(i, j), nx_(i, j), jcd(i + 1, j), nx_(i + 1, j),
                             beta_r, variable(i + 1, j), +1.0);
      const double F_B = arm(lc, lb, jc(i, j), jc(i, j - 1), jcd(i, j), ny_(i, j), jcd(i, j - 1), ny_(i, j - 1),
                             beta_b, variable(i, j - 1), -1.0);
      const double F_T = arm(lc, lt, jc(i, j), jc(i, j + 1), jcd(i, j), ny_(i, j), jcd(i, j + 1), ny_(i, j + 1),
                             beta_t, variable(i, j + 1), +1.0);
      const double F_X = F_L + F_R, F_Y = F_B + F_T;
      double bv = -div[C(i, j, nx)] - F_X - F_Y;  // 951
      // 953-972: Dirichlet lifts.
      if (bc(i - 1, j) == BC_DIR) bv += beta_l * inv_dxdx * pressure(i - 1, j);
      if (bc(i + 1, j) == BC_DIR) bv += beta_r * inv_dxdx * pressure(i + 1, j);
      if (bc(i, j - 1) == BC_DIR) bv += beta_b * inv_dxdx * pressure(i, j - 1);
      if (bc(i, j + 1) == BC_DIR) bv += beta_t * inv_dxdx * pressure(i, j + 1);
      b[row] = bv;
      x0[row] = pressure(i, j);  // 974
    }
  row_ptr[N] = cursor;
  return N;
}

// ---------------------------------------------------------------------------
// A10: CSR SpMV and residual, CSR_MATRIX.h:233-250, 278-298.
void orc_csr_spmv(int N, const int* row_ptr, const int* col, const double* val, const double* x, double* y) {
  for (int r = 0; r < N; ++r) {
    double s = 0;
    for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) s += val[k] * x[col[k]];
    y[r] = s;
  }
}

static double dot(int N, const double* a, const double* b) {  // VECTOR_ND.h:254-265
  double s = 0;
  for (int k = 0; k < N; ++k) s += a[k] * b[k];
  return s;
}

// Stopping rule selector shared by the Krylov restatements:
//   mode 0 = reference absolute rule (CG: r.r < tol^2, CG_METHOD.cpp:52;
//            PCG: r.z < tol^2, PCG_METHOD.cpp:253-254);
//   mode 1 = relative ||r||_2 <= tol * ||b||_2 (the configs' rule).
struct SolveStats {
  int iterations;
  double residual;  // final ||r||_2 (recursive)
};

// A12: CG_METHOD::CGMethod, CG_METHOD.cpp:13-73.  Returns the number of iterations
// run; the reference's own counter (num_iteration) stops one short of it when the
// loop ends by convergence, because the break comes before its increment (:52-56).
int orc_cg(int N, const int* row_ptr, const int* col, const double* val, double* x, const double* b, double tol,
           int max_iteration, int mode, double* out_res) {
  std::vector<double> r(N), p(N), Ap(N);
  orc_csr_spmv(N, row_ptr, col, val, x, Ap.data());
  for (int k = 0; k < N; ++k) r[k] = b[k] - Ap[k];
  p = r;
  const double thr = mode == 0 ? tol * tol : tol * tol * dot(N, b, b);
  double res_old = dot(N, r.data(), p.data()), res_new = res_old;
  int it = 0;
  // As for PCG: the reference tests only after an iteration (CG_METHOD.cpp:38-69).
  if (mode == 1 && res_old < thr) { if (out_res) *out_res = std::sqrt(res_old); return 0; }
  while (it < max_iteration) {
    orc_csr_spmv(N, row_ptr, col, val, p.data(), Ap.data());
    const double alpha = res_old / dot(N, p.data(), Ap.data());
    for (int k = 0; k < N; ++k) {
      x[k] += alpha * p[k];
      r[k] -= alpha * Ap[k];
    }
    res_new = dot(N, r.data(), r.data());
    if (res_new < thr) { ++it; break; }
    const double kk = res_new / res_old;
    for (int k = 0; k < N; ++k) p[k] = r[k] + kk * p[k];
    res_old = res_new;
    ++it;
  }
  if (out_res) *out_res = std::sqrt(res_new);
  return it;
}

// A13: IncompleteCholeskyDecomposition (CSR_MATRIX.h:535-628) for a grid whose
// cells are all full rows in i-fastest order (the Neumann box): IC(0) of the
// 5-point matrix (Meijerink-van der Vorst).  Row c of L holds, in this order,
// the bottom entry L_cb = (1/L_bb) A_bc (CSR_MATRIX.h:567-583), the left entry
// L_cl = (1/L_ll) A_lc (:585-601) and the diagonal L_cc = sqrt(A_cc - (L_cb^2 +
// L_cl^2)) summed in row order (:603-620); the reciprocal-then-multiply order is
// the reference's.  Stored per cell as (Lb, Ll, Ld).
static void ic0_factor(int nx, int ny, const int* row_ptr, const int* col, const double* val,
                       std::vector<double>& Lb, std::vector<double>& Ll, std::vector<double>& Ld) {
  const int N = nx * ny;
  Lb.assign(N, 0.0); Ll.assign(N, 0.0); Ld.assign(N, 0.0);
  auto A = [&](int r, int c) {  // CSR_MATRIX::operator() (search the row)
    for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
      if (col[k] == c) return val[k];
    return 0.0;
  };
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int c = i + j * nx;
      double sum = 0;
      if (j > 0) { Lb[c] = 1.0 / Ld[c - nx] * A(c - nx, c); sum += Lb[c] * Lb[c]; }
      if (i > 0) { Ll[c] = 1.0 / Ld[c - 1] * A(c - 1, c); sum += Ll[c] * Ll[c]; }
      Ld[c] = std::sqrt(A(c, c) - sum);
    }
}

// MultiplicationByMinverse, PCG_METHOD.cpp:70-118: solve L y = b row by row, then
// L^T x = y column by column; each row's sum runs in the CSR order of L (bottom,
// left), the column sums collect row c + nx before row c + 1, and both divide by
// multiplying with 1/L_cc, as the reference does.
static void ic0_apply(int nx, int ny, const std::vector<double>& Lb, const std::vector<double>& Ll,
                      const std::vector<double>& Ld, const double* b, double* x, std::vector<double>& y) {
  const int N = nx * ny;
  for (int c = 0; c < N; ++c) {
    const int i = c % nx;
    double s = 0;
    if (c >= nx) s += Lb[c] * y[c - nx];
    if (i > 0) s += Ll[c] * y[c - 1];
    y[c] = (b[c] - s) * (1.0 / Ld[c]);
  }
  for (int c = N - 1; c >= 0; --c) {
    const int i = c % nx;
    double s = 0;
    if (c + nx < N) s += Lb[c + nx] * x[c + nx];
    if (i + 1 < nx) s += Ll[c + 1] * x[c + 1];
    x[c] = (y[c] - s) * (1.0 / Ld[c]);
  }
}

// The IC(0) factor of a Neumann-box CSR matrix, per cell (test hook of ic0_factor).
void orc_ic0_factor(int nx, int ny, const int* row_ptr, const int* col, const double* val, double* Lb, double* Ll,
                    double* Ld) {
  std::vector<double> b, l, d;
  ic0_factor(nx, ny, row_ptr, col, val, b, l, d);
  std::copy(b.begin(), b.end(), Lb);
  std::copy(l.begin(), l.end(), Ll);
  std::copy(d.begin(), d.end(), Ld);
}

// z = (L L^T)^{-1} r with the IC(0) factor of a Neumann-box CSR matrix (test hook of ic0_apply).
void orc_ic0_solve(int nx, int ny, const int* row_ptr, const int* col, const double* val, const double* r, double* z) {
  std::vector<double> b, l, d, y((size_t)nx * ny);
  ic0_factor(nx, ny, row_ptr, col, val, b, l, d);
  ic0_apply(nx, ny, b, l, d, r, z, y);
}

// A13: PCG_METHOD::PCGMethod with IC(0), PCG_METHOD.cpp:176-264 (the solver the
// live projection path calls).  Requires every cell of the nx x ny box to be a row.
int orc_pcg_ic0(int nx, int ny, const int* row_ptr, const int* col, const double* val, double* x, const double* b,
                double tol, int max_iteration, int mode, double* out_res) {
  const int N = nx * ny;
  std::vector<double> Lb, Ll, Ld, y(N);
  ic0_factor(nx, ny, row_ptr, col, val, Lb, Ll, Ld);
  std::vector<double> r(N), s(N), p(N), Ap(N);
  orc_csr_spmv(N, row_ptr, col, val, x, Ap.data());
  for (int k = 0; k < N; ++k) r[k] = b[k] - Ap[k];
  ic0_apply(nx, ny, Lb, Ll, Ld, r.data(), p.data(), y);
  double res_new = dot(N, r.data(), p.data());
  const double bb = dot(N, b, b);
  int it = 0;
  auto done = [&]() {
    if (mode == 0) return res_new < tol * tol;
    return dot(N, r.data(), r.data()) <= tol * tol * bb;
  };
  // The reference's loop has no test before its first iteration (PCG_METHOD.cpp:206-263);
  // the configs' relative rule (mode 1) accepts a converged initial guess.
  if (mode == 1 && done()) { if (out_res) *out_res = std::sqrt(dot(N, r.data(), r.data())); return 0; }
  while (true) {
    orc_csr_spmv(N, row_ptr, col, val, p.data(), Ap.data());
    const double alpha = res_new / dot(N, p.data(), Ap.data());
    for (int k = 0; k < N; ++k) {
      x[k] += alpha * p[k];
      r[k] -= alpha * Ap[k];
    }
    ic0_apply(nx, ny, Lb, Ll, Ld, r.data(), s.data(), y);
    const double res_old = res_new;
    res_new = dot(N, r.data(), s.data());
    const double beta = res_new / res_old;
    for (int k = 0; k < N; ++k) p[k] = beta * p[k] + s[k];
    ++it;
    if (it > max_iteration) break;
    if (done()) break;
  }
  if (out_res) *out_res = std::sqrt(dot(N, r.data(), r.data()));
  return it;
}

// A13 (threaded variant): Jacobi-PCG, PCG_METHOD.cpp:266-383 with
// MultiplicationByMinverseAsDiagonal (36-68): z = r / diag(A).  The reference
// splits rows over a thread team with barriers; here OpenMP splits the same loops.
int orc_pcg_jacobi(int N, const int* row_ptr, const int* col, const double* val, double* x, const double* b,
                   double tol, int max_iteration, int mode, int threads, double* out_res) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  std::vector<double> r(N), s(N), p(N), Ap(N), dinv(N);
#pragma omp parallel for schedule(static)
  for (int rr = 0; rr < N; ++rr) {
    double d = 0;
    for (int k = row_ptr[rr]; k < row_ptr[rr + 1]; ++k)
      if (col[k] == rr) d = val[k];
    dinv[rr] = 1.0 / d;
  }
  auto spmv = [&](const double* xx, double* yy) {
#pragma omp parallel for schedule(static)
    for (int rr = 0; rr < N; ++rr) {
      double t = 0;
      for (int k = row_ptr[rr]; k < row_ptr[rr + 1]; ++k) t += val[k] * xx[col[k]];
      yy[rr] = t;
    }
  };
  auto pdot = [&](const double* a, const double* c) {
    double t = 0;
#pragma omp parallel for schedule(static) reduction(+ : t)
    for (int k = 0; k < N; ++k) t += a[k] * c[k];
    return t;
  };
  spmv(x, Ap.data());
#pragma omp parallel for schedule(static)
  for (int k = 0; k < N; ++k) { r[k] = b[k] - Ap[k]; p[k] = r[k] * dinv[k]; }
  double res_new = pdot(r.data(), p.data());
  const double bb = pdot(b, b);
  int it = 0;
  while (it < max_iteration) {
    if (mode == 1 && pdot(r.data(), r.data()) <= tol * tol * bb) break;
    spmv(p.data(), Ap.data());
    const double alpha = res_new / pdot(p.data(), Ap.data());
#pragma omp parallel for schedule(static)
    for (int k = 0; k < N; ++k) {
      x[k] += alpha * p[k];
      r[k] -= alpha * Ap[k];
      s[k] = r[k] * dinv[k];
    }
    const double res_old = res_new;
    res_new = pdot(r.data(), s.data());
    ++it;
    if (mode == 0 && res_new < tol * tol) break;
    const double beta = res_new / res_old;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < N; ++k) p[k] = beta * p[k] + s[k];
  }
  if (out_res) *out_res = std::sqrt(pdot(r.data(), r.data()));
  return it;
}

// ---------------------------------------------------------------------------
// Consistent (matrix-free) restatement used by this library.
//
// The face flux of the projection is q_f = beta_f (p_R - p_L + s_f a_f)/h with
// s_f = +1 if the R cell is on the minus side, -1 otherwise, on faces whose two
// cells lie on different sides (PROJECTION_2D.h:1345-1364 sign convention with
// the matrix's side test of POISSON_SOLVER.cpp:883-890), 0 on other faces.
// With D the cell divergence and A0 = -D B G (no-flux walls, no pin):
//   A0 p = -D u/dt + D(B J),  u_proj = u - dt q,  D u_proj = -dt (b - A0 p).
// DEVIATION: jump arms through a wall face (which the reference evaluates with
// ghost values) contribute nothing here; they vanish in every workload whose
// interface does not touch a wall or whose sigma is 0.

// y = A0 x (no-flux walls, no pin), h-scaled: entries beta/h^2.
void orc_apply_A0(int nx, int ny, double h, const double* bx, const double* by, const double* x, double* y) {
  const double ih2 = 1.0 / (h * h);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double xc = x[C(i, j, nx)];
      double s = 0;
      if (i > 0) s += bx[FX(i, j, nx)] * (xc - x[C(i - 1, j, nx)]);
      if (i + 1 < nx) s += bx[FX(i + 1, j, nx)] * (xc - x[C(i + 1, j, nx)]);
      if (j > 0) s += by[FY(i, j, nx)] * (xc - x[C(i, j - 1, nx)]);
      if (j + 1 < ny) s += by[FY(i, j + 1, nx)] * (xc - x[C(i, j + 1, nx)]);
      y[C(i, j, nx)] = s * ih2;
    }
}

// Face jump-gradient J_f = s_f a_f / h (0 off the interface and on walls).
static void jump_faces(int nx, int ny, double h, const double* phi, const double* jc, std::vector<double>& Jx,
                       std::vector<double>& Jy) {
  Jx.assign((size_t)(nx + 1) * ny, 0.0);
  Jy.assign((size_t)nx * (ny + 1), 0.0);
  if (!jc) return;
  for (int j = 0; j < ny; ++j)
    for (int i = 1; i < nx; ++i) {
      const double pL = phi[C(i - 1, j, nx)], pR = phi[C(i, j, nx)];
      if (minus_side(pL) == minus_side(pR)) continue;
      const double a = face_jump(jc[C(i - 1, j, nx)], jc[C(i, j, nx)], pL, pR);
      Jx[FX(i, j, nx)] = (minus_side(pR) ? a : -a) / h;
    }
  for (int j = 1; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double pL = phi[C(i, j - 1, nx)], pR = phi[C(i, j, nx)];
      if (minus_side(pL) == minus_side(pR)) continue;
      const double a = face_jump(jc[C(i, j - 1, nx)], jc[C(i, j, nx)], pL, pR);
      Jy[FY(i, j, nx)] = (minus_side(pR) ? a : -a) / h;
    }
}

// b = -D u/dt + D(B J)  (equals the reference's -div - F away from walls).
void orc_rhs_consistent(int nx, int ny, double h, double dt, const double* phi, const double* jc, const double* bx,
                        const double* by, const double* u, const double* v, double* b) {
  std::vector<double> Jx, Jy;
  jump_faces(nx, ny, h, phi, jc, Jx, Jy);
  const double ih = 1.0 / h, idt = 1.0 / dt;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double d = (ih * (u[FX(i + 1, j, nx)] - u[FX(i, j, nx)]) + ih * (v[FY(i, j + 1, nx)] - v[FY(i, j, nx)])) * idt;
      const double f = (bx[FX(i + 1, j, nx)] * Jx[FX(i + 1, j, nx)] - bx[FX(i, j, nx)] * Jx[FX(i, j, nx)] +
                        by[FY(i, j + 1, nx)] * Jy[FY(i, j + 1, nx)] - by[FY(i, j, nx)] * Jy[FY(i, j, nx)]) *
                       ih;
      b[C(i, j, nx)] = -d + f;
    }
}

// A14 (consistent form): u -= dt * beta_f (p_R - p_L + s a)/h on interior faces
// (wall faces untouched).  Returns max|u| and max|v| after the update
// (PROJECTION_2D.h:1365-1367 tracks them).
void orc_update_consistent(int nx, int ny, double h, double dt, const double* phi, const double* jc,
                           const double* bx, const double* by, const double* p, double* u, double* v,
                           double* max_u, double* max_v) {
  const double ih = 1.0 / h;
  double mu = 0, mv = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 1; i < nx; ++i) {
      const double pL = phi[C(i - 1, j, nx)], pR = phi[C(i, j, nx)];
      double gap = p[C(i, j, nx)] - p[C(i - 1, j, nx)];
      if (jc && minus_side(pL) != minus_side(pR)) {
        const double a = face_jump(jc[C(i - 1, j, nx)], jc[C(i, j, nx)], pL, pR);
        gap += minus_side(pR) ? a : -a;
      }
      double& uf = u[FX(i, j, nx)];
      uf -= dt * gap * bx[FX(i, j, nx)] * ih;
    }
  for (int j = 1; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double pL = phi[C(i, j - 1, nx)], pR = phi[C(i, j, nx)];
      double gap = p[C(i, j, nx)] - p[C(i, j - 1, nx)];
      if (jc && minus_side(pL) != minus_side(pR)) {
        const double a = face_jump(jc[C(i, j - 1, nx)], jc[C(i, j, nx)], pL, pR);
        gap += minus_side(pR) ? a : -a;
      }
      double& vf = v[FY(i, j, nx)];
      vf -= dt * gap * by[FY(i, j, nx)] * ih;
    }
  for (size_t k = 0; k < (size_t)(nx + 1) * ny; ++k) mu = std::max(mu, std::fabs(u[k]));
  for (size_t k = 0; k < (size_t)nx * (ny + 1); ++k) mv = std::max(mv, std::fabs(v[k]));
  if (max_u) *max_u = mu;
  if (max_v) *max_v = mv;
}

// A14 as the reference writes it (diagnostic only):
// UpdateVelocityByPressureGradientVariableDensity, PROJECTION_2D.h:1231-1417 with
// DetermineDensityField / HeavisideFunction (1419-1461, 1506-1531), eps = 1.5h,
// PI = 3.141592; faces whose right/top cell is a ghost are skipped, the left and
// bottom wall faces are updated against the never-written ghost pressure (0).
void orc_update_reference_literal(int nx, int ny, double h, double dt, const double* phi_in, const double* jc_in,
                                  double rho_minus, double rho_plus, const double* p_in, double* u, double* v) {
  const Pad phi = levelset_with_ghosts(nx, ny, phi_in);
  Pad jc(nx, ny), p(nx, ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      jc(i, j) = jc_in ? jc_in[C(i, j, nx)] : 0.0;
      p(i, j) = p_in[C(i, j, nx)];
    }
  const double eps = 1.5 * h, ih = 1.0 / h;
  auto heaviside = [&](double ph) {
    if (ph < -eps) return 0.0;
    if (ph > eps) return 1.0;
    const double oe = 1.0 / eps, op = 1.0 / REF_PI;
    return 0.5 + ph * oe * 0.5 + 0.5 * op * std::sin(REF_PI * ph * oe);
  };
  auto rho_cell = [&](int i, int j) { return phi(i, j) <= 0 ? rho_minus : rho_plus; };
  for (int pass = 0; pass < 2; ++pass) {
    const bool is_x = pass == 0;
    const int fi_end = is_x ? nx : nx - 1, fj_end = is_x ? ny - 1 : ny;
    for (int j = 0; j <= fj_end; ++j)
      for (int i = 0; i <= fi_end; ++i) {
        if (i >= nx || j >= ny) continue;  // bc(i,j) < 0: right/top wall faces skipped
        const int li = is_x ? i - 1 : i, lj = is_x ? j : j - 1;
        const double lsc = phi(i, j), lsn = phi(li, lj);
        const double half = 0.5 * (lsc + lsn);
        const double rho_half = rho_minus + (rho_plus - rho_minus) * heaviside(half);
        const double jcg = (jc(li, lj) * std::fabs(lsc) + jc(i, j) * std::fabs(lsn)) / (std::fabs(lsc) + std::fabs(lsn));
        const double pg = p(i, j) - p(li, lj);
        double& vel = is_x ? u[FX(i, j, nx)] : v[FY(i, j, nx)];
        if (lsc * lsn >= 0) vel -= dt * pg * (1.0 / rho_cell(i, j)) * ih;
        else if (lsc < 0) vel -= dt * (pg + jcg) * (1.0 / rho_half) * ih;
        else if (lsc > 0) vel -= dt * (pg - jcg) * (1.0 / rho_half) * ih;
      }
  }
}

// ---------------------------------------------------------------------------
// Geometric multigrid preconditioner (this library's design; the reference's
// MULTIGRID_METHOD is constant-coefficient and never wired into the projection).
// CPU mirror of the GPU V-cycle, operation for operation, for parity of M r.
//
// Level operator (h^2-scaled): (A' x)_c = sum_f beta_f (x_c - x_nb), singular on constants
// (no-flux walls, no pin on any level).
// The default mode (3, the one the GPU runs) has
//   coarse faces: per fine row or column, the series conductance between the two coarse
//                 centres, 2 / (0.5/b_{k-1} + 1/b_k + 0.5/b_{k+1}), averaged over the two
//                 fine rows that meet the coarse face (coarse_face);
//   P: tensor product of 1-D operator-dependent weights, w on the parent coarse cell and
//      1 - w on its neighbour, w = R_far / (R_near + R_far) from the fine faces' resistances
//      (od_weight; 3/4 for constant coefficients, 1 at a wall); R = P^T.
// Mode 0 (tests only) keeps the plain forms: beta_H = (beta_a + beta_b)/2 and cell-centred
// bilinear P with weights (9,3,3,1)/16.
// Mode bit 2 (mode 7, the GPU's FP32 hierarchy) stores every weight rounded to binary16
// through binary32 (orc_half_via_float), as the FP32 levels do; R = P^T still holds.
// Smoother: red-black Gauss-Seidel over-relaxed by omega on the finest level and by
// omega_coarse on the levels below it (SOR: x <- (1 - w) x + w x_GS, w = 1 plain Gauss-Seidel),
// pre (red, black) x nu, post (black, red) x nu; the reverse colour order after the coarse
// correction keeps M symmetric for every w in (0, 2).
// Bottom: pseudo-inverse of the coarsest operator (constants in its null space), so the
// V-cycle treats the constant mode exactly as the singular fine operator does.

constexpr int MG_DEFAULT_MODE = 3;
struct MGLevel {
  int nx, ny;
  std::vector<double> bx, by, x, r, res;
  std::vector<double> wx, wy;  // per fine cell: interpolation weight on the parent coarse cell (x / y)
};
struct MG {
  std::vector<MGLevel> L;
  int nu_pre, nu_post;
  int mode;                        // bit0: harmonic-normal coarse beta; bit1: operator-dependent P; bit2: binary16 weights
  double omega = 1.0;              // smoother over-relaxation on the finest level
  double omega_coarse = 1.0;       // ... on the coarser levels
  std::vector<double> bottom_inv;  // dense inverse of the coarsest operator
};

static double diag_at(const MGLevel& l, int i, int j) {
  const int nx = l.nx;
  return l.bx[FX(i, j, nx)] + l.bx[FX(i + 1, j, nx)] + l.by[FY(i, j, nx)] + l.by[FY(i, j + 1, nx)];
}
static double offsum(const MGLevel& l, const std::vector<double>& x, int i, int j) {
  const int nx = l.nx, ny = l.ny;
  double s = 0;
  if (i > 0) s += l.bx[FX(i, j, nx)] * x[C(i - 1, j, nx)];
  if (i + 1 < nx) s += l.bx[FX(i + 1, j, nx)] * x[C(i + 1, j, nx)];
  if (j > 0) s += l.by[FY(i, j, nx)] * x[C(i, j - 1, nx)];
  if (j + 1 < ny) s += l.by[FY(i, j + 1, nx)] * x[C(i, j + 1, nx)];
  return s;
}
static void rb_sweep(MGLevel& l, int color, double omega) {
  for (int j = 0; j < l.ny; ++j)
    for (int i = 0; i < l.nx; ++i) {
      if (((i + j) & 1) != color) continue;
      double& x = l.x[C(i, j, l.nx)];
      const double gs = (l.r[C(i, j, l.nx)] + offsum(l, l.x, i, j)) / diag_at(l, i, j);
      x = omega == 1.0 ? gs : omega * gs + (1.0 - omega) * x;
    }
}
static void residual(MGLevel& l) {
  for (int j = 0; j < l.ny; ++j)
    for (int i = 0; i < l.nx; ++i)
      l.res[C(i, j, l.nx)] = l.r[C(i, j, l.nx)] - (diag_at(l, i, j) * l.x[C(i, j, l.nx)] - offsum(l, l.x, i, j));
}
// Interpolation weights.  Fine cell i (1-D) has parent P = i/2 and neighbour
// Q = P +- 1 (towards the side of i inside its parent; clamped at walls).  The
// weight w on P (1-w on Q) is 3/4 for bilinear P, or, operator-dependent, the
// ratio of "resistances" along the fine row: with beta_k the fine face k and
// fine cell centres at integer positions (face k at k-1/2), the coarse centres
// sit on faces 2P+1 and 2Q+1, so
//   even i=2P: R_near = 0.5/beta_{2P+1}, R_far = 1/beta_{2P} + 0.5/beta_{2P-1}
//   odd  i=2P+1: R_near = 0.5/beta_{2P+1}, R_far = 1/beta_{2P+2} + 0.5/beta_{2P+3}
//   w = R_far / (R_near + R_far)      (= 3/4 for constant beta; 1 at a wall)
static double od_weight(const double* bface, int stride, int i, int nf) {
  // bface[k*stride] = beta on fine face k (k = 0..nf), walls 0.
  auto b = [&](int k) { return bface[(size_t)k * stride]; };
  const int P = i >> 1;
  double near = 0.5 / b(2 * P + 1 < nf ? 2 * P + 1 : nf);  // face 2P+1 is interior whenever the parent has 2 cells
  double far;
  if ((i & 1) == 0) {
    if (P == 0) return 1.0;
    far = 1.0 / b(2 * P) + 0.5 / b(2 * P - 1);
  } else {
    if (2 * P + 2 >= nf) return 1.0;
    far = 1.0 / b(2 * P + 2) + 0.5 / b(2 * P + 3);
  }
  if (!(near < 1e300) || !(far < 1e300)) return far >= near ? 1.0 : 0.0;
  return far / (near + far);
}
// w rounded to binary32, then to binary16 (round to nearest, ties to even; binary16
// subnormals below 2^-14), as a double.  For |w| <= 65504.
double orc_half_via_float(double w) {
  const float f = (float)w;
  if (f == 0.0f || !std::isfinite(f)) return f;
  int e;
  std::frexp(f, &e);                           // |f| in [2^(e-1), 2^e)
  const int q = std::max(e - 1, -14) - 10;     // the binary16 quantum of f's binade is 2^q
  return std::ldexp(std::nearbyint(std::ldexp((double)f, -q)), q);
}
static void build_weights(MGLevel& f, int op_dep, bool half) {
  f.wx.assign((size_t)f.nx * f.ny, 0.75);
  f.wy.assign((size_t)f.nx * f.ny, 0.75);
  for (int j = 0; j < f.ny; ++j)
    for (int i = 0; i < f.nx; ++i) {
      const int P = i >> 1, Q = P + ((i & 1) ? 1 : -1);
      const int Pj = j >> 1, Qj = Pj + ((j & 1) ? 1 : -1);
      double wx = 0.75, wy = 0.75;
      if (op_dep) {
        wx = od_weight(&f.bx[FX(0, j, f.nx)], 1, i, f.nx);
        wy = od_weight(&f.by[FY(i, 0, f.nx)], f.nx, j, f.ny);
      }
      if (Q < 0 || Q > (f.nx >> 1) - 1) wx = 1.0;
      if (Qj < 0 || Qj > (f.ny >> 1) - 1) wy = 1.0;
      f.wx[C(i, j, f.nx)] = half ? orc_half_via_float(wx) : wx;
      f.wy[C(i, j, f.nx)] = half ? orc_half_via_float(wy) : wy;
    }
}
// 1-D weight of coarse index I for fine index i given the parent weight w.
static inline double w1(int i, int I, int n, double w) {
  const int P = i >> 1;
  int Q = P + ((i & 1) ? 1 : -1);
  Q = Q < 0 ? 0 : (Q > n - 1 ? n - 1 : Q);
  double s = 0;
  if (P == I) s += w;
  if (Q == I) s += 1.0 - w;
  return s;
}
// r_H = P^T res (exact transpose of prolong_add).
static void restrict_(const MGLevel& f, MGLevel& c) {
  for (int J = 0; J < c.ny; ++J)
    for (int I = 0; I < c.nx; ++I) {
      double s = 0;
      for (int j = 2 * J - 1; j <= 2 * J + 2; ++j) {
        if (j < 0 || j >= f.ny) continue;
        for (int i = 2 * I - 1; i <= 2 * I + 2; ++i) {
          if (i < 0 || i >= f.nx) continue;
          const double wy = w1(j, J, c.ny, f.wy[C(i, j, f.nx)]);
          const double wx = w1(i, I, c.nx, f.wx[C(i, j, f.nx)]);
          if (wx == 0 || wy == 0) continue;
          s += wx * wy * f.res[C(i, j, f.nx)];
        }
      }
      c.r[C(I, J, c.nx)] = s;
    }
}
static void prolong_add(MGLevel& f, const MGLevel& c) {
  for (int j = 0; j < f.ny; ++j)
    for (int i = 0; i < f.nx; ++i) {
      const int I = i >> 1, J = j >> 1;
      int I2 = I + ((i & 1) ? 1 : -1), J2 = J + ((j & 1) ? 1 : -1);
      I2 = std::min(std::max(I2, 0), c.nx - 1);
      J2 = std::min(std::max(J2, 0), c.ny - 1);
      const double wx = f.wx[C(i, j, f.nx)], wy = f.wy[C(i, j, f.nx)];
      const double e = wx * wy * c.x[C(I, J, c.nx)] + (1.0 - wx) * wy * c.x[C(I2, J, c.nx)] +
                       wx * (1.0 - wy) * c.x[C(I, J2, c.nx)] + (1.0 - wx) * (1.0 - wy) * c.x[C(I2, J2, c.nx)];
      f.x[C(i, j, f.nx)] += e;
    }
}

static bool dense_inverse(int n, std::vector<double>& a /* n*n row-major, replaced by inverse */) {
  // Gauss-Jordan with partial pivoting.
  std::vector<double> inv((size_t)n * n, 0.0);
  for (int k = 0; k < n; ++k) inv[(size_t)k * n + k] = 1.0;
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int r = k + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + k]) > std::fabs(a[(size_t)piv * n + k])) piv = r;
    if (a[(size_t)piv * n + k] == 0) return false;
    if (piv != k)
      for (int c = 0; c < n; ++c) {
        std::swap(a[(size_t)k * n + c], a[(size_t)piv * n + c]);
        std::swap(inv[(size_t)k * n + c], inv[(size_t)piv * n + c]);
      }
    const double d = 1.0 / a[(size_t)k * n + k];
    for (int c = 0; c < n; ++c) { a[(size_t)k * n + c] *= d; inv[(size_t)k * n + c] *= d; }
    for (int r = 0; r < n; ++r) {
      if (r == k) continue;
      const double f = a[(size_t)r * n + k];
      if (f == 0) continue;
      for (int c = 0; c < n; ++c) {
        a[(size_t)r * n + c] -= f * a[(size_t)k * n + c];
        inv[(size_t)r * n + c] -= f * inv[(size_t)k * n + c];
      }
    }
  }
  a.swap(inv);
  return true;
}

// Number of levels: coarsen while both sizes are even and the grid has more
// than bottom_max cells, up to max_levels (if > 0).
int orc_mg_num_levels(int nx, int ny, int max_levels, int bottom_max) {
  int n = 1;
  while ((nx % 2 == 0) && (ny % 2 == 0) && (long)nx * ny > bottom_max && (max_levels <= 0 || n < max_levels)) {
    nx /= 2; ny /= 2; ++n;
  }
  return n;
}

// Coarse face coefficient from the fine faces (bit0 of mode selects the rule):
//   mode&1 == 0: beta_H = (beta_a + beta_b)/2, the two fine faces on the coarse face;
//   mode&1 == 1: per fine row, the series conductance between the two coarse
//                centres (0.5/b_{k-1} + 1/b_k + 0.5/b_{k+1})^-1 * 2, then averaged.
static double coarse_face(const double* b, int stride, int k, int nf) {
  auto bb = [&](int q) { return b[(size_t)q * stride]; };
  if (k <= 0 || k >= nf) return 0.0;
  const double r = 0.5 / bb(k - 1) + 1.0 / bb(k) + 0.5 / bb(k + 1);
  return 2.0 / r;
}

void* orc_mg_create(int nx, int ny, const double* bx, const double* by, int max_levels, int bottom_max, int nu_pre,
                    int nu_post, int mode, double omega, double omega_coarse) {
  if (nu_pre < 0 || nu_post < 0 || !(omega > 0.0 && omega < 2.0) || !(omega_coarse > 0.0 && omega_coarse < 2.0))
    return nullptr;
  MG* mg = new MG;
  mg->nu_pre = nu_pre;
  mg->nu_post = nu_post;
  mg->mode = mode;
  mg->omega = omega;
  mg->omega_coarse = omega_coarse;
  const int nl = orc_mg_num_levels(nx, ny, max_levels, bottom_max);
  mg->L.resize(nl);
  for (int l = 0; l < nl; ++l) {
    MGLevel& lv = mg->L[l];
    lv.nx = nx >> l; lv.ny = ny >> l;
    lv.bx.assign((size_t)(lv.nx + 1) * lv.ny, 0.0);
    lv.by.assign((size_t)lv.nx * (lv.ny + 1), 0.0);
    lv.x.assign((size_t)lv.nx * lv.ny, 0.0);
    lv.r = lv.x; lv.res = lv.x;
    if (l == 0) {
      std::memcpy(lv.bx.data(), bx, sizeof(double) * lv.bx.size());
      std::memcpy(lv.by.data(), by, sizeof(double) * lv.by.size());
    } else {
      MGLevel& f = mg->L[l - 1];
      build_weights(f, (mode >> 1) & 1, (mode >> 2) & 1);
      for (int J = 0; J < lv.ny; ++J)
        for (int I = 0; I <= lv.nx; ++I) {
          double v;
          if (mode & 1)
            v = 0.5 * (coarse_face(&f.bx[FX(0, 2 * J, f.nx)], 1, 2 * I, f.nx) +
                       coarse_face(&f.bx[FX(0, 2 * J + 1, f.nx)], 1, 2 * I, f.nx));
          else
            v = 0.5 * (f.bx[FX(2 * I, 2 * J, f.nx)] + f.bx[FX(2 * I, 2 * J + 1, f.nx)]);
          lv.bx[FX(I, J, lv.nx)] = v;
        }
      for (int J = 0; J <= lv.ny; ++J)
        for (int I = 0; I < lv.nx; ++I) {
          double v;
          if (mode & 1)
            v = 0.5 * (coarse_face(&f.by[FY(2 * I, 0, f.nx)], f.nx, 2 * J, f.ny) +
                       coarse_face(&f.by[FY(2 * I + 1, 0, f.nx)], f.nx, 2 * J, f.ny));
          else
            v = 0.5 * (f.by[FY(2 * I, 2 * J, f.nx)] + f.by[FY(2 * I + 1, 2 * J, f.nx)]);
          lv.by[FY(I, J, lv.nx)] = v;
        }
    }
  }
  // Bottom: pseudo-inverse of the singular (no-flux) coarsest operator,
  // A+ = (A + a 1 1^T)^-1 - 1 1^T / (a n^2), with a n = mean diagonal.
  MGLevel& b = mg->L.back();
  const int n = b.nx * b.ny;
  std::vector<double> A((size_t)n * n, 0.0);
  double dsum = 0;
  for (int j = 0; j < b.ny; ++j)
    for (int i = 0; i < b.nx; ++i) dsum += diag_at(b, i, j);
  const double alpha = dsum / n / n;
  for (size_t q = 0; q < A.size(); ++q) A[q] = alpha;
  for (int j = 0; j < b.ny; ++j)
    for (int i = 0; i < b.nx; ++i) {
      const int c = (int)C(i, j, b.nx);
      A[(size_t)c * n + c] += diag_at(b, i, j);
      if (i > 0) A[(size_t)c * n + c - 1] -= b.bx[FX(i, j, b.nx)];
      if (i + 1 < b.nx) A[(size_t)c * n + c + 1] -= b.bx[FX(i + 1, j, b.nx)];
      if (j > 0) A[(size_t)c * n + c - b.nx] -= b.by[FY(i, j, b.nx)];
      if (j + 1 < b.ny) A[(size_t)c * n + c + b.nx] -= b.by[FY(i, j + 1, b.nx)];
    }
  if (!dense_inverse(n, A)) { delete mg; return nullptr; }
  const double corr = 1.0 / (alpha * n * n);
  for (size_t q = 0; q < A.size(); ++q) A[q] -= corr;
  mg->bottom_inv.swap(A);
  return mg;
}

void orc_mg_destroy(void* h) { delete (MG*)h; }
int orc_mg_levels(void* h) { return (int)((MG*)h)->L.size(); }

static void vcycle(MG* mg, int l) {
  MGLevel& lv = mg->L[l];
  if (l + 1 == (int)mg->L.size()) {
    const int n = lv.nx * lv.ny;
    for (int a = 0; a < n; ++a) {
      double s = 0;
      for (int c = 0; c < n; ++c) s += mg->bottom_inv[(size_t)a * n + c] * lv.r[c];
      lv.x[a] = s;
    }
    return;
  }
  std::fill(lv.x.begin(), lv.x.end(), 0.0);
  const double om = l == 0 ? mg->omega : mg->omega_coarse;
  for (int s = 0; s < mg->nu_pre; ++s) { rb_sweep(lv, 0, om); rb_sweep(lv, 1, om); }
  residual(lv);
  MGLevel& cv = mg->L[l + 1];
  restrict_(lv, cv);
  vcycle(mg, l + 1);
  prolong_add(lv, cv);
  for (int s = 0; s < mg->nu_post; ++s) { rb_sweep(lv, 1, om); rb_sweep(lv, 0, om); }
}

// The hierarchy's pieces, for tests that rebuild the cycle from them: what = 0: y = A'_l x
// on level l; 1: y = P_l x (x on level l + 1, y on level l); 2: y = P_l^T x (the
// restriction); 3: y = the bottom pseudo-inverse times x (l ignored).  Returns -1 for an
// invalid level.
int orc_mg_piece(void* h, int what, int l, const double* x, double* y) {
  MG* mg = (MG*)h;
  const int L = (int)mg->L.size();
  if (what == 3) {
    MGLevel& b = mg->L.back();
    const int n = b.nx * b.ny;
    for (int a = 0; a < n; ++a) {
      double s = 0;
      for (int c = 0; c < n; ++c) s += mg->bottom_inv[(size_t)a * n + c] * x[c];
      y[a] = s;
    }
    return 0;
  }
  if (l < 0 || l >= L || (what != 0 && l + 1 >= L)) return -1;
  MGLevel& f = mg->L[l];
  const size_t nf = (size_t)f.nx * f.ny;
  if (what == 0) {
    const std::vector<double> xv(x, x + nf);
    for (int j = 0; j < f.ny; ++j)
      for (int i = 0; i < f.nx; ++i) y[C(i, j, f.nx)] = diag_at(f, i, j) * xv[C(i, j, f.nx)] - offsum(f, xv, i, j);
    return 0;
  }
  MGLevel& c = mg->L[l + 1];
  const size_t nc = (size_t)c.nx * c.ny;
  if (what == 1) {
    std::vector<double> sx = f.x;
    std::fill(f.x.begin(), f.x.end(), 0.0);
    std::vector<double> cx = c.x;
    std::memcpy(c.x.data(), x, nc * sizeof(double));
    prolong_add(f, c);
    std::memcpy(y, f.x.data(), nf * sizeof(double));
    f.x.swap(sx);
    c.x.swap(cx);
    return 0;
  }
  if (what == 2) {
    std::vector<double> sres = f.res, cr = c.r;
    std::memcpy(f.res.data(), x, nf * sizeof(double));
    restrict_(f, c);
    std::memcpy(y, c.r.data(), nc * sizeof(double));
    f.res.swap(sres);
    c.r.swap(cr);
    return 0;
  }
  return -1;
}

// z = M r_scaled: one V-cycle on A' x = r' (r' = h^2 r for the fine operator).
void orc_mg_vcycle(void* h, const double* rs, double* z) {
  MG* mg = (MG*)h;
  MGLevel& f = mg->L[0];
  std::memcpy(f.r.data(), rs, sizeof(double) * f.r.size());
  vcycle(mg, 0);
  std::memcpy(z, f.x.data(), sizeof(double) * f.x.size());
}

// MG-preconditioned CG on the consistent operator A0 (singular, no-flux walls),
// x updated in place, stops when ||r||_2 <= rtol ||b||_2 and ||r||_inf <= atol
// (either test disabled by a non-positive tolerance).  Returns iterations.
int orc_mg_pcg(int nx, int ny, double h, const double* bx, const double* by, double* x, const double* b, double rtol,
               double atol, int max_iteration, int max_levels, int nu, int mode, double omega, double omega_coarse,
               double* out_res2, double* out_resinf) {
  const int N = nx * ny;
  std::vector<double> bxs(bx, bx + (size_t)(nx + 1) * ny), bys(by, by + (size_t)nx * (ny + 1));
  void* mg = orc_mg_create(nx, ny, bxs.data(), bys.data(), max_levels, 64, nu, nu, mode, omega, omega_coarse);
  if (!mg) return -1;
  std::vector<double> r(N), z(N), p(N), q(N), rs(N), bm(b, b + N);
  // A0 is singular (constants); solve with the range component of b, i.e. b minus its mean.
  double mean = 0;
  for (int k = 0; k < N; ++k) mean += b[k];
  mean /= N;
  for (int k = 0; k < N; ++k) bm[k] -= mean;
  orc_apply_A0(nx, ny, h, bx, by, x, q.data());
  for (int k = 0; k < N; ++k) r[k] = bm[k] - q[k];
  const double bn = std::sqrt(dot(N, bm.data(), bm.data()));
  auto rinf = [&]() { double m = 0; for (int k = 0; k < N; ++k) m = std::max(m, std::fabs(r[k])); return m; };
  auto conv = [&]() {
    const bool ok2 = rtol <= 0 || std::sqrt(dot(N, r.data(), r.data())) <= rtol * bn;
    const bool okinf = atol <= 0 || rinf() <= atol;
    return ok2 && okinf;
  };
  int it = 0;
  double rz_old = 0;
  while (!conv() && it < max_iteration) {
    for (int k = 0; k < N; ++k) rs[k] = h * h * r[k];
    orc_mg_vcycle(mg, rs.data(), z.data());
    double zm = 0;  // drop the preconditioner's constant (null-space) component
    for (int k = 0; k < N; ++k) zm += z[k];
    zm /= N;
    for (int k = 0; k < N; ++k) z[k] -= zm;
    const double rz = dot(N, r.data(), z.data());
    if (it == 0) p = z;
    else {
      const double beta = rz / rz_old;
      for (int k = 0; k < N; ++k) p[k] = z[k] + beta * p[k];
    }
    rz_old = rz;
    orc_apply_A0(nx, ny, h, bx, by, p.data(), q.data());
    const double alpha = rz / dot(N, p.data(), q.data());
    for (int k = 0; k < N; ++k) { x[k] += alpha * p[k]; r[k] -= alpha * q[k]; }
    ++it;
  }
  if (out_res2) *out_res2 = std::sqrt(dot(N, r.data(), r.data()));
  if (out_resinf) *out_resinf = rinf();
  orc_mg_destroy(mg);
  return it;
}

// ---------------------------------------------------------------------------
// A1: the whole projection, PROJECTION_2D::Solve(dt), PROJECTION_2D.h:235-242:
// density -> divergence -> jump (sigma*kappa) -> pressure (reference-literal CSR
// with the pin, solved by `solver`) -> consistent face update.
//   solver: 0 = IC(0)-PCG (live path), 1 = CG, 2 = Jacobi-PCG (threads), 3 = MG-PCG on A0.
//   tol: relative ||r||_2/||b||_2 (mode 1).  p_out: pressure as solved (pinned),
//   u, v updated in place.  stats[0..3] = iterations, ||r||_2, max|u|, max|v|.
int orc_project(int nx, int ny, double h, double dt, const double* phi, double rho_minus, double rho_plus,
                double sigma, double* u, double* v, int solver, double tol, int max_iteration, int threads,
                double* p_out, double* stats) {
  const size_t N = (size_t)nx * ny;
  std::vector<double> inv_rho(N), div(N), kappa(N), jc(N);
  orc_density(nx, ny, phi, rho_minus, rho_plus, nullptr, inv_rho.data());
  orc_divergence(nx, ny, h, dt, u, v, div.data());
  if (orc_curvature(nx, ny, h, phi, kappa.data()) != 0) return -1;
  for (size_t k = 0; k < N; ++k) jc[k] = sigma * kappa[k];
  std::vector<double> bx((size_t)(nx + 1) * ny), by((size_t)nx * (ny + 1));
  orc_face_beta(nx, ny, phi, rho_minus, rho_plus, bx.data(), by.data());
  std::vector<double> x(N, 0.0);
  int iters = 0;
  double res = 0;
  if (solver == 3) {
    std::vector<double> b(N);
    orc_rhs_consistent(nx, ny, h, dt, phi, jc.data(), bx.data(), by.data(), u, v, b.data());
    iters = orc_mg_pcg(nx, ny, h, bx.data(), by.data(), x.data(), b.data(), tol, 0.0, max_iteration, 0, 2, MG_DEFAULT_MODE,
                       1.0, 1.0, &res, nullptr);
    const double x0 = x[0];
    for (size_t k = 0; k < N; ++k) x[k] -= x0;  // report with the reference's pin p(0,0) = 0
  } else {
    std::vector<int> row_ptr(N + 1), col(5 * N), row_of(N);
    std::vector<double> val(5 * N), b(N), x0(N);
    const int rows = orc_assemble_reference(nx, ny, h, nullptr, nullptr, phi, inv_rho.data(), div.data(), jc.data(),
                                            nullptr, nullptr, nullptr, row_ptr.data(), col.data(), val.data(),
                                            b.data(), x0.data(), row_of.data());
    std::vector<double> xs(rows, 0.0);
    if (solver == 0) iters = orc_pcg_ic0(nx, ny, row_ptr.data(), col.data(), val.data(), xs.data(), b.data(), tol, max_iteration, 1, &res);
    else if (solver == 1) iters = orc_cg(rows, row_ptr.data(), col.data(), val.data(), xs.data(), b.data(), tol, max_iteration, 1, &res);
    else iters = orc_pcg_jacobi(rows, row_ptr.data(), col.data(), val.data(), xs.data(), b.data(), tol, max_iteration, 1, threads, &res);
    for (size_t k = 0; k < N; ++k) x[k] = row_of[k] >= 0 ? xs[row_of[k]] : 0.0;  // VectorToGrid
  }
  double mu = 0, mv = 0;
  orc_update_consistent(nx, ny, h, dt, phi, jc.data(), bx.data(), by.data(), x.data(), u, v, &mu, &mv);
  if (p_out) std::memcpy(p_out, x.data(), sizeof(double) * N);
  if (stats) { stats[0] = iters; stats[1] = res; stats[2] = mu; stats[3] = mv; }
  return 0;
}

// ---------------------------------------------------------------------------
// Adjoint of the consistent projection (no counterpart in the reference).
// Forward: A0 p = -D u/dt + D(B J), u_proj = u - dt B (G p + J).
// Given g = dL/du_proj on interior faces:
//   A0 mu = -D(B g)                          (one extra symmetric solve)
//   dL/du    = g - G mu                      (interior faces; walls 0)
//   dL/dbeta = -dt (g - G mu) * (G p + J)    (interior faces; walls 0)
// p is the forward pressure (the only checkpoint).  The solve uses MG-PCG to
// ||r||_2 <= rtol ||rhs||_2.
int orc_adjoint(int nx, int ny, double h, double dt, const double* phi, const double* jc, const double* bx,
                const double* by, const double* p, const double* gu, const double* gv, double rtol, double* du,
                double* dv, double* dbx, double* dby, double* mu_out) {
  const size_t N = (size_t)nx * ny;
  std::vector<double> rhs(N), mu(N, 0.0), Jx, Jy;
  const double ih = 1.0 / h;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      auto gxf = [&](int ii) { return (ii > 0 && ii < nx) ? bx[FX(ii, j, nx)] * gu[FX(ii, j, nx)] : 0.0; };
      auto gyf = [&](int jj) { return (jj > 0 && jj < ny) ? by[FY(i, jj, nx)] * gv[FY(i, jj, nx)] : 0.0; };
      rhs[C(i, j, nx)] = -((gxf(i + 1) - gxf(i)) * ih + (gyf(j + 1) - gyf(j)) * ih);
    }
  const int it =
      orc_mg_pcg(nx, ny, h, bx, by, mu.data(), rhs.data(), rtol, 0.0, 100000, 0, 2, MG_DEFAULT_MODE, 1.0, 1.0, nullptr, nullptr);
  if (it < 0) return -1;
  jump_faces(nx, ny, h, phi, jc, Jx, Jy);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const size_t f = FX(i, j, nx);
      if (i == 0 || i == nx) { du[f] = 0; dbx[f] = 0; continue; }
      const double t = gu[f] - (mu[C(i, j, nx)] - mu[C(i - 1, j, nx)]) * ih;
      const double w = (p[C(i, j, nx)] - p[C(i - 1, j, nx)]) * ih + Jx[f];
      du[f] = t;
      dbx[f] = -dt * t * w;
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t f = FY(i, j, nx);
      if (j == 0 || j == ny) { dv[f] = 0; dby[f] = 0; continue; }
      const double t = gv[f] - (mu[C(i, j, nx)] - mu[C(i, j - 1, nx)]) * ih;
      const double w = (p[C(i, j, nx)] - p[C(i, j - 1, nx)]) * ih + Jy[f];
      dv[f] = t;
      dby[f] = -dt * t * w;
    }
  if (mu_out) std::memcpy(mu_out, mu.data(), sizeof(double) * N);
  return it;
}

// Forward of the consistent projection with explicit face coefficients (for the
// adjoint's finite-difference checks): u,v in place, p out (mean-zero).
int orc_project_beta(int nx, int ny, double h, double dt, const double* phi, const double* jc, const double* bx,
                     const double* by, double* u, double* v, double rtol, double* p_out) {
  const size_t N = (size_t)nx * ny;
  std::vector<double> b(N), x(N, 0.0);
  orc_rhs_consistent(nx, ny, h, dt, phi, jc, bx, by, u, v, b.data());
  const int it =
      orc_mg_pcg(nx, ny, h, bx, by, x.data(), b.data(), rtol, 0.0, 100000, 0, 2, MG_DEFAULT_MODE, 1.0, 1.0, nullptr, nullptr);
  if (it < 0) return -1;
  double m = 0;
  for (size_t k = 0; k < N; ++k) m += x[k];
  m /= (double)N;
  for (size_t k = 0; k < N; ++k) x[k] -= m;
  orc_update_consistent(nx, ny, h, dt, phi, jc, bx, by, x.data(), u, v, nullptr, nullptr);
  if (p_out) std::memcpy(p_out, x.data(), sizeof(double) * N);
  return it;
}

// ---------------------------------------------------------------------------
// Level-set re-initialisation, the sussmanstyle_reinitialization branch of
// EULERIAN_FLUID_SOLVER_2D::Reinitialization (EULERIAN_FLUID_SOLVER_2D.h:2365-2400):
//   s = phi0 / sqrt(phi0^2 + dx^2) from the level set at entry (:2380-2383),
//   fictitious dt = dx / scaling_number (:2374), then `iterations` calls of
//   ADVECTION_2D::ReinitializationBySussman (ADVECTION_2D.cpp:415-421) =
//   ADVECTION_METHOD_2D::WENO5thReinitialization (ADVECTION_METHOD_2D.h:805-1076).
// One step reads a ghost copy g of phi (FIELD_STRUCTURE_2D::FillGhostCellsFrom,
// FIELD_STRUCTURE_2D.h:528-582: all three ghosts of a face row take the second
// interior value, e.g. g(-3..-1, j) = phi(1, j); corners are never read), forms
// the WENO5 one-sided derivatives, applies the reference's Godunov branch
// selection, updates phi -= dt s (H - 1), and clamps a sign flip to
// +-epsilon.  The arithmetic keeps the reference's expression forms (e.g.
// v*7/6 is (v*7)/6) so that a build without contraction reproduces it.
namespace {
struct Pad3 {
  int nx, ny;
  std::vector<double> a;
  Pad3(int nx_, int ny_) : nx(nx_), ny(ny_), a((size_t)(nx_ + 6) * (ny_ + 6), 0.0) {}
  double& operator()(int i, int j) { return a[(size_t)(i + 3) + (size_t)(j + 3) * (nx + 6)]; }
  double operator()(int i, int j) const { return a[(size_t)(i + 3) + (size_t)(j + 3) * (nx + 6)]; }
};
Pad3 reinit_ghosts(int nx, int ny, const double* phi) {
  Pad3 g(nx, ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) g(i, j) = phi[C(i, j, nx)];
  for (int j = 0; j < ny; ++j)
    for (int k = 1; k <= 3; ++k) {
      g(-k, j) = phi[C(1, j, nx)];
      g(nx - 1 + k, j) = phi[C(nx - 2, j, nx)];
    }
  for (int i = 0; i < nx; ++i)
    for (int k = 1; k <= 3; ++k) {
      g(i, ny - 1 + k) = phi[C(i, ny - 2, nx)];
      g(i, -k) = phi[C(i, 1, nx)];
    }
  return g;
}
// WENO5 derivative from the five one-sided differences v1..v5 (ADVECTION_METHOD_2D.h:844-929).
double weno5(double v1, double v2, double v3, double v4, double v5, double epsilon) {
  const double s1 = (double)13 / 12 * (v1 - 2 * v2 + v3) * (v1 - 2 * v2 + v3) +
                    (double)1 / 4 * (v1 - 4 * v2 + 3 * v3) * (v1 - 4 * v2 + 3 * v3);
  const double s2 = (double)13 / 12 * (v2 - 2 * v3 + v4) * (v2 - 2 * v3 + v4) + (double)1 / 4 * (v2 - v4) * (v2 - v4);
  const double s3 = (double)13 / 12 * (v3 - 2 * v4 + v5) * (v3 - 2 * v4 + v5) +
                    (double)1 / 4 * (3 * v3 - 4 * v4 + v5) * (3 * v3 - 4 * v4 + v5);
  const double a1 = (double)1 / 10 * (double)1 / ((epsilon + s1) * (epsilon + s1));
  const double a2 = (double)6 / 10 * (double)1 / ((epsilon + s2) * (epsilon + s2));
  const double a3 = (double)3 / 10 * (double)1 / ((epsilon + s3) * (epsilon + s3));
  const double sum = a1 + a2 + a3;
  const double w1 = a1 / sum, w2 = a2 / sum, w3 = a3 / sum;
  return w1 * (v1 * (double)1 / 3 - v2 * (double)7 / 6 + v3 * (double)11 / 6) +
         w2 * (v2 * (-(double)1 / 6) + v3 * (double)5 / 6 + v4 * (double)1 / 3) +
         w3 * (v3 * (double)1 / 3 + v4 * (double)5 / 6 - v5 * (double)1 / 6);
}
// phi_x^-, phi_x^+, phi_y^-, phi_y^+ at (i, j) of the ghost copy (ADVECTION_METHOD_2D.h:817-841).
void weno5_derivs(const Pad3& g, int i, int j, double ih, double epsilon, double* d) {
  const double mx1 = ih * (g(i - 2, j) - g(i - 3, j)), mx2 = ih * (g(i - 1, j) - g(i - 2, j)),
               mx3 = ih * (g(i, j) - g(i - 1, j)), mx4 = ih * (g(i + 1, j) - g(i, j)),
               mx5 = ih * (g(i + 2, j) - g(i + 1, j));
  const double px1 = ih * (g(i + 3, j) - g(i + 2, j)), px2 = ih * (g(i + 2, j) - g(i + 1, j)),
               px3 = ih * (g(i + 1, j) - g(i, j)), px4 = ih * (g(i, j) - g(i - 1, j)),
               px5 = ih * (g(i - 1, j) - g(i - 2, j));
  const double my1 = ih * (g(i, j - 2) - g(i, j - 3)), my2 = ih * (g(i, j - 1) - g(i, j - 2)),
               my3 = ih * (g(i, j) - g(i, j - 1)), my4 = ih * (g(i, j + 1) - g(i, j)),
               my5 = ih * (g(i, j + 2) - g(i, j + 1));
  const double py1 = ih * (g(i, j + 3) - g(i, j + 2)), py2 = ih * (g(i, j + 2) - g(i, j + 1)),
               py3 = ih * (g(i, j + 1) - g(i, j)), py4 = ih * (g(i, j) - g(i, j - 1)),
               py5 = ih * (g(i, j - 1) - g(i, j - 2));
  d[0] = weno5(mx1, mx2, mx3, mx4, mx5, epsilon);
  d[1] = weno5(px1, px2, px3, px4, px5, epsilon);
  d[2] = weno5(my1, my2, my3, my4, my5, epsilon);
  d[3] = weno5(py1, py2, py3, py4, py5, epsilon);
}
// The reference's Godunov branch selection (ADVECTION_METHOD_2D.h:936-1060): the
// Hamiltonian |grad phi| from the one-sided derivatives and the sign s; NaN
// when no branch applies (then the reference leaves phi unchanged).
double godunov_h(double s, double mxd, double pxd, double myd, double pyd) {
  const double smx = s * mxd, spx = s * pxd, smy = s * myd, spy = s * pyd;
  // the y part given the x derivative used (xd) and whether x enters at all
  auto ypart = [&](bool hasx, double xd) -> double {
    const double x2 = hasx ? xd * xd : 0.0;
    if ((spy >= 0) && (smy >= 0)) return hasx ? std::sqrt(x2 + myd * myd) : std::sqrt(myd * myd);
    if ((spy <= 0) && (smy <= 0)) return hasx ? std::sqrt(x2 + pyd * pyd) : std::sqrt(pyd * pyd);
    if ((spy > 0) && (smy < 0)) return hasx ? std::sqrt(x2) : 0.0;
    if ((spy < 0) && (smy > 0)) {
      const double ss = s * (std::fabs(pyd) - std::fabs(myd)) / (pyd - myd);
      if (ss > 0) return hasx ? std::sqrt(x2 + myd * myd) : std::sqrt(myd * myd);
      return hasx ? std::sqrt(x2 + pyd * pyd) : std::sqrt(pyd * pyd);
    }
    return NAN;
  };
  if ((spx >= 0) && (smx >= 0)) return ypart(true, mxd);
  if ((spx <= 0) && (smx <= 0)) return ypart(true, pxd);
  if ((spx > 0) && (smx < 0)) return ypart(false, 0.0);
  if ((spx < 0) && (smx > 0)) {
    const double sss = s * (std::fabs(pxd) - std::fabs(mxd)) / (pxd - mxd);
    return sss > 0 ? ypart(true, mxd) : ypart(true, pxd);
  }
  return NAN;
}
}  // namespace

void orc_weno5_derivatives(int nx, int ny, double h, const double* phi, double epsilon, double* dxm, double* dxp,
                           double* dym, double* dyp) {
  const Pad3 g = reinit_ghosts(nx, ny, phi);
  const double ih = 1.0 / h;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double d[4];
      weno5_derivs(g, i, j, ih, epsilon, d);
      const size_t k = C(i, j, nx);
      dxm[k] = d[0];
      dxp[k] = d[1];
      dym[k] = d[2];
      dyp[k] = d[3];
    }
}

void orc_reinit_sussman(int nx, int ny, double h, const double* phi_in, int iterations, double scaling,
                        double epsilon, double* phi_out) {
  const size_t N = (size_t)nx * ny;
  std::vector<double> sgn(N), phi(phi_in, phi_in + N);
  for (size_t k = 0; k < N; ++k) sgn[k] = phi[k] / std::sqrt(phi[k] * phi[k] + h * h);
  const double dt = h / scaling, ih = 1.0 / h;
  for (int it = 0; it < iterations; ++it) {
    const Pad3 g = reinit_ghosts(nx, ny, phi.data());
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t k = C(i, j, nx);
        double d[4];
        weno5_derivs(g, i, j, ih, epsilon, d);
        const double s = sgn[k];
        const double H = godunov_h(s, d[0], d[1], d[2], d[3]);
        if (H == H) phi[k] = phi[k] - dt * s * (H - 1);
        if (s * phi[k] < 0) {
          if (s < 0) phi[k] = -epsilon;
          if (s > 0) phi[k] = epsilon;
        }
      }
  }
  std::copy(phi.begin(), phi.end(), phi_out);
}

// Level-set advection, one forward-Euler stage: ADVECTION_2D::Solve_Levelset
// (ADVECTION_2D.cpp:165-200, use_mac_grid + use_5th_weno) =
// ADVECTION_METHOD_2D::WENO5th on a scalar field with MAC velocities
// (ADVECTION_METHOD_2D.h:508-618): the ghost copy and WENO5 derivatives of the
// re-initialisation, the cell velocity as the mean of its two faces,
// u = (u(i+1,j) + u(i,j)) * 0.5, v = (v(i,j+1) + v(i,j)) * 0.5, and the upwind
// side by the sign of each component (u > 0: phi_x^-, else phi_x^+; same in y):
// phi -= dt (u phi_x + v phi_y).  u is (nx+1) x ny, v nx x (ny+1).
void orc_advect_levelset(int nx, int ny, double h, double dt, const double* phi, const double* u, const double* v,
                         double epsilon, double* out) {
  const Pad3 g = reinit_ghosts(nx, ny, phi);
  const double ih = 1.0 / h;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double d[4];
      weno5_derivs(g, i, j, ih, epsilon, d);
      const double uv = (u[FX(i + 1, j, nx)] + u[FX(i, j, nx)]) * 0.5;
      const double vv = (v[FY(i, j + 1, nx)] + v[FY(i, j, nx)]) * 0.5;
      const double dx = uv > 0 ? d[0] : d[1], dy = vv > 0 ? d[2] : d[3];
      out[C(i, j, nx)] = phi[C(i, j, nx)] - dt * (uv * dx + vv * dy);
    }
}

// Velocity advection, one forward-Euler stage: ADVECTION_2D::Solve_Velocity
// (ADVECTION_2D.cpp:95-160, use_3rd_eno_v, the scripts' choice) =
// ADVECTION_METHOD_2D::ENO3rd on each MAC component (ADVECTION_METHOD_2D.h:2038-2232).
// Both components are advected by the old velocities.  Each face field
// (u: (nx+1) x ny, v: nx x (ny+1)) gets a 3-cell ghost copy by the rule of
// FillGhostCellsFrom (second interior value, FIELD_STRUCTURE_2D.h:528-582); the
// one-sided derivatives are third-order ENO (divided differences 1/dx, 1/(2dx),
// (1/dx)/3, choice by the smaller |difference|); the advecting velocity at an
// x-face is (u, mean of the four surrounding y-faces), at a y-face (mean of the
// four surrounding x-faces, v), reading the other component's ghost copy; the
// upwind side follows each component's sign.  Every face is updated, walls too.
namespace {
struct PadG {  // a field of fnx x fny values with a 3-cell ghost ring
  int nx, ny;
  std::vector<double> a;
  PadG(int nx_, int ny_) : nx(nx_), ny(ny_), a((size_t)(nx_ + 6) * (ny_ + 6), 0.0) {}
  double& operator()(int i, int j) { return a[(size_t)(i + 3) + (size_t)(j + 3) * (nx + 6)]; }
  double operator()(int i, int j) const { return a[(size_t)(i + 3) + (size_t)(j + 3) * (nx + 6)]; }
};
PadG face_ghosts(int fnx, int fny, const double* f) {
  PadG g(fnx, fny);
  for (int j = 0; j < fny; ++j)
    for (int i = 0; i < fnx; ++i) g(i, j) = f[(size_t)i + (size_t)j * fnx];
  for (int j = 0; j < fny; ++j)
    for (int k = 1; k <= 3; ++k) {
      g(-k, j) = f[(size_t)1 + (size_t)j * fnx];
      g(fnx - 1 + k, j) = f[(size_t)(fnx - 2) + (size_t)j * fnx];
    }
  for (int i = 0; i < fnx; ++i)
    for (int k = 1; k <= 3; ++k) {
      g(i, fny - 1 + k) = f[(size_t)i + (size_t)(fny - 2) * fnx];
      g(i, -k) = f[(size_t)i + (size_t)1 * fnx];
    }
  return g;
}
// ENO3 one-sided derivatives (minus, plus) from six consecutive values
// f(-3..3) of a line, with spacing dx (ADVECTION_METHOD_2D.h:2054-2126).
void eno3(const double* f /* f[0] = value at -3 ... f[6] at +3 */, double dx, double* m, double* p) {
  const double o1 = 1.0 / dx, o2 = 1.0 / (2 * dx), o3 = o1 * 1 / 3;
  const double d1n3 = o1 * (f[1] - f[0]), d1n2 = o1 * (f[2] - f[1]), d1n1 = o1 * (f[3] - f[2]);
  const double d10 = o1 * (f[4] - f[3]), d1p1 = o1 * (f[5] - f[4]), d1p2 = o1 * (f[6] - f[5]);
  const double d2n2 = o2 * (d1n2 - d1n3), d2n1 = o2 * (d1n1 - d1n2), d20 = o2 * (d10 - d1n1);
  const double d2p1 = o2 * (d1p1 - d10), d2p2 = o2 * (d1p2 - d1p1);
  const double d3n2 = o3 * (d2n1 - d2n2), d3n1 = o3 * (d20 - d2n1), d30 = o3 * (d2p1 - d20), d3p1 = o3 * (d2p2 - d2p1);
  if (std::fabs(d2n1) <= std::fabs(d20)) {
    if (std::fabs(d3n2) <= std::fabs(d3n1)) *m = d1n1 + d2n1 * dx + 2 * d3n2 * dx * dx;
    else *m = d1n1 + d2n1 * dx + 2 * d3n1 * dx * dx;
  } else {
    if (std::fabs(d3n1) <= std::fabs(d30)) *m = d1n1 + d20 * dx - d3n1 * dx * dx;
    else *m = d1n1 + d20 * dx - d30 * dx * dx;
  }
  if (std::fabs(d20) <= std::fabs(d2p1)) {
    if (std::fabs(d3n1) <= std::fabs(d30)) *p = d10 - d20 * dx - d3n1 * dx * dx;
    else *p = d10 - d20 * dx - d30 * dx * dx;
  } else {
    if (std::fabs(d30) <= std::fabs(d3p1)) *p = d10 - d2p1 * dx + 2 * d30 * dx * dx;
    else *p = d10 - d2p1 * dx + 2 * d3p1 * dx * dx;
  }
}
void eno3_advect_component(const PadG& g, const PadG& gu, const PadG& gv, bool xcomp, double h, double dt,
                           double* out) {
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      double lx[7], ly[7], mx, px, my, py;
      for (int k = 0; k < 7; ++k) {
        lx[k] = g(i - 3 + k, j);
        ly[k] = g(i, j - 3 + k);
      }
      eno3(lx, h, &mx, &px);
      eno3(ly, h, &my, &py);
      double uv, vv;
      if (xcomp) {
        uv = gu(i, j);
        vv = (gv(i, j) + gv(i, j + 1) + gv(i - 1, j) + gv(i - 1, j + 1)) * 0.25;
      } else {
        uv = (gu(i, j - 1) + gu(i + 1, j - 1) + gu(i, j) + gu(i + 1, j)) * 0.25;
        vv = gv(i, j);
      }
      const double dx = uv > 0 ? mx : px, dy = vv > 0 ? my : py;
      out[(size_t)i + (size_t)j * g.nx] = g(i, j) - dt * (uv * dx + vv * dy);
    }
}
}  // namespace

void orc_advect_velocity(int nx, int ny, double h, double dt, const double* u, const double* v, double* u_out,
                         double* v_out) {
  const PadG gu = face_ghosts(nx + 1, ny, u), gv = face_ghosts(nx, ny + 1, v);
  eno3_advect_component(gu, gu, gv, true, h, dt, u_out);
  eno3_advect_component(gv, gu, gv, false, h, dt, v_out);
}

// Viscosity, air-water delta-function formulation: VISCOSITY_2D::ApplyViscosity
// (VISCOSITY_2D.h:460-548, the threaded form the solver calls at
// EULERIAN_FLUID_SOLVER_2D.h:2086-2089).  An explicit update of both MAC
// components from the old velocities (temp copies whose ghosts are 0): at an
// x-face, level-set averages at the face, its upper/lower corners and the two
// cells; density and viscosities from the mollified Heaviside
// (HeavisideFunction VISCOSITY_2D.h:688-707, PI = 3.141592, width epsilon;
// value = minus + (plus - minus) H, VISCOSITY_2D.h:750-790, 834-871); then
//   u += 2 coef/dx (mu_r/dx (u_E - u) - mu_c/dx (u - u_W))
//      + coef/dy (mu_cu (du/dy_N + dv/dx_N) - mu_cd (du/dy_S + dv/dx_S)),  coef = dt/rho_half,
// and the mirror formula at y-faces.  Every face is updated, walls too.  The
// level set's ghosts are linear extrapolations (levelset_with_ghosts: the
// reference leaves the four corner ghosts unfilled).
namespace {
double heaviside(double phi, double eps) {
  const double oe = (double)1 / eps, op = (double)1 / REF_PI;
  if (phi < -eps) return 0;
  if (phi > eps) return 1;
  return (double)0.5 + phi * oe * (double)0.5 + (double)0.5 * op * std::sin(REF_PI * phi * oe);
}
}  // namespace

void orc_apply_viscosity(int nx, int ny, double h, double dt, const double* phi_in, double rho_minus,
                         double rho_plus, double mu_minus, double mu_plus, double epsilon, double* u, double* v) {
  const Pad phi = levelset_with_ghosts(nx, ny, phi_in);
  const std::vector<double> u0(u, u + (size_t)(nx + 1) * ny), v0(v, v + (size_t)nx * (ny + 1));
  auto U = [&](int i, int j) { return (i < 0 || i > nx || j < 0 || j >= ny) ? 0.0 : u0[FX(i, j, nx)]; };
  auto V = [&](int i, int j) { return (i < 0 || i >= nx || j < 0 || j > ny) ? 0.0 : v0[FY(i, j, nx)]; };
  auto rho = [&](double p) { return rho_minus + (rho_plus - rho_minus) * heaviside(p, epsilon); };
  auto mu = [&](double p) { return mu_minus + (mu_plus - mu_minus) * heaviside(p, epsilon); };
  const double ox = 1.0 / h, oy = 1.0 / h;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const double ph = (double)0.5 * (phi(i - 1, j) + phi(i, j));
      const double pu = (double)0.25 * (phi(i - 1, j) + phi(i, j) + phi(i - 1, j + 1) + phi(i, j + 1));
      const double pd = (double)0.25 * (phi(i - 1, j - 1) + phi(i, j - 1) + phi(i - 1, j) + phi(i, j));
      const double mr = mu(phi(i, j)), mc = mu(phi(i - 1, j)), mcu = mu(pu), mcd = mu(pd);
      const double coef = dt * ((double)1 / rho(ph));
      const double first = (double)2 * coef * ox * (mr * ox * (U(i + 1, j) - U(i, j)) - mc * ox * (U(i, j) - U(i - 1, j)));
      const double second = coef * oy * (mcu * (oy * (U(i, j + 1) - U(i, j)) + ox * (V(i, j + 1) - V(i - 1, j + 1))) -
                                         mcd * (oy * (U(i, j) - U(i, j - 1)) + ox * (V(i, j) - V(i - 1, j))));
      u[FX(i, j, nx)] = u0[FX(i, j, nx)] + (first + second);
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double ph = (double)0.5 * (phi(i, j - 1) + phi(i, j));
      const double pu = (double)0.25 * (phi(i, j - 1) + phi(i + 1, j - 1) + phi(i, j) + phi(i + 1, j));
      const double pl = (double)0.25 * (phi(i - 1, j - 1) + phi(i, j - 1) + phi(i - 1, j) + phi(i, j));
      const double mu_u = mu(phi(i, j)), mc = mu(phi(i, j - 1)), mcu = mu(pu), mcl = mu(pl);
      const double coef = dt * ((double)1 / rho(ph));
      const double first = (double)2 * coef * oy * (mu_u * oy * (V(i, j + 1) - V(i, j)) - mc * oy * (V(i, j) - V(i, j - 1)));
      const double second = coef * ox * (mcu * (oy * (U(i + 1, j) - U(i + 1, j - 1)) + ox * (V(i + 1, j) - V(i, j))) -
                                         mcl * (oy * (U(i, j) - U(i, j - 1)) + ox * (V(i, j) - V(i - 1, j))));
      v[FY(i, j, nx)] = v0[FY(i, j, nx)] + (first + second);
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
