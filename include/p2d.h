/*
 * p2d -- MI355X (gfx950) pressure projection for 2D_Engine, C ABI.
 *
 * Replaces the projection path of the reference:
 *   PROJECTION_2D::Solve(dt)            PROJECTION_2D.h:235-257
 *     DetermineProjectionDensity        PROJECTION_2D.h:259-393
 *     DetermineDivergence               PROJECTION_2D.h:395-478
 *     DetermineJumpConditionField       PROJECTION_2D.h:760-866 (+ LEVELSET_2D.h:898-990)
 *     DeterminePressure                 PROJECTION_2D.h:480-509
 *       POISSON_SOLVER::Solve           POISSON_SOLVER.cpp:127-145
 *       BuildLinearSystemNodeJumpConditionVaribleCoefficient  POISSON_SOLVER.cpp:714-976
 *       LINEAR_SOLVER::Solve (CG/PCG)   LINEAR_SOLVER.h:31-35, CG_METHOD.cpp:13-73, PCG_METHOD.cpp:176-383
 *       VectorToGrid                    POISSON_SOLVER.cpp:1593-1603
 *     UpdateVelocity                    PROJECTION_2D.h:754-758, 1231-1417
 *
 * Layout: nx x ny cells of spacing h (dx == dy, as POISSON_SOLVER.cpp:872 assumes),
 * arrays i-fastest without ghosts: cell (i,j) -> i + j*nx; x-faces u (nx+1) x ny,
 * face (i,j) -> i + j*(nx+1) between cells (i-1,j),(i,j); y-faces v nx x (ny+1),
 * face (i,j) -> i + j*nx between cells (i,j-1),(i,j).  This is the reference's
 * ARRAY_2D order (ARRAY_2D.h:140-146) with the ghost ring stripped; the MAC grids
 * are EULERIAN_FLUID_SOLVER_2D.h:357 (u) and :363 (v).  Walls are no-flux
 * (Neumann_Boundary_Condition, PROJECTION_2D.h:1010-1047): wall-normal faces are
 * never modified.  "Minus" side = phi <= 0 (air in air_bubble_rising,
 * PROJECTION_2D.h:330-343).
 *
 * Pointers are host or device memory as selected by `mem`.  A context is not
 * thread-safe: call it from one thread (the reference runs the solve on thread 0
 * only, PROJECTION_2D.h:250-254).  Every call is ordered on the context's stream
 * and returns after its results are available.  Inside a call, the multigrid's
 * bottom pseudo-inverse runs on a second stream that the context owns, forked
 * from and joined back into the context's stream by events.
 */
#ifndef P2D_H
#define P2D_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (the reference has none: it exit()s or returns silently, see
 * LINEAR_SOLVER.cpp:12-40, CG_METHOD.cpp:38-73, LEVELSET_2D.h:869-872). */
enum {
  P2D_OK = 0,
  P2D_NOT_CONVERGED = 1, /* iteration caps reached before the tolerances */
  P2D_NAN = 2,           /* NaN/Inf or a breakdown (p.Ap <= 0) in the solve, or a zero level-set gradient */
  P2D_BAD_ARG = 3,       /* invalid sizes, pointers or options */
  P2D_HIP_ERROR = 4,     /* a HIP runtime call failed */
  P2D_NO_DEVICE = 5      /* no usable gfx950 device */
};

enum { P2D_MEM_HOST = 0, P2D_MEM_DEVICE = 1 };

/* Largest array the kernels address, in bytes: p2d_create rejects grids with
 * nx * ny * sizeof(double) above it (at most 536,870,911 cells; 23170^2 fits). */
#define P2D_MAX_ARRAY_BYTES 0xFFFFFFFFull

enum { P2D_FP64 = 0, P2D_MIXED = 1 }; /* solver precision */
enum { P2D_PREC_F64 = 0, P2D_PREC_F32 = 1 }; /* element type of test-hook buffers */

typedef struct p2d_config {
  int nx, ny;          /* cells per side (>= 2) */
  double h;            /* cell size */
  int precision;       /* P2D_FP64: FP64 MG-PCG; P2D_MIXED: FP32 MG-PCG inside FP64 refinement */
  int mg_max_levels;   /* 0: coarsen until <= 64 cells or an odd size */
  int mg_pre_sweeps;   /* red-black Gauss-Seidel sweeps before restriction, 1..3 (default 3) */
  int mg_post_sweeps;  /* ... after prolongation, 1..3 (default 3; equal counts keep M symmetric) */
  /* Stopping rules: every enabled one must hold for the true (FP64) residual r = b - A p of
   * the unscaled system (entries beta/h^2, b = -div - F as POISSON_SOLVER.cpp:951 builds it).
   * At least one of rtol, div_tol, abs_tol must be > 0. */
  double rtol;         /* ||r||_2 <= rtol ||b||_2 (mean-free b; 0 disables).  The configurations' rule */
  double div_tol;      /* the predicted max|D u_proj| = dt max|r| <= div_tol (0 disables); for
                        * p2d_solve_pressure, max|r| <= div_tol */
  int max_iterations;  /* cap on inner Krylov iterations per solve (PROJECTION_2D's `max_iteration`) */
  int max_outer;       /* cap on FP64 refinement passes */
  double inner_rtol;   /* mixed: relative residual reduction of each FP32 inner solve (default 1e-4) */
  double abs_tol;      /* ||r||_2 <= abs_tol (0 disables, the default): PROJECTION_2D's `tolerance`
                        * with the reference's absolute rule, res < tolerance^2 on r.r
                        * (CG_METHOD.cpp:52; its IC(0)-PCG tests r.z, PCG_METHOD.cpp:253) */
  double mg_omega;     /* over-relaxation of the red-black smoother on the finest level, 0 < omega < 2
                        * (SOR; 1 = Gauss-Seidel; default 1.4).  The V-cycle stays symmetric for every omega */
  double mg_omega_coarse; /* ... on the coarser levels (default 1.5) */
} p2d_config;

typedef struct p2d_stats {
  int status;               /* P2D_* code of the solve */
  int iterations;           /* inner Krylov iterations, all passes */
  int outer;                /* FP64 refinement passes */
  double rel_residual;      /* final true ||b - A p||_2 / ||b||_2 (FP64) */
  double max_abs_residual;  /* final true ||b - A p||_inf (FP64, unscaled operator) */
  double max_div;           /* max |D u_proj| over cells, FP64 (after the update) */
  double max_u, max_v;      /* max |u|, |v| after the update (PROJECTION_2D max_velocity_x/y) */
} p2d_stats;

typedef struct p2d_ctx p2d_ctx;

/* Fill `cfg` with defaults for an nx x ny grid of spacing h. */
void p2d_config_default(p2d_config* cfg, int nx, int ny, double h);

/* PROJECTION_2D constructor + InitializeFromBlock (PROJECTION_2D.h:107-232):
 * allocates every device buffer for the grid once.  P2D_BAD_ARG, before any device
 * call, for invalid options, nx * ny * 8 > P2D_MAX_ARRAY_BYTES, or sizes that do not
 * halve down to a bottom grid of at most 1024 cells. */
int p2d_create(const p2d_config* cfg, p2d_ctx** out);
int p2d_destroy(p2d_ctx* ctx);

/* Run on a caller-owned HIP stream (hipStream_t); NULL restores the context's own. */
int p2d_set_stream(p2d_ctx* ctx, void* hip_stream);

/* Fluid state, the fields the caller writes before Solve (EULERIAN_FLUID_SOLVER_2D.h:506-515):
 * water_levelset phi (nx*ny), air_density = rho_minus (phi <= 0),
 * water_density = rho_plus (phi > 0), surface_tension sigma.  Face beta = 1/rho_f
 * (POISSON_SOLVER.cpp:743-746) and the jump sigma*kappa (PROJECTION_2D.h:830-834)
 * are rebuilt from these inside every p2d_project. */
int p2d_set_fluid(p2d_ctx* ctx, const double* phi, double rho_minus, double rho_plus, double sigma, int mem);

/* Explicit coefficients instead of (phi, rho, sigma): face beta (walls ignored),
 * the level set (sides of the jump) and the jump on the solution jc (may be NULL = 0).
 * Used by the adjoint's beta-gradient checks; p2d_project then skips assembly. */
int p2d_set_coefficients(p2d_ctx* ctx, const double* beta_x, const double* beta_y, const double* phi,
                         const double* jc, int mem);

/* PROJECTION_2D::Solve(dt): u, v projected in place, pressure kept in the context. */
int p2d_project(p2d_ctx* ctx, double* u, double* v, double dt, p2d_stats* stats, int mem);

/* pressure_field after the last projection.  mean_zero = 1: mean removed;
 * 0: shifted so p(0,0) = 0, the reference's pin (POISSON_SOLVER.cpp:789-807). */
int p2d_get_pressure(p2d_ctx* ctx, double* p, int mean_zero, int mem);

/* PROJECTION_2D::DetermineDivergence(dt) (air-water, PROJECTION_2D.h:395-414): the cell field
 * div = (D u) / dt that the projection's right-hand side is built from (nx*ny; the
 * reference's divergence_field, in its operation order). */
int p2d_divergence(p2d_ctx* ctx, const double* u, const double* v, double dt, double* div, int mem);

/* PROJECTION_2D::DeterminePressure() (PROJECTION_2D.h:480-509, called directly by the solver at
 * EULERIAN_FLUID_SOLVER_2D.h:1513, 1681): the ghost-fluid Poisson solve A p = -div - F
 * (POISSON_SOLVER::Solve, POISSON_SOLVER.cpp:127-145; A and F of the current fluid, as in
 * p2d_project) for a caller-written divergence_field div (nx*ny, already divided by dt).  The
 * pressure is kept in the context (p2d_get_pressure) and, when p is not NULL, written to p in
 * the reference's gauge p(0,0) = 0.  The no-flux operator is singular on constants, so the mean
 * of -div - F is removed first; for a compatible right-hand side (zero sum: any div = D u / dt
 * of wall-tight faces) the result is the reference's pinned solution.  Stops on the
 * configuration's rules, with div_tol bounding max|r|.  Velocities are not touched. */
int p2d_solve_pressure(p2d_ctx* ctx, const double* div, double* p, p2d_stats* stats, int mem);

/* Coefficients built by the last p2d_project / p2d_set_coefficients (any may be NULL):
 * beta_x ((nx+1)*ny), beta_y (nx*(ny+1)), jc (nx*ny) = sigma*kappa. */
int p2d_get_coefficients(p2d_ctx* ctx, double* beta_x, double* beta_y, double* jc, int mem);

/* Adjoint of the last projection (no reference counterpart).  Given g = dL/du_proj,
 * dL/dv_proj on interior faces, returns dL/du_in, dL/dv_in and dL/dbeta_x, dL/dbeta_y
 * (wall entries 0).  One extra solve with the same symmetric operator; only the forward
 * pressure is kept from the forward pass.  Outputs may be NULL.  P2D_BAD_ARG when the
 * context's last solve was not a p2d_project. */
int p2d_adjoint(p2d_ctx* ctx, const double* gu, const double* gv, double* du, double* dv, double* dbeta_x,
                double* dbeta_y, p2d_stats* stats, int mem);

/* Level-set re-initialisation, in place on phi (nx*ny cells, FP64): the
 * sussmanstyle_reinitialization branch of EULERIAN_FLUID_SOLVER_2D::Reinitialization
 * (EULERIAN_FLUID_SOLVER_2D.h:2365-2400): s = phi0/sqrt(phi0^2 + h^2), then `iterations`
 * steps of ADVECTION_2D::ReinitializationBySussman (ADVECTION_2D.cpp:415-421 ->
 * ADVECTION_METHOD_2D::WENO5thReinitialization, ADVECTION_METHOD_2D.h:805-1076) with the
 * fictitious dt = h/scaling and WENO epsilon `epsilon` (the reference's scripts: 25 or 10
 * iterations, scaling 5, epsilon 1e-6).  Needs nx, ny >= 3; independent of the fluid state. */
int p2d_reinitialize(p2d_ctx* ctx, double* phi, int iterations, double scaling, double epsilon, int mem);

/* One forward-Euler stage of the level-set advection, in place on phi (nx*ny, FP64):
 * ADVECTION_2D::Solve_Levelset with use_mac_grid and use_5th_weno (ADVECTION_2D.cpp:165-200 ->
 * ADVECTION_METHOD_2D::WENO5th, ADVECTION_METHOD_2D.h:508-618) by the MAC velocity u ((nx+1)*ny
 * x-faces), v (nx*(ny+1) y-faces): phi -= dt (u phi_x + v phi_y), cell velocity = mean of its
 * faces, upwind WENO5 derivatives with `epsilon` (the scripts: 1e-6).  The caller's RK stages
 * combine these stages as the reference's time integrator does.  Needs nx, ny >= 3. */
int p2d_advect_levelset(p2d_ctx* ctx, double* phi, const double* u, const double* v, double dt, double epsilon,
                        int mem);

/* One forward-Euler stage of the velocity advection, in place on the MAC velocity u
 * ((nx+1)*ny), v (nx*(ny+1)), both components by the old velocity:
 * ADVECTION_2D::Solve_Velocity with use_3rd_eno_v (ADVECTION_2D.cpp:95-160 ->
 * ADVECTION_METHOD_2D::ENO3rd, ADVECTION_METHOD_2D.h:2038-2232).  Every face, walls
 * included, is advected as in the reference; its boundary conditions are applied after.
 * Needs nx, ny >= 3. */
int p2d_advect_velocity(p2d_ctx* ctx, double* u, double* v, double dt, int mem);

/* Explicit viscosity step of the air-water delta-function formulation, in place on the MAC
 * velocity u ((nx+1)*ny), v (nx*(ny+1)): VISCOSITY_2D::ApplyViscosity (VISCOSITY_2D.h:460-548,
 * called at EULERIAN_FLUID_SOLVER_2D.h:2086-2089).  Uses the level set and the densities of the
 * last p2d_set_fluid; mu_minus / mu_plus are the viscosities on the phi <= 0 / phi > 0 sides
 * (air / water for air_bubble_rising) and epsilon the mollification width (the reference's
 * epsilon_for_mollification = 1.5 dx, EULERIAN_FLUID_SOLVER_2D.h:347).  Every face is updated,
 * walls included, as in the reference. */
int p2d_apply_viscosity(p2d_ctx* ctx, double* u, double* v, double dt, double mu_minus, double mu_plus,
                        double epsilon, int mem);

/* Parameters of p2d_advance: the AIR_WATER_SIMULATION block of the reference's script
 * (2D_script.sim:265-345) as the solver reads them. */
typedef struct p2d_step_params {
  double gravity_x, gravity_y;  /* ApplyGraivity: every face += dt g (default (0, -9.8)) */
  double rho_minus, rho_plus;   /* densities where phi <= 0 / phi > 0 */
  double sigma;                 /* surface tension of the projection's jump */
  double mu_minus, mu_plus;     /* viscosities; both 0 = no viscosity step */
  double mollification;         /* epsilon_for_mollification; <= 0: 1.5 h */
  double weno_epsilon;          /* WENO epsilon of advection and re-initialisation (1e-6) */
  int reinit_iterations;        /* Sussman steps per Reinitialization call (25); 0 = none */
  double reinit_scaling;        /* scaling_number_for_reinitialization (5) */
  int rk_order;                 /* order_for_time_advancing (EULERIAN_FLUID_SOLVER_2D.h:1142-1797):
                                 * 1 forward Euler, 2 two stages averaged 1/2-1/2, 3 TVD-RK3 (3, the
                                 * scripts' value) */
} p2d_step_params;

void p2d_step_params_default(p2d_step_params* params);

/* The time step the CFL rule allows for the velocity (u, v) and level set phi:
 * EULERIAN_FLUID_SOLVER_2D::CFLOneTimeStep for the air-water solver
 * (EULERIAN_FLUID_SOLVER_2D.h:3517-3637; SIMULATION_WORLD::DetermineTimeStep calls it before
 * every step).  Advection max|u|/h + max|v|/h, gravity sqrt(9.8/h), viscosity
 * max(mu/rho) (4/h^2), surface tension sqrt(sigma max|curvature| / (min rho h^2)), combined as
 * cfl / (0.5 (c + w + sqrt((c + w)^2 + 4 g^2 + 4 s^2))).  *dt = 0 when that denominator is 0
 * (the caller then keeps its own dt).  phi may be NULL when params->sigma is 0.
 * P2D_NAN when the curvature is undefined (a vanishing gradient). */
int p2d_cfl_dt(p2d_ctx* ctx, const double* phi, const double* u, const double* v, const p2d_step_params* params,
               double cfl_number, double* dt, int mem);

/* p2d_cfl_dt without allocation (no counterpart in the reference): for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  No context: the grid is nx x ny cells of size h; phi (nx*ny, may be NULL
 * when params->sigma is 0), u ((nx+1)*ny), v (nx*(ny+1)) and dt (one double) are DEVICE
 * pointers.  Enqueues four kernels on hip_stream (NULL: the null stream) and returns without
 * waiting; *dt then holds p2d_cfl_dt's value, or NaN where p2d_cfl_dt returns P2D_NAN.  Its
 * reductions use a static device array of the library (1.7 MB per device, kept until the process
 * ends), so it serves grids whose (nx+1) x (ny+1) band needs at most P2D_NOALLOC_MAX_BLOCKS
 * blocks of 256 x 4 (any grid up to 8192 x 8192); larger ones get P2D_BAD_ARG.  Calls that overlap
 * on one device race on that array: order them on one stream. */
#define P2D_NOALLOC_MAX_BLOCKS 69632
int p2d_cfl_dt_noalloc(int nx, int ny, double h, const double* phi, const double* u, const double* v,
                       const p2d_step_params* params, double cfl_number, double* dt, void* hip_stream);

/* One time step of the air-water solver, in place on phi (nx*ny), u ((nx+1)*ny), v (nx*(ny+1)):
 * EULERIAN_FLUID_SOLVER_2D::AdvanceOneTimeStepThread with order_for_time_advancing =
 * params->rk_order (EULERIAN_FLUID_SOLVER_2D.h:1142-1797, no vortex sheet).  Order 3: level set
 * by three WENO5 advection stages (p2d_advect_levelset) combined 3/4-1/4 and 1/3-2/3;
 * re-initialisation by three p2d_reinitialize calls combined the same way; velocity by three
 * SolveForRK stages (ENO3 advection, gravity, viscosity, wall velocities 0, p2d_project with the
 * new level set's fluid) combined the same way.  Order 2: two stages of each, averaged 1/2-1/2;
 * order 1: one stage of each.  stats: the last projection's (status P2D_NOT_CONVERGED if any
 * projection did not converge).  The caller chooses dt (CFL). */
int p2d_advance(p2d_ctx* ctx, double* phi, double* u, double* v, double dt, const p2d_step_params* params,
                p2d_stats* stats, int mem);

/* Unit normals of the level set phi (nx*ny), LEVELSET_2D::ComputeNormals (LEVELSET_2D.h:593-620):
 * central differences over linear-extrapolation ghosts, forward differences where the central
 * gradient vanishes, (0, 0) where both vanish. */
int p2d_levelset_normals(p2d_ctx* ctx, const double* phi, double* normal_x, double* normal_y, int mem);

/* Test hook: y = A x with the consistent operator -D B G (no-flux walls, no pin,
 * entries beta/h^2) of the current coefficients, in FP64 or FP32 (prec). */
int p2d_apply_stencil(p2d_ctx* ctx, const void* x, void* y, int prec, int mem);

/* Test hook: z = M r for the multigrid preconditioner of the current coefficients on the
 * h^2-scaled operator (one V-cycle from z = 0), FP64 or FP32 (prec). */
int p2d_vcycle(p2d_ctx* ctx, const void* r_scaled, void* z, int prec, int mem);

/* Number of multigrid levels of the current configuration. */
int p2d_mg_levels(p2d_ctx* ctx);

/* Bench hook: average duration in ms of `reps` back-to-back launches of one kernel at
 * the finest level, timed with HIP events on the context's stream.
 * which: 0 = FP32 direction update + stencil + p.q (k_dir<float>), 1 = FP64 of the same,
 *        2 = FP32 fused Krylov update + pre-smoothing + restriction (k_mg_pre<float,true>),
 *        3 = FP32 prolongation + post-smoothing + r.z (k_mg_post<float,true>),
 *        4 = true residual in double-double with the FP32 correction (k_correct_resid),
 *        5 = one level-set re-initialisation step of the fluid's level set (k_reinit_step). */
int p2d_time_kernel(p2d_ctx* ctx, int which, int reps, double* avg_ms);

/* Bench hook: from now on (enable = 1) time every launch of kernel `which` inside the
 * solves with a pair of HIP events on the context's stream, or stop (0); both reset the
 * totals.  which: p2d_time_kernel's 0..4, and 5 = FP64 true residual, 6 = FP64 fused
 * pre-smoothing (k_mg_pre<double,true>), 7 = FP64 post-smoothing (k_mg_post<double,true>). */
int p2d_kernel_timing(p2d_ctx* ctx, int which, int enable);

/* Totals of the timed launches of kernel `which` since the last reset. */
int p2d_kernel_time(p2d_ctx* ctx, int which, double* total_ms, int* launches);

const char* p2d_status_string(int status);

/* ---------------------------------------------------------------------------------------
 * Semi-implicit viscosity: the linear solve of POISSON_SOLVER::SolveForViscosity and the
 * oil-water driver around it.  A p2d_visc object serves one grid of nx*ny cells with spacings
 * dx, dy (the oil-water grids need neither square cells nor multigrid-friendly sizes, so it
 * is separate from p2d_ctx).  Component 0 is the x-face grid ((nx+1)*ny), 1 the y-face grid
 * (nx*(ny+1)), row-major, i fastest. */
typedef struct p2d_visc p2d_visc;

enum {
  P2D_VISC_REFERENCE = 0, /* out-of-grid neighbours couple to node 0, as the reference's CSR does */
  P2D_VISC_DIRICHLET = 1  /* out-of-grid neighbours are zero velocity */
};

int p2d_visc_create(int nx, int ny, double dx, double dy, p2d_visc** out);
int p2d_visc_destroy(p2d_visc* visc);
int p2d_visc_set_stream(p2d_visc* visc, void* hip_stream);

/* POISSON_SOLVER::SolveForViscosity(velocity, coef_1..coef_5, bc, explicit_term)
 * (POISSON_SOLVER.cpp:175-184) with the system of
 * BuildLinearSystemNodeForSemiImplicitViscosity (POISSON_SOLVER.cpp:620-712): at node (i, j)
 *   (1 + (c1 + c2 + c3 + c4)/dx^2 + c5) x - c2/dx^2 x_W - c1/dx^2 x_E - c4/dy^2 x_S - c3/dy^2 x_N
 *     = explicit_term.
 * The reference's boundary-condition field keeps 0 = BC_FULL in its ghost ring (VISCOSITY_2D.h:
 * 647-663 visits only the grid's own nodes), so an out-of-grid neighbour is unknown 0 there:
 * boundary = P2D_VISC_REFERENCE reproduces that, P2D_VISC_DIRICHLET takes zero velocity.
 * velocity: initial guess in, solution out.  The system is not symmetric; it is solved by
 * Jacobi-preconditioned BiCGStab in FP64 to the reference's rule ||r||_2 < tolerance (the
 * reference runs CG).  stats: iterations, rel_residual = ||r||/||b||, max_abs_residual = ||r||_2.
 * P2D_NOT_CONVERGED after max_iterations; P2D_NAN on a breakdown. */
int p2d_visc_solve(p2d_visc* visc, int component, double* velocity, const double* c1, const double* c2,
                   const double* c3, const double* c4, const double* c5, const double* explicit_term,
                   int boundary, double tolerance, int max_iterations, p2d_stats* stats, int mem);

/* The OIL_WATER_SIMULATION values VISCOSITY_2D reads (2D_script.sim:345-433). */
typedef struct p2d_oil_water_params {
  double x_min;          /* x of the first cell centre: the radius r of the axisymmetric form */
  double Re;             /* Reynolds number (0.94983) */
  double rho_oil, rho_water, mu_oil, mu_water; /* 0.905, 0.995, 6.01, 0.01 */
  double mollification;  /* epsilon_for_mollification; <= 0: 1.5 dx */
  double tolerance;      /* PROJECTION tolerance (1e-6), absolute on ||r||_2 */
  int max_iterations;    /* PROJECTION max_iteration (10000) */
  int boundary;          /* P2D_VISC_REFERENCE (default) or P2D_VISC_DIRICHLET */
} p2d_oil_water_params;

void p2d_oil_water_params_default(p2d_oil_water_params* params);

/* VISCOSITY_2D::ApplyViscosity(epsilon, dt) for oil_water_simulation with
 * semi_implicit_approach and dimensionless_form (VISCOSITY_2D.h:546-640): mollified
 * dimensionless density and viscosity, the five coefficient fields and the explicit term of
 * the x component, its solve, then the y component's (whose explicit term reads u before and
 * after the x solve) and its solve; u and v in place.  phi: nx*ny cells. */
int p2d_visc_oil_water(p2d_visc* visc, const double* phi, double* u, double* v, double dt,
                       const p2d_oil_water_params* params, p2d_stats* stats_x, p2d_stats* stats_y, int mem);

#ifdef __cplusplus
}
#endif
#endif /* P2D_H */
