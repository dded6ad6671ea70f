"""ctypes front end of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests, the smoke check and the benchmark's CPU-baseline leg import this
module, and only as the checker / CPU baseline.  The product library never calls
it.  Every function wraps one entry point of oracle/p2d_oracle.cpp, which cites
the reference lines it restates.

Array layout: cells (ny, nx) row-major (i fastest); x-faces (ny, nx+1); y-faces
(ny+1, nx).  All arrays are float64.

Parity pins (how each part is checked, see tests/test_oracle_*.py):
  curvature        closed-form circle / Rayleigh-Taylor curvature, O(h^2)
  face beta        closed form (1/rho same side, harmonic-weighted across)
  assembly + jump  LFK test 3 (POISSON_EQUATION_TEST.h:315-343, 409-423,
                   490-507, 608-624) converging under refinement; operator
                   invariants (symmetry, A*1 = pin row only); balanced-force
                   static drop (p_in - p_out = -a exactly, u_proj = 0)
  Krylov solvers   CG, IC(0)-PCG, Jacobi-PCG and MG-PCG agree on the mean-zero
                   solution; IC(0) factor reproduces A on its pattern
  MG V-cycle       symmetry <Mx,y> = <x,My>, constants in its null space, bounded
                   spectrum of M A and iteration counts (constant and 1000:1 coefficients)
  adjoint          central finite differences and the dot-product identity
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.c_double
_i = ctypes.c_int
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def build():
    """Compile oracle/liboracle.so from oracle/p2d_oracle.cpp (make)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        # P2D_ORACLE_LIB: another build of the same source (e.g. the sanitizer build of
        # `make -C oracle asan`)
        path = os.environ.get("P2D_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        sig = {
            "orc_density": (None, [_i, _i, _pd, _d, _d, _pd, _pd]),
            "orc_divergence": (None, [_i, _i, _d, _d, _pd, _pd, _pd]),
            "orc_curvature": (_i, [_i, _i, _d, _pd, _pd]),
            "orc_normals": (None, [_i, _i, _d, _pd, _pd, _pd]),
            "orc_face_beta": (None, [_i, _i, _pd, _d, _d, _pd, _pd]),
            "orc_assemble_reference": (_i, [_i, _i, _d, _pi, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd,
                                            _pi, _pi, _pd, _pd, _pd, _pi]),
            "orc_csr_spmv": (None, [_i, _pi, _pi, _pd, _pd, _pd]),
            "orc_cg": (_i, [_i, _pi, _pi, _pd, _pd, _pd, _d, _i, _i, _pd]),
            "orc_pcg_ic0": (_i, [_i, _i, _pi, _pi, _pd, _pd, _pd, _d, _i, _i, _pd]),
            "orc_pcg_jacobi": (_i, [_i, _pi, _pi, _pd, _pd, _pd, _d, _i, _i, _i, _pd]),
            "orc_ic0_factor": (None, [_i, _i, _pi, _pi, _pd, _pd, _pd, _pd]),
            "orc_ic0_solve": (None, [_i, _i, _pi, _pi, _pd, _pd, _pd]),
            "orc_apply_A0": (None, [_i, _i, _d, _pd, _pd, _pd, _pd]),
            "orc_rhs_consistent": (None, [_i, _i, _d, _d, _pd, _pd, _pd, _pd, _pd, _pd, _pd]),
            "orc_update_consistent": (None, [_i, _i, _d, _d, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd]),
            "orc_update_reference_literal": (None, [_i, _i, _d, _d, _pd, _pd, _d, _d, _pd, _pd, _pd]),
            "orc_half_via_float": (_d, [_d]),
            "orc_mg_num_levels": (_i, [_i, _i, _i, _i]),
            "orc_mg_create": (_vp, [_i, _i, _pd, _pd, _i, _i, _i, _i, _i, _d, _d]),
            "orc_mg_piece": (_i, [_vp, _i, _i, _pd, _pd]),
            "orc_mg_destroy": (None, [_vp]),
            "orc_mg_levels": (_i, [_vp]),
            "orc_mg_vcycle": (None, [_vp, _pd, _pd]),
            "orc_mg_pcg": (_i, [_i, _i, _d, _pd, _pd, _pd, _pd, _d, _d, _i, _i, _i, _i, _d, _d, _pd, _pd]),
            "orc_project": (_i, [_i, _i, _d, _d, _pd, _d, _d, _d, _pd, _pd, _i, _d, _i, _i, _pd, _pd]),
            "orc_adjoint": (_i, [_i, _i, _d, _d, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _d, _pd, _pd, _pd, _pd, _pd]),
            "orc_project_beta": (_i, [_i, _i, _d, _d, _pd, _pd, _pd, _pd, _pd, _pd, _d, _pd]),
            "orc_num_threads": (_i, []),
            "orc_weno5_derivatives": (None, [_i, _i, _d, _pd, _d, _pd, _pd, _pd, _pd]),
            "orc_reinit_sussman": (None, [_i, _i, _d, _pd, _i, _d, _d, _pd]),
            "orc_advect_levelset": (None, [_i, _i, _d, _d, _pd, _pd, _pd, _d, _pd]),
            "orc_advect_velocity": (None, [_i, _i, _d, _d, _pd, _pd, _pd, _pd]),
            "orc_apply_viscosity": (None, [_i, _i, _d, _d, _pd, _d, _d, _d, _d, _d, _pd, _pd]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.float64:
        return a.ctypes.data_as(_pd)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_pi)
    raise TypeError(a.dtype)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


# ----------------------------------------------------------------------------- per-step kernels
def density(phi, rho_minus, rho_plus):
    phi = _f64(phi)
    ny, nx = phi.shape
    rho = np.empty_like(phi)
    inv = np.empty_like(phi)
    lib().orc_density(nx, ny, _p(phi), rho_minus, rho_plus, _p(rho), _p(inv))
    return rho, inv


def divergence(u, v, h, dt=1.0):
    u, v = _f64(u), _f64(v)
    ny, nx1 = u.shape
    nx = nx1 - 1
    out = np.empty((ny, nx))
    lib().orc_divergence(nx, ny, h, dt, _p(u), _p(v), _p(out))
    return out


def curvature(phi, h):
    phi = _f64(phi)
    ny, nx = phi.shape
    out = np.empty_like(phi)
    if lib().orc_curvature(nx, ny, h, _p(phi), _p(out)) != 0:
        raise FloatingPointError("zero level-set gradient (reference exit(0) path)")
    return out


def normals(phi, h):
    phi = _f64(phi)
    ny, nx = phi.shape
    a, b = np.empty_like(phi), np.empty_like(phi)
    lib().orc_normals(nx, ny, h, _p(phi), _p(a), _p(b))
    return a, b


def weno5_derivatives(phi, h, epsilon=1e-6):
    """phi_x^-, phi_x^+, phi_y^-, phi_y^+ of the re-initialisation step (reference ghost rule)."""
    phi = _f64(phi)
    ny, nx = phi.shape
    out = [np.empty_like(phi) for _ in range(4)]
    lib().orc_weno5_derivatives(nx, ny, h, _p(phi), epsilon, *[_p(o) for o in out])
    return tuple(out)


def reinit_sussman(phi, h, iterations=25, scaling=5.0, epsilon=1e-6):
    """Sussman WENO5 re-initialisation (sussmanstyle_reinitialization) of the level set."""
    phi = _f64(phi)
    ny, nx = phi.shape
    out = np.empty_like(phi)
    lib().orc_reinit_sussman(nx, ny, h, _p(phi), int(iterations), float(scaling), float(epsilon), _p(out))
    return out


def advect_levelset(phi, u, v, h, dt, epsilon=1e-6):
    """One forward-Euler WENO5 advection stage of the level set by the MAC velocity (u, v)."""
    phi, u, v = _f64(phi), _f64(u), _f64(v)
    ny, nx = phi.shape
    out = np.empty_like(phi)
    lib().orc_advect_levelset(nx, ny, h, dt, _p(phi), _p(u), _p(v), float(epsilon), _p(out))
    return out


def advect_velocity(u, v, h, dt):
    """One forward-Euler ENO3 advection stage of the MAC velocity by itself (both components
    by the old velocity)."""
    u, v = _f64(u), _f64(v)
    ny, nx1 = u.shape
    nx = nx1 - 1
    uo, vo = np.empty_like(u), np.empty_like(v)
    lib().orc_advect_velocity(nx, ny, h, dt, _p(u), _p(v), _p(uo), _p(vo))
    return uo, vo


def apply_viscosity(phi, u, v, h, dt, rho_minus, rho_plus, mu_minus, mu_plus, epsilon=None):
    """Explicit air-water viscosity update of the MAC velocity (mollified density and
    viscosity, width epsilon = 1.5 h by default); returns new arrays."""
    phi = _f64(phi)
    ny, nx = phi.shape
    uo, vo = _f64(u).copy(), _f64(v).copy()
    eps = 1.5 * h if epsilon is None else epsilon
    lib().orc_apply_viscosity(nx, ny, h, dt, _p(phi), rho_minus, rho_plus, mu_minus, mu_plus, eps, _p(uo), _p(vo))
    return uo, vo


def cfl_dt(phi, u, v, h, cfl_number=0.5, rho_minus=1.0, rho_plus=1000.0, sigma=0.0, mu_minus=0.0, mu_plus=0.0):
    """EULERIAN_FLUID_SOLVER_2D::CFLOneTimeStep, air-water branch (EULERIAN_FLUID_SOLVER_2D.h:
    3517-3637): advection c_f, gravity g_f = sqrt(9.8/dy), viscosity v_f, surface tension s_f
    from the largest |curvature|, CFL = (c_f + v_f + sqrt((c_f + v_f)^2 + 4 g_f^2 + 4 s_f^2))/2,
    dt = cfl_number / CFL (0 when CFL is 0)."""
    c_f = np.abs(u).max() * (1.0 / h) + np.abs(v).max() * (1.0 / h)
    g_f = np.sqrt(9.8 * (1.0 / h))
    v_f = max(mu_plus / rho_plus, mu_minus / rho_minus) * (2 * (1.0 / (h * h)) + 2 * (1.0 / (h * h)))
    kmax = np.abs(curvature(phi, h)).max() if sigma != 0.0 else 0.0
    s_f = np.sqrt(sigma * kmax / (min(rho_plus, rho_minus) * (h * h)))
    cfl = 0.5 * (c_f + v_f + np.sqrt((c_f + v_f) ** 2 + 4 * g_f ** 2 + 4 * s_f ** 2))
    return 0.0 if cfl == 0 else float(cfl_number / cfl)


def rk3(y, stage):
    """The third-order TVD Runge-Kutta combination the solver uses for the level set, its
    re-initialisation and the velocity (EULERIAN_FLUID_SOLVER_2D.h:1700-1795):
    y1 = L(L(y)); y2 = 3/4 y + 1/4 y1; return 1/3 y + 2/3 L(y2).  y is an array or a tuple
    of arrays; stage maps it to the same structure."""
    tup = isinstance(y, tuple)
    ys = y if tup else (y,)
    comb = lambda a, x, b, z: tuple(a * xi + b * zi for xi, zi in zip(x, z))
    as_t = lambda r: r if tup else (r,)
    y1 = as_t(stage(stage(y)))
    y2 = comb(0.75, ys, 0.25, y1)
    y3 = as_t(stage(y2 if tup else y2[0]))
    out = comb(1 / 3, ys, 2 / 3, y3)
    return out if tup else out[0]


def rk(order, y, stage):
    """The time-advancing combination of order 1, 2 or 3 (order_for_time_advancing,
    EULERIAN_FLUID_SOLVER_2D.h:1142-1797, no vortex sheet): 1 is one stage (:1191, :1193,
    :1197); 2 is two stages averaged, 1/2 y + 1/2 L(L(y)) (:1308-1313, :1324-1329,
    :1349-1360); 3 is rk3."""
    if order == 3:
        return rk3(y, stage)
    if order == 1:
        return stage(y)
    if order != 2:
        raise ValueError("order must be 1, 2 or 3")
    tup = isinstance(y, tuple)
    ys = y if tup else (y,)
    y1 = stage(stage(y))
    y1 = y1 if tup else (y1,)
    out = tuple(0.5 * a + 0.5 * b for a, b in zip(ys, y1))
    return out if tup else out[0]


def advance(phi, u, v, h, dt, gravity=(0.0, -9.8), rho_minus=1.0, rho_plus=1000.0, sigma=0.0, mu_minus=0.0,
            mu_plus=0.0, mollification=None, weno_epsilon=1e-6, reinit_iterations=25, reinit_scaling=5.0,
            solver=None, tol=1e-13, rk_order=3):
    """One air-water time step composed of the restated stages,
    EULERIAN_FLUID_SOLVER_2D::AdvanceOneTimeStepThread with order rk_order (default 3) and no
    vortex sheet (EULERIAN_FLUID_SOLVER_2D.h:1142-1797): rk of the level-set advection stage by
    the step's initial velocity, rk of the Reinitialization call, then rk of SolveForRK
    (EULERIAN_FLUID_SOLVER_2D.h:1857-1877: ENO3 advection, ApplyGraivity :3249-3269,
    viscosity, wall velocities 0 :2227-2260, PROJECTION_2D::Solve with the new level set).
    Returns (phi, u, v)."""
    u0, v0 = _f64(u).copy(), _f64(v).copy()
    p = rk(rk_order, _f64(phi).copy(), lambda q: advect_levelset(q, u0, v0, h, dt, weno_epsilon))
    if reinit_iterations > 0:
        p = rk(rk_order, p, lambda q: reinit_sussman(q, h, reinit_iterations, reinit_scaling, weno_epsilon))
    eps = 1.5 * h if mollification is None or mollification <= 0 else mollification
    solver = SOLVER_IC0 if solver is None else solver

    def stage(uv):
        a, b = advect_velocity(uv[0], uv[1], h, dt)
        a = a + dt * gravity[0]
        b = b + dt * gravity[1]
        if mu_minus != 0.0 or mu_plus != 0.0:
            a, b = apply_viscosity(p, a, b, h, dt, rho_minus, rho_plus, mu_minus, mu_plus, eps)
        a[:, 0] = 0.0
        a[:, -1] = 0.0
        b[0, :] = 0.0
        b[-1, :] = 0.0
        _, a, b, _ = project(p, rho_minus, rho_plus, sigma, a, b, h, dt, solver=solver, tol=tol)
        return a, b

    a, b = rk(rk_order, (u0, v0), stage)
    return p, a, b


def face_beta(phi, rho_minus, rho_plus):
    phi = _f64(phi)
    ny, nx = phi.shape
    bx = np.empty((ny, nx + 1))
    by = np.empty((ny + 1, nx))
    lib().orc_face_beta(nx, ny, _p(phi), rho_minus, rho_plus, _p(bx), _p(by))
    return bx, by


# ----------------------------------------------------------------------------- reference assembly
BC_FULL, BC_DIR, BC_NEUM = 0, -1, -4


class CSR:
    def __init__(self, n, row_ptr, col, val):
        self.n, self.row_ptr, self.col, self.val = n, row_ptr, col, val

    def matvec(self, x):
        x = _f64(x)
        y = np.empty(self.n)
        lib().orc_csr_spmv(self.n, _p(self.row_ptr), _p(self.col), _p(self.val), _p(x), _p(y))
        return y

    def to_dense(self):
        A = np.zeros((self.n, self.n))
        for r in range(self.n):
            for k in range(self.row_ptr[r], self.row_ptr[r + 1]):
                A[r, self.col[k]] += self.val[k]
        return A

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val.copy(), self.col.copy(), self.row_ptr.copy()), shape=(self.n, self.n))


def assemble_reference(phi, variable, div, h, jc=None, jcd=None, normals_xy=None, bc_pad=None, p_pad=None):
    """Reference-literal CSR system (POISSON_SOLVER.cpp:714-976). Returns (CSR, b, x0, row_of_cell)."""
    phi, variable, div = _f64(phi), _f64(variable), _f64(div)
    ny, nx = phi.shape
    N = nx * ny
    row_ptr = np.zeros(N + 1, np.int32)
    col = np.zeros(5 * N, np.int32)
    val = np.zeros(5 * N)
    b = np.zeros(N)
    x0 = np.zeros(N)
    row_of = np.zeros(N, np.int32)
    nrx, nry = (None, None) if normals_xy is None else (_f64(normals_xy[0]), _f64(normals_xy[1]))
    bcp = None if bc_pad is None else np.ascontiguousarray(bc_pad, np.int32)
    n = lib().orc_assemble_reference(nx, ny, h, _p(bcp), _p(_f64(p_pad)), _p(phi), _p(variable), _p(div),
                                     _p(_f64(jc)), _p(_f64(jcd)), _p(nrx), _p(nry), _p(row_ptr), _p(col),
                                     _p(val), _p(b), _p(x0), _p(row_of))
    nnz = int(row_ptr[n])
    return CSR(n, row_ptr[: n + 1].copy(), col[:nnz].copy(), val[:nnz].copy()), b[:n].copy(), x0[:n].copy(), row_of


def cg(A, b, x0=None, tol=1e-10, max_iteration=100000, relative=True):
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    res = _d()
    it = lib().orc_cg(A.n, _p(A.row_ptr), _p(A.col), _p(A.val), _p(x), _p(_f64(b)), tol, max_iteration,
                      1 if relative else 0, ctypes.byref(res))
    return x, it, res.value


def pcg_ic0(A, b, nx, ny, x0=None, tol=1e-10, max_iteration=100000, relative=True):
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    res = _d()
    it = lib().orc_pcg_ic0(nx, ny, _p(A.row_ptr), _p(A.col), _p(A.val), _p(x), _p(_f64(b)), tol, max_iteration,
                           1 if relative else 0, ctypes.byref(res))
    return x, it, res.value


def ic0_factor(A, nx, ny):
    """IC(0) factor of a Neumann-box CSR matrix (CSR_MATRIX.h:535-628): per cell the
    bottom, left and diagonal entries of its row of L, each (ny, nx)."""
    Lb, Ll, Ld = (np.empty((ny, nx)) for _ in range(3))
    lib().orc_ic0_factor(nx, ny, _p(A.row_ptr), _p(A.col), _p(A.val), _p(Lb), _p(Ll), _p(Ld))
    return Lb, Ll, Ld


def ic0_solve(A, r, nx, ny):
    """z = (L L^T)^-1 r by the reference's two triangular solves (PCG_METHOD.cpp:70-118)."""
    z = np.empty(A.n)
    lib().orc_ic0_solve(nx, ny, _p(A.row_ptr), _p(A.col), _p(A.val), _p(_f64(r)), _p(z))
    return z


def pcg_jacobi(A, b, x0=None, tol=1e-10, max_iteration=100000, relative=True, threads=0):
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    res = _d()
    it = lib().orc_pcg_jacobi(A.n, _p(A.row_ptr), _p(A.col), _p(A.val), _p(x), _p(_f64(b)), tol, max_iteration,
                              1 if relative else 0, threads, ctypes.byref(res))
    return x, it, res.value


# ----------------------------------------------------------------------------- consistent operator
def apply_A0(bx, by, x, h):
    bx, by, x = _f64(bx), _f64(by), _f64(x)
    ny, nx = x.shape
    y = np.empty_like(x)
    lib().orc_apply_A0(nx, ny, h, _p(bx), _p(by), _p(x), _p(y))
    return y


def rhs_consistent(phi, jc, bx, by, u, v, h, dt=1.0):
    phi = _f64(phi)
    ny, nx = phi.shape
    b = np.empty((ny, nx))
    lib().orc_rhs_consistent(nx, ny, h, dt, _p(phi), _p(_f64(jc)), _p(_f64(bx)), _p(_f64(by)), _p(_f64(u)),
                             _p(_f64(v)), _p(b))
    return b


def update_consistent(phi, jc, bx, by, p, u, v, h, dt=1.0):
    phi = _f64(phi)
    ny, nx = phi.shape
    u2, v2 = _f64(u).copy(), _f64(v).copy()
    mu, mv = _d(), _d()
    lib().orc_update_consistent(nx, ny, h, dt, _p(phi), _p(_f64(jc)), _p(_f64(bx)), _p(_f64(by)), _p(_f64(p)),
                                _p(u2), _p(v2), ctypes.byref(mu), ctypes.byref(mv))
    return u2, v2, mu.value, mv.value


def update_reference_literal(phi, jc, rho_minus, rho_plus, p, u, v, h, dt=1.0):
    phi = _f64(phi)
    ny, nx = phi.shape
    u2, v2 = _f64(u).copy(), _f64(v).copy()
    lib().orc_update_reference_literal(nx, ny, h, dt, _p(phi), _p(_f64(jc)), rho_minus, rho_plus, _p(_f64(p)),
                                       _p(u2), _p(v2))
    return u2, v2


# ----------------------------------------------------------------------------- multigrid mirror
# mode bit0: harmonic-normal coarse beta; bit1: operator-dependent interpolation;
# bit2: weights rounded to binary16 through binary32 (the GPU's FP32 hierarchy)
MG_DEFAULT_MODE = 3
MG_MODE_F32 = MG_DEFAULT_MODE | 4


def half_via_float(w):
    """w rounded to binary32, then to binary16 (nearest, ties to even), as a float."""
    return lib().orc_half_via_float(float(w))


class MG:
    """The V-cycle preconditioner (orc_mg_create): omega over-relaxes the red-black smoother on
    the finest level, omega_coarse (default: omega) on the levels below it."""

    def __init__(self, bx, by, max_levels=0, bottom_max=64, nu_pre=2, nu_post=2, mode=MG_DEFAULT_MODE, omega=1.0,
                 omega_coarse=None):
        self.bx, self.by = _f64(bx), _f64(by)
        ny, nx1 = self.bx.shape
        self.nx, self.ny = nx1 - 1, ny
        omc = omega if omega_coarse is None else omega_coarse
        self.h = lib().orc_mg_create(self.nx, self.ny, _p(self.bx), _p(self.by), max_levels, bottom_max,
                                     nu_pre, nu_post, mode, float(omega), float(omc))
        if not self.h:
            raise RuntimeError("invalid sweeps or omega, or a singular bottom operator")

    @property
    def levels(self):
        return lib().orc_mg_levels(self.h)

    def vcycle(self, rs):
        rs = _f64(rs)
        z = np.empty_like(rs)
        lib().orc_mg_vcycle(self.h, _p(rs), _p(z))
        return z

    def piece(self, what, level, x, n_out):
        """orc_mg_piece: 0 A'_level x, 1 P_level x, 2 P_level^T x, 3 bottom pseudo-inverse x (flat arrays)."""
        x = _f64(np.ravel(x))
        y = np.empty(n_out)
        if lib().orc_mg_piece(self.h, what, level, _p(x), _p(y)) != 0:
            raise ValueError("invalid level")
        return y

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mg_destroy(self.h)
            self.h = None


def mg_num_levels(nx, ny, max_levels=0, bottom_max=64):
    return lib().orc_mg_num_levels(nx, ny, max_levels, bottom_max)


def mg_pcg(bx, by, b, h, x0=None, rtol=1e-10, atol=0.0, max_iteration=10000, max_levels=0, nu=2, mode=None,
           omega=1.0, omega_coarse=None):
    mode = MG_DEFAULT_MODE if mode is None else mode
    b = _f64(b)
    ny, nx = b.shape
    x = np.zeros_like(b) if x0 is None else _f64(x0).copy()
    r2, rinf = _d(), _d()
    it = lib().orc_mg_pcg(nx, ny, h, _p(_f64(bx)), _p(_f64(by)), _p(x), _p(b), rtol, atol, max_iteration,
                          max_levels, nu, mode, float(omega), float(omega if omega_coarse is None else omega_coarse),
                          ctypes.byref(r2), ctypes.byref(rinf))
    if it < 0:
        raise RuntimeError("MG setup failed")
    return x, it, r2.value, rinf.value


# ----------------------------------------------------------------------------- whole projection
SOLVER_IC0, SOLVER_CG, SOLVER_JACOBI, SOLVER_MG = 0, 1, 2, 3


def project(phi, rho_minus, rho_plus, sigma, u, v, h, dt=1.0, solver=SOLVER_IC0, tol=1e-10, max_iteration=100000,
            threads=0):
    """PROJECTION_2D::Solve(dt) restated (PROJECTION_2D.h:235-242). Returns (p_pinned, u, v, stats)."""
    phi = _f64(phi)
    ny, nx = phi.shape
    u2, v2 = _f64(u).copy(), _f64(v).copy()
    p = np.empty((ny, nx))
    st = np.zeros(4)
    rc = lib().orc_project(nx, ny, h, dt, _p(phi), rho_minus, rho_plus, sigma, _p(u2), _p(v2), solver, tol,
                           max_iteration, threads, _p(p), _p(st))
    if rc != 0:
        raise FloatingPointError("curvature denominator zero")
    return p, u2, v2, dict(iterations=int(st[0]), residual=st[1], max_u=st[2], max_v=st[3])


def project_beta(phi, jc, bx, by, u, v, h, dt=1.0, rtol=1e-13):
    phi = _f64(phi)
    ny, nx = phi.shape
    u2, v2 = _f64(u).copy(), _f64(v).copy()
    p = np.empty((ny, nx))
    it = lib().orc_project_beta(nx, ny, h, dt, _p(phi), _p(_f64(jc)), _p(_f64(bx)), _p(_f64(by)), _p(u2), _p(v2),
                                rtol, _p(p))
    if it < 0:
        raise RuntimeError("MG setup failed")
    return p, u2, v2


def adjoint(phi, jc, bx, by, p, gu, gv, h, dt=1.0, rtol=1e-13):
    phi = _f64(phi)
    ny, nx = phi.shape
    du = np.empty((ny, nx + 1))
    dv = np.empty((ny + 1, nx))
    dbx = np.empty_like(du)
    dby = np.empty_like(dv)
    mu = np.empty((ny, nx))
    it = lib().orc_adjoint(nx, ny, h, dt, _p(phi), _p(_f64(jc)), _p(_f64(bx)), _p(_f64(by)), _p(_f64(p)),
                           _p(_f64(gu)), _p(_f64(gv)), rtol, _p(du), _p(dv), _p(dbx), _p(dby), _p(mu))
    if it < 0:
        raise RuntimeError("MG setup failed")
    return du, dv, dbx, dby, mu


def num_threads():
    return lib().orc_num_threads()
