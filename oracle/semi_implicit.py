"""Oracle (test infrastructure only): the semi-implicit viscosity solve of the oil-water
solver, restated in numpy/scipy from the reference.

- oil_water_x / oil_water_y: VISCOSITY_2D::ApplyViscosity, oil_water_simulation with
  semi_implicit_approach and dimensionless_form (VISCOSITY_2D.h:546-640): mollified
  dimensionless density and viscosity on the face and corner level-set averages, the five
  coefficient fields and the explicit term of each MAC component.
- build_literal: POISSON_SOLVER::BuildLinearSystemNodeForSemiImplicitViscosity
  (POISSON_SOLVER.cpp:620-712) as the reference runs it.  SetupBoundaryConditionsForVelocity
  (VISCOSITY_2D.h:647-663) only visits the component grid's own nodes, so the one-wide ghost
  ring of the boundary-condition field keeps its initial value 0 = BC_FULL; after
  AssignSequentialindicesToFullCells renumbers the interior, a ghost neighbour therefore reads
  as unknown 0 and its coefficient couples the row to unknown 0 (the first node).  The
  Dirichlet branches never fire.
- build_dirichlet: the same operator with the ghost neighbours taken as Dirichlet zero
  velocity, which is what SetupBoundaryConditionsForVelocity sets out to do.
- cg_literal: CG_METHOD::CGMethod (CG_METHOD.cpp:13-80) with its absolute stopping rule
  ||r||^2 < tolerance^2.

Only tests import this module.
"""
import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

REF_PI = 3.141592


def _ls_ghosts(phi):
    """The level set with one ring of linearly extrapolated ghosts (the reference's
    FillGhostCellsContinuousDerivativesFrom on water_levelset.arr)."""
    ny, nx = phi.shape
    p = np.zeros((ny + 2, nx + 2))
    p[1:-1, 1:-1] = phi
    p[1:-1, 0] = 2.0 * phi[:, 0] - phi[:, 1]
    p[1:-1, -1] = 2.0 * phi[:, -1] - phi[:, -2]
    p[0, 1:-1] = 2.0 * phi[0, :] - phi[1, :]
    p[-1, 1:-1] = 2.0 * phi[-1, :] - phi[-2, :]
    p[0, 0] = 2.0 * p[0, 1] - p[0, 2]
    p[0, -1] = 2.0 * p[0, -2] - p[0, -3]
    p[-1, 0] = 2.0 * p[-1, 1] - p[-1, 2]
    p[-1, -1] = 2.0 * p[-1, -2] - p[-1, -3]
    return p


def heaviside(phi, eps):
    """VISCOSITY_2D::HeavisideFunction (VISCOSITY_2D.h:667-686), PI = 3.141592."""
    oe, op = 1.0 / eps, 1.0 / REF_PI
    out = 0.5 + phi * oe * 0.5 + 0.5 * op * np.sin(REF_PI * phi * oe)
    out = np.where(phi < -eps, 0.0, out)
    return np.where(phi > eps, 1.0, out)


def _dimless(phi, eps, oil, water):
    """oil/water + (1 - oil/water) H, VISCOSITY_2D.h:776-781 and 817-822."""
    return oil / water + (1.0 - oil / water) * heaviside(phi, eps)


def _pad0(a):
    """One ring of zero ghosts around a MAC component."""
    out = np.zeros((a.shape[0] + 2, a.shape[1] + 2))
    out[1:-1, 1:-1] = a
    return out


def oil_water_x(phi, u, v, dx, dy, x_min, dt, Re, rho_oil, rho_water, mu_oil, mu_water, eps):
    """Coefficients c1..c5 and the explicit term of the x component (VISCOSITY_2D.h:553-591).
    x_min is the cell grid's x_min; the x-face grid's is x_min - dx/2."""
    ny, nx = phi.shape
    P = _ls_ghosts(phi)
    Vg = _pad0(v)
    i = np.arange(nx + 1)[None, :]
    pw = P[1:-1, 0:nx + 1]     # phi(i-1, j)
    pe = P[1:-1, 1:nx + 2]     # phi(i, j)
    pwn, pen = P[2:, 0:nx + 1], P[2:, 1:nx + 2]
    pws, pes = P[:-2, 0:nx + 1], P[:-2, 1:nx + 2]
    half = 0.5 * (pw + pe)
    cu = 0.25 * (pw + pe + pwn + pen)
    cd = 0.25 * (pw + pe + pws + pes)
    rho = _dimless(half, eps, rho_oil, rho_water)
    mu_h = _dimless(half, eps, mu_oil, mu_water)
    mu_cu, mu_cd = _dimless(cu, eps, mu_oil, mu_water), _dimless(cd, eps, mu_oil, mu_water)
    mu_r, mu_c = _dimless(pe, eps, mu_oil, mu_water), _dimless(pw, eps, mu_oil, mu_water)
    gx = x_min - 0.5 * dx
    one_over_r = 1.0 / (gx + (i + 0.5) * dx)
    one_over_r2 = one_over_r * one_over_r
    c1 = one_over_r * (gx + (i + 1) * dx) * 2.0 * dt / (Re * rho) * mu_r
    c2 = one_over_r * (gx + i * dx) * 2.0 * dt / (Re * rho) * mu_c
    c3 = dt / (Re * rho) * mu_cu
    c4 = dt / (Re * rho) * mu_cd
    c5 = 2.0 * dt / (Re * rho) * mu_h * one_over_r2
    odx, ody = 1.0 / dx, 1.0 / dy
    v_n, v_nw = Vg[2:-1, 1:nx + 2], Vg[2:-1, 0:nx + 1]    # v(i, j+1), v(i-1, j+1)
    v_s, v_sw = Vg[1:-2, 1:nx + 2], Vg[1:-2, 0:nx + 1]    # v(i, j), v(i-1, j)
    expl = u + dt / (Re * rho) * (ody * (mu_cu * odx * (v_n - v_nw) - mu_cd * odx * (v_s - v_sw)))
    return c1, c2, c3, c4, c5, expl


def oil_water_y(phi, u_old, u_new, v, dx, dy, x_min, dt, Re, rho_oil, rho_water, mu_oil, mu_water, eps):
    """Coefficients and explicit term of the y component (VISCOSITY_2D.h:593-640): the
    explicit term reads the x velocity before (u_old) and after (u_new) the x solve."""
    ny, nx = phi.shape
    P = _ls_ghosts(phi)
    Ug_old, Ug_new = _pad0(u_old), _pad0(u_new)
    i = np.arange(nx)[None, :]
    ps = P[0:ny + 1, 1:-1]      # phi(i, j-1)
    pn = P[1:ny + 2, 1:-1]      # phi(i, j)
    pse, pne = P[0:ny + 1, 2:], P[1:ny + 2, 2:]
    psw, pnw = P[0:ny + 1, :-2], P[1:ny + 2, :-2]
    half = 0.5 * (ps + pn)
    cu = 0.25 * (ps + pse + pn + pne)
    cl = 0.25 * (psw + ps + pnw + pn)
    rho = _dimless(half, eps, rho_oil, rho_water)
    mu_h = _dimless(half, eps, mu_oil, mu_water)
    mu_cu, mu_cl = _dimless(cu, eps, mu_oil, mu_water), _dimless(cl, eps, mu_oil, mu_water)
    mu_u, mu_c = _dimless(pn, eps, mu_oil, mu_water), _dimless(ps, eps, mu_oil, mu_water)
    r = x_min + i * dx
    one_over_r = 1.0 / r
    c1 = one_over_r * dt / (Re * rho) * mu_cu * (r + 0.5 * dx)
    c2 = one_over_r * dt / (Re * rho) * mu_cl * (r - 0.5 * dx)
    c3 = 2.0 * dt / (Re * rho) * mu_u
    c4 = 2.0 * dt / (Re * rho) * mu_c
    c5 = np.zeros_like(c1)
    odx, ody = 1.0 / dx, 1.0 / dy
    # rows j = 0..ny of the y-face grid; the x-face ghost row j = -1 and j = ny are zero
    uo_e, uo_es = Ug_old[1:ny + 2, 2:nx + 2], Ug_old[0:ny + 1, 2:nx + 2]     # u(i+1, j), u(i+1, j-1)
    uo_c, uo_cs = Ug_old[1:ny + 2, 1:nx + 1], Ug_old[0:ny + 1, 1:nx + 1]     # u(i, j), u(i, j-1)
    un_c, un_e = Ug_new[1:ny + 2, 1:nx + 1], Ug_new[1:ny + 2, 2:nx + 2]
    un_cs, un_es = Ug_new[0:ny + 1, 1:nx + 1], Ug_new[0:ny + 1, 2:nx + 2]
    expl = v + dt / (Re * rho) * ((odx * (mu_cu * ody * (uo_e - uo_es) - mu_cl * ody * (uo_c - uo_cs))) +
                                  one_over_r * mu_h * ody * (0.5 * (un_c + un_e) - 0.5 * (un_cs + un_es)))
    return c1, c2, c3, c4, c5, expl


def _build(c1, c2, c3, c4, c5, vel, expl, dx, dy, literal):
    ny, nx = vel.shape
    n = nx * ny
    idx = lambda i, j: j * nx + i
    inv_dxdx, inv_dydy = 1.0 / (dx * dx), 1.0 / (dy * dy)
    rows, cols, vals = [], [], []
    b = np.empty(n)
    for j in range(ny):
        for i in range(nx):
            k = idx(i, j)
            coef = 0.0
            for (ii, jj, c, inv) in ((i - 1, j, c2, inv_dxdx), (i + 1, j, c1, inv_dxdx),
                                     (i, j - 1, c4, inv_dydy), (i, j + 1, c3, inv_dydy)):
                inside = 0 <= ii < nx and 0 <= jj < ny
                if inside or literal:
                    coef += c[j, i]
                    rows.append(k)
                    cols.append(idx(ii, jj) if inside else 0)
                    vals.append(-c[j, i] * inv)
            if not literal:
                for (ii, jj, c) in ((i - 1, j, c2), (i + 1, j, c1), (i, j - 1, c4), (i, j + 1, c3)):
                    if not (0 <= ii < nx and 0 <= jj < ny):
                        coef += c[j, i]
            if coef == 0:
                coef = 1.0
            rows.append(k)
            cols.append(k)
            vals.append(1.0 + inv_dxdx * coef + c5[j, i])
            b[k] = expl[j, i]
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    return A, b, vel.reshape(-1).copy()


def build_literal(c1, c2, c3, c4, c5, vel, expl, dx, dy):
    """The reference's CSR system (A, b, x0) for one component, ghost neighbours = unknown 0."""
    return _build(c1, c2, c3, c4, c5, vel, expl, dx, dy, True)


def build_dirichlet(c1, c2, c3, c4, c5, vel, expl, dx, dy):
    """The same operator with zero-velocity Dirichlet ghosts."""
    return _build(c1, c2, c3, c4, c5, vel, expl, dx, dy, False)


def solve_direct(A, b):
    return spla.spsolve(A.tocsc(), b)


def cg_literal(A, b, x0, tolerance, max_iteration):
    """CG_METHOD::CGMethod restated: r = b - A x; p = r; stop when r.r < tol^2."""
    x = x0.copy()
    r = b - A @ x
    p = r.copy()
    res_old = r @ p
    it = 0
    while it < max_iteration:
        Ap = A @ p
        alpha = res_old / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        res_new = r @ r
        if res_new < tolerance * tolerance:
            break
        p = r + (res_new / res_old) * p
        res_old = res_new
        it += 1
    return x, it, math.sqrt(res_new)
