"""Pins of the CPU oracle against what the mathematics fixes (the reference ships
no golden vectors; its only result pins are closed-form experiments).

- LFK test 3 (POISSON_EQUATION_TEST.h:315-343, 409-423, 490-507, 608-624;
  Poisson_Solver_Test.cpp:11-86): the reference-literal GFM assembly with jumps
  in the solution and in the flux converges to the closed form under refinement.
  A flipped jump sign, a dropped b-term or a transposed beta breaks convergence.
- Balanced force: u = v = 0 with a constant jump a gives p_out - p_in = a
  exactly and u_proj = 0.
- Curvature: circle kappa = -1/r and the Rayleigh-Taylor closed form, O(h^2).
"""
import math

import numpy as np
import pytest


def _lfk3(O, n, b_terms=True, flip_a=False):
    """Poisson_Solver_Test.cpp:11-86 on an n x n node grid over [0,1]^2 with Dirichlet ghosts."""
    dx = 1.0 / (n - 1)
    x = np.arange(n) * dx
    X, Y = np.meshgrid(x, x)
    phi = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2) - 0.25
    g = np.exp(-X ** 2 - Y ** 2)
    f = np.where(phi <= 0, 8 * (X ** 2 + Y ** 2 - 1) * g, 0.0)
    a = -g * (-1 if flip_a else 1)
    b = 8 * (2 * X ** 2 + 2 * Y ** 2 - X - Y) * g if b_terms else None
    beta_arg = np.where(phi <= 0, 0.5, 1.0)      # passed as "density"
    variable = 1.0 / beta_arg                     # POISSON_SOLVER.cpp:131-138
    bc = np.full((n + 2, n + 2), O.BC_DIR, np.int32)
    bc[1:-1, 1:-1] = O.BC_FULL
    p_pad = np.zeros((n + 2, n + 2))              # true solution is 0 outside the circle
    nrm = O.normals(phi, dx)
    A, rhs, x0, row_of = O.assemble_reference(phi, variable, f, dx, jc=a, jcd=b, normals_xy=nrm, bc_pad=bc,
                                               p_pad=p_pad)
    sol, it, _ = O.pcg_ic0(A, rhs, n, n, tol=1e-13)
    u = sol.reshape(n, n)
    exact = np.where(phi <= 0, g, 0.0)
    return math.sqrt(np.sum((u - exact) ** 2) * dx * dx), A, rhs


def test_lfk3_converges_under_refinement(O):
    errs = [_lfk3(O, n)[0] for n in (41, 81, 161)]
    # LFK ghost-fluid method: first order or better in the solution
    assert errs[0] / errs[1] > 1.6 and errs[1] / errs[2] > 1.6, errs
    assert errs[2] < 2e-3, errs


def test_lfk3_detects_planted_sign_error(O):
    good = _lfk3(O, 81)[0]
    bad_a = _lfk3(O, 81, flip_a=True)[0]
    no_b = _lfk3(O, 81, b_terms=False)[0]
    assert bad_a > 20 * good and no_b > 5 * good, (good, bad_a, no_b)


def test_reference_matrix_invariants(O, W):
    w = W.circle(32)
    rho, inv = O.density(w.phi, w.rho_minus, w.rho_plus)
    div = O.divergence(w.u, w.v, w.h)
    jc = w.sigma * O.curvature(w.phi, w.h)
    A, b, x0, row_of = O.assemble_reference(w.phi, inv, div, w.h, jc=jc)
    D = A.to_dense()
    assert np.array_equal(D, D.T)                       # symmetric (bitwise)
    ones = D @ np.ones(A.n)
    assert np.all(np.abs(ones[1:]) <= 1e-9 * np.abs(D).max())   # A 1 = 0 except the pinned row
    assert ones[0] > 0                                   # the pin (POISSON_SOLVER.cpp:789-807)
    assert np.all(np.linalg.eigvalsh(D) > 0)             # SPD after pinning
    # beta on interior faces equals the closed form 1/rho_f (rho_f = theta-weighted)
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    i, j = 10, 7
    pL, pR = w.phi[j, i - 1], w.phi[j, i]
    rL = w.rho_minus if pL <= 0 else w.rho_plus
    rR = w.rho_minus if pR <= 0 else w.rho_plus
    rho_f = (abs(pL) * rL + abs(pR) * rR) / (abs(pL) + abs(pR))
    assert abs(bx[j, i] - 1.0 / rho_f) <= 1e-15 / rho_f
    # off-diagonal entry of the CSR equals -beta/h^2
    r, c = row_of[j * 32 + i], row_of[j * 32 + i - 1]
    assert abs(D[r, c] + bx[j, i] / w.h ** 2) <= 1e-12 * bx[j, i] / w.h ** 2


def test_jump_rhs_is_antisymmetric_and_matches_consistent_form(O, W):
    w = W.make("drops64", 128)
    rho, inv = O.density(w.phi, w.rho_minus, w.rho_plus)
    jc = w.sigma * O.curvature(w.phi, w.h)
    zero = np.zeros((w.n, w.n))
    _, b_ref, _, _ = O.assemble_reference(w.phi, inv, zero, w.h, jc=jc)
    assert abs(b_ref.sum()) <= 1e-12 * np.abs(b_ref).sum()          # sum_c F_c = 0
    assert np.abs(b_ref).max() > 0
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    div = O.divergence(w.u, w.v, w.h)
    _, b_full, _, _ = O.assemble_reference(w.phi, inv, div, w.h, jc=jc)
    b_cons = O.rhs_consistent(w.phi, jc, bx, by, w.u, w.v, w.h).ravel()
    assert np.abs(b_full - b_cons).max() <= 1e-12 * np.abs(b_full).max()


def test_balanced_force_static_drop(O, W):
    """Constant jump a on a circle, u = v = 0: p_out - p_in = a, u_proj = 0."""
    w = W.circle(64)
    a = 0.37
    jc = np.full((w.n, w.n), a)
    rho, inv = O.density(w.phi, w.rho_minus, w.rho_plus)
    zero = np.zeros((w.n, w.n))
    A, b, _, _ = O.assemble_reference(w.phi, inv, zero, w.h, jc=jc)
    p, it, _ = O.pcg_ic0(A, b, w.n, w.n, tol=1e-14)
    p = p.reshape(w.n, w.n)
    pin, pout = p[w.phi <= 0], p[w.phi > 0]
    assert np.ptp(pin) < 1e-9 and np.ptp(pout) < 1e-9
    assert abs((pout.mean() - pin.mean()) - a) < 1e-9
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    u, v, _, _ = O.update_consistent(w.phi, jc, bx, by, p, np.zeros_like(w.u), np.zeros_like(w.v), w.h)
    assert np.abs(u).max() < 1e-9 and np.abs(v).max() < 1e-9


@pytest.mark.parametrize("R", [0.25, 0.3])
def test_curvature_circle(O, W, R):
    errs = []
    for n in (64, 128, 256):
        h = 1.0 / n
        x = (np.arange(n) + 0.5) * h
        X, Y = np.meshgrid(x, x)
        r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
        phi = r - R
        k = O.curvature(phi, h)
        band = np.abs(phi) < 2 * h
        errs.append(np.abs(k[band] + 1.0 / r[band]).max())
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0, errs     # O(h^2)
    assert errs[2] < 1e-3


def test_curvature_rayleigh_taylor(O):
    errs = []
    for n in (64, 128, 256):
        h = 1.0 / n
        x = (np.arange(n) + 0.5) * h
        X, Y = np.meshgrid(x, x)
        A, kk = 0.01, 8 * math.pi
        phi = Y - 0.5 - A * np.cos(kk * X)
        exact = -A * kk ** 2 * np.cos(kk * X) / (1 + (A * kk * np.sin(kk * X)) ** 2) ** 1.5
        k = O.curvature(phi, h)
        inner = (slice(2, -2), slice(2, -2))
        errs.append(np.abs(k[inner] - exact[inner]).max())
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0, errs


def test_curvature_clamp_and_linear(O):
    n, h = 32, 1.0 / 32
    x = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(x, x)
    k = O.curvature(0.3 * X - 0.7 * Y + 0.1, h)
    assert np.abs(k).max() < 1e-9                      # linear phi: zero everywhere, corners included
    tiny = np.sqrt((X - 0.5 - h / 2) ** 2 + (Y - 0.5 - h / 2) ** 2) - 0.1 * h
    k = O.curvature(tiny, h)
    assert np.abs(k).max() <= 1.0 / h and np.isclose(np.abs(k).max(), 1.0 / h)
