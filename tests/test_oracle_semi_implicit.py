"""The oracle's semi-implicit viscosity pieces (oracle/semi_implicit.py), pinned on the CPU.

- The literal system (BuildLinearSystemNodeForSemiImplicitViscosity as the reference runs it,
  POISSON_SOLVER.cpp:620-712): with square cells every row sums to 1 + c5, because the
  out-of-grid couplings land on node 0; with all coefficients zero the reference's
  `coef_ijk == 0 -> 1` gives the diagonal 1 + 1/dx^2.
- The Dirichlet variant: with constant coefficients it is I - c Laplacian with zero walls, so
  the discrete sine modes are eigenvectors with eigenvalue 1 + c (4/h^2)(sin^2 + sin^2).
- The oil-water coefficients (VISCOSITY_2D.h:553-640) in a single fluid (H = 1, dimensionless
  density and viscosity 1) against their hand-written closed forms.
- The restated CG converges to the direct solution where the reference's CG converges.
"""
import numpy as np
import pytest


def _coef(shape, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.1, 2.0, shape) for _ in range(4)] + [rng.uniform(0.0, 0.5, shape)]


def test_literal_rows_sum_to_one_plus_c5(O):
    from semi_implicit import build_literal
    ny, nx, h = 9, 12, 0.1
    c = _coef((ny, nx), 1)
    A, b, x0 = build_literal(*c, np.zeros((ny, nx)), np.ones((ny, nx)), h, h)
    rows = np.asarray(A.sum(axis=1)).ravel()
    assert np.allclose(rows, 1.0 + c[4].ravel(), rtol=1e-12, atol=1e-12)
    # the couplings that leave the grid go to column 0
    A0 = A.tocsc()[:, 0].toarray().ravel()
    border = np.zeros((ny, nx), bool)
    border[0, :] = border[-1, :] = True
    border[:, 0] = border[:, -1] = True
    assert np.all(A0.reshape(ny, nx)[~border] == 0)
    assert np.all(A0.reshape(ny, nx)[border][1:] < 0)


def test_zero_coefficients_give_the_reference_diagonal(O):
    from semi_implicit import build_literal
    ny, nx, h = 5, 6, 0.25
    z = [np.zeros((ny, nx))] * 5
    A, b, _ = build_literal(*z, np.zeros((ny, nx)), np.ones((ny, nx)), h, h)
    assert np.allclose(A.toarray(), np.eye(nx * ny) * (1 + 1 / h ** 2))


def test_dirichlet_sine_modes_are_eigenvectors(O):
    from semi_implicit import build_dirichlet
    ny, nx, h, c = 10, 14, 0.07, 0.3
    cs = [np.full((ny, nx), c)] * 4 + [np.zeros((ny, nx))]
    A, _, _ = build_dirichlet(*cs, np.zeros((ny, nx)), np.zeros((ny, nx)), h, h)
    for (kx, ky) in ((1, 1), (3, 2), (nx, ny)):
        x = np.sin(np.pi * kx * (np.arange(nx) + 1) / (nx + 1))[None, :] * \
            np.sin(np.pi * ky * (np.arange(ny) + 1) / (ny + 1))[:, None]
        lam = 1 + c * 4 / h ** 2 * (np.sin(np.pi * kx / (2 * (nx + 1))) ** 2 + np.sin(np.pi * ky / (2 * (ny + 1))) ** 2)
        assert np.allclose(A @ x.ravel(), lam * x.ravel(), rtol=1e-12, atol=1e-12 * lam)


def test_oil_water_single_fluid_closed_forms(O):
    from semi_implicit import oil_water_x, oil_water_y
    ny, nx, dx, dy, x0, dt, Re = 6, 5, 0.2, 0.25, 0.3, 1e-3, 0.9
    phi = np.full((ny, nx), 10.0)  # all water: H = 1, density and viscosity 1
    rng = np.random.default_rng(2)
    u, v = rng.standard_normal((ny, nx + 1)), rng.standard_normal((ny + 1, nx))
    kw = dict(dx=dx, dy=dy, x_min=x0, dt=dt, Re=Re, rho_oil=0.9, rho_water=1.0, mu_oil=6.0, mu_water=0.01, eps=0.3)
    c1, c2, c3, c4, c5, ex = oil_water_x(phi, u, v, **kw)
    i = np.arange(nx + 1)[None, :]
    r = x0 + i * dx                  # the x-face grid node i + 1/2 cell: x0 - dx/2 + (i + 1/2) dx
    assert np.allclose(c1, (x0 + (i + 0.5) * dx) / r * 2 * dt / Re, rtol=1e-13)
    assert np.allclose(c2, (x0 + (i - 0.5) * dx) / r * 2 * dt / Re, rtol=1e-13)
    assert np.allclose(c3, dt / Re) and np.allclose(c4, dt / Re)
    assert np.allclose(c5, 2 * dt / Re / r ** 2, rtol=1e-13)
    vg = np.pad(v, 1)
    dvdx_n = (vg[2:-1, 1:nx + 2] - vg[2:-1, 0:nx + 1]) / dx
    dvdx_s = (vg[1:-2, 1:nx + 2] - vg[1:-2, 0:nx + 1]) / dx
    assert np.allclose(ex, u + dt / Re * (dvdx_n - dvdx_s) / dy, rtol=1e-13, atol=1e-15)
    c1, c2, c3, c4, c5, ex = oil_water_y(phi, u, u, v, **kw)
    i = np.arange(nx)[None, :]
    r = x0 + i * dx
    assert np.allclose(c1, dt / Re * (r + dx / 2) / r, rtol=1e-13)
    assert np.allclose(c2, dt / Re * (r - dx / 2) / r, rtol=1e-13)
    assert np.allclose(c3, 2 * dt / Re) and np.allclose(c4, 2 * dt / Re) and np.all(c5 == 0)


def test_literal_cg_converges_to_the_direct_solution(O):
    from semi_implicit import build_literal, cg_literal, oil_water_x, solve_direct
    nx, ny = 16, 24
    dx, dy = 1.26 / nx, 3.14 / ny
    x0 = 0.02 + 0.5 * dx
    xs = x0 + np.arange(nx) * dx
    phi = np.repeat((xs - 0.7)[None, :], ny, axis=0)
    rng = np.random.default_rng(3)
    u, v = rng.uniform(-1, 1, (ny, nx + 1)), rng.uniform(-1, 1, (ny + 1, nx))
    c = oil_water_x(phi, u, v, dx, dy, x0, 1e-5, 0.94983, 0.905, 0.995, 6.01, 0.01, 1.5 * dx)
    A, b, xi = build_literal(*c[:5], u, c[5], dx, dy)
    xd = solve_direct(A, b)
    xc, it, res = cg_literal(A, b, xi, 1e-10, 10000)
    assert res < 1e-10 and it < 10000
    assert np.abs(xc - xd).max() <= 1e-9 * np.abs(xd).max()
    assert np.abs(A @ xd - b).max() <= 1e-12 * np.abs(b).max()
