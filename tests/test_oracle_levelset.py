"""Pins of the oracle's level-set re-initialisation (Sussman, WENO5) and normals.

The reference holds no vectors for these; what the mathematics fixes:
- WENO5 one-sided derivatives are fifth-order accurate on a smooth field
  without critical points (a wrong stencil coefficient or smoothness weight
  drops the order);
- the reference's branch selection of the Hamiltonian equals the textbook
  Godunov flux (max/min form, Sussman-Smereka-Osher 1994) on every input;
- a signed-distance plane is a fixed point away from the walls' ghost band;
- re-initialising a scaled distance restores |grad phi| = 1 near the interface
  without moving it, and never flips a sign;
- unit normals of a circle are radial.
"""
import numpy as np
import pytest


def _grid(n):
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(x, x)
    return h, X, Y


def test_weno5_fifth_order(O):
    errs = []
    for n in (32, 64):
        h, X, Y = _grid(n)
        f = np.exp(0.7 * X + 0.4 * Y)
        dxm, dxp, dym, dyp = O.weno5_derivatives(f, h)
        band = (slice(6, n - 6), slice(6, n - 6))
        ex, ey = 0.7 * f[band], 0.4 * f[band]
        errs.append(max(np.abs(dxm[band] - ex).max(), np.abs(dxp[band] - ex).max(),
                        np.abs(dym[band] - ey).max(), np.abs(dyp[band] - ey).max()))
    assert errs[1] < 2e-7
    assert errs[0] / errs[1] > 2 ** 4.5, errs


def _godunov_textbook(s, a, b, c, d):
    """|grad phi| by the Godunov flux: a, b = D-x, D+x; c, d = D-y, D+y."""
    pos = s > 0
    hx = np.where(pos, np.maximum(np.maximum(a, 0) ** 2, np.minimum(b, 0) ** 2),
                  np.maximum(np.minimum(a, 0) ** 2, np.maximum(b, 0) ** 2))
    hy = np.where(pos, np.maximum(np.maximum(c, 0) ** 2, np.minimum(d, 0) ** 2),
                  np.maximum(np.minimum(c, 0) ** 2, np.maximum(d, 0) ** 2))
    return np.sqrt(hx + hy)


@pytest.mark.parametrize("seed", [0, 1])
def test_reinit_step_is_godunov(O, seed):
    n = 48
    h, X, Y = _grid(n)
    rng = np.random.default_rng(seed)
    phi = (np.hypot(X - 0.45, Y - 0.55) - 0.3) * (1.5 + 0.5 * np.sin(5 * X)) + 0.02 * rng.standard_normal((n, n))
    eps = 1e-6
    a, b, c, d = O.weno5_derivatives(phi, h, eps)
    s = phi / np.sqrt(phi ** 2 + h * h)
    dt = h / 5.0
    ref = phi - dt * s * (_godunov_textbook(s, a, b, c, d) - 1)
    ref = np.where(s * ref < 0, np.where(s < 0, -eps, eps), ref)
    got = O.reinit_sussman(phi, h, iterations=1, scaling=5.0, epsilon=eps)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(phi).max()
    assert np.array_equal(np.sign(got), np.sign(phi))


def test_reinit_plane_is_fixed_point(O):
    n = 64
    h, X, Y = _grid(n)
    phi = X - 0.37
    got = O.reinit_sussman(phi, h, iterations=25)
    # the walls' ghost rule perturbs a band of 3 (stencil) + 25 * 1/5 (travel) cells
    assert np.abs(got - phi)[:, 9:n - 9].max() <= 1e-12


def test_reinit_restores_distance(O):
    n = 128
    h, X, Y = _grid(n)
    d = np.hypot(X - 0.5, Y - 0.5) - 0.25
    got = O.reinit_sussman(2.0 * d, h, iterations=25)
    near = np.abs(d) < 3 * h
    assert np.abs(got - d)[near].max() < 0.15 * h
    gy, gx = np.gradient(got, h)
    band = (np.abs(d) < 3 * h) & (np.abs(d) > h)
    assert np.abs(np.hypot(gx, gy)[band] - 1).max() < 0.05
    assert np.array_equal(np.sign(got), np.sign(d))
    # scaled plane: slope 2 relaxes to 1 around the interface, which stays put
    phi = 2.0 * (X - 0.4)
    got = O.reinit_sussman(phi, h, iterations=25)
    mid = slice(9, n - 9)
    zero = np.argmin(np.abs(got[n // 2, mid])) + 9
    assert abs((zero + 0.5) * h - 0.4) <= h


def test_normals_of_circle_are_radial(O):
    errs = []
    for n in (48, 96):
        h, X, Y = _grid(n)
        r = np.hypot(X - 0.5, Y - 0.5)
        nx, ny = O.normals(r - 0.3, h)
        mask = (r > 0.15) & (r < 0.4)
        errs.append(max(np.abs(nx - (X - 0.5) / r)[mask].max(), np.abs(ny - (Y - 0.5) / r)[mask].max()))
        assert np.abs(np.hypot(nx, ny) - 1).max() < 1e-14
    assert errs[1] < 1e-3 and errs[0] / errs[1] > 3.0, errs  # second order


def test_advection_linear_field_closed_form(O):
    """phi = a x + b y under a uniform velocity: phi - dt (u a + v b) exactly, away
    from the walls' ghost band (WENO5 is exact on linear data)."""
    n = 48
    h, X, Y = _grid(n)
    phi = 0.7 * X - 0.3 * Y + 0.1
    for uu, vv in ((0.5, 0.25), (-0.5, 0.25), (0.5, -0.25), (-0.5, -0.25)):
        u = np.full((n, n + 1), uu)
        v = np.full((n + 1, n), vv)
        got = O.advect_levelset(phi, u, v, h, 0.1 * h)
        ref = phi - 0.1 * h * (uu * 0.7 + vv * -0.3)
        band = (slice(3, n - 3), slice(3, n - 3))
        assert np.abs(got - ref)[band].max() <= 1e-15


def test_advection_upwind_side_and_face_average(O):
    """On a varying face velocity the stage equals -dt (u_c D u + v_c D v) with the
    cell velocity the mean of its faces and the upwind one-sided WENO5 derivative."""
    n = 40
    h, X, Y = _grid(n)
    phi = np.sin(3 * X) * np.cos(2 * Y) + 0.3 * X
    xf = np.arange(n + 1) * h
    yf = (np.arange(n) + 0.5) * h
    u = np.sin(5 * xf)[None, :] * np.cos(4 * yf)[:, None]                 # (n, n+1)
    v = np.cos(3 * ((np.arange(n) + 0.5) * h))[None, :] * np.sin(6 * np.arange(n + 1) * h)[:, None] - 0.2  # (n+1, n)
    dt = 0.3 * h
    a, b, c, d = O.weno5_derivatives(phi, h)
    uc = 0.5 * (u[:, 1:] + u[:, :-1])
    vc = 0.5 * (v[1:, :] + v[:-1, :])
    ref = phi - dt * (uc * np.where(uc > 0, a, b) + vc * np.where(vc > 0, c, d))
    got = O.advect_levelset(phi, u, v, h, dt)
    assert np.abs(got - ref).max() <= 1e-15 * np.abs(phi).max() + 1e-16
    assert (uc > 0).any() and (uc < 0).any() and (vc > 0).any() and (vc < 0).any()


def test_velocity_advection_closed_forms(O):
    """ENO3 is exact on quadratics; under the uniform field u = U, v = V the stage is
    u - dt (U u_x + V u_y) exactly, away from the ghost band; a shear u = y, v = 0
    leaves u unchanged (u_x = 0, v = 0) and moves v by nothing (v = 0 everywhere)."""
    n = 40
    h = 1.0 / n
    xf = np.arange(n + 1) * h
    yc = (np.arange(n) + 0.5) * h
    XU, YU = np.meshgrid(xf, yc)                      # x-faces (n, n+1)
    for U, V in ((0.7, 0.3), (-0.7, 0.3), (0.7, -0.3), (-0.7, -0.3)):
        u = np.full((n, n + 1), U)
        v = np.full((n + 1, n), V)
        uo, vo = O.advect_velocity(u, v, h, 0.2 * h)
        assert np.abs(uo - U).max() <= 1e-15 and np.abs(vo - V).max() <= 1e-15
    # a quadratic x-component advected by a uniform field given through v only
    u = 0.3 + 0.2 * XU + 0.1 * YU ** 2
    v = np.full((n + 1, n), 0.4)
    uo, _ = O.advect_velocity(u, v, h, 0.2 * h)
    ux, uy = 0.2, 0.2 * YU
    ref = u - 0.2 * h * (u * ux + 0.4 * uy)
    band = (slice(4, n - 4), slice(4, n - 3))
    assert np.abs(uo - ref)[band].max() <= 1e-13


def test_velocity_advection_eno_is_third_order(O):
    """Derivative error of the stage on a smooth field: (u - u_out) / dt against
    u u_x + v_c u_y at the x-faces, third order under refinement."""
    errs = []
    for n in (32, 64):
        h = 1.0 / n
        xf = np.arange(n + 1) * h
        yc = (np.arange(n) + 0.5) * h
        XU, YU = np.meshgrid(xf, yc)
        f = lambda x, y: 0.5 + 0.2 * np.sin(2 * x + 1) * np.cos(3 * y)
        u = f(XU, YU)
        XV, YV = np.meshgrid((np.arange(n) + 0.5) * h, np.arange(n + 1) * h)
        v = np.full_like(XV, 0.3)
        dt = 1e-3 * h
        uo, _ = O.advect_velocity(u, v, h, dt)
        ux = 0.2 * 2 * np.cos(2 * XU + 1) * np.cos(3 * YU)
        uy = -0.2 * 3 * np.sin(2 * XU + 1) * np.sin(3 * YU)
        band = (slice(5, n - 5), slice(5, n - 4))
        errs.append(np.abs((u - uo) / dt - (u * ux + 0.3 * uy))[band].max())
    assert errs[0] / errs[1] > 2 ** 2.5, errs
