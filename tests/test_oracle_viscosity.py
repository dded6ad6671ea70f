"""Pins of the oracle's explicit air-water viscosity update (VISCOSITY_2D.h:460-548).

What the mathematics fixes:
- rigid rotation carries no viscous stress: the update vanishes in the interior
  for any density and viscosity distribution;
- single fluid, u = a y^2 + x^2-type fields: the update is dt mu/rho times the
  exact second derivatives (the discrete operator is exact on quadratics);
- the mollified density at a face follows H of the face's mean level set:
  H = 0 below -eps, 1 above eps, 1/2 on the interface, and H(-phi) = 1 - H(phi)
  (swapping the fluids and the sign of phi leaves the update unchanged).
"""
import numpy as np


def _faces(n):
    h = 1.0 / n
    xf, yc = np.arange(n + 1) * h, (np.arange(n) + 0.5) * h
    XU, YU = np.meshgrid(xf, yc)                    # x-faces (n, n+1)
    xc, yf = (np.arange(n) + 0.5) * h, np.arange(n + 1) * h
    XV, YV = np.meshgrid(xc, yf)                    # y-faces (n+1, n)
    X, Y = np.meshgrid(xc, yc)
    return h, XU, YU, XV, YV, X, Y


def test_rigid_rotation_has_no_viscous_update(O):
    n = 48
    h, XU, YU, XV, YV, X, Y = _faces(n)
    u, v = -(YU - 0.5), (XV - 0.5)
    phi = np.hypot(X - 0.45, Y - 0.55) - 0.2
    uo, vo = O.apply_viscosity(phi, u, v, h, 0.01, 1.0, 1000.0, 1.8e-5, 1.1e-3)
    inner_u = (slice(1, n - 1), slice(2, n - 1))
    inner_v = (slice(2, n - 1), slice(1, n - 1))
    assert np.abs(uo - u)[inner_u].max() <= 1e-16 and np.abs(vo - v)[inner_v].max() <= 1e-16


def test_single_fluid_quadratics(O):
    n = 40
    h, XU, YU, XV, YV, X, Y = _faces(n)
    phi = -np.ones((n, n))                          # all minus side
    dt, rho, mu = 0.02, 2.0, 0.3
    a, b, c, d = 0.7, -0.4, 0.5, 0.9
    u = a * XU ** 2 + b * YU ** 2                   # u_xx = 2a, u_yy = 2b
    v = c * XV ** 2 + d * YV ** 2                   # v_xx = 2c, v_yy = 2d
    uo, vo = O.apply_viscosity(phi, u, v, h, dt, rho, 5.0, mu, 7.0)
    # x-face: 2 mu u_xx + mu (u_yy + v_xy) ; y-face: 2 mu v_yy + mu (u_xy + v_xx)
    ru = dt * mu / rho * (2 * 2 * a + 2 * b)
    rv = dt * mu / rho * (2 * 2 * d + 2 * c)
    inner_u = (slice(1, n - 1), slice(2, n - 1))
    inner_v = (slice(2, n - 1), slice(1, n - 1))
    assert np.abs((uo - u)[inner_u] - ru).max() <= 1e-12
    assert np.abs((vo - v)[inner_v] - rv).max() <= 1e-12


def test_mollified_density_and_swap_symmetry(O):
    n = 64
    h, XU, YU, XV, YV, X, Y = _faces(n)
    phi = X - 0.5 + 0.3 * h                          # interface between cell centres
    dt, mu = 1e-3, 0.1
    u, v = XU ** 2, np.zeros_like(XV)               # update = 4 dt mu / rho_face at x-faces
    rm, rp = 1.0, 50.0
    uo, _ = O.apply_viscosity(phi, u, v, h, dt, rm, rp, mu, mu)
    rho_face = 4 * dt * mu / (uo - u)[n // 2, 2:n - 1]
    ph = 0.5 * (phi[n // 2, 1:n - 2] + phi[n // 2, 2:n - 1])  # face i has cells i-1, i
    eps = 1.5 * h
    assert np.allclose(rho_face[ph < -eps], rm, rtol=1e-12)
    assert np.allclose(rho_face[ph > eps], rp, rtol=1e-12)
    mid = np.abs(ph) <= eps
    assert mid.sum() >= 2 and np.all(np.diff(rho_face[mid]) > 0)
    # swapping the fluids and the sign of phi leaves the update unchanged
    phi2 = np.hypot(X - 0.5, Y - 0.5) - 0.25
    rng = np.random.default_rng(2)
    u2, v2 = rng.uniform(-1, 1, XU.shape), rng.uniform(-1, 1, XV.shape)
    a = O.apply_viscosity(phi2, u2, v2, h, 1e-3, 1.0, 1000.0, 2e-5, 1e-3)
    b = O.apply_viscosity(-phi2, u2, v2, h, 1e-3, 1000.0, 1.0, 1e-3, 2e-5)
    for x, y in zip(a, b):
        assert np.abs(x - y).max() <= 1e-12 * np.abs(x).max()
