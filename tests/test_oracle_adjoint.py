"""The CPU reference gradient of the projection (the reference has no adjoint).

Pinned by central finite differences of the oracle's forward projection
(step 1e-6, FP64) along seeded random directions, for both the input face
velocities and the face coefficients beta, with loss 0.5 ||u_proj - u_t||^2.
Also the projection is an idempotent map (P(P u) = P u) whose velocity
Jacobian is self-adjoint for beta = 1: <P a, b> = <a, P b>.
"""
import numpy as np
import pytest


def _setup(O, W, n=32):
    w = W.make("drops64", n)
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    jc = w.sigma * O.curvature(w.phi, w.h)
    ut, vt = W.target_velocity(n)
    return w, bx, by, jc, ut, vt


def _interior(w, rng):
    du = rng.standard_normal(w.u.shape); du[:, 0] = du[:, -1] = 0
    dv = rng.standard_normal(w.v.shape); dv[0] = dv[-1] = 0
    return du, dv


def _loss(O, w, jc, bx, by, u, v, ut, vt):
    _, up, vp = O.project_beta(w.phi, jc, bx, by, u, v, w.h, rtol=1e-14)
    return 0.5 * (np.sum((up - ut) ** 2) + np.sum((vp - vt) ** 2))


def test_adjoint_matches_central_differences(O, W):
    w, bx, by, jc, ut, vt = _setup(O, W)
    p, up, vp = O.project_beta(w.phi, jc, bx, by, w.u, w.v, w.h, rtol=1e-14)
    gu, gv = up - ut, vp - vt
    gu[:, 0] = gu[:, -1] = 0
    gv[0] = gv[-1] = 0
    du, dv, dbx, dby, _ = O.adjoint(w.phi, jc, bx, by, p, gu, gv, w.h, rtol=1e-14)
    rng = np.random.default_rng(3)
    eps = 1e-6
    for _ in range(4):
        a, b = _interior(w, rng)
        fd = (_loss(O, w, jc, bx, by, w.u + eps * a, w.v + eps * b, ut, vt) -
              _loss(O, w, jc, bx, by, w.u - eps * a, w.v - eps * b, ut, vt)) / (2 * eps)
        an = np.sum(du * a) + np.sum(dv * b)
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
        # beta direction (relative perturbation, interior faces)
        sa, sb = _interior(w, rng)
        sa *= bx; sb *= by
        fd = (_loss(O, w, jc, bx + eps * sa, by + eps * sb, w.u, w.v, ut, vt) -
              _loss(O, w, jc, bx - eps * sa, by - eps * sb, w.u, w.v, ut, vt)) / (2 * eps)
        an = np.sum(dbx * sa) + np.sum(dby * sb)
        assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)


def test_projection_idempotent_and_self_adjoint(O, W):
    w, bx, by, jc, ut, vt = _setup(O, W)
    ones_x = (bx > 0) * 1.0
    ones_y = (by > 0) * 1.0
    zero = np.zeros_like(jc)
    _, u1, v1 = O.project_beta(w.phi, zero, ones_x, ones_y, w.u, w.v, w.h, rtol=1e-14)
    _, u2, v2 = O.project_beta(w.phi, zero, ones_x, ones_y, u1, v1, w.h, rtol=1e-14)
    assert np.abs(u2 - u1).max() < 1e-10 and np.abs(v2 - v1).max() < 1e-10
    rng = np.random.default_rng(4)
    a = _interior(w, rng)
    b = _interior(w, rng)
    _, pa_u, pa_v = O.project_beta(w.phi, zero, ones_x, ones_y, a[0], a[1], w.h, rtol=1e-14)
    _, pb_u, pb_v = O.project_beta(w.phi, zero, ones_x, ones_y, b[0], b[1], w.h, rtol=1e-14)
    lhs = np.sum(pa_u * b[0]) + np.sum(pa_v * b[1])
    rhs = np.sum(a[0] * pb_u) + np.sum(a[1] * pb_v)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
