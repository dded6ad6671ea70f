"""Pins of the oracle parts that agreement on the converged pressure cannot reach.

- The IC(0) factor (IncompleteCholeskyDecomposition, CSR_MATRIX.h:535-628): L L^T equals A
  on A's sparsity pattern, the defining property of the zero-fill incomplete Cholesky
  factor; and the factor equals a dense textbook IC(0) (left-looking Cholesky restricted
  to A's lower pattern) computed here.
- The preconditioner solve (MultiplicationByMinverse, PCG_METHOD.cpp:70-118) equals
  (L L^T)^-1 r by dense triangular solves.
- The Krylov loops (PCG_METHOD::PCGMethod, PCG_METHOD.cpp:176-264; CG_METHOD::CGMethod,
  CG_METHOD.cpp:13-73): iterate by iterate against a dense textbook restatement
  (Hestenes-Stiefel) on 8^2 grids, and the iteration count under the reference's absolute
  rule with thresholds placed between two iterations' residuals.
- The reference-literal velocity update (UpdateVelocityByPressureGradientVariableDensity,
  PROJECTION_2D.h:1231-1417, HeavisideFunction :1419-1441, DetermineDensityField
  :1482-1500): closed forms for a linear pressure, a face-aligned interface, an exactly zero
  level set and a face inside the mollification band.
"""
import numpy as np
import pytest
import scipy.linalg as sla

REF_PI = 3.141592  # MACROS.h:5


def _system(O, w):
    rho, inv = O.density(w.phi, w.rho_minus, w.rho_plus)
    div = O.divergence(w.u, w.v, w.h)
    jc = w.sigma * O.curvature(w.phi, w.h)
    A, b, _, _ = O.assemble_reference(w.phi, inv, div, w.h, jc=jc)
    return A, b


def _dense_ic0(A):
    """Textbook IC(0): Cholesky's recurrences with every entry outside A's pattern dropped."""
    n = A.shape[0]
    L = np.zeros_like(A)
    for k in range(n):
        L[k, k] = np.sqrt(A[k, k] - L[k, :k] @ L[k, :k])
        for i in range(k + 1, n):
            if A[i, k] != 0.0:
                L[i, k] = (A[i, k] - L[i, :k] @ L[k, :k]) / L[k, k]
    return L


def _sparse_L(Lb, Ll, Ld):
    ny, nx = Ld.shape
    n = nx * ny
    L = np.zeros((n, n))
    for j in range(ny):
        for i in range(nx):
            c = i + j * nx
            L[c, c] = Ld[j, i]
            if j > 0:
                L[c, c - nx] = Lb[j, i]
            if i > 0:
                L[c, c - 1] = Ll[j, i]
    return L


@pytest.fixture(scope="module", params=[("circle", 8, 8), ("drops64", 7, 5)])
def small_system(request, O, W):
    name, nx, ny = request.param
    w = W.make(name, max(nx, ny))
    if (nx, ny) != (w.n, w.n):   # a rectangular cut of the workload (GFM system of its own)
        w.phi = np.ascontiguousarray(w.phi[:ny, :nx])
        w.u = np.ascontiguousarray(w.u[:ny, :nx + 1])
        w.v = np.ascontiguousarray(w.v[:ny + 1, :nx])
        w.u[:, 0] = w.u[:, -1] = 0.0
        w.v[0, :] = w.v[-1, :] = 0.0
    A, b = _system(O, w)
    return A, b, nx, ny


def test_ic0_factor_equals_dense_textbook(O, small_system):
    A, b, nx, ny = small_system
    L = _sparse_L(*O.ic0_factor(A, nx, ny))
    Lt = _dense_ic0(A.to_dense())
    assert np.abs(L - Lt).max() <= 1e-13 * np.abs(Lt).max()


def test_ic0_factor_reproduces_A_on_its_pattern(O, W):
    """(L L^T)_ij = A_ij wherever A_ij != 0, on a 16^2 GFM matrix with the 1000:1 jump."""
    w = W.circle(16)
    A, _ = _system(O, w)
    Ad = A.to_dense()
    L = _sparse_L(*O.ic0_factor(A, 16, 16))
    M = L @ L.T
    pat = Ad != 0.0
    assert np.abs((M - Ad)[pat]).max() <= 1e-13 * np.abs(Ad).max()
    # the fill it drops sits exactly on the two anti-diagonal neighbours (c, c - nx + 1)
    fill = (M != 0.0) & ~pat
    r, c = np.nonzero(fill)
    assert set(np.abs(r - c).tolist()) == {15}


def test_ic0_solve_equals_dense_triangular_solves(O, small_system):
    A, b, nx, ny = small_system
    L = _dense_ic0(A.to_dense())
    r = np.random.default_rng(5).standard_normal(A.n)
    z = O.ic0_solve(A, r, nx, ny)
    zt = sla.solve_triangular(L.T, sla.solve_triangular(L, r, lower=True), lower=False)
    assert np.abs(z - zt).max() <= 1e-12 * np.abs(zt).max()


def _dense_pcg(Ad, b, minv, iters):
    """Hestenes-Stiefel PCG from x0 = 0; returns the iterates and r.z after each iteration."""
    x = np.zeros_like(b)
    r = b.copy()
    z = minv(r)
    p = z.copy()
    rz = r @ z
    xs, rzs = [], []
    for _ in range(iters):
        Ap = Ad @ p
        alpha = rz / (p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z = minv(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        xs.append(x.copy())
        rzs.append(rz)
    return xs, rzs


def _separated(seq, m):
    """True when the m-th value (1-based) lies well below every earlier one."""
    return seq[m - 1] < 0.5 * min(seq[: m - 1])


def _threshold_between(seq, m):
    """A tolerance whose square lies halfway (geometrically) between the m-th value of seq
    and the smallest earlier one, so the absolute rule first holds after iteration m."""
    return float(np.sqrt(np.sqrt(seq[m - 1] * min(seq[: m - 1]))))


def test_pcg_ic0_iterates_match_dense_textbook(O, small_system):
    A, b, nx, ny = small_system
    Ad = A.to_dense()
    L = _dense_ic0(Ad)
    minv = lambda r: sla.solve_triangular(L.T, sla.solve_triangular(L, r, lower=True), lower=False)
    K = 12
    xs, rzs = _dense_pcg(Ad, b, minv, K)
    for k in range(1, K + 1):
        # max_iteration = k - 1 runs k iterations: the cap is tested after the increment
        x, it, _ = O.pcg_ic0(A, b, nx, ny, tol=0.0, relative=False, max_iteration=k - 1)
        assert it == k
        assert np.linalg.norm(x - xs[k - 1]) <= 1e-11 * np.linalg.norm(xs[k - 1]), k
    # absolute rule r.z < tol^2 (PCG_METHOD.cpp:253-254) with the threshold between iterations
    stops = [m for m in range(2, K + 1) if _separated(rzs, m)]
    assert len(stops) >= 3
    for m in stops[:4]:
        _, it, _ = O.pcg_ic0(A, b, nx, ny, tol=_threshold_between(rzs, m), relative=False)
        assert it == m


def test_cg_iterates_match_dense_textbook(O, small_system):
    A, b, nx, ny = small_system
    Ad = A.to_dense()
    # Unpreconditioned CG on these 1000:1 matrices (condition ~1e5-1e6) loses orthogonality
    # after ~10 iterations, where rounding differences grow; the first 8 agree to 1e-12.
    K = 8
    xs, rrs = _dense_pcg(Ad, b, lambda r: r, K)
    for k in range(1, K + 1):
        x, it, _ = O.cg(A, b, tol=0.0, relative=False, max_iteration=k)
        assert it == k
        assert np.linalg.norm(x - xs[k - 1]) <= 1e-12 * np.linalg.norm(xs[k - 1]), k
    # absolute rule r.r < tol^2 (CG_METHOD.cpp:52): the loop ends in the m-th iteration
    stops = [m for m in range(2, K + 1) if _separated(rrs, m)]
    assert stops
    for m in stops[:4]:
        _, it, _ = O.cg(A, b, tol=_threshold_between(rrs, m), relative=False)
        assert it == m


def test_pcg_ic0_planted_factor_mistake_is_caught(O, small_system):
    """A transposed coupling (L_cb taken from the left neighbour) breaks L L^T = A."""
    A, b, nx, ny = small_system
    Lb, Ll, Ld = O.ic0_factor(A, nx, ny)
    Ad = A.to_dense()
    bad = _sparse_L(Ll.copy(), Lb.copy(), Ld)   # the planted mistake: bottom and left swapped
    M = bad @ bad.T
    assert np.abs((M - Ad)[Ad != 0.0]).max() > 1e-6 * np.abs(Ad).max()


# ----------------------------------------------------------------------------- literal update
def _grid(n):
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    return h, x


def _rand_faces(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (n, n + 1)), rng.uniform(-1, 1, (n + 1, n))


def test_literal_update_single_fluid_linear_pressure(O):
    """One fluid (phi > 0 everywhere), p = a x + b y: interior faces lose dt a / rho and
    dt b / rho; the left and bottom wall faces are updated against the ghost pressure 0
    (boundary_condition_array >= 0 there, PROJECTION_2D.h:1336-1339); the right and top
    wall faces are skipped (their cell is a ghost)."""
    n, rho, dt, a, bb = 8, 7.0, 0.3, 2.5, -1.25
    h, x = _grid(n)
    p = a * x[None, :] + bb * x[:, None]
    phi = np.ones((n, n))
    u, v = _rand_faces(n, 1)
    u2, v2 = O.update_reference_literal(phi, None, 1.0, rho, p, u, v, h, dt)
    tol = 1e-13
    np.testing.assert_allclose(u2[:, 1:n] - u[:, 1:n], -dt * a / rho, rtol=0, atol=tol)
    np.testing.assert_allclose(v2[1:n, :] - v[1:n, :], -dt * bb / rho, rtol=0, atol=tol)
    np.testing.assert_allclose(u2[:, 0] - u[:, 0], -dt * p[:, 0] / rho / h, rtol=0, atol=tol)
    np.testing.assert_allclose(v2[0, :] - v[0, :], -dt * p[0, :] / rho / h, rtol=0, atol=tol)
    assert np.array_equal(u2[:, n], u[:, n]) and np.array_equal(v2[n, :], v[n, :])


@pytest.mark.parametrize("flip", [False, True])
def test_literal_update_face_aligned_interface(O, flip):
    """phi = +-(x - 1/2): the interface sits on face 4, whose mean level set is 0, so the
    mollified density is (rho- + rho+)/2 (H(0) = 1/2).  With p = 0 and a constant jump a,
    only that face moves: by +dt a / (rho_half h) when its right cell is on the plus side
    (the '- jump' branch, PROJECTION_2D.h:1349-1353), by -dt a / (rho_half h) otherwise."""
    n, dt, a, rm, rp = 8, 0.5, 0.3, 1.0, 1000.0
    h, x = _grid(n)
    s = -1.0 if flip else 1.0
    phi = np.tile(s * (x - 0.5), (n, 1))
    jc = np.full((n, n), a)
    u, v = _rand_faces(n, 2)
    u2, v2 = O.update_reference_literal(phi, jc, rm, rp, np.zeros((n, n)), u, v, h, dt)
    du = u2 - u
    expect = s * dt * a / ((rm + rp) / 2) / h
    np.testing.assert_allclose(du[:, 4], expect, rtol=0, atol=1e-15)
    du[:, 4] = 0
    assert np.abs(du).max() == 0 and np.abs(v2 - v).max() == 0


def test_literal_update_zero_levelset_uses_the_right_cell(O):
    """phi = x - x_3 is exactly 0 at cell 3: both faces of that cell take the same-side branch
    (phi_c phi_l >= 0) with the density of their right cell, rho- on face 3 (phi <= 0 is the
    minus side, DetermineProjectionDensity) and rho+ on face 4."""
    n, dt, a, rm, rp = 8, 0.25, 1.5, 1.0, 1000.0
    h, x = _grid(n)
    phi = np.tile(x - x[3], (n, 1))
    p = np.tile(a * x, (n, 1))
    u, v = _rand_faces(n, 3)
    u2, _ = O.update_reference_literal(phi, None, rm, rp, p, u, v, h, dt)
    du = u2 - u
    np.testing.assert_allclose(du[:, 1:4], -dt * a / rm, rtol=0, atol=1e-14)
    np.testing.assert_allclose(du[:, 4:n], -dt * a / rp, rtol=0, atol=1e-15)


def test_literal_update_inside_the_mollification_band(O):
    """phi = 2 (x - 29/64) puts face 4's mean level set at eps/2 (eps = 1.5 h): there the
    reference's Heaviside is 1/2 + 1/4 + sin(PI/2)/(2 PI) with its PI = 3.141592, and the face
    moves by dt a / (rho_half h) with rho_half = rho- + (rho+ - rho-) H."""
    n, dt, a, rm, rp = 8, 0.5, 0.3, 1.0, 1000.0
    h, x = _grid(n)
    phi = np.tile(2 * (x - 29 / 64), (n, 1))
    assert phi[0, 3] == -1 / 32 and phi[0, 4] == 7 / 32
    jc = np.full((n, n), a)
    u, v = _rand_faces(n, 4)
    u2, _ = O.update_reference_literal(phi, jc, rm, rp, np.zeros((n, n)), u, v, h, dt)
    H = 0.75 + 0.5 / REF_PI
    rho_half = rm + (rp - rm) * H
    np.testing.assert_allclose((u2 - u)[:, 4], dt * a / rho_half / h, rtol=1e-12, atol=0)
    # with the true pi the value moves by ~1e-7 relative: the truncated constant is pinned
    H_true = 0.75 + 0.5 / np.pi
    assert abs((rm + (rp - rm) * H_true) / rho_half - 1) > 1e-8
