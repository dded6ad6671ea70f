"""Oracle solvers and the multigrid mirror, pinned by agreement and by properties.

The converged pressure is fixed by the linear system up to a constant, so the
restatements of CG (CG_METHOD.cpp:13-73), IC(0)-PCG (PCG_METHOD.cpp:176-264),
Jacobi-PCG (PCG_METHOD.cpp:266-383) and the MG-PCG of this library must agree
on the mean-zero pressure.  The V-cycle must be a symmetric positive-definite
linear operator that reproduces constants, and must converge in a bounded
number of iterations on the constant-coefficient problem.
"""
import numpy as np
import pytest


def _system(O, w):
    rho, inv = O.density(w.phi, w.rho_minus, w.rho_plus)
    div = O.divergence(w.u, w.v, w.h)
    jc = w.sigma * O.curvature(w.phi, w.h)
    return O.assemble_reference(w.phi, inv, div, w.h, jc=jc), jc


def test_solvers_agree_on_mean_zero_pressure(O, W):
    w = W.circle(64)
    (A, b, x0, _), jc = _system(O, w)
    x_ic0, it0, _ = O.pcg_ic0(A, b, w.n, w.n, tol=1e-13)
    x_cg, it1, _ = O.cg(A, b, tol=1e-13)
    x_jac, it2, _ = O.pcg_jacobi(A, b, tol=1e-13, threads=2)
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    rhs = O.rhs_consistent(w.phi, jc, bx, by, w.u, w.v, w.h)
    x_mg, it3, _, _ = O.mg_pcg(bx, by, rhs, w.h, rtol=1e-13)
    ref = x_ic0 - x_ic0.mean()
    for x in (x_cg, x_jac, x_mg.ravel()):
        xm = x - x.mean()
        assert np.linalg.norm(xm - ref) <= 1e-10 * np.linalg.norm(ref)
    assert it3 < it0 < it2 < it1      # MG < IC(0) < Jacobi < CG iterations
    assert abs(x_ic0[0]) < 1e-6 * np.abs(x_ic0).max()   # the pin holds p(0,0) near 0


def test_reference_absolute_stopping_rule(O, W):
    """mode 0: CG stops at r.r < tol^2 (CG_METHOD.cpp:52); PCG at r.z < tol^2 (PCG_METHOD.cpp:253-254)."""
    w = W.circle(32)
    (A, b, _, _), _ = _system(O, w)
    x, it, res = O.cg(A, b, tol=1e-6, relative=False)
    assert res < 1e-6
    r = b - A.matvec(x)
    assert np.linalg.norm(r) < 1e-5
    x, it, res = O.pcg_ic0(A, b, 32, 32, tol=1e-6, relative=False, max_iteration=3)
    assert it == 4   # "if(num_iteration > max_iteration) break" runs max_iteration+1 iterations


@pytest.fixture(scope="module")
def mg_case(O, W):
    w = W.make("drops64", 64)
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    return w, bx, by, O.MG(bx, by)


def _dense_A_scaled(bx, by):
    ny, nx = bx.shape[0], bx.shape[1] - 1
    n = nx * ny
    A = np.zeros((n, n))
    for j in range(ny):
        for i in range(nx):
            c = i + j * nx
            d = bx[j, i] + bx[j, i + 1] + by[j, i] + by[j + 1, i]
            A[c, c] = d
            if i > 0: A[c, c - 1] = -bx[j, i]
            if i + 1 < nx: A[c, c + 1] = -bx[j, i + 1]
            if j > 0: A[c, c - nx] = -by[j, i]
            if j + 1 < ny: A[c, c + nx] = -by[j + 1, i]
    return A


@pytest.mark.parametrize("omega", [1.0, 1.4])
def test_vcycle_symmetric_positive_linear(O, mg_case, omega):
    w, bx, by, mg = mg_case
    if omega != 1.0:
        mg = O.MG(bx, by, omega=omega)
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal((2, w.n, w.n))
    x -= x.mean()
    y -= y.mean()
    Mx, My = mg.vcycle(x), mg.vcycle(y)
    assert abs(np.vdot(Mx, y) - np.vdot(x, My)) <= 1e-12 * abs(np.vdot(Mx, y))
    assert np.vdot(x, Mx) > 0 and np.vdot(y, My) > 0
    assert np.linalg.norm(mg.vcycle(2 * x - 3 * y) - (2 * Mx - 3 * My)) <= 1e-12 * np.linalg.norm(Mx)


@pytest.mark.parametrize("nu", [1, 2, 3])
@pytest.mark.parametrize("omega", [1.0, 1.4])
def test_smoother_is_textbook_sor(O, W, nu, omega):
    """The smoother, pinned against a dense textbook restatement: on a two-level hierarchy
    (16^2 over an 8^2 bottom) the coarse correction of the nu-sweep cycle is the nu = 0
    cycle M0 = P A_c^+ P^T, so the oracle's cycle must equal nu red-black SOR half-sweeps
    (red, black) from x = 0, x += M0 (r - A x), then nu (black, red) half-sweeps, where a
    half-sweep sets every cell c of its colour to (1 - omega) x_c + omega (r_c + sum_f beta_f
    x_nb) / diag_c (the colours do not couple within a half-sweep)."""
    w = W.circle(16)
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    A = _dense_A_scaled(bx, by)
    n = A.shape[0]
    m0 = O.MG(bx, by, nu_pre=0, nu_post=0)
    mg = O.MG(bx, by, nu_pre=nu, nu_post=nu, omega=omega)
    assert m0.levels == 2 and mg.levels == 2
    M0 = np.stack([m0.vcycle(np.eye(n)[k].reshape(16, 16)).ravel() for k in range(n)], axis=1)
    jj, ii = np.divmod(np.arange(n), 16)
    colour = (ii + jj) & 1
    d = np.diag(A)

    def half(x, r, c):
        gs = (r - (A @ x - d * x)) / d
        return np.where(colour == c, (1.0 - omega) * x + omega * gs, x)

    def cycle(r):
        x = np.zeros(n)
        for _ in range(nu):
            x = half(half(x, r, 0), r, 1)
        x = x + M0 @ (r - A @ x)
        for _ in range(nu):
            x = half(half(x, r, 1), r, 0)
        return x

    rng = np.random.default_rng(5)
    for _ in range(3):
        r = rng.standard_normal(n)
        ref = cycle(r)
        got = mg.vcycle(r.reshape(16, 16)).ravel()
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("omega,omega_coarse", [(1.0, 1.0), (1.4, 1.5), (1.2, 1.7)])
def test_three_level_cycle_is_textbook(O, W, omega, omega_coarse):
    """The whole V-cycle, pinned against a dense textbook recursion built from the hierarchy's own
    pieces (orc_mg_piece: the level operators A'_l, the prolongations P_l, the restrictions and the
    bottom pseudo-inverse) on 32^2 -> 16^2 -> 8^2: the restriction is exactly P^T, and the cycle
    is nu SOR red-black half-sweep pairs with omega on the finest level and omega_coarse below it,
    the coarse correction by recursion, then nu (black, red) pairs."""
    w = W.circle(32)
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    nu = 2
    mg = O.MG(bx, by, nu_pre=nu, nu_post=nu, omega=omega, omega_coarse=omega_coarse)
    assert mg.levels == 3
    side = [32, 16, 8]
    n = [s * s for s in side]
    A = [np.stack([mg.piece(0, l, np.eye(n[l])[k], n[l]) for k in range(n[l])], axis=1) for l in range(2)]
    P = [np.stack([mg.piece(1, l, np.eye(n[l + 1])[k], n[l]) for k in range(n[l + 1])], axis=1) for l in range(2)]
    R = [np.stack([mg.piece(2, l, np.eye(n[l])[k], n[l + 1]) for k in range(n[l])], axis=1) for l in range(2)]
    B = np.stack([mg.piece(3, 0, np.eye(n[2])[k], n[2]) for k in range(n[2])], axis=1)
    for l in range(2):
        assert np.array_equal(R[l], P[l].T)
    assert np.allclose(A[0], _dense_A_scaled(bx, by), rtol=0, atol=1e-12 * np.abs(A[0]).max())

    def cycle(l, r):
        if l == 2:
            return B @ r
        wl = omega if l == 0 else omega_coarse
        jj, ii = np.divmod(np.arange(n[l]), side[l])
        colour = (ii + jj) & 1
        d = np.diag(A[l])

        def half(x, c):
            gs = (r - (A[l] @ x - d * x)) / d
            return np.where(colour == c, (1.0 - wl) * x + wl * gs, x)
        x = np.zeros(n[l])
        for _ in range(nu):
            x = half(half(x, 0), 1)
        x = x + P[l] @ cycle(l + 1, P[l].T @ (r - A[l] @ x))
        for _ in range(nu):
            x = half(half(x, 1), 0)
        return x

    rng = np.random.default_rng(9)
    for _ in range(3):
        r = rng.standard_normal(n[0])
        ref = cycle(0, r)
        got = mg.vcycle(r.reshape(32, 32)).ravel()
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


def test_sor_cuts_iterations(O, W):
    """Over-relaxation (omega = 1.4) needs fewer MG-PCG iterations than Gauss-Seidel on the
    1000:1 workloads (a regression bound, not a pin)."""
    for name, n in [("rt", 256), ("drops64", 256)]:
        w = W.make(name, n)
        bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
        b = -O.divergence(w.u, w.v, w.h)
        _, it1, _, _ = O.mg_pcg(bx, by, b, w.h, rtol=1e-12, nu=3)
        _, it14, _, _ = O.mg_pcg(bx, by, b, w.h, rtol=1e-12, nu=3, omega=1.4)
        assert it14 < it1, (name, it1, it14)
        if name == "drops64":   # many small interfaces: the coarse levels gain from more relaxation
            _, it15, _, _ = O.mg_pcg(bx, by, b, w.h, rtol=1e-12, nu=3, omega=1.4, omega_coarse=1.5)
            assert it15 <= it14, (name, it14, it15)


def test_half_rounding_matches_numpy(O):
    """orc_half_via_float (how the FP32 hierarchy stores its weights) equals numpy's binary32 ->
    binary16 casts (nearest, ties to even) on random weights, log-spread values down to the
    binary16 subnormals, exact binary16 ties and the ends of [0, 1]."""
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.random(4000), np.exp(rng.uniform(np.log(1e-9), 0.0, 4000)),
                        np.arange(1, 2048) * 2.0 ** -11 + 2.0 ** -12,     # ties between binary16 neighbours
                        np.arange(1, 512) * 2.0 ** -25,                   # around the subnormal quantum 2^-24
                        [0.0, 1.0, 0.99999, 0.9999999, 6.1e-5, 2.0 ** -14, 2.0 ** -24, 2.0 ** -25]])
    got = np.array([O.half_via_float(v) for v in x])
    ref = np.float16(np.float32(x)).astype(np.float64)
    assert np.array_equal(got, ref)


def test_vcycle_with_binary16_weights(O, mg_case):
    """Mode 7 (the FP32 hierarchy): every weight rounded to binary16.  The cycle stays a
    symmetric positive linear operator (R = P^T with the rounded weights) and differs from
    mode 3 by the rounding only."""
    w, bx, by, mg = mg_case
    mg16 = O.MG(bx, by, mode=O.MG_MODE_F32)
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal((2, w.n, w.n))
    x -= x.mean()
    y -= y.mean()
    Mx, My = mg16.vcycle(x), mg16.vcycle(y)
    assert abs(np.vdot(Mx, y) - np.vdot(x, My)) <= 1e-12 * abs(np.vdot(Mx, y))
    assert np.vdot(x, Mx) > 0 and np.vdot(y, My) > 0
    d = np.linalg.norm(Mx - mg.vcycle(x)) / np.linalg.norm(Mx)
    assert 0 < d < 1e-2, d


def test_preconditioned_spectrum(O, W):
    """On mean-zero vectors M is SPD and M A' has a bounded spread; constants are in M's null space."""
    w = W.make("drops64", 32)
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    A = _dense_A_scaled(bx, by)
    mg = O.MG(bx, by)
    n = A.shape[0]
    M = np.stack([mg.vcycle(np.eye(n)[k].reshape(w.n, w.n)).ravel() for k in range(n)], axis=1)
    assert np.abs(M - M.T).max() <= 1e-12 * np.abs(M).max()
    Q = np.eye(n) - 1.0 / n                       # projector onto mean-zero vectors
    MQ = Q @ M @ Q
    # the null-space mode is not amplified beyond the rest of the spectrum (a pinned
    # coarse solve would scale it by ~N / pin)
    one = np.ones(n)
    assert np.linalg.norm(M @ one) / np.sqrt(n) <= np.linalg.norm(MQ, 2)
    ev = np.linalg.eigvalsh(0.5 * (MQ + MQ.T))
    assert ev[1] > 0                               # positive definite on the complement of constants
    lam = np.sort(np.real(np.linalg.eigvals(MQ @ A)))[1:]   # drop the zero of the null space
    assert lam[0] > 0.3 and lam[-1] / lam[0] < 40, (lam[0], lam[-1])


def test_mg_pcg_iterations_constant_coefficient(O):
    n, h = 128, 1.0 / 128
    bx = np.ones((n, n + 1)); bx[:, 0] = bx[:, -1] = 0
    by = np.ones((n + 1, n)); by[0] = by[-1] = 0
    rng = np.random.default_rng(2)
    b = rng.standard_normal((n, n)); b -= b.mean()
    x, it, r2, _ = O.mg_pcg(bx, by, b, h, rtol=1e-10)
    assert it <= 12, it
    assert np.linalg.norm(b - O.apply_A0(bx, by, x, h)) <= 1e-10 * np.linalg.norm(b)


@pytest.mark.parametrize("name,n,cap", [("circle", 128, 20), ("rt", 128, 20), ("drops64", 256, 40)])
def test_mg_pcg_iterations_high_contrast(O, W, name, n, cap):
    """1000:1 density: the operator-dependent interpolation keeps iteration counts bounded."""
    w = W.make(name, n)
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    jc = w.sigma * O.curvature(w.phi, w.h)
    b = O.rhs_consistent(w.phi, jc, bx, by, w.u, w.v, w.h)
    x, it, _, _ = O.mg_pcg(bx, by, b, w.h, rtol=1e-12)
    assert it <= cap, it
    x2, it2, _, _ = O.mg_pcg(bx, by, b, w.h, rtol=1e-12, mode=0)   # plain averaging + bilinear
    assert it2 > it
