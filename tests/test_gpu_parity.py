"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on seeded inputs.

Bars: coefficient assembly and the FP64 stencil to FP64 roundoff; the converged
mean-zero pressure within 1e-9 relative of the oracle's IC(0)-PCG solve of the
reference-literal CSR system; FP64 max|D u_proj| <= 1e-10 computed on the CPU
from the GPU's velocities.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _coeffs(O, w):
    bx, by = O.face_beta(w.phi, w.rho_minus, w.rho_plus)
    jc = w.sigma * O.curvature(w.phi, w.h)
    return bx, by, jc


@pytest.mark.parametrize("name,n", [("circle", 64), ("drops64", 128), ("rt", 128)])
def test_coefficients_match_oracle(P2D, O, W, name, n):
    w = W.make(name, n)
    print(f"case coefficients {name} n={n}", flush=True)
    ctx = P2D.Context(n, n, w.h)
    ctx.set_fluid(w.phi, w.rho_minus, w.rho_plus, w.sigma)
    ctx.project(w.u, w.v, allow_not_converged=True)
    bx, by, jc = ctx.coefficients()
    rbx, rby, rjc = _coeffs(O, w)
    assert np.array_equal(bx, rbx) or np.abs(bx - rbx).max() <= 1e-15 * np.abs(rbx).max()
    assert np.array_equal(by, rby) or np.abs(by - rby).max() <= 1e-15 * np.abs(rby).max()
    scale = max(np.abs(rjc).max(), 1e-300)
    assert np.abs(jc - rjc).max() <= 1e-13 * scale


@pytest.mark.parametrize("prec", [0, 1])
def test_stencil_matches_oracle(P2D, O, W, prec):
    w = W.make("drops64", 128)
    print(f"case stencil prec={prec}", flush=True)
    bx, by, jc = _coeffs(O, w)
    ctx = P2D.Context(w.n, w.n, w.h)
    ctx.set_coefficients(bx, by, w.phi, jc)
    x = np.random.default_rng(5).standard_normal((w.n, w.n))
    ref = O.apply_A0(bx, by, x, w.h)
    if prec == 0:
        y = ctx.apply_stencil(x, prec=0)
        assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()
    else:
        y = ctx.apply_stencil(x.astype(np.float32), prec=1)
        assert np.abs(y - ref).max() <= 1e-5 * np.abs(ref).max()


# omega, omega_coarse: the smoother's over-relaxation on the finest level and below it (1: Gauss-Seidel)
@pytest.mark.parametrize("prec,nu,omega,omc", [(0, 1, 1.0, 1.0), (0, 2, 1.0, 1.0), (0, 3, 1.0, 1.0), (1, 2, 1.0, 1.0),
                                               (1, 3, 1.0, 1.0), (0, 1, 1.4, 1.4), (0, 3, 1.4, 1.5), (1, 3, 1.4, 1.5),
                                               (0, 3, 1.7, 1.2), (1, 2, 0.8, 1.6)])
def test_vcycle_matches_oracle(P2D, O, W, prec, nu, omega, omc):
    w = W.make("drops64", 256)
    print(f"case vcycle prec={prec} nu={nu} omega={omega} omega_coarse={omc}", flush=True)
    bx, by, jc = _coeffs(O, w)
    ctx = P2D.Context(w.n, w.n, w.h, mg_pre_sweeps=nu, mg_post_sweeps=nu, mg_omega=omega, mg_omega_coarse=omc)
    ctx.set_coefficients(bx, by, w.phi, jc)
    # the FP32 hierarchy stores its weights in binary16: its oracle is mode 7
    mg = O.MG(bx, by, nu_pre=nu, nu_post=nu, mode=O.MG_DEFAULT_MODE if prec == 0 else O.MG_MODE_F32, omega=omega,
              omega_coarse=omc)
    assert mg.levels == ctx.mg_levels
    r = np.random.default_rng(6).standard_normal((w.n, w.n))
    r -= r.mean()
    ref = mg.vcycle(r)
    if prec == 0:
        z = ctx.vcycle(r, prec=0)
        assert np.linalg.norm(z - ref) <= 1e-11 * np.linalg.norm(ref)
    else:
        z = ctx.vcycle(r.astype(np.float32), prec=1)
        assert np.linalg.norm(z - ref) <= 1e-4 * np.linalg.norm(ref)


# Rectangular grids: hierarchies whose single-workgroup cycle has three levels
# (192x128, 160x96), two levels over a 135-cell bottom whose pseudo-inverse
# stays in global memory (120x72), and only the 576-cell bottom (128x18).
@pytest.mark.parametrize("nx,ny,prec,omega", [(192, 128, 0, 1.0), (160, 96, 0, 1.0), (160, 96, 1, 1.0),
                                              (120, 72, 0, 1.0), (120, 72, 1, 1.0), (128, 18, 0, 1.0),
                                              (160, 96, 0, 1.4), (160, 96, 1, 1.4), (120, 72, 0, 1.4)])
def test_vcycle_rectangular_matches_oracle(P2D, O, nx, ny, prec, omega):
    """omega on every level (omega_coarse = omega)."""
    print(f"case vcycle rectangular {nx}x{ny} prec={prec} omega={omega}", flush=True)
    h = 1.0 / nx
    X, Y = np.meshgrid((np.arange(nx) + 0.5) * h, (np.arange(ny) + 0.5) * h)
    phi = np.hypot(X - 0.4, Y - 0.45 * ny / nx) - 0.3 * ny / nx
    bx, by = O.face_beta(phi, 1.0, 1000.0)
    ctx = P2D.Context(nx, ny, h, mg_pre_sweeps=2, mg_post_sweeps=2, mg_omega=omega, mg_omega_coarse=omega)
    ctx.set_coefficients(bx, by, phi, np.zeros_like(phi))
    mg = O.MG(bx, by, nu_pre=2, nu_post=2, mode=O.MG_DEFAULT_MODE if prec == 0 else O.MG_MODE_F32, omega=omega)
    assert mg.levels == ctx.mg_levels
    r = np.random.default_rng(7).standard_normal((ny, nx))
    r -= r.mean()
    ref = mg.vcycle(r)
    if prec == 0:
        z = ctx.vcycle(r, prec=0)
        assert np.linalg.norm(z - ref) <= 1e-11 * np.linalg.norm(ref)
    else:
        z = ctx.vcycle(r.astype(np.float32), prec=1)
        assert np.linalg.norm(z - ref) <= 1e-4 * np.linalg.norm(ref)


@pytest.mark.parametrize("prec", [0, 1])
def test_vcycle_tiled_setup_partial_tiles(P2D, O, prec):
    """608 x 512 (311,296 cells): the finest level is set up by the tiled setup sweep, whose
    64-column tiles do not divide 608, so its last tile column is partial and reads faces
    beyond the grid as walls; the cycle (default over-relaxation) still matches the oracle.
    (A partial last tile row cannot occur: a height that is not a multiple of 16 halves at most
    three times, which leaves no legal bottom for a grid of 512^2 cells or more.)"""
    nx, ny = 608, 512
    print(f"case vcycle tiled setup {nx}x{ny} prec={prec}", flush=True)
    h = 1.0 / nx
    X, Y = np.meshgrid((np.arange(nx) + 0.5) * h, (np.arange(ny) + 0.5) * h)
    phi = np.minimum(np.hypot(X - 0.93, Y - 0.41) - 0.05, np.hypot(X - 0.3, Y - 0.5) - 0.12)
    bx, by = O.face_beta(phi, 1.0, 1000.0)
    ctx = P2D.Context(nx, ny, h)
    ctx.set_coefficients(bx, by, phi, np.zeros_like(phi))
    om, omc = ctx.cfg.mg_omega, ctx.cfg.mg_omega_coarse
    mg = O.MG(bx, by, nu_pre=3, nu_post=3, mode=O.MG_DEFAULT_MODE if prec == 0 else O.MG_MODE_F32, omega=om,
              omega_coarse=omc)
    assert mg.levels == ctx.mg_levels == 6   # 608x512 ... 19x16
    r = np.random.default_rng(11).standard_normal((ny, nx))
    r -= r.mean()
    ref = mg.vcycle(r)
    if prec == 0:
        z = ctx.vcycle(r, prec=0)
        assert np.linalg.norm(z - ref) <= 1e-11 * np.linalg.norm(ref)
    else:
        z = ctx.vcycle(r.astype(np.float32), prec=1)
        assert np.linalg.norm(z - ref) <= 1e-4 * np.linalg.norm(ref)


@pytest.mark.parametrize("name,n,prec", [("circle", 128, 0), ("circle", 128, 1), ("rt", 256, 0), ("rt", 256, 1),
                                         ("drops64", 256, 0), ("drops64", 256, 1)])
def test_projection_matches_oracle(P2D, O, W, name, n, prec):
    w = W.make(name, n)
    print(f"case projection {name} n={n} prec={prec}", flush=True)
    ctx = P2D.Context(n, n, w.h, precision=prec, rtol=1e-12, div_tol=1e-10)
    ctx.set_fluid(w.phi, w.rho_minus, w.rho_plus, w.sigma)
    u, v, st = ctx.project(w.u, w.v, w.dt)
    p = ctx.pressure(mean_zero=True)
    solver = O.SOLVER_IC0 if n <= 128 else O.SOLVER_MG
    p_ref, u_ref, v_ref, _ = O.project(w.phi, w.rho_minus, w.rho_plus, w.sigma, w.u, w.v, w.h, solver=solver,
                                       tol=1e-13)
    p_ref = p_ref - p_ref.mean()
    rel = np.linalg.norm(p - p_ref) / np.linalg.norm(p_ref)
    div = np.abs(O.divergence(u, v, w.h)).max()
    print(f"  iters={st['iterations']} outer={st['outer']} p_rel={rel:.2e} div={div:.2e} gpu_div={st['max_div']:.2e}")
    assert rel <= 1e-9
    assert div <= 1e-10
    assert abs(div - st["max_div"]) <= 1e-12
    # velocity: the oracle's consistent update with the oracle pressure
    du = np.abs(u - u_ref).max() / np.abs(u_ref).max()
    dv = np.abs(v - v_ref).max() / np.abs(v_ref).max()
    assert du <= 1e-8 and dv <= 1e-8
    # walls untouched
    assert np.all(u[:, 0] == 0) and np.all(u[:, -1] == 0) and np.all(v[0] == 0) and np.all(v[-1] == 0)


@pytest.mark.parametrize("prec,rho_m,rho_p", [(0, 1.0, 1000.0), (1, 1.0, 1000.0), (1, 1000.0, 1.0)])
def test_projection_rectangular_matches_oracle(P2D, O, W, prec, rho_m, rho_p):
    """A 384 x 320 grid (h = 1/384, the hierarchy ends at a 6 x 5 bottom) with two drops, surface
    tension and seeded velocities: the GPU projection against the oracle's MG-PCG solve; light
    drops in a heavy fluid, and heavy drops in a light one."""
    nx, ny = 384, 320
    h = 1.0 / nx
    print(f"case projection rectangular {nx}x{ny} prec={prec} rho {rho_m}/{rho_p}", flush=True)
    x = (np.arange(nx) + 0.5) * h
    y = (np.arange(ny) + 0.5) * h
    X, Y = np.meshgrid(x, y)
    phi = np.minimum(np.hypot(X - 0.4, Y - 0.45) - 0.15, np.hypot(X - 0.72, Y - 0.3) - 0.08)
    nu, nv = ny * (nx + 1), (ny + 1) * nx
    r = W.uniform(11, 5, nu + nv)
    u = (2.0 * r[:nu] - 1.0).reshape(ny, nx + 1)
    v = (2.0 * r[nu:] - 1.0).reshape(ny + 1, nx)
    u, v = W.zero_walls(u, v)
    sigma = 0.0728
    ctx = P2D.Context(nx, ny, h, precision=prec, rtol=1e-12, div_tol=1e-10)
    ctx.set_fluid(phi, rho_m, rho_p, sigma)
    up, vp, st = ctx.project(u, v, 1.0)
    p = ctx.pressure(mean_zero=True)
    p_ref, u_ref, v_ref, _ = O.project(phi, rho_m, rho_p, sigma, u, v, h, solver=O.SOLVER_MG, tol=1e-15)
    p_ref = p_ref - p_ref.mean()
    rel = np.linalg.norm(p - p_ref) / np.linalg.norm(p_ref)
    div = np.abs(O.divergence(up, vp, h)).max()
    print(f"  iters={st['iterations']} outer={st['outer']} p_rel={rel:.2e} div={div:.2e}")
    assert st["status"] == 0
    assert rel <= 1e-9
    assert div <= 1e-10
    du = np.abs(up - u_ref).max() / np.abs(u_ref).max()
    dv = np.abs(vp - v_ref).max() / np.abs(v_ref).max()
    assert du <= 1e-8 and dv <= 1e-8
    assert np.all(up[:, 0] == 0) and np.all(up[:, -1] == 0) and np.all(vp[0] == 0) and np.all(vp[-1] == 0)


@pytest.mark.parametrize("prec", [0, 1])
def test_projection_single_phase(P2D, O, W, prec):
    """No interface: the level set is negative everywhere (with a nonzero gradient, so the
    curvature of the surface-tension term is defined), every face has beta = 1 / rho_minus and no
    jump term enters; the solve is the plain Neumann Poisson problem."""
    n = 160
    h = 1.0 / n
    print(f"case projection single phase {n}^2 prec={prec}", flush=True)
    X, Y = np.meshgrid((np.arange(n) + 0.5) * h, (np.arange(n) + 0.5) * h)
    phi = -0.5 - 0.2 * X - 0.1 * Y
    u, v = W.uniform_velocity(n, 13)
    ctx = P2D.Context(n, n, h, precision=prec, rtol=1e-12, div_tol=1e-10)
    ctx.set_fluid(phi, 1.0, 1000.0, 0.0728)
    up, vp, st = ctx.project(u, v, 1.0)
    bx, by, jc = ctx.coefficients()
    assert np.all(bx[:, 1:-1] == 1.0) and np.all(by[1:-1] == 1.0)
    p = ctx.pressure(mean_zero=True)
    p_ref, _, _, _ = O.project(phi, 1.0, 1000.0, 0.0728, u, v, h, solver=O.SOLVER_MG, tol=1e-15)
    p_ref = p_ref - p_ref.mean()
    rel = np.linalg.norm(p - p_ref) / np.linalg.norm(p_ref)
    div = np.abs(O.divergence(up, vp, h)).max()
    print(f"  iters={st['iterations']} outer={st['outer']} p_rel={rel:.2e} div={div:.2e}")
    assert st["status"] == 0
    assert rel <= 1e-9
    assert div <= 1e-10


@pytest.mark.parametrize("prec", [0, 1])
def test_projection_matches_golden(P2D, W, prec):
    """The config-0 case against the committed oracle fixture (no oracle run on the GPU box)."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "circle128.npz"))
    w = W.circle(128)
    print(f"case golden circle 128 prec={prec}", flush=True)
    ctx = P2D.Context(128, 128, w.h, precision=prec, rtol=1e-12, div_tol=1e-10)
    ctx.set_fluid(w.phi, w.rho_minus, w.rho_plus, w.sigma)
    u, v, st = ctx.project(w.u, w.v, w.dt)
    p = ctx.pressure(mean_zero=True)
    assert np.linalg.norm(p - g["p"]) <= 1e-9 * np.linalg.norm(g["p"])
    scale = max(np.abs(g["u"]).max(), np.abs(g["v"]).max())
    assert max(np.abs(u - g["u"]).max(), np.abs(v - g["v"]).max()) <= 1e-8 * scale


def test_vcycle_2048_with_two_block_coarse_level(P2D, O, W):
    """At 2048^2 the coarse levels take all three tile shapes: the first one (1024^2) 12-wavefront
    tiles, about two per CU, and the smaller ones 8-wavefront tiles; the cycle (with the default
    over-relaxation) still matches the oracle."""
    w = W.rt(2048)
    print("case vcycle 2048 prec=1", flush=True)
    bx, by, jc = _coeffs(O, w)
    ctx = P2D.Context(w.n, w.n, w.h)
    ctx.set_coefficients(bx, by, w.phi, jc)
    om, omc = ctx.cfg.mg_omega, ctx.cfg.mg_omega_coarse
    assert (om, omc) == (1.4, 1.5)
    mg = O.MG(bx, by, nu_pre=3, nu_post=3, omega=om, omega_coarse=omc)
    mg16 = O.MG(bx, by, nu_pre=3, nu_post=3, mode=O.MG_MODE_F32, omega=om, omega_coarse=omc)
    assert mg.levels == ctx.mg_levels
    r = np.random.default_rng(8).standard_normal((w.n, w.n))
    r -= r.mean()
    ref = mg.vcycle(r)
    ref16 = mg16.vcycle(r)
    z = ctx.vcycle(r.astype(np.float32), prec=1)
    assert np.linalg.norm(z - ref16) <= 1e-4 * np.linalg.norm(ref16)
    z64 = ctx.vcycle(r, prec=0)
    assert np.linalg.norm(z64 - ref) <= 1e-11 * np.linalg.norm(ref)
