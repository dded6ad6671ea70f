"""Seeded workload generators: determinism, brute-force equality of the culled
circle level set, wall handling and FP32 representability."""
import numpy as np


def test_uniform_is_deterministic_and_in_range(W):
    a = W.uniform(7, 3, 1000)
    b = W.uniform(7, 3, 1000)
    assert np.array_equal(a, b) and a.min() >= 0 and a.max() < 1
    assert not np.array_equal(a, W.uniform(8, 3, 1000))
    assert abs(a.mean() - 0.5) < 0.05


def test_circle_levelset_equals_brute_force(W):
    c, r = W.random_circles(40, 11, 0.1, 0.9, 0.02, 0.08)
    n = 96
    phi = W.levelset_circles(n, c, r, tile=16)
    x, y = W.cell_centers(n)
    brute = np.min(np.stack([np.sqrt((x - c[k, 0]) ** 2 + (y - c[k, 1]) ** 2) - r[k] for k in range(len(r))]), axis=0)
    assert np.array_equal(phi, brute)


def test_workload_shapes_and_walls(W):
    for name, n in (("circle", 32), ("drops64", 32), ("rt", 32), ("drops512", 64)):
        w = W.make(name, n)
        assert w.phi.shape == (n, n) and w.u.shape == (n, n + 1) and w.v.shape == (n + 1, n)
        assert np.all(w.u[:, 0] == 0) and np.all(w.u[:, -1] == 0)
        assert np.all(w.v[0] == 0) and np.all(w.v[-1] == 0)
        assert (w.phi <= 0).any() and (w.phi > 0).any()
    w = W.make("drops512", 64)
    assert np.array_equal(w.u.astype(np.float32).astype(np.float64), w.u)


def test_band_limited_field_statistics(W):
    u, v = W.band_limited_velocity(128, 1)
    core = u[:, 1:-1]
    assert 0.8 < np.sqrt(np.mean(core ** 2)) < 1.2


def test_drops512_dt_is_its_cfl_step(W, O):
    """drops512 at 4096^2 projects at the step of the reference's CFL rule for its own state
    (CFLOneTimeStep, EULERIAN_FLUID_SOLVER_2D.h:3517-3637) at CFL 0.5, as the oracle restates it."""
    w = W.drops512(4096)
    dt = O.cfl_dt(w.phi, w.u, w.v, w.h, cfl_number=0.5, rho_minus=w.rho_minus, rho_plus=w.rho_plus,
                  sigma=w.sigma)
    assert w.dt == W.DROPS512_CFL_DT
    assert abs(dt - w.dt) <= 1e-14 * dt
    assert W.drops512(64).dt == 1.0
