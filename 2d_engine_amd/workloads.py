"""Seeded synthetic inputs for the projection workloads.

Grid convention (shared with the C ABI): n x n cells on [0,1]^2, h = 1/n, cell
(i,j) at ((i+1/2)h, (j+1/2)h); x-faces at (i h, (j+1/2)h), shape (n, n+1);
y-faces at ((i+1/2)h, j h), shape (n+1, n).  Wall-normal faces are zeroed
before projection (EULERIAN_FLUID_SOLVER_2D.h:2240-2261).  In the reference's
own terms this is base_grid_resolution=(n,n), base_grid_min=(h/2,h/2),
base_grid_max=(1-h/2,1-h/2) (GRID_STRUCTURE_2D.h:88-124), so its node spacing
dx = (x_max-x_min)/(n-1) equals h.

Random numbers: a counter-based splitmix64 (Steele, Lea, Flood 2014) -> 53-bit
uniforms, vectorised in numpy.  Each draw site uses its own stream id so the
inputs of one workload do not depend on the sizes of another.

Workloads (density ratio 1000:1 everywhere, rho = 1 on the phi <= 0 side):
  circle      one circle r=0.25 at (0.5,0.5), sigma=0.0728, u,v ~ U[-1,1] seed 0
  drops64     64 circles c~U[.1,.9]^2, r~U[.02,.08] (seed 1), sigma=0.0728,
              band-limited velocity (Fourier |k|<=16, N(0,1)/|k| amplitudes,
              unit RMS) + U[-0.1,0.1] noise
  rt          Rayleigh-Taylor phi = y - 0.5 - 0.01 cos(8 pi x), sigma=0,
              u,v ~ U[-1,1] seed 2
  drops512    512 circles c~U[.05,.95]^2, r~U[.005,.03] (seed 4), sigma=0.0728,
              u,v ~ FP32 U[-1,1] (values representable in FP32); at 4096^2 its dt is
              the CFL step of that state (DROPS512_CFL_DT), elsewhere 1

The reference projects only at the step its CFL rule allows (CFLOneTimeStep,
EULERIAN_FLUID_SOLVER_2D.h:3517-3637, before every step: SIMULATION_WORLD.h:1194-1207).
The other workloads use dt = 1: for them dt only scales the pressure (sigma = 0 or
small jumps).  For drops512 at dt = 1 the surface tension of r = 0.005 drops drives
|u_proj| to ~1.5e3, where the rounding of u alone leaves max|D u| ~ 1.5e-9; at its CFL
step |u_proj| stays O(1).
"""
import math

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(seed, stream, count, offset=0):
    """`count` uniforms in [0,1) from stream (seed, stream): splitmix64 output k = mix(s0 + (k+1)*GOLD)."""
    with np.errstate(over="ignore"):
        s0 = _mix(np.array([np.uint64(seed) * _GOLD + np.uint64(stream) * _M1], dtype=np.uint64))[0]
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = _mix(s0 + k * _GOLD)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, stream, count):
    """Box-Muller standard normals from two uniform blocks of the same stream."""
    u1 = uniform(seed, stream, count)
    u2 = uniform(seed, stream, count, offset=count)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


def cell_centers(n):
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    return c[None, :], c[:, None]  # x (1,n), y (n,1)


def zero_walls(u, v):
    u[:, 0] = 0.0
    u[:, -1] = 0.0
    v[0, :] = 0.0
    v[-1, :] = 0.0
    return u, v


def uniform_velocity(n, seed, lo=-1.0, hi=1.0, fp32=False):
    nu = n * (n + 1)
    r = uniform(seed, 7, 2 * nu)
    w = lo + (hi - lo) * r
    if fp32:
        w = w.astype(np.float32).astype(np.float64)
    u = w[:nu].reshape(n, n + 1).copy()
    v = w[nu:].reshape(n + 1, n).copy()
    return zero_walls(u, v)


def levelset_circles(n, centers, radii, tile=32):
    """phi = min_k (|x - c_k| - r_k) at cell centres, exact (brute-force equal).

    Circles that cannot be the minimum anywhere in a tile (their lower bound over
    the tile exceeds some circle's upper bound) are culled per tile; every kept
    (cell, circle) pair is evaluated with the same numpy arithmetic, so the result
    equals the brute-force minimum bit for bit.
    """
    h = 1.0 / n
    cx, cy = centers[:, 0], centers[:, 1]
    phi = np.empty((n, n))
    for j0 in range(0, n, tile):
        j1 = min(n, j0 + tile)
        ys = (np.arange(j0, j1) + 0.5) * h
        for i0 in range(0, n, tile):
            i1 = min(n, i0 + tile)
            xs = (np.arange(i0, i1) + 0.5) * h
            # distance bounds from each centre to the tile's box of cell centres
            dx_lo = np.maximum(0.0, np.maximum(xs[0] - cx, cx - xs[-1]))
            dy_lo = np.maximum(0.0, np.maximum(ys[0] - cy, cy - ys[-1]))
            dx_hi = np.maximum(np.abs(cx - xs[0]), np.abs(cx - xs[-1]))
            dy_hi = np.maximum(np.abs(cy - ys[0]), np.abs(cy - ys[-1]))
            lo = np.sqrt(dx_lo ** 2 + dy_lo ** 2) - radii
            hi = np.sqrt(dx_hi ** 2 + dy_hi ** 2) - radii
            keep = np.nonzero(lo <= hi.min() + 1e-12)[0]
            best = None
            X, Y = xs[None, :], ys[:, None]
            for k in keep:
                d = np.sqrt((X - cx[k]) ** 2 + (Y - cy[k]) ** 2) - radii[k]
                best = d if best is None else np.minimum(best, d)
            phi[j0:j1, i0:i1] = best
    return phi


def random_circles(count, seed, c_lo, c_hi, r_lo, r_hi):
    c = c_lo + (c_hi - c_lo) * uniform(seed, 1, 2 * count).reshape(count, 2)
    r = r_lo + (r_hi - r_lo) * uniform(seed, 2, count)
    return c, r


def band_limited_velocity(n, seed, kmax=16, noise=0.1):
    """Fourier modes 1 <= |k| <= kmax, amplitudes N(0,1)/|k|, unit RMS, + U[-noise, noise]."""
    h = 1.0 / n
    out = []
    ks = [(kx, ky) for ky in range(0, kmax + 1) for kx in range(-kmax, kmax + 1)
          if 1 <= math.hypot(kx, ky) <= kmax and not (ky == 0 and kx < 0)]
    K = len(ks)
    for comp, (shape, xoff, yoff) in enumerate((((n, n + 1), 0.0, 0.5), ((n + 1, n), 0.5, 0.0))):
        g = normal(seed, 10 + comp, 2 * K).reshape(2, K)
        kk = np.array([math.hypot(a, b) for a, b in ks])
        a, b = g[0] / kk, g[1] / kk
        ny_, nx_ = shape
        x = (np.arange(nx_) + xoff) * h
        y = (np.arange(ny_) + yoff) * h
        A = np.zeros((kmax + 1, 2 * kmax + 1))
        B = np.zeros_like(A)
        for m, (kx, ky) in enumerate(ks):
            A[ky, kx + kmax] = a[m]
            B[ky, kx + kmax] = b[m]
        kxs = np.arange(-kmax, kmax + 1)
        kys = np.arange(0, kmax + 1)
        Cx, Sx = np.cos(2 * math.pi * np.outer(x, kxs)), np.sin(2 * math.pi * np.outer(x, kxs))
        Cy, Sy = np.cos(2 * math.pi * np.outer(y, kys)), np.sin(2 * math.pi * np.outer(y, kys))
        # sum a cos(X+Y) + b sin(X+Y) with X = 2 pi kx x, Y = 2 pi ky y
        f = Cy @ (A @ Cx.T + B @ Sx.T) + Sy @ (B @ Cx.T - A @ Sx.T)
        f /= math.sqrt(float(np.mean(f * f)))
        f += noise * (2.0 * uniform(seed, 20 + comp, f.size).reshape(f.shape) - 1.0)
        out.append(np.ascontiguousarray(f))
    return zero_walls(out[0], out[1])


class Workload:
    """Inputs of one projection: level set, densities, surface tension, velocities."""

    def __init__(self, name, n, phi, rho_minus, rho_plus, sigma, u, v, dt=1.0, extra=None):
        self.name, self.n, self.phi = name, n, phi
        self.rho_minus, self.rho_plus, self.sigma = rho_minus, rho_plus, sigma
        self.u, self.v, self.dt = u, v, dt
        self.h = 1.0 / n
        self.extra = extra or {}


def circle(n=128):
    x, y = cell_centers(n)
    phi = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2) - 0.25
    u, v = uniform_velocity(n, 0)
    return Workload("circle", n, np.ascontiguousarray(phi), 1.0, 1000.0, 0.0728, u, v)


def drops64(n=1024):
    c, r = random_circles(64, 1, 0.1, 0.9, 0.02, 0.08)
    phi = levelset_circles(n, c, r)
    u, v = band_limited_velocity(n, 1)
    return Workload("drops64", n, phi, 1.0, 1000.0, 0.0728, u, v)


def rt(n=2048):
    x, y = cell_centers(n)
    phi = np.ascontiguousarray(np.broadcast_to(y - 0.5 - 0.01 * np.cos(8.0 * math.pi * x), (n, n)))
    u, v = uniform_velocity(n, 2)
    return Workload("rt", n, phi, 1.0, 1000.0, 0.0, u, v)


# dt of drops512 at 4096^2: CFLOneTimeStep of the workload's own phi, u, v at CFL number 0.5,
# rho 1 / 1000, sigma 0.0728, gravity 9.8 and no viscosity.  Its parts: advection
# c = (max|u| + max|v|) / h = 8192.0, gravity g = sqrt(9.8 / h) = 200.4 and surface tension
# s = sqrt(sigma max|kappa| / (min rho h^2)) = 70730 with max|kappa| at the clamp 1/h (the
# kinks of the union of circles); dt = 0.5 / (0.5 (c + sqrt(c^2 + 4 g^2 + 4 s^2))).  The
# oracle's restatement (CPU test) and p2d_cfl_dt (GPU test) both reproduce it.
DROPS512_CFL_DT = 6.671550069261913e-06


def drops512(n=4096):
    c, r = random_circles(512, 4, 0.05, 0.95, 0.005, 0.03)
    phi = levelset_circles(n, c, r)
    u, v = uniform_velocity(n, 4, fp32=True)
    dt = DROPS512_CFL_DT if n == 4096 else 1.0
    return Workload("drops512", n, phi, 1.0, 1000.0, 0.0728, u, v, dt=dt, extra={"cfl_number": 0.5})


def target_velocity(n, seed=3):
    """u_target ~ U[-1,1] (seed 3) for the adjoint loss 0.5 ||u_proj - u_target||^2."""
    return uniform_velocity(n, seed)


BUILDERS = {"circle": circle, "drops64": drops64, "rt": rt, "drops512": drops512}


def make(name, n=None):
    f = BUILDERS[name]
    return f() if n is None else f(n)
