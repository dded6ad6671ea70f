"""p2d_cfl_dt_noalloc (the entry point that allocates nothing) against p2d_cfl_dt, through the
C ABI: two calls in a row, the second on a smaller grid, both equal to the context-based entry
point's dt; an undefined curvature gives NaN where p2d_cfl_dt gives P2D_NAN, and the next call
is finite again (the error flag in the shared static scratch is cleared per call); a grid beyond
the static scratch is refused with P2D_BAD_ARG."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(W, name, n):
    w = W.make(name, n)
    return w


def _noalloc(N, torch, w, params, cfl, stream):
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    phi, u, v = T(w.phi), T(w.u), T(w.v)
    out = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    ny, nx = w.phi.shape
    rc = N.lib.p2d_cfl_dt_noalloc(nx, ny, w.h, phi.data_ptr(), u.data_ptr(), v.data_ptr(), ctypes.byref(params),
                                  cfl, out.data_ptr(), stream.cuda_stream)
    assert rc == N.P2D_OK
    stream.synchronize()
    return float(out.cpu().item())


def _reference(P2D, w, cfl, **kw):
    ny, nx = w.phi.shape
    ctx = P2D.Context(nx, ny, w.h)
    return ctx.cfl_dt(w.phi, w.u, w.v, cfl, **kw)


def test_noalloc_cfl_matches_the_context_entry_point(P2D):
    import torch
    N = P2D._native
    W = P2D.workloads
    stream = torch.cuda.current_stream(torch.device("cuda:0"))
    kw = dict(rho_minus=1.0, rho_plus=1000.0, sigma=0.0728, mu_minus=1.78e-5, mu_plus=1.137e-3)
    params = P2D.Context._step_params(kw)
    for name, n in (("drops64", 512), ("circle", 128)):      # second call on a smaller grid
        w = _case(W, name, n)
        print(f"case cfl noalloc {name} n={n}", flush=True)
        got = _noalloc(N, torch, w, params, 0.5, stream)
        ref = _reference(P2D, w, 0.5, **kw)
        assert ref > 0 and abs(got - ref) <= 1e-15 * ref, (got, ref)
    # a rectangular grid without surface tension (phi not read)
    w = _case(W, "rt", 256)
    w.phi = np.ascontiguousarray(w.phi[:96, :])
    w.u = np.ascontiguousarray(w.u[:96, :])
    w.v = np.ascontiguousarray(w.v[:97, :])
    kw0 = dict(kw, sigma=0.0)
    got = _noalloc(N, torch, w, P2D.Context._step_params(kw0), 0.9, stream)
    ref = _reference(P2D, w, 0.9, **kw0)
    assert abs(got - ref) <= 1e-15 * ref, (got, ref)


def test_noalloc_cfl_error_flag_and_limit(P2D):
    import torch
    N = P2D._native
    W = P2D.workloads
    stream = torch.cuda.current_stream(torch.device("cuda:0"))
    kw = dict(rho_minus=1.0, rho_plus=1000.0, sigma=0.0728)
    params = P2D.Context._step_params(kw)
    w = _case(W, "circle", 64)
    flat = _case(W, "circle", 64)
    flat.phi = np.zeros_like(flat.phi)                       # zero gradient: curvature undefined
    print("case cfl noalloc zero gradient", flush=True)
    with pytest.raises(P2D.P2DError) as e:
        _reference(P2D, flat, 0.5, **kw)
    assert e.value.status == N.P2D_NAN
    assert np.isnan(_noalloc(N, torch, flat, params, 0.5, stream))
    got = _noalloc(N, torch, w, params, 0.5, stream)          # the flag does not stick
    ref = _reference(P2D, w, 0.5, **kw)
    assert abs(got - ref) <= 1e-15 * ref
    # beyond the static scratch (65 x 4097 blocks of 256 x 4 > P2D_NOALLOC_MAX_BLOCKS): refused on
    # the host before any launch, so the one-element buffers are never read
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    p = dummy.data_ptr()
    assert N.lib.p2d_cfl_dt_noalloc(16384, 16384, 1.0 / 16384, p, p, p, ctypes.byref(params), 0.5, p,
                                    stream.cuda_stream) == N.P2D_BAD_ARG
