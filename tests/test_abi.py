"""The C-ABI library loads and exports every symbol include/p2d.h declares (no GPU needed)."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "p2d.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(p2d_[a-z_0-9]+)\s*\(", src)))


def test_header_symbols_exported():
    lib = os.path.join(ROOT, "2d_engine_amd", "libp2d.so")
    assert os.path.exists(lib), "build first: __graft_entry__.build()"
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (p2d_\w+)", out))
    declared = _declared()
    assert len(declared) >= 15
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    L = ctypes.CDLL(lib)
    for s in declared:
        getattr(L, s)


def test_binding_matches_header(P2D):
    from importlib import import_module
    N = import_module("2d_engine_amd._native")
    assert sorted(N.EXPORTS) == _declared()


def test_bad_arguments_rejected_without_compute(P2D):
    N = P2D._native
    cfg = N.default_config(1, 8, 0.1)          # nx < 2
    h = ctypes.c_void_p()
    assert N.lib.p2d_create(ctypes.byref(cfg), ctypes.byref(h)) == N.P2D_BAD_ARG
    cfg = N.default_config(8, 8, 0.125)
    cfg.rtol = 0.0
    cfg.div_tol = 0.0                           # no stopping rule
    assert N.lib.p2d_create(ctypes.byref(cfg), ctypes.byref(h)) == N.P2D_BAD_ARG
    cfg = N.default_config(8, 8, 0.125)
    cfg.inner_rtol = 1.5
    assert N.lib.p2d_create(ctypes.byref(cfg), ctypes.byref(h)) == N.P2D_BAD_ARG
    for om in (0.0, 2.0, float("nan")):         # SOR needs 0 < omega < 2
        for field in ("mg_omega", "mg_omega_coarse"):
            cfg = N.default_config(8, 8, 0.125)
            setattr(cfg, field, om)
            assert N.lib.p2d_create(ctypes.byref(cfg), ctypes.byref(h)) == N.P2D_BAD_ARG
    assert N.status_string(N.P2D_NAN) == "NaN or breakdown"


def test_array_size_limit_checked_before_any_device_call(P2D):
    """The tile kernels use 32-bit byte offsets: a grid whose FP64 cell array reaches 2^32
    bytes is refused with P2D_BAD_ARG before the library touches a device, while one just
    below the limit gets past every argument check (here, with no GPU, to P2D_NO_DEVICE)."""
    N = P2D._native
    h = ctypes.c_void_p()
    for n in (23552, 24576):                    # 23*1024 and 24*1024: legal bottoms, > 4 GiB arrays
        cfg = N.default_config(n, n, 1.0 / n)
        assert n * n * 8 > 0xFFFFFFFF
        assert N.lib.p2d_create(ctypes.byref(cfg), ctypes.byref(h)) == N.P2D_BAD_ARG
    import torch
    if torch.cuda.device_count() > 0:
        return   # the grid below the limit would allocate ~230 GB on a device
    cfg = N.default_config(32768, 15360, 1.0 / 32768)   # 503 M cells, bottom 32 x 15
    assert 32768 * 15360 * 8 <= 0xFFFFFFFF
    assert N.lib.p2d_create(ctypes.byref(cfg), ctypes.byref(h)) == N.P2D_NO_DEVICE


def test_status_codes_match_header():
    src = open(os.path.join(ROOT, "include", "p2d.h")).read()
    codes = dict(re.findall(r"(P2D_[A-Z_]+)\s*=\s*(\d+)", src))
    from importlib import import_module
    N = import_module("2d_engine_amd._native")
    for k in ("P2D_OK", "P2D_NOT_CONVERGED", "P2D_NAN", "P2D_BAD_ARG", "P2D_HIP_ERROR", "P2D_NO_DEVICE"):
        assert int(codes[k]) == getattr(N, k)


def test_cpp_header_compiles(tmp_path):
    """include/p2d_projection.hpp (the C++ host side) and p2d.h compile as C++17 without HIP headers."""
    import subprocess
    inc = os.path.join(ROOT, "include")
    src = tmp_path / "t.cpp"
    src.write_text('#include "p2d_projection.hpp"\nint main() { p2d::PROJECTION_2D p(64, 64, 1.0 / 64); return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", inc, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_struct_layouts_match_ctypes(tmp_path, P2D):
    """p2d.h compiles as C99, and every struct the ctypes binding mirrors has the same size
    and field offsets as the C compiler gives it (a mismatch would corrupt every call)."""
    import ctypes
    N = P2D._native
    structs = {"p2d_config": N.P2DConfig, "p2d_stats": N.P2DStats, "p2d_step_params": N.P2DStepParams,
               "p2d_oil_water_params": N.P2DOilWaterParams}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "p2d.h"', 'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_parameter_defaults_and_visc_argument_checks(P2D):
    """The step and oil-water defaults are the reference script's values, and the viscosity
    solver rejects bad grids before touching the GPU."""
    import ctypes
    N = P2D._native
    prm = N.P2DStepParams()
    N.lib.p2d_step_params_default(ctypes.byref(prm))
    assert (prm.gravity_x, prm.gravity_y) == (0.0, -9.8)
    assert prm.reinit_iterations == 25 and prm.reinit_scaling == 5.0 and prm.weno_epsilon == 1e-6
    assert prm.rk_order == 3                    # the scripts' order_for_time_advancing
    ow = N.P2DOilWaterParams()
    N.lib.p2d_oil_water_params_default(ctypes.byref(ow))
    assert (ow.Re, ow.rho_oil, ow.rho_water, ow.mu_oil, ow.mu_water) == (0.94983, 0.905, 0.995, 6.01, 0.01)
    assert ow.tolerance == 1e-6 and ow.max_iterations == 10000 and ow.boundary == N.VISC_REFERENCE
    h = ctypes.c_void_p()
    for nx, ny, dx, dy in ((1, 8, 0.1, 0.1), (8, 1, 0.1, 0.1), (8, 8, 0.0, 0.1), (8, 8, 0.1, -1.0)):
        assert N.lib.p2d_visc_create(nx, ny, dx, dy, ctypes.byref(h)) == N.P2D_BAD_ARG
    assert N.lib.p2d_visc_solve(None, 0, None, None, None, None, None, None, None, 0, 1e-6, 10, None, 0) == \
        N.P2D_BAD_ARG
