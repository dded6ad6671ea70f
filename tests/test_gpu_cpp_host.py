"""The C++ host side of the drop-in (include/p2d_projection.hpp): the driver
tests/cpp/projection_demo (built next to libp2d.so by the package Makefile) runs
the config-0 case through p2d::PROJECTION_2D and must match the golden fixture."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DEMO = os.path.join(HERE, "cpp", "projection_demo")


@pytest.mark.parametrize("prec", [0, 1])
def test_cpp_host_matches_golden(W, tmp_path, prec):
    if not os.path.exists(DEMO):
        pytest.fail("tests/cpp/projection_demo is not built (run __graft_entry__.build())")
    w = W.circle(128)
    for name, a in (("phi", w.phi), ("u", w.u), ("v", w.v)):
        np.ascontiguousarray(a, dtype=np.float64).tofile(str(tmp_path / f"{name}.bin"))
    print(f"case C++ host circle 128 prec={prec}", flush=True)
    r = subprocess.run([DEMO, str(tmp_path), "128", str(prec), str(w.rho_minus), str(w.rho_plus), str(w.sigma)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    g = np.load(os.path.join(HERE, "golden", "circle128.npz"))
    p = np.fromfile(str(tmp_path / "p_out.bin")).reshape(128, 128)
    assert p[0, 0] == 0.0            # pressure_field in the reference's gauge
    p = p - p.mean()
    u = np.fromfile(str(tmp_path / "u_out.bin")).reshape(128, 129)
    v = np.fromfile(str(tmp_path / "v_out.bin")).reshape(129, 128)
    assert np.linalg.norm(p - g["p"]) <= 1e-9 * np.linalg.norm(g["p"])
    scale = max(np.abs(g["u"]).max(), np.abs(g["v"]).max())
    assert max(np.abs(u - g["u"]).max(), np.abs(v - g["v"]).max()) <= 1e-8 * scale
