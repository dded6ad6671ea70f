"""Golden fixtures (tests/golden, written by tests/golden/make_golden.py).

The config-0 case (circle 128^2): the seeded generators must still produce the
recorded inputs, and the oracle must still reproduce the recorded mean-zero
pressure and projected velocities (IC(0)-PCG to relative residual 1e-13).
"""
import hashlib
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "circle128.npz")


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def test_golden_inputs_are_reproduced(W):
    g = np.load(GOLDEN)
    w = W.circle(128)
    assert _digest(w.phi) == str(g["phi_sha256"])
    assert _digest(w.u) == str(g["u_sha256"])
    assert _digest(w.v) == str(g["v_sha256"])


def test_oracle_reproduces_golden(O, W):
    g = np.load(GOLDEN)
    w = W.circle(128)
    p, u, v, st = O.project(w.phi, w.rho_minus, w.rho_plus, w.sigma, w.u, w.v, w.h, solver=O.SOLVER_IC0, tol=1e-13)
    p = p - p.mean()
    assert np.linalg.norm(p - g["p"]) <= 1e-12 * np.linalg.norm(g["p"])
    assert np.abs(u - g["u"]).max() <= 1e-12 and np.abs(v - g["v"]).max() <= 1e-12
    assert st["iterations"] == int(g["iterations"])
