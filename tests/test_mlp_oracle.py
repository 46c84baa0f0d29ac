"""Pins the closure-network oracle (oracle/mlp_oracle.py): its analytic
gradient against central finite differences of its own loss, in float64, on
every parameter block; and the layout/size of the parameter vector the C ABI
declares."""
import re
import os

import numpy as np

from oracle import mlp_oracle as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_parameter_count_matches_header():
    h = open(os.path.join(ROOT, "include", "lbcuda_amd.h")).read()
    n = int(re.search(r"#define LBM_MLP_NPARAM (\d+)", h).group(1))
    assert n == M.NPARAM == 4 * 64 + 64 + 64 * 64 + 64 + 3 * 64 + 3


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(3).astype(np.float64) + 0.05 * rng.standard_normal(M.NPARAM)
    x = rng.standard_normal((37, 4))
    t = rng.standard_normal((37, 3))
    _, g = M.loss_grad(p, x, t)
    # a sample of every block: W1, b1, W2, b2, W3, b3
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SIZES])
    picks = np.concatenate([rng.integers(offs[k], offs[k + 1], 6) for k in range(len(M.SIZES))])
    h = 1e-6
    for i in picks:
        e = np.zeros(M.NPARAM)
        e[i] = h
        fd = (M.loss(p + e, x, t) - M.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_linear_output_layer_fits_exactly():
    """With W3/b3 solved by least squares on fixed hidden features, the
    gradient w.r.t. W3 and b3 vanishes (a closed-form optimum)."""
    rng = np.random.default_rng(2)
    p = M.init_params(5).astype(np.float64)
    x = rng.standard_normal((50, 4))
    t = rng.standard_normal((50, 3))
    _, (_, _, a2) = M.forward(p, x)
    A = np.hstack([a2, np.ones((50, 1))])
    sol, *_ = np.linalg.lstsq(A, t, rcond=None)
    w = M.unpack(p)
    w["W3"] = sol[:64].T
    w["b3"] = sol[64]
    _, g = M.loss_grad(M.pack(w), x, t)
    d = M.unpack(g)
    assert np.abs(d["W3"]).max() < 1e-10 and np.abs(d["b3"]).max() < 1e-10
