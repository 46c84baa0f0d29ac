"""Learned near-contact closure on the GPU: the network kernels against the
float64 CPU oracle (oracle/mlp_oracle.py, pinned by finite differences),
training with the library's own kernels, and the closure inside the
pair-force kernel against the same network evaluated by the batch kernel.

Tolerances: forward 2e-6 absolute (FP32, tanh of inputs of order 1);
gradients 1e-4 relative to the largest component (FP32 sums over the batch).
"""
import math
import os

import numpy as np
import pytest

from lbcuda_amd import lib as L
from lbcuda_amd import closure as C
from lbcuda_amd.abi import default_config, INIT_RHO, WET_NEUTRAL
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def random_problem(n, seed):
    rng = np.random.default_rng(seed)
    p = (M.init_params(seed) + 0.05 * rng.standard_normal(M.NPARAM)).astype(np.float32)
    x = rng.standard_normal((n, 4)).astype(np.float32)
    t = rng.standard_normal((n, 3)).astype(np.float32)
    return p, x, t


def test_forward_matches_oracle():
    p, x, _ = random_problem(300, 1)
    y_gpu = L.mlp_forward(p, x)
    y_cpu, _ = M.forward(p.astype(np.float64), x.astype(np.float64))
    assert np.abs(y_gpu - y_cpu).max() < 2e-6


@pytest.mark.parametrize("n", [1, 64, 200])
def test_gradient_matches_oracle(n):
    p, x, t = random_problem(n, 2 + n)
    loss_gpu, g_gpu = L.mlp_loss_grad(p, x, t)
    loss_cpu, g_cpu = M.loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(loss_gpu - loss_cpu) <= 1e-5 * abs(loss_cpu)
    assert np.abs(g_gpu - g_cpu).max() <= 1e-4 * np.abs(g_cpu).max()


def test_training_fits_a_smooth_map():
    rng = np.random.default_rng(5)
    x = rng.uniform(-1.5, 1.5, (1024, 4)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0]) * x[:, 1], 0.5 * x[:, 2] ** 2 - x[:, 3], np.tanh(x[:, 0] + x[:, 3])], 1)
    p0 = C.glorot(3)
    l0 = M.loss(p0.astype(np.float64), x, t)
    p, l1 = L.mlp_train(p0, x, t.astype(np.float32), epochs=300, batch=64, lr=3e-3, seed=7)
    print(f"training loss {l0:.3e} -> {l1:.3e}")
    assert l1 < 2e-3 and l1 < 0.01 * l0
    # the loss the training reports is the oracle's loss of the returned weights
    assert abs(l1 - M.loss(p.astype(np.float64), x, t)) < 1e-4 * max(l1, 1e-6) + 1e-7


def test_closure_force_in_pair_kernel():
    """Two spheres with prescribed velocities close from gap 1 to about 0.6
    while sliding.  Trajectories and fluid do not depend on the pair forces
    (1/m = 0), so the force on sphere 1 with the closure on minus the force
    with it off is the closure term alone: -F_n n - F_t t and torque
    T (n x t) of the network evaluated on the pair's features."""
    n = 40
    R = 5.0
    kw = dict(sigma=0.0, beta=0.0, with_particles=1, n_particles=2, radius=R, lrotate=1, rmass=0.0, max_vel=1.0,
              wetting=WET_NEUTRAL, hertz=0, lubrication=0, closure=1)
    pos = np.zeros((13, 2), order="F")
    pos[0:3, 0] = (14.2, 20.1, 19.9)
    pos[0:3, 1] = (14.2 + 11.0, 20.1, 19.9)
    pos[3:6, 0] = (0.002, 0.001, 0.0)
    pos[3:6, 1] = (-0.002, -0.0005, 0.0)
    pos[9, :] = 1.0
    rr = np.ones((n, n, n), np.float32, order="F")
    rb = np.zeros_like(rr)
    rng = np.random.default_rng(9)
    prm = (C.glorot(4) + 0.1 * rng.standard_normal(L.MLP_NPARAM)).astype(np.float32)
    mean = np.array([0.1, 0.002, 0.001, 5.0], np.float32)
    scale = np.array([0.1, 0.002, 0.001, 2.0], np.float32)
    oscale = np.array([0.05, 0.02, 0.1], np.float32)
    lo = np.array([0.0, -1.0, 0.0, 1.0], np.float32)
    hi = np.array([1.0, 1.0, 1.0, 10.0], np.float32)
    block = np.concatenate([prm, mean, scale, oscale, lo, hi, [1.0]]).astype(np.float32)
    forces, positions = [], []
    for on in (False, True):
        cfg = default_config(n, **kw)
        b = L.Box(cfg)
        b.set_particles(pos)
        b.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
        if on:
            b.set_closure(block)
        assert b.step(100) == 0
        p, f = b.particles()
        forces.append(f[:, 0].copy())
        positions.append(p.copy())
        b.close()
    assert np.array_equal(positions[0], positions[1])
    p = positions[0]
    x1 = p[0:3, 0] - p[3:6, 0]     # positions the last pair-force pass saw
    x2 = p[0:3, 1] - p[3:6, 1]
    d = x2 - x1
    r = np.linalg.norm(d)
    nv = d / r
    u = p[3:6, 0] - p[3:6, 1]
    vn = u @ nv
    tv = u - vn * nv
    vt = np.linalg.norm(tv)
    tv /= vt
    feat = np.array([(r - 2 * R) / R, vn, vt, R], np.float32)
    assert 0.05 < feat[0] < 0.2
    y = L.mlp_forward(prm, ((feat - mean) / scale)[None, :])[0] * oscale
    expect = -y[0] * nv - y[1] * tv
    got = forces[1][0:3] - forces[0][0:3]
    print("closure force", got, "expected", expect)
    assert np.abs(got - expect).max() <= 1e-5 * (1 + np.abs(expect).max())
    torque = forces[1][6:9] - forces[0][6:9]
    assert np.abs(torque - y[2] * np.cross(nv, tv)).max() <= 1e-5 * (1 + abs(y[2]))


def test_pair_run_smoke():
    """The data generator runs a small coarse/fine pair (R = 3, 2x) and
    returns finite coefficients with the gap closing to about 0.1/R."""
    c = C.run_pair(3.0, 0.0, 1, sample_every=20)
    f = C.run_pair(3.0, 0.0, 2, sample_every=80)
    for r in (c, f):
        assert np.all(np.isfinite(r["Cn"])) and np.all(np.isfinite(r["Bn"])) and r["gap"][-1] < 0.1
    # an approaching pair is pushed apart while a fluid layer separates it
    far = c["gap"] > 0.5
    assert c["Bn"][far].mean() > 0 and f["Bn"][f["gap"] > 0.5].mean() > 0


@pytest.fixture(scope="module")
def trained():
    """The closure trained with lbm_mlp_train on the committed pair-approach
    data, every third gap bin of each run held out."""
    train, held = C.split(C.load())
    block, loss = C.fit(train)
    print(f"closure training loss {loss:.3e} on {len(train)} bins")
    return block, train, held


def test_learned_closure_gain_on_held_out_bins(trained):
    """On gap bins it has not seen, coarse force + learned correction is
    closer to the 4x-resolved force than the coarse force alone and than the
    coarse force + the reference's lubrication term."""
    block, train, held = trained
    e = C.evaluate(block, held)
    tot = {k: float(np.sqrt((v ** 2).sum())) for k, v in e.items()}
    print({k: np.round(v, 4).tolist() for k, v in e.items()}, tot)
    assert tot["learned"] < 0.7 * tot["none"]
    assert e["learned"][0] < 0.7 * e["lubrication"][0]
    # normal force and torque each improve
    assert e["learned"][0] < e["none"][0] and e["learned"][2] < e["none"][2]


@pytest.mark.parametrize("angle", [0.0, 40.0])
def test_learned_closure_in_solver(trained, angle):
    """The trained block set on a coarse box (R = 6): the pair-force kernel
    evaluates the network on each step's features and adds it to the
    hydrodynamic force.  Per held-out gap bin, the medians of the total
    force coefficients move towards the 4x-resolved ones."""
    block, train, held = trained
    R = 6.0
    run_id = 2 + int(angle > 0)
    rows = held[held[:, 10] == run_id]
    assert len(rows) >= 5
    r = C.run_pair(R, angle, 1, closure=block)
    bins = C.gap_bins(r, R)
    got = np.array([[np.median(r[k][bins == int(b)]) for k in ("Cn", "Ct", "CT")] for b in rows[:, 11]])
    fine, coarse = rows[:, 7:10], rows[:, 4:7]
    err_on = np.sqrt(((got - fine) ** 2).sum())
    err_off = np.sqrt(((coarse - fine) ** 2).sum())
    print(f"angle {angle}: |C - C_fine| without closure {err_off:.4f}, with {err_on:.4f}")
    assert err_on < 0.8 * err_off
