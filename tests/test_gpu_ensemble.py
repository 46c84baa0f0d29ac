"""Ensemble of independent boxes advanced together (one persistent fluid
launch per step over a work list of all boxes' tiles, one launch per
particle kernel for all boxes' particles) against the same boxes stepped one
by one through lbm_step with the same tile and z chunk.

The per-box work is the same arithmetic in a different launch shape: node
types and particle trajectories must agree to FP32/FP64 rounding, and
populations to the few-ulp tolerance of the single-box tile tests."""
import numpy as np
import pytest

from lbcuda_amd.abi import KERNEL_FUSED, INIT_RHO
from lbcuda_amd import workloads as wl
from lbcuda_amd.lib import Box, Ensemble

pytestmark = pytest.mark.gpu

ULP_TOL = 4e-7


def small_ensemble():
    cfgs, inputs = wl.ensemble_workload(nbox=5, seed=3, edge_range=(28, 44), spheres_range=(1, 3),
                                        radius_range=(4.0, 6.0))
    for b, (rr, rb, pos) in enumerate(inputs):
        rng = np.random.default_rng(b)
        pos[3:6] = 0.05 * rng.uniform(-1, 1, (3, pos.shape[1]))
        pos[6:9] = 0.01 * rng.uniform(-1, 1, (3, pos.shape[1]))
    return cfgs, inputs


@pytest.mark.parametrize("tile,zc", [(0, 7), (1, 16)])
def test_ensemble_matches_boxes_stepped_alone(tile, zc):
    print(f"ensemble tile={tile} zc={zc}", flush=True)
    cfgs, inputs = small_ensemble()
    ens = Ensemble(cfgs)
    ens.set_tile(tile, zc)
    for b, (rr, rb, pos) in zip(ens.boxes, inputs):
        b.set_particles(pos)
        b.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    assert ens.step(12) == 0
    for i, (cfg, (rr, rb, pos)) in enumerate(zip(cfgs, inputs)):
        print(f"  box {i} {cfg.nx}x{cfg.ny}x{cfg.nz} spheres {cfg.n_particles} R {cfg.radius:.2f}", flush=True)
        g = Box(cfg)
        g.set_kernel(KERNEL_FUSED, tile, zc)
        g.set_particles(pos)
        g.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
        assert g.step(12) == 0
        e = ens.boxes[i]
        e.npart = g.npart
        assert np.array_equal(e.flags(), g.flags())
        fluid = (g.flags() == 1)[..., None].repeat(19, -1)
        for c in (0, 1):
            err = float(np.abs(e.pops(c) - g.pops(c))[fluid].max())
            assert err < ULP_TOL, (i, c, err)
        ep, ef = e.particles()
        gp, gf = g.particles()
        assert np.abs(ep - gp).max() < 1e-9
        assert np.all(np.abs(ef - gf) <= 1e-6 * (1 + np.abs(gf)))
        g.close()
    ens.close()


def test_ensemble_autotune_keeps_state():
    print("ensemble autotune", flush=True)
    cfgs, inputs = small_ensemble()
    ens = Ensemble(cfgs)
    for b, (rr, rb, pos) in zip(ens.boxes, inputs):
        b.set_particles(pos)
        b.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    assert ens.step(2) == 0
    before = [b.pops(0) for b in ens.boxes]
    t, zc, ms = ens.autotune(reps=2)
    assert 0 <= t <= 4 and zc >= 1 and ms > 0
    for b, p in zip(ens.boxes, before):
        assert np.array_equal(b.pops(0), p)
    assert ens.step(2) == 0
    ens.close()
