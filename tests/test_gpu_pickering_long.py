"""The headline workload over its full length: 256^3 two-component colour
gradient + 1280 spheres of radius 5.5, 2000 steps, in FP32 and in shifted
FP16 population storage (the FP64 oracle covers the first 100 steps in
test_gpu_pickering.py; 2000 steps would take it about twenty minutes).

Checked after the run: no device error was raised (density range, force,
velocity, out of box, storage range, neighbour list), every sphere still has
exactly one centre node and the node types are the four a sphere writes,
the particle state is finite and within the velocity guard, each colour's
fluid mass has moved by less than a percent (fill and erase move mass
between the fluid and the spheres' nodes), and the two storage formats
coarsen alike (the runs are chaotic, their statistics are not).
Measured: colour mass change 4.7e-4 / 6.3e-4 (FP32) and 2.0e-4 / 3.2e-4
(FP16), red fraction 0.5000 and interface fraction 0.2049 in both."""
import numpy as np
import pytest

from lbcuda_amd.abi import INIT_RHO, STORE_F32, STORE_F16
from lbcuda_amd import workloads as wl
from lbcuda_amd.lib import Box

pytestmark = pytest.mark.gpu

STEPS = 2000


def run(storage, rr, rb, pos):
    cfg = wl.pickering_config(256, 1280, storage=storage)
    g = Box(cfg)
    g.set_particles(pos)
    g.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    s0 = g.sums()
    for s in range(0, STEPS, 500):
        assert g.step(500) == 0
        print(f"  storage {storage}: step {s + 500}", flush=True)
    s1 = g.sums()
    fl = g.flags()
    p, f = g.particles()
    gr, gb = g.rho()
    g.close()
    assert int((fl == 5).sum()) == 1280
    assert set(np.unique(fl).tolist()) <= {1, 2, 4, 5}
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(f))
    assert np.all(np.sqrt((p[3:6] ** 2).sum(0)) <= cfg.max_vel)
    dm = [abs(s1[k] - s0[k]) / s0[k] for k in (0, 1)]
    fluid = fl == 1
    ph = (gr[fluid] - gb[fluid]) / (gr[fluid] + gb[fluid])
    red, interface = float((ph > 0).mean()), float((np.abs(ph) < 0.5).mean())
    d = p[0:3] - pos[0:3]
    disp = float(np.abs(d - 256.0 * np.rint(d / 256.0)).max())   # minimum image
    print(f"  storage {storage}: colour mass change {dm[0]:.2e} / {dm[1]:.2e}, red fraction {red:.4f}, "
          f"interface fraction {interface:.4f}, largest displacement {disp:.3f}", flush=True)
    assert max(dm) < 1e-2
    return red, interface


def test_pickering_256_2000_steps_fp32_and_fp16():
    print(f"pickering 256^3 + 1280 spheres x {STEPS} in FP32 and FP16", flush=True)
    rr, rb, pos = wl.pickering_inputs(256, 1280)
    red32, if32 = run(STORE_F32, rr, rb, pos)
    red16, if16 = run(STORE_F16, rr, rb, pos)
    assert abs(red16 - red32) < 0.02
    assert abs(if16 / if32 - 1.0) < 0.15
