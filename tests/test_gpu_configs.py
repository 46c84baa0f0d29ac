"""The workload configurations at their full sizes on the GPU.

Each case runs the stated box and step count on the GPU and checks what holds
at any size (decay laws, conservation, invariants of the node types, agreement
between storage formats); the FP64 oracle is run beside it on the same input
for as many steps as it finishes in seconds at that size, and the GPU state
is compared with it there.

  C1  64^3 single-component BGK, Taylor-Green start u0 = 0.01, tau = 1,
      1000 steps, FP32 on the GPU (the FP64 run is the oracle's): kinetic
      energy decay against exp(-4 nu k^2 t) and against the oracle.
  C2  128^3 two-component spinodal decomposition, rho_R = 0.5 +- 0.05 U,
      sigma = 0.02, beta = 0.7, 5000 steps in FP32, FP16 and BF16 storage:
      mass per colour, phase-field error against the FP64 oracle after 100
      steps, and the coarsened structure of the three runs after 5000.
  C3  128^3 BGK, a sphere of radius 5.5 translating at 0.01 along x with
      moving bounce-back and fill/erase, 4000 steps: the node-type
      invariants at every sample, the force history against the oracle over
      the first 100 steps, and a steady drag after that.
"""
import math

import numpy as np
import pytest

from lbcuda_amd.abi import (default_config, INIT_RHO, INIT_VELFIELD, STORE_F16, STORE_BF16, STORE_F32,
                            WET_NEUTRAL)
from lbcuda_amd import workloads as wl
from lbcuda_amd.lib import Box
from oracle.pyoracle import OracleBox

pytestmark = pytest.mark.gpu

EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1])


def kinetic_energy(f):
    rho = f.sum(-1)
    u = (f * EX).sum(-1) / rho
    v = (f * EY).sum(-1) / rho
    return 0.5 * float((rho * (u * u + v * v)).sum())


def test_c1_taylor_green_64_1000_steps():
    print("C1 taylor-green 64^3 x 1000", flush=True)
    n, steps = 64, 1000
    rr, rb, vel = wl.taylor_green(n, u0=0.01)
    cfg = default_config(n)
    g = Box(cfg)
    g.init_fields(INIT_VELFIELD, rho_r=rr, rho_b=rb, vel=vel)
    o = OracleBox(cfg, 64)
    o.init_fields(INIT_VELFIELD, rho_r=rr, rho_b=rb, vel=vel)
    e0 = kinetic_energy(g.pops(0))
    nu, k = (1.0 - 0.5) / 3.0, 2.0 * math.pi / n
    for t in (250, 500, 750, 1000):
        assert g.step(250) == 0 and o.step(250) == 0
        eg, eo = kinetic_energy(g.pops(0)), kinetic_energy(o.pops(0))
        rate = -math.log(eg / e0) / t
        print(f"  step {t}: KE gpu/oracle - 1 = {eg / eo - 1.0:.2e}, decay rate / 4 nu k^2 - 1 = "
              f"{rate / (4.0 * nu * k * k) - 1.0:.2e}", flush=True)
        # FP32 against FP64: the energy decays by e^-6.4 over the run, and the
        # float weights (exact sum, second moment 1 + 7.5e-9) shift the decay
        # rate by 2e-5 of itself; the difference reaches 9e-4 at step 1000
        assert abs(eg / eo - 1.0) < 2e-3, (t, eg, eo)
        # exact Navier-Stokes decay rate; the oracle's lattice rate is 0.08 %
        # above it at k = 2 pi/64 (higher-order lattice viscosity terms)
        assert abs(rate / (4.0 * nu * k * k) - 1.0) < 2e-3, t
    fg, fo = g.pops(0), o.pops(0)
    ug = (fg * EX).sum(-1) / fg.sum(-1)
    uo = (fo * EX).sum(-1) / fo.sum(-1)
    assert np.abs(fg - fo).max() < 1e-5 and np.abs(ug - uo).max() < 5e-7
    g.close()
    o.close()


def phase(rr, rb):
    return (rr - rb) / (rr + rb)


@pytest.fixture(scope="module")
def spinodal_128():
    n = 128
    rr, rb = wl.spinodal_rho((n, n, n), seed=2024)
    return n, rr, rb


def test_c2_spinodal_128_against_oracle_100_steps(spinodal_128):
    print("C2 spinodal 128^3: oracle comparison over 100 steps", flush=True)
    n, rr, rb = spinodal_128
    o = OracleBox(default_config(n, sigma=0.02, beta=0.7), 64)
    o.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    assert o.step(100) == 0
    orr, ob = o.rho()
    o.close()
    ph_o = phase(orr, ob)
    errs = {}
    for storage in (STORE_F32, STORE_F16, STORE_BF16):
        g = Box(default_config(n, sigma=0.02, beta=0.7, storage=storage, store_rho0=0.5))
        g.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
        assert g.step(100) == 0
        gr, gb = g.rho()
        g.close()
        rel = max(np.abs(gr - orr).max(), np.abs(gb - ob).max()) / float(np.abs(orr + ob).max())
        dph = float(np.abs(phase(gr, gb) - ph_o).max())
        l2 = float(np.sqrt(((gr - orr) ** 2 + (gb - ob) ** 2).sum() / ((orr ** 2 + ob ** 2).sum())))
        print(f"  storage {storage}: density error max {rel:.2e} L2 {l2:.2e}, phase-field error {dph:.2e}",
              flush=True)
        errs[storage] = (rel, l2)
    # FP32: rounding over 100 steps of a spinodal field (4.0e-5 measured).
    # FP16: L2 1.4e-4; the largest error over 2.1 M nodes is 2.1e-3 (1e-3
    # holds at 32^3, test_low_precision_storage_error): the spinodal
    # instability amplifies the storage rounding of the early, mixed field,
    # and which node carries the largest error changes with the seed
    # (5.4e-3 for seed 7, L2 unchanged).  BF16: 1.7e-2 max, 1.1e-3 L2.
    assert errs[STORE_F32][0] < 1e-4
    assert errs[STORE_F16][0] < 3e-3 and errs[STORE_F16][1] < 1e-3
    assert errs[STORE_BF16][0] < 2e-2


def test_c2_spinodal_128_5000_steps(spinodal_128):
    print("C2 spinodal 128^3 x 5000 in FP32/FP16/BF16", flush=True)
    n, rr, rb = spinodal_128
    stats, drift = {}, {}
    for storage in (STORE_F32, STORE_F16, STORE_BF16):
        g = Box(default_config(n, sigma=0.02, beta=0.7, storage=storage, store_rho0=0.5))
        g.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
        s0 = g.sums()
        assert g.step(5000) == 0
        s1 = g.sums()
        gr, gb = g.rho()
        g.close()
        dm = max(abs(s1[0] - s0[0]) / s0[0], abs(s1[1] - s0[1]) / s0[1])
        ph = phase(gr, gb)
        red = float((ph > 0).mean())
        interface = float((np.abs(ph) < 0.5).mean())
        stats[storage] = (red, interface)
        print(f"  storage {storage}: mass drift {dm:.2e}, red fraction {red:.4f}, interface fraction "
              f"{interface:.4f}", flush=True)
        drift[storage] = dm
        # the mixture has separated: most nodes are in one phase
        assert interface < 0.5
    # FP32: the float weights sum to exactly 1 (the nearest floats to 1/3,
    # 1/18, 1/36 would add 1.5e-8 of the mass per step, 7.5e-5 here, as the
    # reference's FP32 build does); 3.8e-7 measured
    assert drift[STORE_F32] < 5e-6
    # 16-bit storage: rounding noise makes a colour's density dip below zero
    # where the other phase fills the node, and the MINDENS clamp of the
    # density pass (kernels_fluid_CG.CUF:443-446) turns that noise into a mass
    # gain: 4.9e-4 (FP16) and 5.3e-3 (BF16) after 5000 steps
    assert drift[STORE_F16] < 2e-3 and drift[STORE_BF16] < 5e-2
    # the three runs coarsen alike (the trajectories are chaotic, their
    # statistics are not)
    # (interface fraction: FP16 +2 %, BF16 +12 % measured)
    for storage, tol in ((STORE_F16, 0.05), (STORE_BF16, 0.2)):
        assert abs(stats[storage][0] - stats[STORE_F32][0]) < 0.02
        assert abs(stats[storage][1] / stats[STORE_F32][1] - 1.0) < tol


def test_c3_translating_sphere_128_4000_steps():
    print("C3 translating sphere 128^3 x 4000", flush=True)
    n = 128
    cfg = default_config(n, sigma=0.0, beta=0.0, with_particles=1, n_particles=1, rmass=0.0, max_vel=1.0,
                         wetting=WET_NEUTRAL)
    pos = np.zeros((13, 1), order="F")
    pos[0:3, 0] = (64.3, 64.5, 64.5)   # node crossings 80 steps before each sample
    pos[3, 0] = 0.01
    pos[9, 0] = 1.0
    rr = np.ones((n, n, n), np.float32)
    rb = np.zeros_like(rr)
    g = Box(cfg)
    g.set_particles(pos)
    g.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    o = OracleBox(cfg, 64)
    o.set_particles(pos)
    o.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    for _ in range(5):
        assert g.step(20) == 0 and o.step(20) == 0
        _, gf = g.particles()
        _, of = o.particles()
        assert np.all(np.abs(gf - of) <= 1e-4 * (1 + np.abs(of)))
        assert (g.flags() == o.flags()).all()
    o.close()
    drag = []
    for s in range(39):
        assert g.step(100) == 0
        fl = g.flags()
        # one sphere of radius 5.5: 739 non-fluid nodes, one centre node
        assert int((fl != 1).sum()) == 739 and int((fl == 5).sum()) == 1
        p, f = g.particles()
        drag.append(f[0, 0])
    p, _ = g.particles()
    g.close()
    assert abs(p[0, 0] - (64.3 + 0.01 * 4000)) < 1e-9
    drag = np.array(drag)
    print(f"  drag: first {drag[0]:.3e}, last {drag[-1]:.3e}", flush=True)
    # resisting and finite; the fluid, with no momentum sink, is dragged along,
    # so the resistance falls as the box fluid picks up speed
    assert np.all(drag < 0) and np.all(np.isfinite(drag))
    assert abs(drag[-1]) < abs(drag[0])
