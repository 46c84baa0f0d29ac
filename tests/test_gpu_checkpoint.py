"""Checkpoint images (lbm_checkpoint_save / lbm_checkpoint_load).

A box saved after some steps and loaded into a fresh context of the same
configuration continues bit for bit as the saved box does: populations,
densities, node flags and the particle state after further steps are equal,
for FP32 and shifted FP16 storage, with moving, rotating spheres whose node
crossings (erase and fill) fall both before and after the save, and for both
collide-stream variants.  Images of another configuration or particle count
are refused.
"""
import numpy as np
import pytest

from lbcuda_amd.abi import (default_config, INIT_RHO, KERNEL_FUSED, KERNEL_TWOPASS, STORE_F16, STORE_F32,
                            WET_JANUS, WET_NEUTRAL)
from lbcuda_amd import workloads as wl
from lbcuda_amd.lib import Box, LbmError

pytestmark = pytest.mark.gpu

N = 40
NPART = 4


def case(storage, wetting, seed=21):
    cfg = default_config(N, sigma=0.02, beta=0.7, storage=storage, with_particles=1, n_particles=NPART,
                         lrotate=1, max_vel=1.0, wetting=wetting)
    pos = wl.random_spheres((N, N, N), NPART, 5.5, seed=seed)
    rng = np.random.default_rng(seed)
    pos[3:6] = 0.08 * rng.uniform(-1, 1, (3, NPART))
    pos[6:9] = 0.01 * rng.uniform(-1, 1, (3, NPART))
    rr, rb = wl.spinodal_rho((N, N, N), seed=seed)
    return cfg, pos, rr, rb


def state(g):
    rr, rb = g.rho()
    pos, frc = g.particles()
    return {"pops_r": g.pops(0), "pops_b": g.pops(1), "rho_r": rr, "rho_b": rb, "flags": g.flags(),
            "pos": pos, "force": frc}


@pytest.mark.parametrize("variant", [KERNEL_FUSED, KERNEL_TWOPASS])
@pytest.mark.parametrize("storage,wetting", [(STORE_F32, WET_NEUTRAL), (STORE_F16, WET_JANUS)])
def test_checkpoint_continues_bit_for_bit(storage, wetting, variant):
    print(f"checkpoint storage={storage} wetting={wetting} variant={variant}", flush=True)
    cfg, pos, rr, rb = case(storage, wetting)
    a = Box(cfg)
    a.set_kernel(variant)
    a.set_particles(pos)
    a.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    assert a.step(9) == 0
    image = a.checkpoint()
    assert a.step(11) == 0
    want = state(a)
    a.close()

    b = Box(cfg)
    b.set_kernel(variant)
    b.set_particles(pos)          # allocates the particle buffers; the image overwrites them
    b.restore(image)
    assert b.step(11) == 0
    got = state(b)
    b.close()
    # the spheres' nearest nodes changed: erase and fill passes ran
    assert np.any(np.rint(want["pos"][0:3]) != np.rint(pos[0:3]))
    bad = {k: int(np.count_nonzero(got[k] != want[k])) for k in want}
    assert not any(bad.values()), bad


def test_checkpoint_refuses_other_boxes():
    print("checkpoint refusals", flush=True)
    cfg, pos, rr, rb = case(STORE_F32, WET_NEUTRAL)
    a = Box(cfg)
    a.set_particles(pos)
    a.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    a.step(2)
    image = a.checkpoint()
    a.close()

    other = default_config(N, sigma=0.03, beta=0.7, with_particles=1, n_particles=NPART, lrotate=1,
                           max_vel=1.0, wetting=WET_NEUTRAL)
    b = Box(other)
    b.set_particles(pos)
    with pytest.raises(LbmError) as e:
        b.restore(image)
    assert e.value.code == 1
    b.close()

    c = Box(cfg)               # particles not set: no particle buffers of that count
    with pytest.raises(LbmError) as e:
        c.restore(image)
    assert e.value.code == 9
    c.close()

    d = Box(cfg)
    d.set_particles(pos)
    with pytest.raises(LbmError) as e:
        d.restore(image[: image.size // 2])
    assert e.value.code == 1
    d.restore(image)
    assert d.step(1) == 0
    d.close()
