"""The single-pass (fused) collide-stream kernel: every tile shape and z-chunk
length of its tile table gives the same populations as the two-kernel step
(density pre-pass + collide-push) on the same input, to FP32 rounding: the
instantiations are compiled separately (1024-lane tiles run under a tighter
register budget) and agree to an ulp or two, not always bit for bit.  The
fused kernel reads neighbour phase fields from an LDS ring filled from a halo
of re-summed nodes; a wrong halo index, a stale ring slot or a wrong periodic
wrap changes populations by 1e-3 or more, far above the tolerance here.  Tile
entries from 10 on keep two planes of loads in flight (a different pipeline
through the same arithmetic); they are covered the same way.

Also: the autotuner leaves the state untouched and picks a valid tile.
"""
import numpy as np
import pytest

from lbcuda_amd.abi import default_config, INIT_RHO, KERNEL_FUSED, KERNEL_TWOPASS, WET_NEUTRAL
from lbcuda_amd import workloads as wl
from lbcuda_amd.lib import Box, tile_shapes

pytestmark = pytest.mark.gpu


def run(shape, steps, variant, tile=0, zc=0, seed=7, pos=None, cfg_kw=None):
    cfg = default_config(shape[0], sigma=0.02, beta=0.7, **(cfg_kw or {}))
    cfg.nx, cfg.ny, cfg.nz = shape
    g = Box(cfg)
    if variant == KERNEL_FUSED:
        g.set_kernel(KERNEL_FUSED, tile, zc)
    else:
        g.set_kernel(KERNEL_TWOPASS)
    if pos is not None:
        g.set_particles(pos)
    rr, rb = wl.spinodal_rho(shape, seed=seed)
    g.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    assert g.step(steps) == 0
    out = (g.pops(0), g.pops(1), g.flags())
    g.close()
    return out


def valid_tiles(nx):
    return [i for i, (tx, ty, wrap) in enumerate(tile_shapes()) if not wrap or tx == nx]


# FP32 rounding of one step, accumulated over 4 steps, for values <= 1.  The
# fused kernel forms the collision with per-node constants and 1-ulp
# reciprocals, the two-kernel step in the reference's operation order, so they
# differ by a few ulps per step (4e-7 measured), not bit for bit.
ULP_TOL = 1e-6


@pytest.mark.parametrize("shape", [(64, 24, 10), (70, 37, 11), (128, 12, 9), (256, 8, 5)])
def test_all_tiles_agree(shape):
    print(f"tiles shape={shape}", flush=True)
    two = run(shape, 4, KERNEL_TWOPASS)
    for t in valid_tiles(shape[0]):
        for zc in sorted({1, 3, shape[2]}):
            got = run(shape, 4, KERNEL_FUSED, t, zc)
            err = max(float(np.abs(got[c] - two[c]).max()) for c in (0, 1))
            print(f"  tile {t} {tile_shapes()[t]} zc {zc}: max |fused - two-pass| = {err:.2e}", flush=True)
            assert err < ULP_TOL, (t, zc, err)


def test_fused_equals_twopass_with_static_particles():
    """Particles that do not cross nodes: the fused step reads surface and
    interior densities from the arrays the particle kernels fill, so it must
    match the two-kernel step exactly."""
    print("static particles fused vs two-pass", flush=True)
    n = 48
    pos = wl.random_spheres((n, n, n), 4, 5.5, seed=5)
    pos[3:9] = 0.0
    kw = dict(with_particles=1, n_particles=4, lrotate=1, rmass=0.0, max_vel=1.0, wetting=WET_NEUTRAL)
    two = run((n, n, n), 6, KERNEL_TWOPASS, pos=pos, cfg_kw=kw)
    fluid = (two[2] == 1)[..., None].repeat(19, -1)
    tiles = valid_tiles(n)
    for t in tiles[:3] + [t for t in tiles if t >= 10][:2]:
        got = run((n, n, n), 6, KERNEL_FUSED, t, 7, pos=pos, cfg_kw=kw)
        assert np.array_equal(got[2], two[2])
        for c in (0, 1):
            err = float(np.abs(got[c][fluid] - two[c][fluid]).max())
            print(f"  tile {t} colour {c}: max |fused - two-pass| on fluid nodes = {err:.2e}", flush=True)
            assert err < ULP_TOL, (t, c, err)


def test_autotune_keeps_state():
    print("autotune", flush=True)
    n = 64
    cfg = default_config(n, sigma=0.02, beta=0.7)
    g = Box(cfg)
    rr, rb = wl.spinodal_rho((n, n, n), seed=3)
    g.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    g.step(2)
    before = g.pops(0)
    t, zc, ms = g.autotune(reps=2)
    assert 0 <= t < len(tile_shapes()) and 1 <= zc <= n and ms > 0
    assert g.get_kernel() == (KERNEL_FUSED, t, zc)
    assert np.array_equal(g.pops(0), before)
    assert g.step(3) == 0
    g.close()
