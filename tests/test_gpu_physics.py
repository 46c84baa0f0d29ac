"""Physics of the moving-wall bounce-back and the fill/erase passes on the
GPU, against a published closed form rather than the oracle.

A free sphere pulled by a constant external force through a periodic box of
fluid (single colour): the whole system accelerates at F / (m_p + m_f), and
the sphere settles at a velocity relative to the mean fluid velocity at
which the hydrodynamic drag F m_f / (m_p + m_f) balances.  For a cubic array
of spheres the drag is Stokes' law times Hasimoto's factor (J. Fluid Mech.
5, 317 (1959)): F = 6 pi mu R U / (1 - 1.7601 c^(1/3) + c - 1.5593 c^2),
c = 4 pi R^3 / (3 L^3).  The staircase sphere of bounce-back links has a
hydrodynamic radius a little off R, so the measured factor agrees within
a few per cent (1.5 % here), not exactly; a sign error in the moving-wall
term or a missing momentum transfer changes it by far more.
"""
import math

import numpy as np
import pytest

from lbcuda_amd.abi import default_config, INIT_RHO, WET_NEUTRAL, KERNEL_FUSED, KERNEL_TWOPASS
from lbcuda_amd.lib import Box

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [KERNEL_FUSED, KERNEL_TWOPASS])
def test_settling_sphere_follows_hasimoto_drag(variant):
    print(f"settling sphere variant={variant}", flush=True)
    n, R, fext, mp = 48, 5.5, 5e-3, 1500.0
    cfg = default_config(n, sigma=0.0, beta=0.0, with_particles=1, n_particles=1, rmass=1.0 / mp,
                         max_vel=1.0, wetting=WET_NEUTRAL, ext_force=(fext, 0.0, 0.0))
    g = Box(cfg)
    if variant != KERNEL_FUSED:
        g.set_kernel(variant)
    pos = np.zeros((13, 1), order="F")
    pos[0:3, 0] = (24.3, 24.0, 23.8)
    pos[9, 0] = 1.0
    g.set_particles(pos)
    g.init_fields(INIT_RHO, rho_r=np.ones((n, n, n), np.float32), rho_b=np.zeros((n, n, n), np.float32))
    assert g.step(3000) == 0
    ks = []
    for _ in range(30):
        assert g.step(100) == 0
        s = g.sums()
        mf = s[0] + s[1]
        p, _ = g.particles()
        urel = p[3, 0] - s[2] / mf
        fdrag = fext * mf / (mf + mp)
        mu = (1.0 - 0.5) / 3.0
        ks.append(fdrag / (6.0 * math.pi * mu * R * urel))
    c = 4.0 * math.pi * R ** 3 / (3.0 * n ** 3)
    k_has = 1.0 / (1.0 - 1.7601 * c ** (1.0 / 3.0) + c - 1.5593 * c * c)
    k = float(np.mean(ks))
    print(f"drag factor measured {k:.4f} (spread {np.ptp(ks):.4f}), Hasimoto {k_has:.4f}, ratio {k / k_has:.4f}",
          flush=True)
    p, _ = g.particles()
    g.close()
    assert p[0, 0] > 24.3 + 1.0                 # it crossed nodes on the way
    # measured 1.014 (fused) and 1.016 (two-kernel): a hydrodynamic radius
    # about 1.5 % above R
    assert 0.95 < k / k_has < 1.05
