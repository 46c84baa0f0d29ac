"""Pins the density-ratio colour-gradient model of the CPU oracle (the
reference's DIFFDENS build: equilCG, kernels_fluid_CG.CUF:40-49; alphaCG,
main.CUF:1077-1083; the psi/xi correction of the BGK step, :755-792) with
facts that do not come from its own formulas:

* the equilibrium's moments have the closed forms of a colour-dependent
  sound speed: sum f = rho, sum e f = rho u, sum e_i e_j f =
  rho ((1 - alpha_k)/2 delta_ij + u_i u_j);
* with densR = 1 the rest weights balance the pressures of the two pure
  phases at their nominal densities: densR (1 - alpha_R) = densB (1 - alpha_B);
* the correction terms neither create mass nor momentum: a periodic box keeps
  each colour's mass and the total momentum to rounding while the terms act
  (a missing psi or xi term changes the mass by O(1e-4) per step here);
* densR = densB = 1 gives alpha_R = alpha_B = 1/3, the plain D3Q19 weights.

FP64 oracle runs on small boxes (seconds)."""
import numpy as np
import pytest

from lbcuda_amd.abi import default_config, INIT_DROPLET
from oracle.pyoracle import OracleBox

E = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0],
              [0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1],
              [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1]], dtype=np.float64).T

U0 = (0.01, -0.02, 0.005)


def droplet(dens_b, n=24, u0=U0, **kw):
    cfg = default_config(n, sigma=0.05, beta=0.7, diffdens=1, dens_r=1.0, dens_b=dens_b, **kw)
    o = OracleBox(cfg, 64)
    o.init_fields(INIT_DROPLET, u0=u0)
    return o


def second_moment(f):
    return np.einsum("l,li,lj->ij", f, E, E)


@pytest.mark.parametrize("dens_b", [4.0, 1.0])
def test_equilibrium_moments_and_pressure_balance(dens_b):
    n = 24
    o = droplet(dens_b, n)
    pr, pb = o.pops(0), o.pops(1)
    c = n // 2 - 1              # inside the droplet: pure red, rho = densR
    red = pr[c, c, c]
    blue = pb[0, 0, 0]          # far corner: pure blue, rho = densB
    assert pb[c, c, c].sum() == 0.0 and pr[0, 0, 0].sum() == 0.0
    u = np.array(U0, dtype=np.float32).astype(np.float64)   # u0 crosses the ABI as float
    cs2 = []
    for f, rho in ((red, 1.0), (blue, dens_b)):
        assert abs(f.sum() - rho) < 1e-14
        np.testing.assert_allclose(f @ E, rho * u, atol=1e-15)
        m2 = second_moment(f) - rho * np.outer(u, u)
        # isotropic: (1 - alpha)/2 rho on the diagonal, nothing off it
        assert np.abs(m2 - np.diag(np.diag(m2))).max() < 1e-15
        assert np.ptp(np.diag(m2)) < 1e-15
        cs2.append(m2[0, 0] / rho)
    # pressures of the pure phases agree: rho_R cs_R^2 = rho_B cs_B^2
    assert abs(1.0 * cs2[0] - dens_b * cs2[1]) < 1e-15
    assert abs(cs2[0] - 1.0 / 3.0) < 1e-15        # alpha_R = p0 = 1/3
    if dens_b == 1.0:
        assert abs(cs2[1] - 1.0 / 3.0) < 1e-15


def test_correction_terms_conserve_mass_and_momentum():
    """Random colour fields (seeded, no symmetry), unequal viscosities: each
    colour's mass and the total momentum stay put to rounding.  (Densities
    stay positive here, so the MINDENS clamp of init_rho*_CG, which does
    remove mass where a colour goes negative, never acts.)"""
    from lbcuda_amd.abi import INIT_RHO
    from lbcuda_amd import workloads as wl
    n = 20
    cfg = default_config(n, sigma=0.05, beta=0.7, diffdens=1, dens_r=1.0, dens_b=4.0, tau_r=0.8, tau_b=1.3)
    rr, _ = wl.spinodal_rho((n, n, n), seed=5, mean=0.5, amp=0.2)
    rb = (4.0 * (1.0 - rr)).astype(np.float32)
    o = OracleBox(cfg, 64)
    o.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)

    def totals():
        pr, pb = o.pops(0), o.pops(1)
        return pr.sum(), pb.sum(), ((pr + pb) @ E).sum(axis=(0, 1, 2)), min(pr.sum(-1).min(), pb.sum(-1).min())
    m0r, m0b, j0, _ = totals()
    for _ in range(4):
        assert o.step(5) == 0
        mr, mb, j, lo = totals()
        assert lo > 0.0
        assert abs(mr - m0r) < 1e-12 * m0r and abs(mb - m0b) < 1e-12 * m0b
        np.testing.assert_allclose(j, j0, atol=1e-12 * (m0r + m0b))
    # the correction terms were active: velocities and density gradients grew
    assert np.abs(j0).max() > 0 or np.abs(((o.pops(0) + o.pops(1)) @ E)).max() > 1e-4


def test_the_correction_is_what_separates_the_models():
    """The same droplet with diffdens off and on at equal densities differs
    only through the psi/xi terms, which vanish where the densities are
    uniform: far from the interface the two runs agree to rounding."""
    n = 24
    a = droplet(1.0, n, u0=(0.0, 0.0, 0.0))
    cfg = default_config(n, sigma=0.05, beta=0.7)
    b = OracleBox(cfg, 64)
    b.init_fields(INIT_DROPLET)
    assert np.array_equal(a.pops(0), b.pops(0))   # alpha = 1/3: equilCG = equil
    a.step(3)
    b.step(3)
    d = np.abs(a.pops(0) - b.pops(0)).max(axis=-1)
    assert d.max() > 1e-7                          # near the interface they differ
    assert d[0:3, 0:3, 0:3].max() < 1e-15          # the corner stays uniform in both
