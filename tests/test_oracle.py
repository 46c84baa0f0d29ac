"""Pins the CPU oracle (oracle/lbm_oracle.cpp) with facts that do not come
from its own formulas: conservation laws, closed-form flows, an independent
numpy BGK, lattice-point counts and the reference's own invariants.

The reference ships no golden vectors or expected outputs (its Tests/ hold
input decks only), so these are the anchors of every parity claim."""
import math

import numpy as np
import pytest

from lbcuda_amd.abi import default_config, INIT_RHO, INIT_VELFIELD, INIT_DROPLET, WET_NEUTRAL
from lbcuda_amd import workloads as wl
from oracle.pyoracle import OracleBox

EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def test_lattice_is_isotropic():
    """D3Q19 weights and velocities of dimensions_m.CUF:201-219 satisfy the
    moment identities the LB method needs (sum w = 1, sum w e = 0,
    sum w e e = I/3, opposite pairs at (2l-1, 2l))."""
    E = np.stack([EX, EY, EZ], 1)
    assert abs(W.sum() - 1) < 1e-15
    assert np.allclose((W[:, None] * E).sum(0), 0)
    assert np.allclose(np.einsum("l,li,lj->ij", W, E, E), np.eye(3) / 3)
    for l in range(1, 19, 2):
        assert (E[l] == -E[l + 1]).all()
    # Leclaire B_l with xi = 0.5 (dimensions_m.CUF:222-228): sum B_l = sum w (e.n)^2 = 1/3
    B = np.array([-2 / 9] + [1 / 54] * 6 + [1 / 27] * 12)
    assert abs(B.sum() - 1 / 3) < 1e-15


def ball_count(r2):
    g = np.arange(-12, 13)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return int(((X * X + Y * Y + Z * Z) <= r2).sum())


@pytest.mark.parametrize("radius", [4.0, 5.5, 8.0])
def test_sphere_template(radius):
    """spherical_template (main.CUF:1990-2130): the non-fluid set is exactly
    the lattice ball |r|^2 <= R^2, and a node is a surface node iff one of its
    18 D3Q19 neighbours lies outside that ball."""
    cfg = default_config(32, radius=radius, with_particles=1)
    o = OracleBox(cfg, 64)
    rmax, t = o.template()
    assert rmax == int(radius + 1)
    assert int((t != 1).sum()) == ball_count(radius * radius)
    g = np.arange(-rmax, rmax + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    inside = (X * X + Y * Y + Z * Z) <= radius * radius
    surf = np.zeros_like(inside)
    for l in range(1, 19):
        sh = np.roll(np.roll(np.roll(inside, -EX[l], 0), -EY[l], 1), -EZ[l], 2)
        surf |= inside & ~sh
    expect = np.where(inside, np.where(surf, 2, 4), 1)
    expect[rmax, rmax, rmax] = 5
    assert (t == expect).all()
    if radius == 5.5:
        assert ((t == 2).sum(), (t == 4).sum()) == (410, 328)


def run_oracle(cfg, init, steps, bits=64, **kw):
    o = OracleBox(cfg, bits)
    o.init_fields(init, **kw)
    rc = o.step(steps)
    return o, rc


def test_mass_conserved_per_colour():
    """Perturbation and recolouring move mass between directions only
    (time_step_CG, kernels_fluid_CG.CUF:803-837): per-colour mass is constant."""
    n = 16
    rr, rb = wl.spinodal_rho((n, n, n), seed=7)
    cfg = default_config(n, sigma=0.02, beta=0.7)
    o = OracleBox(cfg, 64)
    o.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    m0 = (o.pops(0).sum(), o.pops(1).sum())
    assert o.step(60) == 0
    m1 = (o.pops(0).sum(), o.pops(1).sum())
    assert abs(m1[0] - m0[0]) < 1e-10 * m0[0]
    assert abs(m1[1] - m0[1]) < 1e-10 * m0[1]


def momentum(o):
    p = o.pops(0) + o.pops(1)
    return np.array([(p * EX).sum(), (p * EY).sum(), (p * EZ).sum()])


def test_momentum_conserved():
    n = 16
    rr, rb = wl.spinodal_rho((n, n, n), seed=9)
    cfg = default_config(n, sigma=0.05, beta=0.7)
    o = OracleBox(cfg, 64)
    o.init_fields(INIT_RHO, u0=(0.01, -0.02, 0.005), rho_r=rr, rho_b=rb)
    p0 = momentum(o)
    o.step(40)
    assert np.allclose(momentum(o), p0, atol=1e-11, rtol=0)


def test_rest_state_is_fixed_point():
    """Uniform colours at rest: grad phi = 0, f = f_eq, the step is identity."""
    n = 8
    rr = np.full((n, n, n), 0.7, np.float32)
    rb = np.full((n, n, n), 0.3, np.float32)
    o, rc = run_oracle(default_config(n), INIT_RHO, 10, rho_r=rr, rho_b=rb)
    pr = o.pops(0)
    for l in range(19):
        assert np.allclose(pr[..., l], np.float32(0.7) * W[l], atol=1e-12)


def numpy_bgk(f, tau, steps):
    """Independent single-component BGK (pull form, numpy roll)."""
    for _ in range(steps):
        rho = f.sum(-1)
        u = np.stack([(f * EX).sum(-1), (f * EY).sum(-1), (f * EZ).sum(-1)], -1) / rho[..., None]
        eu = 3.0 * (u[..., 0:1] * EX + u[..., 1:2] * EY + u[..., 2:3] * EZ)
        usq = 1.5 * (u * u).sum(-1)[..., None]
        feq = rho[..., None] * W * (1 + eu + 0.5 * eu * eu - usq)
        f = f + (feq - f) / tau
        for l in range(19):
            f[..., l] = np.roll(f[..., l], (EX[l], EY[l], EZ[l]), (0, 1, 2))
    return f


def test_single_colour_equals_independent_bgk():
    """With rhoB = 0 the CG step reduces to BGK with omega = 1/tauR
    (kernels_fluid_CG.CUF:694-702, 805): compare with a numpy BGK."""
    n = 8
    rng = np.random.default_rng(3)
    rr = (1.0 + 0.05 * rng.uniform(-1, 1, (n, n, n))).astype(np.float32)
    rb = np.zeros_like(rr)
    vel = (0.02 * rng.uniform(-1, 1, (3, n, n, n))).astype(np.float32)
    tau = float(np.float32(0.8))   # lbm_config stores tau as float
    cfg = default_config(n, tau_r=tau, tau_b=tau)
    o = OracleBox(cfg, 64)
    o.init_fields(INIT_VELFIELD, rho_r=rr, rho_b=rb, vel=np.asfortranarray(vel))
    f0 = o.pops(0).copy()
    o.step(6)
    ref = numpy_bgk(f0.copy(), tau, 6)
    assert np.abs(o.pops(0) - ref).max() < 1e-13
    assert np.abs(o.pops(1)).max() == 0.0


def test_streaming_moves_by_e_and_wraps():
    """Push streaming (stream*_CG, :929-1015) plus the halo pushes
    (stream_CG_x/_y/_z, :1017-1141): an excess in direction l moves by e_l
    per step and re-enters through the opposite face."""
    n = 6
    rr = np.ones((n, n, n), np.float32)
    rb = np.zeros_like(rr)
    cfg = default_config(n, tau_r=1e6, tau_b=1e6)   # omega ~ 1e-6: collision ~ identity
    for l in (1, 4, 7, 12, 17):
        o = OracleBox(cfg, 64)
        o.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
        f = o.pops(0)
        f[4, 1, 5, l] += 0.01
        o.set_pops(f, o.pops(1))
        o.step(3)
        d = o.pops(0)[..., l] - W[l]
        i, j, k = np.unravel_index(np.argmax(d), d.shape)
        assert (i, j, k) == ((4 + 3 * EX[l]) % n, (1 + 3 * EY[l]) % n, (5 + 3 * EZ[l]) % n)
        assert abs(d.max() - 0.01) < 1e-6


def test_taylor_green_decay_rate():
    """2-D Taylor-Green vortex is an exact Navier-Stokes solution: kinetic
    energy decays as exp(-4 nu k^2 t), nu = (tau - 1/2)/3."""
    n = 48
    rr, rb, vel = wl.taylor_green(n, u0=0.01)
    o = OracleBox(default_config(n), 64)
    o.init_fields(INIT_VELFIELD, rho_r=rr, rho_b=rb, vel=vel)

    def ke():
        f = o.pops(0)
        rho = f.sum(-1)
        u = (f * EX).sum(-1) / rho
        v = (f * EY).sum(-1) / rho
        return 0.5 * (rho * (u * u + v * v)).sum()

    e0 = ke()
    o.step(60)
    e1 = ke()
    nu = (1.0 - 0.5) / 3.0
    k = 2 * math.pi / n
    rate = -math.log(e1 / e0) / 60
    assert abs(rate / (4 * nu * k * k) - 1) < 0.02


def test_recolouring_keeps_interface_sharp():
    """beta > 0 pushes colour up the phase gradient (recolouring,
    kernels_fluid_CG.CUF:817-837): a smooth planar interface gets sharper,
    with beta = 0 it only diffuses."""
    n = 24

    def width(beta):
        z = np.arange(n)
        prof = np.tanh((np.minimum(z - n / 4, 3 * n / 4 - z)) / 4.0)
        rr = np.broadcast_to(0.5 * (1 + prof), (n, n, n)).astype(np.float32).copy()
        rb = (1 - rr).astype(np.float32)
        o = OracleBox(default_config(n, sigma=0.01, beta=beta), 64)
        o.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
        o.step(100)
        a, b = o.pops(0).sum(-1), o.pops(1).sum(-1)
        phi = (a - b) / (a + b)
        return int((np.abs(phi[0, 0, :]) < 0.9).sum())

    assert width(0.7) < width(0.0)


def test_droplet_laplace_law():
    """A droplet relaxes to Laplace pressure dp = 2 sigma / R (pressure
    cs^2 rho), the surface tension the perturbation step (A = 9/4 omega
    sigma, kernels_fluid_CG.CUF:803-815) is built to give."""
    n = 40
    sigma = 0.05
    cfg = default_config(n, sigma=sigma, beta=0.7)
    o = OracleBox(cfg, 64)
    o.init_fields(INIT_DROPLET)
    o.step(600)
    a, b = o.pops(0).sum(-1), o.pops(1).sum(-1)
    rho = a + b
    phi = (a - b) / rho
    inside = rho[phi > 0.9].mean()
    outside = rho[phi < -0.9].mean()
    vol = (phi > 0).sum()
    R = (3 * vol / (4 * math.pi)) ** (1 / 3)
    dp = (inside - outside) / 3.0
    assert abs(dp * R / (2 * sigma) - 1) < 0.15


def particle_case(n=32, npart=3, seed=11, lrotate=1, rmass=1.0 / 1500.0, vel=0.05):
    cfg = default_config(n, sigma=0.02, beta=0.7, with_particles=1, n_particles=npart, lrotate=lrotate,
                         rmass=rmass, max_vel=1.0, wetting=WET_NEUTRAL)
    pos = wl.random_spheres((n, n, n), npart, 5.5, seed=seed)
    rng = np.random.default_rng(seed)
    pos[3:6] = vel * rng.uniform(-1, 1, (3, npart))
    if lrotate:
        pos[6:9] = 0.01 * rng.uniform(-1, 1, (3, npart))
    rr, rb = wl.spinodal_rho((n, n, n), seed=seed)
    return cfg, pos, rr, rb


def test_particle_volume_invariant():
    """Every particle owns totSphere = 739 non-fluid nodes and one centre
    node (write_output.f90:1093-1124), also after the particles move."""
    cfg, pos, rr, rb = particle_case(npart=3)
    o = OracleBox(cfg, 64)
    o.set_particles(pos)
    o.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    fl = o.flags()
    assert (fl != 1).sum() == 739 * 3 and (fl == 5).sum() == 3
    assert o.step(25) == 0
    fl = o.flags()
    assert (fl != 1).sum() == 739 * 3 and (fl == 5).sum() == 3
    p, _ = o.particles()
    assert np.abs(p[0:3] - pos[0:3]).max() > 0.5   # they did move across nodes


def test_sphere_at_rest_feels_no_force():
    """A sphere centred on a node in fluid at rest: by symmetry the
    bounce-back force and torque vanish (partBB_CG, PART:916-973)."""
    n = 24
    cfg = default_config(n, with_particles=1, n_particles=1, lrotate=1, max_vel=1.0, wetting=WET_NEUTRAL)
    pos = np.zeros((13, 1), order="F")
    pos[0:3, 0] = 12.0
    pos[9, 0] = 1.0
    rr = np.ones((n, n, n), np.float32)
    rb = np.zeros_like(rr)
    o = OracleBox(cfg, 64)
    o.set_particles(pos)
    o.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    assert o.step(5) == 0
    _, f = o.particles()
    assert np.abs(f[0:12]).max() < 1e-10


def test_big_force_guard_raises():
    """nve_lf stops when |F|/m > 0.5 (kernels_fluid_PART.CUF:1532-1537); an
    external force of 1000 with m = 1500 trips it on the first step."""
    n = 24
    cfg = default_config(n, with_particles=1, n_particles=1, max_vel=10.0, wetting=WET_NEUTRAL,
                         ext_force=(1000.0, 0.0, 0.0))
    pos = np.zeros((13, 1), order="F")
    pos[0:3, 0] = 12.0
    pos[9, 0] = 1.0
    rr = np.ones((n, n, n), np.float32)
    o = OracleBox(cfg, 64)
    o.set_particles(pos)
    o.init_fields(INIT_RHO, rho_r=rr, rho_b=np.zeros_like(rr))
    assert o.step(1) == 5
    assert o.status()[0] == 5


def test_fp32_and_fp64_instantiations_agree():
    cfg, pos, rr, rb = particle_case(npart=2)
    outs = []
    for bits in (32, 64):
        o = OracleBox(cfg, bits)
        o.set_particles(pos)
        o.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
        assert o.step(10) == 0
        outs.append((o.pops(0), o.particles()[0]))
    assert np.abs(outs[0][0] - outs[1][0]).max() < 1e-5
    assert np.abs(outs[0][1] - outs[1][1]).max() < 1e-4
