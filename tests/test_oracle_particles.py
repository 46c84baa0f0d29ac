"""Pins the particle half of the CPU oracle (oracle/lbm_oracle.cpp) with facts
that do not come from its own formulas: an exact momentum balance between the
fluid and a free particle across node crossings, the Hertz law through its
potential, the Reynolds lubrication law with the lattice viscosity, the
moment of inertia of a solid sphere and the axis and sense of its rotation,
and the geometry of the two wetting modes.

All runs are FP64 oracle runs on small boxes (a few seconds each)."""
import math

import numpy as np

from lbcuda_amd.abi import default_config, INIT_RHO, WET_NEUTRAL, WET_JANUS
from oracle.pyoracle import OracleBox

EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1])
E = np.stack([EX, EY, EZ], 1).astype(np.float64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def sphere(x, v=(0.0, 0.0, 0.0), q=(1.0, 0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0)):
    p = np.zeros(13)
    p[0:3], p[3:6], p[6:9], p[9:13] = x, v, omega, q
    return p


def uniform_box(cfg, pos, rho_r=1.0, rho_b=0.0):
    n = cfg.nx
    o = OracleBox(cfg, 64)
    o.set_particles(np.asfortranarray(np.stack(pos, 1)))
    o.init_fields(INIT_RHO, rho_r=np.full((n, n, n), rho_r, np.float32),
                  rho_b=np.full((n, n, n), rho_b, np.float32))
    return o


def test_momentum_balance_at_node_crossings():
    """A free sphere (m = 2000) moving through fluid at rest, crossing nodes.
    Per step the fluid momentum changes by exactly
        -F_bb - T_fe - M_skipped,
    F_bb the link momentum the particle received (partBB_CG), T_fe its
    fill/erase transfer (make_fluid/del_fluid, eqs. 18, 25-26 of PRE 83,
    046707), M_skipped the momentum of the nodes erased this step that lie
    farther than rmax from the particle's previous centre: del_fluid measures
    the distance from the old centre and returns beyond sphereMax = rmax^2
    (kernels_fluid_PART.CUF:400-406, main.CUF:2109), so those nodes leave the
    fluid without a transfer.  The balance closes to round-off at every step,
    which pins the sign and size of the link exchange and of both transfers
    (a dropped or sign-flipped transfer leaves a residual of order 1 at each
    crossing).  Fluid momentum is summed over the nodes that are fluid in the
    step that produced the populations; the flags the oracle reports after a
    step are those of the step before (its two time levels alternate)."""
    n, m, rmax = 24, 2000.0, 6
    cfg = default_config(n, with_particles=1, n_particles=1, lrotate=0, rmass=1.0 / m, max_vel=1.0,
                         wetting=WET_NEUTRAL)
    o = uniform_box(cfg, [sphere((12.3, 11.8, 12.1), v=(0.05, -0.03, 0.02))])
    snap = [(o.flags(), o.pops(0) + o.pops(1), o.particles())]
    for _ in range(21):
        assert o.step(1) == 0
        snap.append((o.flags(), o.pops(0) + o.pops(1), o.particles()))
    g = np.arange(n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    crossings = 0
    for s in range(2, 20):
        fl_before, fl_after = snap[s][0], snap[s + 1][0]
        p_before, p_after = snap[s - 1][1], snap[s][1]
        _, f_before = snap[s - 1][2]
        _, f_after = snap[s][2]
        dpf = (p_after[fl_after == 1] @ E).sum(0) - (p_before[fl_before == 1] @ E).sum(0)
        fbb = f_after[3:6, 0]
        t_fe = f_after[0:3, 0] - 0.5 * (f_after[3:6, 0] + f_before[3:6, 0])
        erased = (fl_before == 1) & (fl_after != 1)
        c_old = np.rint(snap[s - 2][2][0][0:3, 0]) - 1
        d = [(A - c + n // 2) % n - n // 2 for A, c in zip((X, Y, Z), c_old)]
        skipped = erased & ((d[0] ** 2 + d[1] ** 2 + d[2] ** 2) > rmax * rmax)
        m_skipped = (p_before[skipped] @ E).sum(0)
        resid = dpf + fbb + t_fe + m_skipped
        assert np.abs(resid).max() < 1e-10, (s, resid)
        if erased.any():
            crossings += 1
            assert np.abs(t_fe).max() > 0.1          # the transfers are not trivially zero
    assert crossings >= 2


def pair_forces_after_two_steps(cfg, pos):
    """Pair force on each particle in the second step: F minus the link part
    (F_bb(t) + F_bb(t-1)) / 2 (addForce_particle_bounce_back); valid when no
    node changed type (no fill/erase transfer)."""
    o = uniform_box(cfg, pos)
    assert o.step(1) == 0
    _, f1 = o.particles()
    assert o.step(1) == 0
    x2, f2 = o.particles()
    return f2[0:3] - 0.5 * (f2[3:6] + f1[3:6]), x2


def test_hertz_force_is_minus_gradient_of_its_potential():
    """Soft-core repulsion (addHertzian, kernels_fluid_PART.CUF:1385-1503):
    the Hertz law F = -dU/dr of U = k (r_min - r)^(5/2) between r_cap and
    r_min, constant inside r_cap, zero beyond r_min; equal and opposite on
    the two spheres and along the line of centres.  Particles are held fixed
    (1/m = 0) in fluid at rest, so the pair force is what remains of F after
    the link part.  Radius 4: the input check wants centres at least 2R + 1
    apart (main.CUF:1216), which leaves the Hertz range r < 10.5 open."""
    k, rmin, rcap = 10.0, 10.5, 9.4
    U = lambda r: k * (rmin - r) ** 2.5
    for r, expect in ((10.2, None), (9.6, None), (9.2, "cap"), (10.8, "zero")):
        cfg = default_config(32, radius=4.0, with_particles=1, n_particles=2, lrotate=0, rmass=0.0,
                             max_vel=1.0, wetting=WET_NEUTRAL, hertz=1, hz_k=k, hz_rmin=rmin, hz_rcap=rcap,
                             lubrication=0)
        fp, _ = pair_forces_after_two_steps(cfg, [sphere((16.0 - r / 2, 16.2, 15.9)),
                                                  sphere((16.0 + r / 2, 16.2, 15.9))])
        assert np.abs(fp[:, 0] + fp[:, 1]).max() < 1e-12          # equal and opposite
        assert np.abs(fp[1:3, :]).max() < 1e-12                     # along the line of centres
        h = 1e-6
        if expect is None:
            grad = (U(r + h) - U(r - h)) / (2 * h)                   # dU/dr < 0: repulsive
            assert abs(fp[0, 1] - (-grad)) < 1e-6 * abs(grad)
        elif expect == "cap":
            grad = (U(rcap + h) - U(rcap - h)) / (2 * h)
            assert abs(fp[0, 1] - (-grad)) < 1e-6 * abs(grad)
        else:
            assert abs(fp[0, 1]) < 1e-12


def test_lubrication_follows_reynolds_law_with_lattice_viscosity():
    """Near-contact lubrication (kernels_fluid_PART.CUF:1463-1501): the
    Reynolds squeeze-film force between two spheres, F = 6 pi mu a^2 v_n / h
    with the reduced radius a = R1 R2 / (R1 + R2) = R/2 and the lattice
    dynamic viscosity mu = rho c_s^2 (tau - 1/2), in the reference's form
    shifted to vanish at a gap of one lattice unit, F = k_l 6 pi mu a^2 v_n
    (1/h - 1), with h clamped at h_cap; it pushes approaching spheres apart.
    The input check keeps gaps at 1 or more (main.CUF:1216), so the spheres
    start one lattice unit apart and close in at prescribed speeds (1/m = 0,
    no node crossed): the force of the second step acts at h = 1 - v_rel."""
    R, tau, kl, hc, hcap = 5.5, 0.8, 1.0, 1.0, 0.1
    mu = 1.0 * (1.0 / 3.0) * (tau - 0.5)
    a = R * R / (R + R)
    for vrel in (0.1, 0.4, 0.8, 0.95):
        cfg = default_config(32, with_particles=1, n_particles=2, lrotate=0, rmass=0.0, max_vel=1.0,
                             wetting=WET_NEUTRAL, hertz=0, lubrication=1, lub_k=kl, lub_rmin=hc,
                             lub_rcap=hcap, tau_r=tau, tau_b=tau)
        r0 = 2 * R + 1.0
        fp, _ = pair_forces_after_two_steps(cfg, [sphere((16.0 - r0 / 2, 16.1, 15.8), v=(vrel / 2, 0, 0)),
                                                  sphere((16.0 + r0 / 2, 16.1, 15.8), v=(-vrel / 2, 0, 0))])
        h = max(r0 - vrel - 2 * R, hcap)
        expect = -kl * 6 * math.pi * mu * a * a * vrel * (1.0 / h - 1.0 / hc)
        assert abs(fp[0, 0] - expect) < 1e-6 * abs(expect), (vrel, fp[0, 0], expect)   # config is FP32
        assert abs(fp[0, 0] + fp[0, 1]) < 1e-12
        assert np.abs(fp[1:3, :]).max() < 1e-12
        assert fp[0, 0] < 0                                      # pushes the approaching spheres apart


def quat_mul(a, b):
    return np.array([a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
                     a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
                     a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
                     a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]])


def axis_quat(axis, angle):
    u = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.concatenate([[math.cos(angle / 2)], math.sin(angle / 2) * u])


def test_rigid_body_rotation_under_torque():
    """time_step_force_particles / rot_midstep_impl_lf (PART:1098-1166,
    1621-1827) for a sphere under a constant external torque in fluid at rest:
    the body angular velocity grows by T_body / I per step with the moment of
    inertia of a solid sphere, I = 2/5 m R^2; the torque reaches the body frame
    through the transpose of the orientation's rotation matrix; the
    orientation turns about the lab axis of the torque, in its sense, by the
    accumulated angle sum(omega) up to O(omega^2) per step.  The sphere starts
    turned by 90 degrees about z, so a torque about lab x is about body -y:
    a transposed matrix or a wrong quaternion product turns it about the
    wrong axis."""
    n, m, R, T = 24, 1500.0, 5.5, 0.5
    I = 2.0 / 5.0 * m * R * R
    q0 = axis_quat((0, 0, 1), math.pi / 2)
    cfg = default_config(n, with_particles=1, n_particles=1, lrotate=1, rmass=1.0 / m, max_vel=1.0,
                         wetting=WET_NEUTRAL, ext_torque=(T, 0.0, 0.0))
    o = uniform_box(cfg, [sphere((12.0, 12.0, 12.0), q=q0)])
    x, _ = o.particles()
    angle = 0.0
    for step in range(40):
        om_before = x[6:9, 0].copy()
        assert o.step(1) == 0
        x, f = o.particles()
        # lab torque T_lab = F[6:9]; body torque = R(q_n)^T T_lab, R(q_n) a 90 degree turn about z
        tb = np.array([f[7, 0], -f[6, 0], f[8, 0]])
        assert np.allclose(x[6:9, 0] - om_before, tb / I, rtol=1e-6, atol=1e-15)   # 1/m is FP32 in the config
        angle += -x[7, 0]                         # omega^{n+1/2} about body -y, i.e. lab x
    q = x[9:13, 0]
    assert abs(np.linalg.norm(q) - 1) < 1e-12
    qrel = quat_mul(q, np.array([q0[0], -q0[1], -q0[2], -q0[3]]))   # rotation since the start, lab frame
    turn = 2 * math.atan2(np.linalg.norm(qrel[1:4]), qrel[0])
    axis = qrel[1:4] / np.linalg.norm(qrel[1:4])
    assert axis[0] > 1 - 1e-6                     # about +x, the torque's axis and sense
    assert angle > 0 and abs(turn - angle) < 0.05 * angle


def test_neutral_wetting_of_a_uniform_mixture_is_neutral():
    """compute_densities_wall (PART:267-297, 864-894): a surface node takes
    the weight-averaged densities of its fluid D3Q27 neighbours, so in a
    uniform mixture at rest it takes the mixture's densities, the phase field
    has no gradient at the particle, and the whole state stays at rest: the
    populations remain f = w rho and the particle feels no force."""
    n = 24
    cfg = default_config(n, sigma=0.05, beta=0.7, with_particles=1, n_particles=1, lrotate=1, rmass=0.0,
                         max_vel=1.0, wetting=WET_NEUTRAL)
    o = uniform_box(cfg, [sphere((12.0, 12.0, 12.0))], rho_r=0.6, rho_b=0.4)
    assert o.step(5) == 0
    fl = o.flags()
    rr, rb = o.rho()
    surf = fl == 2
    assert np.abs(rr[surf] - 0.6).max() < 1e-6 and np.abs(rb[surf] - 0.4).max() < 1e-6
    fluid = fl == 1
    for c, rho in ((0, 0.6), (1, 0.4)):
        p = o.pops(c)[fluid]
        assert np.abs(p - np.float32(rho) * W).max() < 1e-7
    _, f = o.particles()
    assert np.abs(f[0:12, 0]).max() < 1e-7


def test_janus_wetting_follows_the_body_axis():
    """compute_virtual_rho (PART:299-352): a surface node is red (1, 0) on the
    half of the sphere the body x axis points into and blue (0, 1) on the
    other half.  With the identity orientation the split is by the sign of
    x - x_c; turned 90 degrees about z the body x axis is lab +y and the split
    is by y - y_c."""
    n = 24
    c = np.array([12.0, 12.0, 12.0])
    for q, axis in (((1.0, 0.0, 0.0, 0.0), 0), (tuple(axis_quat((0, 0, 1), math.pi / 2)), 1)):
        cfg = default_config(n, sigma=0.05, beta=0.7, with_particles=1, n_particles=1, lrotate=1, rmass=0.0,
                             max_vel=1.0, wetting=WET_JANUS)
        o = uniform_box(cfg, [sphere(tuple(c), q=q)], rho_r=0.5, rho_b=0.5)
        assert o.step(1) == 0
        fl = o.flags()
        rr, rb = o.rho()
        idx = np.argwhere(fl == 2)
        side = (idx[:, axis] + 1) - c[axis]
        red = rr[fl == 2] > 0.5
        assert (rr[fl == 2] + rb[fl == 2] == 1).all()
        assert red[side > 0].all() and not red[side < 0].any()
