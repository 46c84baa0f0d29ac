"""Synthetic inputs of the benchmark and parity workloads.

Every random field comes from a counter-based generator (SplitMix64 of
seed XOR index), so the same (seed, index) gives the same number on any host
and in any order.  The reference's initType 6 uses the compiler's
RANDOM_NUMBER with its default seed (main.CUF:2629-2641), which is not
reproducible; the form of the field (rhoR = mean + amp*U, rhoB = 1 - rhoR,
main.CUF:2653-2660) is kept.
"""
import math

import numpy as np

from .abi import default_config, INIT_RHO, INIT_VELFIELD, WET_NEUTRAL  # noqa: F401

_M64 = (1 << 64) - 1


def splitmix64(x):
    """SplitMix64 finaliser on a uint64 numpy array (vectorised)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, idx):
    """U[0,1) doubles for the counters idx under seed."""
    z = splitmix64(np.uint64(seed & _M64) ^ np.asarray(idx, dtype=np.uint64))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def spinodal_rho(shape, seed=1234, mean=0.5, amp=0.05):
    """rhoR = mean + amp*U(-1,1) per node, rhoB = 1 - rhoR; (nx,ny,nz) float32
    Fortran-ordered (node counter = x + nx*(y + ny*z))."""
    nx, ny, nz = shape
    idx = np.arange(nx * ny * nz, dtype=np.uint64)
    u = uniform01(seed, idx) * 2.0 - 1.0
    rr = (mean + amp * u).astype(np.float32)
    rr = rr.reshape((nx, ny, nz), order="F")
    rb = (np.float32(1.0) - rr).astype(np.float32)
    return np.asfortranarray(rr), np.asfortranarray(rb)


def taylor_green(n, u0=0.01, rho0=1.0):
    """2-D Taylor-Green vortex in an n^3 box, wavelength = box:
    u = u0 (sin kx cos ky, -cos kx sin ky, 0), k = 2 pi / n, at node
    coordinates x = 0..n-1.  Returns rhoR, rhoB (=0), vel (3,n,n,n)."""
    k = 2.0 * math.pi / n
    x = np.arange(n, dtype=np.float64)
    X, Y, _ = np.meshgrid(x, x, x, indexing="ij")
    vel = np.zeros((3, n, n, n), dtype=np.float32, order="F")
    vel[0] = u0 * np.sin(k * X) * np.cos(k * Y)
    vel[1] = -u0 * np.cos(k * X) * np.sin(k * Y)
    rr = np.full((n, n, n), rho0, dtype=np.float32, order="F")
    rb = np.zeros((n, n, n), dtype=np.float32, order="F")
    return rr, rb, np.asfortranarray(vel)


def random_spheres(shape, count, radius, seed=4321, min_dist=None, max_tries=2000000):
    """Seeded uniform positions in [1, n] (lattice units, 1-based) with centre
    distance >= 2R+1 under the minimum image (SetupParts, main.CUF:1213-1221).
    Returns pos (13, count): x, v = 0, omega = 0, q = (1,0,0,0)."""
    nx, ny, nz = shape
    L = np.array([nx, ny, nz], dtype=np.float64)
    dmin = 2.0 * radius + 1.0 if min_dist is None else float(min_dist)
    cell = dmin
    ncell = np.maximum((L // cell).astype(int), 1)
    grid = {}
    pts = []
    counter = 0
    while len(pts) < count:
        if counter > max_tries:
            raise RuntimeError("random_spheres: could not place all spheres")
        u = uniform01(seed, np.arange(counter * 3, counter * 3 + 3, dtype=np.uint64))
        counter += 1
        p = 1.0 + u * (L - 1.0)
        c = tuple((np.floor((p - 1.0) / (L / ncell))).astype(int) % ncell)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    key = ((c[0] + dx) % ncell[0], (c[1] + dy) % ncell[1], (c[2] + dz) % ncell[2])
                    for q in grid.get(key, ()):
                        d = p - q
                        d -= L * np.round(d / L)
                        # SetupParts measures the plain distance; the minimum
                        # image also keeps periodic neighbours apart
                        if np.sqrt((d * d).sum()) < dmin or np.sqrt(((p - q) ** 2).sum()) < dmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            pts.append(p)
            grid.setdefault(c, []).append(p)
    pos = np.zeros((13, count), dtype=np.float64, order="F")
    pos[0:3] = np.array(pts).T
    pos[9] = 1.0
    return pos


def eul2q(phi, psi, theta):
    """eul2q of main.CUF:1352-1370 (called as eul2q(e1, e3, e2), :1284)."""
    cy, sy = math.cos(phi * 0.5), math.sin(phi * 0.5)
    cp, sp = math.cos(psi * 0.5), math.sin(psi * 0.5)
    cr, sr = math.cos(theta * 0.5), math.sin(theta * 0.5)
    return (cy * cp * cr + sy * sp * sr, cy * cp * sr - sy * sp * cr,
            sy * cp * sr + cy * sp * cr, sy * cp * cr - cy * sp * sr)


def pickering_config(n=256, nparticles=1280, storage=0):
    """Two-component CG + colloids (the headline workload): sigma 0.02,
    beta 0.7, tau 1, R = 5.5, neutral wetting (90 degree contact angle),
    Hertz + lubrication, FP64 particle state."""
    return default_config(n, sigma=0.02, beta=0.7, storage=storage, with_particles=1,
                          n_particles=nparticles, radius=5.5, lrotate=1, max_vel=0.3,
                          wetting=WET_NEUTRAL, hertz=1, lubrication=1, lub_k=1.0,
                          lub_rmin=1.0, lub_rcap=0.1)


def pickering_inputs(n=256, nparticles=1280, seed=2024):
    """50/50 random colour field (0.5 +- 0.05 U) and seeded sphere positions."""
    rr, rb = spinodal_rho((n, n, n), seed=seed)
    pos = random_spheres((n, n, n), nparticles, 5.5, seed=seed + 1)
    return rr, rb, pos


def ensemble_workload(nbox=64, seed=77, storage=0, edge_range=(48, 96), spheres_range=(2, 16),
                      radius_range=(4.0, 8.0)):
    """Batched ensemble of independent CG boxes: box edges uniform in
    edge_range (each axis drawn separately, seeded), 2..16 spheres per box at
    seeded non-overlapping positions, one sphere radius per box drawn from
    radius_range (the sphere template is a per-box table).  Each box carries
    its own 50/50 random colour field.  Returns (configs, inputs) with
    inputs[i] = (rho_r, rho_b, pos)."""
    cfgs, inputs = [], []
    for b in range(nbox):
        u = uniform01(seed * 1000003 + b, np.arange(6, dtype=np.uint64))
        lo, hi = edge_range
        nx, ny, nz = (int(lo + math.floor(u[k] * (hi - lo + 1))) for k in range(3))
        ns = int(spheres_range[0] + math.floor(u[3] * (spheres_range[1] - spheres_range[0] + 1)))
        radius = float(radius_range[0] + u[4] * (radius_range[1] - radius_range[0]))
        cfg = default_config(nx, sigma=0.02, beta=0.7, storage=storage, with_particles=1, n_particles=ns,
                             radius=radius, lrotate=1, max_vel=0.3, wetting=WET_NEUTRAL, hertz=1,
                             lubrication=1, lub_k=1.0, lub_rmin=1.0, lub_rcap=0.1)
        cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
        rr, rb = spinodal_rho((nx, ny, nz), seed=seed * 7919 + b)
        pos = random_spheres((nx, ny, nz), ns, radius, seed=seed * 104729 + b)
        cfgs.append(cfg)
        inputs.append((rr, rb, pos))
    return cfgs, inputs
