"""Learned near-contact closure: training data from the solver, training on
the GPU, and the closure block the pair-force kernel reads.

Data.  Two equal spheres approach each other in a periodic single-colour box
(CG with rho_B = 0, no surface tension: plain BGK, tau = 1) with prescribed
velocities (1/m = 0), once at the working resolution ("coarse", radius R)
and once at 4x resolution ("fine", radius 4R, box scaled by 4, speed / 4, so
that the Reynolds number and the path in units of R are the same and the
fine run takes 16x the steps).  Runs are head-on and oblique (relative
velocity at an angle to the initial line of centres, so the pair slides
past at closest approach); the coarse gap goes from about 3 down to 0.1
lattice units after a warm-up path of 2.  Pair forces (Hertz, lubrication) are off, so the measured
force is hydrodynamic only.

Along each run the force on sphere 1 is split into a normal part F_n
(repulsive > 0), a tangential part F_t (opposing the relative sliding > 0)
and the torque T about n x t, and made dimensionless with the Stokes scales
6 pi mu R |u| (forces) and 8 pi mu R^2 |u| (torque).  Both runs are reduced
to medians in bins of the coarse-unit gap (0.1 wide, 0.35 to 3); the
training target of a bin is the fine coefficient minus the coarse one, in
coarse force units: what the coarse lattice misses.  Features (gap/R, v_n,
v_t, R) are those lbm_set_closure documents.

The network (4 -> 64 -> 64 -> 3) is trained with the library's own GPU
kernels (lbm_mlp_train); lbm_set_closure then applies it in the pair-force
kernel for pairs with gap < gap_max inside the feature box of the data, the
reference's lubrication term elsewhere (kernels_fluid_PART.CUF:1463-1501).
"""
import math
import os

import numpy as np

from .abi import default_config, INIT_RHO, WET_NEUTRAL
from . import lib as L

MU = (1.0 - 0.5) / 3.0          # rho * nu at rho = 1, tau = 1
GAP_START, GAP_END = 3.0, 0.1   # coarse lattice units
FINE = 4


WARM = 2.0                      # coarse path length run before sampling


def _path(R, angle_deg):
    """Coarse-unit path of one run: separation r12(s) from sphere 1 to
    sphere 2 after a relative displacement s, the unit direction e_u of the
    relative velocity of sphere 1 w.r.t. sphere 2, and the total path
    length (closest approach, gap GAP_END, at its end)."""
    a = math.radians(angle_deg)
    e_u = np.array([math.cos(a), math.sin(a), 0.0])
    dmin = 2 * R + GAP_END
    if angle_deg == 0.0:
        total = GAP_START + WARM - GAP_END
        r0 = np.array([dmin + total, 0.0, 0.0])
    else:
        total = math.sqrt((2 * R + GAP_START) ** 2 - dmin ** 2) + WARM
        perp = np.array([-math.sin(a), math.cos(a), 0.0]) * dmin
        r0 = perp + total * e_u
    return r0, e_u, total


SPEED = 4e-3                    # relative speed at the coarse resolution (Re = U R / nu <= 0.2)


def run_pair(R, angle_deg, scale, speed=SPEED, sample_every=None, storage=0, closure=None):
    """One pair-approach run at `scale` x resolution (radius R*scale, box and
    path scaled, relative speed speed/scale).  Returns per-sample arrays:
    path s (coarse units), gap/R, v_n/|u|, v_t/|u| and the dimensionless
    C_n, C_t, C_T of the total force (and Bn, Bt, BT of the link part
    alone).  `closure`: a closure block applied in the pair-force kernel."""
    Rs = R * scale
    us = speed / scale
    r0, e_u, total = _path(R, angle_deg)
    # sphere 1 starts at the origin and moves +s/2 e_u, sphere 2 starts at
    # r0 and moves -s/2 e_u: the box holds both paths plus R of fluid
    # beyond each sphere
    pts = np.array([np.zeros(3), 0.5 * total * e_u, r0, r0 - 0.5 * total * e_u])
    lo, hi = pts.min(0), pts.max(0)
    dims = [int(math.ceil((hi[k] - lo[k] + 4 * R) * scale)) for k in range(3)]
    dims = [d + (d % 2) for d in dims]
    cfg = default_config(dims[0], sigma=0.0, beta=0.0, storage=storage, with_particles=1, n_particles=2,
                         radius=float(Rs), lrotate=1, rmass=0.0, max_vel=1.0, wetting=WET_NEUTRAL, hertz=0,
                         lubrication=0, closure=int(closure is not None))
    cfg.nx, cfg.ny, cfg.nz = dims
    box = L.Box(cfg)
    if closure is not None:
        box.set_closure(closure)
    n3 = np.array(dims, dtype=np.float64)
    origin = n3 / 2.0 + 0.5 - 0.5 * (lo + hi) * scale
    pos = np.zeros((13, 2), order="F")
    pos[0:3, 0] = origin
    pos[0:3, 1] = origin + r0 * scale
    pos[3:6, 0], pos[3:6, 1] = 0.5 * us * e_u, -0.5 * us * e_u
    pos[9, :] = 1.0
    box.set_particles(pos)
    rr = np.ones(dims, np.float32, order="F")
    rb = np.zeros(dims, np.float32, order="F")
    box.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    nsteps = int(round(total * scale / us))
    every = sample_every or max(1, nsteps // 3000)
    out = {k: [] for k in ("s", "gap", "vn", "vt", "Cn", "Ct", "CT", "Bn", "Bt", "BT")}
    done = 0
    while done < nsteps:
        k = min(every, nsteps - done)
        rc = box.step(k)
        if rc:
            raise RuntimeError(f"pair run stopped with {box.status()}")
        done += k
        p, f = box.particles()
        d = p[0:3, 1] - p[0:3, 0]
        d -= n3 * np.round(d / n3)
        r = float(np.linalg.norm(d))
        n = d / r
        u = p[3:6, 0] - p[3:6, 1]
        vn = float(u @ n)
        tv = u - vn * n
        vt = float(np.linalg.norm(tv))
        t = tv / vt if vt > 1e-14 else np.zeros(3)
        F1, T1 = f[0:3, 0], f[6:9, 0]
        B1, BT1 = f[3:6, 0], f[9:12, 0]     # link (bounce-back) part alone
        umag = float(np.linalg.norm(u))
        out["s"].append(done * us / scale)
        out["gap"].append((r - 2 * Rs) / Rs)
        out["vn"].append(vn / umag)
        out["vt"].append(vt / umag)
        out["Cn"].append(-float(F1 @ n) / (6 * math.pi * MU * Rs * umag))
        out["Ct"].append(-float(F1 @ t) / (6 * math.pi * MU * Rs * umag))
        out["CT"].append((float(T1 @ np.cross(n, t)) if vt > 1e-14 else 0.0) / (8 * math.pi * MU * Rs * Rs * umag))
        out["Bn"].append(-float(B1 @ n) / (6 * math.pi * MU * Rs * umag))
        out["Bt"].append(-float(B1 @ t) / (6 * math.pi * MU * Rs * umag))
        out["BT"].append((float(BT1 @ np.cross(n, t)) if vt > 1e-14 else 0.0) / (8 * math.pi * MU * Rs * Rs * umag))
    box.close()
    return {k: np.array(v) for k, v in out.items()}


GAP_BIN = 0.1        # coarse lattice units
GAP_TRAIN_MIN = 0.35  # below this the 4x run itself has < 1.4 nodes in the gap
OUTLIER = 20.0        # |coefficient| above this: a fill/erase or shared-node spike


def gap_bins(run, R):
    """Bin index (GAP_BIN wide from GAP_TRAIN_MIN, coarse units) of every
    sample of a run after the warm-up path; -1 outside the binned range and
    for non-finite samples (a node uncovered with no fluid node within two
    belts gets zero density, as in the reference, and poisons the run from
    there on; at 4x resolution that happens below a coarse gap of 0.1)."""
    g = run["gap"] * R
    ok = np.isfinite(g) & (run["s"] > WARM)
    b = np.floor((np.where(ok, g, -1.0) - GAP_TRAIN_MIN) / GAP_BIN).astype(int)
    nb = int(round((GAP_START - GAP_TRAIN_MIN) / GAP_BIN))
    return np.where(ok & (b >= 0) & (b < nb), b, -1)


def bin_runs(c, f, R, speed, run_id):
    """Median coefficients of a coarse/fine run pair in bins of the coarse
    gap (GAP_BIN wide, from GAP_TRAIN_MIN to GAP_START), after the warm-up
    path.  Medians, not means: fill/erase events put impulses of tens of
    Stokes units on single steps.  Bins with a median beyond OUTLIER (an
    artefact of surface nodes shared by the two spheres that lasts for the
    whole bin) are dropped.  Rows follow COLS."""
    rows = []
    bc, bf = gap_bins(c, R), gap_bins(f, R)
    keys = ("gap", "vn", "vt", "Cn", "Ct", "CT")
    for b in range(int(round((GAP_START - GAP_TRAIN_MIN) / GAP_BIN))):
        mc, mf = bc == b, bf == b
        if not (mc.any() and mf.any()):
            continue
        vc = {k: float(np.median(c[k][mc])) for k in keys}
        vf = {k: float(np.median(f[k][mf])) for k in keys}
        coef = [vc["Cn"], vc["Ct"], vc["CT"], vf["Cn"], vf["Ct"], vf["CT"]]
        if max(abs(x) for x in coef) > OUTLIER:
            continue
        rows.append([vc["gap"], vc["vn"] * speed, vc["vt"] * speed, R] + coef + [run_id, b])
    return rows


def dataset(radii=(4.0, 6.0, 8.0), angles=(0.0, 40.0), speed=SPEED, storage=0, log=None):
    """Coarse and fine pair runs for every radius and angle, binned by
    bin_runs; returns the rows (columns COLS)."""
    rows = []
    for ri, R in enumerate(radii):
        for ai, ang in enumerate(angles):
            if log:
                log(f"pair runs R={R} angle={ang}")
            c = run_pair(R, ang, 1, speed, storage=storage)
            f = run_pair(R, ang, FINE, speed, storage=storage)
            rows += bin_runs(c, f, R, speed, ri * len(angles) + ai)
    return np.array(rows)


COLS = ["gap_R", "vn", "vt", "R", "Cn_c", "Ct_c", "CT_c", "Cn_f", "Ct_f", "CT_f", "run", "bin"]
RUN_KEYS = ("s", "gap", "vn", "vt", "Cn", "Ct", "CT")


def targets(data):
    """Closure targets in coarse force units: (fine - coarse) coefficients
    times the Stokes scales of the coarse run."""
    R = data[:, 3]
    u = np.hypot(data[:, 1], data[:, 2])
    sF = 6 * math.pi * MU * R * u
    sT = 8 * math.pi * MU * R * R * u
    return np.stack([(data[:, 7] - data[:, 4]) * sF, (data[:, 8] - data[:, 5]) * sF,
                     (data[:, 9] - data[:, 6]) * sT], axis=1)


def lubrication_normal(data, lub_k=1.0, lub_rmin=1.0, lub_rcap=0.1):
    """Normal force of the reference's lubrication term (addHertzian,
    kernels_fluid_PART.CUF:1463-1501) at the data's gaps and normal
    velocities, in coarse force units (repulsive > 0)."""
    R = data[:, 3]
    rpar = 2 * R
    r = rpar + data[:, 0] * R
    rr = np.maximum(r, rpar + lub_rcap)
    fac = lub_k * 6 * math.pi * R ** 4 / rpar ** 2
    f = MU * fac * data[:, 1] * (1.0 / (rr - rpar) - 1.0)
    return np.where(r <= rpar + lub_rmin, f, 0.0)


DATA = os.path.join(os.path.dirname(__file__), "data", "closure_pairs.npz")


def load(path=DATA):
    """The committed pair-approach dataset (rows of COLS), generated on an
    MI355X by `python -m lbcuda_amd.closure generate`."""
    return np.load(path)["data"]


def split(data, hold_every=3):
    """Training rows and held-out rows (every hold_every-th gap bin of each
    run)."""
    keep = np.ones(len(data), bool)
    for run in np.unique(data[:, 10]):
        idx = np.where(data[:, 10] == run)[0]
        keep[idx[1::hold_every]] = False
    return data[keep], data[~keep]


def normalisation(x, y):
    mean = x.mean(0)
    scale = x.std(0)
    scale[scale == 0] = 1.0
    oscale = np.abs(y).max(0)
    oscale[oscale == 0] = 1.0
    return mean, scale, oscale


def glorot(seed=0):
    """Glorot-uniform initial weights, zero biases, in the parameter layout
    of LBM_MLP_NPARAM."""
    rng = np.random.default_rng(seed)
    parts = []
    for shape in ((64, 4), (64,), (64, 64), (64,), (3, 64), (3,)):
        if len(shape) == 2:
            lim = math.sqrt(6.0 / (shape[0] + shape[1]))
            parts.append(rng.uniform(-lim, lim, shape).ravel())
        else:
            parts.append(np.zeros(shape))
    return np.concatenate(parts).astype(np.float32)


def fit(data, epochs=3000, batch=128, lr=2e-3, seed=1, gap_max=GAP_START):
    """Train the closure network with the GPU kernels.  Returns (block,
    final normalised loss)."""
    x = data[:, 0:4]
    y = targets(data)
    mean, scale, oscale = normalisation(x, y)
    xn = ((x - mean) / scale).astype(np.float32)
    yn = (y / oscale).astype(np.float32)
    p, loss = L.mlp_train(glorot(seed), xn, yn, epochs, batch, lr, seed)
    lo, hi = x.min(0), x.max(0)
    block = np.concatenate([p, mean, scale, oscale, lo, hi, [gap_max]]).astype(np.float32)
    assert block.size == L.MLP_CLOSURE_NFLOAT
    return block, loss


def evaluate(block, data):
    """Error of the coarse force with and without corrections against the
    4x-resolved force on the data rows, per output (normal force, tangential
    force, torque), as RMS in coarse force units: no correction, the
    reference's lubrication term (normal only) and the learned closure."""
    y = targets(data)
    pred = predict(block, data)
    lub = lubrication_normal(data)
    rms = lambda e: np.sqrt(np.mean(e ** 2, axis=0))
    none = rms(y)
    lubr = rms(y - np.stack([lub, np.zeros_like(lub), np.zeros_like(lub)], 1))
    learned = rms(y - pred)
    return {"none": none, "lubrication": lubr, "learned": learned}


def predict(block, data):
    """Closure output (coarse force units) for the data's features."""
    p = block[:L.MLP_NPARAM]
    mean = block[L.MLP_NPARAM:L.MLP_NPARAM + 4]
    scale = block[L.MLP_NPARAM + 4:L.MLP_NPARAM + 8]
    oscale = block[L.MLP_NPARAM + 8:L.MLP_NPARAM + 11]
    xn = ((data[:, 0:4] - mean) / scale).astype(np.float32)
    return L.mlp_forward(p, xn) * oscale


def main(argv=None):
    """python -m lbcuda_amd.closure generate OUT.npz  -- run the pair
    approaches on the GPU and save the binned data (columns COLS)."""
    import argparse
    import sys
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["generate"])
    ap.add_argument("out")
    ap.add_argument("--radii", default="4,6,8")
    ap.add_argument("--angles", default="0,40")
    a = ap.parse_args(argv)
    t0 = time.time()
    radii = tuple(float(x) for x in a.radii.split(","))
    angles = tuple(float(x) for x in a.angles.split(","))
    d = dataset(radii, angles, log=lambda m: print(f"[closure] {time.time() - t0:7.1f}s {m}", file=sys.stderr,
                                                  flush=True))
    np.savez(a.out, data=d, cols=np.array(COLS), speed=SPEED, radii=np.array(radii), angles=np.array(angles))
    print(f"[closure] {d.shape[0]} rows -> {a.out} in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
