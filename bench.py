"""Benchmark of the colour-gradient LB step with colloids on one MI355X.

Workload (the headline case): 256^3 periodic two-component colour-gradient
fluid (sigma 0.02, beta 0.7, tau 1, 50/50 random colour field) with 1280
rigid spheres of radius 5.5 (neutral wetting, Hertz + lubrication, FP64
particle state).  A "step" is one full time step of fluid + particles with all
inputs resident in HBM.

Printed: one JSON line.  value = MLUPS over interior nodes (N^3 per step) of
the FP32-stored path; the FP16-stored path is timed the same way and reported
beside it (metric: MLUPS, % of HBM peak, FP16-store vs FP32 speedup).
roofline: algorithmic bytes of the collide-stream kernel per launch (2 colours
x 19 populations x read+write x bytes per value + 1 flag byte, per node) over
its average duration measured with HIP events on the library stream.
cpu_baseline: the CPU oracle (FP32 instantiation, OpenMP) on the same box for
a few steps, on this host's cores.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(storage_tag):
    """HBM bytes per launch of the fused collide-stream kernel, from the
    committed rocprofv3 counter summary (profiles/k_fused_256_pickering.md:
    FETCH_SIZE and WRITE_SIZE scaled by the same counters of a copy kernel of
    known size), else None."""
    path = os.path.join(ROOT, "profiles", "k_fused_256_pickering.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0] == storage_tag and cells[1].startswith("k_fused"):
            try:
                return float(cells[5])
            except ValueError:
                return None
    return None


def run_gpu(args, storage, rr, rb, pos):
    from lbcuda_amd.lib import Box
    from lbcuda_amd.abi import INIT_RHO
    from lbcuda_amd import workloads as wl
    cfg = wl.pickering_config(args.n, args.particles, storage=storage)
    box = Box(cfg)
    if args.particles:
        box.set_particles(pos)
    box.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    if not args.no_autotune:
        t, zc, ms = box.autotune(reps=3)
        log(f"storage={storage}: autotune picked tile {t} z-chunk {zc} ({ms:.3f} ms per launch)")
    log(f"storage={storage}: warmup {args.warmup} steps")
    rc = box.step(args.warmup)
    if rc:
        raise RuntimeError(f"warmup raised {box.status()}")
    t0 = time.perf_counter()
    rc = box.step(args.steps)       # returns after the stream has drained
    t1 = time.perf_counter()
    if rc:
        raise RuntimeError(f"timed steps raised {box.status()}")
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    mlups = args.n ** 3 / (ms_per_step * 1e-3) / 1e6
    log(f"storage={storage}: {ms_per_step:.3f} ms/step, {mlups:.0f} MLUPS")
    # kernel timing: events around every collide-stream launch
    box.kernel_timing(True)
    box.step(args.prof_steps)
    tot_ms, nl = box.get_kernel_timing()
    box.kernel_timing(False)
    name, bpl = box.kernel_info()
    avg_ms = tot_ms / max(nl, 1)
    achieved = bpl * args.n ** 3 / (avg_ms * 1e-3) / 1e9
    sums = box.sums()
    box.close()
    return dict(ms_per_step=ms_per_step, mlups=mlups, kernel=name, bytes_per_lu=bpl, kernel_ms=avg_ms,
                achieved_gbs=achieved, mass=sums[0:2].tolist())


def run_ensemble(args):
    """64 independent CG boxes of uneven size (edges 48..96, 2..16 spheres of
    radius 4..8 each), one persistent fluid launch per step."""
    from lbcuda_amd.lib import Ensemble
    from lbcuda_amd.abi import INIT_RHO
    from lbcuda_amd import workloads as wl
    cfgs, inputs = wl.ensemble_workload(64, storage=0)
    ens = Ensemble(cfgs)
    for b, (rr, rb, pos) in zip(ens.boxes, inputs):
        b.set_particles(pos)
        b.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    tile, zc, _ = ens.autotune(3)
    ens.step(5)
    steps = max(10, args.steps // 4)
    t0 = time.perf_counter()
    rc = ens.step(steps)
    dt = time.perf_counter() - t0
    ens.close()
    if rc:
        raise RuntimeError(f"ensemble step returned {rc}")
    nodes = sum(c.nx * c.ny * c.nz for c in cfgs)
    log(f"ensemble: 64 boxes, {nodes} nodes, {dt / steps * 1e3:.3f} ms/step")
    return {"boxes": 64, "nodes": nodes, "particles": sum(c.n_particles for c in cfgs), "steps": steps,
            "ms_per_step": round(dt / steps * 1e3, 4), "mlups": round(nodes * steps / dt / 1e6, 1),
            "tile": tile, "zchunk": zc}


def cpu_baseline(args, rr, rb, pos):
    from oracle.pyoracle import OracleBox, lib as orc_lib
    from lbcuda_amd.abi import INIT_RHO
    from lbcuda_amd import workloads as wl
    cfg = wl.pickering_config(args.n, args.particles, storage=0)
    o = OracleBox(cfg, 32)
    if args.particles:
        o.set_particles(pos)
    o.init_fields(INIT_RHO, rho_r=rr, rho_b=rb)
    o.step(1)
    t0 = time.perf_counter()
    o.step(args.cpu_steps)
    dt = time.perf_counter() - t0
    cores = orc_lib().orc_num_threads()
    o.close()
    mlups = args.n ** 3 * args.cpu_steps / dt / 1e6
    log(f"cpu oracle fp32: {dt:.1f} s for {args.cpu_steps} steps on {cores} threads, {mlups:.2f} MLUPS")
    return dict(value=round(mlups, 3), unit="MLUPS", cores=cores, kind="port",
                sample=f"{args.n}^3 CG + {args.particles} colloids, {args.cpu_steps} steps after 1 warm-up step, "
                       f"FP32 OpenMP restatement of the reference step")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--particles", type=int, default=1280)
    ap.add_argument("--prof-steps", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs one box on one GPU; multi-GPU runs are not part of it "
              "(use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    from lbcuda_amd import workloads as wl
    log(f"inputs: {args.n}^3, {args.particles} spheres")
    rr, rb, pos = wl.pickering_inputs(args.n, args.particles)

    from lbcuda_amd.lib import copy_bandwidth, stream_bandwidth
    copy_gbs = copy_bandwidth(2 << 30, 10)
    log(f"device copy kernel: {copy_gbs:.0f} GB/s (read + write)")
    stream_gbs = stream_bandwidth(args.n ** 3, 10)
    log(f"19-plane stream with the push's x shifts: {stream_gbs:.0f} GB/s (read + write)")
    r32 = run_gpu(args, 0, rr, rb, pos)
    r16 = None if args.no_fp16 else run_gpu(args, 1, rr, rb, pos)
    ens = None if args.no_ensemble else run_ensemble(args)
    cpu = None if args.no_cpu else cpu_baseline(args, rr, rb, pos)

    traffic = traffic_from_profiles("f32")
    traffic16 = traffic_from_profiles("f16")
    out = {
        "metric": "MLUPS @256^3 CG+1280 colloids on 1 GPU; % HBM peak; FP16-store vs FP32 speedup",
        "value": round(r32["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"pickering emulsion {args.n}^3 two-component colour-gradient D3Q19 + "
                               f"{args.particles} spheres R=5.5",
                   "box": [args.n] * 3, "particles": args.particles, "storage": "f32",
                   "sigma": 0.02, "beta": 0.7, "tau": 1.0},
        "roofline": {"bound": "hbm", "achieved": round(r32["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r32["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": r32["kernel"], "bytes_per_lu": r32["bytes_per_lu"],
                     "kernel_ms": round(r32["kernel_ms"], 4)},
        "mlups_ref_style": round(r32["mlups"] * ((args.n + 2) / args.n) ** 3, 1),
        "copy_kernel_gbs": round(copy_gbs, 1),
        "stream19_gbs": round(stream_gbs, 1),
        "cpu_baseline": cpu,
        "ensemble": ens,
    }
    if r16:
        out["fp16_store"] = {"mlups": round(r16["mlups"], 1), "ms_per_step": round(r16["ms_per_step"], 4),
                             "speedup_vs_f32": round(r16["mlups"] / r32["mlups"], 3),
                             "kernel_speedup_vs_f32": round(r32["kernel_ms"] / r16["kernel_ms"], 3),
                             "kernel_achieved_gbs": round(r16["achieved_gbs"], 1),
                             "kernel_frac": round(r16["achieved_gbs"] / HBM_PEAK_GBS, 4),
                             "kernel_ms": round(r16["kernel_ms"], 4), "kernel": r16["kernel"],
                             "traffic": traffic16}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
