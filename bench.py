"""Benchmark: Mpoint-iterations/s of the q-LSKUM pseudo-time iteration.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2_2m]

A step is one fpi_solver iteration over the whole cloud.  Default workload:
the 2M-point ellipsoid cloud (120k wall, 40k outer, 1.84M interior, k=25),
FP64 explicit, Mach 0.84, AoA 3.06 deg, CFL 0.2, freestream start; the cloud
is generated from its seed and resident in HBM before the timed region.
W warmup iterations, then K timed iterations, each side bracketed by a
device synchronize.

Reported beside the throughput:
  roofline      the fused interior flux-residual kernel (the dominant
                kernel): algorithmic bytes per launch (the reference's
                eval_flux_residual traffic, 388 + 4*(6 + list entries) bytes
                per point, for the points that launch processes) / its mean
                launch time, from HIP events recorded on the library's stream
                around every launch of the timed iterations, against 8 TB/s;
                traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE and
                WRITE_SIZE when a summary exists in profiles/;
  cpu_baseline  the C restatement of the reference (bit-identical to the
                reference Fortran on the golden fixtures), one core, one
                iteration of the same cloud from the GPU's state.
  mixed         the same W + K iterations in the mixed-precision form (FP32
                q-variables, gradients and split fluxes): its rate, its
                speed-up over FP64 and its wall-Cp error against the FP64
                state after the same iterations.
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG = importlib.import_module("3d-mfcfd_amd")
pc = PKG.pointcloud
solver = PKG.solver
roofline = importlib.import_module("3d-mfcfd_amd.roofline")
postprocess = PKG.postprocess

METRIC = "Mpoint-iter/s on 1 GPU, FP64 vs mixed, %HBM roofline, wall-Cp err vs FP64"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(config):
    """HBM bytes per interior flux launch from a committed rocprof summary."""
    path = os.path.join(ROOT, "profiles", f"flux_interior_{config}_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def cpu_baseline(cloud, cfg, state, t):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    o = orc.Oracle(cloud, cfg["mach"], cfg["aoa"], cfg["cfl"], prim=state["prim"])
    o.delUp[:] = state["delUp"]
    o.delUn[:] = state["delUn"]
    t0 = time.perf_counter()
    o.fpi_solver(t, cfg["mode"] == "implicit")
    dt = time.perf_counter() - t0
    return {"value": cloud.n / dt / 1e6, "unit": "Mpoint-iter/s", "cores": 1, "kind": "port",
            "sample": f"1 iteration (t={t}) of the same {cloud.n}-point cloud from the GPU state, "
                      f"C restatement of the reference on one host core, {dt:.1f} s"}


def copy_probe(mib):
    """A plain device copy of known size (reads and writes mib MiB each
    launch, three launches) for scaling the HBM traffic counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    log(f"copy probe: 3 launches of {mib} MiB read + {mib} MiB written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c2_2m", choices=sorted(pc.CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--phase-iters", type=int, default=3)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this size (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build runs "
              f"the q-LSKUM iteration on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    cfg = pc.CONFIGS[args.config]
    log(f"generating {args.config}")
    t0 = time.perf_counter()
    cloud = pc.config_cloud(args.config, workers=16)
    log(f"cloud {cloud.counts()} in {time.perf_counter() - t0:.1f} s")
    prim0 = np.tile(pc.freestream_prim(cfg["mach"], cfg["aoa"]), (cloud.n, 1))
    cnt, lists = solver.split_stencils(cloud)
    alias = solver.alias_order(cloud)
    s = solver.Solver(cloud, cfg["mach"], cfg["aoa"], cfg["cfl"], mode=cfg["mode"],
                      split=(cnt, lists), alias=alias)
    s.upload("prim", prim0)
    log(f"bound; sweep schedule {s.sweep_info()}")

    W, K = args.warmup, args.steps
    if W > 0:
        s.run(1, W)
    s.synchronize()
    log(f"warmup {W} done")
    s.kernel_timing(True)
    t0 = time.perf_counter()
    rn, _ = s.run(W + 1, K)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    flux_total_ms, flux_launches = s.kernel_time()
    s.kernel_timing(False)
    ms_per_step = elapsed / K * 1000.0
    value = cloud.n * K / elapsed / 1e6
    log(f"timed {K} iterations: {ms_per_step:.3f} ms/iter, {value:.2f} Mpoint-iter/s, "
        f"res_new {rn[0]:.6e} -> {rn[-1]:.6e}")

    phases = s.time_phases(W + K + 1, args.phase_iters)
    log("phases (ms): " + ", ".join(f"{k} {v:.3f}" for k, v in phases.items()))
    t_next = W + K + args.phase_iters + 1
    by = roofline.iteration_bytes(cloud.status, cnt, cloud.nbhs, cfg["mode"])
    cls = s.point_class_map()
    flux_bytes = int(roofline.flux_bytes_per_point(cloud.status, cnt, cloud.nbhs)[cls == 0].sum())
    flux_ms = flux_total_ms / max(flux_launches, 1)
    log(f"fused interior flux kernel: {flux_launches} launches, mean {flux_ms:.4f} ms, "
        f"{int(np.sum(cls == 0))} points, {flux_bytes} algorithmic bytes per launch")
    achieved = flux_bytes / (flux_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": roofline.HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / roofline.HBM_PEAK_GBS, 5),
            "traffic": read_traffic(args.config),
            "kernel": "k_flux_fused<0,true> (interior points on the fused path)",
            "bytes_per_launch": flux_bytes, "launch_ms": round(flux_ms, 4),
            "launches_timed": flux_launches}
    step_gbs = by["total"] / (ms_per_step * 1e-3) / 1e9

    prim_fp64 = s.download("prim")
    mixed = None
    if not args.no_mixed:
        # same start, same W + K iterations, mixed precision
        m = solver.Solver(cloud, cfg["mach"], cfg["aoa"], cfg["cfl"], mode=cfg["mode"],
                          precision="mixed", split=(cnt, lists), alias=alias)
        m.upload("prim", prim0)
        if W > 0:
            m.run(1, W)
        m.synchronize()
        t0 = time.perf_counter()
        rn_m, _ = m.run(W + 1, K)
        m.synchronize()
        el_m = time.perf_counter() - t0
        prim_m = m.download("prim")
        m.close()
        wall = cloud.status == pc.STATUS_WALL
        cp64 = postprocess.wall_cp(prim_fp64[wall], cfg["mach"])
        cpm = postprocess.wall_cp(prim_m[wall], cfg["mach"])
        v_m = cloud.n * K / el_m / 1e6
        mixed = {"value": round(v_m, 3), "ms_per_step": round(el_m / K * 1000.0, 4),
                 "speedup_vs_fp64": round(v_m / value, 3),
                 "wall_cp_err_vs_fp64": float(np.max(np.abs(cpm - cp64)) / np.max(np.abs(cp64))),
                 "res_new_last": float(rn_m[-1]), "res_new_last_fp64": float(rn[-1])}
        log(f"mixed: {mixed}")

    cpu = None
    if not args.no_cpu:
        state = {nm: s.download(nm) for nm in ("prim", "delUp", "delUn")}
        s.close()
        log("cpu baseline: one oracle iteration on one core")
        cpu = cpu_baseline(cloud, cfg, state, t_next)
        log(f"cpu baseline {cpu['value']:.4f} Mpoint-iter/s")
    else:
        s.close()

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpoint-iter/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.config}: ellipsoid cloud {cloud.n} points "
                               f"(wall {cfg['n_wall']}, outer {cfg['n_outer']}, interior "
                               f"{cfg['n_interior']}, k=25), FP64 {cfg['mode']}, Mach "
                               f"{cfg['mach']}, AoA {cfg['aoa']} deg, CFL {cfg['cfl']}",
                   "points": cloud.n, "mode": cfg["mode"], "precision": "fp64"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "phases_ms": {k: round(v, 4) for k, v in phases.items()},
        "step_algorithmic_gbs": round(step_gbs, 2),
        "step_hbm_frac": round(step_gbs / roofline.HBM_PEAK_GBS, 5),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
