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


def class_counts(cls, status):
    """Points per flux kernel class: fused interior, fused general (and how many
    of those are interior points whose lists the spill altered), list-by-list."""
    return {"fused_interior": int(np.sum(cls == 0)), "fused_general": int(np.sum(cls == 1)),
            "fused_general_interior": int(np.sum((cls == 1) & (status == pc.STATUS_INTERIOR))),
            "listwise": int(np.sum(cls == 2))}


def bench_batch(args):
    """64 independent clouds (log-uniform 20k..500k points, aspect 2..6,
    Mach 0.3..0.85, AoA -5..10 deg) in one context, the iteration captured
    as a HIP graph, launch shapes from the autotuner."""
    cfg = pc.BATCH
    plan = pc.batch_plan()
    log(f"generating {len(plan)} clouds")
    t0 = time.perf_counter()
    clouds = pc.batch_clouds(plan, workers=16)
    n = int(sum(c.n for c in clouds))
    log(f"{n} points in {time.perf_counter() - t0:.1f} s")
    machs = [p["mach"] for p in plan]
    aoas = [p["aoa"] for p in plan]
    p0 = np.concatenate([np.tile(pc.freestream_prim(m, a), (c.n, 1)) for c, m, a in zip(clouds, machs, aoas)])
    W, K = args.warmup, args.steps
    legs = {}
    for precision in (["fp64"] if args.no_mixed else ["fp64", "mixed"]):
        bs = solver.BatchSolver(clouds, machs, aoas, cfg["cfl"], mode=cfg["mode"], precision=precision)
        bs.upload("prim", p0)
        if W > 0:
            bs.run_graph(1, W)
        # shapes tuned on the state the timed iterations start from
        tuning = None if args.no_tune else bs.autotune()
        bs.synchronize()
        t0 = time.perf_counter()
        _, inst = bs.run_graph(W + 1, K)
        bs.synchronize()
        el = time.perf_counter() - t0
        # the dominant kernel, timed with events over two eager iterations
        bs.kernel_timing(True)
        bs.run(W + K + 1, 2)
        fms, fl = bs.kernel_time()
        bs.kernel_timing(False)
        cls = bs.point_class_map()
        prim = bs.download("prim")
        bs.close()
        legs[precision] = dict(el=el, inst=inst, tuning=tuning, flux_ms=fms / max(fl, 1), cls=cls, prim=prim)
        log(f"{precision}: {el / K * 1000:.3f} ms/iter, {n * K / el / 1e6:.2f} Mpoint-iter/s")
    f64 = legs["fp64"]
    value = n * K / f64["el"] / 1e6
    status = np.concatenate([c.status for c in clouds])
    cnt = np.concatenate([solver.split_stencils(c)[0] for c in clouds], axis=1)
    nbhs = np.concatenate([c.nbhs for c in clouds])
    flux_bytes = int(roofline.flux_bytes_per_point(status, cnt, nbhs)[f64["cls"] == 0].sum())
    achieved = flux_bytes / (f64["flux_ms"] * 1e-3) / 1e9
    last = f64["inst"][-1]
    mixed = None
    if "mixed" in legs:
        mx = legs["mixed"]
        offs = np.concatenate([[0], np.cumsum([c.n for c in clouds])])
        errs = []
        for b, c in enumerate(clouds):
            w = offs[b] + np.flatnonzero(c.status == pc.STATUS_WALL)
            a = postprocess.wall_cp(f64["prim"][w], machs[b])
            m = postprocess.wall_cp(mx["prim"][w], machs[b])
            errs.append(float(np.max(np.abs(m - a))))
        vm = n * K / mx["el"] / 1e6
        mixed = {"value": round(vm, 3), "ms_per_step": round(mx["el"] / K * 1000, 4),
                 "speedup_vs_fp64": round(vm / value, 3), "wall_cp_err_vs_fp64_max": max(errs),
                 "tuning": mx["tuning"]}
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpoint-iter/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": round(f64["el"] / K * 1000.0, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"c5_batch: {len(clouds)} ellipsoid clouds, {n} points in all "
                               f"(log-uniform {cfg['n_min']}..{cfg['n_max']}), aspect "
                               f"{cfg['aspect'][0]}..{cfg['aspect'][1]}, Mach {cfg['mach'][0]}..{cfg['mach'][1]}, "
                               f"AoA {cfg['aoa'][0]}..{cfg['aoa'][1]} deg, FP64 {cfg['mode']}, CFL {cfg['cfl']}, "
                               f"one HIP graph per iteration",
                   "points": n, "clouds": len(clouds), "mode": cfg["mode"], "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": roofline.HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / roofline.HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "k_flux_fused<0,true,double> (interior points on the fused path)",
                     "bytes_per_launch": flux_bytes, "launch_ms": round(f64["flux_ms"], 4)},
        "cpu_baseline": None,
        "mixed": mixed,
        "tuning": f64["tuning"],
        "instance_residual_last": {"min": float(last.min()), "max": float(last.max()),
                                   "finite": int(np.isfinite(last).sum())},
        "point_classes": class_counts(f64["cls"], status),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c2_2m", choices=sorted(pc.CONFIGS) + ["c5_batch"])
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--phase-iters", type=int, default=3)
    ap.add_argument("--cp-iters", type=int, default=0,
                    help="iterations at which FP64 and mixed wall Cp are compared (default: the "
                         "configuration's own count for explicit runs, 6 for implicit ones)")
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

    if args.config == "c5_batch":
        bench_batch(args)
        return
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
    log("warmup")
    if W > 0:
        s.run(1, W)
    s.synchronize()
    # shapes tuned on the state the timed iterations start from
    tuning = None
    if not args.no_tune:
        log("autotune")
        tuning = s.autotune()
        log(f"autotune: {tuning}")
    log(f"warmup {W} done; timed iterations")
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

    # accuracy: both forms from the same start to the same iteration count
    n_done = W + K + args.phase_iters
    # Implicit runs compare at iteration 6: the reference updates outer
    # points explicitly (U = Ubar - dt*R) with the implicit run's CFL-50 local
    # step, and that update amplifies a far-field disturbance about 1.8x per
    # iteration, in FP64 and in the oracle alike.  On the 8M cloud the FP64
    # pressures turn non-finite from iteration 14, the mixed form's (seeded
    # at FP32 rounding) from iteration 9.
    n_cp = args.cp_iters if args.cp_iters > 0 else (cfg["iters"] if cfg["mode"] == "explicit" else 6)
    n_cp = max(n_cp, n_done)
    if not args.no_mixed and n_cp > n_done:
        log(f"FP64 to iteration {n_cp} for the wall-Cp comparison")
        s.run(n_done + 1, n_cp - n_done)
    else:
        n_cp = n_done
    t_next = n_cp + 1
    prim_fp64 = s.download("prim")
    mixed = None
    if not args.no_mixed:
        # same start, same W + K iterations, mixed precision
        m = solver.Solver(cloud, cfg["mach"], cfg["aoa"], cfg["cfl"], mode=cfg["mode"],
                          precision="mixed", split=(cnt, lists), alias=alias)
        m.upload("prim", prim0)
        log("mixed leg: warmup")
        if W > 0:
            m.run(1, W)
        m.synchronize()
        log("mixed leg: autotune")
        tuning_m = m.autotune() if not args.no_tune else None
        log("mixed leg: timed iterations")
        t0 = time.perf_counter()
        rn_m, _ = m.run(W + 1, K)
        m.synchronize()
        el_m = time.perf_counter() - t0
        phases_m = m.time_phases(W + K + 1, args.phase_iters) if args.phase_iters > 0 else None
        if n_cp > n_done:
            log(f"mixed to iteration {n_cp} for the wall-Cp comparison")
            m.run(n_done + 1, n_cp - n_done)
        prim_m = m.download("prim")
        m.close()
        wall = cloud.status == pc.STATUS_WALL
        cp64 = postprocess.wall_cp(prim_fp64[wall], cfg["mach"])
        cpm = postprocess.wall_cp(prim_m[wall], cfg["mach"])
        v_m = cloud.n * K / el_m / 1e6
        mixed = {"value": round(v_m, 3), "ms_per_step": round(el_m / K * 1000.0, 4),
                 "speedup_vs_fp64": round(v_m / value, 3),
                 # Cp is already normalised by the dynamic pressure: the
                 # absolute difference is the error; the relative one is
                 # only meaningful once the wall flow has developed
                 "wall_cp_err_vs_fp64": float(np.max(np.abs(cpm - cp64))),
                 "wall_cp_iterations": n_cp,
                 "nonfinite_points": [int((~np.isfinite(prim_fp64)).any(1).sum()),
                                      int((~np.isfinite(prim_m)).any(1).sum())],
                 "wall_cp_err_rel_max_cp": float(np.max(np.abs(cpm - cp64)) / np.max(np.abs(cp64))),
                 "res_new_last": float(rn_m[-1]), "res_new_last_fp64": float(rn[-1]),
                 "phases_ms": {k: round(v, 4) for k, v in phases_m.items()} if phases_m else None,
                 "tuning": tuning_m}
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
        "tuning": tuning,
        "phases_ms": {k: round(v, 4) for k, v in phases.items()},
        "point_classes": class_counts(cls, cloud.status),
        "step_algorithmic_gbs": round(step_gbs, 2),
        "step_hbm_frac": round(step_gbs / roofline.HBM_PEAK_GBS, 5),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
