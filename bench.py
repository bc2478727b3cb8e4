"""Benchmark: Mpoint-iterations/s of the q-LSKUM pseudo-time iteration.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3_8m]

A step is one fpi_solver iteration over the whole cloud.  Default workload:
the 8M-point ellipsoid cloud (400k wall, 120k outer, 7.48M interior, k=25,
Morton order), FP64 implicit LU-SGS, Mach 0.84, AoA 3.06 deg, CFL 50, from
the freestream, generated from its seed and resident in HBM before the
timed region; its mixed-precision twin is reported beside it.

Only finite states are timed.  At CFL 50 the reference's own canonical
implicit pairing diverges (tests/golden/cfl50_20k_history.npz: the far
field grows until the pressures turn non-finite), and on the 8M cloud the
FP64 run turns non-finite at iteration 14, the mixed one at iteration 9-10.  So
the state after W warmup iterations is saved, and the K timed iterations run
in blocks of at most --block iterations, each block starting from that saved
state (restored by host copies outside the timed region): with the
defaults, iterations 4..7 five times.  Explicit configurations (c2_2m) time
their K iterations in one block.

Reported beside the throughput:
  roofline      the dominant kernel -- the sync-free sweep kernel for an
                implicit run (algorithmic bytes per launch: 224 + 8 per list
                entry for every swept point, roofline.sweep_bytes_per_point),
                the fused interior flux kernel for an explicit one -- divided
                by its mean launch time from HIP events recorded on the
                library's stream around every launch of the timed
                iterations, against 8 TB/s; traffic = HBM bytes per launch
                from the rocprofv3 FETCH_SIZE and WRITE_SIZE summary in
                profiles/ when one exists, valu from the rocprofv3 SQ
                counter summary when one exists;
  cpu_baseline  the C restatement of the reference (bit-identical to the
                reference Fortran on the golden fixtures) on one core, on a
                bounded sample of the same workload;
  mixed         the same iterations in the mixed-precision form (FP32
                q-variables, gradients and split fluxes): its rate, its
                speed-up over FP64, and its wall Cp, CL/CD and residual drop
                against the FP64 state after the same iterations;
  c2_2m         (implicit default only) the 2M-point explicit FP64
                configuration's rate, roofline and mixed speed-up.
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


def read_profile_values(name):
    """key = value lines of a committed rocprof summary in profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    vals = {}
    if not os.path.exists(path):
        return vals
    with open(path) as f:
        for line in f:
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                try:
                    vals[k.strip()] = float(v.strip())
                except ValueError:
                    pass
    return vals


def cpu_baseline_sample(cfg, n_sample):
    """The C oracle on one core: a bounded sample of the configuration --
    the same generator and flow on n_sample points, timed over its first two
    iterations from the freestream."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    n_wall = int(round(n_sample * cfg["n_wall"] / (cfg["n_wall"] + cfg["n_outer"] + cfg["n_interior"])))
    n_outer = int(round(n_sample * cfg["n_outer"] / (cfg["n_wall"] + cfg["n_outer"] + cfg["n_interior"])))
    cl = pc.ellipsoid_cloud(n_wall, n_outer, n_sample - n_wall - n_outer, seed=cfg["seed"], order=cfg["order"],
                            workers=16)
    prim0 = np.tile(pc.freestream_prim(cfg["mach"], cfg["aoa"]), (cl.n, 1))
    o = orc.Oracle(cl, cfg["mach"], cfg["aoa"], cfg["cfl"], prim=prim0)
    t0 = time.perf_counter()
    o.run(1, 2, cfg["mode"])
    dt = time.perf_counter() - t0
    return {"value": round(2 * cl.n / dt / 1e6, 5), "unit": "Mpoint-iter/s", "cores": 1, "kind": "port",
            "sample": f"iterations 1-2 of a {cl.n}-point cloud from the same generator and flow "
                      f"({cfg['mode']}, CFL {cfg['cfl']}), C restatement of the reference on one host core, "
                      f"{dt:.1f} s"}


def cpu_baseline_batch(clouds, plan):
    """The C oracle on one core for the batch: its median-size cloud, the
    first two iterations from the freestream (each cloud iterates on its own;
    the batch's rate is the same work cloud after cloud)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    b = int(np.argsort([c.n for c in clouds])[len(clouds) // 2])
    cl, p = clouds[b], plan[b]
    prim0 = np.tile(pc.freestream_prim(p["mach"], p["aoa"]), (cl.n, 1))
    o = orc.Oracle(cl, p["mach"], p["aoa"], pc.BATCH["cfl"], prim=prim0)
    t0 = time.perf_counter()
    o.run(1, 2, pc.BATCH["mode"])
    dt = time.perf_counter() - t0
    return {"value": round(2 * cl.n / dt / 1e6, 5), "unit": "Mpoint-iter/s", "cores": 1, "kind": "port",
            "sample": f"iterations 1-2 of batch cloud {b} ({cl.n} points, the median size, Mach {p['mach']:.3f}, "
                      f"AoA {p['aoa']:.2f} deg), C restatement of the reference on one host core, {dt:.1f} s"}


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


STATE = ("prim", "delUp", "delUn")


def run_leg(cloud, cfg, split, alias, precision, W, K, block, tune, phase_iters, n_cp):
    """W warmup iterations, then K timed iterations in blocks of at most
    `block`, each from the state after the warmup; then the state at
    iteration n_cp (from the same saved state) for the accuracy comparison."""
    implicit = cfg["mode"] != "explicit"
    s = solver.Solver(cloud, cfg["mach"], cfg["aoa"], cfg["cfl"], mode=cfg["mode"], precision=precision,
                      split=split, alias=alias)
    s.upload("prim", np.tile(pc.freestream_prim(cfg["mach"], cfg["aoa"]), (cloud.n, 1)))
    if implicit:
        log(f"{precision}: sweep plan {s.sweep_info()}")
    hist0 = np.zeros(0)
    if W > 0:
        hist0, _ = s.run(1, W)
    s.synchronize()
    saved = {nm: s.download(nm) for nm in STATE}

    def restore():
        for nm in STATE:
            s.upload(nm, saved[nm])
        s.synchronize()

    tuning = None
    if tune:
        log(f"{precision}: autotune")
        tuning = s.autotune()
        restore()
    blocks = [K] if not implicit else [min(block, K - b0) for b0 in range(0, K, block)]
    elapsed = 0.0
    kernel = "sweeps" if implicit else "flux"
    s.kernel_timing(True, kernel)
    rn_all = []
    for bi, nb in enumerate(blocks):
        if bi > 0:
            restore()
        s.synchronize()
        t0 = time.perf_counter()
        rn, _ = s.run(W + 1, nb)
        s.synchronize()
        elapsed += time.perf_counter() - t0
        rn_all.append(rn)
    k_ms, k_launches = s.kernel_time()
    s.kernel_timing(False)
    log(f"{precision}: timed {K} iterations in blocks {blocks}: {elapsed / K * 1000:.3f} ms/iter, "
        f"{cloud.n * K / elapsed / 1e6:.2f} Mpoint-iter/s; {kernel} kernel {k_launches} launches, "
        f"mean {k_ms / max(k_launches, 1):.4f} ms")
    phases = None
    done = W + K                       # explicit: iterations run so far
    if phase_iters > 0:
        if implicit:
            restore()
            phases = s.time_phases(W + 1, min(phase_iters, block))
        else:
            phases = s.time_phases(done + 1, phase_iters)
            done += phase_iters
    # the comparison state (from the saved state for implicit runs)
    if implicit:
        restore()
        rn_cp, _ = s.run(W + 1, n_cp - W)
    elif n_cp > done:
        s.run(done + 1, n_cp - done)
    prim = s.download("prim")
    cls = s.point_class_map()
    hist = np.concatenate([hist0, rn_cp]) if implicit else None
    s.close()
    return dict(elapsed=elapsed, ms=elapsed / K * 1000.0, value=cloud.n * K / elapsed / 1e6, tuning=tuning,
                kernel_ms=k_ms / max(k_launches, 1), kernel_launches=k_launches, phases=phases, prim=prim,
                cls=cls, blocks=blocks, rn=np.concatenate(rn_all), hist=hist)


def bench_config(args, name, steps, warmup, mixed=True, cpu=True):
    cfg = pc.CONFIGS[name]
    implicit = cfg["mode"] != "explicit"
    log(f"generating {name}")
    t0 = time.perf_counter()
    cloud = pc.config_cloud(name, workers=16)
    log(f"cloud {cloud.counts()} in {time.perf_counter() - t0:.1f} s")
    cnt, lists = solver.split_stencils(cloud)
    alias = solver.alias_order(cloud)
    W, K = warmup, steps
    block = args.block if implicit else K
    # implicit: compare at the last iteration of a block; explicit: at the
    # configuration's own iteration count
    n_cp = (W + block) if implicit else (args.cp_iters if args.cp_iters > 0 else cfg["iters"])
    n_cp = max(n_cp, W + 1) if implicit else max(n_cp, W + K + args.phase_iters)
    legs = {}
    for precision in (["fp64", "mixed"] if mixed else ["fp64"]):
        legs[precision] = run_leg(cloud, cfg, (cnt, lists), alias, precision, W, K, block, not args.no_tune,
                                  args.phase_iters, n_cp)
    f64 = legs["fp64"]
    status = cloud.status
    if implicit:
        per_pt = roofline.sweep_bytes_per_point(status, cnt, cloud.nbhs)
        k_bytes = int(per_pt.sum())
        kname = "k_sweep_async<FORWARD, UW, double> (forward and backward sync-free LU-SGS sweeps)"
        prof = read_profile_values(f"sweep_{name}_traffic.txt")
        valu = read_profile_values(f"sweep_{name}_valu.txt")
    else:
        k_bytes = int(roofline.flux_bytes_per_point(status, cnt, cloud.nbhs)[f64["cls"] == 0].sum())
        kname = "k_flux_fused<0, true, double> (interior points on the fused path)"
        prof = read_profile_values(f"flux_interior_{name}_traffic.txt")
        valu = read_profile_values(f"flux_interior_{name}_valu.txt")
    achieved = k_bytes / (f64["kernel_ms"] * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": roofline.HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / roofline.HBM_PEAK_GBS, 5),
            "traffic": prof.get("hbm_bytes_per_launch"), "kernel": kname, "bytes_per_launch": k_bytes,
            "launch_ms": round(f64["kernel_ms"], 4), "launches_timed": f64["kernel_launches"]}
    if valu:
        roof["valu"] = {k: valu[k] for k in sorted(valu)}
    by = roofline.iteration_bytes(status, cnt, cloud.nbhs, cfg["mode"])
    step_gbs = by["total"] / (f64["ms"] * 1e-3) / 1e9
    mixed_out = None
    if "mixed" in legs:
        mx = legs["mixed"]
        wall = status == pc.STATUS_WALL
        cp64 = postprocess.wall_cp(f64["prim"][wall], cfg["mach"])
        cpm = postprocess.wall_cp(mx["prim"][wall], cfg["mach"])
        areas = postprocess.wall_areas(cloud)
        fc64 = postprocess.force_coefficients(cloud, f64["prim"], cfg["mach"], cfg["aoa"], areas=areas)
        fcm = postprocess.force_coefficients(cloud, mx["prim"], cfg["mach"], cfg["aoa"], areas=areas)
        mixed_out = {"value": round(mx["value"], 3), "ms_per_step": round(mx["ms"], 4),
                     "speedup_vs_fp64": round(mx["value"] / f64["value"], 3),
                     "wall_cp_err_vs_fp64": float(np.max(np.abs(cpm - cp64))),
                     "compared_at_iteration": n_cp,
                     "CL": [fc64["CL"], fcm["CL"]], "CD": [fc64["CD"], fcm["CD"]],
                     "CL_err": abs(fcm["CL"] - fc64["CL"]), "CD_err": abs(fcm["CD"] - fc64["CD"]),
                     "nonfinite_points": [int((~np.isfinite(f64["prim"])).any(1).sum()),
                                          int((~np.isfinite(mx["prim"])).any(1).sum())],
                     "kernel_ms": round(mx["kernel_ms"], 4),
                     "phases_ms": {k: round(v, 4) for k, v in mx["phases"].items()} if mx["phases"] else None,
                     "tuning": mx["tuning"]}
        if implicit:
            # residual drop over the compared iterations (res_new at n_cp over
            # res_old, the residual of iteration 2: fpi_solver.f90:38-43)
            d64 = f64["hist"][n_cp - 1] / f64["hist"][1]
            dm = mx["hist"][n_cp - 1] / mx["hist"][1]
            mixed_out["residual_drop"] = [float(d64), float(dm)]
            mixed_out["residual_drop_err_rel"] = float(abs(dm - d64) / abs(d64))
        log(f"mixed: {mixed_out}")
    cpu_out = None
    if cpu:
        log("cpu baseline: the oracle on one core, bounded sample")
        cpu_out = cpu_baseline_sample(cfg, args.cpu_sample_points if implicit else 250000)
        log(f"cpu baseline {cpu_out}")
    timing = (f"warmup iterations 1-{W}, state saved; timed iterations {W + 1}-{W + block} "
              f"{len(f64['blocks'])} times (blocks {f64['blocks']}), each from the saved state "
              f"(restores outside the timed region)") if implicit else f"iterations {W + 1}-{W + K}"
    return {
        "metric": METRIC, "value": round(f64["value"], 3), "unit": "Mpoint-iter/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(f64["ms"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{name}: ellipsoid cloud {cloud.n} points (wall {cfg['n_wall']}, outer "
                               f"{cfg['n_outer']}, interior {cfg['n_interior']}, k=25, {cfg['order']} order), "
                               f"FP64 {cfg['mode']}, Mach {cfg['mach']}, AoA {cfg['aoa']} deg, CFL {cfg['cfl']}",
                   "points": cloud.n, "mode": cfg["mode"], "precision": "fp64", "timing": timing},
        "roofline": roof,
        "cpu_baseline": cpu_out,
        "mixed": mixed_out,
        "tuning": f64["tuning"],
        "phases_ms": {k: round(v, 4) for k, v in f64["phases"].items()} if f64["phases"] else None,
        "point_classes": class_counts(f64["cls"], status),
        "step_algorithmic_gbs": round(step_gbs, 2),
        "step_hbm_frac": round(step_gbs / roofline.HBM_PEAK_GBS, 5),
    }


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
    prof = read_profile_values("flux_interior_c5_batch_traffic.txt")
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
                     "unit": "GB/s", "frac": round(achieved / roofline.HBM_PEAK_GBS, 5),
                     "traffic": prof.get("hbm_bytes_per_launch"),
                     "kernel": "k_flux_fused<0,true,double> (interior points on the fused path)",
                     "bytes_per_launch": flux_bytes, "launch_ms": round(f64["flux_ms"], 4)},
        "cpu_baseline": None if args.no_cpu else cpu_baseline_batch(clouds, plan),
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3_8m", choices=sorted(pc.CONFIGS) + ["c5_batch"])
    ap.add_argument("--block", type=int, default=4,
                    help="implicit runs: timed iterations per block, each block from the saved state")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the 2M explicit line of the default run")
    ap.add_argument("--cpu-sample-points", type=int, default=250000)
    ap.add_argument("--phase-iters", type=int, default=2)
    ap.add_argument("--cp-iters", type=int, default=0,
                    help="explicit runs: iteration at which FP64 and mixed wall Cp are compared "
                         "(default: the configuration's own count)")
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
    out = bench_config(args, args.config, args.steps, args.warmup, mixed=not args.no_mixed,
                       cpu=not args.no_cpu)
    if args.config == "c3_8m" and not args.no_extra:
        c2 = bench_config(args, "c2_2m", args.steps, 5, mixed=not args.no_mixed, cpu=False)
        out["c2_2m"] = {k: c2[k] for k in ("value", "ms_per_step", "roofline", "step_hbm_frac")}
        out["c2_2m"]["workload"] = c2["config"]["workload"]
        if c2["mixed"]:
            out["c2_2m"]["mixed"] = {k: c2["mixed"][k] for k in ("value", "speedup_vs_fp64",
                                                                 "wall_cp_err_vs_fp64", "CL_err", "CD_err")}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
