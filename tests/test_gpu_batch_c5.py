"""The 64-cloud batch workload (pointcloud.batch_plan) on a subset that
spans its sizes: the largest cloud and the smallest eight.

  * FP64 and mixed batches, run through the captured graph, against each
    cloud run by itself in the same precision: bit for bit;
  * the two smallest clouds against the oracle after 3 iterations (residual
    history 1e-10 relative, primitive state 1e-9 per component);
  * mixed against FP64: wall Cp within 1e-4 of the largest |Cp| per cloud.
"""

import sys

import numpy as np
import pytest

import orc
from conftest import PKG

pytestmark = pytest.mark.gpu
solver = PKG.solver
pc = PKG.pointcloud
pp = PKG.postprocess

ITERS, ORACLE_ITERS = 10, 3


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def per_component_err(a, b):
    return max(float(np.max(np.abs(a[:, c] - b[:, c])) / max(np.max(np.abs(b[:, c])), 1e-300))
               for c in range(b.shape[1]))


def subset():
    plan = pc.batch_plan()
    order = np.argsort([p["n_wall"] + p["n_outer"] + p["n_interior"] for p in plan])
    pick = [int(order[-1])] + [int(b) for b in order[:8]]
    sub = [plan[b] for b in pick]
    return pick, sub, pc.batch_clouds(sub, workers=16)


def freestream(cls, sub):
    return [np.tile(pc.freestream_prim(p["mach"], p["aoa"]), (c.n, 1)) for c, p in zip(cls, sub)]


def test_batch_workload_subset():
    pick, sub, cls = subset()
    cfl = pc.BATCH["cfl"]
    machs = [p["mach"] for p in sub]
    aoas = [p["aoa"] for p in sub]
    p0 = freestream(cls, sub)
    note(f"[gpu] batch workload: clouds {pick}, sizes {[c.n for c in cls]}")
    final = {}
    for precision in ("fp64", "mixed"):
        bs = solver.BatchSolver(cls, machs, aoas, cfl, mode=pc.BATCH["mode"], precision=precision)
        bs.upload("prim", np.concatenate(p0))
        rn_a, inst_a = bs.run_graph(1, ORACLE_ITERS)
        early = bs.split(bs.download("prim"))
        rn_b, inst_b = bs.run_graph(ORACLE_ITERS + 1, ITERS - ORACLE_ITERS)
        final[precision] = bs.split(bs.download("prim"))
        inst = np.concatenate([inst_a, inst_b])
        bs.close()
        for b, cl in enumerate(cls):
            s = solver.Solver(cl, machs[b], aoas[b], cfl, mode=pc.BATCH["mode"], precision=precision)
            s.upload("prim", p0[b])
            rn, _ = s.run(1, ITERS)
            p = s.download("prim")
            s.close()
            assert np.array_equal(final[precision][b], p), (precision, pick[b])
            nz = rn != 0
            assert np.max(np.abs(inst[nz, b] - rn[nz]) / rn[nz]) < 1e-13, (precision, pick[b])
        note(f"      {precision}: every cloud equals its own run")
        if precision == "fp64":
            # the two smallest clouds against the oracle
            for b in (1, 2):
                o = orc.Oracle(cls[b], machs[b], aoas[b], cfl, prim=p0[b])
                rn_o, _ = o.run(1, ORACLE_ITERS, False)
                nz = rn_o != 0
                err = np.max(np.abs(inst_a[nz, b] - rn_o[nz]) / rn_o[nz])
                perr = per_component_err(early[b], o.prim.reshape(-1, 5))
                note(f"      cloud {pick[b]} vs oracle: residual {err:.2e}, prim {perr:.2e}")
                assert err < 1e-10 and perr < 1e-9
    for b, cl in enumerate(cls):
        w = cl.status == pc.STATUS_WALL
        cp64 = pp.wall_cp(final["fp64"][b][w], machs[b])
        cpm = pp.wall_cp(final["mixed"][b][w], machs[b])
        err = float(np.max(np.abs(cpm - cp64)) / np.max(np.abs(cp64)))
        assert err < 1e-4, (pick[b], err)
    note("      mixed wall Cp within 1e-4 of FP64 on every cloud")
