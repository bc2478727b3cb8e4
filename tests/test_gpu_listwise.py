"""The list-by-list flux kernel inside the fused order, FP64 and mixed.

A point whose conn plus the list entries that do not follow conn's order
exceed the 64 entries of a walk takes the list-by-list kernel
(k_flux_residual_lists) even in the fused flux order.  No generated cloud
has such a point, so the split lists of the explicit_blocks fixture are
handed over with three interior points' xpos, xneg and ypos lists replaced
by 22 points each that are not in their conn (as a host's own
preprocessing could produce): 66 extra walk entries.  Then:
  * those points run list by list (point_classes), in FP64 and in mixed;
  * FP64 fused order against the oracle on the same lists: residual history
    1e-10 relative, final primitive state 1e-9 per component;
  * FP64 fused order against the reference-order ("lists") path: 1e-12;
  * mixed against FP64: wall Cp within 1e-4 of the largest |Cp|.
"""

import sys

import numpy as np
import pytest

import orc
from conftest import PKG, load_fixture

pytestmark = pytest.mark.gpu
solver = PKG.solver
pc = PKG.pointcloud
pp = PKG.postprocess

ITERS = 6


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def crafted():
    cl, d = load_fixture("explicit_blocks")
    cnt, lists = solver.split_stencils(cl)
    cnt, lists = cnt.copy(), lists.copy()
    rng = np.random.default_rng(11)
    interior = np.nonzero(cl.status == pc.STATUS_INTERIOR)[0]
    picked = []
    for p in rng.permutation(interior):
        if len(picked) == 3:
            break
        conn = set(int(c) for c in cl.conn[p, :cl.nbhs[p]])
        pool = np.array([k + 1 for k in interior if k != p and (k + 1) not in conn], dtype=np.int32)
        for dr in (0, 1, 2):
            lists[dr, 22 * p:22 * p + 22] = rng.choice(pool, 22, replace=False)
            cnt[dr, p] = 22
        picked.append(int(p))
    return cl, d, (cnt, lists), picked


def per_component_err(a, b):
    return max(float(np.max(np.abs(a[:, c] - b[:, c])) / np.max(np.abs(b[:, c]))) for c in range(b.shape[1]))


def run(cl, d, split, precision="fp64", flux_order="fused"):
    s = solver.Solver(cl, float(d["mach"]), float(d["aoa"]), float(d["cfl"]), mode="explicit",
                      precision=precision, flux_order=flux_order, split=split)
    s.upload("prim", d["prim0"])
    classes = s.point_classes()
    rn, _ = s.run(1, ITERS)
    prim = s.download("prim")
    s.close()
    return classes, rn, prim


def test_list_by_list_points_in_the_fused_order():
    cl, d, split, picked = crafted()
    note(f"[gpu] list-by-list points {picked} in the fused order, FP64 and mixed")
    cls64, rn64, p64 = run(cl, d, split)
    clsm, rnm, pm = run(cl, d, split, precision="mixed")
    _, rnl, pl = run(cl, d, split, flux_order="lists")
    assert cls64["listwise"] >= 3 and clsm["listwise"] >= 3
    # the oracle on the same lists
    o = orc.Oracle(cl, float(d["mach"]), float(d["aoa"]), float(d["cfl"]), prim=d["prim0"], preprocess=False)
    cnt, lists = split
    for dr, nm in enumerate(["xpos", "xneg", "ypos", "yneg", "zpos", "zneg"]):
        getattr(o, nm + "_nbhs")[:] = cnt[dr]
        getattr(o, nm + "_conn")[:] = lists[dr][:len(getattr(o, nm + "_conn"))]
    o.alias[:], o.point_with_alias[:] = solver.alias_order(cl)
    rno, _ = o.run(1, ITERS, False)
    po = o.prim.reshape(-1, 5)
    nz = rno != 0
    err = float(np.max(np.abs(rn64[nz] - rno[nz]) / rno[nz]))
    perr = per_component_err(p64, po)
    note(f"      FP64 vs oracle: residual {err:.2e}, prim {perr:.2e}")
    assert err < 1e-10 and perr < 1e-9
    assert np.max(np.abs(rn64[nz] - rnl[nz]) / rnl[nz]) < 1e-12
    assert per_component_err(p64, pl) < 1e-12
    w = cl.status == pc.STATUS_WALL
    cp64 = pp.wall_cp(p64[w], float(d["mach"]))
    cpm = pp.wall_cp(pm[w], float(d["mach"]))
    cp_err = float(np.max(np.abs(cpm - cp64)) / np.max(np.abs(cp64)))
    note(f"      mixed vs FP64: wall Cp {cp_err:.2e}")
    assert cp_err < 1e-4 and np.all(np.isfinite(pm))
