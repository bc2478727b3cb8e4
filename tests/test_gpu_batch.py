"""Batches of clouds and the graph-captured iteration.

A batch lays several clouds out one after the other and iterates them with
the same launches; nothing couples the clouds, and every point's arithmetic
is the one it gets alone, so each cloud's state must equal, bit for bit, a
run of that cloud by itself with its own Mach and angle of attack.  The
graph path replays one captured iteration; every replay is checked against
the eagerly launched iterations.
"""

import sys

import numpy as np
import pytest

from conftest import PKG, load_fixture

pytestmark = pytest.mark.gpu
solver = PKG.solver
pc = PKG.pointcloud


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def clouds():
    a, _ = load_fixture("explicit_blocks")
    b, _ = load_fixture("explicit_morton")
    c = pc.ellipsoid_cloud(800, 400, 8800, seed=31, axes=(0.5, 0.1, 0.05), order="morton")
    d = pc.ellipsoid_cloud(1200, 600, 13200, seed=32, axes=(0.5, 0.25, 0.075), order="blocks")
    return [a, b, c, d], [0.6, 0.84, 0.5, 0.75], [2.0, 3.06, -4.0, 8.0]


def prim0(cl, mach, aoa):
    return np.tile(pc.freestream_prim(mach, aoa), (cl.n, 1))


@pytest.mark.parametrize("mode,cfl,iters", [("explicit", 0.2, 25), ("implicit", 1.0, 12)])
def test_batch_equals_separate_runs(mode, cfl, iters):
    cls, machs, aoas = clouds()
    note(f"[gpu] batch of {len(cls)} clouds ({[c.n for c in cls]}), {mode}, graph replay")
    bs = solver.BatchSolver(cls, machs, aoas, cfl, mode=mode)
    k, counts = bs.batch_info()
    assert k == len(cls) and list(counts) == [c.n for c in cls]
    bs.upload("prim", np.concatenate([prim0(c, m, a) for c, m, a in zip(cls, machs, aoas)]))
    _, inst = bs.run_graph(1, iters)
    prims = bs.split(bs.download("prim"))
    bs.close()
    for b, (cl, m, a) in enumerate(zip(cls, machs, aoas)):
        s = solver.Solver(cl, m, a, cfl, mode=mode)
        s.upload("prim", prim0(cl, m, a))
        rn, _ = s.run(1, iters)
        p = s.download("prim")
        s.close()
        assert np.array_equal(prims[b], p), f"cloud {b}"
        nz = rn != 0
        assert np.max(np.abs(inst[nz, b] - rn[nz]) / rn[nz]) < 1e-13, f"cloud {b}"


def test_graph_replays_match_eager_iterations():
    cl, d = load_fixture("implicit_morton")
    mach, aoa, cfl = float(d["mach"]), float(d["aoa"]), float(d["cfl"])
    note("[gpu] graph replays vs eager iterations, implicit fixture")
    eager = solver.Solver(cl, mach, aoa, cfl, mode="implicit")
    eager.upload("prim", d["prim0"])
    rn_e, _ = eager.run(1, 20)
    pe = eager.download("prim")
    eager.close()
    g = solver.Solver(cl, mach, aoa, cfl, mode="implicit")
    g.upload("prim", d["prim0"])
    rn1, inst1 = g.run_graph(1, 8)
    mid = g.download("prim")          # eager work and a synchronisation between replays
    assert np.all(np.isfinite(mid))
    rn2, inst2 = g.run_graph(9, 12)
    pg = g.download("prim")
    g.close()
    assert np.array_equal(np.concatenate([rn1, rn2]), rn_e)
    assert np.array_equal(pg, pe)
    inst = np.concatenate([inst1[:, 0], inst2[:, 0]])
    nz = rn_e != 0
    assert np.max(np.abs(inst[nz] - rn_e[nz]) / rn_e[nz]) < 1e-13


def test_autotuned_shapes_leave_results_unchanged():
    """The tuner picks block sizes, register caps and the sweep grid; none
    changes any point's arithmetic, so tuned and default runs agree bit for
    bit (implicit, 20k shell cloud)."""
    note("[gpu] autotune: tuned vs default launch shapes")
    cl = pc.ellipsoid_cloud(1600, 800, 17600, seed=21, order="morton")
    mach, aoa, cfl = 0.84, 3.06, 1.0
    p0 = prim0(cl, mach, aoa)
    out = []
    for tune in (False, True):
        s = solver.Solver(cl, mach, aoa, cfl, mode="implicit")
        s.upload("prim", p0)
        if tune:
            t = s.autotune(2)
            note(f"      tuning {t}")
            assert t["flux_interior_block"] in (64, 128, 256)
            assert t["flux_interior_ms"] <= t["flux_interior_default_ms"]
        rn, _ = s.run(1, 6)
        out.append((rn, s.download("prim")))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])



def test_mixed_flux_variants_agree():
    """Every compiled variant of the mixed form's fused flux kernels (block
    64/128/256 x register caps of 1..6 waves per SIMD; the 5- and 6-wave caps
    spill a few registers) does the same FP32 arithmetic per point, so runs
    under each shape agree bit for bit (explicit, 20k shell cloud)."""
    note("[gpu] mixed flux kernel variants: every shape vs the default")
    cl = pc.ellipsoid_cloud(1600, 800, 17600, seed=21, order="morton")
    mach, aoa, cfl = 0.84, 3.06, 0.5
    p0 = prim0(cl, mach, aoa)
    ref = None
    for block, waves in [(0, 0), (64, 1), (128, 3), (256, 4), (128, 5), (256, 6), (64, 6)]:
        s = solver.Solver(cl, mach, aoa, cfl, mode="explicit", precision="mixed")
        s.upload("prim", p0)
        s.set_tuning({"flux_interior_block": block, "flux_interior_waves": waves,
                      "flux_wall_block": block, "flux_wall_waves": waves,
                      "qderiv_block": 0, "qderiv_waves": 0, "sweep_div": 0})
        rn, _ = s.run(1, 8)
        got = (rn, s.download("prim"))
        s.close()
        if ref is None:
            ref = got
            continue
        note(f"      block {block} waves {waves}")
        assert np.array_equal(got[0], ref[0])
        assert np.array_equal(got[1], ref[1])
