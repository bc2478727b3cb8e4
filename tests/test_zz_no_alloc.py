"""mfcfd_run_noalloc (no counterpart in the reference) against mfcfd_run.

The no-allocation run keeps its per-iteration residual history in a fixed
device array shared by every call on the device.  Two calls in a row -- a
20k-point cloud in the implicit form for 12 iterations, then a smaller
15k-point cloud in the explicit form for 5 -- must give exactly mfcfd_run's
residual history, residue history and final state on fresh contexts with
the same inputs (the same kernels run in the same order: bit for bit).  The
entry point also refuses runs longer than its array and an implicit run
whose sweep plan is not built yet, before it launches anything.
"""

import sys

import numpy as np
import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu
solver = PKG.solver
pc = PKG.pointcloud

MACH, AOA = 0.84, 3.06


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fresh(cl, mode, cfl):
    s = solver.Solver(cl, MACH, AOA, cfl, mode=mode)
    s.upload("prim", np.tile(pc.freestream_prim(MACH, AOA), (cl.n, 1)))
    return s


def compare(cl, mode, cfl, iters):
    ref = fresh(cl, mode, cfl)
    rn_ref, rs_ref = ref.run(1, iters)
    prim_ref = ref.download("prim")
    ref.close()
    s = fresh(cl, mode, cfl)
    if mode != "explicit":
        s.sweep_info()   # builds the sweep plan, which allocates
    rn, rs = s.run_noalloc(1, iters)
    prim = s.download("prim")
    res = s.residual()
    s.close()
    assert np.array_equal(rn, rn_ref)
    assert np.array_equal(rs, rs_ref)
    assert np.array_equal(prim, prim_ref)
    assert res["res_new"] == rn_ref[-1] and res["iteration"] == iters
    note(f"      {cl.n} points, {mode}, {iters} iterations: res_new {rn[-1]:.6e}, identical to mfcfd_run")


def test_noalloc_run_matches_run_on_two_calls():
    note("[gpu] mfcfd_run_noalloc: 20k implicit x 12, then 15k explicit x 5")
    big = pc.ellipsoid_cloud(1600, 800, 17600, seed=21, order="morton")
    small = pc.ellipsoid_cloud(1200, 600, 13200, seed=21, order="morton")
    compare(big, "implicit", 2.0, 12)
    compare(small, "explicit", 0.2, 5)


def test_noalloc_run_rejects_what_it_cannot_do_without_allocating():
    note("[gpu] mfcfd_run_noalloc: refused calls")
    cl = pc.ellipsoid_cloud(1200, 600, 13200, seed=21, order="morton")
    s = fresh(cl, "implicit", 2.0)
    with pytest.raises(solver.MfcfdError, match="sweep plan"):
        s.run_noalloc(1, 2)
    s.sweep_info()
    with pytest.raises(solver.MfcfdError, match="exceeds"):
        s.run_noalloc(1, solver.NOALLOC_MAX_ITERS + 1)
    rn, _ = s.run_noalloc(1, 2)
    assert np.all(np.isfinite(rn))
    s.close()
