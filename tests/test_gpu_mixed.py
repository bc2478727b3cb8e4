"""The mixed-precision form against FP64.

Mixed: q, qm, dq, the gradient right-hand sides and the split fluxes in
FP32; geometry, the 3x3 solves, Cramer weights, flux sums, D_inv, state
update and residual in FP64.  Its bar is the wall pressure coefficient
within 1e-4 (relative to the largest |Cp| on the wall) of the FP64 result,
and a residual history that follows the FP64 one.
"""

import sys

import numpy as np
import pytest

from conftest import GOLDEN, PKG, load_fixture

pytestmark = pytest.mark.gpu
solver = PKG.solver
pc = PKG.pointcloud
pp = PKG.postprocess


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(cl, mach, aoa, cfl, iters, mode="explicit", precision="fp64", prim0=None):
    s = solver.Solver(cl, mach, aoa, cfl, mode=mode, precision=precision)
    s.upload("prim", prim0 if prim0 is not None else np.tile(pc.freestream_prim(mach, aoa), (cl.n, 1)))
    rn, _ = s.run(1, iters)
    prim = s.download("prim")
    s.close()
    return rn, prim


def cp_err(prim_a, prim_b, cl, mach):
    w = cl.status == pc.STATUS_WALL
    a, b = pp.wall_cp(prim_a[w], mach), pp.wall_cp(prim_b[w], mach)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_mixed_entry_points_follow_fp64():
    """One iteration's entry points from the reference's own inputs."""
    note("[gpu] mixed entry points, fixture explicit_morton")
    cl, d = load_fixture("explicit_morton")
    s = solver.Solver(cl, float(d["mach"]), float(d["aoa"]), float(d["cfl"]), precision="mixed")
    s.upload("prim", d["start_prim"])
    s.compute_conserved_vector()
    s.eval_q_variables()
    q = s.download("q")
    assert np.max(np.abs(q - d["step_q"]) / np.max(np.abs(d["step_q"]), axis=0)) < 1e-6
    s.eval_q_derivatives()
    qm = s.download("qm")
    assert np.max(np.abs(qm - d["step_qm"])) / np.max(np.abs(d["step_qm"])) < 1e-6
    s.upload("q", d["step_q"])
    s.upload("qm", d["step_qm"])
    s.upload("dq", d["step_dq"])
    s.eval_flux_residual()
    fr = s.download("flux_res")
    ref = d["step_flux_res"]
    err = np.max(np.abs(fr - ref), axis=0) / np.max(np.abs(ref), axis=0)
    note(f"      flux residual rel err per component {err}")
    assert np.all(err < 1e-3)
    s.close()


@pytest.mark.parametrize("mode,cfl,iters", [("explicit", 0.2, 60), ("implicit", 2.0, 30)])
def test_mixed_follows_fp64_on_stretched_cloud(mode, cfl, iters):
    note(f"[gpu] mixed vs FP64, 20k shell cloud, {mode}")
    cl = pc.ellipsoid_cloud(1600, 800, 17600, seed=21, order="morton")
    mach, aoa = 0.84, 3.06
    r64, p64 = run(cl, mach, aoa, cfl, iters, mode)
    r32, p32 = run(cl, mach, aoa, cfl, iters, mode, "mixed")
    # from iteration 2: the first residual from the freestream is zero or
    # round-off in both forms (round-off of FP32 fluxes in the mixed one)
    nz = r64 != 0
    nz[0] = False
    herr = float(np.max(np.abs(r32[nz] - r64[nz]) / r64[nz]))
    e = cp_err(p32, p64, cl, mach)
    note(f"      residual history max rel diff {herr:.2e}, wall Cp rel err {e:.2e}")
    assert herr < 1e-2
    assert e < 1e-4


def test_mixed_wall_cp_on_the_50k_history():
    """1000 iterations of the 50k configuration: the mixed wall Cp against
    the reference Fortran's own FP64 final wall state."""
    d = np.load(f"{GOLDEN}/c1_50k_history.npz", allow_pickle=False)
    cl = pc.config_cloud("c1_50k")
    mach = float(d["mach"])
    note(f"[gpu] mixed, c1_50k, {int(d['iterations'])} iterations vs the reference's wall state")
    rn, prim = run(cl, mach, float(d["aoa"]), float(d["cfl"]), int(d["iterations"]), precision="mixed")
    wi = d["wall_index"]
    cp, cp_ref = pp.wall_cp(prim[wi], mach), pp.wall_cp(d["wall_prim"], mach)
    e = float(np.max(np.abs(cp - cp_ref)) / np.max(np.abs(cp_ref)))
    ref = d["hist"][:, 0]
    nz = ref != 0
    herr = float(np.max(np.abs(rn[nz] - ref[nz]) / ref[nz]))
    note(f"      wall Cp rel err {e:.2e}, residual history max rel diff {herr:.2e}")
    assert e < 1e-4
