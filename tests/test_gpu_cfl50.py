"""HIP implicit path at CFL 50 against the reference's own run.

tests/golden/cfl50_20k_history.npz is the reference Fortran's run of the
canonical implicit pairing at Mach 0.84, AoA 3.06 deg, CFL 50 on a 20k-point
shell cloud (far field at radius 15): residual history and far-field
pressure trace of every iteration up to its first non-finite one (170).

The run amplifies round-off: the oracle started from the same state times
(1 + 1e-15 noise) tracks the reference's residual to 1.3e-11 at iteration
90, 3.7e-10 at 100, 6e-9 at 120, 4e-7 at 130 and 4e-3 at 150, and turns
non-finite at iteration 166 instead of 170.  The device's exp/log/erf
differ from glibc's by ulps, a perturbation of the same size injected every
iteration.  So the FP64 path is held to the fixture's 1e-10 bar through
iteration 80, to the perturbed oracle's envelope after that, and its first
non-finite iteration to the window the perturbation opens (160..175).
Both sweep schedules run.  The mixed form (FP32 fluxes) is compared with the
FP64 path at iteration 30, before the far field grows: wall Cp, CL/CD and
the residual drop.
"""

import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, PKG, check_zero_residuals

pytestmark = pytest.mark.gpu
pc = PKG.pointcloud
solver = PKG.solver
postprocess = PKG.postprocess


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup():
    d = dict(np.load(os.path.join(GOLDEN, "cfl50_20k_history.npz"), allow_pickle=False))
    n = int(d["n"])
    n_wall, n_outer = int(0.08 * n), int(0.04 * n)
    cl = pc.ellipsoid_cloud(n_wall, n_outer, n - n_wall - n_outer, seed=int(d["seed"]), order="morton")
    prim0 = np.tile(pc.freestream_prim(float(d["mach"]), float(d["aoa"])), (cl.n, 1))
    return cl, d, prim0


# relative residual envelope of a 1e-15 start perturbation (oracle), with a
# factor 30 of margin: (last iteration, bound)
ENVELOPE = [(80, 1e-10), (100, 1e-8), (120, 2e-7), (130, 1e-5), (140, 1e-3)]


@pytest.mark.parametrize("sweep", ["async", "levels"])
def test_cfl50_history_against_reference(sweep):
    note(f"[gpu] CFL 50, 20k shell cloud, FP64 implicit, {sweep} sweeps")
    cl, d, prim0 = _setup()
    mach, aoa, cfl = float(d["mach"]), float(d["aoa"]), float(d["cfl"])
    ref = d["hist"][:, 0]
    first_ref = int(d["first_nonfinite"])
    s = solver.Solver(cl, mach, aoa, cfl, mode="implicit", sweep_schedule=sweep)
    s.upload("prim", prim0)
    rn, _ = s.run(1, 160)
    assert ref[0] == 0.0
    check_zero_residuals(rn[:1], float(prim0[0, 0]), cl.n, True)
    lo = 1
    for last, bound in ENVELOPE:
        err = np.max(np.abs(rn[lo:last] - ref[lo:last]) / np.abs(ref[lo:last]))
        note(f"      iterations {lo + 1}..{last}: max rel err {err:.2e} (bound {bound:.0e})")
        assert err < bound, (last, err)
        lo = last
    prim = s.download("prim")
    assert np.all(np.isfinite(rn)) and np.all(np.isfinite(prim))
    outer = cl.status == pc.STATUS_OUTER
    dev = np.max(np.abs(prim[outer, 4] - 1.0 / 1.4))
    note(f"      far-field |p - p_inf| at 160: {dev:.4f} (reference {d['trace'][159, 0]:.4f})")
    assert dev > 0.05   # the far field has grown, as in the reference
    # the first non-finite iteration, within the window round-off opens
    first = 0
    for t in range(161, 176):
        r, _ = s.run(t, 1)
        if not np.isfinite(r[0]):
            first = t
            break
    note(f"      first non-finite iteration {first} (reference {first_ref})")
    assert 160 < first <= 175
    s.close()


def test_cfl50_mixed_against_fp64():
    """C4's comparison on the CFL-50 run: mixed against FP64 at iteration 30."""
    note("[gpu] CFL 50, 20k shell cloud, mixed vs FP64 at iteration 30")
    cl, d, prim0 = _setup()
    mach, aoa, cfl = float(d["mach"]), float(d["aoa"]), float(d["cfl"])
    out = {}
    for precision in ("fp64", "mixed"):
        s = solver.Solver(cl, mach, aoa, cfl, mode="implicit", precision=precision)
        s.upload("prim", prim0)
        rn, _ = s.run(1, 30)
        out[precision] = (rn, s.download("prim"))
        s.close()
    rn64, p64 = out["fp64"]
    rnm, pm = out["mixed"]
    ref = d["hist"][:30, 0]
    assert np.max(np.abs(rn64[1:] - ref[1:]) / ref[1:]) < 1e-10
    wall = cl.status == pc.STATUS_WALL
    cp64 = postprocess.wall_cp(p64[wall], mach)
    cpm = postprocess.wall_cp(pm[wall], mach)
    cp_err = float(np.max(np.abs(cpm - cp64)))
    areas = postprocess.wall_areas(cl)
    f64 = postprocess.force_coefficients(cl, p64, mach, aoa, areas=areas)
    fm = postprocess.force_coefficients(cl, pm, mach, aoa, areas=areas)
    cl64, cd64, clm, cdm = f64["CL"], f64["CD"], fm["CL"], fm["CD"]
    drop64 = rn64[-1] / rn64[1]
    dropm = rnm[-1] / rnm[1]
    note(f"      wall Cp err {cp_err:.2e}, CL {cl64:.6e} / {clm:.6e}, CD {cd64:.6e} / {cdm:.6e}, "
         f"residual drop {drop64:.6e} / {dropm:.6e}")
    assert cp_err < 1e-4
    assert abs(clm - cl64) < 1e-4 and abs(cdm - cd64) < 1e-4
    assert abs(dropm - drop64) / drop64 < 1e-3
