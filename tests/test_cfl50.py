"""The canonical implicit form at CFL 50, pinned on the reference itself.

tests/golden/cfl50_20k_history.npz holds the reference Fortran's own run
(built by flang, tests/golden/make_golden.py --cfl50) of the implicit LU-SGS
pairing (implicit_forward_and_reverse_sweeps.f90 + state_update.f90) at
Mach 0.84, AoA 3.06 deg, CFL 50 on a 20k-point shell cloud with its far
field at radius 15: the residual history and, per iteration, the largest
far-field pressure deviation |p - pr_inf| over the outer points and the
number of non-finite pressures, up to the first non-finite iteration.

The outer points are updated explicitly, U = Ubar - delt*R
(state_update.f90:51-60), with the implicit run's CFL-50 local step, and the
sweeps never touch them (implicit_forward_and_reverse_sweeps.f90:30-52).  In
the reference's own run that update amplifies the far-field disturbance
until the pressures turn non-finite.  These tests pin that on the reference's
data and check the oracle against it bit for bit on its first iterations
(the whole run is 170 one-core iterations, too long for the CPU suite).
"""

import os

import numpy as np
import pytest

import orc
from conftest import GOLDEN, PKG

pc = PKG.pointcloud

PATH = os.path.join(GOLDEN, "cfl50_20k_history.npz")


def _fixture():
    return dict(np.load(PATH, allow_pickle=False))


def cfl50_cloud(d):
    n = int(d["n"])
    n_wall, n_outer = int(0.08 * n), int(0.04 * n)
    return pc.ellipsoid_cloud(n_wall, n_outer, n - n_wall - n_outer, seed=int(d["seed"]), order="morton")


def test_reference_diverges_at_cfl50_from_the_far_field():
    d = _fixture()
    hist, tr = d["hist"][:, 0], d["trace"]
    first = int(d["first_nonfinite"])
    # finite for 169 iterations, non-finite from the 170th
    assert first == 170 and len(hist) == first
    assert np.all(np.isfinite(hist[:first - 1])) and not np.isfinite(hist[first - 1])
    assert np.all(tr[:first - 1, 2] == 0) and tr[first - 1, 2] > 0
    # the far-field pressure deviation: small and bounded for the first 120
    # iterations, then growing geometrically (x1.2 - x1.3 per iteration)
    # until it is O(1) just before the pressures break down
    dev = tr[:, 0]
    assert np.max(dev[:120]) < 2e-4
    growth = dev[126:150] / dev[125:149]
    assert np.all(growth > 1.15) and np.all(growth < 1.4)
    assert dev[first - 2] > 0.3
    # the density residual stays small while the far field grows: the
    # residual sums wall and interior points only (state_update.f90:39-44,
    # 69-74), and the outer points carry the instability
    assert np.max(hist[100:first - 5]) < 1e-5


def test_oracle_matches_reference_at_cfl50():
    d = _fixture()
    cl = cfl50_cloud(d)
    import importlib, sys
    sys.path.insert(0, os.path.join(GOLDEN))
    mg = importlib.import_module("make_golden")
    assert mg.cloud_digest(cl) == str(d["digest"]), "generator output changed"
    prim0 = np.tile(pc.freestream_prim(float(d["mach"]), float(d["aoa"])), (cl.n, 1))
    o = orc.Oracle(cl, float(d["mach"]), float(d["aoa"]), float(d["cfl"]), prim=prim0)
    iters = 12
    outer = cl.status == pc.STATUS_OUTER
    for t in range(1, iters + 1):
        o.fpi_solver(t, True)
        assert o.res_new == d["hist"][t - 1, 0], t
        dev = np.max(np.abs(o.prim[outer, 4] - 1.0 / 1.4))
        assert dev == d["trace"][t - 1, 0], t
