"""Inputs the library refuses with MFCFD_EINVAL, through the C ABI.

Where the reference divides by a zero determinant (a coplanar or collinear
least-squares stencil, an empty direction list) or carries a NaN of prim
into every stencil (tests/test_failure_modes.py shows both on the oracle),
the library returns a clean error that names the point, and leaves the
device state as it was.
"""

import sys

import numpy as np
import pytest

from conftest import PKG, degenerate_cloud, load_fixture

pytestmark = pytest.mark.gpu
solver = PKG.solver
pc = PKG.pointcloud


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_bind_refuses_an_empty_direction_list():
    note("[gpu] bind of a cloud with an empty xpos list")
    cl, p, split = degenerate_cloud("empty")
    with pytest.raises(solver.MfcfdError, match=f"point {p + 1}: list 0, which its flux walks, is empty"):
        solver.Solver(cl, 0.6, 2.0, 0.2, split=split)


def test_bind_refuses_a_singular_stencil():
    """Moving the xpos neighbours of one point onto its tan1 axis makes that
    list's matrix rank one; the stencils of the moved points and of their
    own neighbours change too, so the point named is the first singular one
    in device order, which need not be that point."""
    note("[gpu] bind of a cloud with a collinear xpos stencil")
    cl, p, _ = degenerate_cloud("collinear")
    with pytest.raises(solver.MfcfdError, match=r"point \d+: the least-squares matrix of its \w+ stencil is singular"):
        solver.Solver(cl, 0.6, 2.0, 0.2)


def test_upload_refuses_a_state_the_iteration_cannot_start_from():
    note("[gpu] upload of prim with NaN, negative density, zero pressure")
    cl, d = load_fixture("explicit_blocks")
    s = solver.Solver(cl, float(d["mach"]), float(d["aoa"]), float(d["cfl"]))
    prim0 = np.array(d["prim0"], dtype=np.float64)
    s.upload("prim", prim0)
    before = s.download("prim")
    for r, c, v in [(7, 2, np.nan), (11, 0, -1.0), (12, 4, 0.0), (13, 1, np.inf)]:
        bad = prim0.copy()
        bad[r, c] = v
        with pytest.raises(solver.MfcfdError, match=f"prim of point {r + 1} "):
            s.upload("prim", bad)
        assert np.array_equal(s.download("prim"), before)
    rn, _ = s.run(1, 3)
    assert np.all(np.isfinite(rn))
    s.close()
