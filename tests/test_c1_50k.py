"""The 50k-point explicit configuration against the reference's own history.

tests/golden/c1_50k_history.npz holds the residual history of 1000
iterations of the reference Fortran (flang -O2, one core) on the 50k-point
ellipsoid cloud (4000 wall, 2000 outer, 44000 interior; Mach 0.6, AoA 2,
CFL 0.2) and the final primitive state of its wall points.  The cloud itself
is not stored: the generator is deterministic, and its digest is checked
before anything is compared.
"""

import numpy as np
import pytest

import orc
from conftest import GOLDEN, PKG

pc = PKG.pointcloud
FILE = f"{GOLDEN}/c1_50k_history.npz"


def _digest(cl):
    import hashlib

    h = hashlib.sha256()
    for a in (cl.x, cl.y, cl.z, cl.tan1, cl.tan2, cl.nor, cl.status, cl.nbhs, cl.conn, cl.min_dist):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def c1():
    d = np.load(FILE, allow_pickle=False)
    cl = pc.config_cloud("c1_50k")
    assert _digest(cl) == str(d["digest"]), "generator no longer reproduces the 50k cloud"
    return cl, d


def wall_cp(prim, mach):
    # pressure coefficient against the freestream p = 1/1.4, rho = 1
    return (prim[:, 4] - 1.0 / 1.4) / (0.5 * mach * mach)


def test_oracle_follows_reference_history(c1):
    """First 12 iterations of the C oracle, bit for bit."""
    cl, d = c1
    prim0 = np.tile(pc.freestream_prim(float(d["mach"]), float(d["aoa"])), (cl.n, 1))
    o = orc.Oracle(cl, float(d["mach"]), float(d["aoa"]), float(d["cfl"]), prim=prim0)
    rn, _ = o.run(1, 12, False)
    assert np.array_equal(rn, d["hist"][:12, 0])


@pytest.mark.gpu
def test_gpu_1000_iterations_match_reference(c1):
    cl, d = c1
    mach = float(d["mach"])
    print(f"[gpu] c1_50k: {cl.n} points, 1000 iterations vs reference history", flush=True)
    s = PKG.solver.Solver(cl, mach, float(d["aoa"]), float(d["cfl"]))
    s.upload("prim", np.tile(pc.freestream_prim(mach, float(d["aoa"])), (cl.n, 1)))
    iters = int(d["iterations"])
    rn, _ = s.run(1, iters)
    prim = s.download("prim")
    s.close()
    ref = d["hist"][:, 0]
    nz = ref != 0.0
    assert np.all(rn[~nz] == 0.0)
    err = np.max(np.abs(rn[nz] - ref[nz]) / np.abs(ref[nz]))
    print(f"      residual history max rel err {err:.3e}")
    assert err < 1e-10
    wi = d["wall_index"]
    assert np.array_equal(wi, np.flatnonzero(cl.status == pc.STATUS_WALL))
    cp, cp_ref = wall_cp(prim[wi], mach), wall_cp(d["wall_prim"], mach)
    cerr = np.max(np.abs(cp - cp_ref)) / np.max(np.abs(cp_ref))
    print(f"      wall Cp max rel err {cerr:.3e}")
    assert cerr < 1e-9
    scale = np.max(np.abs(d["wall_prim"]), axis=0)
    assert np.max(np.abs(prim[wi] - d["wall_prim"]) / scale) < 1e-9
