"""The 2M-point explicit configuration (Mach 0.84, AoA 3.06 deg, CFL 0.2):
500 iterations on the GPU, then iteration 501 on the GPU and in the oracle
from the GPU's state after 500, compared: res_new within 1e-10 relative,
primitive state within 1e-9 of each component's largest magnitude.  One
oracle iteration of 2M points takes about 40 s on one core.
"""

import sys
import time

import numpy as np
import pytest

import orc
from conftest import PKG

pytestmark = pytest.mark.gpu
solver = PKG.solver
pc = PKG.pointcloud


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_step_501_of_the_2m_explicit_run_against_the_oracle():
    c = pc.CONFIGS["c2_2m"]
    t0 = time.time()
    cl = pc.config_cloud("c2_2m", workers=16)
    note(f"[gpu] 2M explicit: cloud of {cl.n} points in {time.time() - t0:.0f} s")
    s = solver.Solver(cl, c["mach"], c["aoa"], c["cfl"], mode="explicit")
    s.upload("prim", np.tile(pc.freestream_prim(c["mach"], c["aoa"]), (cl.n, 1)))
    rn, _ = s.run(1, 500)
    assert np.all(np.isfinite(rn))
    prim500 = s.download("prim")
    res501, _ = s.fpi_solver(501)
    prim501 = s.download("prim")
    s.close()
    note(f"      GPU: res_new(500) {rn[-1]:.6e}, res_new(501) {res501:.6e}")
    t0 = time.time()
    o = orc.Oracle(cl, c["mach"], c["aoa"], c["cfl"], prim=prim500)
    o.fpi_solver(501, False)
    note(f"      oracle step 501 in {time.time() - t0:.0f} s: res_new {o.res_new:.6e}")
    err = abs(res501 - o.res_new) / o.res_new
    po = o.prim.reshape(-1, 5)
    perr = max(float(np.max(np.abs(prim501[:, r] - po[:, r])) / np.max(np.abs(po[:, r]))) for r in range(5))
    note(f"      res_new rel err {err:.2e}, prim err {perr:.2e}")
    assert err < 1e-10 and perr < 1e-9
