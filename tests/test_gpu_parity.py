"""HIP path against the reference's golden fixtures and the C oracle.

Tolerances (the device exp/log/erf are ocml's, the reference's glibc's, so
results agree to a few ulps per operation, not bit for bit):
  * one entry point from the reference's own inputs: 1e-11 relative to the
    array's largest magnitude;
  * residual history: 1e-10 relative per iteration (absolute 1e-300 where
    the reference residual is exactly 0);
  * final conserved/primitive state: 1e-9 relative to each component's
    largest magnitude.
"""

import sys

import numpy as np
import pytest

import orc
from conftest import FIXTURE_NAMES, PKG, check_zero_residuals, load_fixture

pytestmark = pytest.mark.gpu
solver = PKG.solver
pc = PKG.pointcloud


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b)) if b.size else 1.0
    if scale == 0.0:
        return float(np.max(np.abs(a - b))) if a.size else 0.0
    return float(np.max(np.abs(a - b)) / scale)


def per_component_err(a, b):
    a = a.reshape(a.shape[0], -1)
    b = b.reshape(b.shape[0], -1)
    return max(rel_err(a[:, c], b[:, c]) for c in range(a.shape[1]))


def make_solver(name, **kw):
    cl, d = load_fixture(name)
    s = solver.Solver(cl, float(d["mach"]), float(d["aoa"]), float(d["cfl"]),
                      mode=str(d["mode"]), **kw)
    s.upload("prim", d["prim0"])
    return cl, d, s


ORDERS = ["fused", "lists"]
# the reference-order flux kernel runs with the level-set sweep schedule,
# the fused kernel with the sync-free one
SWEEP_OF = {"fused": "async", "lists": "levels"}


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("name", FIXTURE_NAMES)
def test_entry_points_match_reference(name, order):
    note(f"[gpu] entry points, fixture {name}, flux order {order}")
    cl, d, s = make_solver(name, flux_order=order, sweep_schedule=SWEEP_OF[order])
    s.upload("prim", d["start_prim"])
    s.upload("delUp", d["start_delUp"])
    s.upload("delUn", d["start_delUn"])
    s.compute_conserved_vector()
    assert per_component_err(s.download("U"), d["step_U"]) < 1e-13
    s.eval_q_variables()
    assert per_component_err(s.download("q"), d["step_q"]) < 1e-13
    s.eval_q_derivatives()
    assert per_component_err(s.download("qm"), d["step_qm"]) < 1e-13
    assert per_component_err(s.download("dq"), d["step_dq"]) < 1e-11
    s.timestep_delt()
    assert rel_err(s.download("delt"), d["step_delt"]) < 1e-13
    # flux residual from the reference's own q, qm, dq
    s.upload("q", d["step_q"])
    s.upload("qm", d["step_qm"])
    s.upload("dq", d["step_dq"])
    s.upload("delt", d["step_delt"])
    s.eval_flux_residual()
    fr = s.download("flux_res")
    assert per_component_err(fr, d["step_flux_res"]) < 1e-11
    s.upload("flux_res", d["step_flux_res"])
    s.implicit_forward_backward_sweeps()
    assert per_component_err(s.download("delUn"), d["step_delUn"]) < 1e-11
    assert per_component_err(s.download("delUp"), d["step_delUp"]) < 1e-11
    s.upload("delUn", d["step_delUn"])
    s.state_update()
    assert per_component_err(s.download("prim"), d["step_prim"]) < 1e-13
    r = s.residual()
    assert abs(r["res_new"] - float(d["step_res_new"])) <= 1e-12 * abs(float(d["step_res_new"]))
    s.close()


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("name", FIXTURE_NAMES)
def test_history_and_final_state_match_reference(name, order):
    note(f"[gpu] residual history, fixture {name}, flux order {order}")
    cl, d, s = make_solver(name, flux_order=order, sweep_schedule=SWEEP_OF[order])
    iters = int(d["iterations"])
    rn, rs = s.run(1, iters)
    ref = d["hist"][:, 0]
    nz = ref != 0.0
    check_zero_residuals(rn[~nz], float(d["prim0"][0, 0]), cl.n, str(d["mode"]) != "explicit")
    assert np.max(np.abs(rn[nz] - ref[nz]) / np.abs(ref[nz])) < 1e-10
    for nm in ("prim", "U"):
        assert per_component_err(s.download(nm), d["final_" + nm]) < 1e-9, nm
    s.close()


def test_fpi_solver_matches_run_and_is_deterministic():
    note("[gpu] fpi_solver vs run, determinism")
    name = "explicit_morton"
    cl, d, s1 = make_solver(name)
    h1 = [s1.fpi_solver(t)[0] for t in range(1, 11)]
    p1 = s1.download("prim")
    s1.close()
    _, _, s2 = make_solver(name)
    h2, _ = s2.run(1, 10)
    p2 = s2.download("prim")
    s2.close()
    assert np.array_equal(np.array(h1), h2)
    assert np.array_equal(p1, p2)


def test_against_oracle_on_stretched_cloud():
    """A 20k-point stretched-shell cloud (the production generator), 15
    iterations, HIP vs the C oracle step for step."""
    note("[gpu] 20k shell cloud vs oracle")
    cl = pc.ellipsoid_cloud(1600, 800, 17600, seed=21, order="morton")
    mach, aoa, cfl = 0.84, 3.06, 0.2
    prim0 = np.tile(pc.freestream_prim(mach, aoa), (cl.n, 1))
    o = orc.Oracle(cl, mach, aoa, cfl, prim=prim0)
    rn_o, _ = o.run(1, 15, False)
    s = solver.Solver(cl, mach, aoa, cfl)
    s.upload("prim", prim0)
    rn_g, _ = s.run(1, 15)
    nz = rn_o != 0
    assert np.all(rn_g[~nz] == 0.0)
    assert np.max(np.abs(rn_g[nz] - rn_o[nz]) / rn_o[nz]) < 1e-10
    assert per_component_err(s.download("prim"), o.prim) < 1e-9
    s.close()


def test_fused_and_list_orders_agree_on_stretched_cloud():
    """The fused kernel evaluates the same fluxes and applies each list's
    least-squares solve edge by edge; over 10 iterations on a stretched
    Morton-ordered cloud it tracks the list-by-list kernel to 1e-12."""
    note("[gpu] fused vs list order, 30k shell cloud")
    cl = pc.ellipsoid_cloud(2400, 1200, 26400, seed=23, order="morton")
    mach, aoa, cfl = 0.84, 3.06, 0.2
    prim0 = np.tile(pc.freestream_prim(mach, aoa), (cl.n, 1))
    out = {}
    for order in ORDERS:
        s = solver.Solver(cl, mach, aoa, cfl, flux_order=order)
        s.upload("prim", prim0)
        rn, _ = s.run(1, 10)
        out[order] = (rn, s.download("prim"), s.point_classes())
        s.close()
    note(f"      point classes {out['fused'][2]}")
    assert out["fused"][2]["fused_interior"] > 0.9 * cl.counts()["interior"]
    a, b = out["fused"][0], out["lists"][0]
    nz = b != 0
    assert np.max(np.abs(a[nz] - b[nz]) / b[nz]) < 1e-12
    assert per_component_err(out["fused"][1], out["lists"][1]) < 1e-12


def test_sweep_schedules_on_stretched_cloud():
    """Implicit LU-SGS on a 20k-point stretched-shell cloud: the sync-free
    sweeps against the C oracle (8 iterations, same bounds as the fixtures)
    and against the level-set schedule (same dependencies, each list's solve
    applied edge by edge instead of on the summed products: 1e-12)."""
    note("[gpu] implicit sweeps, 20k shell cloud: async vs levels vs oracle")
    cl = pc.ellipsoid_cloud(1600, 800, 17600, seed=22, order="morton")
    mach, aoa, cfl, iters = 0.84, 3.06, 2.0, 8
    prim0 = np.tile(pc.freestream_prim(mach, aoa), (cl.n, 1))
    o = orc.Oracle(cl, mach, aoa, cfl, prim=prim0)
    rn_o, _ = o.run(1, iters, True)
    out = {}
    for sweep in ("async", "levels"):
        s = solver.Solver(cl, mach, aoa, cfl, mode="implicit", sweep_schedule=sweep)
        s.upload("prim", prim0)
        rn, _ = s.run(1, iters)
        out[sweep] = (rn, s.download("prim"))
        note(f"      {sweep}: sweep levels {s.sweep_info()}")
        s.close()
    nz = rn_o != 0
    for sweep in out:
        rn = out[sweep][0]
        err_r = float(np.max(np.abs(rn[nz] - rn_o[nz]) / rn_o[nz]))
        err_p = per_component_err(out[sweep][1], o.prim)
        note(f"      {sweep} vs oracle: residual {err_r:.2e}, prim {err_p:.2e}")
        check_zero_residuals(rn[~nz], float(prim0[0, 0]), cl.n, True)
        assert err_r < 1e-10, sweep
        assert err_p < 1e-9, sweep
    a, b = out["async"][0], out["levels"][0]
    err_r = float(np.max(np.abs(a[nz] - b[nz]) / b[nz]))
    err_p = per_component_err(out["async"][1], out["levels"][1])
    note(f"      async vs levels: residual {err_r:.2e}, prim {err_p:.2e}")
    assert err_r < 1e-12
    assert err_p < 1e-12


def test_bind_rejects_bad_input():
    note("[gpu] argument checks")
    cl, d = load_fixture("explicit_blocks")
    bad = pc.PointCloud(cl.x, cl.y, cl.z, cl.tan1, cl.tan2, cl.nor, cl.status, cl.nbhs,
                        cl.conn.copy(), cl.min_dist)
    bad.conn[5, 3] = cl.n + 7
    with pytest.raises(solver.MfcfdError, match="out of range"):
        solver.Solver(bad, 0.6, 2.0, 0.2)
