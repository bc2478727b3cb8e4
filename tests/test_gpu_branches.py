"""HIP path on the branches the canonical configurations never reach, against
the reference's own fixtures (written by the reference compiled with flang):

  explicit_supersonic / implicit_supersonic
      status 5/6 points: the state update's nearest-neighbour copy
      (state_update.f90:80-114) and, implicit, the sweeps' delU copy
      (implicit_forward_and_reverse_sweeps.f90:38-52, 69-83);
  explicit_power2_inner1 / implicit_power2_inner1
      least-squares weights 1/dist**2 (the kernels compiled for power /= 0)
      and one second-order q-derivative iteration (q_derivatives.f90:116-201);
  implicit_backup
      the alternative pairing: type-ordered sweeps over wall, interior, outer
      and supersonic points with the outer D_inv and sweep terms
      (implicit_forward_and_reverse_sweeps_backup.f90:16-141), and
      state_update_implicit.f90.  Its outer points turn non-finite in the
      second iteration in the reference too; non-finite entries must match
      position for position and the finite ones the usual tolerances.

Tolerances as in test_gpu_parity.py (device exp/log/erf are ocml's, the
reference's glibc's): one entry point 1e-11..1e-13 relative to the array's
largest finite magnitude per component; residual history 1e-10 relative;
final state 1e-9 per component.
"""

import sys

import numpy as np
import pytest

from conftest import BACKUP_FIXTURES, BRANCH_FIXTURES, PKG, check_zero_residuals, fixture_params, load_fixture

pytestmark = pytest.mark.gpu
solver = PKG.solver

NAMES = BRANCH_FIXTURES + BACKUP_FIXTURES
# (flux order, sweep schedule) pairs: the default fused order with the
# sync-free sweeps, and the reference's list order with the level sets
ORDERS = [("fused", "async"), ("lists", "levels")]


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def component_err(a, b):
    """Largest difference relative to each component's largest finite
    magnitude; non-finite entries must coincide."""
    a = np.asarray(a, dtype=np.float64).reshape(len(a), -1)
    b = np.asarray(b, dtype=np.float64).reshape(len(b), -1)
    fa, fb = np.isfinite(a), np.isfinite(b)
    assert np.array_equal(fa, fb), f"{int((fa != fb).sum())} entries differ in finiteness"
    worst = 0.0
    for c in range(a.shape[1]):
        m = fb[:, c]
        if not m.any():
            continue
        scale = np.max(np.abs(b[m, c]))
        d = np.max(np.abs(a[m, c] - b[m, c]))
        worst = max(worst, d / scale if scale > 0 else d)
    return worst


def make_solver(name, **kw):
    cl, d = load_fixture(name)
    power, inner = fixture_params(d)
    s = solver.Solver(cl, float(d["mach"]), float(d["aoa"]), float(d["cfl"]), mode=str(d["mode"]),
                      power=power, inner_iterations=inner, **kw)
    s.upload("prim", d["prim0"])
    return cl, d, s


@pytest.mark.parametrize("order,sweep", ORDERS)
@pytest.mark.parametrize("name", NAMES)
def test_branch_entry_points_match_reference(name, order, sweep):
    note(f"[gpu] entry points, fixture {name}, flux order {order}, sweeps {sweep}")
    cl, d, s = make_solver(name, flux_order=order, sweep_schedule=sweep)
    s.upload("prim", d["start_prim"])
    s.upload("delUp", d["start_delUp"])
    s.upload("delUn", d["start_delUn"])
    s.compute_conserved_vector()
    assert component_err(s.download("U"), d["step_U"]) < 1e-13
    s.eval_q_variables()
    assert component_err(s.download("q"), d["step_q"]) < 1e-13
    s.eval_q_derivatives()
    assert component_err(s.download("qm"), d["step_qm"]) < 1e-13
    assert component_err(s.download("dq"), d["step_dq"]) < 1e-11
    s.timestep_delt()
    assert component_err(s.download("delt")[:, None], d["step_delt"][:, None]) < 1e-13
    # flux residual and sweeps from the reference's own inputs
    for nm in ("q", "qm", "dq", "delt"):
        s.upload(nm, d["step_" + nm])
    s.eval_flux_residual()
    assert component_err(s.download("flux_res"), d["step_flux_res"]) < 1e-11
    s.upload("flux_res", d["step_flux_res"])
    s.implicit_forward_backward_sweeps()
    assert component_err(s.download("delUp"), d["step_delUp"]) < 1e-11
    assert component_err(s.download("delUn"), d["step_delUn"]) < 1e-11
    s.upload("delUp", d["step_delUp"])
    s.upload("delUn", d["step_delUn"])
    s.state_update()
    assert component_err(s.download("prim"), d["step_prim"]) < 1e-12
    assert component_err(s.download("delUp"), d["step_delUp_after"]) < 1e-12
    r = s.residual()
    ref = float(d["step_res_new"])
    assert abs(r["res_new"] - ref) <= 1e-12 * abs(ref)
    s.close()


@pytest.mark.parametrize("order,sweep", ORDERS)
@pytest.mark.parametrize("name", NAMES)
def test_branch_history_and_final_state_match_reference(name, order, sweep):
    note(f"[gpu] residual history, fixture {name}, flux order {order}, sweeps {sweep}")
    cl, d, s = make_solver(name, flux_order=order, sweep_schedule=sweep)
    iters = int(d["iterations"])
    rn, _ = s.run(1, iters)
    ref = d["hist"][:, 0]
    nz = ref != 0.0
    check_zero_residuals(rn[~nz], float(d["prim0"][0, 0]), cl.n, str(d["mode"]) != "explicit")
    assert np.max(np.abs(rn[nz] - ref[nz]) / np.abs(ref[nz])) < 1e-10
    for nm in ("prim", "U", "delUp", "delUn"):
        assert component_err(s.download(nm), d["final_" + nm]) < 1e-9, nm
    s.close()


def test_supersonic_fixtures_hold_supersonic_points():
    """the branch fixtures exercise what they are named for"""
    for name in ("explicit_supersonic", "implicit_supersonic", "implicit_backup"):
        cl, d = load_fixture(name)
        c = cl.counts()
        assert c["supersonic_inlet"] > 10 and c["supersonic_outlet"] > 10, name
    for name in ("explicit_power2_inner1", "implicit_power2_inner1"):
        cl, d = load_fixture(name)
        assert fixture_params(d) == (2.0, 1)
