"""The C restatement against the reference Fortran's own outputs.

The golden fixtures were written by the reference compiled with flang
(tests/golden/make_golden.py).  The oracle uses the same glibc libm and the
same expression order without fused multiply-add, so the bar here is
bit-for-bit equality for every array and every residual.
"""

import numpy as np
import pytest

import orc
from conftest import ALL_FIXTURES, BACKUP_FIXTURES, FIXTURE_NAMES, fixture_params, load_fixture


def _oracle(name, prim=None):
    cl, d = load_fixture(name)
    power, inner = fixture_params(d)
    o = orc.Oracle(cl, float(d["mach"]), float(d["aoa"]), float(d["cfl"]), power=power,
                   inner_iterations=inner, prim=d["prim0"] if prim is None else prim)
    return cl, d, o


@pytest.mark.parametrize("name", ALL_FIXTURES)
def test_split_stencils_and_alias_bitwise(name):
    # generate_connectivity.f90 (with the 22-column spill) and implicit_aliasing.f90
    cl, d, o = _oracle(name)
    cnt, lists = o.split_lists()
    assert np.array_equal(cnt, d["split_counts"])
    assert np.array_equal(lists, d["split_lists"])
    assert np.array_equal(o.alias, d["alias"])
    assert np.array_equal(o.point_with_alias, d["point_with_alias"])


@pytest.mark.parametrize("name", ALL_FIXTURES)
def test_residual_history_and_final_state_bitwise(name):
    cl, d, o = _oracle(name)
    iters = int(d["iterations"])
    rn, rs = o.run(1, iters, str(d["mode"]))
    assert np.array_equal(rn, d["hist"][:, 0])
    assert np.array_equal(rs, d["hist"][:, 1])
    # equal_nan: the alternative pairing's outer points are non-finite after
    # its second iteration, in the reference as here
    for nm in ("prim", "U", "flux_res", "delUp", "delUn"):
        assert np.array_equal(getattr(o, nm), d["final_" + nm], equal_nan=True), nm


@pytest.mark.parametrize("name", ALL_FIXTURES)
def test_each_entry_point_bitwise(name):
    """Start from the reference's own state before its last iteration and
    compare the output of every entry point (fpi_solver.f90:27-36)."""
    cl, d, o = _oracle(name, prim=None)
    o.prim[:] = d["start_prim"]
    o.delUp[:] = d["start_delUp"]
    o.delUn[:] = d["start_delUn"]
    eq = lambda a, b: np.array_equal(a, b, equal_nan=True)
    o.call("compute_conserved_vector")
    assert eq(o.U, d["step_U"])
    o.call("eval_q_variables")
    assert eq(o.q, d["step_q"])
    o.call("eval_q_derivatives")
    assert eq(o.qm, d["step_qm"])
    assert eq(o.dq, d["step_dq"])
    o.call("timestep_delt")
    assert eq(o.delt, d["step_delt"])
    o.call("eval_flux_residual")
    assert eq(o.flux_res, d["step_flux_res"])
    mode = str(d["mode"])
    o.call({"explicit": "explicit_sweep", "implicit": "implicit_sweeps",
            "implicit_backup": "implicit_sweeps_backup"}[mode])
    assert eq(o.delUp, d["step_delUp"])
    assert eq(o.delUn, d["step_delUn"])
    o.call("state_update_implicit" if mode == "implicit_backup" else "state_update")
    assert eq(o.prim, d["step_prim"])
    assert eq(o.delUp, d["step_delUp_after"])
    assert o.res_new == float(d["step_res_new"])


def test_fixtures_cover_the_spill_and_alias_hazards():
    """The Morton fixtures interleave wall/outer points with interior ones,
    so the 22-column spill of wall zneg / outer zpos lists lands in other
    points' columns; every fixture leaves point 1 unswept and ends the alias
    order with zeros."""
    for name in FIXTURE_NAMES:
        cl, d = load_fixture(name)
        n = cl.n
        pwa = d["point_with_alias"]
        assert 1 not in set(pwa.tolist())          # point 1 is never swept
        assert pwa[-1] == 0                         # alias order ends in zeros
        assert d["alias"][0] == 1
    cl, d = load_fixture("explicit_morton")
    st = cl.status
    # boundary point i followed by an interior point i+1: its spill overwrote
    # the first entries of that interior point's zneg/zpos column
    spill = np.flatnonzero((st[:-1] != 0) & (st[1:] == 0))
    assert spill.size > 10
