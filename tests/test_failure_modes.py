"""What the reference does with inputs the library refuses (CPU, oracle).

The oracle restates the reference step for step, so it shows what the
reference computes from:
  * a coplanar/collinear least-squares stencil: the Cramer solve of that
    direction list divides by a zero determinant (interior_flux_dGxpos.f90:
    84-95) and the point's flux residual is NaN;
  * an empty direction list: the same division, 0/0;
  * a NaN in prim: the NaN enters the residual.
The library refuses the first two at bind and the third at upload with
MFCFD_EINVAL (tests/test_gpu_failure_modes.py), where the reference would
carry on with NaN.  The host-side preprocessing entry points still accept
such clouds: they only sort neighbours.
"""

import numpy as np
import pytest

import orc
from conftest import PKG, degenerate_cloud, load_fixture

pc = PKG.pointcloud
solver = PKG.solver
MACH, AOA, CFL = 0.6, 2.0, 0.2


def first_flux(cl, split=None):
    prim0 = np.tile(pc.freestream_prim(MACH, AOA), (cl.n, 1))
    o = orc.Oracle(cl, MACH, AOA, CFL, prim=prim0, preprocess=split is None)
    if split is not None:
        # the host's own split lists, as the reference's preprocessing hands them over
        cnt, lists = split
        for d, nm in enumerate(["xpos", "xneg", "ypos", "yneg", "zpos", "zneg"]):
            getattr(o, nm + "_nbhs")[:] = cnt[d]
            getattr(o, nm + "_conn")[:] = lists[d][:len(getattr(o, nm + "_conn"))]
        o.alias[:], o.point_with_alias[:] = solver.alias_order(cl)
    # the freestream makes every residual zero; perturb so a zero determinant shows
    rng = np.random.default_rng(3)
    o.prim[:] = o.prim * (1.0 + 1e-3 * rng.standard_normal(o.prim.shape))
    for step in ("compute_conserved_vector", "eval_q_variables", "eval_q_derivatives", "timestep_delt",
                 "eval_flux_residual"):
        o.call(step)
    return o


@pytest.mark.parametrize("kind", ["collinear", "empty"])
def test_reference_divides_by_zero_on_a_degenerate_stencil(kind):
    cl, p, split = degenerate_cloud(kind)
    # the host preprocessing drop-ins accept the cloud
    cnt, _ = solver.split_stencils(cl)
    assert cnt[0, p] >= 3
    o = first_flux(cl, split)
    fr = o.flux_res.reshape(cl.n, 5)
    bad = np.nonzero(~np.all(np.isfinite(fr), axis=1))[0]
    assert p in bad
    # the untouched cloud gives a finite residual everywhere
    ref, _ = load_fixture("explicit_blocks")
    assert np.all(np.isfinite(first_flux(ref).flux_res))


def test_reference_carries_nan_in_prim_into_the_residual():
    cl, d = load_fixture("explicit_blocks")
    prim = np.array(d["prim0"], dtype=np.float64)
    prim[7, 2] = np.nan
    o = orc.Oracle(cl, float(d["mach"]), float(d["aoa"]), float(d["cfl"]), prim=prim)
    o.fpi_solver(1, False)
    o.fpi_solver(2, False)
    assert np.isnan(o.res_new)
