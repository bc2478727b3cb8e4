"""Closed-form and brute-force checks of the oracle, independent of the
reference build: each would fail on a dropped term, a wrong sign, index or
column."""

import math

import numpy as np
import pytest

import orc
from conftest import PKG, load_fixture

pc = PKG.pointcloud
RNG = np.random.default_rng(1234)


def random_frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    h = rng.normal(size=3)
    t1 = h - h.dot(n) * n
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    return t1, t2, n


def random_prim(rng):
    return np.array([rng.uniform(0.5, 2.0), *rng.uniform(-1.5, 1.5, size=3), rng.uniform(0.3, 1.5)])


def euler_flux(axis_vec, t1, t2, n, prim):
    """Exact Euler flux through a plane with normal axis_vec, components in
    the (t1, t2, n) frame."""
    rho, u, p = prim[0], prim[1:4], prim[4]
    ua = u.dot(axis_vec)
    e = 2.5 * p + 0.5 * rho * u.dot(u)
    mom = rho * ua * u + p * axis_vec
    return np.array([rho * ua, mom.dot(t1), mom.dot(t2), mom.dot(n), (e + p) * ua])


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_kfvs_half_fluxes_sum_to_euler_flux(axis):
    # split_fluxes.f90: G+ + G- is the full Euler flux along t1, t2 or n
    for _ in range(50):
        t1, t2, n = random_frame(RNG)
        prim = random_prim(RNG)
        gp = orc.flux_half(axis, +1, t1, t2, n, prim)
        gn = orc.flux_half(axis, -1, t1, t2, n, prim)
        ref = euler_flux((t1, t2, n)[axis], t1, t2, n, prim)
        assert np.allclose(gp + gn, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("sign", [+1, -1])
def test_octant_fluxes_sum_to_half_flux(axis, sign):
    # octant_fluxes.f90: wall (A3-) + outer (A3+) quarter fluxes = half flux
    for _ in range(50):
        t1, t2, n = random_frame(RNG)
        prim = random_prim(RNG)
        gw = orc.flux_octant(axis, sign, -1, t1, t2, n, prim)
        go = orc.flux_octant(axis, sign, +1, t1, t2, n, prim)
        gh = orc.flux_half(axis, sign, t1, t2, n, prim)
        assert np.allclose(gw + go, gh, rtol=1e-13, atol=1e-13)


def test_half_flux_limits_and_mass_moment_by_quadrature():
    # positive half-range mass flux = integral over v_a > 0 of v_a * Maxwellian
    t1, t2, n = np.eye(3)
    for _ in range(20):
        prim = random_prim(RNG)
        rho, u, p = prim[0], prim[1:4], prim[4]
        beta = 0.5 * rho / p
        v = np.linspace(0.0, u[0] + 12.0 / math.sqrt(beta), 200001)
        f = rho * math.sqrt(beta / math.pi) * np.exp(-beta * (v - u[0]) ** 2)
        mass = np.trapezoid(v * f, v)
        g = orc.flux_half(0, +1, t1, t2, n, prim)
        assert abs(g[0] - mass) < 1e-8 * max(1.0, abs(mass))
    # supersonic limit: everything goes into the positive half
    prim = np.array([1.0, 40.0, 0.1, -0.2, 0.7])
    g = orc.flux_half(0, -1, t1, t2, n, prim)
    assert np.all(np.abs(g) < 1e-300 + 1e-12)


def test_q_variables_round_trip():
    # q_variables.f90 -> qvariables_to_primitive_variables.f90 is the identity
    for _ in range(200):
        prim = random_prim(RNG)
        rho, u1, u2, u3, pr = prim
        beta = 0.5 * rho / pr
        q = np.array([math.log(rho) + 2.5 * math.log(beta) - beta * (u1 * u1 + u2 * u2 + u3 * u3),
                      2 * beta * u1, 2 * beta * u2, 2 * beta * u3, -2 * beta])
        back = orc.qtilde_to_primitive(q)
        assert np.allclose(back, prim, rtol=2e-14, atol=2e-14)


def _tiny_oracle(prim=None, mach=0.6, aoa=2.0):
    cl, d = load_fixture("explicit_blocks")
    if prim is None:
        prim = np.tile(pc.freestream_prim(mach, aoa), (cl.n, 1))
    return cl, orc.Oracle(cl, mach, aoa, 0.2, prim=prim)


def test_conserved_primitive_round_trip_and_wall_tangency():
    cl, _ = load_fixture("explicit_blocks")
    prim = np.array([random_prim(RNG) for _ in range(cl.n)])
    _, o = _tiny_oracle(prim=prim)
    o.call("compute_conserved_vector")
    # U is the rotated conserved vector
    mom = prim[:, 0:1] * prim[:, 1:4]
    assert np.allclose(o.U[:, 1], np.sum(mom * cl.tan1, axis=1), rtol=1e-14, atol=1e-14)
    assert np.allclose(o.U[:, 3], np.sum(mom * cl.nor, axis=1), rtol=1e-14, atol=1e-14)
    o.delUn[:] = 0.0
    o.flux_res[:] = 0.0
    o.call("state_update")
    inter = cl.status == 0
    assert np.allclose(o.prim[inter], prim[inter], rtol=1e-13, atol=1e-13)
    # wall points lose their normal momentum (state_update.f90:37)
    wall = cl.status == 1
    un = np.sum(o.prim[wall, 1:4] * cl.nor[wall], axis=1)
    assert np.all(np.abs(un) < 1e-14)
    assert o.res_new == 0.0


def test_freestream_gives_zero_flux_residual_exactly():
    # every G_k - G_i vanishes on a uniform state
    cl, o = _tiny_oracle()
    for nm in ("compute_conserved_vector", "eval_q_variables", "eval_q_derivatives",
               "timestep_delt", "eval_flux_residual"):
        o.call(nm)
    assert np.all(o.dq == 0.0)
    assert np.all(o.flux_res == 0.0)


def test_first_implicit_iteration_moves_density_by_round_off_only():
    # From the freestream the flux residual is zero, yet the sweeps see
    # round-off: prim(U) does not reproduce prim bit for bit (the frame
    # rotation and the pressure), so some Gnew - Gold are ulps.  The density
    # corrections that result are nonzero but below half an ulp of the
    # freestream density, which is why the reference's first residual
    # (delta rho after rounding, state_update.f90:39-44) is exactly zero.
    cl, d = load_fixture("implicit_morton")
    o = orc.Oracle(cl, float(d["mach"]), float(d["aoa"]), float(d["cfl"]), prim=d["prim0"])
    rn, _ = o.run(1, 1, True)
    assert rn[0] == 0.0 and float(d["hist"][0, 0]) == 0.0
    assert np.all(o.flux_res == 0.0)
    swept = cl.status != 2
    drho = o.delUn.reshape(-1, 5)[swept, 0]
    rho = float(d["prim0"][0, 0])
    assert np.count_nonzero(drho) > 0
    assert np.max(np.abs(drho)) < 0.5 * np.spacing(rho)


def test_ubar_at_freestream_is_freestream():
    # conserved_vector_Ubar.f90: A3+ + A3- = 1 and the B3 terms cancel
    cl, o = _tiny_oracle(mach=0.84, aoa=3.06)
    o.call("compute_conserved_vector")
    o.delUn[:] = 0.0
    o.flux_res[:] = 0.0
    o.delt[:] = 0.1
    prim0 = o.prim.copy()
    o.call("state_update")
    outer = cl.status == 2
    assert np.allclose(o.prim[outer], prim0[outer], rtol=1e-13, atol=1e-14)


def test_wls_gradient_exact_for_linear_fields():
    cl, o = _tiny_oracle()
    xyz = cl.xyz()
    grad = RNG.normal(size=(5, 3))
    o.q[:] = xyz @ grad.T + RNG.normal(size=5)
    o.call("eval_q_derivatives")
    # dq(c, r, k) = d q_r / d x_c
    assert np.allclose(o.dq, np.broadcast_to(grad, (cl.n, 5, 3)), rtol=1e-8, atol=1e-8)
    nb = cl.conn[:, :25] - 1
    full = np.concatenate([o.q[:, None, :], o.q[nb]], axis=1)
    assert np.array_equal(o.qm[:, :, 0], full.max(axis=1))
    assert np.array_equal(o.qm[:, :, 1], full.min(axis=1))


def test_timestep_brute_force():
    cl, _ = load_fixture("explicit_blocks")
    prim = np.array([random_prim(RNG) for _ in range(cl.n)])
    _, o = _tiny_oracle(prim=prim)
    o.call("timestep_delt")
    xyz = cl.xyz()
    nb = cl.conn[:, :25] - 1
    d = np.sqrt(np.sum((xyz[nb] - xyz[:, None, :]) ** 2, axis=2))
    p = prim[nb]
    a = np.sqrt(np.sum(p[..., 1:4] ** 2, axis=2)) + 3.0 * np.sqrt(p[..., 4] / p[..., 0])
    expect = np.minimum(1.0, np.min(0.2 * d / a, axis=1))
    assert np.allclose(o.delt, expect, rtol=1e-15, atol=0)


def test_limiter_threshold_is_single_precision_literal():
    # limiters.f90:26 compares with the REAL(4) literal 10e-6
    assert float(np.float32(10e-6)) == 9.999999747378751e-06
    cl, o = _tiny_oracle()
    o.q[:] = 1.0
    o.qm[:, :, 0] = 2.0
    o.qm[:, :, 1] = 0.0
    phi = np.zeros(5)
    for delta, expect_one in ((9.99999e-6, True), (-9.999999747378751e-06, True),
                              (1.0000001e-5, False), (0.5, False), (-0.3, False)):
        qt = np.full(5, 1.0 + delta)
        orc.lib().orc_venkat_limiter(orc.C.byref(o.s), orc._dptr(qt), orc._dptr(phi), 3)
        if expect_one:
            assert np.all(phi == 1.0)
        else:
            assert np.all((phi > 0.0) & (phi <= 1.0))


def _split_bruteforce(cl):
    """Split stencils re-derived from generate_connectivity.f90 with the
    22-column spill expressed as a flat array."""
    n = cl.n
    xyz = cl.xyz()
    cnt = np.zeros((6, n), dtype=np.int32)
    lists = np.zeros((6, 22 * n + 32), dtype=np.int32)

    def push(lst, i, k):
        cnt[lst, i] += 1
        lists[lst, i * 22 + cnt[lst, i] - 1] = k

    order = [np.flatnonzero(cl.status == s) for s in (0, 1, 2)]
    for st, idx in zip((0, 1, 2), order):
        for i in idx:
            if st == 0:
                cnt[:, i] = 0
            else:
                cnt[[0, 1, 2, 3, 5], i] = 0
            for k1 in cl.conn[i, :cl.nbhs[i]]:
                d = xyz[k1 - 1] - xyz[i]
                s, t, m = d.dot(cl.tan1[i]), d.dot(cl.tan2[i]), d.dot(cl.nor[i])
                if s <= 0: push(0, i, k1)
                if (s > 0) if st == 0 else (s >= 0): push(1, i, k1)
                if t <= 0: push(2, i, k1)
                if (t > 0) if st == 0 else (t >= 0): push(3, i, k1)
                if st == 0:
                    if m <= 0: push(4, i, k1)
                    if m > 0: push(5, i, k1)
                elif st == 1:
                    push(5, i, k1)
                else:
                    push(4, i, k1)
    return cnt, lists[:, :22 * n]


def test_split_stencils_brute_force():
    cl, _ = load_fixture("explicit_morton")
    o = orc.Oracle(cl, 0.84, 3.06, 0.2)
    cnt, lists = o.split_lists()
    bc, bl = _split_bruteforce(cl)
    assert np.array_equal(cnt, bc)
    assert np.array_equal(lists, bl)


def test_alias_bfs_brute_force():
    cl, _ = load_fixture("implicit_morton")
    o = orc.Oracle(cl, 0.84, 3.06, 2.0)
    n = cl.n
    alias = np.zeros(n, dtype=np.int64)
    pwa = np.zeros(n, dtype=np.int64)
    alias[0] = 1
    pwa[0] = 1
    count = 0
    for i in range(n):
        p = pwa[i]
        if p == 0:
            continue
        for k in cl.conn[p - 1, :cl.nbhs[p - 1]]:
            if alias[k - 1] == 0:
                count += 1
                alias[k - 1] = count
                pwa[count - 1] = k
    assert np.array_equal(o.alias, alias)
    assert np.array_equal(o.point_with_alias, pwa)
