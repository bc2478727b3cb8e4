"""Deck formats, restart round trip and wall post-processing (CPU)."""

import math
import os
import tempfile

import numpy as np

from conftest import PKG, load_fixture

pc = PKG.pointcloud
pp = PKG.postprocess


def test_input_deck_and_solution_round_trip():
    """3d_input_data (point_preprocessor.f90:30-36) and the solution files
    (initial_conditions.f90:26, post_processor.f90:34) written and read back
    bit for bit; a primal-solution file restarts as stored-solution."""
    cl, d = load_fixture("explicit_blocks")
    rng = np.random.default_rng(5)
    prim = d["prim0"] * (1.0 + 1e-3 * rng.standard_normal(d["prim0"].shape))
    with tempfile.TemporaryDirectory() as wd:
        deck = os.path.join(wd, "3d_input_data")
        pc.write_input_deck(deck, cl)
        back = pc.read_input_deck(deck)
        for nm in ("x", "y", "z", "tan1", "tan2", "nor", "status", "nbhs", "conn", "min_dist"):
            assert np.array_equal(getattr(back, nm), getattr(cl, nm)), nm
        sol = os.path.join(wd, "primal-solution.dat")
        pc.write_solution(sol, cl, prim)
        assert np.array_equal(pc.read_solution(sol, cl.n), prim)
        # records wrapped over several lines, as a Fortran list-directed
        # WRITE may leave them, read the same
        wrapped = os.path.join(wd, "wrapped.dat")
        with open(wrapped, "w") as f:
            for row in np.column_stack([cl.x, cl.y, cl.z, prim])[:10]:
                s = ["%.17g" % v for v in row]
                f.write(" " + " ".join(s[:4]) + "\n " + " ".join(s[4:]) + "\n")
        assert np.array_equal(pc.read_solution(wrapped, 10), prim[:10])
        # restart: the solution file is a valid stored-solution.dat
        os.rename(sol, os.path.join(wd, "stored-solution.dat"))
        assert np.array_equal(pc.read_solution(os.path.join(wd, "stored-solution.dat"), cl.n), prim)


def test_ellipsoid_area_quadrature():
    # sphere, and the closed form of a prolate spheroid (a > b = c)
    assert abs(pp.ellipsoid_area(1.0, 1.0, 1.0, n=400) - 4.0 * math.pi) < 1e-4
    a, b = 2.0, 1.0
    e = math.sqrt(1.0 - b * b / (a * a))
    exact = 2.0 * math.pi * b * b * (1.0 + a / (b * e) * math.asin(e))
    assert abs(pp.ellipsoid_area(a, b, b, n=400) - exact) < 1e-4 * exact


def test_wall_areas_cover_the_body_and_close():
    cl = pc.ellipsoid_cloud(1600, 800, 17600, seed=4)
    a, b, c = cl.meta["axes"]
    A = pp.wall_areas(cl)
    assert np.all(A > 0)
    total = pp.ellipsoid_area(a, b, c)
    assert abs(A.sum() - total) < 0.02 * total
    w = cl.status == pc.STATUS_WALL
    # a closed surface: the area-weighted normals sum to ~0
    assert np.linalg.norm((cl.nor[w] * A[:, None]).sum(axis=0)) < 1e-3 * total


def test_force_coefficients():
    cl = pc.ellipsoid_cloud(1600, 800, 17600, seed=4)
    mach, aoa = 0.6, 2.0
    prim = np.tile(pc.freestream_prim(mach, aoa), (cl.n, 1))
    w = cl.status == pc.STATUS_WALL
    assert np.all(pp.wall_cp(prim[w], mach) == 0.0)
    # a uniform pressure offset exerts (almost) no net force on a closed body
    prim[:, 4] += 0.05
    f = pp.force_coefficients(cl, prim, mach, aoa)
    cp = 0.05 / (0.5 * mach * mach)
    assert abs(f["CL"]) < 1e-2 * cp and abs(f["CD"]) < 1e-2 * cp
    # suction on the +y half only: lift along +y rotated by aoa, matching the
    # projected area of that half
    prim[:, 4] = 1.0 / 1.4
    up = w & (cl.y > 0)
    prim[up, 4] -= 0.1
    f = pp.force_coefficients(cl, prim, mach, aoa)
    F = f["force"]
    A = pp.wall_areas(cl)
    proj = (cl.nor[w][:, 1] * A * (cl.y[w] > 0)).sum()
    assert abs(F[1] - 0.1 * proj) < 1e-12
    q = 0.5 * mach * mach
    al = math.radians(aoa)
    assert abs(f["CL"] - (F[1] * math.cos(al) - F[0] * math.sin(al)) / (q * f["s_ref"])) < 1e-12
    assert f["CL"] > 0
