"""Wall quantities after an iteration: pressure coefficient and forces.

The reference writes only the primitive state (post_processor.f90:33-39);
it computes no Cp or force coefficients.  These follow the usual
definitions in the reference's non-dimensionalisation (freestream
rho = 1, p = 1/1.4, so the sound speed is 1 and |V_inf| = Mach;
parameter.f90:13-33):

    Cp   = (p - p_inf) / (0.5 * rho_inf * |V_inf|^2)
    F    = - sum_wall (p - p_inf) * n * A          (n: outward body normal)
    CL   = F . (-sin aoa, cos aoa, 0) / (0.5 * rho_inf * |V_inf|^2 * S_ref)
    CD   = F . ( cos aoa, sin aoa, 0) / (0.5 * rho_inf * |V_inf|^2 * S_ref)

with the angle of attack in the x-y plane as in the freestream of
parameter.f90:29-31.  The wall-point areas A are the Voronoi cells of the
wall points on the ellipsoid: each point is mapped to the unit sphere by
u = (x/a, y/b, z/c)/|.|, the spherical Voronoi cell areas are taken there,
and scaled by the ellipsoid's area element for that map,
dA = a b c |(u_x/a, u_y/b, u_z/c)| dOmega.  The default reference area is
the ellipsoid's projection on the x-z plane, pi a c.
"""

from __future__ import annotations

import math

import numpy as np

from .pointcloud import STATUS_WALL

P_INF = 1.0 / 1.4


def wall_cp(prim, mach):
    """Cp of each row of prim (rows are points)."""
    prim = np.asarray(prim, dtype=np.float64)
    return (prim[:, 4] - P_INF) / (0.5 * mach * mach)


def wall_areas(cloud, axes=None):
    """Surface area carried by each wall point, in wall-index order."""
    from scipy.spatial import SphericalVoronoi

    a, b, c = axes if axes is not None else cloud.meta["axes"]
    w = np.flatnonzero(cloud.status == STATUS_WALL)
    p = np.stack([cloud.x[w] / a, cloud.y[w] / b, cloud.z[w] / c], axis=1)
    u = p / np.linalg.norm(p, axis=1, keepdims=True)
    sv = SphericalVoronoi(u, radius=1.0, center=np.zeros(3))
    omega = sv.calculate_areas()
    scale = a * b * c * np.linalg.norm(u / np.array([a, b, c]), axis=1)
    return omega * scale


def force_coefficients(cloud, prim, mach, aoa, axes=None, s_ref=None, areas=None):
    """CL, CD and the force vector from the wall pressures.

    prim holds every point of the cloud (host order)."""
    a, b, c = axes if axes is not None else cloud.meta["axes"]
    w = np.flatnonzero(cloud.status == STATUS_WALL)
    A = wall_areas(cloud, (a, b, c)) if areas is None else np.asarray(areas)
    n = np.asarray(cloud.nor, dtype=np.float64)[w]
    dp = np.asarray(prim, dtype=np.float64)[w, 4] - P_INF
    F = -(dp[:, None] * n * A[:, None]).sum(axis=0)
    q_inf = 0.5 * mach * mach
    s = s_ref if s_ref is not None else math.pi * a * c
    al = math.radians(aoa)
    lift = np.array([-math.sin(al), math.cos(al), 0.0])
    drag = np.array([math.cos(al), math.sin(al), 0.0])
    return {"CL": float(F @ lift / (q_inf * s)), "CD": float(F @ drag / (q_inf * s)),
            "force": F, "s_ref": s}


def ellipsoid_area(a, b, c, n=2000):
    """Surface area of the ellipsoid by quadrature over (theta, phi)."""
    th = (np.arange(n) + 0.5) * math.pi / n
    ph = (np.arange(2 * n) + 0.5) * math.pi / n
    T, P = np.meshgrid(th, ph, indexing="ij")
    st, ct, sp, cp = np.sin(T), np.cos(T), np.sin(P), np.cos(P)
    # |r_theta x r_phi| for r = (a st cp, b st sp, c ct)
    nx = b * c * st * st * cp
    ny = a * c * st * st * sp
    nz = a * b * st * ct
    dA = np.sqrt(nx * nx + ny * ny + nz * nz)
    return float(dA.sum() * (math.pi / n) ** 2)
