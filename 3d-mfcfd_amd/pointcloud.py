"""Synthetic ellipsoid point clouds and the reference's ASCII input formats.

The solver reads a point cloud in which every point carries a status
(0 interior, 1 wall, 2 outer), a local orthonormal frame (tan1, tan2, nor),
a neighbour list ``conn`` (1-based, at most 27 entries) and ``min_dist``
(distance to the nearest neighbour).  This module builds such clouds around
a tri-axial ellipsoid inside a spherical far field and reads/writes the two
text decks the reference consumes:

* ``3d_input_data``  one line per point:
  ``counter status x y z tan1(3) tan2(3) nor(3) nbhs conn(1:nbhs) min_dist``
  (read by point_preprocessor.f90:30-36);
* ``stored-solution.dat``  one line per point: ``x y z prim(1:5)``
  (read by initial_conditions.f90:26-30; the same columns are written by
  post_processor.f90:33-39 as ``primal-solution.dat``).

Geometry (all lengths in body units):
  wall      jittered latitude/longitude grid on the ellipsoid surface, rows
            and points spaced uniformly in arc length; ``nor`` is the outward
            body normal (pointing into the fluid), ``tan1`` follows the
            azimuthal direction and ``tan2 = nor x tan1``;
  interior  shells blended from the ellipsoid (t=0) to the far-field sphere
            (t=1); shell offsets grow geometrically from the wall spacing and
            each shell holds an arc-length latitude/longitude grid sized so the
            tangential spacing matches the radial one, capped at the far-field
            point spacing; 10 % jitter; Cartesian frame (e_x, e_y, e_z);
  outer     golden-spiral points on the far-field sphere, ``nor`` radial
            outward, jittered 10 % tangentially.

Ordering rules keep the reference's out-of-range accesses benign: point 1 is
an outer point (the alias breadth-first search never sweeps point 1) and the
last point is an interior point whose six split stencils all fit the 22
columns of the split-connectivity arrays (so no split list is written past
the end of its array).
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

STATUS_INTERIOR = 0
STATUS_WALL = 1
STATUS_OUTER = 2
STATUS_SUPERSONIC_INLET = 5
STATUS_SUPERSONIC_OUTLET = 6

MAX_CONN = 27        # data_structure.f90:111  conn(27, max_points)
SPLIT_COLS = 22      # data_structure.f90:119-124  *_conn(22, max_points)


@dataclass
class PointCloud:
    """Point arrays in the reference's layout, 0-based rows, 1-based conn.

    ``conn`` has shape (N, 27) and holds 1-based point numbers (0 = unused),
    exactly the column-major ``conn(27, N)`` of data_structure.f90 viewed
    row-per-point.
    """

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    tan1: np.ndarray      # (N, 3)
    tan2: np.ndarray      # (N, 3)
    nor: np.ndarray       # (N, 3)
    status: np.ndarray    # (N,) int32
    nbhs: np.ndarray      # (N,) int32
    conn: np.ndarray      # (N, 27) int32, 1-based
    min_dist: np.ndarray  # (N,)
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def counts(self) -> dict:
        s = self.status
        return {"interior": int((s == 0).sum()), "wall": int((s == 1).sum()),
                "outer": int((s == 2).sum()),
                "supersonic_inlet": int((s == 5).sum()),
                "supersonic_outlet": int((s == 6).sum())}

    def xyz(self) -> np.ndarray:
        return np.stack([self.x, self.y, self.z], axis=1)


def freestream_prim(mach: float, aoa_deg: float) -> np.ndarray:
    """Freestream primitive vector (rho, u1, u2, u3, p).

    Same expressions as parameter.f90:19-33: theta = aoa*pi/180,
    u1 = Mach*cos(theta), u2 = Mach*sin(theta), u3 = 0, p = 1/1.4.
    """
    pi = 4.0 * math.atan(1.0)
    theta = aoa_deg * pi / 180.0
    return np.array([1.0, mach * math.cos(theta), mach * math.sin(theta), 0.0,
                     1.0 / 1.4])


# --------------------------------------------------------------------------
# geometry helpers
# --------------------------------------------------------------------------

def _normalize(v: np.ndarray) -> np.ndarray:
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _golden_directions(m: int, rng: np.random.Generator, jitter: float) -> np.ndarray:
    """m quasi-uniform unit vectors (golden spiral), randomly rotated, jittered."""
    j = np.arange(m, dtype=np.float64)
    zc = 1.0 - 2.0 * (j + 0.5) / m
    r = np.sqrt(np.maximum(0.0, 1.0 - zc * zc))
    phi = j * (math.pi * (3.0 - math.sqrt(5.0))) + rng.uniform(0.0, 2.0 * math.pi)
    d = np.stack([r * np.cos(phi), r * np.sin(phi), zc], axis=1)
    spacing = math.sqrt(4.0 * math.pi / max(m, 1))
    d = d + rng.uniform(-jitter, jitter, size=d.shape) * spacing
    d = _normalize(d)
    # random rotation so that shells do not share the spiral's seam
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, a, b, c = q
    rot = np.array([
        [1 - 2 * (b * b + c * c), 2 * (a * b - c * w), 2 * (a * c + b * w)],
        [2 * (a * b + c * w), 1 - 2 * (a * a + c * c), 2 * (b * c - a * w)],
        [2 * (a * c - b * w), 2 * (b * c + a * w), 1 - 2 * (a * a + b * b)],
    ])
    return d @ rot.T


def _frame_from_normal(nor: np.ndarray, hint: np.ndarray) -> tuple:
    """Right-handed orthonormal frame (tan1, tan2, nor) with tan1 ~ hint."""
    t1 = hint - np.sum(hint * nor, axis=1, keepdims=True) * nor
    t1 = _normalize(t1)
    t2 = np.cross(nor, t1)
    t2 = _normalize(t2)
    return t1, t2


def _ellipsoid_area(a: float, b: float, c: float) -> float:
    p = 1.6075
    return 4.0 * math.pi * ((a ** p * b ** p + a ** p * c ** p + b ** p * c ** p) / 3.0) ** (1.0 / p)


def _surface_point(axes, radius, t, th, ph):
    """Point of the shell surface t in [0,1]: (1-t)*ellipsoid + t*sphere,
    both at polar angle th and azimuth ph."""
    a, b, c = axes
    th, ph = np.broadcast_arrays(th, ph)
    t = np.broadcast_to(t, th.shape)[..., None]
    st, ct, sp, cp = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)
    e = np.stack([a * st * cp, b * st * sp, c * ct], axis=-1)
    d = np.stack([st * cp, st * sp, ct], axis=-1)
    return (1.0 - t) * e + t * radius * d


def _latlong_on_surface(m, axes, radius, t, rng, jitter):
    """m points in a jittered latitude/longitude grid that is uniform in
    arc length on the shell surface t: rows equally spaced along the
    (azimuth-averaged) meridian, points equally spaced along each row.
    Returns the points and their (theta, phi) parameters."""
    if m <= 0:
        return np.zeros((0, 3)), np.zeros(0), np.zeros(0)
    th_f = np.linspace(0.0, math.pi, 2049)
    ph_s = np.linspace(0.0, 2.0 * math.pi, 16, endpoint=False)
    pts = _surface_point(axes, radius, t, th_f[:, None], ph_s[None, :])
    seg = np.linalg.norm(np.diff(pts, axis=0), axis=-1).mean(axis=1)
    s_m = np.concatenate([[0.0], np.cumsum(seg)])
    length = s_m[-1]
    # area ~ sum over rows of perimeter * ds; pick the spacing h that gives m
    th_mid = 0.5 * (th_f[1:] + th_f[:-1])
    ph_f = np.linspace(0.0, 2.0 * math.pi, 257)
    ring = _surface_point(axes, radius, t, th_mid[::16, None], ph_f[None, :])
    perim_c = np.linalg.norm(np.diff(ring, axis=1), axis=-1).sum(axis=1)
    area = np.sum(np.interp(th_mid, th_mid[::16], perim_c) * seg)
    h = math.sqrt(area / m)
    n_rows = max(1, int(round(length / h)))
    s_rows = (np.arange(n_rows) + 0.5) * length / n_rows
    th_rows = np.interp(s_rows, s_m, th_f)
    ring = _surface_point(axes, radius, t, th_rows[:, None], ph_f[None, :])
    seglen = np.linalg.norm(np.diff(ring, axis=1), axis=-1)
    perim = seglen.sum(axis=1)
    raw = perim / perim.sum() * m
    cnt = np.maximum(1, np.floor(raw).astype(np.int64))
    while cnt.sum() < m:
        cnt[np.argmax(raw - cnt)] += 1
    while cnt.sum() > m:
        k = np.argmax(np.where(cnt > 1, cnt - raw, -np.inf))
        cnt[k] -= 1
    dth = math.pi / n_rows
    ths, phs = [], []
    for r in range(n_rows):
        k = int(cnt[r])
        cum = np.concatenate([[0.0], np.cumsum(seglen[r])])
        u = (np.arange(k) + 0.5 + rng.uniform(-jitter, jitter, size=k)) / k
        u = (u + rng.uniform(0.0, 1.0 / k)) % 1.0
        ph = np.interp(u * perim[r], cum, ph_f)
        # jitter the row position by a fraction of the local row spacing
        j = rng.uniform(-jitter, jitter, size=k)
        lo = th_rows[max(r - 1, 0)] if r > 0 else 0.0
        hi = th_rows[min(r + 1, n_rows - 1)] if r < n_rows - 1 else math.pi
        th = th_rows[r] + j * (0.5 * (hi - lo) if n_rows > 1 else dth)
        ths.append(np.clip(th, 1e-9, math.pi - 1e-9))
        phs.append(ph)
    th = np.concatenate(ths)
    ph = np.concatenate(phs)
    return _surface_point(axes, radius, t, th, ph), th, ph


def _wall_points(n_wall: int, axes, rng, jitter):
    """Jittered latitude/longitude grid on the ellipsoid, uniform in arc
    length so that the thin rim is sampled as finely as the faces."""
    a, b, c = axes
    p, th, ph = _latlong_on_surface(n_wall, axes, 1.0, 0.0, rng, jitter)
    st, ct = np.sin(th), np.cos(th)
    sp, cp = np.sin(ph), np.cos(ph)
    nor = _normalize(np.stack([st * cp / a, st * sp / b, ct / c], axis=1))
    hint = np.stack([-a * sp, b * cp, np.zeros_like(sp)], axis=1)
    t1, t2 = _frame_from_normal(nor, hint)
    return p, t1, t2, nor


def _shell_plan(n_interior: int, h_wall: float, rho_e: float, radius: float,
                aspect: float = 1.0, h_cap: float = float("inf")):
    """Shell offsets (distance from the wall) and point counts.

    Shell spacing grows geometrically from the wall spacing, h_s = h_w*g^s,
    until it reaches h_cap (the far-field point spacing) and stays there, so
    the last shells match the far-field points and every outer point sees
    interior neighbours as well as far-field ones.  The growth g is chosen
    by bisection so that shells holding aspect*4*pi*rho^2/h^2 points each
    (aspect 1: tangential spacing equal to the radial one) sum to
    n_interior; the shells stop half a spacing short of the far field.
    """
    span = radius - rho_e

    def plan(g):
        deltas, hs = [], []
        d = 0.0
        h = min(h_wall, h_cap)
        while True:
            d_next = d + h
            if d_next > span - 0.5 * h:
                break
            deltas.append(d_next)
            hs.append(h)
            d = d_next
            h = min(h * g, h_cap)
        deltas = np.array(deltas)
        hs = np.array(hs)
        rho = rho_e + deltas
        m = aspect * 4.0 * math.pi * rho * rho / (hs * hs)
        return deltas, hs, m

    lo, hi = 1.0005, 3.0
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        _, _, m = plan(mid)
        if m.sum() > n_interior:
            lo = mid
        else:
            hi = mid
    deltas, hs, m = plan(hi)
    if len(deltas) == 0:
        deltas, hs, m = np.array([0.5 * span]), np.array([span]), np.array([1.0])
    raw = m / m.sum() * n_interior
    cnt = np.floor(raw).astype(np.int64)
    while cnt.sum() < n_interior:
        cnt[np.argmax(raw - cnt)] += 1
    return deltas, hs, cnt


def _interior_points(n_interior, axes, radius, h_wall, rng, jitter, aspect, radial_jitter,
                     h_cap):
    a, b, c = axes
    rho_e = math.sqrt(_ellipsoid_area(a, b, c) / (4.0 * math.pi))
    deltas, hs, cnt = _shell_plan(n_interior, h_wall, rho_e, radius, aspect, h_cap)
    span = radius - rho_e
    pts = []
    for d, h, m in zip(deltas, hs, cnt):
        m = int(m)
        if m == 0:
            continue
        t0 = d / span
        _, th, ph = _latlong_on_surface(m, axes, radius, t0, rng, jitter)
        t = (d + rng.uniform(-radial_jitter, radial_jitter, size=m) * h) / span
        t = np.clip(t, 1e-6, 1.0 - 1e-6)
        pts.append(_surface_point(axes, radius, t, th, ph))
    p = np.concatenate(pts, axis=0) if pts else np.zeros((0, 3))
    return p, len(deltas)


def _outer_points(n_outer, radius, rng, jitter):
    d = _golden_directions(n_outer, rng, jitter)
    hint = np.stack([-d[:, 1], d[:, 0], np.zeros(n_outer)], axis=1)
    small = np.linalg.norm(hint, axis=1) < 1e-8
    hint[small] = np.array([1.0, 0.0, 0.0])
    t1, t2 = _frame_from_normal(d, hint)
    return radius * d, t1, t2, d


def _lattice_interior(n_interior, axes, radius, rng, jitter, margin=0.5):
    """Jittered cubic-lattice interior between the body and the far field.

    The lattice spacing is chosen so that about n_interior nodes fall inside
    the sphere of the given radius and outside the body (kept clear of both
    surfaces by margin*spacing); the node set is then trimmed to exactly
    n_interior points at random.
    """
    a, b, c = axes
    vol = 4.0 / 3.0 * math.pi * (radius ** 3 - a * b * c)
    h = (vol / n_interior) ** (1.0 / 3.0)
    for _ in range(60):
        g = np.arange(-radius, radius + h, h)
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        q = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        q = q + rng.uniform(-jitter, jitter, size=q.shape) * h
        r = np.linalg.norm(q, axis=1)
        # clearance from the ellipsoid measured in the scaled metric
        e = np.sqrt((q[:, 0] / (a + margin * h)) ** 2 + (q[:, 1] / (b + margin * h)) ** 2
                    + (q[:, 2] / (c + margin * h)) ** 2)
        keep = (r < radius - margin * h) & (e > 1.0)
        if keep.sum() >= n_interior:
            q = q[keep]
            sel = np.sort(rng.choice(q.shape[0], n_interior, replace=False))
            return q[sel], h
        h *= 0.97
    raise RuntimeError("lattice interior could not reach the requested size")


def morton_codes(p: np.ndarray, lo: float = -15.0, hi: float = 15.0, bits: int = 21) -> np.ndarray:
    """63-bit Morton (Z-order) codes, 21 bits per axis over [lo, hi]^3."""
    scale = (2 ** bits - 1) / (hi - lo)
    q = np.clip(np.floor((p - lo) * scale), 0, 2 ** bits - 1).astype(np.uint64)

    def spread(v):
        v = v & np.uint64(0x1FFFFF)
        v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
        v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
        v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
        v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
        return v

    return spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1)) | (spread(q[:, 2]) << np.uint64(2))


def _knn(p: np.ndarray, k: int, workers: int):
    from scipy.spatial import cKDTree
    tree = cKDTree(p)
    dist, idx = tree.query(p, k=k + 1, workers=workers)
    n = p.shape[0]
    self_col = idx == np.arange(n)[:, None]
    has_self = self_col.any(axis=1)
    # drop self (normally column 0); rows without self drop their last column
    keep = ~self_col
    keep[~has_self, -1] = False
    nb = idx[keep].reshape(n, k)
    nd = dist[keep].reshape(n, k)
    return nb.astype(np.int64), nd


def _split_counts(p, nb, rows, chunk=262144):
    """Per-row counts of (del <= 0, del > 0) along x, y, z (Cartesian frame)."""
    le = np.empty((rows.size, 3), dtype=np.int32)
    gt = np.empty((rows.size, 3), dtype=np.int32)
    for s0 in range(0, rows.size, chunk):
        r = rows[s0:s0 + chunk]
        d = p[nb[r]] - p[r][:, None, :]
        c = (d <= 0.0).sum(axis=1)
        le[s0:s0 + chunk] = c
        gt[s0:s0 + chunk] = nb.shape[1] - c
    return le, gt


def _repair_interior(p, status, axes, radius, k, workers, min_split, max_passes=40,
                     seed=0):
    """Nudge interior points whose Cartesian half-stencils are too small.

    A half-stencil with fewer than 3 points makes the 3x3 least-squares
    system singular (and the reference then divides by a zero determinant).
    Such points sit where a curved shell puts every same-shell neighbour on
    one side; moving the point toward the starved side by a growing fraction
    of its nearest-neighbour distance (plus a small random component that
    breaks ties between passes) fixes them.  After each pass only the rows
    whose k-nearest lists can have changed are queried again.  Points are
    kept outside the body and inside the far field.
    """
    from scipy.spatial import cKDTree
    a, b, c = axes
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    interior = np.flatnonzero(status == STATUS_INTERIOR)
    is_interior = status == STATUS_INTERIOR
    nb, nd = _knn(p, k, workers)
    row_of = np.full(p.shape[0], -1, dtype=np.int64)
    row_of[interior] = np.arange(interior.size)
    le, gt = _split_counts(p, nb, interior)
    for it in range(max_passes):
        bad = (le < min_split) | (gt < min_split)
        rows = np.flatnonzero(bad.any(axis=1))
        if rows.size == 0:
            return nb, nd, 0
        idx = interior[rows]
        frac = 0.3 + 0.05 * it
        step = frac * nd[idx, 0][:, None]
        move = np.where(le[rows] < min_split, 1.0, 0.0) - np.where(gt[rows] < min_split, 1.0, 0.0)
        move = move + rng.uniform(-0.3, 0.3, size=move.shape)
        newp = p[idx] + step * move
        # a move blocked by the body or the far field pulls the point halfway
        # toward the centroid of its neighbours instead
        inside = (newp[:, 0] / a) ** 2 + (newp[:, 1] / b) ** 2 + (newp[:, 2] / c) ** 2 <= 1.0
        outside = np.linalg.norm(newp, axis=1) >= radius
        blocked = inside | outside
        if blocked.any():
            cen = p[nb[idx[blocked]]].mean(axis=1)
            newp[blocked] = 0.5 * p[idx[blocked]] + 0.5 * cen
        inside = (newp[:, 0] / a) ** 2 + (newp[:, 1] / b) ** 2 + (newp[:, 2] / c) ** 2 <= 1.0
        outside = np.linalg.norm(newp, axis=1) >= radius
        ok = ~(inside | outside)
        moved = idx[ok]
        if moved.size == 0:
            break
        old = p[moved].copy()
        p[moved] = newp[ok]
        # rows to refresh: the moved points, every row that listed one of
        # them, and every row close enough to list one at its new position
        tree = cKDTree(p)
        reach = nd[moved, -1] * 1.5 + np.linalg.norm(p[moved] - old, axis=1)
        near = tree.query_ball_point(p[moved], reach, workers=workers)
        touched = np.zeros(p.shape[0], dtype=bool)
        touched[moved] = True
        for lst in near:
            touched[lst] = True
        touched |= np.isin(nb, moved).any(axis=1)
        rows_q = np.flatnonzero(touched)
        dist, ind = tree.query(p[rows_q], k=k + 1, workers=workers)
        self_col = ind == rows_q[:, None]
        has_self = self_col.any(axis=1)
        keep = ~self_col
        keep[~has_self, -1] = False
        nb[rows_q] = ind[keep].reshape(-1, k)
        nd[rows_q] = dist[keep].reshape(-1, k)
        # only rows whose point or neighbour list changed need new counts
        ri = row_of[rows_q]
        sel = ri >= 0
        le_q, gt_q = _split_counts(p, nb, rows_q[sel])
        le[ri[sel]] = le_q
        gt[ri[sel]] = gt_q
    left = int(((le < min_split) | (gt < min_split)).any(axis=1).sum())
    return nb, nd, left


def _interior_split_ok(p, nb, rows):
    """Rows whose six Cartesian half-stencils all fit in 22 columns."""
    d = p[nb[rows]] - p[rows][:, None, :]
    ok = np.ones(len(rows), dtype=bool)
    for ax in range(3):
        le = (d[:, :, ax] <= 0.0).sum(axis=1)
        gt = (d[:, :, ax] > 0.0).sum(axis=1)
        ok &= (le <= SPLIT_COLS) & (gt <= SPLIT_COLS)
    return ok


def ellipsoid_cloud(n_wall: int, n_outer: int, n_interior: int, k: int = 25,
                    seed: int = 0, axes=(0.5, 0.167, 0.05), radius: float = 15.0,
                    order: str = "blocks", jitter: float = 0.1,
                    shell_aspect: float = 1.0, radial_jitter: float | None = None,
                    min_split: int = 4, interior: str = "shells",
                    workers: int = -1) -> PointCloud:
    """Build an ellipsoid cloud.

    order = "blocks": outer points, then wall, then interior (generation order).
    order = "morton": Z-order over [-15, 15]^3 (21 bits per axis), then an
    outer point is swapped into slot 1 and an interior point into slot N.
    In both orders the last point is an interior point whose split stencils
    fit in 22 columns.
    """
    if k < 1 or k > MAX_CONN:
        raise ValueError(f"k must be in 1..{MAX_CONN}")
    if n_outer < 1 or n_interior < 1 or n_wall < 1:
        raise ValueError("every point class needs at least one point")
    rng = np.random.Generator(np.random.PCG64(seed))
    a, b, c = axes
    pw, t1w, t2w, nw = _wall_points(n_wall, axes, rng, jitter)
    h_wall = math.sqrt(_ellipsoid_area(a, b, c) / n_wall)
    rj = jitter if radial_jitter is None else radial_jitter
    if interior == "shells":
        h_cap = math.sqrt(4.0 * math.pi * radius * radius / n_outer)
        pi_, n_shells = _interior_points(n_interior, axes, radius, h_wall, rng, jitter,
                                         shell_aspect, rj, h_cap)
    elif interior == "lattice":
        pi_, _h = _lattice_interior(n_interior, axes, radius, rng, jitter)
        n_shells = 0
    else:
        raise ValueError(f"unknown interior layout {interior!r}")
    po, t1o, t2o, no = _outer_points(n_outer, radius, rng, jitter)

    eye = np.eye(3)
    p = np.concatenate([po, pw, pi_], axis=0)
    t1 = np.concatenate([t1o, t1w, np.tile(eye[0], (n_interior, 1))], axis=0)
    t2 = np.concatenate([t2o, t2w, np.tile(eye[1], (n_interior, 1))], axis=0)
    nr = np.concatenate([no, nw, np.tile(eye[2], (n_interior, 1))], axis=0)
    status = np.concatenate([np.full(n_outer, STATUS_OUTER), np.full(n_wall, STATUS_WALL),
                             np.full(n_interior, STATUS_INTERIOR)]).astype(np.int32)
    n = p.shape[0]
    if k >= n:
        raise ValueError("k must be smaller than the number of points")

    if min_split > 0:
        nb, nd, left = _repair_interior(p, status, axes, radius, k, workers, min_split,
                                        seed=seed)
        if left:
            raise RuntimeError(f"{left} interior points keep a half-stencil below {min_split}")
    else:
        nb, nd = _knn(p, k, workers)

    # ---- ordering -------------------------------------------------------
    if order == "blocks":
        perm = np.arange(n)
    elif order == "morton":
        perm = np.argsort(morton_codes(p, -radius, radius), kind="stable")
        first_outer = int(np.flatnonzero(status[perm] == STATUS_OUTER)[0])
        perm[[0, first_outer]] = perm[[first_outer, 0]]
    else:
        raise ValueError(f"unknown order {order!r}")
    # last slot: an interior point whose split lists fit (prefer the one
    # already there, else the last eligible interior point in the order)
    cand = np.flatnonzero(status[perm] == STATUS_INTERIOR)
    ok = _interior_split_ok(p, nb, perm[cand])
    if not ok.any():
        raise RuntimeError("no interior point with split stencils of <= 22 entries")
    pos = int(cand[np.flatnonzero(ok)[-1]])
    if pos != n - 1:
        perm[[pos, n - 1]] = perm[[n - 1, pos]]

    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    nb_new = inv[nb[perm]]
    conn = np.zeros((n, MAX_CONN), dtype=np.int32)
    conn[:, :k] = (nb_new + 1).astype(np.int32)
    cloud = PointCloud(
        x=np.ascontiguousarray(p[perm, 0]), y=np.ascontiguousarray(p[perm, 1]),
        z=np.ascontiguousarray(p[perm, 2]),
        tan1=np.ascontiguousarray(t1[perm]), tan2=np.ascontiguousarray(t2[perm]),
        nor=np.ascontiguousarray(nr[perm]), status=np.ascontiguousarray(status[perm]),
        nbhs=np.full(n, k, dtype=np.int32), conn=conn,
        min_dist=np.ascontiguousarray(nd[perm, 0]),
        meta={"seed": seed, "axes": tuple(axes), "radius": radius, "order": order,
              "k": k, "n_shells": n_shells, "jitter": jitter, "shell_aspect": shell_aspect,
              "radial_jitter": rj, "interior": interior,
              "n_wall": n_wall, "n_outer": n_outer, "n_interior": n_interior},
    )
    return cloud


def fixture_cloud(n_interior: int, radius: float = 0.9, axes=(0.5, 0.3, 0.2),
                  wall_refine: float = 1.5, seed: int = 0, order: str = "blocks",
                  k: int = 25) -> PointCloud:
    """Small cloud for golden fixtures: lattice interior, matched boundaries.

    The wall and far-field point counts follow the lattice spacing h (wall
    spacing h/wall_refine, far-field spacing h), so every half-stencil of
    every point holds several neighbours even at a few hundred points.
    """
    a, b, c = axes
    vol = 4.0 / 3.0 * math.pi * (radius ** 3 - a * b * c)
    h = (vol / n_interior) ** (1.0 / 3.0)
    n_outer = int(round(4.0 * math.pi * radius * radius / (h * h)))
    n_wall = int(round(wall_refine * wall_refine * _ellipsoid_area(a, b, c) / (h * h)))
    return ellipsoid_cloud(n_wall, n_outer, n_interior, k=k, seed=seed, axes=axes,
                           radius=radius, order=order, interior="lattice")


# --------------------------------------------------------------------------
# text decks
# --------------------------------------------------------------------------

def _fmt(v: np.ndarray) -> np.ndarray:
    return np.char.mod("%.17g", v)


def write_input_deck(path: str, cloud: PointCloud) -> None:
    """Write ``3d_input_data`` (point_preprocessor.f90:30-36 column order)."""
    n = cloud.n
    with open(path, "w") as f:
        chunk = 200000
        for s in range(0, n, chunk):
            e = min(n, s + chunk)
            cols = [np.char.mod("%d", np.arange(s + 1, e + 1)),
                    np.char.mod("%d", cloud.status[s:e]),
                    _fmt(cloud.x[s:e]), _fmt(cloud.y[s:e]), _fmt(cloud.z[s:e])]
            for arr in (cloud.tan1, cloud.tan2, cloud.nor):
                for c in range(3):
                    cols.append(_fmt(arr[s:e, c]))
            cols.append(np.char.mod("%d", cloud.nbhs[s:e]))
            lines = []
            head = cols[0]
            for c in cols[1:]:
                head = np.char.add(np.char.add(head, " "), c)
            md = _fmt(cloud.min_dist[s:e])
            for r in range(e - s):
                nb = cloud.nbhs[s + r]
                lines.append(head[r] + " " + " ".join(map(str, cloud.conn[s + r, :nb]))
                             + " " + md[r] + "\n")
            f.writelines(lines)


def read_input_deck(path: str) -> PointCloud:
    """Read ``3d_input_data`` written by :func:`write_input_deck` or by hand."""
    xs, ys, zs, t1, t2, nr, st, nbh, conn, md = [], [], [], [], [], [], [], [], [], []
    with open(path) as f:
        toks = f.read().split()
    i = 0
    while i < len(toks):
        i += 1  # counter
        st.append(int(toks[i])); i += 1
        xs.append(float(toks[i])); ys.append(float(toks[i + 1])); zs.append(float(toks[i + 2])); i += 3
        t1.append([float(t) for t in toks[i:i + 3]]); i += 3
        t2.append([float(t) for t in toks[i:i + 3]]); i += 3
        nr.append([float(t) for t in toks[i:i + 3]]); i += 3
        nb = int(toks[i]); i += 1
        if nb < 0 or nb > MAX_CONN:
            raise ValueError(f"point {len(st)}: nbhs={nb} outside 0..{MAX_CONN}")
        row = np.zeros(MAX_CONN, dtype=np.int32)
        row[:nb] = [int(t) for t in toks[i:i + nb]]; i += nb
        conn.append(row); nbh.append(nb)
        md.append(float(toks[i])); i += 1
    return PointCloud(np.array(xs), np.array(ys), np.array(zs), np.array(t1), np.array(t2),
                      np.array(nr), np.array(st, dtype=np.int32), np.array(nbh, dtype=np.int32),
                      np.array(conn, dtype=np.int32).reshape(-1, MAX_CONN), np.array(md))


def write_solution(path: str, cloud: PointCloud, prim: np.ndarray) -> None:
    """Write ``x y z prim(1:5)`` per point (stored-solution.dat / primal-solution.dat)."""
    prim = np.asarray(prim, dtype=np.float64).reshape(cloud.n, 5)
    data = np.column_stack([cloud.x, cloud.y, cloud.z, prim])
    np.savetxt(path, data, fmt="%.17g")


def read_list_directed(path: str, per_record: int, n: int | None = None) -> np.ndarray:
    """Values of a file of list-directed records, ``per_record`` values each.

    Like a Fortran list-directed READ, values are taken in order whatever
    the line breaks (a Fortran list-directed WRITE may wrap one record over
    several lines); separators are blanks or commas and ``r*c`` stands for
    r copies of c."""
    vals = []
    with open(path) as f:
        for tok in f.read().replace(",", " ").split():
            if "*" in tok:
                r, c = tok.split("*", 1)
                vals.extend([float(c)] * int(r))
            else:
                vals.append(float(tok))
    if len(vals) % per_record:
        raise ValueError(f"{path}: {len(vals)} values do not form records of {per_record}")
    data = np.array(vals, dtype=np.float64).reshape(-1, per_record)
    if n is not None and data.shape[0] != n:
        raise ValueError(f"{path}: expected {n} records, found {data.shape[0]}")
    return data


def read_solution(path: str, n: int | None = None) -> np.ndarray:
    """Read the prim columns of stored-solution.dat / primal-solution.dat -> (N, 5)
    (records x, y, z, prim(1:5): initial_conditions.f90:26, post_processor.f90:34)."""
    return np.ascontiguousarray(read_list_directed(path, 8, n)[:, 3:8])


def save_npz(path: str, cloud: PointCloud) -> None:
    np.savez_compressed(path, x=cloud.x, y=cloud.y, z=cloud.z, tan1=cloud.tan1,
                        tan2=cloud.tan2, nor=cloud.nor, status=cloud.status,
                        nbhs=cloud.nbhs, conn=cloud.conn, min_dist=cloud.min_dist)


def load_npz(path: str) -> PointCloud:
    d = np.load(path, allow_pickle=False)
    return PointCloud(d["x"], d["y"], d["z"], d["tan1"], d["tan2"], d["nor"],
                      d["status"].astype(np.int32), d["nbhs"].astype(np.int32),
                      d["conn"].astype(np.int32), d["min_dist"])


# --------------------------------------------------------------------------
# the configurations the solver is exercised on
# --------------------------------------------------------------------------

CONFIGS = {
    # 50k explicit: M 0.6, AoA 2 deg, CFL 0.2, 1000 iterations
    "c1_50k": dict(n_wall=4000, n_outer=2000, n_interior=44000, seed=1, order="blocks",
                   mach=0.6, aoa=2.0, cfl=0.2, iters=1000, mode="explicit"),
    # 2M explicit: M 0.84, AoA 3.06 deg, CFL 0.2, 500 iterations
    "c2_2m": dict(n_wall=120000, n_outer=40000, n_interior=1840000, seed=2, order="blocks",
                  mach=0.84, aoa=3.06, cfl=0.2, iters=500, mode="explicit"),
    # 8M implicit LU-SGS: M 0.84, AoA 3.06 deg, CFL 50, 300 iterations, Morton order
    "c3_8m": dict(n_wall=400000, n_outer=120000, n_interior=7480000, seed=3, order="morton",
                  mach=0.84, aoa=3.06, cfl=50.0, iters=300, mode="implicit"),
}


def config_cloud(name: str, workers: int = -1) -> PointCloud:
    c = CONFIGS[name]
    return ellipsoid_cloud(c["n_wall"], c["n_outer"], c["n_interior"], seed=c["seed"],
                           order=c["order"], workers=workers)


# batched configuration: many independent clouds iterated together
BATCH = dict(n_clouds=64, n_min=20000, n_max=500000, aspect=(2.0, 6.0), mach=(0.3, 0.85),
             aoa=(-5.0, 10.0), seed=5, cfl=0.2, mode="explicit")


def batch_plan(n_clouds: int = BATCH["n_clouds"], seed: int = BATCH["seed"],
               n_min: int = BATCH["n_min"], n_max: int = BATCH["n_max"]):
    """Sizes (log-uniform in [n_min, n_max]), body axes (aspect a:b uniform
    in 2..6 with a = 0.5 and thickness c = 0.3 b, as the 0.5/0.167/0.05
    body), Mach and angle of attack of each cloud, all from one seed.  Each
    cloud has 8% wall, 4% outer and 88% interior points, the proportions of
    the 50k configuration."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_clouds)).astype(np.int64)
    aspect = rng.uniform(*BATCH["aspect"], n_clouds)
    mach = rng.uniform(*BATCH["mach"], n_clouds)
    aoa = rng.uniform(*BATCH["aoa"], n_clouds)
    plan = []
    for b in range(n_clouds):
        n_wall = int(round(0.08 * n[b]))
        n_outer = int(round(0.04 * n[b]))
        b_ax = 0.5 / aspect[b]
        plan.append(dict(n_wall=n_wall, n_outer=n_outer, n_interior=int(n[b]) - n_wall - n_outer,
                         axes=(0.5, b_ax, 0.3 * b_ax), mach=float(mach[b]), aoa=float(aoa[b]),
                         seed=seed * 1000 + b))
    return plan


def batch_clouds(plan, workers: int = -1):
    return [ellipsoid_cloud(p["n_wall"], p["n_outer"], p["n_interior"], seed=p["seed"], axes=p["axes"],
                            order="morton", workers=workers) for p in plan]


if __name__ == "__main__":  # pragma: no cover - manual deck writer
    import argparse
    ap = argparse.ArgumentParser(description="write a synthetic ellipsoid deck")
    ap.add_argument("--wall", type=int, required=True)
    ap.add_argument("--outer", type=int, required=True)
    ap.add_argument("--interior", type=int, required=True)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--order", default="blocks")
    ap.add_argument("--mach", type=float, default=0.6)
    ap.add_argument("--aoa", type=float, default=2.0)
    ap.add_argument("--out", default=".")
    args = ap.parse_args()
    cl = ellipsoid_cloud(args.wall, args.outer, args.interior, seed=args.seed, order=args.order)
    os.makedirs(args.out, exist_ok=True)
    write_input_deck(os.path.join(args.out, "3d_input_data"), cl)
    prim = np.tile(freestream_prim(args.mach, args.aoa), (cl.n, 1))
    write_solution(os.path.join(args.out, "stored-solution.dat"), cl, prim)
    print(cl.counts(), cl.meta)
