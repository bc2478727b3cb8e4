"""Algorithmic HBM bytes of each kernel, from the cloud's real stencil sizes.

Counting rule: every per-point array a kernel reads is counted once per point
it is read for (neighbour gathers are assumed served on chip), every array it
writes once; FP64 = 8 bytes, int32 = 4 bytes.  The sums use the actual split
list lengths, so wall and outer points (whose z stencil runs over the full
neighbour list) are counted as they are.

Per point (FP64 path):
  conserved+q      prim 40 + frame 72 read, U 40 + q 40 written          = 192
  qderiv+timestep  pos 32, nbhs 4, conn 4*nbhs, q 40, prim 40 read,
                   qm 80, dq 120, delt 8 written                        = 324 + 4*nbhs
  flux residual    index 4, pos 32, frame 72, q 40, qm 80, dq 120 read,
                   4 bytes per list count and per list entry read,
                   flux_res 40 written                                  = 388 + 4*(counts+entries)
  explicit sweep   delt 8, flux_res 40 read, delUn 40, delUp 40 written = 128
  state update     status 4, respos 4, U 40, delUn 40, frame 72 read,
                   prim 40, delUp 40 written (outer: prim, frame, delt,
                   flux_res read, prim written)                         = 240 (outer 204)
"""

from __future__ import annotations

import numpy as np

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_VALU_TFLOPS = 78.6        # MI355X FP64 vector


def flux_bytes_per_point(status, split_counts, nbhs):
    """Algorithmic bytes of eval_flux_residual for each point (array)."""
    status = np.asarray(status)
    cnt = np.asarray(split_counts, dtype=np.int64)
    nbhs = np.asarray(nbhs, dtype=np.int64)
    base = 4 + 32 + 72 + 40 + 80 + 120 + 40
    interior = base + 4 * 6 + 4 * cnt.sum(axis=0)
    wall = base + 4 * 4 + 4 * cnt[:4].sum(axis=0) + 4 + 4 * nbhs
    outer = base + 4 * 5 + 4 * cnt[:5].sum(axis=0)
    return np.where(status == 0, interior, np.where(status == 1, wall,
                                                     np.where(status == 2, outer, 0)))


def sweep_entries(status, split_counts, nbhs):
    """List entries one sweep walks per point: every direction list (wall:
    four split lists and conn for z), whether the entry is taken or not."""
    status = np.asarray(status)
    cnt = np.asarray(split_counts, dtype=np.int64)
    nbhs = np.asarray(nbhs, dtype=np.int64)
    return np.where(status == 0, cnt.sum(axis=0), cnt[:4].sum(axis=0) + nbhs)


def sweep_bytes_per_point(status, split_counts, nbhs):
    """Algorithmic bytes of one sweep (LdelU or UdelU pass) for each point:
    index 4, pos 32, frame 72, alias 4, dinv 8, list counts 24, per list
    entry the entry and the neighbour's alias (8), flux_res or delUp 40
    read, the new delta 40 written = 224 + 8*entries; 0 for points the
    canonical sweeps do not visit (outer)."""
    status = np.asarray(status)
    swept = (status == 0) | (status == 1)
    return np.where(swept, 224 + 8 * sweep_entries(status, split_counts, nbhs), 0)


def iteration_bytes(status, split_counts, nbhs, mode="explicit"):
    """Per-kernel algorithmic bytes of one whole iteration (dict of totals)."""
    status = np.asarray(status)
    nbhs = np.asarray(nbhs, dtype=np.int64)
    fb = flux_bytes_per_point(status, split_counts, nbhs)
    n = status.size
    out = {
        "conserved_q": 192 * n,
        "qderiv_timestep": int(np.sum(324 + 4 * nbhs)),
        "flux_interior": int(fb[status == 0].sum()),
        "flux_boundary": int(fb[status != 0].sum()),
        "state_update": int(240 * np.sum((status == 0) | (status == 1)) + 204 * np.sum(status == 2)),
    }
    if mode == "explicit":
        out["sweeps"] = 128 * n
    else:
        cnt = np.asarray(split_counts, dtype=np.int64)
        swept = (status == 0) | (status == 1)
        entries = np.where(status == 0, cnt.sum(axis=0), cnt[:4].sum(axis=0) + nbhs)
        # D_inv: index, pos, frame, prim, delt, counts, entries; dinv written.
        dinv = 4 + 32 + 72 + 40 + 8 + 24 + 4 * entries + 8
        # each sweep: index, pos, frame, alias, dinv, counts, entries, the
        # neighbours' alias (4 B/entry), flux_res or delUp read, delta written
        sweep = 4 + 32 + 72 + 4 + 8 + 24 + 8 * entries + 40 + 40
        out["sweeps"] = int(np.sum((dinv + 2 * sweep)[swept]) + 80 * n)
    out["total"] = int(sum(out.values()))
    return out
