"""Shared test setup.

Markers:
  gpu   needs an AMD GPU and the built HIP library (run on the GPU box).
Everything else runs on the CPU: the oracle against the golden fixtures,
closed-form checks, the host-side preprocessing and the C-ABI export list.
"""

import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

PKG = importlib.import_module("3d-mfcfd_amd")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


def load_fixture(name):
    """Golden fixture -> (PointCloud, dict of arrays)."""
    d = dict(np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False))
    pc = PKG.pointcloud
    cl = pc.PointCloud(d["x"], d["y"], d["z"], d["tan1"], d["tan2"], d["nor"],
                       d["status"].astype(np.int32), d["nbhs"].astype(np.int32),
                       d["conn"].astype(np.int32), d["min_dist"])
    return cl, d


# the canonical configurations (power 0, no inner iterations, statuses 0/1/2)
FIXTURE_NAMES = ["explicit_blocks", "explicit_morton", "implicit_morton"]
# branches the canonical configurations never reach: supersonic points
# (status 5/6), power = 2 with one inner q-derivative iteration
BRANCH_FIXTURES = ["explicit_supersonic", "implicit_supersonic", "explicit_power2_inner1",
                   "implicit_power2_inner1"]
# the alternative implicit pairing (..._backup.f90 + state_update_implicit.f90)
BACKUP_FIXTURES = ["implicit_backup"]
ALL_FIXTURES = FIXTURE_NAMES + BRANCH_FIXTURES + BACKUP_FIXTURES


def fixture_params(d):
    """(power, inner_iterations) a fixture was written with"""
    power = float(d["power"]) if "power" in d else 0.0
    inner = int(d["inner_iterations"]) if "inner_iterations" in d else 0
    return power, inner


@pytest.fixture(scope="session")
def pkg():
    return PKG
