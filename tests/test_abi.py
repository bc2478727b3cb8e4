"""The C ABI library without a GPU: it loads, exports every function that
include/mfcfd_hip.h declares, reports its ABI version, and its host-side
preprocessing drop-ins (generate_split_stencils, aliasing) reproduce the
reference's arrays bit for bit."""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import FIXTURE_NAMES, PKG, ROOT, load_fixture

solver = PKG.solver
HEADER = os.path.join(ROOT, "include", "mfcfd_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(mfcfd_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(solver.LIB_PATH)
    missing = [nm for nm in names if not hasattr(lib, nm)]
    assert not missing, missing
    # the Python mirror binds exactly these
    assert sorted(solver.EXPORTS) == names


def test_abi_version_matches_header():
    ver = int(re.search(r"#define MFCFD_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    assert solver.lib().mfcfd_abi_version() == ver


@pytest.mark.parametrize("name", FIXTURE_NAMES)
def test_host_preprocessing_matches_reference(name):
    cl, d = load_fixture(name)
    cnt, lists = solver.split_stencils(cl)
    assert np.array_equal(cnt, d["split_counts"])
    m = d["split_lists"].shape[1]   # the reference's 22*N
    assert np.array_equal(lists[:, :m], d["split_lists"])
    assert not lists[:, m:].any()   # nothing spills past the last point
    alias, pwa = solver.alias_order(cl)
    assert np.array_equal(alias, d["alias"])
    assert np.array_equal(pwa, d["point_with_alias"])
