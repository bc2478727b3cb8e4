"""Host preprocessing under AddressSanitizer and UndefinedBehaviorSanitizer.

Collected only when named (tests/test_sanitized.py runs it in a Python
process that preloads the gcc ASan runtime): the sanitizer build of
preprocess.cpp (tests/sanitize/Makefile) reproduces the reference's split
stencils and alias order on every golden fixture, and its sweep plans --
canonical and alternative drivers -- keep their invariants: each visited
point is executed once, every entry names a point of the cloud and a
direction, and every entry that reads this sweep's value (a wait entry)
reads a point executed earlier.  A sanitizer report aborts the process.
"""

import ctypes as C
import os

import numpy as np
import pytest

from conftest import ALL_FIXTURES, load_fixture

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libpreprocess_san.so")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")

TAKE_WAIT, TAKE_COPY, TAKE_SHIFT = 8, 7, 4


def lib():
    assert "libasan" in open("/proc/self/maps").read(), "the ASan runtime is not loaded"
    L = C.CDLL(LIB)
    L.san_split_stencils.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _ip]
    L.san_aliasing.argtypes = [C.c_int, _ip, _ip, _ip, _ip]
    L.san_sweep_plan.argtypes = [C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp, C.c_int,
                                 C.c_longlong, _ip, _lp, _up, _lp]
    L.san_sweep_plan.restype = C.c_int
    return L


def arrays(cl):
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
    i = lambda a: np.ascontiguousarray(a, dtype=np.int32).ravel()
    return (f(cl.x), f(cl.y), f(cl.z), f(cl.tan1), f(cl.tan2), f(cl.nor), i(cl.status), i(cl.nbhs),
            i(cl.conn))


def test_the_oracle_loaded_is_the_sanitizer_build():
    import orc
    assert orc.LIB_PATH == os.environ.get("ORC_LIB") and orc.LIB_PATH.endswith("liborc_san.so")
    L = orc.lib()
    assert hasattr(L, "__asan_report_load8")   # instrumented: resolves through its ASan dependency


@pytest.mark.parametrize("name", ALL_FIXTURES)
def test_preprocessing_under_sanitizers(name):
    L = lib()
    cl, d = load_fixture(name)
    n = cl.n
    x, y, z, t1, t2, nor, status, nbhs, conn = arrays(cl)
    cnt = np.zeros(6 * n, dtype=np.int32)
    lists = np.zeros(6 * (22 * n + 32), dtype=np.int32)
    L.san_split_stencils(n, x, y, z, t1, t2, nor, status, nbhs, conn, cnt, lists)
    m = d["split_lists"].shape[1]
    assert np.array_equal(cnt.reshape(6, n), d["split_counts"])
    assert np.array_equal(lists.reshape(6, -1)[:, :m], d["split_lists"])
    alias = np.zeros(n, dtype=np.int32)
    pwa = np.zeros(n, dtype=np.int32)
    L.san_aliasing(n, nbhs, conn, alias, pwa)
    assert np.array_equal(alias, d["alias"]) and np.array_equal(pwa, d["point_with_alias"])

    for backup in (0, 1):
        cap = 200 * n
        seq = np.zeros(2 * n, dtype=np.int32)
        take_ptr = np.zeros(2 * (n + 1), dtype=np.int64)
        take = np.zeros(2 * cap, dtype=np.uint32)
        sizes = np.zeros(6, dtype=np.int64)
        rc = L.san_sweep_plan(n, status, nbhs, conn, cnt, lists, alias, pwa, x, y, z, backup, cap, seq,
                              take_ptr, take, sizes)
        assert rc == 0
        for dr in range(2):
            ns, ntp, nt = sizes[3 * dr:3 * dr + 3]
            s = seq[dr * n:dr * n + ns]
            tp = take_ptr[dr * (n + 1):dr * (n + 1) + ntp]
            tk = take[dr * cap:dr * cap + nt]
            assert len(np.unique(s)) == ns and ntp == ns + 1 and tp[0] == 0 and tp[-1] == nt
            assert np.all(np.diff(tp) >= 0)
            pos = np.full(n, -1)
            pos[s] = np.arange(ns)
            k = (tk >> TAKE_SHIFT).astype(np.int64)
            dirn = tk & 7
            assert np.all(k < n) and np.all((dirn < 6) | (dirn == TAKE_COPY))
            owner = np.repeat(np.arange(ns), np.diff(tp))
            wait = (tk & TAKE_WAIT) != 0
            assert np.all(pos[k[wait & (dirn != TAKE_COPY)]] < owner[wait & (dirn != TAKE_COPY)])
            assert np.all(pos[k[wait & (dirn != TAKE_COPY)]] >= 0)
