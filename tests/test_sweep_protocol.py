"""The sync-free sweep's progress protocol, modelled on the CPU.

k_sweep_async (kernels_sweep.hip) runs a resident grid of workgroups of eight
waves.  Each round, thread 0 of a workgroup takes eight consecutive execution
positions with one atomic add on a ticket counter, a workgroup barrier
follows, wave w works on position base + w -- it waits, for every wait entry
of its take list, until the entry's point has published -- then publishes,
and the workgroup meets at the barrier again before its next round.  A
ticket past the last position ends the workgroup.

This model replays that protocol step by step on real sweep plans (SweepPlan
of preprocess.cpp, through the shim in tests/sanitize): both directions, the
canonical and the alternative driver, grids of 1 to 64 workgroups, under
random interleavings of the workgroups' steps.  Every run must complete: at
every moment some step can make progress.  A wait ends only once its
producer has published, so a completed run also shows that no position
needs a value that is never produced.  A planted mistake -- one consumer
moved ahead of its producer in the execution order -- must deadlock a
one-workgroup grid, so the check is not vacuous.
"""

import ctypes as C
import os
import random
import shutil
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT, load_fixture

pc = PKG.pointcloud
solver = PKG.solver
SAN = os.path.join(ROOT, "tests", "sanitize")
LIB = os.path.join(SAN, "_build", "libpreprocess_plain.so")
TAKE_WAIT, TAKE_SHIFT = 8, 4
WAVES = 8

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


MAX_LATE = 4


def lib():
    if not os.path.exists(LIB):
        if shutil.which("make") is None:
            pytest.skip("make is not available")
        r = subprocess.run(["make", "-s", "-C", SAN, "_build/libpreprocess_plain.so"], capture_output=True,
                           text=True)
        assert r.returncode == 0, r.stdout + r.stderr
    L = C.CDLL(LIB)
    L.san_sweep_plan.argtypes = [C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp, C.c_int,
                                 C.c_longlong, _ip, _lp, _up, _lp]
    L.san_sweep_plan.restype = C.c_int
    L.san_sweep_plan_late.argtypes = L.san_sweep_plan.argtypes + [
        np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")]
    L.san_sweep_plan_late.restype = C.c_int
    return L


def plans(cl, backup):
    """Per direction: list of the dependency positions of every position."""
    n = cl.n
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
    i = lambda a: np.ascontiguousarray(a, dtype=np.int32).ravel()
    cnt, lists = solver.split_stencils(cl)
    alias, pwa = solver.alias_order(cl)
    cap = 200 * n
    seq = np.zeros(2 * n, dtype=np.int32)
    take_ptr = np.zeros(2 * (n + 1), dtype=np.int64)
    take = np.zeros(2 * cap, dtype=np.uint32)
    sizes = np.zeros(6, dtype=np.int64)
    rc = lib().san_sweep_plan(n, i(cl.status), i(cl.nbhs), i(cl.conn), i(cnt), i(lists), i(alias), i(pwa),
                              f(cl.x), f(cl.y), f(cl.z), int(backup), cap, seq, take_ptr, take, sizes)
    assert rc == 0
    out = []
    for dr in range(2):
        ns, ntp, nt = (int(v) for v in sizes[3 * dr:3 * dr + 3])
        s = seq[dr * n:dr * n + ns]
        tp = take_ptr[dr * (n + 1):dr * (n + 1) + ntp]
        tk = take[dr * cap:dr * cap + nt]
        pos = np.full(n, -1, dtype=np.int64)
        pos[s] = np.arange(ns)
        deps = []
        for u in range(ns):
            e = tk[tp[u]:tp[u + 1]]
            w = e[(e & TAKE_WAIT) != 0] >> TAKE_SHIFT
            deps.append(sorted(set(int(v) for v in pos[w.astype(np.int64)])))
        out.append(deps)
    return out


def simulate(deps, groups, seed):
    """Random interleaving of the ticket/barrier/poll protocol.  Returns the
    number of positions published before no step could proceed."""
    rng = random.Random(seed)
    P = len(deps)
    missing = [len(d) for d in deps]
    consumers = [[] for _ in range(P)]
    for u, d in enumerate(deps):
        for v in d:
            consumers[v].append(u)
    published = [False] * P
    assigned = [False] * P
    ctr = 0
    taking = list(range(groups))      # workgroups about to take a ticket
    round_left = {}                   # workgroup -> positions of its round not yet published
    owner = {}
    ready = []                        # assigned, unpublished positions whose dependencies all published
    done = 0
    while True:
        actions = len(taking) + len(ready)
        if actions == 0:
            return done
        a = rng.randrange(actions)
        if a < len(taking):
            g = taking.pop(a)
            base = WAVES * ctr
            ctr += 1
            if base >= P:
                continue              # past the end: the workgroup ends
            mine = [u for u in range(base, min(base + WAVES, P))]
            round_left[g] = len(mine)
            for u in mine:
                assigned[u] = True
                owner[u] = g
                if missing[u] == 0:
                    ready.append(u)
        else:
            k = a - len(taking)
            ready[k], ready[-1] = ready[-1], ready[k]
            u = ready.pop()
            published[u] = True
            done += 1
            for c in consumers[u]:
                missing[c] -= 1
                if missing[c] == 0 and assigned[c] and not published[c]:
                    ready.append(c)
            g = owner[u]
            round_left[g] -= 1
            if round_left[g] == 0:    # the barrier: the workgroup's next round
                taking.append(g)


def check_plan(deps):
    # the plan's own guarantee: every wait reads an earlier position
    assert all(v < u for u, d in enumerate(deps) for v in d)
    for groups in (1, 3, 64):
        for seed in (1, 2):
            assert simulate(deps, groups, seed) == len(deps), (groups, seed)


@pytest.mark.parametrize("name,backup", [("implicit_morton", False), ("implicit_backup", True),
                                         ("implicit_supersonic", False)])
def test_protocol_completes_on_fixture_plans(name, backup):
    cl, _ = load_fixture(name)
    for deps in plans(cl, backup):
        check_plan(deps)


def test_protocol_completes_on_a_20k_plan():
    cl = pc.ellipsoid_cloud(1600, 800, 17600, seed=22, order="morton")
    fwd, bwd = plans(cl, False)
    assert sum(len(d) for d in fwd) > 0 and sum(len(d) for d in bwd) > 0
    check_plan(fwd)
    check_plan(bwd)


def test_a_consumer_ahead_of_its_producer_deadlocks():
    cl, _ = load_fixture("implicit_morton")
    deps = plans(cl, False)[0]
    P = len(deps)
    # a consumer whose producer sits at least one round earlier, moved to
    # the very front of the order
    u = next(u for u, d in enumerate(deps) if d and d[-1] >= WAVES)
    order = [u] + [v for v in range(P) if v != u]
    where = {v: i for i, v in enumerate(order)}
    planted = [sorted(where[v] for v in deps[o]) for o in order]
    assert simulate(planted, 1, 1) < P


@pytest.mark.parametrize("name,backup", [("implicit_morton", False), ("implicit_backup", True),
                                         ("implicit_supersonic", False)])
def test_late_entries_read_the_dependency_that_executes_last(name, backup):
    """Each position's last late[u] entries are wait entries of one point, the
    wait dependency that executes last; no other entry of the position reads
    it; a position sets none apart only when it has no regular wait entry or
    that point fills more than MAX_LATE entries.  The sweep kernel sums the
    other entries before it waits for these."""
    cl, _ = load_fixture(name)
    n = cl.n
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
    i = lambda a: np.ascontiguousarray(a, dtype=np.int32).ravel()
    cnt, lists = solver.split_stencils(cl)
    alias, pwa = solver.alias_order(cl)
    cap = 200 * n
    seq = np.zeros(2 * n, dtype=np.int32)
    take_ptr = np.zeros(2 * (n + 1), dtype=np.int64)
    take = np.zeros(2 * cap, dtype=np.uint32)
    sizes = np.zeros(6, dtype=np.int64)
    late = np.zeros(2 * n, dtype=np.uint8)
    rc = lib().san_sweep_plan_late(n, i(cl.status), i(cl.nbhs), i(cl.conn), i(cnt), i(lists), i(alias), i(pwa),
                                   f(cl.x), f(cl.y), f(cl.z), int(backup), cap, seq, take_ptr, take, sizes, late)
    assert rc == 0
    set_apart = 0
    for dr in range(2):
        ns, ntp, nt = (int(v) for v in sizes[3 * dr:3 * dr + 3])
        s_ = seq[dr * n:dr * n + ns]
        tp = take_ptr[dr * (n + 1):dr * (n + 1) + ntp]
        tk = take[dr * cap:dr * cap + nt].astype(np.int64)
        lt = late[dr * n:dr * n + ns]
        pos = np.full(n, -1, dtype=np.int64)
        pos[s_] = np.arange(ns)
        for u in range(ns):
            e = tk[tp[u]:tp[u + 1]]
            regular_wait = ((e & TAKE_WAIT) != 0) & ((e & 7) != 7)
            nl = int(lt[u])
            if not regular_wait.any():
                assert nl == 0
                continue
            k = e >> TAKE_SHIFT
            k_last = k[regular_wait][np.argmax(pos[k[regular_wait]])]
            reads_last = regular_wait & (k == k_last)
            if nl == 0:
                assert reads_last.sum() > MAX_LATE, u
                continue
            set_apart += 1
            assert 1 <= nl <= MAX_LATE
            assert reads_last.sum() == nl
            assert reads_last[len(e) - nl:].all()
    assert set_apart > 0

