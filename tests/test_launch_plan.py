"""The launch plan of a build (gaussium_amd/csrc/gx_plan.h), compiled for the
host with g++: which class combinations run, in what order, on which stream.

* Timed (every active combination has an autotuned time, side streams in
  use): longest first, each onto the stream with the least time queued; the
  caller's stream (-1) is the lane after the side streams.  Checked against
  schedules worked by hand, and on random cases against the list-scheduling
  bound (makespan <= total / lanes + longest job) and the invariants (every
  active combination once, non-increasing times).
* Untimed (any active time 0, or nside <= 1): class order, round robin over
  the side streams; nside = 0 puts everything on the caller's stream.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT

NCLS = 6                       # pair classes (00) (10) (11) (20) (21) (22)
NCOMBO = NCLS * (NCLS + 1) // 2
XY = [(X, Y) for X in range(NCLS) for Y in range(X + 1)]


@pytest.fixture(scope='module')
def plan():
    build = os.path.join(ROOT, 'tests', 'host_harness', '_build')
    os.makedirs(build, exist_ok=True)
    so = os.path.join(build, 'plan_host.%d.so' % os.getpid())  # one per process (parallel test workers)
    subprocess.run(['g++', '-O1', '-std=c++17', '-shared', '-fPIC', '-I', os.path.join(ROOT, 'gaussium_amd', 'csrc'),
                    os.path.join(ROOT, 'tests', 'host_harness', 'plan_host.cpp'), '-o', so], check=True)
    lib = ctypes.CDLL(so)
    os.remove(so)
    lib.plan_case.restype = ctypes.c_int

    def run(n_items, tms, mode, nside):
        n = np.ascontiguousarray(n_items, dtype=np.int32)
        t = np.ascontiguousarray(tms, dtype=np.float64)
        out = np.zeros(4 * NCOMBO, dtype=np.int32)
        k = lib.plan_case(NCLS, n.ctypes.data_as(ctypes.c_void_p), len(n), t.ctypes.data_as(ctypes.c_void_p), len(t),
                          mode, nside, out.ctypes.data_as(ctypes.c_void_p))
        return [tuple(int(x) for x in out[4 * i:4 * i + 4]) for i in range(k)]
    return run


def _six(ms, mode=1):
    """The six s/p combinations of a 220-function cluster with times `ms`."""
    n = np.zeros(NCOMBO, dtype=np.int32)
    n[:6] = 10
    t = np.zeros(2 * NCOMBO)
    t[mode * NCOMBO:mode * NCOMBO + 6] = ms
    return n, t


def test_timed_plan_matches_hand_schedule(plan):
    # 113 -> 0, 89 -> 1, 51 -> 2, 36 -> caller (empty), 28 -> caller (36 < 51),
    # 26 -> 2 (51 < 64)
    n, t = _six([113, 89, 51, 36, 28, 26])
    assert plan(n, t, 1, 3) == [(0, 0, 0, 0), (1, 1, 0, 1), (2, 1, 1, 2), (3, 2, 0, -1), (4, 2, 1, -1), (5, 2, 2, 2)]
    # the order follows the times, not the classes; ties keep class order and
    # go to the lowest lane, side streams before the caller's
    n, t = _six([5, 40, 5, 40, 20, 20])
    assert plan(n, t, 1, 3) == [(1, 1, 0, 0), (3, 2, 0, 1), (4, 2, 1, 2), (5, 2, 2, -1), (0, 0, 0, 2), (2, 1, 1, -1)]


def test_untimed_plans_round_robin_in_class_order(plan):
    n, t = _six([113, 89, 0.0, 36, 28, 26])          # one combination never timed
    assert plan(n, t, 1, 3) == [(k, *XY[k], k % 3) for k in range(6)]
    n, t = _six([113, 89, 51, 36, 28, 26], mode=1)   # timed in the mixed mode only
    assert plan(n, t, 0, 3) == [(k, *XY[k], k % 3) for k in range(6)]
    n, t = _six([113, 89, 51, 36, 28, 26])
    assert plan(n, t, 1, 0) == [(k, *XY[k], -1) for k in range(6)]   # no side streams (capture, timing)
    assert plan(n, t, 1, 1) == [(k, *XY[k], 0) for k in range(6)]    # one side stream: nothing to balance


def test_inactive_combinations_are_skipped(plan):
    n, t = _six([113, 89, 51, 36, 28, 26])
    n[2] = 0
    n[20] = 3                                         # (22|22) active but untimed
    assert [s[0] for s in plan(n, t, 1, 3)] == [0, 1, 3, 4, 5, 20]
    t[NCOMBO + 20] = 7.0
    got = plan(n, t, 1, 3)
    assert [s[0] for s in got] == [0, 1, 3, 4, 5, 20]
    assert all((s[1], s[2]) == XY[s[0]] for s in got)


@pytest.mark.parametrize('seed', range(20))
def test_random_timed_plans_hold_list_scheduling_bound(plan, seed):
    rng = np.random.default_rng(seed)
    n = (rng.random(NCOMBO) < 0.7).astype(np.int32) * rng.integers(1, 100, NCOMBO).astype(np.int32)
    n[0] = 1
    t = np.zeros(2 * NCOMBO)
    t[:NCOMBO] = rng.lognormal(3.0, 1.0, NCOMBO)
    nside = int(rng.integers(2, 5))
    got = plan(n, t, 0, nside)
    act = [k for k in range(NCOMBO) if n[k] > 0]
    assert sorted(s[0] for s in got) == act
    ms = [t[s[0]] for s in got]
    assert all(a >= b for a, b in zip(ms, ms[1:]))
    load = {}
    for s in got:
        assert -1 <= s[3] < nside
        load[s[3]] = load.get(s[3], 0.0) + t[s[0]]
    lanes = nside + 1
    assert max(load.values()) <= sum(ms) / lanes + max(ms) + 1e-9
