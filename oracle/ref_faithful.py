"""CPU baseline: the reference's ERI + J/K path restated in pure Python, timed on a sample.

TEST / BASELINE INFRASTRUCTURE ONLY (``bench.py``'s ``cpu_baseline`` leg).

What is timed follows the reference:
* ERI build: every canonical basis-function quartet evaluated by the
  recursive Obara-Saika scheme without a cross-quartet memo
  (integrals/twoelectronrepulsion/obara_saika_scheme.py:17-195), farmed out
  with ``multiprocessing.Pool(procs).starmap`` over the key list
  (matrixelements/electron_repulsion_matrix.py:35-39);
* dense fill: N^4 ``sort_index`` + dictionary lookups (:43-45,
  common/symmetry.py:186-197);
* one J/K (Fock) build: the per-element Python loop of
  hartreefock/fock_matrix.py:12-22.

The full workload at 220 functions needs ~3e8 quartets, so each part is
timed on a seeded sample and scaled (labelled "extrapolated"):
ERI time = mean(worker seconds per sampled quartet) x total quartets / procs
+ measured Pool start-up; fill and Fock loops are timed on a subset of rows.
"""
import itertools
import multiprocessing as mp
import os
import time

import numpy as np

from .basis import sort_index
from .os_eri import eri_contracted

_BASIS = None


def _init(basis):
    global _BASIS
    _BASIS = basis


def _eri_timed(i, j, k, l):
    t0 = time.perf_counter()
    v = eri_contracted(_BASIS[i], _BASIS[j], _BASIS[k], _BASIS[l], memo=False)
    return v, time.perf_counter() - t0


def sample_canonical(n, count, rng):
    out = []
    for _ in range(count):
        a, b, c, d = (int(x) for x in rng.integers(0, n, 4))
        a, b, c, d = sort_index(a, b, c, d)
        out.append((a, b, c, d))
    return out


def n_canonical(n):
    npair = n * (n + 1) // 2
    return npair * (npair + 1) // 2


def time_eri(basis, seconds=12.0, procs=None, seed=0, batch=64):
    """Worker-seconds per quartet (mean over a random canonical sample) and Pool start-up."""
    procs = procs or len(os.sched_getaffinity(0))
    rng = np.random.default_rng(seed)
    n = len(basis)
    t0 = time.perf_counter()
    pool = mp.get_context('fork').Pool(procs, initializer=_init, initargs=(basis,))
    pool.starmap(_eri_timed, [(0, 0, 0, 0)] * procs)
    startup = time.perf_counter() - t0
    per = []
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < seconds:
        keys = sample_canonical(n, batch * procs, rng)
        per += [t for _, t in pool.starmap(_eri_timed, keys)]
    wall = time.perf_counter() - t_start
    pool.close()
    pool.join()
    return {'procs': procs, 'sampled_quartets': len(per), 'worker_s_per_quartet': float(np.mean(per)),
            'sample_wall_s': wall, 'pool_startup_s': startup}


def time_dense_fill(n, rows=2):
    """Seconds per element of the reference's N^4 sort_index + dict fill (timed on `rows` leading indices)."""
    keys = {}
    for a, b in itertools.product(range(n), repeat=2):
        if a <= b:
            keys[(a, b)] = 0.0
    t0 = time.perf_counter()
    cnt = 0
    for a in range(rows):
        for b, c, d in itertools.product(range(n), repeat=3):
            k = sort_index(a, b, c, d)
            _ = keys.get((k[0], k[1]), 0.0)
            cnt += 1
    return (time.perf_counter() - t0) / cnt


def time_fock_loop(n, rows=3, seed=0):
    """Seconds per (i<=j) element of the reference's restricted Fock loop on an N^4 array."""
    rng = np.random.default_rng(seed)
    m = min(n, 64)
    eri = rng.random((m, m, m, m))
    p = rng.random((m, m))
    t0 = time.perf_counter()
    done = 0
    for i in range(rows):
        g = 0.0
        for a, b in itertools.product(range(m), repeat=2):
            g += p.item(a, b) * (eri.item(i, i, a, b) - 0.5 * eri.item(i, b, a, i))
        done += 1
    per_inner = (time.perf_counter() - t0) / (done * m * m)
    return per_inner * n * n  # one (i, j) element sums over N^2 terms


def baseline(basis, seconds=12.0, procs=None, seed=0):
    """Extrapolated reference CPU time of one ERI build + dense fill + one Fock build."""
    n = len(basis)
    e = time_eri(basis, seconds, procs, seed)
    nq = n_canonical(n)
    eri_s = e['worker_s_per_quartet'] * nq / e['procs'] + e['pool_startup_s']
    fill_s = time_dense_fill(n) * n ** 4
    fock_s = time_fock_loop(n) * n * (n + 1) // 2
    total = eri_s + fill_s + fock_s
    return {'nbf': n, 'canonical_quartets': nq, 'eri_s': eri_s, 'fill_s': fill_s, 'fock_s': fock_s,
            'total_s': total, 'quartets_per_s': nq / total, 'eri_quartets_per_s': nq / eri_s, **e,
            'extrapolated': True}
