"""The generated VRR/HRR class code, compiled for the host with g++, against the C oracle.

Every class up to (dd|dd) with random centres, exponents and contractions;
the oracle is the independent C restatement of the reference's recursive
Obara-Saika scheme (oracle/c/os_eri.c).  Tolerance 1e-10 relative to the
largest integral of the quartet.
"""
import ctypes
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from tests.conftest import ROOT

HARNESS = os.path.join(ROOT, 'tests', 'host_harness')
BUILD = os.path.join(ROOT, 'tests', 'host_harness', '_build')

COMPS = {0: [(0, 0, 0)], 1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
         2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]}


@pytest.fixture(scope='module')
def host_lib():
    build = os.path.join(BUILD, 'p%d' % os.getpid())  # one per process (parallel test workers)
    os.makedirs(build, exist_ok=True)
    gen = os.path.join(build, 'eri_gen.h')
    subprocess.run([sys.executable, os.path.join(ROOT, 'gaussium_amd', 'csrc', 'gen_eri.py'), gen, '2'],
                   check=True)
    sys.path.insert(0, ROOT)
    from gaussium_amd.build import generate_rys
    rys_dir = generate_rys(os.path.join(BUILD, 'rys'))  # shared: about 30 s to make, kept while current
    so = os.path.join(build, 'liberi_host.so')
    subprocess.run(['g++', '-O1', '-std=c++17', '-shared', '-fPIC', '-I', build, '-I', os.path.dirname(rys_dir),
                    '-I', os.path.join(ROOT, 'gaussium_amd', 'csrc'),
                    os.path.join(HARNESS, 'eri_host.cpp'), '-o', so], check=True)
    lib = ctypes.CDLL(so)
    shutil.rmtree(build)
    lib.host_quartet.restype = ctypes.c_int
    lib.host_quartet_prog.restype = ctypes.c_int
    lib.host_quartet_grad.restype = ctypes.c_int
    lib.host_quartet_rys.restype = ctypes.c_int
    return lib


def _shell(rng, l, center):
    n = int(rng.integers(1, 4))
    a = np.exp(rng.uniform(np.log(0.1), np.log(20.0), n))
    w = rng.uniform(0.1, 1.0, n)
    return l, np.asarray(center, float), a, w


def _oracle(shells):
    """Dense block via the C oracle: one pseudo-basis function per component."""
    from oracle import eri_c

    class P:
        pass

    basis = []
    for (l, c, a, w) in shells:
        for comp in COMPS[l]:
            b = P()
            b.coordinates = tuple(c)
            b.integral_exponents = comp
            b.normalisation = 1.0
            prims = []
            for ai, wi in zip(a, w):
                g = P()
                g.exponent = ai
                g.contraction = wi
                g.normalisation = 1.0
                prims.append(g)
            b.primitive_gaussian_array = prims
            basis.append(b)
    offs = np.cumsum([0] + [len(COMPS[s[0]]) for s in shells])
    keys = []
    for i in range(offs[0], offs[1]):
        for j in range(offs[1], offs[2]):
            for k in range(offs[2], offs[3]):
                for m in range(offs[3], offs[4]):
                    keys.append((i, j, k, m))
    return eri_c.eri_keys(basis, keys)


CLASSES = [(la, lb, lc, ld) for la in range(3) for lb in range(la + 1)
           for lc in range(3) for ld in range(lc + 1)
           if (la * (la + 1) // 2 + lb) >= (lc * (lc + 1) // 2 + ld)]


@pytest.mark.parametrize('entry', ['host_quartet', 'host_quartet_prog', 'host_quartet_rys'])
@pytest.mark.parametrize('cls', CLASSES, ids=lambda c: '%d%d%d%d' % c)
def test_generated_class_matches_oracle(host_lib, cls, entry):
    """``host_quartet``: the code path the device takes for the class (per-lane
    class code, in passes over the ket components for the large classes
    without ket HRR, or the cooperative program for the large classes);
    ``host_quartet_prog``: the cooperative program of every class, run level by
    level with a check that each source slot is written before it is read;
    ``host_quartet_rys``: Rys quadrature (tabulated roots and weights, 2D
    integrals per root and axis)."""
    rng = np.random.default_rng(sum(v * 10 ** i for i, v in enumerate(cls)) + 7)
    for trial in range(2):
        centers = rng.uniform(-1.5, 1.5, (4, 3))
        if trial == 1:
            centers[1] = centers[0]  # same-centre pair exercises the zero-skips
        shells = [_shell(rng, l, c) for l, c in zip(cls, centers)]
        l = np.array([s[0] for s in shells], dtype=np.int32)
        ctr = np.ascontiguousarray(np.concatenate([s[1] for s in shells]))
        n = np.array([len(s[2]) for s in shells], dtype=np.int32)
        a = np.ascontiguousarray(np.concatenate([s[2] for s in shells]))
        w = np.ascontiguousarray(np.concatenate([s[3] for s in shells]))
        nout = int(np.prod([len(COMPS[x]) for x in cls]))
        out = np.zeros(nout)
        rc = getattr(host_lib, entry)(l.ctypes.data_as(ctypes.c_void_p), ctr.ctypes.data_as(ctypes.c_void_p),
                                   n.ctypes.data_as(ctypes.c_void_p), a.ctypes.data_as(ctypes.c_void_p),
                                   w.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
        assert rc == 0
        ref = _oracle(shells)
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(out - ref).max() <= 1e-10 * scale, (cls, trial, np.abs(out - ref).max())


GRAD_CLASSES = [c for c in CLASSES if sum(c) <= 5] + [(2, 1, 2, 1), (2, 2, 2, 2)]


@pytest.mark.parametrize('cls', GRAD_CLASSES, ids=lambda c: '%d%d%d%d' % c)
@pytest.mark.parametrize('entry', ['host_quartet_grad', 'host_quartet_grad_rys'])
def test_gradient_program_matches_finite_differences(host_lib, cls, entry):
    """d/dX_i (ab|cd) for X = A, B, C from the gradient program (raised and
    lowered classes with 2 alpha weights; ``host_quartet_grad_rys``: the Rys
    2D integrals and their centre derivatives) against a 4-point central
    difference of the oracle's integrals with the centre displaced (h = 1e-3,
    error O(h^4))."""
    getattr(host_lib, entry).restype = ctypes.c_int
    rng = np.random.default_rng(sum(v * 10 ** i for i, v in enumerate(cls)) + 31)
    centers = rng.uniform(-1.2, 1.2, (4, 3))
    shells = [_shell(rng, l, c) for l, c in zip(cls, centers)]
    l = np.array([s[0] for s in shells], dtype=np.int32)
    ctr = np.ascontiguousarray(np.concatenate([s[1] for s in shells]))
    n = np.array([len(s[2]) for s in shells], dtype=np.int32)
    a = np.ascontiguousarray(np.concatenate([s[2] for s in shells]))
    w = np.ascontiguousarray(np.concatenate([s[3] for s in shells]))
    nout = int(np.prod([len(COMPS[x]) for x in cls]))
    out = np.zeros(9 * nout)
    rc = getattr(host_lib, entry)(l.ctypes.data_as(ctypes.c_void_p), ctr.ctypes.data_as(ctypes.c_void_p),
                                  n.ctypes.data_as(ctypes.c_void_p), a.ctypes.data_as(ctypes.c_void_p),
                                  w.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    got = out.reshape(nout, 3, 3)
    h = 1e-3
    scale = max(1.0, np.abs(_oracle(shells)).max())
    for X in range(3):
        for i in range(3):
            vals = []
            for step in (2, 1, -1, -2):
                sh = [list(s) for s in shells]
                c = sh[X][1].copy()
                c[i] += step * h
                sh[X][1] = c
                vals.append(_oracle([tuple(s) for s in sh]))
            fd = (8 * (vals[1] - vals[2]) - (vals[0] - vals[3])) / (12 * h)
            err = np.abs(got[:, X, i] - fd).max()
            assert err <= 1e-8 * scale, (cls, X, i, err)
    # translational invariance fixes the D centre: the three derivatives must
    # not sum to zero unless D is on one of the other centres, so only check
    # that the D derivative from a finite difference equals -(A + B + C)
    fdD = []
    for i in range(3):
        vals = []
        for step in (2, 1, -1, -2):
            sh = [list(s) for s in shells]
            c = sh[3][1].copy()
            c[i] += step * h
            sh[3][1] = c
            vals.append(_oracle([tuple(s) for s in sh]))
        fdD.append((8 * (vals[1] - vals[2]) - (vals[0] - vals[3])) / (12 * h))
    for i in range(3):
        assert np.abs(-got[:, :, i].sum(axis=1) - fdD[i]).max() <= 1e-8 * scale


# one-pass classes, then the three-pass ones (one program per centre)
LANE_GRAD_CLASSES = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (2, 0, 0, 0), (2, 1, 0, 0),
                     (1, 1, 1, 0), (2, 0, 1, 0), (2, 0, 2, 0), (2, 1, 1, 0), (2, 2, 0, 0)]


@pytest.mark.parametrize('cls', LANE_GRAD_CLASSES, ids=lambda c: '%d%d%d%d' % c)
def test_lane_gradient_code_matches_program(host_lib, cls):
    """The per-lane gradient code (straight-line VRR, renamed HRR, contraction
    with D_k = f (4 P_ab P_cd - P_ac P_bd - P_ad P_bc)) equals the gradient
    program's derivative integrals contracted with the same D_k (the full
    program, itself checked against finite differences above; for the
    three-pass classes this checks the per-centre programs against it)."""
    host_lib.host_grad_lane.restype = ctypes.c_int
    rng = np.random.default_rng(sum(v * 10 ** i for i, v in enumerate(cls)) + 53)
    centers = rng.uniform(-1.2, 1.2, (4, 3))
    shells = [_shell(rng, l, c) for l, c in zip(cls, centers)]
    l = np.array([s[0] for s in shells], dtype=np.int32)
    ctr = np.ascontiguousarray(np.concatenate([s[1] for s in shells]))
    n = np.array([len(s[2]) for s in shells], dtype=np.int32)
    a = np.ascontiguousarray(np.concatenate([s[2] for s in shells]))
    w = np.ascontiguousarray(np.concatenate([s[3] for s in shells]))
    NA, NB, NC, ND = (len(COMPS[x]) for x in cls)
    nout = NA * NB * NC * ND
    d = np.zeros(9 * nout)
    p_ = lambda v: v.ctypes.data_as(ctypes.c_void_p)
    assert host_lib.host_quartet_grad(p_(l), p_(ctr), p_(n), p_(a), p_(w), p_(d)) == 0
    d = d.reshape(NA, NB, NC, ND, 9)
    blocks = [rng.standard_normal(k) for k in (NA * NB, NC * ND, NA * NC, NA * ND, NB * NC, NB * ND)]
    Pab, Pcd, Pac, Pad, Pbc, Pbd = (b.reshape(s) for b, s in zip(
        blocks, [(NA, NB), (NC, ND), (NA, NC), (NA, ND), (NB, NC), (NB, ND)]))
    f = 0.5
    D = f * (4 * np.einsum('ab,cd->abcd', Pab, Pcd) - np.einsum('ac,bd->abcd', Pac, Pbd)
             - np.einsum('ad,bc->abcd', Pad, Pbc))
    want = np.einsum('abcd,abcdj->j', D, d)
    pb = np.ascontiguousarray(np.concatenate(blocks))
    out = np.zeros(9)
    assert host_lib.host_grad_lane(p_(l), p_(ctr), p_(n), p_(a), p_(w), p_(pb), ctypes.c_double(f), p_(out)) == 0
    assert np.abs(out - want).max() <= 1e-11 * max(1.0, np.abs(want).max())
