"""GPU parity with d shells: every class up to (dd|dd), including the
wave-cooperative ones, against the C oracle; J/K with d shells; the
255-function polarised cluster through size-independent properties.

Bars as in test_gpu_eri / test_gpu_jk: 1e-10 against the oracle with the
exact Boys function, 1e-8 against the reference's truncated Boys series.
"""
import numpy as np
import pytest
import torch

from tests.helpers import canonical_sample

pytestmark = pytest.mark.gpu


def _d_basis():
    """Three centres, contracted S/L/P/D shells (two-primitive D included)."""
    from gaussium_amd.basis import expand_shells
    a = [['S', [0.4, 9.0], [0.7, 1.7], [0.3, 0.35]], ['L', [0.5, 1.2, 0.6], [0.8, 0.3, 0.5]],
         ['D', [0.6, 1.4], [0.5, 0.45]]]
    b = [['S', [0.9, 2.5], [0.4, 0.5]], ['P', [1.0, 0.7]], ['D', [1.0, 0.8]]]
    c = [['S', [0.5, 3.0], [0.6, 0.6]], ['D', [0.4, 2.2], [0.8, 0.9]]]
    basis = expand_shells(a, (0.0, 0.0, 0.0)) + expand_shells(b, (1.3, -0.4, 0.6)) + \
        expand_shells(c, (-0.5, 1.1, -0.9))
    return basis


def _check(got, fn):
    from oracle import eri_c
    ref = fn()
    with eri_c.exact_boys():
        exact = fn()
    assert np.abs(got - exact).max() < 1e-10
    assert np.abs(got - ref).max() < 1e-8


def test_dense_with_d_shells_matches_oracle():
    from gaussium_amd.engine import Engine
    from oracle import eri_c
    basis = _d_basis()
    assert len(basis) == 28
    eng = Engine(basis)
    got = eng.eri_dense().cpu().numpy()
    _check(got, lambda: eri_c.eri_dense(basis))


def test_quartets_with_d_shells_match_oracle():
    from gaussium_amd.engine import Engine
    from oracle import eri_c
    basis = _d_basis()
    eng = Engine(basis)
    keys = canonical_sample(len(basis), 3000, np.random.default_rng(21))
    got = eng.eri_quartets(keys)
    _check(got, lambda: eri_c.eri_keys(basis, keys))


def test_jk_with_d_shells_matches_oracle():
    from gaussium_amd.engine import Engine
    from oracle import eri_c
    from oracle.scf import jk_from_dense
    basis = _d_basis()
    eng = Engine(basis)
    n = len(basis)
    c = np.random.default_rng(8).standard_normal((n, n // 2))
    P = c @ c.T / n
    J, K = eng.jk(torch.from_numpy(P).cuda())
    with eri_c.exact_boys():
        Jr, Kr = jk_from_dense(eri_c.eri_dense(basis), P)
    assert np.abs(J.cpu().numpy() - Jr).max() < 1e-10
    assert np.abs(K.cpu().numpy() - Kr).max() < 1e-10


def test_jk_polarised_cluster_matches_oracle():
    """3 + 2 atom polarised cluster (d on the heavy atoms), with and without a cutoff."""
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    from oracle import eri_c
    from oracle.scf import jk_from_dense
    nuc, basis, ne, info = workloads.cluster(3, 2, 5.0, polarised=True, seed=9)
    n = len(basis)
    assert n == 49
    eng = Engine(basis, nuc)
    c = np.random.default_rng(4).standard_normal((n, n // 2))
    P = c @ c.T / n
    with eri_c.exact_boys():
        Jr, Kr = jk_from_dense(eri_c.eri_dense(basis), P)
    J, K = eng.jk(torch.from_numpy(P).cuda())
    assert np.abs(J.cpu().numpy() - Jr).max() < 1e-10
    assert np.abs(K.cpu().numpy() - Kr).max() < 1e-10
    J, K = eng.jk(torch.from_numpy(P).cuda(), cutoff=1e-11)
    assert np.abs(J.cpu().numpy() - Jr).max() < 1e-9
    assert np.abs(K.cpu().numpy() - Kr).max() < 1e-9
    J, K = eng.jk(torch.from_numpy(P).cuda(), cutoff=1e-11, fp32_bound=1e-4, mixed=True)
    assert np.abs(J.cpu().numpy() - Jr).max() < 1e-6
    assert np.abs(K.cpu().numpy() - Kr).max() < 1e-6


def test_jk_255bf_rows_match_oracle():
    """255-function polarised cluster (d shells): selected J and K elements
    against the oracle's integrals for those elements, J and K symmetric, and
    the mixed build within 1e-6 of FP64."""
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    from oracle import eri_c
    nuc, basis, ne, info = workloads.cluster_255bf()
    n = len(basis)
    assert n == 255
    eng = Engine(basis, nuc)
    rng = np.random.default_rng(6)
    c = rng.standard_normal((n, ne // 2))
    P = c @ c.T / n
    Pd = torch.from_numpy(P).cuda()
    J, K = eng.jk(Pd)
    J, K = J.cpu().numpy(), K.cpu().numpy()
    assert np.abs(J - J.T).max() < 1e-12 and np.abs(K - K.T).max() < 1e-12
    # elements on d functions (heavy atoms carry xx..yz last in their block)
    d_idx = [i for i, b in enumerate(basis) if sum(b.integral_exponents) == 2]
    picks = [(d_idx[0], d_idx[3]), (d_idx[7], 0), (5, d_idx[-1])]
    kl = np.array([(k, l) for k in range(n) for l in range(n)], dtype=np.int64)
    with eri_c.exact_boys():
        for i, j in picks:
            keys_j = np.column_stack([np.full(len(kl), i), np.full(len(kl), j), kl[:, 0], kl[:, 1]])
            keys_k = np.column_stack([np.full(len(kl), i), kl[:, 0], np.full(len(kl), j), kl[:, 1]])
            vj = eri_c.eri_keys(basis, keys_j.astype(np.int32))
            vk = eri_c.eri_keys(basis, keys_k.astype(np.int32))
            pf = P[kl[:, 0], kl[:, 1]]
            assert abs(J[i, j] - vj @ pf) < 1e-10
            assert abs(K[i, j] - vk @ pf) < 1e-10
    Jm, Km = eng.jk(Pd, cutoff=1e-11, fp32_bound=1e-4, mixed=True)
    assert np.abs(Jm.cpu().numpy() - J).max() < 1e-6
    assert np.abs(Km.cpu().numpy() - K).max() < 1e-6


def _one_electron_parity(basis, nuc):
    from gaussium_amd.engine import Engine
    from oracle.basis import Nuclei as ONuc
    from oracle.boys import exact
    from oracle.one_electron import kinetic_matrix, nuclear_matrix, overlap_matrix
    eng = Engine(basis, nuc)
    S, T, V = (x.cpu().numpy() for x in eng.one_electron())
    onuc = [ONuc(n.element, n.charge, n.mass, n.coordinates) for n in nuc]
    errs = (np.abs(S - overlap_matrix(basis)).max(), np.abs(T - kinetic_matrix(basis)).max())
    with exact():
        ev = np.abs(V - nuclear_matrix(basis, onuc)).max()
    print('S %.1e T %.1e V %.1e' % (errs[0], errs[1], ev))
    assert errs[0] < 1e-10 and errs[1] < 1e-10 and ev < 1e-10


def test_one_electron_with_d_shells_matches_oracle():
    """S, T, V of contracted S/L/P/D shells on three centres (d_xx .. d_yz,
    one- and two-primitive D) against the oracle, 1e-10."""
    from gaussium_amd.basis import Nuclei
    nuc = [Nuclei('A', 3.0, 6.0, (0.0, 0.0, 0.0)), Nuclei('B', 2.0, 4.0, (1.3, -0.4, 0.6)),
           Nuclei('C', 1.0, 1.0, (-0.5, 1.1, -0.9))]
    _one_electron_parity(_d_basis(), nuc)


def test_one_electron_polarised_cluster_matches_oracle():
    from gaussium_amd import workloads
    nuc, basis, ne, info = workloads.cluster(3, 2, 5.0, polarised=True, seed=9)
    _one_electron_parity(basis, nuc)
