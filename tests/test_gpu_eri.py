"""GPU parity of the ERI kernels (through libgx.so) against the oracle.

Bars: ERIs within 1e-10 absolute of the CPU oracle (the C restatement of the
reference's Obara-Saika recursion) evaluated with an exact Boys function, and
within 1e-8 of the same oracle with the reference's truncated Boys series
(integrals/boys_function.py:28-54 drops series terms below 1e-10; on the
66-function cluster that alone moves integrals by up to 4e-9); the
reference's own 4-significant-digit HeH+ values.
"""
import numpy as np
import pytest

from tests.helpers import canonical_sample, heh_sto3g_mock, load_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def Engine():
    from gaussium_amd.engine import Engine
    return Engine


def _both(fn):
    """Oracle values with the reference's Boys series and with the exact Boys function."""
    from oracle import eri_c
    ref = fn()
    with eri_c.exact_boys():
        exact = fn()
    return ref, exact


def _check(got, fn, tol=1e-10, tol_series=1e-8):
    ref, exact = _both(fn)
    assert np.abs(got - exact).max() < tol
    assert np.abs(got - ref).max() < tol_series


def test_heh_reference_values(Engine):
    """tests/testsfast/matrixelements/test_electron_repulsion_matrix.py:124-146 of the reference."""
    from oracle import eri_c
    basis = heh_sto3g_mock()
    eng = Engine(basis)
    keys = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 0), (1, 1, 1, 1)]
    want = [1.3072, 0.4373, 0.1773, 0.6057, 0.3118, 0.7746]
    got = eng.eri_quartets(keys)
    for g, w in zip(got, want):
        np.testing.assert_approx_equal(g, w, 4)
    _check(got, lambda: eri_c.eri_keys(basis, keys))


@pytest.mark.parametrize('system', [('HeH+.mol', 'STO-3G.gbs'), ('HeH+.mol', '6-311+GPP.gbs'),
                                    ('C2H4.mol', '3-21G.gbs')])
def test_dense_matches_oracle(Engine, system):
    from oracle import eri_c
    nuc, basis, _ = load_system(*system)
    eng = Engine(basis, nuc)
    got = eng.eri_dense().cpu().numpy()
    _check(got, lambda: eri_c.eri_dense(basis))


def test_triatomic_all_unique_quartets(Engine):
    from gaussium_amd import workloads
    from oracle import eri_c
    from oracle.basis import canonical_keys
    nuc, basis, ne, info = workloads.triatomic_7bf()
    eng = Engine(basis, nuc)
    keys = np.array(canonical_keys(len(basis)), dtype=np.int32)
    assert len(keys) == 406
    got = eng.eri_quartets(keys)
    _check(got, lambda: eri_c.eri_keys(basis, keys))
    dense = eng.eri_dense().cpu().numpy()
    _check(dense, lambda: eri_c.eri_dense(basis))


def test_cluster_66bf_sampled_quartets(Engine):
    from gaussium_amd import workloads
    from oracle import eri_c
    nuc, basis, ne, info = workloads.cluster_66bf()
    eng = Engine(basis, nuc)
    keys = canonical_sample(len(basis), 4000, np.random.default_rng(11))
    got = eng.eri_quartets(keys)
    _check(got, lambda: eri_c.eri_keys(basis, keys))


def test_dense_symmetry_and_schwarz(Engine):
    """Size-independent properties: 8-fold symmetry and |(ij|kl)| <= sqrt((ij|ij)(kl|kl))."""
    from gaussium_amd import workloads
    nuc, basis, ne, info = workloads.cluster_66bf()
    eng = Engine(basis, nuc)
    d = eng.eri_dense().cpu().numpy()
    assert np.array_equal(d, d.transpose(1, 0, 2, 3))
    assert np.array_equal(d, d.transpose(2, 3, 0, 1))
    n = d.shape[0]
    diag = np.einsum('ijij->ij', d)
    bound = np.sqrt(np.abs(diag)).reshape(n * n)
    m = d.reshape(n * n, n * n)
    assert np.all(np.abs(m) <= np.outer(bound, bound) * (1 + 1e-12) + 1e-14)


def test_dense_repulsion_matrix_with_point_group(Engine):
    """``TwoElectronRepulsionMatrixOS`` with the ``Symmetry`` of C2H4 (D2h,
    oriented by ``MoleculeFactory``): the integrals the group forbids are
    exactly 0.0 (as the reference's ``calculate_integral`` returns them), the
    others match the oracle, and ``calculate_integral`` agrees; the RHF
    energy of ``start(symmetry=True)`` equals that without symmetry."""
    import os
    from gaussium_amd.basis import read_basis_set_file, read_mol_file
    from gaussium_amd.hartree_fock import TwoElectronRepulsionMatrixOS
    from gaussium_amd.main import start
    from gaussium_amd.symmetry import MoleculeFactory, Symmetry
    from oracle import eri_c
    from tests.conftest import ROOT
    g = os.path.join(ROOT, 'tests', 'golden')
    nuclei, _, _ = read_mol_file(os.path.join(g, 'C2H4.mol'))
    nuclei, pg = MoleculeFactory(symmetry=True).create(nuclei)
    assert pg.label == 'D_{2h}'
    basis = read_basis_set_file(os.path.join(g, 'STO-3G.gbs'), nuclei)
    sym = Symmetry(pg, basis)
    op = TwoElectronRepulsionMatrixOS(basis, symmetry=sym, engine=Engine(basis, nuclei))
    got = op.create_repulsion_matrix()
    mask = sym.zero_mask()
    assert mask.sum() > 0 and np.all(got[mask] == 0.0)
    ref = eri_c.eri_dense(basis)
    assert np.abs(got - ref).max() < 1e-10
    for q in [(0, 1, 2, 3), (1, 1, 5, 7), (2, 4, 6, 8)]:
        v = op.calculate_integral(*q)
        assert (v == 0.0) if mask[q] else abs(v - ref[q]) < 1e-10
    e_sym = start(os.path.join(g, 'C2H4.mol'), os.path.join(g, 'STO-3G.gbs'), 'RHF', 1, symmetry=True)
    e_c1 = start(os.path.join(g, 'C2H4.mol'), os.path.join(g, 'STO-3G.gbs'), 'RHF', 1)
    assert abs(e_sym - e_c1) < 1e-9
