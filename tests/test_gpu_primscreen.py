"""Primitive-pair screening (gx_set_prim_screen).

A primitive pair (a, b) is dropped when S_ab S_max < eps, S_ab the Schwarz
value of the scaled primitive product.  By the Schwarz inequality every
primitive quartet left out is below eps, so a contracted integral moves by
less than eps x (its number of primitive quartets).  Checked here:

* that bound itself, with an eps (1e-2) large enough to drop primitive pairs
  of a contracted S/L/P/D basis on three close centres (the dense tensor with
  and without screening), and that eps = 0 reproduces the oracle at 1e-10;
* the 220-function cluster: the default threshold keeps a fraction of the
  primitive pairs and leaves J and K (cutoff 1e-11, FP64 and mixed) within
  1e-11 of the unscreened build.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _d_system():
    from gaussium_amd.basis import Nuclei
    from tests.test_gpu_dshell import _d_basis
    nuc = [Nuclei('A', 3.0, 6.0, (0.0, 0.0, 0.0)), Nuclei('B', 2.0, 4.0, (1.3, -0.4, 0.6)),
           Nuclei('C', 1.0, 1.0, (-0.5, 1.1, -0.9))]
    return _d_basis(), nuc


def _nprim_quartets(basis):
    """Primitive quartets of each bf quartet: product of the four contraction lengths."""
    k = np.array([len(b.primitive_gaussian_array) for b in basis], dtype=np.float64)
    return k[:, None, None, None] * k[None, :, None, None] * k[None, None, :, None] * k[None, None, None, :]


def test_screening_bound_on_dense_tensor():
    from gaussium_amd.engine import Engine
    from oracle import eri_c
    basis, nuc = _d_system()
    full = Engine(basis, nuc, prim_eps=0.0)
    ref = full.eri_dense().cpu().numpy()
    with eri_c.exact_boys():
        exact = eri_c.eri_dense(basis)
    assert np.abs(ref - exact).max() < 1e-10
    eps = 1e-2
    cut = Engine(basis, nuc, prim_eps=eps)
    assert cut.n_primitive_pairs < full.n_primitive_pairs
    got = cut.eri_dense().cpu().numpy()
    # per element: below eps times its primitive-quartet count (pair products
    # of one shell pair share the same screened primitive set)
    assert np.all(np.abs(got - ref) <= eps * _nprim_quartets(basis) + 1e-15)
    dflt = Engine(basis, nuc)
    assert np.abs(dflt.eri_dense().cpu().numpy() - ref).max() < 1e-14


def test_cluster_220bf_jk_unchanged_by_default_screening():
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    nuc, basis, ne, info = workloads.cluster_220bf()
    n = len(basis)
    full = Engine(basis, nuc, prim_eps=0.0)
    eng = Engine(basis, nuc)
    kept, total = eng.n_primitive_pairs, full.n_primitive_pairs
    print('primitive pairs kept %d of %d' % (kept, total), flush=True)
    assert kept < 0.6 * total
    c = np.random.default_rng(1).standard_normal((n, ne // 2))
    P = torch.from_numpy(c @ c.T / n).cuda()
    for kw in (dict(cutoff=1e-11), dict(cutoff=1e-11, fp32_bound=1e-4, mixed=True)):
        J0, K0 = full.jk(P, **kw)
        J1, K1 = eng.jk(P, **kw)
        assert (J1 - J0).abs().max().item() < 1e-11, kw
        assert (K1 - K0).abs().max().item() < 1e-11, kw
