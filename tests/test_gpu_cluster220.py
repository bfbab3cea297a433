"""Oracle parity on the benchmark's 40-atom, 220-function cluster.

* J and K elements (s and p functions, on and off the diagonal) at cutoff 0
  against the oracle's integrals for those elements (the rows (ij|kl) and
  (ik|jl) over all k, l, exact Boys function) contracted with the same
  density, as hartreefock/fock_matrix.py:17-19 of the reference sums them;
  bar 1e-10.
* the benchmark's Schwarz cutoff (1e-11) moves J and K by less than 1e-8.
* at the converged SCF density, the mixed-precision build's energy
  functional E[P] = 1/2 tr P (2H + J - K/2) is within 1e-6 Eh of the FP64
  build's.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def cluster():
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    nuc, basis, ne, info = workloads.cluster_220bf()
    assert len(basis) == 220
    return nuc, basis, ne, Engine(basis, nuc)


def test_cluster_220bf_jk_elements_match_oracle(cluster):
    from oracle import eri_c
    nuc, basis, ne, eng = cluster
    n = len(basis)
    rng = np.random.default_rng(12)
    c = rng.standard_normal((n, ne // 2))
    P = c @ c.T / n
    Pd = torch.from_numpy(P).cuda()
    J, K = eng.jk(Pd)
    J, K = J.cpu().numpy(), K.cpu().numpy()
    assert np.abs(J - J.T).max() < 1e-12 and np.abs(K - K.T).max() < 1e-12
    p_idx = [i for i, b in enumerate(basis) if sum(b.integral_exponents) == 1]
    s_idx = [i for i, b in enumerate(basis) if sum(b.integral_exponents) == 0]
    picks = [(p_idx[0], p_idx[0]), (p_idx[5], p_idx[40]), (s_idx[3], p_idx[-1]), (s_idx[0], s_idx[-1])]
    kl = np.array([(k, l) for k in range(n) for l in range(n)], dtype=np.int64)
    pf = P[kl[:, 0], kl[:, 1]]
    worst = 0.0
    with eri_c.exact_boys():
        for i, j in picks:
            keys_j = np.column_stack([np.full(len(kl), i), np.full(len(kl), j), kl[:, 0], kl[:, 1]])
            keys_k = np.column_stack([np.full(len(kl), i), kl[:, 0], np.full(len(kl), j), kl[:, 1]])
            vj = eri_c.eri_keys(basis, keys_j.astype(np.int32)) @ pf
            vk = eri_c.eri_keys(basis, keys_k.astype(np.int32)) @ pf
            worst = max(worst, abs(J[i, j] - vj), abs(K[i, j] - vk))
            assert abs(J[i, j] - vj) < 1e-10, (i, j, J[i, j], vj)
            assert abs(K[i, j] - vk) < 1e-10, (i, j, K[i, j], vk)
    print('220 bf: max |GPU - oracle| over %d J and %d K elements = %.2e' % (len(picks), len(picks), worst))
    Jc, Kc = eng.jk(Pd, cutoff=1e-11)
    dj = np.abs(Jc.cpu().numpy() - J).max()
    dk = np.abs(Kc.cpu().numpy() - K).max()
    print('220 bf: cutoff 1e-11 moves J by %.2e, K by %.2e' % (dj, dk))
    assert dj < 1e-8 and dk < 1e-8


def test_cluster_220bf_mixed_energy_at_converged_density(cluster):
    from gaussium_amd.basis import nuclear_repulsion
    from gaussium_amd.hartree_fock import RestrictedHF
    nuc, basis, ne, eng = cluster
    hf = RestrictedHF(nuc, None, ne, engine=eng, schwarz_cutoff=1e-11, stabilizer='second_order', max_iter=5000)
    e = hf.energies()[0] + nuclear_repulsion(nuc)
    assert hf.scf_method.converged
    P = hf.scf_method.density
    H = hf.core_hamiltonian
    energy = {}
    for mode, kw in (('fp64', {}), ('mixed', dict(fp32_bound=1e-4, mixed=True))):
        J, K = eng.jk(P, cutoff=1e-11, **kw)
        energy[mode] = 0.5 * float(torch.sum(P * (2.0 * H + J - 0.5 * K))) + nuclear_repulsion(nuc)
    print('220 bf: SCF %.10f (%d evaluations); E[P] FP64 %.10f, mixed %.10f' % (
        e, hf.scf_method.iterations, energy['fp64'], energy['mixed']))
    assert abs(energy['fp64'] - e) < 1e-9
    assert abs(energy['mixed'] - energy['fp64']) < 1e-6
