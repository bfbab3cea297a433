"""Converged RHF on the seeded random clusters, with the opt-in second-order
SCF (``stabilizer='second_order'``), and the gradients built on it.

The reference's loop (hartreefock/scf_procedure.py:24-42 with the DIIS of
diismethod/diis.py:30-112) cycles on these clusters; the stabilised SCF is
compared with the oracle's restatement of the same option.

* 66-function cluster: total energy within 1e-8 Eh of the oracle SCF (dense
  oracle ERIs with the exact Boys function, cutoff 0); the mixed-precision
  build within 1e-6 Eh of FP64 at convergence.
* 66-function cluster: the analytic gradient at the converged density
  against central differences of converged SCF energies (h = 1e-4 bohr) on
  all 36 coordinates, within 1e-6 Eh/bohr (the displaced SCFs start from
  the converged density, so they stay on the same solution).
* 255-function cluster with d shells: converged energy and gradient; the
  same central differences on 6 coordinates, translational and rotational
  invariance.
"""
import numpy as np
import pytest

from tests.helpers import oracle_rhf

pytestmark = pytest.mark.gpu


def _surface(workload, **kw):
    from gaussium_amd.gradient import RHFSurface
    nuc, basis, ne, info = workload()
    kw.setdefault('max_iter', 3000)
    return RHFSurface(nuc, basis, ne, stabilizer='second_order', **kw)


@pytest.fixture(scope='module')
def surface_66():
    from gaussium_amd import workloads
    return _surface(workloads.cluster_66bf)


def test_cluster_66bf_energy_matches_oracle():
    from gaussium_amd import workloads
    from gaussium_amd.basis import nuclear_repulsion
    from gaussium_amd.hartree_fock import RestrictedHF
    nuc, basis, ne, info = workloads.cluster_66bf()
    hf = RestrictedHF(nuc, basis, ne, stabilizer='second_order', max_iter=3000)
    e = hf.energies()[0] + nuclear_repulsion(nuc)
    assert hf.scf_method.converged
    e_ref, it = oracle_rhf(nuc, basis, ne, stabilizer='second_order', max_iter=3000)
    print('66 bf: GPU %.10f (%d evaluations)  oracle %.10f (%d)' % (e, hf.scf_method.iterations, e_ref, it))
    assert abs(e - e_ref) < 1e-8
    # the same solution in mixed precision with the benchmark's screening
    hm = RestrictedHF(nuc, basis, ne, stabilizer='second_order', max_iter=3000, schwarz_cutoff=1e-11,
                      fp32_bound=1e-4, precision='mixed')
    em = hm.energies()[0] + nuclear_repulsion(nuc)
    assert hm.scf_method.converged
    assert abs(em - e) < 1e-6


def _fd_check(s, coords, h=1e-4, bar=1e-6):
    e0, g = s.energy_and_gradient(s.xyz0)
    assert s.last.scf_method.converged
    p0 = s.last.scf_method.density.clone()
    worst = 0.0
    for atom, axis in coords:
        xp, xm = s.xyz0.copy(), s.xyz0.copy()
        xp[atom, axis] += h
        xm[atom, axis] -= h
        fd = (s.energy(xp, initial_density=p0) - s.energy(xm, initial_density=p0)) / (2 * h)
        worst = max(worst, abs(g[atom, axis] - fd))
        assert abs(g[atom, axis] - fd) < bar, (atom, axis, g[atom, axis], fd)
    print('max |analytic - central difference| = %.2e Eh/bohr over %d coordinates' % (worst, len(coords)))
    return e0, g, p0


def test_cluster_66bf_gradient_matches_scf_finite_differences(surface_66):
    """Config: 12-atom cluster, all 36 coordinates, h = 1e-4 bohr."""
    s = surface_66
    _fd_check(s, [(a, x) for a in range(len(s.nuclei)) for x in range(3)])


@pytest.fixture(scope='module')
def surface_255():
    from gaussium_amd import workloads
    return _surface(workloads.cluster_255bf)


def test_cluster_255bf_energy_and_gradient(surface_255):
    """30 atoms with d shells: converged RHF energy, gradient against central
    differences of converged energies on 6 coordinates (heavy and light
    atoms, every axis), and its invariances."""
    s = surface_255
    e0, g, p0 = _fd_check(s, [(0, 0), (3, 1), (7, 2), (15, 0), (22, 1), (29, 2)])
    assert np.abs(g.sum(axis=0)).max() < 1e-8
    torque = np.cross(s.xyz0, g).sum(axis=0)
    assert np.abs(torque).max() < 1e-7
    # the mixed-precision build of the same energy
    from gaussium_amd import workloads
    sm = _surface(workloads.cluster_255bf, schwarz_cutoff=1e-11, precision='mixed', fp32_bound=1e-4)
    em = sm.energy(s.xyz0, initial_density=p0)
    assert abs(em - e0) < 1e-6


def test_cluster_66bf_mixed_gradient(surface_66):
    """Mixed-precision gradient (gx_grad_ex with GX_JK_MIXED: quartets with
    Schwarz product below 1e-4 in FP32) at the converged density of the
    66-function cluster: within 1e-6 Eh/bohr of the FP64 gradient."""
    s = surface_66
    s.energy_and_gradient(s.xyz0)
    hf = s.last
    eng = s.engine
    no = s.electrons // 2
    c = hf.scf_method.orbital_coefficients[:, :no]
    W = 2.0 * (c * hf.scf_method.orbital_energies[:no]) @ c.T
    P = hf.scf_method.density
    g64 = eng.grad(P, W, cutoff=1e-11).cpu().numpy()
    gmx = eng.grad(P, W, cutoff=1e-11, fp32_bound=1e-4, mixed=True).cpu().numpy()
    rows = eng.jk_class_counts(cutoff=1e-11, fp32_bound=1e-4, mixed=True)
    assert sum(r[5] for r in rows) > 0  # FP32 quartets exist
    err = np.abs(gmx - g64).max()
    print('66 bf: max |g_mixed - g_fp64| = %.2e Eh/bohr' % err)
    assert err < 1e-6
    assert np.abs(gmx.sum(axis=0)).max() < 1e-8


def test_mixed_gradient_all_fp32_path():
    """Every quartet of the 7-function triatomic in FP32 (bound 1e30): the FP32
    primitive loop (pair records with the exponent of A, FP32 Boys table)
    agrees with FP64 to single precision, at a random symmetric P and W."""
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    import torch
    nuc, basis, ne, info = workloads.triatomic_7bf()
    n = len(basis)
    rng = np.random.default_rng(4)
    a = rng.standard_normal((n, n))
    b = rng.standard_normal((n, n))
    P = torch.from_numpy((a + a.T) / n).cuda()
    W = torch.from_numpy((b + b.T) / n).cuda()
    eng = Engine(basis, nuc)
    g64 = eng.grad(P, W).cpu().numpy()
    g32 = eng.grad(P, W, fp32_bound=1e30, mixed=True).cpu().numpy()
    scale = np.abs(g64).max()
    err = np.abs(g32 - g64).max()
    print('all-FP32 gradient: max |g32 - g64| = %.2e (max |g| %.2e)' % (err, scale))
    assert err < 1e-4 * max(scale, 1.0)
    assert err > 0.0  # the FP32 path ran
