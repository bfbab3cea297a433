"""Host-side logic without a GPU: C-ABI exports, shell packing, readers, flop model."""
import ctypes
import os
import re

import numpy as np
import pytest

from tests.conftest import ROOT
from tests.helpers import golden, heh_sto3g_mock


def _header_symbols():
    with open(os.path.join(ROOT, 'include', 'gx.h')) as f:
        text = f.read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(gx_[a-z_0-9]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    """libgx.so loads on the CPU host and exports each function of include/gx.h (no compute call)."""
    from gaussium_amd import _lib
    lib = _lib.load()
    syms = _header_symbols()
    assert 'gx_jk' in syms and 'gx_eri_dense' in syms
    for name in syms:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, 'ctypes signature missing for %s' % name


def test_last_error_reports_bad_arguments():
    """Error path of the C ABI: a null context is rejected with a message, no GPU touched."""
    from gaussium_amd import _lib
    lib = _lib.load()
    assert lib.gx_destroy(None) == -1
    assert b'null' in lib.gx_last_error()
    assert lib.gx_set_stream(None, None) == -1
    assert lib.gx_jk(None, None, None, None, 0.0, 0.0, 0) == -1
    with pytest.raises(_lib.GxError):
        _lib.check(lib.gx_eri_dense(None, None), 'gx_eri_dense')


def test_pack_shells_reference_order_and_scales():
    from gaussium_amd.basis import expand_shells, pack_shells
    shells = [['S', [0.3, 4.0], [0.6, 0.9]], ['L', [0.2, 1.5, 0.5], [0.8, 0.3, 0.7]],
              ['D', [0.6, 1.1], [0.5, 0.35]], ['P', [1.0, 0.5]]]
    basis = expand_shells(shells, (0.0, 1.0, 2.0))
    sh = pack_shells(basis)
    assert list(sh.l) == [0, 0, 1, 2, 1]
    assert list(sh.bf_off) == [0, 1, 2, 5, 11]
    assert list(sh.nprim) == [2, 2, 2, 2, 1]
    # d components carry 1/sqrt((2l-1)!!...) relative to xx: xy/xx = sqrt(3)
    d = sh.bf_scale[5:11]
    assert np.allclose(d[:3], 1.0) and np.allclose(d[3:], np.sqrt(3.0))
    # the packed radial coefficients reproduce n_bf * c * N_prim of every component
    for s in range(sh.nshell):
        for k in range({0: 1, 1: 3, 2: 6}[sh.l[s]]):
            b = basis[sh.bf_off[s] + k]
            w = [b.normalisation * g.contraction * g.normalisation for g in b.primitive_gaussian_array]
            got = sh.coefs[sh.prim_off[s]:sh.prim_off[s] + sh.nprim[s]] * sh.bf_scale[sh.bf_off[s] + k]
            assert np.allclose(got, w, rtol=1e-13)


def test_pack_shells_rejects_incomplete_shell():
    from gaussium_amd.basis import expand_shells, pack_shells
    basis = expand_shells([['P', [1.0, 0.5]]], (0, 0, 0))
    with pytest.raises(ValueError):
        pack_shells(basis[:2])


def test_mock_basis_packs_with_given_normalisation():
    from gaussium_amd.basis import pack_shells
    sh = pack_shells(heh_sto3g_mock())
    assert sh.nshell == 2 and list(sh.l) == [0, 0]
    assert sh.coefs[0] == pytest.approx(0.15432897 * 3.9336432656254527)


def test_readers_match_oracle():
    from gaussium_amd.basis import read_basis_set_file, read_mol_file
    from oracle.basis import read_basis_set_file as orb, read_mol_file as orm
    for mol, gbs in (('HeH+.mol', '6-311+GPP.gbs'), ('C2H4.mol', '3-21G.gbs')):
        n1, e1, m1 = read_mol_file(golden(mol))
        n2, e2, m2 = orm(golden(mol))
        assert (e1, m1) == (e2, m2)
        b1, b2 = read_basis_set_file(golden(gbs), n1), orb(golden(gbs), n2)
        assert len(b1) == len(b2)
        for x, y in zip(b1, b2):
            assert x.integral_exponents == y.integral_exponents
            assert x.normalisation == pytest.approx(y.normalisation, rel=1e-14)
            assert [g.exponent for g in x.primitive_gaussian_array] == [g.exponent for g in y.primitive_gaussian_array]


def test_workload_sizes():
    from gaussium_amd import pack_shells, workloads
    expect = {'triatomic_7bf': (7, 5, 10), 'cluster_220bf': (220, 140, 140), 'cluster_66bf': (66, 42, 42),
              'cluster_255bf': (255, 120, 104)}
    for name, (nbf, nsh, ne) in expect.items():
        nuc, basis, electrons, info = getattr(workloads, name)()
        assert len(basis) == nbf and pack_shells(basis).nshell == nsh and electrons == ne


def test_flop_model_per_class_totals():
    """Per-class K=1 totals (primitive work + HRR) of the canonical model."""
    from gaussium_amd.flops import hrr_flops, prim_flops
    table = {(0, 0, 0, 0): 48, (1, 0, 0, 0): 69, (1, 0, 1, 0): 164, (1, 1, 0, 0): 152, (1, 1, 1, 0): 517,
             (1, 1, 1, 1): 1662, (2, 0, 0, 0): 128, (2, 1, 1, 0): 1180, (2, 2, 2, 2): 38976}
    for c, want in table.items():
        assert prim_flops(*c) + hrr_flops(*c) == want, c


def test_canonical_counts():
    from oracle.basis import canonical_keys
    from oracle.ref_faithful import n_canonical
    assert len(canonical_keys(7)) == 406 == n_canonical(7)
    assert n_canonical(220) == 295500205


def _o2_sto3g_oracle_matrices():
    from oracle import eri_c
    from oracle.basis import nuclear_repulsion, read_basis_set_file, read_mol_file
    from oracle.one_electron import kinetic_matrix, nuclear_matrix, overlap_matrix
    from tests.helpers import golden
    nuc, ne, mult = read_mol_file(golden('O2.mol'))
    basis = read_basis_set_file(golden('STO-3G.gbs'), nuc)
    with eri_c.exact_boys():
        eri = eri_c.eri_dense(basis)
    S = overlap_matrix(basis)
    H = kinetic_matrix(basis) + nuclear_matrix(basis, nuc)
    return nuc, ne, mult, basis, eri, S, H, nuclear_repulsion(nuc)


def test_blocked_scf_host_logic_matches_oracle():
    """The device-side blocked SCF classes (torch; CPU tensors here) driven by
    the oracle's blocked Fock sums reach the oracle GUHF energy and the
    reference's published O2/STO-3G value."""
    import torch
    from gaussium_amd.scf import BlockedLinearAlgebra, BlockedUnrestrictedSCF
    from oracle.scf import blocked_spin_basis_set, fock_blocked, guhf
    nuc, ne, mult, basis, eri, S, H, e_nuc = _o2_sto3g_oracle_matrices()
    z = np.zeros_like(H)
    H2 = np.block([[H, z], [z, H]])
    S2 = np.block([[S, z], [z, S]])
    eri2 = blocked_spin_basis_set(eri)

    class OracleBlockedFock:
        core_hamiltonian = torch.from_numpy(H2)

        def create(self, p2):
            return torch.from_numpy(fock_blocked(H2, eri2, p2.numpy()))

    la = BlockedLinearAlgebra(torch.from_numpy(S2))
    scf = BlockedUnrestrictedSCF(la, ne, mult, torch.from_numpy(S2), OracleBlockedFock())
    e, _, c = scf.begin_iterations(*la.diagonalize(torch.from_numpy(H2)))
    e_ref, it = guhf(H, S, ne, mult, lambda p2: fock_blocked(H2, eri2, p2))
    assert abs(e - e_ref) < 1e-10, (e, e_ref)
    assert scf.iterations == it
    np.testing.assert_approx_equal(e + e_nuc, -147.634028141, 6)
    n = H.shape[0]
    assert bool(torch.all(c[:n, n:] == 0)) and bool(torch.all(c[n:, :n] == 0))


def test_unrestricted_scf_host_logic_matches_oracle():
    import torch
    from gaussium_amd.scf import LinearAlgebra, PopleNesbetBerthier
    from oracle.scf import jk_from_dense, uhf
    nuc, ne, mult, basis, eri, S, H, e_nuc = _o2_sto3g_oracle_matrices()

    def fb(pa, pb):
        ja, ka = jk_from_dense(eri, pa)
        jb, kb = jk_from_dense(eri, pb)
        return H + ja + jb - ka, H + ja + jb - kb

    class OracleFock:
        core_hamiltonian = torch.from_numpy(H)

        def create(self, pa, pb):
            fa, fb_ = fb(pa.numpy(), pb.numpy())
            return torch.from_numpy(fa), torch.from_numpy(fb_)

    la = LinearAlgebra(torch.from_numpy(S))
    scf = PopleNesbetBerthier(la, ne, mult, OracleFock())
    e = scf.begin_iterations(*la.diagonalize(torch.from_numpy(H)))[0]
    e_ref, it = uhf(H, S, ne, mult, fb)
    assert abs(e - e_ref) < 1e-10, (e, e_ref)
    np.testing.assert_approx_equal(e + e_nuc, -147.634028141, 6)


def _c2h4_oracle_matrices():
    from oracle import eri_c
    from oracle.basis import Nuclei, nuclear_repulsion
    from oracle.one_electron import kinetic_matrix, nuclear_matrix, overlap_matrix
    from tests.helpers import load_system
    nuc, basis, ne = load_system('C2H4.mol', '3-21G.gbs')
    nuc = [Nuclei(n.element, n.charge, n.mass, n.coordinates) for n in nuc]
    with eri_c.exact_boys():
        eri = eri_c.eri_dense(basis)
    S = overlap_matrix(basis)
    H = kinetic_matrix(basis) + nuclear_matrix(basis, nuc)
    return ne, eri, S, H, nuclear_repulsion(nuc)


class _OracleFock:
    """The product's Fock-builder interface over the oracle's dense sums."""

    def __init__(self, eri, H):
        import torch
        self.eri = eri
        self.H = H
        self.core_hamiltonian = torch.from_numpy(H)
        self.builds = 0

    def create(self, p):
        import torch
        from oracle.scf import jk_from_dense
        self.builds += 1
        J, K = jk_from_dense(self.eri, p.numpy())
        return torch.from_numpy(self.H + J - 0.5 * K)


def test_second_order_scf_host_logic_matches_oracle():
    """The device-side SecondOrderSCF (torch; CPU tensors here) driven by the
    oracle's dense Fock sums reaches the oracle's second-order result and the
    reference's published C2H4/3-21G energy."""
    import torch
    from gaussium_amd.scf import LinearAlgebra, RestrictedSCF
    from oracle.scf import jk_from_dense, rhf
    ne, eri, S, H, e_nuc = _c2h4_oracle_matrices()
    fock = _OracleFock(eri, H)
    la = LinearAlgebra(torch.from_numpy(S))
    scf = RestrictedSCF(la, ne, torch.from_numpy(S), fock, stabilizer='second_order')
    e, eps, c = scf.begin_iterations(*la.diagonalize(fock.core_hamiltonian))
    assert scf.converged
    e_ref, _, _, it = rhf(H, S, ne, lambda p: H + jk_from_dense(eri, p)[0] - 0.5 * jk_from_dense(eri, p)[1],
                          stabilizer='second_order')
    assert abs(e - e_ref) < 1e-10, (e, e_ref)
    np.testing.assert_approx_equal(e + e_nuc, -77.600460844, 6)
    # W from the returned canonical orbitals equals P F P / 2 at the solution
    no = ne // 2
    co = c[:, :no]
    W = 2.0 * (co * eps[:no]) @ co.T
    P = scf.density
    assert torch.allclose(W, 0.5 * P @ scf.fock @ P, atol=1e-7)


def test_second_order_hessian_matches_gradient_differences():
    """The Newton stage's orbital Hessian-vector product equals the central
    difference of the orbital gradient 4 C_vir^T F C_occ along C exp(t K)."""
    import torch
    from gaussium_amd.scf import LinearAlgebra, RestrictedSCF, SecondOrderSCF, _rotate
    ne, eri, S, H, e_nuc = _c2h4_oracle_matrices()
    fock = _OracleFock(eri, H)
    la = LinearAlgebra(torch.from_numpy(S))
    scf = RestrictedSCF(la, ne, torch.from_numpy(S), fock, max_iter=6)
    scf.begin_iterations(*la.diagonalize(fock.core_hamiltonian))  # a non-stationary point
    no = ne // 2
    so = SecondOrderSCF()
    c, eo, ev, _, _ = so._canonical(scf.orbital_coefficients, scf.fock, no)
    co, cv = c[:, :no], c[:, no:]
    diag = 4.0 * (ev[:, None] - eo[None, :])
    k = torch.from_numpy(np.random.default_rng(0).standard_normal(diag.shape))
    d = 2.0 * (cv @ k @ co.T)
    hk = diag * k + 4.0 * (cv.T @ (fock.create(d + d.T) - fock.core_hamiltonian) @ co)

    def grad_at(t):
        # the gradient in the frame of the rotated orbitals, whose t-derivative is H k
        ct = _rotate(c, t * k, no)
        return so._gradient(ct, fock.create(2.0 * ct[:, :no] @ ct[:, :no].T), no)

    h = 1e-4
    fd = (grad_at(h) - grad_at(-h)) / (2 * h)
    err = (fd - hk).abs().max()
    print('Hessian-vector product vs difference of gradients: %.2e (|Hk| %.2e)' % (float(err), float(hk.abs().max())))
    assert float(err) < 1e-6 * max(1.0, float(hk.abs().max()))


def test_bfgs_minimize_converges_and_never_climbs():
    """The optimiser behind start(..., geometry_optimization=...): it reaches
    the minimum of a Rosenbrock valley, and with a gradient that points uphill
    (every line search fails) it stops at the start instead of accepting an
    uphill point."""
    from gaussium_amd.geometry import bfgs_minimize

    def rosen(x):
        e = (1 - x[0]) ** 2 + 10 * (x[1] - x[0] ** 2) ** 2
        g = np.array([-2 * (1 - x[0]) - 40 * x[0] * (x[1] - x[0] ** 2), 20 * (x[1] - x[0] ** 2)])
        return e, g

    x, e, hist = bfgs_minimize(rosen, [-1.0, 1.5], gtol=1e-8, etol=0.0, max_steps=500)
    assert np.abs(x - 1.0).max() < 1e-6 and e < 1e-12
    assert all(b[0] <= a[0] for a, b in zip(hist, hist[1:]))
    calls = []

    def wrong_way(x):
        e, g = rosen(x)
        calls.append(e)
        return e, -g

    x0 = np.array([-1.0, 1.5])
    x, e, hist = bfgs_minimize(wrong_way, x0, max_steps=50)
    assert e == rosen(x0)[0] and np.array_equal(x, x0)
    assert len(hist) == 1 and len(calls) > 1
