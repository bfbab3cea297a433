"""Drop-in RHF driver and the reference's operator call sites.

* ``TwoElectronRepulsionMatrixOS`` -- same constructor and methods as
  matrixelements/electron_repulsion_matrix.py:9-53 of the reference
  (``calculate_integral(i, j, k, l)``, ``create_repulsion_matrix()``), backed
  by the gx class kernels instead of a ``multiprocessing.Pool``.
* ``FockMatrixRestricted`` -- the Fock-builder plugin of
  hartreefock/fock_matrix.py:5-22 (``create(P) -> F``, ``.core_hamiltonian``),
  computing G = J - K/2 by the direct fused J/K kernel; no N^4 tensor.
* ``RestrictedHF`` -- hartreefock/hartree_fock.py:18-84: S, T, V on the
  device, canonical orthogonalisation, core-H guess, DIIS SCF.
* ``FockMatrixUnrestricted`` / ``UnrestrictedHF`` -- fock_matrix.py:25-55,
  hartree_fock.py:108-116 (Pople-Nesbet-Berthier).
* ``BlockedFockMatrixUnrestricted`` / ``BlockedHartreeFock`` --
  fock_matrix.py:56-81, hartree_fock.py:119-139 ('GUHF': the unrestricted SCF
  on the 2N spin-blocked basis with DIIS).

Options the reference does not have (``schwarz_cutoff``, ``fp32_bound``,
``precision``, ``max_iter``, ``threshold``, ``stabilizer``,
``initial_density``) default to the reference's behaviour: no screening, all
FP64, no iteration cap, |dE| < 1e-12, the reference's DIIS, the core-H
guess.  ``stabilizer='second_order'`` (RHF) replaces the DIIS loop by
``scf.SecondOrderSCF`` (optimal damping, L-BFGS, trust-region Newton).
"""
import time

import numpy as np
import torch

from .engine import Engine
from .scf import BlockedLinearAlgebra, BlockedUnrestrictedSCF, LinearAlgebra, PopleNesbetBerthier, RestrictedSCF


class TwoElectronRepulsionMatrixOS:
    """ERI operator with the reference's interface (electron_repulsion_matrix.py:9-53)."""

    def __init__(self, basis_set_array, symmetry=None, processes=1, engine=None):
        self.basis_set_array = basis_set_array
        self.matrix_size = len(basis_set_array)
        self.symmetry = symmetry
        self.processes = processes  # accepted for signature parity; the GPU path ignores it
        self.engine = engine if engine is not None else Engine(basis_set_array)

    def calculate_integral(self, i, j, k, l):
        if self.symmetry is not None and hasattr(self.symmetry, 'none_zero_integral') \
                and not self.symmetry.none_zero_integral((i, j, k, l)):
            return 0.0
        return float(self.engine.eri_quartets([(i, j, k, l)])[0])

    def create_repulsion_matrix(self):
        """Dense [N,N,N,N] numpy tensor (small N; the SCF path never forms it).
        With a ``Symmetry`` of a point group other than C_1, the integrals it
        forces to zero are exactly 0.0 (electron_repulsion_matrix.py:18-26)."""
        out = self.engine.eri_dense().cpu().numpy()
        if self.symmetry is not None and hasattr(self.symmetry, 'zero_mask'):
            out[self.symmetry.zero_mask()] = 0.0
        return out


class FockMatrixRestricted:
    """F = H + J - K/2 from the direct fused J/K build (fock_matrix.py:12-22)."""

    def __init__(self, core_hamiltonian, engine, schwarz_cutoff=0.0, fp32_bound=0.0, precision='fp64'):
        self.core_hamiltonian = core_hamiltonian
        self.engine = engine
        self.schwarz_cutoff = schwarz_cutoff
        self.fp32_bound = fp32_bound
        self.mixed = precision == 'mixed'
        self.matrix_size = core_hamiltonian.shape[0]
        self.builds = 0

    def create(self, density_matrix):
        J, K = self.engine.jk(density_matrix, cutoff=self.schwarz_cutoff, fp32_bound=self.fp32_bound,
                              mixed=self.mixed)
        self.builds += 1
        return self.core_hamiltonian + J - 0.5 * K


class FockMatrixUnrestricted:
    """F_a = H + J(P_a + P_b) - K(P_a), F_b = H + J(P_a + P_b) - K(P_b)
    (hartreefock/fock_matrix.py:25-55).  FP64: one two-density build
    (``gx_jk_uhf``: the integrals are evaluated once for both spins); mixed
    precision: one build per spin density (J is linear, so J(P_a) + J(P_b)
    is the Coulomb matrix of the total)."""

    def __init__(self, core_hamiltonian, engine, schwarz_cutoff=0.0, fp32_bound=0.0, precision='fp64'):
        self.core_hamiltonian = core_hamiltonian
        self.engine = engine
        self.schwarz_cutoff = schwarz_cutoff
        self.fp32_bound = fp32_bound
        self.mixed = precision == 'mixed'
        self.builds = 0

    def create(self, density_matrix_alph, density_matrix_beta):
        if self.mixed:
            kw = dict(cutoff=self.schwarz_cutoff, fp32_bound=self.fp32_bound, mixed=True)
            ja, ka = self.engine.jk(density_matrix_alph, **kw)
            jb, kb = self.engine.jk(density_matrix_beta, **kw)
            j = ja + jb
        else:
            j, ka, kb = self.engine.jk_uhf(density_matrix_alph, density_matrix_beta, cutoff=self.schwarz_cutoff)
        self.builds += 1
        return self.core_hamiltonian + j - ka, self.core_hamiltonian + j - kb


class BlockedFockMatrixUnrestricted:
    """F = H + J - K on the spin-blocked 2N basis (hartreefock/fock_matrix.py:56-81,
    over the blocked tensor of matrixelements/transformations.py:29-71).

    With spin blocks P_aa, P_bb of the blocked density the diagonal blocks are
    H + J(P_aa + P_bb) - K(P_aa) and H + J(P_aa + P_bb) - K(P_bb): one
    two-density direct build (``gx_jk_uhf``) in FP64, two single-density
    builds in mixed precision; no 2N tensor is formed.  The spin-off-diagonal
    blocks of F are -K of the transposed off-diagonal density blocks; the
    blocked SCF keeps those density blocks exactly zero (its orbitals are
    spin-pure), and a density with nonzero ones is rejected."""

    def __init__(self, core_hamiltonian, engine, schwarz_cutoff=0.0, fp32_bound=0.0, precision='fp64'):
        self.core_hamiltonian = core_hamiltonian  # blockdiag(H, H)
        self.engine = engine
        self.schwarz_cutoff = schwarz_cutoff
        self.fp32_bound = fp32_bound
        self.mixed = precision == 'mixed'
        self.builds = 0

    def create(self, density_matrix):
        n = density_matrix.shape[0] // 2
        if bool(torch.any(density_matrix[:n, n:] != 0)) or bool(torch.any(density_matrix[n:, :n] != 0)):
            raise ValueError('blocked Fock build: the density has alpha-beta blocks; only spin-blocked '
                             'densities (spin-pure orbitals) are supported')
        pa = density_matrix[:n, :n].contiguous()
        pb = density_matrix[n:, n:].contiguous()
        if self.mixed:
            kw = dict(cutoff=self.schwarz_cutoff, fp32_bound=self.fp32_bound, mixed=True)
            ja, ka = self.engine.jk(pa, **kw)
            jb, kb = self.engine.jk(pb, **kw)
            j = ja + jb
        else:
            j, ka, kb = self.engine.jk_uhf(pa, pb, cutoff=self.schwarz_cutoff)
        self.builds += 1
        fock = self.core_hamiltonian.clone()
        fock[:n, :n] += j - ka
        fock[n:, n:] += j - kb
        return fock


class RestrictedHF:
    """Restricted Hartree-Fock (hartree_fock.py:18-84) on one GPU."""

    def __init__(self, nuclei_array, basis_set_array, electrons, symmetry=None, processes=1,
                 schwarz_cutoff=0.0, fp32_bound=0.0, precision='fp64', max_iter=None, engine=None,
                 verbose=False, threshold=1e-12, density_threshold=None, stabilizer=None, initial_density=None):
        self.nuclei_array = nuclei_array
        self.basis_set_array = basis_set_array
        self.electrons = electrons
        self.symmetry = symmetry
        self.verbose = verbose
        t0 = time.perf_counter()
        self.engine = engine if engine is not None else Engine(basis_set_array, nuclei_array)
        if engine is not None and basis_set_array is not None:
            self.engine.set_basis(basis_set_array, nuclei_array)
            self.engine.set_nuclei(nuclei_array)
        S, T, V = self.engine.one_electron()
        self.orbital_overlap = S
        self.kinetic_energy = T
        self.nuclear_attraction = V
        self.core_hamiltonian = T + V
        self.linear_algebra = LinearAlgebra(S)
        self.fock_builder = FockMatrixRestricted(self.core_hamiltonian, self.engine, schwarz_cutoff,
                                                 fp32_bound, precision)
        self.schwarz_cutoff = schwarz_cutoff
        self.scf_method = RestrictedSCF(self.linear_algebra, electrons, S, self.fock_builder,
                                        threshold=threshold, max_iter=max_iter,
                                        density_threshold=density_threshold, stabilizer=stabilizer)
        self.initial_density = initial_density
        self.setup_time = time.perf_counter() - t0
        self._repulsion = None

    @property
    def repulsion(self):
        """Dense ERI tensor (numpy), formed only on request (post-HF consumers, small N)."""
        if self._repulsion is None:
            self._repulsion = self.engine.eri_dense().cpu().numpy()
        return self._repulsion

    def initial_guess(self):
        """Core-Hamiltonian guess (hartree_fock.py:62-73); with ``initial_density``
        (e.g. the converged density of a nearby geometry) the orbitals of
        F(initial_density) instead."""
        if self.initial_density is not None:
            p0 = torch.as_tensor(self.initial_density, dtype=torch.float64, device=self.core_hamiltonian.device)
            return self.linear_algebra.diagonalize(self.fock_builder.create(p0))
        return self.linear_algebra.diagonalize(self.core_hamiltonian)

    @property
    def smallest_overlap_eigenvalue(self):
        return self.linear_algebra.smallest_overlap_eigenvalue

    def begin_scf(self):
        e0, c0 = self.initial_guess()
        t0 = time.perf_counter()
        energy, eps, c = self.scf_method.begin_iterations(e0, c0)
        self.scf_time = time.perf_counter() - t0
        if self.verbose:
            print('SCF converged in %d iterations, %.3f s' % (self.scf_method.iterations, self.scf_time))
        return energy, eps, c

    def energies(self):
        return self.begin_scf()[0], 0.0

    def gradient(self):
        """Analytic nuclear gradient dE_total/dR [natom, 3] (device tensor) at the
        converged SCF: the density and the energy-weighted density
        W = 2 sum_occ e_i C_i C_i^T of the last iteration (runs the SCF first if
        it has not run), with the Fock builder's screening and precision."""
        if not hasattr(self.scf_method, 'density'):
            self.begin_scf()
        eps = self.scf_method.orbital_energies
        c = self.scf_method.orbital_coefficients[:, :self.electrons // 2]
        W = 2.0 * (c * eps[:self.electrons // 2]) @ c.T
        fb = self.fock_builder
        return self.engine.grad(self.scf_method.density, W, cutoff=self.schwarz_cutoff, fp32_bound=fb.fp32_bound,
                                mixed=fb.mixed)


class UnrestrictedHF:
    """Unrestricted Hartree-Fock (hartree_fock.py:87-122, Pople-Nesbet-Berthier
    iterations of scf_procedure.py:45-89) on one GPU."""

    def __init__(self, nuclei_array, basis_set_array, electrons, multiplicity, symmetry=None, processes=1,
                 schwarz_cutoff=0.0, fp32_bound=0.0, precision='fp64', max_iter=None, engine=None,
                 threshold=1e-12):
        self.nuclei_array = nuclei_array
        self.basis_set_array = basis_set_array
        self.electrons = electrons
        self.multiplicity = multiplicity
        self.symmetry = symmetry
        self.engine = engine if engine is not None else Engine(basis_set_array, nuclei_array)
        if engine is not None and basis_set_array is not None:
            self.engine.set_basis(basis_set_array, nuclei_array)
            self.engine.set_nuclei(nuclei_array)
        S, T, V = self.engine.one_electron()
        self.orbital_overlap = S
        self.core_hamiltonian = T + V
        self.linear_algebra = LinearAlgebra(S)
        self.fock_builder = FockMatrixUnrestricted(self.core_hamiltonian, self.engine, schwarz_cutoff,
                                                   fp32_bound, precision)
        self.scf_method = PopleNesbetBerthier(self.linear_algebra, electrons, multiplicity, self.fock_builder,
                                              threshold=threshold, max_iter=max_iter)

    def initial_guess(self):
        return self.linear_algebra.diagonalize(self.core_hamiltonian)

    def begin_scf(self):
        e0, c0 = self.initial_guess()
        return self.scf_method.begin_iterations(e0, c0)

    def energies(self):
        return self.begin_scf()[0], 0.0


class BlockedHartreeFock:
    """'GUHF' (hartree_fock.py:119-139): the unrestricted SCF on the spin-blocked
    basis, S and H as blockdiag(S, S) and blockdiag(H, H), blocked canonical
    orthogonalisation, DIIS (scf_procedure.py:92-120), on one GPU."""

    def __init__(self, nuclei_array, basis_set_array, electrons, multiplicity, symmetry=None, processes=1,
                 schwarz_cutoff=0.0, fp32_bound=0.0, precision='fp64', max_iter=None, engine=None,
                 threshold=1e-12):
        self.nuclei_array = nuclei_array
        self.basis_set_array = basis_set_array
        self.electrons = electrons
        self.multiplicity = multiplicity
        self.symmetry = symmetry
        self.engine = engine if engine is not None else Engine(basis_set_array, nuclei_array)
        if engine is not None and basis_set_array is not None:
            self.engine.set_basis(basis_set_array, nuclei_array)
            self.engine.set_nuclei(nuclei_array)
        S, T, V = self.engine.one_electron()
        self.orbital_overlap = torch.block_diag(S, S)
        self.core_hamiltonian = torch.block_diag(T + V, T + V)
        self.linear_algebra = BlockedLinearAlgebra(self.orbital_overlap)
        self.fock_builder = BlockedFockMatrixUnrestricted(self.core_hamiltonian, self.engine, schwarz_cutoff,
                                                          fp32_bound, precision)
        self.scf_method = BlockedUnrestrictedSCF(self.linear_algebra, electrons, multiplicity,
                                                 self.orbital_overlap, self.fock_builder, threshold=threshold,
                                                 max_iter=max_iter)

    def initial_guess(self):
        return self.linear_algebra.diagonalize(self.core_hamiltonian)

    def begin_scf(self):
        e0, c0 = self.initial_guess()
        return self.scf_method.begin_iterations(e0, c0)

    def energies(self):
        return self.begin_scf()[0], 0.0
