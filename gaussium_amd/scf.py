"""Restricted Hartree-Fock SCF on the device (FP64 torch tensors).

Mirrors the reference's classes and their numerical conventions:

* ``density_matrix_restricted``  P = 2 C_occ C_occ^T (matrixelements/density_matrix.py:4-9)
* ``TotalEnergy.restricted``     1/2 sum P (H + F) (hartreefock/total_energy.py:16-35)
* ``LinearAlgebra``              X = U s^-1/2 without cutoff, X^T F X, eigh,
                                 ascending order, C = X C' (hartreefock/linear_algebra.py:4-31)
* ``DIIS``                       e = S P F - F P S in the orthonormal basis; signed
                                 on/off tests any(e < 0.1) / all(e < 1e-6); reset
                                 after 8 vectors; pop the oldest vector when B is
                                 singular (diismethod/diis.py:30-112).  B is also
                                 treated as singular when cond(B) > 1e12: with very
                                 few basis functions the error matrices are
                                 parallel, B is singular to working precision and
                                 a plain solve returns arbitrary coefficients.
* ``RestrictedSCF``              core-H guess, stop at |dE| < 1e-12
                                 (hartreefock/scf_procedure.py:17-42); ``max_iter``
                                 adds the iteration cap the reference lacks.
* ``PopleNesbetBerthier``        unrestricted SCF (scf_procedure.py:45-89)
* ``BlockedUnrestrictedSCF``     unrestricted SCF on the spin-blocked 2N basis with
                                 ``BlockedLinearAlgebra`` and DIIS (scf_procedure.py:92-120,
                                 linear_algebra.py:36-98)
"""
import numpy as np
import torch

DIIS_COND_LIMIT = 1e12


def density_matrix_restricted(c, electrons):
    occ = c[:, :electrons // 2]
    return 2.0 * (occ @ occ.T)


def density_matrix_unrestricted(c, electrons):
    """P = C_occ C_occ^T for one spin (matrixelements/density_matrix.py:12-17)."""
    occ = c[:, :electrons]
    return occ @ occ.T


def blocked_density_matrix(c, electrons_alph, electrons_beta):
    """Spin-blocked P (matrixelements/density_matrix.py:20-36): alpha columns
    [0, n_a) and beta columns [N, N + n_b) of the 2N x 2N coefficients."""
    half = c.shape[0] // 2
    a = c[:, :electrons_alph]
    b = c[:, half:half + electrons_beta]
    return a @ a.T + b @ b.T


class TotalEnergy:
    def __init__(self, core_hamiltonian):
        self.core_hamiltonian = core_hamiltonian

    def restricted(self, density_matrix, hamiltonian):
        return 0.5 * float(torch.sum(density_matrix * (self.core_hamiltonian + hamiltonian)))

    def unrestricted(self, density_alpha, density_beta, fock_alpha, fock_beta):
        """1/2 sum (P_a + P_b) H + P_a F_a + P_b F_b (hartreefock/total_energy.py:37-63)."""
        return 0.5 * float(torch.sum((density_alpha + density_beta) * self.core_hamiltonian
                                     + density_alpha * fock_alpha + density_beta * fock_beta))


class LinearAlgebra:
    def __init__(self, orbital_overlap):
        w, u = torch.linalg.eigh(orbital_overlap)
        order = torch.argsort(w)
        w, u = w[order], u[:, order]
        self.smallest_overlap_eigenvalue = float(w[0])
        self.transformation_matrix = u * w.rsqrt()

    def diagonalize(self, fock_matrix):
        x = self.transformation_matrix
        e, c = torch.linalg.eigh(x.T @ fock_matrix @ x)
        order = torch.argsort(e)
        return e[order], x @ c[:, order]

    def orthonormalize(self, matrix):
        x = self.transformation_matrix
        return x.T @ matrix @ x


class BlockedLinearAlgebra(LinearAlgebra):
    """Spin-blocked linear algebra (hartreefock/linear_algebra.py:36-98): X
    from each spin block of the blocked overlap on its own; X^T F X
    diagonalised per spin block (ascending each), C = X blockdiag(C_a, C_b),
    energies alpha then beta.  The orbitals stay spin-pure."""

    def __init__(self, orbital_overlap):
        n = orbital_overlap.shape[0] // 2
        self.half = n
        x = torch.zeros_like(orbital_overlap)
        smallest = []
        for b in (0, 1):
            w, u = torch.linalg.eigh(orbital_overlap[b * n:(b + 1) * n, b * n:(b + 1) * n])
            order = torch.argsort(w)
            w, u = w[order], u[:, order]
            smallest.append(float(w[0]))
            x[b * n:(b + 1) * n, b * n:(b + 1) * n] = u * w.rsqrt()
        self.smallest_overlap_eigenvalue = min(smallest)
        self.transformation_matrix = x

    def diagonalize(self, fock_matrix):
        n = self.half
        x = self.transformation_matrix
        fo = x.T @ fock_matrix @ x
        c = torch.zeros_like(fock_matrix)
        es = []
        for b in (0, 1):
            e, v = torch.linalg.eigh(fo[b * n:(b + 1) * n, b * n:(b + 1) * n])
            order = torch.argsort(e)
            es.append(e[order])
            c[b * n:(b + 1) * n, b * n:(b + 1) * n] = v[:, order]
        return torch.cat(es), x @ c


class DIIS:
    def __init__(self, overlap, linear_algebra, diis_subspace=8, cond_limit=DIIS_COND_LIMIT):
        self.cond_limit = cond_limit
        self.overlap = overlap
        self.linear_algebra = linear_algebra
        self.fock_array = []
        self.error_array = []
        self.dots = np.zeros((0, 0))  # tr(e_i e_j^T) of the stored error matrices
        self.diis_subspace = diis_subspace
        self.begin = False

    def fock_matrix(self, fock, density):
        error = self.overlap @ density @ fock - fock @ density @ self.overlap
        error = self.linear_algebra.orthonormalize(error)
        if bool(torch.any(error < 0.1)):
            self.begin = True
        if bool(torch.all(error < 1e-6)):
            self.begin = False
        if self.begin:
            # new row of tr(e_i e_j^T): one elementwise product and reduction on the device
            row = (torch.stack(self.error_array + [error]) * error).sum(dim=(1, 2)).cpu().numpy() \
                if self.error_array else (error * error).sum().reshape(1).cpu().numpy()
            n = len(self.error_array)
            dots = np.zeros((n + 1, n + 1))
            dots[:n, :n] = self.dots
            dots[n, :] = row
            dots[:, n] = row
            self.dots = dots
            self.error_array.append(error)
            self.fock_array.append(fock)
            if len(self.fock_array) > self.diis_subspace:
                self.fock_array = []
                self.error_array = []
                self.dots = np.zeros((0, 0))
                return fock
            if len(self.fock_array) > 1:
                coef = self.create_b_matrix()
                fock = torch.zeros_like(fock)
                for c, f in zip(coef.tolist(), self.fock_array):
                    fock = fock + c * f
        return fock

    def create_b_matrix(self):
        # B_ij = tr(e_i e_j^T) (kept incrementally); the (n+1)^2 <= 81-element solve on the host.
        n = len(self.fock_array)
        b = np.ones((n + 1, n + 1))
        b[:n, :n] = self.dots
        b[n, n] = 0.0
        rhs = np.zeros(n + 1)
        rhs[n] = 1.0
        try:
            if self.cond_limit is not None and np.linalg.cond(b) > self.cond_limit:
                raise np.linalg.LinAlgError('ill-conditioned DIIS matrix')
            return np.linalg.solve(b, rhs)[:n]
        except np.linalg.LinAlgError:
            self.fock_array.pop(0)
            self.error_array.pop(0)
            self.dots = self.dots[1:, 1:]
            return self.create_b_matrix()


class RestrictedSCF:
    def __init__(self, linear_algebra, electrons, overlap, hamiltonian_matrix_factory, threshold=1e-12,
                 max_iter=None, use_diis=True, density_threshold=None):
        self.linear_algebra = linear_algebra
        self.electrons = electrons
        self.hamiltonian_matrix_factory = hamiltonian_matrix_factory
        self.threshold = threshold
        self.max_iter = max_iter
        self.calculate = TotalEnergy(hamiltonian_matrix_factory.core_hamiltonian)
        self.diis = DIIS(overlap, linear_algebra) if use_diis else None
        self.iterations = 0
        self.energy_history = []
        # optional second stop test, max |P_n - P_{n-1}| below this (gradients
        # are first order in the density error, energies second order)
        self.density_threshold = density_threshold

    def begin_iterations(self, orbital_energies, orbital_coefficients):
        previous = 0.0
        previous_density = None
        while True:
            density = density_matrix_restricted(orbital_coefficients, self.electrons)
            fock = self.hamiltonian_matrix_factory.create(density)
            total = self.calculate.restricted(density, fock)
            delta = previous - total
            previous = total
            self.iterations += 1
            self.energy_history.append(total)
            converged = abs(delta) < self.threshold
            if converged and self.density_threshold is not None:
                converged = previous_density is not None and \
                    float((density - previous_density).abs().max()) < self.density_threshold
            previous_density = density
            self.converged = converged
            if converged:
                break
            if self.max_iter is not None and self.iterations >= self.max_iter:
                break
            if self.diis is not None:
                fock = self.diis.fock_matrix(fock, density)
            orbital_energies, orbital_coefficients = self.linear_algebra.diagonalize(fock)
        self.density = density
        self.fock = fock
        self.orbital_energies = orbital_energies
        self.orbital_coefficients = orbital_coefficients
        return total, orbital_energies, orbital_coefficients


class PopleNesbetBerthier:
    """Unrestricted SCF (hartreefock/scf_procedure.py:45-89): n_a = (n + m - 1) / 2,
    n_b = (n - m + 1) / 2 electrons; the first iteration takes the alpha density
    from every orbital of the guess and an empty beta density, as the
    reference does; no DIIS; stop at |dE| < threshold (plus the optional
    iteration cap)."""

    def __init__(self, linear_algebra, electrons, multiplicity, hamiltonian_matrix_factory, threshold=1e-12,
                 max_iter=None):
        self.linear_algebra = linear_algebra
        self.electrons = electrons
        self.hamiltonian_matrix_factory = hamiltonian_matrix_factory
        self.threshold = threshold
        self.max_iter = max_iter
        self.calculate = TotalEnergy(hamiltonian_matrix_factory.core_hamiltonian)
        self.electrons_alph = (electrons + multiplicity - 1) // 2
        self.electrons_beta = (electrons - multiplicity + 1) // 2
        self.iterations = 0
        self.converged = False

    def begin_iterations(self, orbital_energies, orbital_coefficients):
        ca = cb = orbital_coefficients
        ea = eb = orbital_energies
        total = previous = 0.0
        while True:
            if self.iterations == 0:
                pa = density_matrix_unrestricted(ca, ca.shape[0])
                pb = density_matrix_unrestricted(cb, 0)
            else:
                pa = density_matrix_unrestricted(ca, self.electrons_alph)
                pb = density_matrix_unrestricted(cb, self.electrons_beta)
            fa, fb = self.hamiltonian_matrix_factory.create(pa, pb)
            total = self.calculate.unrestricted(pa, pb, fa, fb)
            delta = previous - total
            previous = total
            self.iterations += 1
            self.converged = abs(delta) < self.threshold
            if self.converged or (self.max_iter is not None and self.iterations >= self.max_iter):
                break
            ea, ca = self.linear_algebra.diagonalize(fa)
            eb, cb = self.linear_algebra.diagonalize(fb)
        self.density_alpha, self.density_beta = pa, pb
        return total, ea, eb, ca, cb


class BlockedUnrestrictedSCF:
    """Unrestricted SCF on the spin-blocked basis (hartreefock/scf_procedure.py:92-120):
    n_a = (n + m - 1) / 2, n_b = (n - m + 1) / 2; the first density takes every
    alpha column of the guess and no beta one, as the reference does; DIIS on
    the 2N matrices; E = 1/2 sum P (H + F); stop at |dE| < threshold (plus the
    optional iteration cap)."""

    def __init__(self, linear_algebra, electrons, multiplicity, overlap, hamiltonian_matrix_factory,
                 threshold=1e-12, max_iter=None):
        self.linear_algebra = linear_algebra
        self.electrons = electrons
        self.hamiltonian_matrix_factory = hamiltonian_matrix_factory
        self.threshold = threshold
        self.max_iter = max_iter
        self.calculate = TotalEnergy(hamiltonian_matrix_factory.core_hamiltonian)
        self.electrons_alph = (electrons + multiplicity - 1) // 2
        self.electrons_beta = (electrons - multiplicity + 1) // 2
        self.diis = DIIS(overlap, linear_algebra)
        self.iterations = 0
        self.converged = False

    def begin_iterations(self, orbital_energies, orbital_coefficients):
        half = orbital_coefficients.shape[0] // 2
        total = previous = 0.0
        while True:
            if self.iterations == 0:
                density = blocked_density_matrix(orbital_coefficients, half, 0)
            else:
                density = blocked_density_matrix(orbital_coefficients, self.electrons_alph, self.electrons_beta)
            fock = self.hamiltonian_matrix_factory.create(density)
            total = self.calculate.restricted(density, fock)
            delta = previous - total
            previous = total
            self.iterations += 1
            self.converged = abs(delta) < self.threshold
            if self.converged or (self.max_iter is not None and self.iterations >= self.max_iter):
                break
            fock = self.diis.fock_matrix(fock, density)
            orbital_energies, orbital_coefficients = self.linear_algebra.diagonalize(fock)
        self.density = density
        self.fock = fock
        return total, orbital_energies, orbital_coefficients
