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
                                 adds the iteration cap the reference lacks, and
                                 ``stabilizer='second_order'`` the opt-in
                                 ``SecondOrderSCF`` (damping, L-BFGS, trust-region
                                 Newton) in place of the DIIS loop.
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


def _rotate(c, kappa, nocc):
    """C exp(K) with K antisymmetric, K[vir, occ] = kappa, K[occ, vir] = -kappa^T."""
    n = c.shape[0]
    k = torch.zeros(n, n, dtype=c.dtype, device=c.device)
    k[nocc:, :nocc] = kappa
    k[:nocc, nocc:] = -kappa.T
    return c @ torch.linalg.matrix_exp(k)


def _dot(a, b):
    return float(torch.sum(a * b))


class SecondOrderSCF:
    """Opt-in convergence aid for the closed-shell SCF; the reference has none
    (its loop, hartreefock/scf_procedure.py:24-42, with the DIIS of
    diismethod/diis.py:30-112, cycles on dense random clusters whose core-H
    guess is far from any solution).

    It minimises the reference's energy functional E(P) = 1/2 tr P (H + F(P)),
    P = 2 C_occ C_occ^T, in three stages; every stage evaluates F through the
    caller's Fock builder, so screening and precision are the builder's.

    1. Optimal damping (Cancès and Le Bris, Int. J. Quantum Chem. 79, 82
       (2000)) from the core-H guess: a relaxed density Pt and Ft = F(Pt)
       (exact, F is affine in P) move towards each aufbau density by the
       lam in [0, 1] that minimises the quadratic
       E(Pt + lam D) = E(Pt) + lam tr(Ft D) + lam^2/2 tr((F - Ft) D), D = P - Pt;
       Ft is diagonalised next.  Until max |e| < ``switch`` for the
       orthonormalised commutator e = X^T (S P F - F P S) X of the aufbau
       density, or lam = 0.
    2. Direct minimisation over orbital rotations C exp(K),
       K[vir, occ] = kappa: gradient dE/dkappa_ai = 4 (C_vir^T F C_occ)_ai,
       L-BFGS (``history`` pairs) preconditioned by
       max(4 (e_a - e_i), ``precondition_floor``), steps capped at
       ``max_rotation`` in any element, backtracking line search on the
       energy.  After each step the orbitals are made canonical inside the
       occupied and the virtual block (the L-BFGS pairs are rotated with
       them), so the preconditioner is the diagonal of the orbital Hessian.
       Until max |dE/dkappa| < ``newton_switch``.
    3. Trust-region Newton (Steihaug conjugate gradients, at most ``max_cg``
       Hessian products per step, relative residual min(0.5, |g|^1/2)) with
       the exact orbital Hessian of the exponential parametrisation:
       (H kappa)_ai = 4 (e_a - e_i) kappa_ai + 4 (C_vir^T G(D) C_occ)_ai,
       G(D) = F(D) - H (J - K/2 of D),
       D = 2 (C_vir kappa C_occ^T + C_occ kappa^T C_vir^T).

    Converged when |dE| < the SCF threshold between accepted iterates and
    max |dE/dkappa| < ``gradient_threshold`` (and the SCF's density test, if
    any).  The solution is a stationary point of the same functional as the
    reference's fixed point (F commutes with P) and a local minimum over
    orbital rotations; the orbitals it returns are canonical in the occupied
    and virtual blocks.
    """

    def __init__(self, switch=0.1, newton_switch=1e-4, gradient_threshold=1e-7, history=20, max_rotation=0.3,
                 precondition_floor=0.2, trust_radius=0.5, max_cg=40):
        self.switch = switch
        self.newton_switch = newton_switch
        self.gradient_threshold = gradient_threshold
        self.history = history
        self.max_rotation = max_rotation
        self.precondition_floor = precondition_floor
        self.trust_radius = trust_radius
        self.max_cg = max_cg
        self.stage_iterations = {'damping': 0, 'lbfgs': 0, 'newton': 0}
        self.hessian_products = 0

    # -- pieces ---------------------------------------------------------------
    def _evaluate(self, scf, c):
        p = density_matrix_restricted(c, scf.electrons)
        f = scf.hamiltonian_matrix_factory.create(p)
        e = scf.calculate.restricted(p, f)
        scf.iterations += 1
        scf.energy_history.append(e)
        return p, f, e

    def _canonical(self, c, f, nocc):
        fmo = c.T @ f @ c
        eo, uo = torch.linalg.eigh(fmo[:nocc, :nocc])
        ev, uv = torch.linalg.eigh(fmo[nocc:, nocc:])
        return torch.cat([c[:, :nocc] @ uo, c[:, nocc:] @ uv], dim=1), eo, ev, uo, uv

    @staticmethod
    def _gradient(c, f, nocc):
        return 4.0 * (c[:, nocc:].T @ f @ c[:, :nocc])

    # -- driver -----------------------------------------------------------------
    def run(self, scf, orbital_coefficients):
        nocc = scf.electrons // 2
        la = scf.linear_algebra
        s = scf.overlap
        cap = scf.max_iter

        def out_of_budget():
            return cap is not None and scf.iterations >= cap

        # stage 1: optimal damping
        c = orbital_coefficients
        relaxed = None
        while True:
            p, f, e = self._evaluate(scf, c)
            self.stage_iterations['damping'] += 1
            lam = 1.0
            if relaxed is None:
                relaxed = (p, f)
            else:
                pt, ft = relaxed
                d = p - pt
                slope = _dot(ft, d)
                curv = _dot(f - ft, d)
                lam = 1.0 if curv <= -slope else -slope / curv
                lam = min(max(lam, 0.0), 1.0)
                relaxed = (pt + lam * d, ft + lam * (f - ft))
            err = float(la.orthonormalize(s @ p @ f - f @ p @ s).abs().max())
            if err < self.switch or lam == 0.0 or out_of_budget():
                break
            _, c = la.diagonalize(relaxed[1])
        relaxed = None

        c, eo, ev, _, _ = self._canonical(c, f, nocc)
        g = self._gradient(c, f, nocc)
        pairs = []
        stage = 'lbfgs'
        converged = False
        delta = self.trust_radius
        while not out_of_budget():
            gmax = float(g.abs().max())
            if stage == 'lbfgs' and gmax < self.newton_switch:
                stage = 'newton'
            diag = 4.0 * (ev[:, None] - eo[None, :])
            prec = torch.clamp(diag, min=self.precondition_floor)
            if stage == 'lbfgs':
                cn, pn, fn, en, step = self._lbfgs_step(scf, c, e, g, prec, pairs, nocc)
                self.stage_iterations['lbfgs'] += 1
                accepted = True
            else:
                cn, pn, fn, en, accepted, delta = self._newton_step(scf, c, e, g, diag, prec, delta, nocc)
                self.stage_iterations['newton'] += 1
            if not accepted:
                continue
            gn = self._gradient(cn, fn, nocc)
            if stage == 'lbfgs':
                y = gn - g
                sy = _dot(step, y)
                if sy > 1e-14:
                    pairs.append((step, y, 1.0 / sy))
                    if len(pairs) > self.history:
                        pairs.pop(0)
            de = e - en
            dp = float((pn - p).abs().max())
            c, eo, ev, uo, uv = self._canonical(cn, fn, nocc)
            g = uv.T @ gn @ uo
            pairs = [(uv.T @ a @ uo, uv.T @ b @ uo, r) for a, b, r in pairs]
            p, f, e = pn, fn, en
            if abs(de) < scf.threshold and float(g.abs().max()) < self.gradient_threshold and \
                    (scf.density_threshold is None or dp < scf.density_threshold):
                converged = True
                break
        return converged, p, f, e, torch.cat([eo, ev]), c

    def _lbfgs_step(self, scf, c, e, g, prec, pairs, nocc):
        q = g.clone()
        alphas = []
        for sv, yv, rho in reversed(pairs):
            a = rho * _dot(sv, q)
            q = q - a * yv
            alphas.append(a)
        r = q / prec
        for (sv, yv, rho), a in zip(pairs, reversed(alphas)):
            b = rho * _dot(yv, r)
            r = r + (a - b) * sv
        d = -r
        gd = _dot(g, d)
        if gd >= 0.0:  # not a descent direction: forget the pairs
            pairs.clear()
            d = -g / prec
            gd = _dot(g, d)
        m = float(d.abs().max())
        if m > self.max_rotation:
            d = d * (self.max_rotation / m)
            gd *= self.max_rotation / m
        t = 1.0
        for _ in range(12):
            cn = _rotate(c, t * d, nocc)
            pn, fn, en = self._evaluate(scf, cn)
            if en <= e + 1e-4 * t * gd + 1e-11:
                break
            tq = -gd * t * t / (2.0 * (en - e - gd * t))
            t = min(0.5 * t, max(0.1 * t, tq))
        return cn, pn, fn, en, t * d

    def _newton_step(self, scf, c, e, g, diag, prec, delta, nocc):
        co, cv = c[:, :nocc], c[:, nocc:]
        h = scf.hamiltonian_matrix_factory.core_hamiltonian

        def hess(k):
            d = 2.0 * (cv @ k @ co.T)
            self.hessian_products += 1
            return diag * k + 4.0 * (cv.T @ (scf.hamiltonian_matrix_factory.create(d + d.T) - h) @ co)

        def to_boundary(x, d):
            xx, xd, dd = _dot(x, x), _dot(x, d), _dot(d, d)
            return x + ((-xd + (xd * xd + dd * (delta * delta - xx)) ** 0.5) / dd) * d

        gnorm = float(g.norm())
        eta = min(0.5, gnorm ** 0.5)
        x = torch.zeros_like(g)
        r = g.clone()
        z = r / prec
        d = -z
        rz = _dot(r, z)
        hit = False
        for _ in range(self.max_cg):
            hd = hess(d)
            dhd = _dot(d, hd)
            if dhd <= 0.0:
                x, hit = to_boundary(x, d), True
                break
            alpha = rz / dhd
            xn = x + alpha * d
            if float(xn.norm()) >= delta:
                x, hit = to_boundary(x, d), True
                break
            x = xn
            r = r + alpha * hd
            if float(r.norm()) < eta * gnorm:
                break
            z = r / prec
            rzn = _dot(r, z)
            d = -z + (rzn / rz) * d
            rz = rzn
        predicted = _dot(g, x) + 0.5 * _dot(x, hess(x))
        cn = _rotate(c, x, nocc)
        pn, fn, en = self._evaluate(scf, cn)
        de = en - e
        rho = de / predicted if predicted < -1e-11 else 1.0
        if rho < 0.25:
            delta *= 0.25
        elif rho > 0.75 and hit:
            delta = min(2.0 * delta, 2.0)
        return cn, pn, fn, en, de < 1e-11, delta


class RestrictedSCF:
    def __init__(self, linear_algebra, electrons, overlap, hamiltonian_matrix_factory, threshold=1e-12,
                 max_iter=None, use_diis=True, density_threshold=None, stabilizer=None):
        self.linear_algebra = linear_algebra
        self.electrons = electrons
        self.overlap = overlap
        self.hamiltonian_matrix_factory = hamiltonian_matrix_factory
        self.threshold = threshold
        self.max_iter = max_iter
        self.calculate = TotalEnergy(hamiltonian_matrix_factory.core_hamiltonian)
        # ``stabilizer``: None keeps the reference's DIIS loop; 'second_order'
        # (or a SecondOrderSCF) replaces it
        if stabilizer == 'second_order':
            stabilizer = SecondOrderSCF()
        elif stabilizer is not None and not isinstance(stabilizer, SecondOrderSCF):
            raise ValueError('unknown SCF stabilizer %r (None or "second_order")' % (stabilizer,))
        self.stabilizer = stabilizer
        self.diis = DIIS(overlap, linear_algebra) if use_diis and stabilizer is None else None
        self.iterations = 0
        self.energy_history = []
        self.converged = False
        # optional second stop test, max |P_n - P_{n-1}| below this (gradients
        # are first order in the density error, energies second order)
        self.density_threshold = density_threshold

    def begin_iterations(self, orbital_energies, orbital_coefficients):
        if self.stabilizer is not None:
            self.converged, density, fock, total, orbital_energies, orbital_coefficients = \
                self.stabilizer.run(self, orbital_coefficients)
            self.density = density
            self.fock = fock
            self.orbital_energies = orbital_energies
            self.orbital_coefficients = orbital_coefficients
            return total, orbital_energies, orbital_coefficients
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
