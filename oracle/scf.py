"""Restricted Hartree-Fock with DIIS, restated from the reference.

* density  P = 2 sum_{i < n/2} c_i c_i^T        ``matrixelements/density_matrix.py:4-9``
* Fock     F = H + G, G_ij = sum_ab P_ab [(ij|ab) - 1/2 (ib|aj)]
                                               ``hartreefock/fock_matrix.py:5-22``
* energy   E = 1/2 sum_ij P_ij (H_ij + F_ij)   ``hartreefock/total_energy.py:16-35``
* X = U s^-1/2 (canonical, no cutoff), diagonalise via X^T F X, eigh,
  ascending sort, C = X C'                     ``hartreefock/linear_algebra.py:4-31``
* DIIS (signed on/off tests, 8-vector reset, pop-oldest on singular B)
                                               ``diismethod/diis.py:30-112``
* SCF loop: core-H guess, stop at |dE| < 1e-12 ``hartreefock/scf_procedure.py:17-42``,
                                               ``hartreefock/hartree_fock.py:45-73``

Test infrastructure only; see ``oracle/__init__.py``.
"""
import numpy as np


def density_matrix_restricted(c, electrons):
    """density_matrix.py:4-9."""
    occ = c[:, :electrons // 2]
    return 2.0 * occ @ occ.T


def jk_from_dense(eri, p):
    """J_ij = sum_kl (ij|kl) P_kl and K_ij = sum_kl (ik|jl) P_kl from a dense tensor."""
    j = np.einsum('ijkl,kl->ij', eri, p, optimize=True)
    k = np.einsum('ikjl,kl->ij', eri, p, optimize=True)
    return j, k


def fock_restricted_loop(h, eri, p):
    """F = H + G with the reference's element loop (fock_matrix.py:12-22); small N only."""
    n = h.shape[0]
    g = np.zeros_like(h)
    for i in range(n):
        for j in range(i, n):
            s = 0.0
            for a in range(n):
                for b in range(n):
                    s += p[a, b] * (eri[i, j, a, b] - 0.5 * eri[i, b, a, j])
            g[i, j] = g[j, i] = s
    return h + g


def total_energy_restricted(h, p, f):
    """total_energy.py:16-35."""
    return 0.5 * float(np.sum(p * (h + f)))


class LinearAlgebra:
    """linear_algebra.py:4-31."""

    def __init__(self, s):
        w, u = np.linalg.eigh(s)
        order = np.argsort(w)
        w, u = w[order], u[:, order]
        self.x = u @ np.diag(w ** -0.5)

    def diagonalize(self, f):
        fp = self.x.T @ f @ self.x
        e, c = np.linalg.eigh(fp)
        order = np.argsort(e)
        return e[order], self.x @ c[:, order]

    def orthonormalize(self, m):
        return self.x.T @ m @ self.x


class DIIS:
    """diis.py:30-112.

    ``cond_limit``: None keeps the reference's test (only an exactly singular
    B raises LinAlgError); a number also treats cond(B) > cond_limit as
    singular, which triggers the reference's own recovery (drop the oldest
    vector, diis.py:105-112).  With two basis functions the error matrices
    are all parallel and B is singular to working precision (cond ~ 1e18),
    where LAPACK returns arbitrary coefficients.
    """

    def __init__(self, s, la, subspace=8, cond_limit=None):
        self.cond_limit = cond_limit
        self.s = s
        self.la = la
        self.focks = []
        self.errors = []
        self.subspace = subspace
        self.begin = False

    def fock_matrix(self, f, p):
        e = self.s @ p @ f - f @ p @ self.s
        e = self.la.orthonormalize(e)
        if np.any(e < 0.1):
            self.begin = True
        if np.all(e < 1e-6):
            self.begin = False
        if self.begin:
            self.errors.append(e)
            self.focks.append(f)
            if len(self.focks) > self.subspace:
                self.focks = []
                self.errors = []
                return f
            if len(self.focks) > 1:
                coef = self._coefficients()
                f = np.zeros_like(f)
                for c, fk in zip(coef, self.focks):
                    f = f + c * fk
        return f

    def _coefficients(self):
        n = len(self.focks)
        b = np.zeros((n + 1, n + 1))
        rhs = np.zeros(n + 1)
        rhs[n] = 1.0
        for i in range(n):
            b[i, n] = b[n, i] = 1.0
            for j in range(n):
                b[i, j] = np.trace(self.errors[i] @ self.errors[j].T)
        try:
            if self.cond_limit is not None and np.linalg.cond(b) > self.cond_limit:
                raise np.linalg.LinAlgError('ill-conditioned DIIS matrix')
            return np.linalg.solve(b, rhs)[:n]
        except np.linalg.LinAlgError:
            self.focks.pop(0)
            self.errors.pop(0)
            return self._coefficients()


def _rotate(c, kappa, nocc):
    """C exp(K), K antisymmetric with K[vir, occ] = kappa."""
    from scipy.linalg import expm
    n = c.shape[0]
    k = np.zeros((n, n))
    k[nocc:, :nocc] = kappa
    k[:nocc, nocc:] = -kappa.T
    return c @ expm(k)


class SecondOrderSCF:
    """Opt-in SCF convergence aid; the reference has none (its loop is
    scf_procedure.py:24-42 with the DIIS of diis.py:30-112).  Minimises
    E(P) = 1/2 tr P (H + F(P)), P = 2 C_occ C_occ^T, in three stages:

    1. optimal damping (Cancès and Le Bris, Int. J. Quantum Chem. 79, 82
       (2000)) from the core-H guess until max |X^T (S P F - F P S) X| <
       ``switch`` or lam = 0;
    2. L-BFGS over rotations C exp(K) (gradient 4 C_vir^T F C_occ,
       preconditioner max(4 (e_a - e_i), floor), max-element step cap,
       backtracking line search, canonical orbitals per block after each
       step) until max |g| < ``newton_switch``;
    3. trust-region Newton with Steihaug CG and the exact orbital Hessian
       4 (e_a - e_i) kappa + 4 C_vir^T G(D) C_occ, G(D) = F(D) - H,
       D = 2 (C_vir kappa C_occ^T + transpose).
    Converged when |dE| < threshold and max |g| < ``gradient_threshold``.
    """

    def __init__(self, switch=0.1, newton_switch=1e-4, gradient_threshold=1e-7, history=20, max_rotation=0.3,
                 precondition_floor=0.2, trust_radius=0.5, max_cg=40):
        self.switch = switch
        self.newton_switch = newton_switch
        self.gradient_threshold = gradient_threshold
        self.history = history
        self.max_rotation = max_rotation
        self.floor = precondition_floor
        self.trust_radius = trust_radius
        self.max_cg = max_cg
        self.evaluations = 0

    def run(self, h, s, la, electrons, fock_builder, c, threshold, max_iter):
        nocc = electrons // 2

        def evaluate(c):
            p = density_matrix_restricted(c, electrons)
            f = fock_builder(p)
            self.evaluations += 1
            return p, f, total_energy_restricted(h, p, f)

        def canonical(c, f):
            fmo = c.T @ f @ c
            eo, uo = np.linalg.eigh(fmo[:nocc, :nocc])
            ev, uv = np.linalg.eigh(fmo[nocc:, nocc:])
            return np.hstack([c[:, :nocc] @ uo, c[:, nocc:] @ uv]), eo, ev, uo, uv

        def grad(c, f):
            return 4.0 * (c[:, nocc:].T @ f @ c[:, :nocc])

        def budget():
            return max_iter is not None and self.evaluations >= max_iter

        relaxed = None
        while True:
            p, f, e = evaluate(c)
            lam = 1.0
            if relaxed is None:
                relaxed = (p, f)
            else:
                pt, ft = relaxed
                d = p - pt
                slope = float(np.sum(ft * d))
                curv = float(np.sum((f - ft) * d))
                lam = 1.0 if curv <= -slope else -slope / curv
                lam = min(max(lam, 0.0), 1.0)
                relaxed = (pt + lam * d, ft + lam * (f - ft))
            err = np.abs(la.orthonormalize(s @ p @ f - f @ p @ s)).max()
            if err < self.switch or lam == 0.0 or budget():
                break
            _, c = la.diagonalize(relaxed[1])

        c, eo, ev, _, _ = canonical(c, f)
        g = grad(c, f)
        pairs = []
        stage = 'lbfgs'
        delta = self.trust_radius
        converged = False
        while not budget():
            if stage == 'lbfgs' and np.abs(g).max() < self.newton_switch:
                stage = 'newton'
            diag = 4.0 * (ev[:, None] - eo[None, :])
            prec = np.maximum(diag, self.floor)
            if stage == 'lbfgs':
                q = g.copy()
                alphas = []
                for sv, yv, rho in reversed(pairs):
                    a = rho * np.sum(sv * q)
                    q = q - a * yv
                    alphas.append(a)
                r = q / prec
                for (sv, yv, rho), a in zip(pairs, reversed(alphas)):
                    r = r + (a - rho * np.sum(yv * r)) * sv
                d = -r
                gd = np.sum(g * d)
                if gd >= 0.0:
                    pairs = []
                    d = -g / prec
                    gd = np.sum(g * d)
                m = np.abs(d).max()
                if m > self.max_rotation:
                    d = d * (self.max_rotation / m)
                    gd *= self.max_rotation / m
                t = 1.0
                for _ in range(12):
                    cn = _rotate(c, t * d, nocc)
                    pn, fn, en = evaluate(cn)
                    if en <= e + 1e-4 * t * gd + 1e-11:
                        break
                    tq = -gd * t * t / (2.0 * (en - e - gd * t))
                    t = min(0.5 * t, max(0.1 * t, tq))
                step = t * d
            else:
                co, cv = c[:, :nocc], c[:, nocc:]

                def hess(k):
                    dd = 2.0 * (cv @ k @ co.T)
                    return diag * k + 4.0 * (cv.T @ (fock_builder(dd + dd.T) - h) @ co)

                def to_boundary(x, d):
                    xx, xd, dd_ = np.sum(x * x), np.sum(x * d), np.sum(d * d)
                    return x + ((-xd + np.sqrt(xd * xd + dd_ * (delta * delta - xx))) / dd_) * d

                gnorm = np.linalg.norm(g)
                eta = min(0.5, np.sqrt(gnorm))
                x = np.zeros_like(g)
                r = g.copy()
                z = r / prec
                d = -z
                rz = np.sum(r * z)
                hit = False
                for _ in range(self.max_cg):
                    hd = hess(d)
                    dhd = np.sum(d * hd)
                    if dhd <= 0.0:
                        x, hit = to_boundary(x, d), True
                        break
                    alpha = rz / dhd
                    xn = x + alpha * d
                    if np.linalg.norm(xn) >= delta:
                        x, hit = to_boundary(x, d), True
                        break
                    x = xn
                    r = r + alpha * hd
                    if np.linalg.norm(r) < eta * gnorm:
                        break
                    z = r / prec
                    rzn = np.sum(r * z)
                    d = -z + (rzn / rz) * d
                    rz = rzn
                predicted = np.sum(g * x) + 0.5 * np.sum(x * hess(x))
                cn = _rotate(c, x, nocc)
                pn, fn, en = evaluate(cn)
                rho = (en - e) / predicted if predicted < -1e-11 else 1.0
                if rho < 0.25:
                    delta *= 0.25
                elif rho > 0.75 and hit:
                    delta = min(2.0 * delta, 2.0)
                if not en - e < 1e-11:
                    continue
            gn = grad(cn, fn)
            if stage == 'lbfgs':
                y = gn - g
                sy = np.sum(step * y)
                if sy > 1e-14:
                    pairs.append((step, y, 1.0 / sy))
                    if len(pairs) > self.history:
                        pairs.pop(0)
            de = e - en
            c, eo, ev, uo, uv = canonical(cn, fn)
            g = uv.T @ gn @ uo
            pairs = [(uv.T @ a @ uo, uv.T @ b @ uo, rr) for a, b, rr in pairs]
            p, f, e = pn, fn, en
            if abs(de) < threshold and np.abs(g).max() < self.gradient_threshold:
                converged = True
                break
        return converged, e, np.concatenate([eo, ev]), c


def rhf(h, s, electrons, fock_builder, threshold=1e-12, max_iter=None, diis_cond_limit=None, stabilizer=None,
        initial_density=None):
    """Restricted SCF (scf_procedure.py:24-42, hartree_fock.py:45-73).

    ``fock_builder(P) -> F``.  Returns (E_elec, orbital energies, C, iterations).
    ``max_iter`` is None for the reference behaviour (no cap).
    ``stabilizer='second_order'`` runs ``SecondOrderSCF`` instead of the
    reference's DIIS loop (an option the reference does not have; the
    iteration count is then its number of energy evaluations);
    ``initial_density`` starts from the orbitals of F(initial_density)
    instead of the core-H guess.
    """
    la = LinearAlgebra(s)
    eps, c = la.diagonalize(h if initial_density is None else fock_builder(initial_density))
    if stabilizer == 'second_order':
        so = SecondOrderSCF()
        converged, e, eps, c = so.run(h, s, la, electrons, fock_builder, c, threshold, max_iter)
        if not converged:
            raise RuntimeError('second-order SCF did not converge in %d evaluations' % so.evaluations)
        return e, eps, c, so.evaluations
    if stabilizer is not None:
        raise ValueError(stabilizer)
    diis = DIIS(s, la, cond_limit=diis_cond_limit)
    prev = 0.0
    it = 0
    while True:
        p = density_matrix_restricted(c, electrons)
        f = fock_builder(p)
        e = total_energy_restricted(h, p, f)
        de = prev - e
        prev = e
        it += 1
        if abs(de) < threshold or (max_iter is not None and it >= max_iter):
            break
        f = diis.fock_matrix(f, p)
        eps, c = la.diagonalize(f)
    return e, eps, c, it


def density_matrix_unrestricted(c, electrons):
    """density_matrix.py:12-17: P = sum_{i < n} c_i c_i^T (one spin)."""
    occ = c[:, :electrons]
    return occ @ occ.T


def total_energy_unrestricted(h, pa, pb, fa, fb):
    """total_energy.py:37-63: 1/2 sum (P_a + P_b) H + P_a F_a + P_b F_b."""
    return 0.5 * float(np.sum((pa + pb) * h + pa * fa + pb * fb))


def uhf(h, s, electrons, multiplicity, fock_builder, threshold=1e-12, max_iter=None):
    """Pople-Nesbet-Berthier SCF (scf_procedure.py:45-89, hartree_fock.py:87-122).

    ``fock_builder(P_a, P_b) -> (F_a, F_b)`` with F_a = H + J(P_a + P_b) - K(P_a)
    (fock_matrix.py:25-55).  The first iteration takes the alpha density from
    every orbital of the core-H guess and an empty beta density, as the
    reference does; no DIIS.  Returns (E_elec, iterations)."""
    la = LinearAlgebra(s)
    _, c = la.diagonalize(h)
    ca = cb = c
    na, nb = (electrons + multiplicity - 1) // 2, (electrons - multiplicity + 1) // 2
    prev = 0.0
    it = 0
    while True:
        if it == 0:
            pa, pb = density_matrix_unrestricted(ca, ca.shape[0]), density_matrix_unrestricted(cb, 0)
        else:
            pa, pb = density_matrix_unrestricted(ca, na), density_matrix_unrestricted(cb, nb)
        fa, fb = fock_builder(pa, pb)
        e = total_energy_unrestricted(h, pa, pb, fa, fb)
        de = prev - e
        prev = e
        it += 1
        if abs(de) < threshold or (max_iter is not None and it >= max_iter):
            break
        _, ca = la.diagonalize(fa)
        _, cb = la.diagonalize(fb)
    return e, it


def blocked_spin_basis_set(eri):
    """transformations.py:29-71: spin-orbital (rs|tu) in the blocked order
    (N alpha, then N beta functions): (rs|tu) = (r's'|t'u') of the spatial
    functions when r, s share a spin and t, u share a spin, else 0."""
    n = eri.shape[0]
    out = np.zeros((2 * n,) * 4)
    for sr in (0, 1):
        for st in (0, 1):
            out[sr * n:(sr + 1) * n, sr * n:(sr + 1) * n, st * n:(st + 1) * n, st * n:(st + 1) * n] = eri
    return out


def fock_blocked(h2, eri2, p2):
    """fock_matrix.py:56-81: F = H + J - K on the spin-blocked basis with
    J_ij = sum_ab P_ab (ij|ab) and K_ij = sum_ab P_ab (ib|aj)."""
    j = np.einsum('ijab,ab->ij', eri2, p2, optimize=True)
    k = np.einsum('ibaj,ab->ij', eri2, p2, optimize=True)
    return h2 + j - k


class BlockedLinearAlgebra(LinearAlgebra):
    """linear_algebra.py:36-98: X from each spin block of the blocked overlap
    on its own; diagonalise the two spin blocks of X^T F X separately, each
    in ascending order; C = X blockdiag(C_a, C_b), energies alpha then beta."""

    def __init__(self, s2):
        n = s2.shape[0] // 2
        self.n = n
        self.x = np.zeros_like(s2)
        for b in (0, 1):
            w, u = np.linalg.eigh(s2[b * n:(b + 1) * n, b * n:(b + 1) * n])
            order = np.argsort(w)
            w, u = w[order], u[:, order]
            self.x[b * n:(b + 1) * n, b * n:(b + 1) * n] = u @ np.diag(w ** -0.5)

    def diagonalize(self, f):
        n = self.n
        fp = self.x.T @ f @ self.x
        c = np.zeros_like(f)
        es = []
        for b in (0, 1):
            e, v = np.linalg.eigh(fp[b * n:(b + 1) * n, b * n:(b + 1) * n])
            order = np.argsort(e)
            es.append(e[order])
            c[b * n:(b + 1) * n, b * n:(b + 1) * n] = v[:, order]
        return np.concatenate(es), self.x @ c


def blocked_density_matrix(c, electrons_alph, electrons_beta):
    """density_matrix.py:20-36: sum of c_i c_i^T over the alpha columns
    [0, n_a) and the beta columns [N, N + n_b) of the 2N x 2N coefficients."""
    half = c.shape[0] // 2
    a = c[:, :electrons_alph]
    b = c[:, half:half + electrons_beta]
    return a @ a.T + b @ b.T


def guhf(h, s, electrons, multiplicity, fock_builder, threshold=1e-12, max_iter=None, diis_cond_limit=None):
    """Blocked unrestricted SCF (scf_procedure.py:92-120, hartree_fock.py:119-139).

    ``h``, ``s``: spatial N x N; the SCF runs on blockdiag(h, h) and
    blockdiag(s, s).  ``fock_builder(P2) -> F2`` on the 2N basis.  The first
    density takes every alpha column of the core-H guess and no beta one, as
    the reference does; DIIS on the 2N matrices; E = 1/2 sum P (H + F).
    Returns (E_elec, iterations)."""
    n = h.shape[0]
    z = np.zeros((n, n))
    h2 = np.block([[h, z], [z, h]])
    s2 = np.block([[s, z], [z, s]])
    la = BlockedLinearAlgebra(s2)
    diis = DIIS(s2, la, cond_limit=diis_cond_limit)
    _, c = la.diagonalize(h2)
    na, nb = (electrons + multiplicity - 1) // 2, (electrons - multiplicity + 1) // 2
    e = prev = 0.0
    it = 0
    while True:
        if e == 0:
            p = blocked_density_matrix(c, n, 0)
        else:
            p = blocked_density_matrix(c, na, nb)
        f = fock_builder(p)
        e = total_energy_restricted(h2, p, f)
        de = prev - e
        prev = e
        it += 1
        if abs(de) < threshold or (max_iter is not None and it >= max_iter):
            break
        f = diis.fock_matrix(f, p)
        _, c = la.diagonalize(f)
    return e, it


def molecular_orbitals(eri, c):
    """transformations.py:5-26: (pq|rs) = sum C_mp C_nq C_lr C_ss (mn|ls), one
    N x N slice at a time as the reference does (on a copy)."""
    out = np.array(eri, copy=True)
    n = c.shape[0]
    for r in range(n):
        for s_ in range(n):
            out[r, s_, :, :] = c.T @ out[r, s_, :, :] @ c
    for t in range(n):
        for u in range(n):
            out[:, :, t, u] = c.T @ out[:, :, t, u] @ c
    return out


def mp2_correlation(eri_mo, eps, electrons):
    """moeller_plesset.py:15-23: sum over occupied i, j and virtual a, b of
    [2 (ia|jb)^2 - (ia|jb)(ib|ja)] / (e_i + e_j - e_a - e_b), as explicit loops."""
    nocc = electrons // 2
    n = len(eps)
    corr = 0.0
    for i in range(nocc):
        for j in range(nocc):
            for a in range(nocc, n):
                for b in range(nocc, n):
                    d = eps[i] + eps[j] - eps[a] - eps[b]
                    corr += 2.0 * eri_mo[i, a, j, b] ** 2 / d
                    corr -= eri_mo[i, a, j, b] * eri_mo[i, b, j, a] / d
    return corr
