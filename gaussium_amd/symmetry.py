"""Point-group symmetry of a nuclear framework and the electron-repulsion
integrals it forces to zero.

Replaces factory/molecule_factory.py:20-73, factory/symmetry_factory.py and
common/symmetry.py:18-55 of the reference (``MoleculeFactory(symmetry=True)
.create`` and ``Symmetry(point_group, basis).none_zero_integral``), behind
the same names: ``MoleculeFactory`` returns the centred and oriented nuclei
and a ``PointGroup`` with the reference's labels; ``Symmetry`` answers
``none_zero_integral((i, j, k, l))`` and gives the whole zero pattern of a
dense ERI tensor at once (``zero_mask``), which ``TwoElectronRepulsionMatrixOS``
applies to the GPU tensor.

Elements are found by trying candidate axes and plane normals (the atoms,
the midpoints of atom pairs and the cross products of both, as the
reference does) with numpy: an operation is kept when it maps every nucleus
onto a nucleus of the same charge within ``error``.  The integral test is
exact: an operation g of the group whose matrix, in the oriented frame,
permutes and signs the axes maps a Cartesian Gaussian onto +/- one basis
function (``image``); if g maps the pair {i, j} onto itself and {k, l} onto
itself, (ij|kl) = s (ij|kl) with s the product of the four signs, so the
integral vanishes when s = -1.  Integrals with all four functions on one
centre vanish unless the total exponent is even along every axis (reflection
of both electrons through that centre).  A C_1 molecule has no zero by
symmetry (as in the reference, which then returns True for every index).
"""
import itertools

import numpy as np

__all__ = ['PointGroup', 'MoleculeFactory', 'Symmetry', 'find_elements']


def _unit(v):
    n = np.linalg.norm(v)
    return v / n if n > 1e-8 else None


def _rotation(axis, angle):
    a = np.asarray(axis, float)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def _reflection(normal):
    n = np.asarray(normal, float)
    return np.eye(3) - 2.0 * np.outer(n, n)


def _unique_directions(vectors, tol=1e-3):
    out = []
    for v in vectors:
        u = _unit(np.asarray(v, float))
        if u is None:
            continue
        if all(abs(abs(float(u @ w)) - 1.0) > tol for w in out):
            out.append(u)
    return out


class _Frame:
    def __init__(self, Z, X, error):
        self.Z = np.asarray(Z)
        self.X = np.asarray(X, float)
        self.error = error

    def maps(self, R):
        """R maps every nucleus onto a nucleus of the same charge."""
        Y = self.X @ np.asarray(R).T
        for i in range(len(self.Z)):
            d = np.linalg.norm(self.X - Y[i], axis=1)
            if not np.any((d < self.error) & (self.Z == self.Z[i])):
                return False
        return True


class Elements:
    """Symmetry elements: rotation axes with their highest fold, mirror-plane
    normals, improper axes with their fold, and whether there is inversion."""

    def __init__(self, rotations, reflections, improper, inversion):
        self.rotations = rotations      # [(unit axis, fold)]
        self.reflections = reflections  # [unit normal]
        self.improper = improper        # [(unit axis, fold)]
        self.inversion = inversion


def find_elements(charges, coordinates, error=1e-3, max_fold=8):
    """Symmetry elements of nuclei already centred on their geometric centre."""
    Z = np.asarray(charges)
    X = np.asarray(coordinates, float)
    fr = _Frame(Z, X, error)
    off = [x for x in X if np.linalg.norm(x) > error]
    mids = [0.5 * (a + b) for a, b in itertools.combinations(off, 2)]
    # a mirror plane or two-fold axis that swaps two equal nuclei is normal
    # to, or contains, their difference
    diffs = [X[a] - X[b] for a, b in itertools.combinations(range(len(X)), 2) if Z[a] == Z[b]]
    cands = off + mids
    cross = [np.cross(a, b) for a, b in itertools.combinations(cands, 2)] if len(cands) < 60 else \
        [np.cross(a, b) for a, b in itertools.combinations(off, 2)]
    directions = _unique_directions(cands + diffs + cross + [np.eye(3)[k] for k in range(3)])
    rotations = []
    for u in directions:
        best = 1
        for n in range(max_fold, 1, -1):
            if fr.maps(_rotation(u, 2 * np.pi / n)):
                best = n
                break
        if best > 1:
            rotations.append((u, best))
    reflections = [u for u in directions if fr.maps(_reflection(u))]
    improper = []
    for u, n in rotations:
        for m in (2 * n, n):
            if m > 2 and fr.maps(_reflection(u) @ _rotation(u, 2 * np.pi / m)):
                improper.append((u, m))
                break
    inversion = fr.maps(-np.eye(3))
    return Elements(rotations, reflections, improper, inversion)


class PointGroup:
    """Point group: its label in the reference's notation ('C_{2v}', 'D_{3d}',
    ...; linear molecules 'D_{4h}' / 'C_{4v}' and single atoms 'O_{h}' as the
    reference reports them) and its operations as 3x3 matrices."""

    def __init__(self, label, operations, elements=None):
        self.label = label
        self.operations = operations
        self.elements = elements


def _classify(el, X, error):
    if len(X) == 1:
        return 'O_{h}'
    off = [x for x in X if np.linalg.norm(x) > error]
    if off:
        d = _unit(off[0])
        if all(np.linalg.norm(np.cross(d, x)) < error for x in off):
            return 'D_{4h}' if el.inversion else 'C_{4v}'
    rot = el.rotations
    if sum(1 for _, n in rot if n >= 3) >= 2:
        if not el.inversion:
            return 'T_{d}'
        return 'I_{h}' if any(n == 5 for _, n in rot) else 'O_{h}'
    if not rot:
        if el.reflections:
            return 'C_{s}'
        return 'C_{i}' if el.inversion else 'C_{1}'
    axis, n = max(rot, key=lambda r: r[1])
    perp_c2 = [u for u, m in rot if m % 2 == 0 and abs(float(u @ axis)) < 1e-3]
    sigma_h = any(abs(abs(float(u @ axis)) - 1.0) < 1e-3 for u in el.reflections)
    sigma_v = sum(1 for u in el.reflections if abs(float(u @ axis)) < 1e-3)
    if len(perp_c2) >= n:
        if sigma_h:
            return 'D_{%dh}' % n
        return 'D_{%dd}' % n if sigma_v >= n else 'D_{%d}' % n
    if sigma_h:
        return 'C_{%dh}' % n
    if sigma_v >= n:
        return 'C_{%dv}' % n
    if any(abs(abs(float(u @ axis)) - 1.0) < 1e-3 and m == 2 * n for u, m in el.improper):
        return 'S_{%d}' % (2 * n)
    return 'C_{%d}' % n


def _frame(el, ops):
    """Orthonormal frame (rows x, y, z) in which as many of the operations
    as possible permute and sign the axes (so that they map Cartesian
    functions onto single functions): z and x chosen among the rotation axes
    and mirror normals, perpendicular to each other."""
    dirs = [u for u, _ in el.rotations] + list(el.reflections) + [u for u, _ in el.improper]
    best, best_n = np.eye(3), sum(_signed_permutation(R) is not None for R in ops)
    for z in dirs:
        for x in dirs:
            if abs(float(x @ z)) > 1e-3:
                continue
            F = np.vstack([x, np.cross(z, x), z])
            n = sum(_signed_permutation(F @ R @ F.T) is not None for R in ops)
            if n > best_n:
                best, best_n = F, n
    if best_n == 1 and dirs:  # no pair of perpendicular elements: put one on z
        z = dirs[0]
        t = np.eye(3)[int(np.argmin(np.abs(z)))]
        x = _unit(t - (t @ z) * z)
        best = np.vstack([x, np.cross(z, x), z])
    return best


def _group_operations(el):
    ops = [np.eye(3)]
    for u, n in el.rotations:
        ops += [_rotation(u, 2 * np.pi * k / n) for k in range(1, n)]
    ops += [_reflection(u) for u in el.reflections]
    for u, n in el.improper:
        ops += [_reflection(u) @ _rotation(u, 2 * np.pi * k / n) for k in range(1, n, 2)]
    if el.inversion:
        ops.append(-np.eye(3))
    out = []
    for R in ops:
        if all(np.abs(R - S).max() > 1e-6 for S in out):
            out.append(R)
    return out


class MoleculeFactory:
    """``create(nuclei)`` -> (nuclei centred and oriented, PointGroup), as
    molecule_factory.py:20-73 (``symmetry=False``: the nuclei unchanged and
    C_1)."""

    def __init__(self, symmetry=False, error=1e-3):
        self.symmetry = symmetry
        self.error = error

    def create(self, nuclei_array):
        if not self.symmetry:
            return nuclei_array, PointGroup('C_{1}', [np.eye(3)])
        X = np.array([n.coordinates for n in nuclei_array], float)
        X = X - X.mean(axis=0)
        Z = [n.charge for n in nuclei_array]
        el = find_elements(Z, X, self.error)
        label = _classify(el, X, self.error)
        world = _group_operations(el)
        F = _frame(el, world)
        Xo = X @ F.T
        ops = [F @ R @ F.T for R in world]
        for n, x in zip(nuclei_array, Xo):
            n.coordinates = tuple(float(v) for v in x)
        return nuclei_array, PointGroup(label, ops, el)


def _signed_permutation(R, tol=1e-6):
    """(p, s) with (R r)_i = s_i r_{p_i}, or None if R is not one."""
    p, s = [], []
    for i in range(3):
        k = int(np.argmax(np.abs(R[i])))
        if abs(abs(R[i, k]) - 1.0) > tol or np.abs(np.delete(R[i], k)).max() > tol:
            return None
        p.append(k)
        s.append(1 if R[i, k] > 0 else -1)
    return p, s


class Symmetry:
    """``Symmetry(point_group, basis_set_array)``: which integrals vanish by
    symmetry (common/symmetry.py:9-55).  ``image[g][i] = (index, sign)`` of
    basis function i under operation g (only operations that permute and sign
    the axes of the oriented frame; -1 where the image is not a single basis
    function)."""

    def __init__(self, point_group, basis_set_array, error=1e-3):
        self.point_group = point_group
        self.basis_set = basis_set_array
        self.error = error
        n = len(basis_set_array)
        self.trivial = point_group.label == 'C_{1}' or n == 0
        self._centres = []
        self.centre = np.array([self._centre_id(b.coordinates) for b in basis_set_array], dtype=np.int64)
        key = {}
        for i, b in enumerate(basis_set_array):
            key[(int(self.centre[i]),) + self._key(b.integral_exponents, b)] = i
        self.index, self.sign = [], []
        for R in point_group.operations:
            ps = _signed_permutation(np.asarray(R))
            if ps is None or np.abs(np.asarray(R) - np.eye(3)).max() < 1e-9:
                continue
            p, s = ps
            idx = np.full(n, -1, dtype=np.int64)
            sg = np.zeros(n, dtype=np.int64)
            for i, b in enumerate(basis_set_array):
                c = np.asarray(b.coordinates, float)
                gc = np.array([s[k] * c[p[k]] for k in range(3)])
                e = b.integral_exponents
                ge = tuple(e[p[k]] for k in range(3))
                cid = self._find_centre(gc)
                j = key.get((cid,) + self._key(ge, b)) if cid >= 0 else None
                if j is not None:
                    idx[i] = j
                    sg[i] = int(np.prod([s[k] ** e[p[k]] for k in range(3)]))
            self.index.append(idx)
            self.sign.append(sg)
        self.exps = np.array([b.integral_exponents for b in basis_set_array], dtype=np.int64).reshape(n, 3)

    def _find_centre(self, c):
        for k, d in enumerate(self._centres):
            if np.linalg.norm(np.asarray(c, float) - d) < self.error:
                return k
        return -1

    def _centre_id(self, c):
        k = self._find_centre(c)
        if k >= 0:
            return k
        self._centres.append(np.asarray(c, float))
        return len(self._centres) - 1

    @staticmethod
    def _key(exps, b):
        prims = tuple((round(q.exponent, 10), round(q.contraction, 10)) for q in b.primitive_gaussian_array)
        return (tuple(exps), prims)

    def none_zero_integral(self, index):
        """False when (ij|kl) vanishes by symmetry (symmetry.py:18-55)."""
        if self.trivial:
            return True
        return not bool(self.zero_mask(np.asarray([index], dtype=np.int64))[0])

    def zero_mask(self, keys=None):
        """Boolean mask of the integrals that vanish by symmetry: for index
        quartets ``keys`` [M, 4], or the whole [N, N, N, N] tensor."""
        n = len(self.basis_set)
        if keys is None:
            g = np.indices((n, n, n, n)).reshape(4, -1).T
        else:
            g = np.asarray(keys, dtype=np.int64).reshape(-1, 4)
        i, j, k, l = g.T
        zero = np.zeros(len(g), dtype=bool)
        if self.trivial:
            return zero if keys is not None else zero.reshape(n, n, n, n)
        one = (self.centre[i] == self.centre[j]) & (self.centre[j] == self.centre[k]) & \
              (self.centre[k] == self.centre[l])
        tot = self.exps[i] + self.exps[j] + self.exps[k] + self.exps[l]
        zero |= one & np.any(tot % 2 == 1, axis=1)
        for idx, sg in zip(self.index, self.sign):
            gi, gj, gk, gl = idx[i], idx[j], idx[k], idx[l]
            ok = (gi >= 0) & (gj >= 0) & (gk >= 0) & (gl >= 0)
            bra = ((gi == i) & (gj == j)) | ((gi == j) & (gj == i))
            ket = ((gk == k) & (gl == l)) | ((gk == l) & (gl == k))
            s = sg[i] * sg[j] * sg[k] * sg[l]
            zero |= ok & bra & ket & (s == -1)
        return zero if keys is not None else zero.reshape(n, n, n, n)
