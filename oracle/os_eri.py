"""Obara-Saika two-electron repulsion integrals, restated from the reference.

Follows ``integrals/twoelectronrepulsion/obara_saika_scheme.py``:

* ``eri_contracted``  -- ``ObaraSaika.integrate`` (:17-82): loop over primitive
  quartets, pair quantities p, q, rho, P, Q, W (:52-66), T = rho |PQ|^2
  (:68), prefactor 2 pi^(5/2) / (p q sqrt(p+q)) exp(-ab/p R_AB^2 - cd/q R_CD^2)
  (:69-70), F_L by the reference series then downward recursion (:71-78),
  contraction weight c1 c2 c3 c4 N1 N2 N3 N4 n_i n_j n_k n_l (:34-50).
* ``_os``            -- ``os_begin``/``os_recursive`` (:84-146): lower the
  first non-zero Cartesian index in the order a_x,a_y,a_z,b_x..,d_z; the
  eight-term recursion with the same zero-skips (:131-144).

The recursion keeps the reference's call structure (no shared memo across
primitive quartets); ``memo=True`` caches sub-results inside one primitive
quartet, which changes only the run time, not the arithmetic.

Test infrastructure only; see ``oracle/__init__.py``.
"""
from math import sqrt, pi, exp
import itertools

from .boys import boys_function, boys_function_recursion


def _gpc(a, ra, b, rb):
    s = a + b
    return ((a * ra[0] + b * rb[0]) / s, (a * ra[1] + b * rb[1]) / s, (a * ra[2] + b * rb[2]) / s)


def _dist(r1, r2):
    return sqrt((r1[0] - r2[0]) ** 2 + (r1[1] - r2[1]) ** 2 + (r1[2] - r2[2]) ** 2)


class _Prim:
    __slots__ = ('a', 'r', 'l')

    def __init__(self, a, r, l):
        self.a = a
        self.r = r
        self.l = l


def _lower(l, i, k):
    t = list(l)
    t[i] -= k
    return tuple(t)


class _OSQuartet:
    """One primitive quartet: base values end[m] and the OS recursion."""

    def __init__(self, g1, g2, g3, g4, end, a7, r7, memo):
        self.end = end
        self.a7 = a7
        self.r7 = r7
        self.memo = {} if memo else None
        self.g = (g1, g2, g3, g4)

    def begin(self, m, l1, l2, l3, l4):
        if self.memo is not None:
            key = (m, l1, l2, l3, l4)
            v = self.memo.get(key)
            if v is not None:
                return v
            v = self._begin(m, l1, l2, l3, l4)
            self.memo[key] = v
            return v
        return self._begin(m, l1, l2, l3, l4)

    def _begin(self, m, l1, l2, l3, l4):
        # obara_saika_scheme.py:90-115: pick the first non-zero index.
        for slot, (ls, perm) in enumerate(((l1, (0, 1, 2, 3)), (l2, (1, 0, 3, 2)),
                                           (l3, (2, 3, 0, 1)), (l4, (3, 2, 1, 0)))):
            for r in range(3):
                if ls[r] > 0:
                    return self._rec(r, m, perm, (l1, l2, l3, l4))
        return self.end[m]

    def _rec(self, r, m, perm, ls):
        # obara_saika_scheme.py:117-146 with g1..g4 = the permuted centres.
        i1, i2, i3, i4 = perm
        G = self.g
        a1, a2, a3, a4 = G[i1].a, G[i2].a, G[i3].a, G[i4].a
        L1, L2, L3, L4 = ls[i1], ls[i2], ls[i3], ls[i4]
        a5 = a1 + a2
        a6 = a3 + a4
        r1 = G[i1].r
        r5 = _gpc(a1, r1, a2, G[i2].r)
        g1x1 = _lower(L1, r, 1)
        g1x2 = _lower(L1, r, 2)
        g2x1 = _lower(L2, r, 1)
        g3x1 = _lower(L3, r, 1)
        g4x1 = _lower(L4, r, 1)

        def call(mm, n1, n2, n3, n4):
            # Put the permuted exponent tuples back into slot order.
            slots = [None] * 4
            slots[i1], slots[i2], slots[i3], slots[i4] = n1, n2, n3, n4
            return self.begin(mm, *slots)

        out = 0.0
        if r5[r] != r1[r]:
            out += (r5[r] - r1[r]) * call(m, g1x1, L2, L3, L4)
        if self.r7[r] != r5[r]:
            out += (self.r7[r] - r5[r]) * call(m + 1, g1x1, L2, L3, L4)
        if g1x2[r] >= 0:
            n = g1x1[r] if g1x1[r] != 0 else 1
            out += n * (1 / (2 * a5)) * call(m, g1x2, L2, L3, L4)
            out -= n * (self.a7 / (2 * a5 ** 2)) * call(m + 1, g1x2, L2, L3, L4)
        if g2x1[r] >= 0:
            n = L2[r] if L2[r] != 0 else 1
            out += n * (1 / (2 * a5)) * call(m, g1x1, g2x1, L3, L4)
            out -= n * (self.a7 / (2 * a5 ** 2)) * call(m + 1, g1x1, g2x1, L3, L4)
        if g3x1[r] >= 0:
            n = L3[r] if L3[r] != 0 else 1
            out += n * (1 / (2 * (a5 + a6))) * call(m + 1, g1x1, L2, g3x1, L4)
        if g4x1[r] >= 0:
            n = L4[r] if L4[r] != 0 else 1
            out += n * (1 / (2 * (a5 + a6))) * call(m + 1, g1x1, L2, L3, g4x1)
        return out


def eri_primitive(a1, r1, l1, a2, r2, l2, a3, r3, l3, a4, r4, l4, memo=True):
    """Unnormalised primitive [l1 l2|l3 l4] (obara_saika_scheme.py:52-80)."""
    l_total = sum(l1) + sum(l2) + sum(l3) + sum(l4)
    a5 = a1 + a2
    a6 = a3 + a4
    a7 = (a5 * a6) / (a5 + a6)
    r5 = _gpc(a1, r1, a2, r2)
    r6 = _gpc(a3, r3, a4, r4)
    r7 = _gpc(a5, r5, a6, r6)
    r12 = _dist(r1, r2)
    r34 = _dist(r3, r4)
    r56 = _dist(r5, r6)
    x = (a5 * a6 * r56 ** 2) / (a5 + a6)
    out1 = (2 * pi ** 2.5) / (a5 * a6 * sqrt(a5 + a6))
    out2 = exp(((-a1 * a2 * r12 ** 2) / a5) - ((a3 * a4 * r34 ** 2) / a6))
    f = boys_function(l_total, x)
    end = {l_total: out1 * out2 * f}
    m = l_total
    while m >= 1:
        f = boys_function_recursion(m, x, f)
        m -= 1
        end[m] = out1 * out2 * f
    q = _OSQuartet(_Prim(a1, r1, l1), _Prim(a2, r2, l2), _Prim(a3, r3, l3), _Prim(a4, r4, l4),
                   end, a7, r7, memo)
    return q.begin(0, tuple(l1), tuple(l2), tuple(l3), tuple(l4))


def eri_contracted(bi, bj, bk, bl, memo=True):
    """(ij|kl) of four contracted functions (obara_saika_scheme.py:17-82).

    ``b*`` are duck-typed ``Basis`` objects: ``primitive_gaussian_array``,
    ``coordinates``, ``integral_exponents``, ``normalisation``; primitives
    carry ``contraction``, ``exponent``, ``normalisation``.
    """
    n = bi.normalisation * bj.normalisation * bk.normalisation * bl.normalisation
    li, lj, lk, ll = (tuple(b.integral_exponents) for b in (bi, bj, bk, bl))
    ri, rj, rk, rl = (tuple(b.coordinates) for b in (bi, bj, bk, bl))
    ans = 0.0
    for g1, g2, g3, g4 in itertools.product(bi.primitive_gaussian_array, bj.primitive_gaussian_array,
                                            bk.primitive_gaussian_array, bl.primitive_gaussian_array):
        w = (g1.contraction * g2.contraction * g3.contraction * g4.contraction
             * g1.normalisation * g2.normalisation * g3.normalisation * g4.normalisation * n)
        ans += w * eri_primitive(g1.exponent, ri, li, g2.exponent, rj, lj,
                                 g3.exponent, rk, lk, g4.exponent, rl, ll, memo=memo)
    return ans


def eri_unique(basis, keys=None, memo=True):
    """{(a,b,c,d): value} over the canonical keys (electron_repulsion_matrix.py:28-41)."""
    from .basis import canonical_keys
    if keys is None:
        keys = canonical_keys(len(basis))
    return {k: eri_contracted(basis[k[0]], basis[k[1]], basis[k[2]], basis[k[3]], memo=memo) for k in keys}


def eri_dense_from_unique(values, n):
    """Dense N^4 tensor filled by 8-fold symmetry (electron_repulsion_matrix.py:43-45)."""
    import numpy as np
    out = np.zeros((n, n, n, n))
    for (a, b, c, d), v in values.items():
        for (i, j, k, l) in ((a, b, c, d), (b, a, c, d), (a, b, d, c), (b, a, d, c),
                             (c, d, a, b), (d, c, a, b), (c, d, b, a), (d, c, b, a)):
            out[i, j, k, l] = v
    return out
