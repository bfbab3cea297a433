"""One-electron integrals S, T, V, restated from the reference.

* overlap: ``integrals/orbital_overlap_integral.py:9-36`` (binomial series with
  (2j-1)!! / (2g)^j, (-1)!! = 1) and ``integrals/binomial_coefficients.py:4-16``.
* kinetic: ``matrixelements/kinetic_energy_matrix.py:16-33`` with the
  -1/2 nabla^2 expansion of ``factory/primitive_basis_factory.py:89-122``.
* nuclear attraction: ``integrals/nuclear_attraction_integral.py:10-57``
  (Taketa-Huzinaga-O-ohata explicit sums) and
  ``matrixelements/nuclear_attraction_matrix.py:16-29``.
* matrices are filled on the upper triangle and mirrored
  (``matrixelements/create_matrix.py:18-35``).
* like the reference, the primitive integrals take the primitives' own
  centres (``orbital_overlap_integral.py:22-23``); only the normalisation
  comes from the contracted function.

Test infrastructure only; see ``oracle/__init__.py``.
"""
from math import exp, pi, sqrt, factorial
import itertools
import numpy as np

from .basis import factorial2
from .boys import boys_function


def _comb(n, k):
    return factorial(n) / (factorial(k) * factorial(n - k)) if k <= n else 0


def _binom(j, l1, l2, a, b):
    ans = 0.0
    for k in range(max(0, j - l2), min(j, l1) + 1):
        ans += _comb(l1, k) * _comb(l2, j - k) * a ** (l1 - k) * b ** (l2 + k - j)
    return ans


def _gpc(a, ra, b, rb):
    s = a + b
    return tuple((a * ra[i] + b * rb[i]) / s for i in range(3))


def _s1d(l1, l2, a, b, g):
    s = 0.0
    for j in range((l1 + l2) // 2 + 1):
        s += _binom(2 * j, l1, l2, a, b) * (factorial2(2 * j - 1) / (2 * g) ** j)
    return s


def overlap_prim(a1, ra, l1, a2, rb, l2):
    """Unnormalised primitive overlap (orbital_overlap_integral.py:16-36)."""
    rab2 = sum((ra[i] - rb[i]) ** 2 for i in range(3))
    rp = _gpc(a1, ra, a2, rb)
    g = a1 + a2
    s = 1.0
    for i in range(3):
        s *= _s1d(l1[i], l2[i], rp[i] - ra[i], rp[i] - rb[i], g)
    return (pi / g) ** 1.5 * exp(-a1 * a2 * rab2 / g) * s


def kinetic_prim(a1, ra, l1, a2, rb, l2):
    """Primitive <a| -1/2 nabla^2 |b> via the del_operator expansion (primitive_basis_factory.py:89-122)."""
    l, m, n = l2
    terms = [(a2 * (2 * (l + m + n) + 3), (l, m, n)),
             (-2 * a2 ** 2, (l + 2, m, n)), (-2 * a2 ** 2, (l, m + 2, n)), (-2 * a2 ** 2, (l, m, n + 2))]
    if l >= 2:
        terms.append((-0.5 * l * (l - 1), (l - 2, m, n)))
    if m >= 2:
        terms.append((-0.5 * m * (m - 1), (l, m - 2, n)))
    if n >= 2:
        terms.append((-0.5 * n * (n - 1), (l, m, n - 2)))
    return sum(c * overlap_prim(a1, ra, l1, a2, rb, lc) for c, lc in terms)


def _a_func(l, r, i, l1, l2, pa, pb, pc, g):
    e = 1 / (4 * g)
    out1 = _binom(l, l1, l2, pa, pb)
    out2 = (-1) ** i * factorial(l) * pc ** (l - 2 * r - 2 * i) * e ** (r + i)
    out3 = factorial(r) * factorial(i) * factorial(l - 2 * r - 2 * i)
    return (-1) ** l * out1 * (out2 / out3)


def nuclear_prim(a1, ra, l1, a2, rb, l2, rc):
    """Primitive <a|1/r_C|b> (nuclear_attraction_integral.py:19-57)."""
    rp = _gpc(a1, ra, a2, rb)
    rab2 = sum((ra[i] - rb[i]) ** 2 for i in range(3))
    rpc2 = sum((rp[i] - rc[i]) ** 2 for i in range(3))
    pa = [rp[i] - ra[i] for i in range(3)]
    pb = [rp[i] - rb[i] for i in range(3)]
    pc = [rp[i] - rc[i] for i in range(3)]
    g = a1 + a2
    ans = 0.0
    for l in range(l1[0] + l2[0] + 1):
        for r in range(l // 2 + 1):
            for i in range((l - 2 * r) // 2 + 1):
                o1 = _a_func(l, r, i, l1[0], l2[0], pa[0], pb[0], pc[0], g)
                for m in range(l1[1] + l2[1] + 1):
                    for s in range(m // 2 + 1):
                        for j in range((m - 2 * s) // 2 + 1):
                            o2 = _a_func(m, s, j, l1[1], l2[1], pa[1], pb[1], pc[1], g)
                            for n in range(l1[2] + l2[2] + 1):
                                for t in range(n // 2 + 1):
                                    for k in range((n - 2 * t) // 2 + 1):
                                        o3 = _a_func(n, t, k, l1[2], l2[2], pa[2], pb[2], pc[2], g)
                                        v = (l + m + n) - 2 * (r + s + t) - (i + j + k)
                                        ans += o1 * o2 * o3 * boys_function(v, g * rpc2)
    return ans * (2 * pi / g) * exp(-a1 * a2 * rab2 / g)


def _contract(bi, bj, prim_fn):
    acc = 0.0
    for pa, pb in itertools.product(bi.primitive_gaussian_array, bj.primitive_gaussian_array):
        acc += (pa.normalisation * pb.normalisation * pa.contraction * pb.contraction
                * prim_fn(pa.exponent, tuple(pa.coordinates), tuple(bi.integral_exponents),
                          pb.exponent, tuple(pb.coordinates), tuple(bj.integral_exponents)))
    return bi.normalisation * bj.normalisation * acc


def _matrix(basis, fn):
    n = len(basis)
    out = np.zeros((n, n))
    for i in range(n):
        for j in range(i, n):
            out[i, j] = out[j, i] = fn(basis[i], basis[j])
    return out


def overlap_matrix(basis):
    return _matrix(basis, lambda a, b: _contract(a, b, overlap_prim))


def kinetic_matrix(basis):
    return _matrix(basis, lambda a, b: _contract(a, b, kinetic_prim))


def nuclear_matrix(basis, nuclei):
    def fn(a, b):
        v = 0.0
        for nuc in nuclei:
            rc = tuple(nuc.coordinates)
            v += -nuc.charge * _contract(a, b, lambda *args: nuclear_prim(*args, rc))
        return v
    return _matrix(basis, fn)
