#!/usr/bin/env python3
"""Generate the Rys-quadrature tables and per-class 2D-integral code.

Rys quadrature writes the Boys functions of a primitive quartet as a Gauss
rule, F_m(T) = sum_i w_i u_i^m (i < n, m < 2n), for the weight
exp(-T u) u^(-1/2) / 2 on u in [0, 1].  An ERI class (la lb|lc ld) with
L = la + lb + lc + ld needs n = L / 2 + 1 roots, and factorises per root into
x, y and z "2D" integrals:

  (ab|cd) = pref * sum_i w_i Ix(i) Iy(i) Iz(i)

Per root u and axis, with PA = P - A, QC = Q - C, PQ = P - Q,
rho = pq / (p + q), the 2D integrals I(n, m) (n on the bra, m on the ket)
follow the Obara-Saika recursion of the ERI with [..]^(m+1) -> u [..]:

  C00 = PA - rho/p PQ u,   C0p = QC + rho/q PQ u,
  B10 = (1 - rho/p u) / (2p),   B01 = (1 - rho/q u) / (2q),   B00 = u / (2(p + q))
  I(0,0) = 1,  I(n+1,0) = C00 I(n,0) + n B10 I(n-1,0)
  I(n,m+1) = C0p I(n,m) + m B01 I(n,m-1) + n B00 I(n-1,m)

then the horizontal recursions (a, b+1) = (a+1, b) + AB (a, b) and
(c, d+1) = (c+1, d) + CD (c, d) in each axis.  These are the recursions of
obara_saika_scheme.py:131-146 of the reference with the Boys order carried by
the root instead of by m.

Roots and weights: for T < RYS_TMAX, Chebyshev series on unit intervals of T
(RYS_NCHEB terms), fitted here to values computed in 60-digit arithmetic from
the moments F_m(T) (the Chebyshev algorithm for the recurrence coefficients,
then the eigenvalues of the Jacobi matrix); for T >= RYS_TMAX the large-T
limit u_i = s_i / T, w_i = l_i / (2 sqrt(T)) with (s_i, l_i) the Gauss rule of
the weight s^(-1/2) e^(-s) on [0, inf), whose error at T = 64 is below 1e-15
relative for n <= 5.

Usage:  python gen_rys.py OUT.h [LMAX]
"""
import math
import sys

NMAX_DEFAULT = 5
TMAX = 64
NCHEB = 16


def _mp():
    import mpmath
    mpmath.mp.dps = 60
    return mpmath


def gauss_from_moments(mp, mu, n):
    """Gauss rule (nodes ascending, weights) of n points from moments mu[0..2n-1]
    (Chebyshev algorithm for the recurrence coefficients, Jacobi eigenproblem)."""
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = list(mu)
    alpha = [mu[1] / mu[0]]
    beta = [mu[0]]
    for k in range(1, n):
        s_new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            s_new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha.append(s_new[k + 1] / s_new[k] - sig[k] / sig[k - 1])
        beta.append(s_new[k] / sig[k - 1])
        sig_prev, sig = sig, s_new
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    order = sorted(range(n), key=lambda i: E[i])
    return [E[i] for i in order], [mu[0] * Q[0, i] ** 2 for i in order]


def boys(mp, m, T):
    T = mp.mpf(T)
    if T == 0:
        return mp.mpf(1) / (2 * m + 1)
    return mp.gammainc(m + mp.mpf(1) / 2, 0, T) / (2 * T ** (m + mp.mpf(1) / 2))


def rys_rule(mp, n, T):
    return gauss_from_moments(mp, [boys(mp, k, T) for k in range(2 * n)], n)


def cheb_tables(mp, n):
    """[interval k][quantity q (roots, then weights)][coefficient j]."""
    nodes = [math.cos(math.pi * (j + 0.5) / NCHEB) for j in range(NCHEB)]
    out = []
    for k in range(TMAX):
        vals = []
        for x in nodes:
            u, w = rys_rule(mp, n, mp.mpf(k) + (mp.mpf(x) + 1) / 2)
            vals.append(u + w)
        coef = []
        for q in range(2 * n):
            c = []
            for m in range(NCHEB):
                s = mp.fsum(vals[j][q] * mp.cos(m * mp.pi * (j + mp.mpf(1) / 2) / NCHEB) for j in range(NCHEB))
                c.append(float(2 * s / NCHEB * (mp.mpf(1) / 2 if m == 0 else 1)))
            coef.append(c)
        out.append(coef)
    return out


def asymptotic(mp, n):
    s, l = gauss_from_moments(mp, [mp.gamma(k + mp.mpf(1) / 2) for k in range(2 * n)], n)
    return [float(x) for x in s], [float(x) for x in l]


def fmt(x):
    return repr(float(x))


def cart(l):
    if l == 0:
        return [(0, 0, 0)]
    if l == 1:
        return [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    if l == 2:
        return [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def emit_2d(la, lb, lc, ld):
    """Straight-line 2D integrals of one root and axis: G[((a (lb+1) + b)(lc+1) + c)(ld+1) + d]."""
    nb, nk = la + lb, lc + ld
    L = []
    I = {}
    I[(0, 0)] = '1.0'
    n_tmp = [0]

    def tmp(expr):
        name = 'i%d' % n_tmp[0]
        n_tmp[0] += 1
        L.append('    const double %s = %s;' % (name, expr))
        return name
    for n in range(1, nb + 1):
        e = 'C00 * %s' % I[(n - 1, 0)]
        if n >= 2:
            e += ' + %d.0 * B10 * %s' % (n - 1, I[(n - 2, 0)])
        I[(n, 0)] = tmp(e)
    for m in range(1, nk + 1):
        for n in range(0, nb + 1):
            e = 'C0p * %s' % I[(n, m - 1)]
            if m >= 2:
                e += ' + %d.0 * B01 * %s' % (m - 1, I[(n, m - 2)])
            if n >= 1:
                e += ' + %d.0 * B00 * %s' % (n, I[(n - 1, m - 1)])
            I[(n, m)] = tmp(e)
    # bra HRR: H[(a, b, m)] for a <= la, b <= lb, all m
    H = {}
    for m in range(nk + 1):
        for a in range(nb + 1):
            H[(a, 0, m)] = I[(a, m)]
        for b in range(1, lb + 1):
            for a in range(0, nb + 1 - b):
                H[(a, b, m)] = tmp('%s + AB * %s' % (H[(a + 1, b - 1, m)], H[(a, b - 1, m)]))
    # ket HRR per (a, b): K[(c, d)] from H[(a, b, c + ...)]
    out = []
    for a in range(la + 1):
        for b in range(lb + 1):
            K = {}
            for c in range(nk + 1):
                K[(c, 0)] = H[(a, b, c)]
            for d in range(1, ld + 1):
                for c in range(0, nk + 1 - d):
                    K[(c, d)] = tmp('%s + CD * %s' % (K[(c + 1, d - 1)], K[(c, d - 1)]))
            for c in range(lc + 1):
                for d in range(ld + 1):
                    idx = ((a * (lb + 1) + b) * (lc + 1) + c) * (ld + 1) + d
                    out.append('    G[%d] = %s;' % (idx, K[(c, d)]))
    return L + out


def emit_2d_grad(la, lb, lc, ld):
    """2D integrals of one root and axis and their derivatives with respect to
    the centres A, B, C of this axis: G, GA, GB, GC[((a (lb+1) + b)(lc+1) + c)(ld+1) + d]
    with  GA = 2 alpha E(a+1) - a E(a-1),  GB = 2 beta E(b+1) - b E(b-1),
          GC = 2 gamma E(c+1) - c E(c-1)
    (E the 2D integrals of the raised index; the exponents alpha, beta, gamma
    of the primitives on A, B, C)."""
    nb, nk = la + lb + 1, lc + ld + 1
    L = []
    n_tmp = [0]

    def tmp(expr):
        name = 'i%d' % n_tmp[0]
        n_tmp[0] += 1
        L.append('    const double %s = %s;' % (name, expr))
        return name
    I = {(0, 0): '1.0'}
    for n in range(1, nb + 1):
        e = 'C00 * %s' % I[(n - 1, 0)]
        if n >= 2:
            e += ' + %d.0 * B10 * %s' % (n - 1, I[(n - 2, 0)])
        I[(n, 0)] = tmp(e)
    for m in range(1, nk + 1):
        for n in range(0, nb + 1):
            if n + m > la + lb + lc + ld + 1:
                continue
            e = 'C0p * %s' % I[(n, m - 1)]
            if m >= 2:
                e += ' + %d.0 * B01 * %s' % (m - 1, I[(n, m - 2)])
            if n >= 1:
                e += ' + %d.0 * B00 * %s' % (n, I[(n - 1, m - 1)])
            I[(n, m)] = tmp(e)
    H = {}

    def hb(a, b, m):  # bra HRR, memoised
        key = (a, b, m)
        if key not in H:
            H[key] = I[(a, m)] if b == 0 else tmp('%s + AB * %s' % (hb(a + 1, b - 1, m), hb(a, b - 1, m)))
        return H[key]
    K = {}

    def E(a, b, c, d):  # ket HRR, memoised
        key = (a, b, c, d)
        if key not in K:
            K[key] = hb(a, b, c) if d == 0 else tmp('%s + CD * %s' % (E(a, b, c + 1, d - 1), E(a, b, c, d - 1)))
        return K[key]
    out = []
    for a in range(la + 1):
        for b in range(lb + 1):
            for c in range(lc + 1):
                for d in range(ld + 1):
                    idx = ((a * (lb + 1) + b) * (lc + 1) + c) * (ld + 1) + d
                    out.append('    G[%d] = %s;' % (idx, E(a, b, c, d)))
                    ga = '2.0 * alpha * %s' % E(a + 1, b, c, d)
                    if a:
                        ga += ' - %d.0 * %s' % (a, E(a - 1, b, c, d))
                    gb = '2.0 * beta * %s' % E(a, b + 1, c, d)
                    if b:
                        gb += ' - %d.0 * %s' % (b, E(a, b - 1, c, d))
                    gc = '2.0 * gamma * %s' % E(a, b, c + 1, d)
                    if c:
                        gc += ' - %d.0 * %s' % (c, E(a, b, c - 1, d))
                    out.append('    GA[%d] = %s;' % (idx, ga))
                    out.append('    GB[%d] = %s;' % (idx, gb))
                    out.append('    GC[%d] = %s;' % (idx, gc))
    return L + out


def emit_class(la, lb, lc, ld):
    L = la + lb + lc + ld
    n = L // 2 + 1
    ng = (la + 1) * (lb + 1) * (lc + 1) * (ld + 1)
    # output k (reference component order) -> offsets of its x, y, z factors
    offs = []
    for a in cart(la):
        for b in cart(lb):
            for c in cart(lc):
                for d in cart(ld):
                    o = [((a[i] * (lb + 1) + b[i]) * (lc + 1) + c[i]) * (ld + 1) + d[i] for i in range(3)]
                    offs.append(o[0] | (o[1] << 10) | (o[2] << 20))
    tag = '%d%d%d%d' % (la, lb, lc, ld)
    s = []
    s.append('__device__ const int gx_rys_off_%s[] = {%s};' % (tag, ', '.join(str(x) for x in offs)))
    s.append('template <> struct RysClass<%d, %d, %d, %d> {' % (la, lb, lc, ld))
    s.append('  static constexpr int L = %d, NROOT = %d, NG = %d, NOUT = %d;' % (L, n, ng, len(offs)))
    s.append('  static constexpr int NA = %d, NB = %d, NC = %d, ND = %d;' % (
        len(cart(la)), len(cart(lb)), len(cart(lc)), len(cart(ld))))
    s.append('  static __device__ __forceinline__ const int* off() { return gx_rys_off_%s; }' % tag)
    s.append('  // 2D integrals of one root and axis into G[NG]')
    s.append('  static __device__ __forceinline__ void g2d(double C00, double C0p, double B10, double B01,')
    s.append('                                              double B00, double AB, double CD, double* G) {')
    s.append('    (void)C00; (void)C0p; (void)B10; (void)B01; (void)B00; (void)AB; (void)CD;')
    s += emit_2d(la, lb, lc, ld)
    s.append('  }')
    if (L + 1) // 2 + 1 <= NMAX_DEFAULT:
        s.append('  // the gradient: one more root at L + 1; G, GA, GB, GC[NG] (see emit_2d_grad)')
        s.append('  static constexpr int NROOT_GRAD = %d;' % ((L + 1) // 2 + 1))
        s.append('  static __device__ __forceinline__ void g2d_grad(double C00, double C0p, double B10, double B01,')
        s.append('      double B00, double AB, double CD, double alpha, double beta, double gamma, double* G,')
        s.append('      double* GA, double* GB, double* GC) {')
        s.append('    (void)C00; (void)C0p; (void)B10; (void)B01; (void)B00; (void)AB; (void)CD;')
        s += emit_2d_grad(la, lb, lc, ld)
        s.append('  }')
    s.append('};')
    return '\n'.join(s)


def main():
    out = sys.argv[1]
    lmax = int(sys.argv[2]) if len(sys.argv) > 2 else 2
    nmax = min(NMAX_DEFAULT, 2 * lmax + 1)
    mp = _mp()
    parts = ['// Generated by gen_rys.py -- do not edit.', '#pragma once',
             '#define RYS_NMAX %d' % nmax, '#define RYS_TMAX %d' % TMAX, '#define RYS_NCHEB %d' % NCHEB, '']
    offs = [0]
    tabs = []
    asym_s, asym_l = [], []
    for n in range(1, nmax + 1):
        t = cheb_tables(mp, n)
        flat = [c for k in t for q in k for c in q]
        tabs += flat
        offs.append(offs[-1] + len(flat))
        s, l = asymptotic(mp, n)
        asym_s += s
        asym_l += l
    parts.append('// Chebyshev coefficients: n = 1..RYS_NMAX, [interval][root/weight q < 2n][term]')
    parts.append('__device__ const int gx_rys_cheb_off[] = {%s};' % ', '.join(str(o) for o in offs))
    lines = []
    for i in range(0, len(tabs), 4):
        lines.append('  ' + ', '.join(fmt(x) for x in tabs[i:i + 4]) + ',')
    parts.append('__device__ const double gx_rys_cheb[] = {\n' + '\n'.join(lines) + '\n};')
    parts.append('// large-T rule: n = 1..RYS_NMAX, n entries each, at offset n (n - 1) / 2')
    parts.append('__device__ const double gx_rys_asym_s[] = {%s};' % ', '.join(fmt(x) for x in asym_s))
    parts.append('__device__ const double gx_rys_asym_l[] = {%s};' % ', '.join(fmt(x) for x in asym_l))
    parts.append('')
    parts.append('template <int LA, int LB, int LC, int LD> struct RysClass;')
    pairs = [(a, b) for a in range(lmax + 1) for b in range(a + 1)]
    done = []
    for i, (la, lb) in enumerate(pairs):
        for (lc, ld) in pairs[:i + 1]:
            if (la + lb + lc + ld) // 2 + 1 <= nmax:
                parts.append(emit_class(la, lb, lc, ld))
                done.append((la, lb, lc, ld))
    parts.append('#define GX_FOR_EACH_RYS_CLASS(X) ' + ' '.join('X(%d, %d, %d, %d)' % c for c in done))
    with open(out, 'w') as f:
        f.write('\n'.join(parts) + '\n')


if __name__ == '__main__':
    main()
