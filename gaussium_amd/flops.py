"""Canonical algorithmic FLOP model of one direct J/K build (implementation independent).

Per primitive quartet of class (la lb|lc ld), la >= lb, lc >= ld, L = sum:
  setup 25; Boys 20 + 4L;
  VRR: for every (e, f), 0 <= e <= la+lb, 0 <= f <= lc+ld:
       (0,0): L+1 (prefactor x F_m);
       else for every m <= L-e-f and every Cartesian pair (a of order e, c of
       order f): lower the first non-zero x,y,z index of the bra (the ket if
       e = 0); 2 x terms, terms = 2, +2 if that index minus 1 > 0, +1 if the
       same index on the other side > 0;
  contraction: 2 per [e0|f0] element, la <= e <= la+lb, lc <= f <= lc+ld.
Per contracted shell quartet: HRR, 2 per generated element (bra HRR for
every ket component block, then ket HRR for the final bra components).
Digestion: 12 per unique basis-function quartet.

The counts are FMA = 2 flops and do not depend on how a kernel schedules
the work; roofline fractions use them against the FP64 (or FP32) peak.

Gradient (one derivative shell quartet, d/dA, d/dB, d/dC; d/dD by
translational invariance), per primitive quartet: setup 25 + 3 (the
exponent weights); Boys 20 + 4 (L + 1); one VRR over 0 <= e <= la+lb+1,
0 <= f <= lc+ld+1 without the corner (e, f) = (la+lb+1, lc+ld+1) (no
centre pair is raised together), costed as above with L + 1 in place of L;
contraction of four weighted sets (unweighted for the lowered terms, 2a, 2b,
2c): 2 x 4 per [e0|f0] element with la <= e <= la+lb+1, lc <= f <= lc+ld+1
(corner excluded).  Per contracted shell quartet: the HRR of the three raised
classes (hrr_flops with l_X + 1) and of the unraised class (for the lowered
terms).  Per basis-function quartet: the density factor D (8 flops) and nine
derivatives, each 2 a (a+1|..) - a (a-1|..) contracted with D (3 + 2 flops).
"""
from functools import lru_cache


def ncart(l):
    return (l + 1) * (l + 2) // 2


def cart(l):
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


@lru_cache(maxsize=None)
def prim_flops(la, lb, lc, ld):
    L = la + lb + lc + ld
    f = 25 + 20 + 4 * L
    for e in range(la + lb + 1):
        for g in range(lc + ld + 1):
            if e == 0 and g == 0:
                f += L + 1
                continue
            nm = L - e - g + 1
            for a in cart(e):
                for c in cart(g):
                    if e > 0:
                        i = next(k for k in range(3) if a[k] > 0)
                        terms = 2 + (2 if a[i] - 1 > 0 else 0) + (1 if c[i] > 0 else 0)
                    else:
                        i = next(k for k in range(3) if c[k] > 0)
                        terms = 2 + (2 if c[i] - 1 > 0 else 0) + (1 if a[i] > 0 else 0)
                    f += nm * 2 * terms
    nbra = sum(ncart(e) for e in range(la, la + lb + 1))
    nket = sum(ncart(g) for g in range(lc, lc + ld + 1))
    f += 2 * nbra * nket
    return f


@lru_cache(maxsize=None)
def hrr_flops(la, lb, lc, ld):
    """2 flops per element generated by the bra then the ket HRR."""
    f = 0
    nket_blocks = sum(ncart(g) for g in range(lc, lc + ld + 1))
    # bra: (a, b) built from [e0| for b = 1..lb, a from la .. la+lb-b
    bra = 0
    for b in range(1, lb + 1):
        for a in range(la, la + lb - b + 1):
            bra += ncart(a) * ncart(b)
    f += 2 * bra * nket_blocks
    ket = 0
    for d in range(1, ld + 1):
        for c in range(lc, lc + ld - d + 1):
            ket += ncart(c) * ncart(d)
    f += 2 * ket * ncart(la) * ncart(lb)
    return f


@lru_cache(maxsize=None)
def grad_prim_flops(la, lb, lc, ld):
    """Model flops of one derivative primitive quartet (see the module text)."""
    L = la + lb + lc + ld + 1
    eb, fb = la + lb + 1, lc + ld + 1
    f = 25 + 3 + 20 + 4 * L
    for e in range(eb + 1):
        for g in range(fb + 1):
            if e == eb and g == fb:
                continue
            if e == 0 and g == 0:
                f += L + 1
                continue
            nm = L - e - g + 1
            for a in cart(e):
                for c in cart(g):
                    if e > 0:
                        i = next(k for k in range(3) if a[k] > 0)
                        terms = 2 + (2 if a[i] - 1 > 0 else 0) + (1 if c[i] > 0 else 0)
                    else:
                        i = next(k for k in range(3) if c[k] > 0)
                        terms = 2 + (2 if c[i] - 1 > 0 else 0) + (1 if a[i] > 0 else 0)
                    f += nm * 2 * terms
    nel = 0
    for e in range(la, eb + 1):
        for g in range(lc, fb + 1):
            if e == eb and g == fb:
                continue
            nel += ncart(e) * ncart(g)
    f += 2 * 4 * nel
    return f


def grad_hrr_flops(la, lb, lc, ld):
    return (hrr_flops(la + 1, lb, lc, ld) + hrr_flops(la, lb + 1, lc, ld) + hrr_flops(la, lb, lc + 1, ld)
            + hrr_flops(la, lb, lc, ld))


def grad_model_flops(rows):
    """Model flops of the two-electron gradient (FP64 part, FP32 part) from
    ``gx_jk_class_counts`` rows taken with the gradient's thresholds."""
    f64 = f32 = 0.0
    for la, lb, lc, ld, s64, s32, p64, p32, b64, b32 in rows:
        pf, hf = grad_prim_flops(la, lb, lc, ld), grad_hrr_flops(la, lb, lc, ld)
        f64 += p64 * pf + s64 * hf + (8.0 + 9.0 * 5.0) * b64
        f32 += p32 * pf + s32 * hf + (8.0 + 9.0 * 5.0) * b32
    return f64, f32


def jk_model_flops(rows):
    """Model flops of the FP64 and FP32 parts from ``gx_jk_class_counts`` rows.

    Each row: la, lb, lc, ld, shq64, shq32, primq64, primq32, bfq64, bfq32.
    """
    f64 = f32 = 0.0
    for la, lb, lc, ld, s64, s32, p64, p32, b64, b32 in rows:
        pf, hf = prim_flops(la, lb, lc, ld), hrr_flops(la, lb, lc, ld)
        f64 += p64 * pf + s64 * hf + 12.0 * b64
        f32 += p32 * pf + s32 * hf + 12.0 * b32
    return f64, f32


def check_table():
    """The per-class K=1 totals quoted with the model (setup + Boys + VRR + contraction)."""
    return {c: prim_flops(*c) for c in [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0),
                                        (1, 1, 1, 0), (1, 1, 1, 1), (2, 0, 0, 0), (2, 1, 1, 0),
                                        (2, 2, 2, 2)]}
