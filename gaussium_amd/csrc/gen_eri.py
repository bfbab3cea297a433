#!/usr/bin/env python3
"""Generate straight-line HIP device code for each ERI angular-momentum class.

For a shell quartet class (la lb | lc ld) (la >= lb, lc >= ld) this emits

* ``prim<R>``: one primitive quartet of the Obara-Saika vertical recursion
  (VRR) on centre A (bra) and C (ket), accumulating the contracted
  [e0|f0]^(0) intermediates, la <= |e| <= la+lb, lc <= |f| <= lc+ld;
* ``hrr<R>``: the Head-Gordon-Pople horizontal recursion (HRR) moving
  angular momentum from A to B and from C to D on the contracted
  intermediates, producing (ab|cd) in the reference's component order.

Everything is unrolled into SSA temporaries so that the compiler keeps the
recursion in VGPRs; ``R`` is double (FP64 class) or float (FP32 class).

Recursions (``a``/``c`` Cartesian exponent triples, ``i`` an axis, ``m`` the
Boys order; PA=P-A, WP=W-P, QC=Q-C, WQ=W-Q, AB=A-B, CD=C-D):

  [a+1i|0]^m  = PA_i [a|0]^m + WP_i [a|0]^{m+1}
               + a_i/(2p) ([a-1i|0]^m - rho/p [a-1i|0]^{m+1})
  [a|c+1i]^m  = QC_i [a|c]^m + WQ_i [a|c]^{m+1}
               + c_i/(2q) ([a|c-1i]^m - rho/q [a|c-1i]^{m+1})
               + a_i/(2(p+q)) [a-1i|c-1i]^{m+1}
  (a,b+1i|    = (a+1i,b| + AB_i (a,b|            (likewise on the ket)

These are the recursions of obara_saika_scheme.py:131-146 (reference) in the
VRR+HRR arrangement of head_gordon_pople.py:111-181 (reference).

Usage:  python gen_eri.py OUT.h [LMAX]
"""
import sys

AXES = 'xyz'


def cart(l):
    """Cartesian components of angular momentum l, reference order for l <= 2."""
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


def ncart(l):
    return (l + 1) * (l + 2) // 2


def sub(t, i, k=1):
    t = list(t)
    t[i] -= k
    return tuple(t)


def add(t, i, k=1):
    t = list(t)
    t[i] += k
    return tuple(t)


def pick_axis(t):
    """Axis to lower: the non-zero component with the smallest value (fewest terms)."""
    best = None
    for i in range(3):
        if t[i] > 0 and (best is None or t[i] < t[best]):
            best = i
    return best


class Emitter:
    def __init__(self, R):
        self.lines = []
        self.n = 0
        self.R = R

    def tmp(self, expr):
        name = 't%d' % self.n
        self.n += 1
        self.lines.append('    const %s %s = %s;' % (self.R, name, expr))
        return name


def gen_vrr(la, lb, lc, ld, fsub=None):
    """Body of prim<R>: returns (lines, accumulator layout).  fsub (ld = 0
    only): the ket components to accumulate (all when None)."""
    L = la + lb + lc + ld
    em = Emitter('R')
    memo = {}
    zero = (0, 0, 0)

    def node(e, f, m):
        key = (e, f, m)
        if key in memo:
            return memo[key]
        if min(e) < 0 or min(f) < 0:
            raise AssertionError
        if e == zero and f == zero:
            memo[key] = 'F[%d]' % m
            return memo[key]
        terms = []
        if f == zero:
            i = pick_axis(e)
            em1 = sub(e, i)
            ax = AXES[i]
            terms.append('PA%s * %s' % (ax, node(em1, f, m)))
            terms.append('WP%s * %s' % (ax, node(em1, f, m + 1)))
            if e[i] >= 2:
                em2 = sub(e, i, 2)
                k = e[i] - 1
                a0 = node(em2, f, m)
                a1 = node(em2, f, m + 1)
                terms.append('R(%d) * o2p * (%s - rho_p * %s)' % (k, a0, a1))
        else:
            i = pick_axis(f)
            fm1 = sub(f, i)
            ax = AXES[i]
            terms.append('QC%s * %s' % (ax, node(e, fm1, m)))
            terms.append('WQ%s * %s' % (ax, node(e, fm1, m + 1)))
            if f[i] >= 2:
                fm2 = sub(f, i, 2)
                k = f[i] - 1
                c0 = node(e, fm2, m)
                c1 = node(e, fm2, m + 1)
                terms.append('R(%d) * o2q * (%s - rho_q * %s)' % (k, c0, c1))
            if e[i] >= 1:
                terms.append('R(%d) * o2pq * %s' % (e[i], node(sub(e, i), fm1, m + 1)))
        name = em.tmp(' + '.join(terms))
        memo[key] = name
        return name

    layout = []
    for le in range(la, la + lb + 1):
        for e in cart(le):
            for lf in range(lc, lc + ld + 1):
                for f in cart(lf):
                    if fsub is None or f in fsub:
                        layout.append((e, f))
    # each accumulation right after the value it adds is first available
    for idx, (e, f) in enumerate(layout):
        v = node(e, f, 0)
        em.lines.append('    acc[%d] += %s;' % (idx, v))
    assert max(m for (_, _, m) in memo) <= L
    return em.lines, layout


def gen_hrr(la, lb, lc, ld, layout, fsub=None):
    """Body of hrr<R>: contracted [e0|f0] -> (ab|cd) in reference component order
    (fsub: only the ket components c in it, in their cart(lc) order)."""
    em = Emitter('R')
    index = {ef: k for k, ef in enumerate(layout)}
    zero = (0, 0, 0)
    bra_memo = {}

    def bra(a, b, f):
        key = (a, b, f)
        if key in bra_memo:
            return bra_memo[key]
        if b == zero:
            v = 'acc[%d]' % index[(a, f)]
        else:
            i = [k for k in range(3) if b[k] > 0][0]
            v = em.tmp('%s + AB%s * %s' % (bra(add(a, i), sub(b, i), f), AXES[i], bra(a, sub(b, i), f)))
        bra_memo[key] = v
        return v

    ket_memo = {}

    def ket(a, b, c, d):
        key = (a, b, c, d)
        if key in ket_memo:
            return ket_memo[key]
        if d == zero:
            v = bra(a, b, c)
        else:
            i = [k for k in range(3) if d[k] > 0][0]
            v = em.tmp('%s + CD%s * %s' % (ket(a, b, add(c, i), sub(d, i)), AXES[i], ket(a, b, c, sub(d, i))))
        ket_memo[key] = v
        return v

    out = []
    k = 0
    for a in cart(la):
        for b in cart(lb):
            for c in cart(lc):
                if fsub is not None and c not in fsub:
                    continue
                for d in cart(ld):
                    out.append('    out[%d] = %s;' % (k, ket(a, b, c, d)))
                    k += 1
    return em.lines + out


def classes(lmax):
    pairs = [(a, b) for a in range(lmax + 1) for b in range(a + 1)]
    out = []
    for i, (la, lb) in enumerate(pairs):
        for (lc, ld) in pairs[:i + 1]:
            out.append((la, lb, lc, ld))
    return out


def emit_class(la, lb, lc, ld):
    vrr, layout = gen_vrr(la, lb, lc, ld)
    hrr = gen_hrr(la, lb, lc, ld, layout)
    L = la + lb + lc + ld
    na, nb, nc, nd = ncart(la), ncart(lb), ncart(lc), ncart(ld)
    s = []
    s.append('template <> struct EriClass<%d, %d, %d, %d> {' % (la, lb, lc, ld))
    s.append('  static constexpr int L = %d, NACC = %d, NOUT = %d;' % (L, len(layout), na * nb * nc * nd))
    s.append('  static constexpr int NA = %d, NB = %d, NC = %d, ND = %d;' % (na, nb, nc, nd))
    s.append('  template <typename R>')
    s.append('  static __device__ __forceinline__ void prim(R (&acc)[NACC], const R (&F)[L + 1],')
    s.append('      R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, R QCx, R QCy, R QCz,')
    s.append('      R WQx, R WQy, R WQz, R o2p, R o2q, R o2pq, R rho_p, R rho_q) {')
    s.append('    (void)PAx; (void)PAy; (void)PAz; (void)WPx; (void)WPy; (void)WPz;')
    s.append('    (void)QCx; (void)QCy; (void)QCz; (void)WQx; (void)WQy; (void)WQz;')
    s.append('    (void)o2p; (void)o2q; (void)o2pq; (void)rho_p; (void)rho_q;')
    s += vrr
    s.append('  }')
    s.append('  template <typename R>')
    s.append('  static __device__ __forceinline__ void hrr(const R (&acc)[NACC], R ABx, R ABy, R ABz,')
    s.append('      R CDx, R CDy, R CDz, R (&out)[NOUT]) {')
    s.append('    (void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz;')
    s += hrr
    s.append('  }')
    s.append('};')
    return '\n'.join(s)


# ---------------------------------------------------------------------------
# Per-lane J/K classes without ket HRR (ld = 0) whose contracted
# intermediates exceed JK_PASS_NACC run in passes over the ket components:
# (ab|c0) for the components c of one pass needs only the [e0|c]
# accumulators of those c, and the J/K digestion of the passes adds up.  A
# pass repeats the primitive loop (Boys function and bra recursion) but keeps
# a fraction of the registers: (dd|ds) takes 186 accumulators and 216 outputs
# per lane in one pass (512 registers, 1.4 KB of scratch per lane).
JK_PASS_NACC = 64


def jk_lane_passes(c):
    """Ket-component groups (indices into cart(lc)) of the J/K passes of a
    per-lane class, grouped by the axis their recursion lowers so that a pass
    shares its [e0|p] intermediates; one group when the class runs in one pass."""
    la, lb, lc, ld = c
    comps = cart(lc)
    ne = sum(ncart(e) for e in range(la, la + lb + 1))
    if ld != 0 or is_coop(c) or ne * len(comps) <= JK_PASS_NACC:
        return [list(range(len(comps)))]
    per = max(1, JK_PASS_NACC // ne)
    npass = -(-len(comps) // per)
    order = sorted(range(len(comps)), key=lambda k: (pick_axis(comps[k]), k))
    return [sorted(order[len(order) * i // npass:len(order) * (i + 1) // npass]) for i in range(npass)]


def emit_jk_pass(la, lb, lc, ld, P, group):
    comps = cart(lc)
    fsub = [comps[k] for k in group]
    vrr, layout = gen_vrr(la, lb, lc, ld, fsub)
    hrr = gen_hrr(la, lb, lc, ld, layout, fsub)
    L = la + lb + lc + ld
    na, nb, nd = ncart(la), ncart(lb), ncart(ld)
    sel = ''.join('i == %d ? %d : ' % (i, k) for i, k in enumerate(group[:-1])) + '%d' % group[-1]
    s = []
    s.append('template <> struct EriJKPass<%d, %d, %d, %d, %d> {' % (la, lb, lc, ld, P))
    s.append('  static constexpr int L = %d, NACC = %d, NOUT = %d;' % (L, len(layout), na * nb * len(group) * nd))
    s.append('  static constexpr int NA = %d, NB = %d, NC = %d, ND = %d;' % (na, nb, len(group), nd))
    s.append('  // ket component (index into cart(lc)) of the pass\'s component i')
    s.append('  static __host__ __device__ constexpr int cidx(int i) { return %s; }' % sel)
    s.append('  template <typename R>')
    s.append('  static __device__ __forceinline__ void prim(R (&acc)[NACC], const R (&F)[L + 1],')
    s.append('      R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, R QCx, R QCy, R QCz,')
    s.append('      R WQx, R WQy, R WQz, R o2p, R o2q, R o2pq, R rho_p, R rho_q) {')
    s.append('    (void)PAx; (void)PAy; (void)PAz; (void)WPx; (void)WPy; (void)WPz;')
    s.append('    (void)QCx; (void)QCy; (void)QCz; (void)WQx; (void)WQy; (void)WQz;')
    s.append('    (void)o2p; (void)o2q; (void)o2pq; (void)rho_p; (void)rho_q;')
    s += vrr
    s.append('  }')
    s.append('  template <typename R>')
    s.append('  static __device__ __forceinline__ void hrr(const R (&acc)[NACC], R ABx, R ABy, R ABz,')
    s.append('      R CDx, R CDy, R CDz, R (&out)[NOUT]) {')
    s.append('    (void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz;')
    s += hrr
    s.append('  }')
    s.append('};')
    return '\n'.join(s)


# ---------------------------------------------------------------------------
# Programs for the wave-cooperative evaluator.  Classes whose contracted
# intermediates do not fit a lane's registers are evaluated by a whole wave
# on one shell quartet: the VRR table [e0|f0]^(m) and the HRR table live in
# LDS and the lanes share the entries of each recursion level.  The programs
# list those entries level by level:
#
#   VRR entry (int4):  x = dest | kind << 16 | axis << 17 | k1 << 20 | k2 << 24
#                      y = s0 | s1 << 16, z = s2 | s3 << 16, w = s4
#     kind 0 (bra):  T[dest] = PA_i T[s0] + WP_i T[s1] + k1 o2p (T[s2] - rho_p T[s3])
#     kind 1 (ket):  T[dest] = QC_i T[s0] + WQ_i T[s1] + k1 o2q (T[s2] - rho_q T[s3])
#                              + k2 o2pq T[s4]
#     slot 0 holds 0, slots 1..L+1 hold pref * F_m.
#   accumulator t (t < NACC) adds T[acc[t]] (the [e0|f0]^(0) layout of
#   EriClass::prim).
#   HRR entry (int2):  x = dest | kind << 16 | axis << 17, y = s0 | s1 << 16
#     kind 0: H[dest] = H[s0] + AB_i H[s1];  kind 1: H[dest] = H[s0] + CD_i H[s1]
#     slots 0..NACC-1 hold the contracted accumulators.
#   out[k] = H[outslot[k]] in the reference component order.
# ---------------------------------------------------------------------------
COOP_NACC = 80  # classes with more contracted intermediates than this run one quartet per wave
               # (J/K: the Rys kernel; (pp|pp), 81 intermediates, 32 -> 18 us there at the
               # 255-function cluster) ...
COOP_NACC_KET_S = 200  # ... except those without ket HRR (ld = 0) up to this many (measured on
                       # gfx950: per-lane code, even spilling, beats the wave-per-quartet kernels
                       # there: (dp|ds) 0.38 -> 0.11 ms, (dd|ps) 0.29 -> 0.07, (dd|ds) 0.39 -> 0.24
                       # at the 255-function cluster against the Obara-Saika programs, and
                       # 0.23-0.26 ms for the Rys kernel, whose one quartet per wave does not
                       # suit their many contracted quartets)


def compact_hrr(hentries, nacc, fixed, keep):
    """Physical H slots for the HRR entries (level, (x, y)) written with
    logical slots: a slot is reused once the value it holds has been read for
    the last time at an earlier level (entries of one level run in parallel,
    so a level never writes a slot that the same level reads).  Slots
    0..nacc-1 (the accumulators) and `fixed` keep their positions; values in
    `keep` stay live to the end.  Returns (entries, logical->physical, NH)."""
    import heapq
    hentries = sorted(hentries, key=lambda t: t[0])
    ents = [(lev, x & 0xffff, x >> 16, y & 0xffff, y >> 16) for lev, (x, y) in hentries]
    inf = 1 << 30
    last = {}
    for lev, d, fl, s0, s1 in ents:
        last[s0] = max(last.get(s0, 0), lev)
        last[s1] = max(last.get(s1, 0), lev)
    for k in keep:
        last[k] = inf
    for f in fixed:
        last[f] = inf
    phys = {i: i for i in range(nacc)}
    for f in fixed:
        phys[f] = f
    live = dict(phys)  # logical -> physical
    top = max([nacc - 1] + list(fixed)) + 1
    free = []
    out = []
    levels = sorted(set(e[0] for e in ents))
    k = 0
    for lev in levels:
        for lg in [lg for lg in live if last.get(lg, -1) < lev]:
            heapq.heappush(free, live.pop(lg))
        while k < len(ents) and ents[k][0] == lev:
            _, d, fl, s0, s1 = ents[k]
            if free:
                pd = heapq.heappop(free)
            else:
                pd = top
                top += 1
            phys[d] = pd
            live[d] = pd
            out.append((lev, (pd | fl << 16, phys[s0] | phys[s1] << 16)))
            k += 1
    return out, phys, top


def coop_program(la, lb, lc, ld):
    L = la + lb + lc + ld
    zero = (0, 0, 0)
    memo = {}
    entries = []  # (level, packed int4)
    nslot = [L + 2]

    def node(e, f, m):
        if e == zero and f == zero:
            return 1 + m
        key = (e, f, m)
        if key in memo:
            return memo[key]
        s2 = s3 = s4 = 0
        k1 = k2 = 0
        if f == zero:
            kind = 0
            i = pick_axis(e)
            em1 = sub(e, i)
            s0, s1 = node(em1, f, m), node(em1, f, m + 1)
            if e[i] >= 2:
                k1 = e[i] - 1
                s2, s3 = node(sub(e, i, 2), f, m), node(sub(e, i, 2), f, m + 1)
        else:
            kind = 1
            i = pick_axis(f)
            fm1 = sub(f, i)
            s0, s1 = node(e, fm1, m), node(e, fm1, m + 1)
            if f[i] >= 2:
                k1 = f[i] - 1
                s2, s3 = node(e, sub(f, i, 2), m), node(e, sub(f, i, 2), m + 1)
            if e[i] >= 1:
                k2 = e[i]
                s4 = node(sub(e, i), fm1, m + 1)
        dest = nslot[0]
        nslot[0] += 1
        assert dest < 65536 and k1 < 16 and k2 < 16
        lev = sum(e) + sum(f)
        entries.append((lev, (dest | kind << 16 | i << 17 | k1 << 20 | k2 << 24, s0 | s1 << 16,
                              s2 | s3 << 16, s4)))
        memo[key] = dest
        return dest

    layout = []
    for le in range(la, la + lb + 1):
        for e in cart(le):
            for lf in range(lc, lc + ld + 1):
                for f in cart(lf):
                    layout.append((e, f))
    acc = [node(e, f, 0) for (e, f) in layout]
    entries.sort(key=lambda t: t[0])  # stable: creation order inside a level
    vlev = [0] * (L + 2)
    for lev, _ in entries:
        vlev[lev + 1] += 1
    for k in range(1, L + 2):
        vlev[k] += vlev[k - 1]
    vrr = [e for _, e in entries]

    index = {ef: k for k, ef in enumerate(layout)}
    hslot = [len(layout)]
    hentries = []
    bra_memo = {}

    def bra(a, b, f):
        key = (a, b, f)
        if key in bra_memo:
            return bra_memo[key]
        if b == zero:
            v = index[(a, f)]
        else:
            i = [k for k in range(3) if b[k] > 0][0]
            s0, s1 = bra(add(a, i), sub(b, i), f), bra(a, sub(b, i), f)
            v = hslot[0]
            hslot[0] += 1
            hentries.append((sum(b), (v | 0 << 16 | i << 17, s0 | s1 << 16)))
        bra_memo[key] = v
        return v

    ket_memo = {}

    def ket(a, b, c, d):
        key = (a, b, c, d)
        if key in ket_memo:
            return ket_memo[key]
        if d == zero:
            v = bra(a, b, c)
        else:
            i = [k for k in range(3) if d[k] > 0][0]
            s0, s1 = ket(a, b, add(c, i), sub(d, i)), ket(a, b, c, sub(d, i))
            v = hslot[0]
            hslot[0] += 1
            hentries.append((lb + sum(d), (v | 1 << 16 | i << 17, s0 | s1 << 16)))
        ket_memo[key] = v
        return v

    out = [ket(a, b, c, d) for a in cart(la) for b in cart(lb) for c in cart(lc) for d in cart(ld)]
    assert hslot[0] < 65536
    hentries, phys, nh = compact_hrr(hentries, len(layout), [], set(out))
    out = [phys[o] for o in out]
    hslot[0] = nh
    nhl = lb + ld + 1
    hlev = [0] * (nhl + 1)
    for lev, _ in hentries:
        hlev[lev] += 1  # levels 1..lb+ld -> hlev[lev] counts; prefix below
    hl = [0]
    for lev in range(1, nhl):
        hl.append(hl[-1] + hlev[lev])
    hrr = [e for _, e in hentries]
    return dict(L=L, NACC=len(layout), NOUT=len(out), NT=nslot[0], NH=hslot[0], vrr=vrr, vlev=vlev,
                acc=acc, hrr=hrr, hlev=hl, out=out)


# ---------------------------------------------------------------------------
# Gradient programs.  The nuclear derivative of a contracted quartet on
# centre A is  d/dA_i (ab|cd) = sum_prim [2 alpha_a (a+1i b|cd) - a_i (a-1i b|cd)]
# (likewise B and C; D follows from translational invariance).  One VRR per
# primitive quartet covers every shifted class; its [e0|f0] values are
# accumulated into four weighted sets:
#   set 0: weight 1        (lowered A, B, C terms)
#   set 1: weight 2 alpha_a (raised A),  set 2: 2 alpha_b (raised B),
#   set 3: 2 alpha_c (raised C)
# and HRR runs per set.  For every component k of (la lb|lc ld) and every
# (centre X in A, B, C; axis i) the program lists the H slots of the raised
# and lowered values and the integer factor:  g = H[plus] - n H[minus].
#   accumulator entry (unsigned int): T slot | set << 16
#   gradient term (int2): x = plus, y = minus | n << 16   (minus = the zero slot if n = 0)
# ---------------------------------------------------------------------------
def grad_program(la, lb, lc, ld, centers=(0, 1, 2)):
    """Derivative program of (ab|cd) with respect to the centres in `centers`
    (0 = A, 1 = B, 2 = C): the raised set of each of them (weight 2 alpha)
    and the lowered (unweighted) set they need; 3 terms per centre."""
    L = la + lb + lc + ld + 1
    zero = (0, 0, 0)
    memo = {}
    entries = []
    nslot = [L + 2]

    def node(e, f, m):
        if e == zero and f == zero:
            return 1 + m
        key = (e, f, m)
        if key in memo:
            return memo[key]
        s2 = s3 = s4 = 0
        k1 = k2 = 0
        if f == zero:
            kind = 0
            i = pick_axis(e)
            em1 = sub(e, i)
            s0, s1 = node(em1, f, m), node(em1, f, m + 1)
            if e[i] >= 2:
                k1 = e[i] - 1
                s2, s3 = node(sub(e, i, 2), f, m), node(sub(e, i, 2), f, m + 1)
        else:
            kind = 1
            i = pick_axis(f)
            fm1 = sub(f, i)
            s0, s1 = node(e, fm1, m), node(e, fm1, m + 1)
            if f[i] >= 2:
                k1 = f[i] - 1
                s2, s3 = node(e, sub(f, i, 2), m), node(e, sub(f, i, 2), m + 1)
            if e[i] >= 1:
                k2 = e[i]
                s4 = node(sub(e, i), fm1, m + 1)
        dest = nslot[0]
        nslot[0] += 1
        assert dest < 65536 and k1 < 16 and k2 < 16
        entries.append((sum(e) + sum(f), (dest | kind << 16 | i << 17 | k1 << 20 | k2 << 24, s0 | s1 << 16,
                                          s2 | s3 << 16, s4)))
        memo[key] = dest
        return dest

    def rect(e0, e1, f0, f1):
        return [(le, lf) for le in range(max(e0, 0), e1 + 1) for lf in range(max(f0, 0), f1 + 1)]

    raised = {1: rect(la + 1, la + lb + 1, lc, lc + ld), 2: rect(la, la + lb + 1, lc, lc + ld),
              3: rect(la, la + lb, lc + 1, lc + ld + 1)}
    sets = {0: []}
    for X in centers:
        sets[X + 1] = raised[X + 1]
    low = set()
    if la >= 1 and 0 in centers:
        low |= set(rect(la - 1, la + lb - 1, lc, lc + ld))
    if lb >= 1 and 1 in centers:
        low |= set(rect(la, la + lb - 1, lc, lc + ld))
    if lc >= 1 and 2 in centers:
        low |= set(rect(la, la + lb, lc - 1, lc + ld - 1))
    sets[0] = sorted(low)
    acc = []
    index = {}
    for st in sorted(sets):
        for (le, lf) in sets[st]:
            for e in cart(le):
                for f in cart(lf):
                    index[(st, e, f)] = len(acc)
                    acc.append(node(e, f, 0) | st << 16)
    nacc = len(acc)
    entries.sort(key=lambda t: t[0])
    vlev = [0] * (L + 2)
    for lev, _ in entries:
        vlev[lev + 1] += 1
    for k in range(1, L + 2):
        vlev[k] += vlev[k - 1]
    vrr = [e for _, e in entries]

    hslot = [nacc + 1]  # slot nacc holds 0
    zslot = nacc
    hentries = []
    bra_memo = {}

    def bra(st, a, b, f):
        key = (st, a, b, f)
        if key in bra_memo:
            return bra_memo[key]
        if b == zero:
            v = index[(st, a, f)]
        else:
            i = [k for k in range(3) if b[k] > 0][0]
            s0, s1 = bra(st, add(a, i), sub(b, i), f), bra(st, a, sub(b, i), f)
            v = hslot[0]
            hslot[0] += 1
            hentries.append((sum(b), (v | 0 << 16 | i << 17, s0 | s1 << 16)))
        bra_memo[key] = v
        return v

    ket_memo = {}
    lbmax = lb + 1

    def ket(st, a, b, c, d):
        key = (st, a, b, c, d)
        if key in ket_memo:
            return ket_memo[key]
        if d == zero:
            v = bra(st, a, b, c)
        else:
            i = [k for k in range(3) if d[k] > 0][0]
            s0, s1 = ket(st, a, b, add(c, i), sub(d, i)), ket(st, a, b, c, sub(d, i))
            v = hslot[0]
            hslot[0] += 1
            hentries.append((lbmax + sum(d), (v | 1 << 16 | i << 17, s0 | s1 << 16)))
        ket_memo[key] = v
        return v

    def unit(i):
        return tuple(1 if k == i else 0 for k in range(3))

    terms = []
    for a in cart(la):
        for b in cart(lb):
            for c in cart(lc):
                for d in cart(ld):
                    for X in centers:
                        for i in range(3):
                            if X == 0:
                                plus = ket(1, add(a, i), b, c, d)
                                n = a[i]
                                minus = ket(0, sub(a, i), b, c, d) if n else zslot
                            elif X == 1:
                                plus = ket(2, a, add(b, i), c, d)
                                n = b[i]
                                minus = ket(0, a, sub(b, i), c, d) if n else zslot
                            else:
                                plus = ket(3, a, b, add(c, i), d)
                                n = c[i]
                                minus = ket(0, a, b, sub(c, i), d) if n else zslot
                            terms.append((plus, minus, n))
    assert hslot[0] < 65536
    keep = set(t[0] for t in terms) | set(t[1] for t in terms)
    hentries, phys, nh = compact_hrr(hentries, nacc, [zslot], keep)
    terms = [(phys[pl], phys[mi] | n << 16) for pl, mi, n in terms]
    hslot[0] = nh
    nhl = lbmax + ld + 1
    hcount = [0] * (nhl + 1)
    for lev, _ in hentries:
        hcount[lev] += 1
    hl = [0]
    for lev in range(1, nhl):
        hl.append(hl[-1] + hcount[lev])
    hrr = [e for _, e in hentries]
    return dict(L=L, NACC=nacc, NOUT=len(terms) // (3 * len(centers)), NT=nslot[0], NH=hslot[0], ZSLOT=zslot,
                vrr=vrr, vlev=vlev, acc=acc, hrr=hrr, hlev=hl, NHLEV=nhl - 1, terms=terms, centers=tuple(centers))


def emit_grad_prog(la, lb, lc, ld):
    g = grad_program(la, lb, lc, ld)
    tag = 'g%d%d%d%d' % (la, lb, lc, ld)
    na, nb, nc, nd = ncart(la), ncart(lb), ncart(lc), ncart(ld)

    def arr(name, typ, vals, per_line=8):
        items = ['{%s}' % ', '.join(str(x) for x in v) if isinstance(v, tuple) else str(v) for v in vals]
        if not items:
            items = ['{}']
        lines = [', '.join(items[k:k + per_line]) for k in range(0, len(items), per_line)]
        return '__device__ const %s %s_%s[] = {\n  %s};' % (typ, name, tag, ',\n  '.join(lines))

    s = [arr('gx_vrr', 'int4', g['vrr'], 4), arr('gx_vlev', 'int', g['vlev']),
         arr('gx_acc', 'unsigned int', g['acc'], 12), arr('gx_hrr', 'int2', g['hrr'], 6),
         arr('gx_hlev', 'int', g['hlev']), arr('gx_gterm', 'int2', g['terms'], 6)]
    s.append('template <> struct EriGradProg<%d, %d, %d, %d> {' % (la, lb, lc, ld))
    s.append('  static constexpr int L = %d, NACC = %d, NOUT = %d, NT = %d, NH = %d, ZSLOT = %d;'
             % (g['L'], g['NACC'], g['NOUT'], g['NT'], g['NH'], g['ZSLOT']))
    s.append('  static constexpr int NA = %d, NB = %d, NC = %d, ND = %d;' % (na, nb, nc, nd))
    s.append('  static constexpr int NVLEV = %d, NHLEV = %d;' % (g['L'], g['NHLEV']))
    for fn, typ in (('vrr', 'int4'), ('vlev', 'int'), ('acc', 'unsigned int'), ('hrr', 'int2'),
                    ('hlev', 'int'), ('gterm', 'int2')):
        s.append('  static __device__ __forceinline__ const %s* %s() { return gx_%s_%s; }' % (typ, fn, fn, tag))
    s.append('};')
    return '\n'.join(s)


# ---------------------------------------------------------------------------
# Per-lane gradient code for the small classes: the gradient program written
# out as straight-line code (each VRR slot a register, the HRR slots renamed
# on every write), one shell quartet per lane.  A class whose whole program
# fits the registers runs in one pass; one whose per-centre programs fit runs
# three passes over the primitives (A, B, C derivatives; the lowered set is
# recomputed per pass) instead of spilling; the rest run wave-cooperatively.
GRAD_LANE_NACC = 120  # one pass when the whole program has at most this many weighted intermediates
GRAD_PASS_NACC = 120  # three passes when each per-centre program has at most this many


def grad_lane_passes(c):
    """1 (one pass over all centres), 3 (one pass per centre) or 0 (cooperative)."""
    if grad_program(*c)['NACC'] <= GRAD_LANE_NACC:
        return 1
    if max(grad_program(*c, centers=(x,))['NACC'] for x in range(3)) <= GRAD_PASS_NACC:
        return 3
    return 0


def emit_grad_lane_pass(la, lb, lc, ld, X, centers):
    g = grad_program(la, lb, lc, ld, centers=centers)
    L = g['L']
    na, nb, nc, nd = ncart(la), ncart(lb), ncart(lc), ncart(ld)
    axes = 'xyz'
    lines = []

    def var(slot):
        if slot == 0:
            return 'R(0)'
        if 1 <= slot <= L + 1:
            return 'F[%d]' % (slot - 1)
        return 'v%d' % slot

    # each accumulation follows the definition of the value it adds, so that
    # a VRR value lives only until its last use
    wname = ['', 'w1 * ', 'w2 * ', 'w3 * ']
    acc_of = {}
    for t, e in enumerate(g['acc']):
        acc_of.setdefault(e & 0xffff, []).append('    acc[%d] += %s%s;' % (t, wname[e >> 16], var(e & 0xffff)))
    for slot in sorted(acc_of):
        if slot <= L + 1:
            lines += acc_of[slot]
    for (x, y, z, w) in g['vrr']:
        dest, kind, ax = x & 0xffff, (x >> 16) & 1, (x >> 17) & 3
        k1, k2 = (x >> 20) & 15, (x >> 24) & 15
        s0, s1, s2, s3, s4 = y & 0xffff, y >> 16, z & 0xffff, z >> 16, w
        cA = ('QC' if kind else 'PA') + axes[ax]
        cB = ('WQ' if kind else 'WP') + axes[ax]
        terms = ['%s * %s' % (cA, var(s0)), '%s * %s' % (cB, var(s1))]
        if k1:
            o, r = ('o2q', 'rho_q') if kind else ('o2p', 'rho_p')
            terms.append('R(%d) * %s * (%s - %s * %s)' % (k1, o, var(s2), r, var(s3)))
        if k2:
            terms.append('R(%d) * o2pq * %s' % (k2, var(s4)))
        lines.append('    const R v%d = %s;' % (dest, ' + '.join(terms)))
        lines += acc_of.get(dest, [])
    fin = []
    cur = {t: 'acc[%d]' % t for t in range(g['NACC'])}
    cur[g['ZSLOT']] = 'R(0)'
    for n_, (x, y) in enumerate(g['hrr']):
        dest, kind, ax = x & 0xffff, (x >> 16) & 1, (x >> 17) & 3
        s0, s1 = y & 0xffff, y >> 16
        Xs = ('CD' if kind else 'AB') + axes[ax]
        fin.append('    const R h%d = %s + %s * %s;' % (n_, cur[s0], Xs, cur[s1]))
        cur[dest] = 'h%d' % n_
    nterm = 3 * len(centers)
    k = 0
    for ia in range(na):
        for ib in range(nb):
            for ic in range(nc):
                for id_ in range(nd):
                    fin.append('    {')
                    fin.append('      const R D = f * (R(4) * Pab[%d] * Pcd[%d] - Pac[%d] * Pbd[%d] - Pad[%d] * Pbc[%d]);'
                               % (ia * nb + ib, ic * nd + id_, ia * nc + ic, ib * nd + id_, ia * nd + id_,
                                  ib * nc + ic))
                    for j in range(nterm):
                        gi = 3 * centers[j // 3] + j % 3
                        plus, mword = g['terms'][nterm * k + j]
                        minus, n = mword & 0xffff, mword >> 16
                        if n:
                            fin.append('      g[%d] += D * (%s - R(%d) * %s);' % (gi, cur[plus], n, cur[minus]))
                        else:
                            fin.append('      g[%d] += D * %s;' % (gi, cur[plus]))
                    fin.append('    }')
                    k += 1
    s = []
    s.append('template <> struct EriGradLanePass<%d, %d, %d, %d, %d> {' % (la, lb, lc, ld, X))
    s.append('  static constexpr int L = %d, NACC = %d;' % (L, g['NACC']))
    s.append('  template <typename R>')
    s.append('  static __device__ __forceinline__ void prim(R (&acc)[NACC], const R (&F)[L + 1],')
    s.append('      R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, R QCx, R QCy, R QCz,')
    s.append('      R WQx, R WQy, R WQz, R o2p, R o2q, R o2pq, R rho_p, R rho_q, R w1, R w2, R w3) {')
    s.append('    (void)PAx; (void)PAy; (void)PAz; (void)WPx; (void)WPy; (void)WPz;')
    s.append('    (void)QCx; (void)QCy; (void)QCz; (void)WQx; (void)WQy; (void)WQz;')
    s.append('    (void)o2p; (void)o2q; (void)o2pq; (void)rho_p; (void)rho_q; (void)w1; (void)w2; (void)w3;')
    s += lines
    s.append('  }')
    s.append('  // HRR of the weighted sets and contraction with')
    s.append('  // D_k = f (4 P_ab P_cd - P_ac P_bd - P_ad P_bc):  g[3 X + i] += D_k d(ab|cd)_k / dX_i')
    s.append('  // for the centres X of this pass (%s)' % ', '.join('ABC'[x] for x in centers))
    s.append('  template <typename R>')
    s.append('  static __device__ __forceinline__ void finish(const R (&acc)[NACC], R ABx, R ABy, R ABz,')
    s.append('      R CDx, R CDy, R CDz, const R* Pab, const R* Pcd, const R* Pac, const R* Pad,')
    s.append('      const R* Pbc, const R* Pbd, R f, R (&g)[9]) {')
    s.append('    (void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz;')
    s += fin
    s.append('  }')
    s.append('};')
    return '\n'.join(s)


def emit_grad_lane(la, lb, lc, ld):
    npass = grad_lane_passes((la, lb, lc, ld))
    assert npass in (1, 3)
    parts = []
    if npass == 1:
        parts.append(emit_grad_lane_pass(la, lb, lc, ld, 0, (0, 1, 2)))
    else:
        for X in range(3):
            parts.append(emit_grad_lane_pass(la, lb, lc, ld, X, (X,)))
    na, nb, nc, nd = ncart(la), ncart(lb), ncart(lc), ncart(ld)
    L = la + lb + lc + ld + 1
    parts.append('template <> struct EriGradLane<%d, %d, %d, %d> {\n'
                 '  static constexpr int L = %d, NPASS = %d;\n'
                 '  static constexpr int NA = %d, NB = %d, NC = %d, ND = %d;\n};'
                 % (la, lb, lc, ld, L, npass, na, nb, nc, nd))
    return '\n'.join(parts)


def grad_is_lane(c):
    return grad_lane_passes(c) > 0


def emit_prog(la, lb, lc, ld):
    g = coop_program(la, lb, lc, ld)
    tag = '%d%d%d%d' % (la, lb, lc, ld)
    na, nb, nc, nd = ncart(la), ncart(lb), ncart(lc), ncart(ld)

    def arr(name, typ, vals, per_line=8):
        items = ['{%s}' % ', '.join(str(x) for x in v) if isinstance(v, tuple) else str(v) for v in vals]
        if not items:
            items = ['{}' if typ != 'int' else '0']
        lines = [', '.join(items[k:k + per_line]) for k in range(0, len(items), per_line)]
        return '__device__ const %s %s_%s[] = {\n  %s};' % (typ, name, tag, ',\n  '.join(lines))

    s = [arr('gx_vrr', 'int4', g['vrr'], 4), arr('gx_vlev', 'int', g['vlev']), arr('gx_acc', 'unsigned short', g['acc'], 16),
         arr('gx_hrr', 'int2', g['hrr'], 6), arr('gx_hlev', 'int', g['hlev']),
         arr('gx_out', 'unsigned short', g['out'], 16)]
    s.append('template <> struct EriProg<%d, %d, %d, %d> {' % (la, lb, lc, ld))
    s.append('  static constexpr int L = %d, NACC = %d, NOUT = %d, NT = %d, NH = %d;'
             % (g['L'], g['NACC'], g['NOUT'], g['NT'], g['NH']))
    s.append('  static constexpr int NA = %d, NB = %d, NC = %d, ND = %d;' % (na, nb, nc, nd))
    s.append('  static constexpr int NVLEV = %d, NHLEV = %d;  // VRR levels 1..L, HRR levels 1..LB+LD'
             % (g['L'], lb + ld))
    s.append('  static __device__ __forceinline__ const int4* vrr() { return gx_vrr_%s; }' % tag)
    s.append('  static __device__ __forceinline__ const int* vlev() { return gx_vlev_%s; }' % tag)
    s.append('  static __device__ __forceinline__ const unsigned short* acc() { return gx_acc_%s; }' % tag)
    s.append('  static __device__ __forceinline__ const int2* hrr() { return gx_hrr_%s; }' % tag)
    s.append('  static __device__ __forceinline__ const int* hlev() { return gx_hlev_%s; }' % tag)
    s.append('  static __device__ __forceinline__ const unsigned short* out() { return gx_out_%s; }' % tag)
    s.append('};')
    return '\n'.join(s)


def is_coop(c):
    la, lb, lc, ld = c
    nacc = sum(ncart(e) for e in range(la, la + lb + 1)) * sum(ncart(f) for f in range(lc, lc + ld + 1))
    return nacc > (COOP_NACC_KET_S if ld == 0 else COOP_NACC)


def main():
    out_path = sys.argv[1]
    lmax = int(sys.argv[2]) if len(sys.argv) > 2 else 2
    parts = ['// Generated by gen_eri.py -- do not edit.',
             '#pragma once',
             'template <int LA, int LB, int LC, int LD> struct EriClass;',
             'template <int LA, int LB, int LC, int LD> struct EriProg;',
             'template <int LA, int LB, int LC, int LD> struct EriGradProg;',
             'template <int LA, int LB, int LC, int LD> struct EriGradLane;',
             'template <int LA, int LB, int LC, int LD, int X> struct EriGradLanePass;',
             'template <int LA, int LB, int LC, int LD, int P> struct EriJKPass;',
             '// J/K passes of a per-lane class (EriJKPass<..., 0 .. NPASS - 1> when NPASS > 1)\n'
             'template <int LA, int LB, int LC, int LD> struct EriJKLane {\n'
             '  static constexpr int NPASS = 1;\n};']
    lane = [c for c in classes(lmax) if not is_coop(c)]
    coop = [c for c in classes(lmax) if is_coop(c)]
    for c in lane:
        parts.append(emit_class(*c))
        groups = jk_lane_passes(c)
        if len(groups) > 1:
            parts.append('template <> struct EriJKLane<%d, %d, %d, %d> {\n  static constexpr int NPASS = %d;\n};'
                         % (c + (len(groups),)))
            for P, g in enumerate(groups):
                parts.append(emit_jk_pass(*c, P, g))
    for c in classes(lmax):
        parts.append(emit_prog(*c))
    for c in classes(lmax):
        parts.append(emit_grad_prog(*c))
    glane = [c for c in classes(lmax) if grad_is_lane(c)]
    for c in glane:
        parts.append(emit_grad_lane(*c))
    parts.append('#define GX_ERI_LMAX %d' % lmax)
    parts.append('#define GX_FOR_EACH_CLASS(X) ' + ' '.join('X(%d, %d, %d, %d)' % c for c in classes(lmax)))
    parts.append('#define GX_FOR_EACH_LANE_CLASS(X) ' + ' '.join('X(%d, %d, %d, %d)' % c for c in lane))
    parts.append('#define GX_FOR_EACH_COOP_CLASS(X) ' + ' '.join('X(%d, %d, %d, %d)' % c for c in coop))
    pcs = [(a, b) for a in range(lmax + 1) for b in range(a + 1)]
    parts.append('#define GX_FOR_EACH_PAIRCLASS(X) ' + ' '.join('X(%d, %d)' % c for c in pcs))
    parts.append('#define GX_FOR_EACH_LANE_PAIRCLASS(X) ' + ' '.join(
        'X(%d, %d)' % c for c in pcs if not is_coop(c + c)))
    parts.append('#define GX_FOR_EACH_COOP_PAIRCLASS(X) ' + ' '.join(
        'X(%d, %d)' % c for c in pcs if is_coop(c + c)))
    parts.append('constexpr bool gx_class_is_coop(int la, int lb, int lc, int ld) {\n  return false' + ''.join(
        ' ||\n         (la == %d && lb == %d && lc == %d && ld == %d)' % c for c in coop) + ';\n}')
    parts.append('#define GX_FOR_EACH_GRAD_LANE_CLASS(X) ' + ' '.join('X(%d, %d, %d, %d)' % c for c in glane))
    parts.append('#define GX_FOR_EACH_GRAD_COOP_CLASS(X) ' + ' '.join(
        'X(%d, %d, %d, %d)' % c for c in classes(lmax) if c not in glane))
    parts.append('constexpr bool gx_grad_is_lane(int la, int lb, int lc, int ld) {\n  return false' + ''.join(
        ' ||\n         (la == %d && lb == %d && lc == %d && ld == %d)' % c for c in glane) + ';\n}')
    text = '\n\n'.join(parts) + '\n'
    with open(out_path, 'w') as f:
        f.write(text)


if __name__ == '__main__':
    main()
