/*
 * CPU oracle: Obara-Saika ERIs in plain C, restated from the reference.
 *
 * TEST INFRASTRUCTURE ONLY (never linked into the product library).
 *
 * Follows integrals/twoelectronrepulsion/obara_saika_scheme.py of the
 * reference (paths relative to src/gaussium/):
 *   os_integrate   -- ObaraSaika.integrate (:17-82): primitive-quartet loop,
 *                     p, q, rho, P, Q, W, T, prefactor, Boys F_L by the
 *                     reference series and downward recursion (:52-78).
 *   os_begin/os_rec-- os_begin/os_recursive (:84-146).  Like the reference,
 *                     the recursion continues with the centres in the order
 *                     os_gaussian_factory hands them over (bra/ket or
 *                     within-pair swaps are not undone), so this is an
 *                     independent evaluation path from oracle/os_eri.py,
 *                     which restores slot order.
 *   boys_ref       -- integrals/boys_function.py:4-76.
 *   canonical keys -- matrixelements/electron_repulsion_matrix.py:30-33.
 *
 * Basis functions are passed flat: per function a centre, (lx,ly,lz), a
 * primitive range; per primitive an exponent and a weight w = n_bf * c * N_prim
 * (the product the reference forms at obara_saika_scheme.py:34-50).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

typedef struct {
    double a;      /* exponent */
    double r[3];   /* centre */
    int l[3];      /* Cartesian exponents */
} prim_t;

typedef struct {
    double end[32]; /* prefactor * F_m(T), m = 0..L */
    double a7;      /* rho */
    double r7[3];   /* W */
} quartet_t;

/* 0: the reference's truncated series; 1: the convergent series
 * F_v(x) = e^-x sum_k (2x)^k / ((2v+1)(2v+3)...(2v+2k+1)) in long double for
 * x < 60 and the closed-form asymptote above, accurate to ~1e-16 relative
 * (pinned against mpmath in the tests). */
static int g_exact_boys = 0;

void oracle_set_exact_boys(int on) { g_exact_boys = on; }

static double boys_exact(int v, double x)
{
    /* x >= 60: F_v = Gamma(v+1/2) / (2 x^(v+1/2)) up to the upper incomplete
     * gamma term, < e^-60 x^(v-1/2) / Gamma(v+1/2) relative (< 1e-16 for v <= 12). */
    if (x >= 60.0)
        return (double)(tgammal(v + 0.5L) / (2.0L * powl((long double)x, v + 0.5L)));
    long double term = 1.0L / (2 * v + 1), sum = term;
    for (int k = 1; k < 100000; ++k) {
        term *= 2.0L * x / (2 * v + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-21L) break;
    }
    return (double)(expl(-(long double)x) * sum);
}

static double boys_ref(int v, double x)
{
    if (g_exact_boys) return boys_exact(v, x);
    double ans = 0.0, g_v = tgamma(v + 0.5);
    int i = 0;
    if (x <= 25.0) {
        for (;;) {
            double seq = (g_v / tgamma(v + i + 1.5)) * pow(x, (double)i);
            if (seq < 1e-10) break;
            ans += seq;
            ++i;
        }
        return ans * 0.5 * exp(-x);
    }
    for (;;) {
        double seq = (g_v / tgamma(v - i + 1.5)) * pow(x, (double)(-i));
        if (seq < 1e-10) break;
        ans += seq;
        ++i;
    }
    ans *= 0.5 * exp(-x);
    return (g_v / (2.0 * pow(x, v + 0.5))) - ans;
}

static double os_begin(const quartet_t *q, int m, prim_t g1, prim_t g2, prim_t g3, prim_t g4);

static double os_rec(const quartet_t *q, int r, int m, prim_t g1, prim_t g2, prim_t g3, prim_t g4)
{
    /* g1 is already lowered by one along r (os_gaussian_factory, :154-195). */
    double a5 = g1.a + g2.a, a6 = g3.a + g4.a;
    double r5 = (g1.a * g1.r[r] + g2.a * g2.r[r]) / a5;
    prim_t g5 = g1, g6 = g2, g7 = g3, g8 = g4;
    double out = 0.0;
    g5.l[r] -= 1;
    g6.l[r] -= 1;
    g7.l[r] -= 1;
    g8.l[r] -= 1;
    if (r5 != g1.r[r]) out += (r5 - g1.r[r]) * os_begin(q, m, g1, g2, g3, g4);
    if (q->r7[r] != r5) out += (q->r7[r] - r5) * os_begin(q, m + 1, g1, g2, g3, g4);
    if (g5.l[r] >= 0) {
        double n = g1.l[r] ? g1.l[r] : 1;
        out += n * (1.0 / (2.0 * a5)) * os_begin(q, m, g5, g2, g3, g4);
        out -= n * (q->a7 / (2.0 * a5 * a5)) * os_begin(q, m + 1, g5, g2, g3, g4);
    }
    if (g6.l[r] >= 0) {
        double n = g2.l[r] ? g2.l[r] : 1;
        out += n * (1.0 / (2.0 * a5)) * os_begin(q, m, g1, g6, g3, g4);
        out -= n * (q->a7 / (2.0 * a5 * a5)) * os_begin(q, m + 1, g1, g6, g3, g4);
    }
    if (g7.l[r] >= 0) {
        double n = g3.l[r] ? g3.l[r] : 1;
        out += n * (1.0 / (2.0 * (a5 + a6))) * os_begin(q, m + 1, g1, g2, g7, g4);
    }
    if (g8.l[r] >= 0) {
        double n = g4.l[r] ? g4.l[r] : 1;
        out += n * (1.0 / (2.0 * (a5 + a6))) * os_begin(q, m + 1, g1, g2, g3, g8);
    }
    return out;
}

static double os_begin(const quartet_t *q, int m, prim_t g1, prim_t g2, prim_t g3, prim_t g4)
{
    int r;
    for (r = 0; r < 3; ++r)
        if (g1.l[r] > 0) { g1.l[r] -= 1; return os_rec(q, r, m, g1, g2, g3, g4); }
    for (r = 0; r < 3; ++r)
        if (g2.l[r] > 0) { g2.l[r] -= 1; return os_rec(q, r, m, g2, g1, g4, g3); }
    for (r = 0; r < 3; ++r)
        if (g3.l[r] > 0) { g3.l[r] -= 1; return os_rec(q, r, m, g3, g4, g1, g2); }
    for (r = 0; r < 3; ++r)
        if (g4.l[r] > 0) { g4.l[r] -= 1; return os_rec(q, r, m, g4, g3, g2, g1); }
    return q->end[m];
}

static double dist2(const double *a, const double *b)
{
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

static double os_integrate(const double *ctr, const int *lxyz, const int *poff, const int *np,
                           const double *pexp, const double *pw, int i, int j, int k, int l)
{
    const int idx[4] = {i, j, k, l};
    int n[4], o[4], t, ltot = 0;
    double ans = 0.0;
    for (t = 0; t < 4; ++t) {
        n[t] = np[idx[t]];
        o[t] = poff[idx[t]];
        ltot += lxyz[3 * idx[t]] + lxyz[3 * idx[t] + 1] + lxyz[3 * idx[t] + 2];
    }
    for (int p1 = 0; p1 < n[0]; ++p1)
    for (int p2 = 0; p2 < n[1]; ++p2)
    for (int p3 = 0; p3 < n[2]; ++p3)
    for (int p4 = 0; p4 < n[3]; ++p4) {
        prim_t g[4];
        int pp[4] = {o[0] + p1, o[1] + p2, o[2] + p3, o[3] + p4};
        double w = 1.0;
        quartet_t q;
        for (t = 0; t < 4; ++t) {
            g[t].a = pexp[pp[t]];
            for (int c = 0; c < 3; ++c) {
                g[t].r[c] = ctr[3 * idx[t] + c];
                g[t].l[c] = lxyz[3 * idx[t] + c];
            }
            w *= pw[pp[t]];
        }
        double a5 = g[0].a + g[1].a, a6 = g[2].a + g[3].a;
        double r5[3], r6[3];
        q.a7 = a5 * a6 / (a5 + a6);
        for (int c = 0; c < 3; ++c) {
            r5[c] = (g[0].a * g[0].r[c] + g[1].a * g[1].r[c]) / a5;
            r6[c] = (g[2].a * g[2].r[c] + g[3].a * g[3].r[c]) / a6;
            q.r7[c] = (a5 * r5[c] + a6 * r6[c]) / (a5 + a6);
        }
        double x = a5 * a6 * dist2(r5, r6) / (a5 + a6);
        double out1 = (2.0 * pow(M_PI, 2.5)) / (a5 * a6 * sqrt(a5 + a6));
        double out2 = exp((-g[0].a * g[1].a * dist2(g[0].r, g[1].r) / a5)
                          - (g[2].a * g[3].a * dist2(g[2].r, g[3].r) / a6));
        double f = boys_ref(ltot, x);
        q.end[ltot] = out1 * out2 * f;
        for (int m = ltot; m >= 1; --m) {
            f = (exp(-x) + 2.0 * x * f) / (2 * m - 1);
            q.end[m - 1] = out1 * out2 * f;
        }
        ans += w * os_begin(&q, 0, g[0], g[1], g[2], g[3]);
    }
    return ans;
}

/* Values of the requested (i,j,k,l) quartets. */
int oracle_eri_keys(int nbf, const double *ctr, const int *lxyz, const int *poff, const int *np,
                    const double *pexp, const double *pw, int64_t nkeys, const int32_t *keys,
                    double *out)
{
    (void)nbf;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nkeys; ++q)
        out[q] = os_integrate(ctr, lxyz, poff, np, pexp, pw,
                              keys[4 * q], keys[4 * q + 1], keys[4 * q + 2], keys[4 * q + 3]);
    return 0;
}

/* Dense N^4 tensor: canonical keys evaluated once, scattered to all 8 images. */
int oracle_eri_dense(int nbf, const double *ctr, const int *lxyz, const int *poff, const int *np,
                     const double *pexp, const double *pw, double *out)
{
    const int64_t n = nbf;
    const int64_t npair = n * (n + 1) / 2;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ab = 0; ab < npair; ++ab) {
        /* unrank ab -> (a, b) with a <= b, row-major over a */
        int64_t a = 0, rem = ab;
        while (rem >= n - a) { rem -= n - a; ++a; }
        int64_t b = a + rem;
        for (int64_t c = a; c < n; ++c)
            for (int64_t d = c; d < n; ++d) {
                if (a == c && b > d) continue;
                double v = os_integrate(ctr, lxyz, poff, np, pexp, pw, (int)a, (int)b, (int)c, (int)d);
                const int64_t im[8][4] = {{a, b, c, d}, {b, a, c, d}, {a, b, d, c}, {b, a, d, c},
                                          {c, d, a, b}, {d, c, a, b}, {c, d, b, a}, {d, c, b, a}};
                for (int t = 0; t < 8; ++t)
                    out[((im[t][0] * n + im[t][1]) * n + im[t][2]) * n + im[t][3]] = v;
            }
    }
    return 0;
}

/* Reference Boys series, exposed for the oracle tests. */
double oracle_boys(int v, double x) { return boys_ref(v, x); }
