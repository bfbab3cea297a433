// Host (CPU) build of the generated class code: evaluates contracted shell
// quartets with the same VRR/HRR source the GPU kernels compile, so that the
// generated recursions can be checked against the oracle without a GPU.
// Boys values come from a long-double series here (the device table is
// checked separately on the GPU).
#include <cmath>
#include <cstdint>

#define __device__
#define __host__
#define __forceinline__ inline
struct int2 {
  int x, y;
};
struct int4 {
  int x, y, z, w;
};
#include "eri_gen.h"
using std::fma;
using std::sqrt;
#include "gx_rys.h"

static void boys_host(int L, double T, double scale, double* F) {
  const int M = L;
  long double term = 1.0L / (2 * M + 1), sum = term;
  for (int k = 1; k < 5000; ++k) {
    term *= 2.0L * T / (2 * M + 2 * k + 1);
    sum += term;
    if (term < sum * 1e-22L) break;
  }
  const long double e = expl(-(long double)T);
  long double f = e * sum;
  F[M] = (double)(f * scale);
  for (int m = M; m >= 1; --m) {
    f = (2.0L * T * f + e) / (2 * m - 1);
    F[m - 1] = (double)(f * scale);
  }
}

struct Shell {
  double c[3];
  int n;
  const double* a;
  const double* w;
};

// One contracted quartet through class code K (EriClass, or one J/K pass
// EriJKPass): out[0 .. K::NOUT) in K's component order.
template <class K>
static void quartet_cls(const Shell& A, const Shell& B, const Shell& C, const Shell& D, double* out) {
  double acc[K::NACC] = {0};
  const double two_pi_52 = 34.98683665524972;
  for (int i = 0; i < A.n; ++i)
    for (int j = 0; j < B.n; ++j) {
      const double a = A.a[i], b = B.a[j], p = a + b;
      double P[3], ab2 = 0;
      for (int x = 0; x < 3; ++x) {
        P[x] = (a * A.c[x] + b * B.c[x]) / p;
        ab2 += (A.c[x] - B.c[x]) * (A.c[x] - B.c[x]);
      }
      const double Kab = A.w[i] * B.w[j] * std::exp(-a * b / p * ab2);
      for (int k = 0; k < C.n; ++k)
        for (int l = 0; l < D.n; ++l) {
          const double c = C.a[k], d = D.a[l], q = c + d;
          double Q[3], cd2 = 0, pq2 = 0;
          for (int x = 0; x < 3; ++x) {
            Q[x] = (c * C.c[x] + d * D.c[x]) / q;
            cd2 += (C.c[x] - D.c[x]) * (C.c[x] - D.c[x]);
            pq2 += (P[x] - Q[x]) * (P[x] - Q[x]);
          }
          const double Kcd = C.w[k] * D.w[l] * std::exp(-c * d / q * cd2);
          const double rho = p * q / (p + q);
          const double pref = two_pi_52 / (p * q * std::sqrt(p + q)) * Kab * Kcd;
          double F[K::L + 1];
          boys_host(K::L, rho * pq2, pref, F);
          const double rp = rho / p, rq = rho / q;
          double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
          K::template prim<double>(acc, F, P[0] - A.c[0], P[1] - A.c[1], P[2] - A.c[2], -rp * PQ[0],
                                   -rp * PQ[1], -rp * PQ[2], Q[0] - C.c[0], Q[1] - C.c[1],
                                   Q[2] - C.c[2], rq * PQ[0], rq * PQ[1], rq * PQ[2], 0.5 / p,
                                   0.5 / q, 0.5 / (p + q), rp, rq);
        }
    }
  double o[K::NOUT];
  K::template hrr<double>(acc, A.c[0] - B.c[0], A.c[1] - B.c[1], A.c[2] - B.c[2], C.c[0] - D.c[0],
                          C.c[1] - D.c[1], C.c[2] - D.c[2], o);
  for (int t = 0; t < K::NOUT; ++t) out[t] = o[t];
}

// The J/K passes P.. of a class evaluated in passes over its ket components,
// each pass's outputs put at their place in the class's component order.
template <int LA, int LB, int LC, int LD, int P>
static void quartet_passes(const Shell& A, const Shell& B, const Shell& C, const Shell& D, double* out) {
  if constexpr (P < EriJKLane<LA, LB, LC, LD>::NPASS) {
    using K = EriJKPass<LA, LB, LC, LD, P>;
    double o[K::NOUT];
    quartet_cls<K>(A, B, C, D, o);
    constexpr int NCF = (LC + 1) * (LC + 2) / 2;
    for (int ia = 0; ia < K::NA; ++ia)
      for (int ib = 0; ib < K::NB; ++ib)
        for (int ic = 0; ic < K::NC; ++ic)
          for (int id = 0; id < K::ND; ++id)
            out[((ia * K::NB + ib) * NCF + K::cidx(ic)) * K::ND + id] = o[((ia * K::NB + ib) * K::NC + ic) * K::ND + id];
    quartet_passes<LA, LB, LC, LD, P + 1>(A, B, C, D, out);
  }
}

// A per-lane class the way the J/K kernel evaluates it: in one pass, or in
// passes over the ket components.
template <int LA, int LB, int LC, int LD>
static void quartet(const Shell& A, const Shell& B, const Shell& C, const Shell& D, double* out) {
  if constexpr (EriJKLane<LA, LB, LC, LD>::NPASS > 1)
    quartet_passes<LA, LB, LC, LD, 0>(A, B, C, D, out);
  else
    quartet_cls<EriClass<LA, LB, LC, LD>>(A, B, C, D, out);
}

template <class G>
static bool run_hrr(const Shell& A, const Shell& B, const Shell& C, const Shell& D, double* H, bool* hok);

// gradient programs keep one zero slot right after the accumulators
template <class G, class = void>
struct has_zslot {
  static constexpr int value = 0;
};
template <class G>
struct has_zslot<G, decltype((void)G::ZSLOT)> {
  static constexpr int value = 1;
};
template <class G>
constexpr int zero_slots() {
  return has_zslot<G>::value;
}

// VRR of one primitive quartet by the program's levels into T (every source
// written at an earlier level, every slot written once).
template <class G>
static bool run_vrr(const double* F, const double* PA, const double* QC, const double* WP, const double* WQ,
                    double p, double q, double rp, double rq, double* T, bool* ok) {
  for (int t = 0; t < G::NT; ++t) ok[t] = false;
  T[0] = 0.0;
  ok[0] = true;
  for (int m = 0; m <= G::L; ++m) {
    T[1 + m] = F[m];
    ok[1 + m] = true;
  }
  const int4* vp = G::vrr();
  const int* vl = G::vlev();
  for (int lev = 1; lev <= G::L; ++lev) {
    for (int t = vl[lev]; t < vl[lev + 1]; ++t) {
      const unsigned y = (unsigned)vp[t].y, z = (unsigned)vp[t].z;
      const unsigned src[5] = {y & 0xffffu, y >> 16, z & 0xffffu, z >> 16, (unsigned)vp[t].w};
      for (unsigned v : src)
        if (v >= (unsigned)G::NT || !ok[v]) return false;
    }
    for (int t = vl[lev]; t < vl[lev + 1]; ++t) {
      const unsigned x = (unsigned)vp[t].x, y = (unsigned)vp[t].y, z = (unsigned)vp[t].z;
      const bool ket = (x >> 16) & 1u;
      const unsigned ax = (x >> 17) & 3u;
      const double k1 = (double)((x >> 20) & 15u), k2 = (double)((x >> 24) & 15u);
      const double cA = ket ? QC[ax] : PA[ax], cB = ket ? WQ[ax] : WP[ax];
      const double o = ket ? 0.5 / q : 0.5 / p, r = ket ? rq : rp;
      const unsigned dst = x & 0xffffu;
      if (dst >= (unsigned)G::NT || ok[dst]) return false;
      T[dst] = cA * T[y & 0xffffu] + cB * T[y >> 16] + k1 * o * (T[z & 0xffffu] - r * T[z >> 16]) +
               k2 * (0.5 / (p + q)) * T[vp[t].w];
    }
    for (int t = vl[lev]; t < vl[lev + 1]; ++t) ok[(unsigned)vp[t].x & 0xffffu] = true;
  }
  return true;
}

// The gradient program of the class (EriGradProg): for every component k and
// (centre X in A, B, C; axis i), d/dX_i of the contracted quartet, k-major.
template <int LA, int LB, int LC, int LD>
static int quartet_grad(const Shell& A, const Shell& B, const Shell& C, const Shell& D, double* out) {
  using G = EriGradProg<LA, LB, LC, LD>;
  static double T[G::NT], H[G::NH + G::NACC + G::NT];
  static bool ok[G::NT], hok[G::NH + 1];
  for (int t = 0; t < G::NACC + 1; ++t) H[t] = 0.0;  // slot NACC is the zero slot
  const double two_pi_52 = 34.98683665524972;
  for (int i = 0; i < A.n; ++i)
    for (int j = 0; j < B.n; ++j) {
      const double a = A.a[i], b = B.a[j], p = a + b;
      double P[3], ab2 = 0;
      for (int x = 0; x < 3; ++x) {
        P[x] = (a * A.c[x] + b * B.c[x]) / p;
        ab2 += (A.c[x] - B.c[x]) * (A.c[x] - B.c[x]);
      }
      const double Kab = A.w[i] * B.w[j] * std::exp(-a * b / p * ab2);
      for (int k = 0; k < C.n; ++k)
        for (int l = 0; l < D.n; ++l) {
          const double c = C.a[k], d = D.a[l], q = c + d;
          double Q[3], cd2 = 0, pq2 = 0;
          for (int x = 0; x < 3; ++x) {
            Q[x] = (c * C.c[x] + d * D.c[x]) / q;
            cd2 += (C.c[x] - D.c[x]) * (C.c[x] - D.c[x]);
            pq2 += (P[x] - Q[x]) * (P[x] - Q[x]);
          }
          const double Kcd = C.w[k] * D.w[l] * std::exp(-c * d / q * cd2);
          const double rho = p * q / (p + q);
          const double pref = two_pi_52 / (p * q * std::sqrt(p + q)) * Kab * Kcd;
          double F[G::L + 1];
          boys_host(G::L, rho * pq2, pref, F);
          const double rp = rho / p, rq = rho / q;
          const double PA[3] = {P[0] - A.c[0], P[1] - A.c[1], P[2] - A.c[2]};
          const double QC[3] = {Q[0] - C.c[0], Q[1] - C.c[1], Q[2] - C.c[2]};
          double WP[3], WQ[3];
          for (int x = 0; x < 3; ++x) {
            WP[x] = -rp * (P[x] - Q[x]);
            WQ[x] = rq * (P[x] - Q[x]);
          }
          if (!run_vrr<G>(F, PA, QC, WP, WQ, p, q, rp, rq, T, ok)) return -2;
          const double wt[4] = {1.0, 2.0 * a, 2.0 * b, 2.0 * c};
          for (int t = 0; t < G::NACC; ++t) {
            const unsigned e = G::acc()[t];
            if (!ok[e & 0xffffu]) return -4;
            H[t] += wt[e >> 16] * T[e & 0xffffu];
          }
        }
    }
  if (!run_hrr<G>(A, B, C, D, H, hok)) return -5;
  if (H[G::ZSLOT] != 0.0) return -6;
  for (int t = 0; t < 9 * G::NOUT; ++t) {
    const int2 g = G::gterm()[t];
    const unsigned plus = (unsigned)g.x, minus = (unsigned)g.y & 0xffffu, n = (unsigned)g.y >> 16;
    if (!hok[plus] || !hok[minus]) return -7;
    out[t] = H[plus] - n * H[minus];
  }
  return 0;
}

// The cooperative program of the class (EriProg), run entry by entry in level
// order as the lanes of a wave run it.  Every source slot must have been
// written before it is read (checked).
template <int LA, int LB, int LC, int LD>
static int quartet_prog(const Shell& A, const Shell& B, const Shell& C, const Shell& D, double* out) {
  using G = EriProg<LA, LB, LC, LD>;
  static double T[G::NT], H[G::NH + G::NACC + G::NT];
  static bool ok[G::NT], hok[G::NH];
  for (int t = 0; t < G::NACC; ++t) H[t] = 0.0;
  const double two_pi_52 = 34.98683665524972;
  for (int i = 0; i < A.n; ++i)
    for (int j = 0; j < B.n; ++j) {
      const double a = A.a[i], b = B.a[j], p = a + b;
      double P[3], ab2 = 0;
      for (int x = 0; x < 3; ++x) {
        P[x] = (a * A.c[x] + b * B.c[x]) / p;
        ab2 += (A.c[x] - B.c[x]) * (A.c[x] - B.c[x]);
      }
      const double Kab = A.w[i] * B.w[j] * std::exp(-a * b / p * ab2);
      for (int k = 0; k < C.n; ++k)
        for (int l = 0; l < D.n; ++l) {
          const double c = C.a[k], d = D.a[l], q = c + d;
          double Q[3], cd2 = 0, pq2 = 0;
          for (int x = 0; x < 3; ++x) {
            Q[x] = (c * C.c[x] + d * D.c[x]) / q;
            cd2 += (C.c[x] - D.c[x]) * (C.c[x] - D.c[x]);
            pq2 += (P[x] - Q[x]) * (P[x] - Q[x]);
          }
          const double Kcd = C.w[k] * D.w[l] * std::exp(-c * d / q * cd2);
          const double rho = p * q / (p + q);
          const double pref = two_pi_52 / (p * q * std::sqrt(p + q)) * Kab * Kcd;
          double F[G::L + 1];
          boys_host(G::L, rho * pq2, pref, F);
          const double rp = rho / p, rq = rho / q;
          const double PA[3] = {P[0] - A.c[0], P[1] - A.c[1], P[2] - A.c[2]};
          const double QC[3] = {Q[0] - C.c[0], Q[1] - C.c[1], Q[2] - C.c[2]};
          double WP[3], WQ[3];
          for (int x = 0; x < 3; ++x) {
            WP[x] = -rp * (P[x] - Q[x]);
            WQ[x] = rq * (P[x] - Q[x]);
          }
          for (int t = 0; t < G::NT; ++t) ok[t] = false;
          T[0] = 0.0;
          ok[0] = true;
          for (int m = 0; m <= G::L; ++m) {
            T[1 + m] = F[m];
            ok[1 + m] = true;
          }
          const int4* vp = G::vrr();
          const int* vl = G::vlev();
          for (int lev = 1; lev <= G::L; ++lev) {
            // entries of one level read only earlier levels
            for (int t = vl[lev]; t < vl[lev + 1]; ++t) {
              const unsigned x = (unsigned)vp[t].x, y = (unsigned)vp[t].y, z = (unsigned)vp[t].z;
              const unsigned s[5] = {y & 0xffffu, y >> 16, z & 0xffffu, z >> 16, (unsigned)vp[t].w};
              for (unsigned v : s)
                if (v >= (unsigned)G::NT || !ok[v]) return -2;
            }
            for (int t = vl[lev]; t < vl[lev + 1]; ++t) {
              const unsigned x = (unsigned)vp[t].x, y = (unsigned)vp[t].y, z = (unsigned)vp[t].z;
              const bool ket = (x >> 16) & 1u;
              const unsigned ax = (x >> 17) & 3u;
              const double k1 = (double)((x >> 20) & 15u), k2 = (double)((x >> 24) & 15u);
              const double cA = ket ? QC[ax] : PA[ax], cB = ket ? WQ[ax] : WP[ax];
              const double o = ket ? 0.5 / q : 0.5 / p, r = ket ? rq : rp;
              const unsigned dst = x & 0xffffu;
              if (dst >= (unsigned)G::NT || ok[dst]) return -3;
              T[dst] = cA * T[y & 0xffffu] + cB * T[y >> 16] + k1 * o * (T[z & 0xffffu] - r * T[z >> 16]) +
                       k2 * (0.5 / (p + q)) * T[vp[t].w];
            }
            for (int t = vl[lev]; t < vl[lev + 1]; ++t) ok[(unsigned)vp[t].x & 0xffffu] = true;
          }
          for (int t = 0; t < G::NACC; ++t) {
            if (!ok[G::acc()[t]]) return -4;
            H[t] += T[G::acc()[t]];
          }
        }
    }
  if (!run_hrr<G>(A, B, C, D, H, hok)) return -5;
  for (int t = 0; t < G::NOUT; ++t) {
    if (!hok[G::out()[t]]) return -7;
    out[t] = H[G::out()[t]];
  }
  return 0;
}

// HRR levels as the lanes run them: every source written before the level,
// no slot both read and written within a level, distinct destinations.
template <class G>
static bool run_hrr(const Shell& A, const Shell& B, const Shell& C, const Shell& D, double* H, bool* hok) {
  for (int t = 0; t < G::NH; ++t) hok[t] = t < G::NACC + zero_slots<G>();
  const double AB[3] = {A.c[0] - B.c[0], A.c[1] - B.c[1], A.c[2] - B.c[2]};
  const double CD[3] = {C.c[0] - D.c[0], C.c[1] - D.c[1], C.c[2] - D.c[2]};
  const int2* hp = G::hrr();
  const int* hl = G::hlev();
  static double tmp[1 << 16];
  static int mark[1 << 16];
  static int stamp = 0;
  for (int lev = 0; lev < G::NHLEV; ++lev) {
    ++stamp;
    for (int t = hl[lev]; t < hl[lev + 1]; ++t) {
      const unsigned y = (unsigned)hp[t].y;
      const unsigned s0 = y & 0xffffu, s1 = y >> 16;
      if (s0 >= (unsigned)G::NH || s1 >= (unsigned)G::NH || !hok[s0] || !hok[s1]) return false;
      mark[s0] = mark[s1] = stamp;
    }
    for (int t = hl[lev]; t < hl[lev + 1]; ++t) {
      const unsigned d = (unsigned)hp[t].x & 0xffffu;
      if (d >= (unsigned)G::NH || mark[d] == stamp || mark[d] == -stamp) return false;
      mark[d] = -stamp;
    }
    for (int t = hl[lev]; t < hl[lev + 1]; ++t) {
      const unsigned x = (unsigned)hp[t].x, y = (unsigned)hp[t].y;
      const unsigned ax = (x >> 17) & 3u;
      const double X = ((x >> 16) & 1u) ? CD[ax] : AB[ax];
      tmp[t - hl[lev]] = H[y & 0xffffu] + X * H[y >> 16];
    }
    for (int t = hl[lev]; t < hl[lev + 1]; ++t) {
      const unsigned d = (unsigned)hp[t].x & 0xffffu;
      H[d] = tmp[t - hl[lev]];
      hok[d] = true;
    }
  }
  return true;
}

static void unpack(const int* n, const double* ctr, const double* a, const double* w, Shell* s) {
  int off = 0;
  for (int k = 0; k < 4; ++k) {
    s[k].c[0] = ctr[3 * k];
    s[k].c[1] = ctr[3 * k + 1];
    s[k].c[2] = ctr[3 * k + 2];
    s[k].n = n[k];
    s[k].a = a + off;
    s[k].w = w + off;
    off += n[k];
  }
}

extern "C" int host_quartet_grad(const int* l, const double* ctr, const int* n, const double* a,
                                 const double* w, double* out) {
  Shell s[4];
  unpack(n, ctr, a, w, s);
#define GX_CASE(A, B, C, D) \
  if (l[0] == A && l[1] == B && l[2] == C && l[3] == D) return quartet_grad<A, B, C, D>(s[0], s[1], s[2], s[3], out);
  GX_FOR_EACH_CLASS(GX_CASE)
#undef GX_CASE
  return -1;
}

// Per-lane gradient code (EriGradLane): g[3 X + i] of the quartet contracted
// with D_k = f (4 P_ab P_cd - P_ac P_bd - P_ad P_bc); blocks in `pb` in the
// order P_ab, P_cd, P_ac, P_ad, P_bc, P_bd.
template <int LA, int LB, int LC, int LD, int X>
static void quartet_grad_lane_pass(const Shell& A, const Shell& B, const Shell& C, const Shell& D,
                                   const double* pb, double f, double (&g)[9]) {
  using G = EriGradLanePass<LA, LB, LC, LD, X>;
  using Q = EriGradLane<LA, LB, LC, LD>;
  double acc[G::NACC] = {0};
  const double two_pi_52 = 34.98683665524972;
  for (int i = 0; i < A.n; ++i)
    for (int j = 0; j < B.n; ++j) {
      const double a = A.a[i], b = B.a[j], p = a + b;
      double P[3], ab2 = 0;
      for (int x = 0; x < 3; ++x) {
        P[x] = (a * A.c[x] + b * B.c[x]) / p;
        ab2 += (A.c[x] - B.c[x]) * (A.c[x] - B.c[x]);
      }
      const double Kab = A.w[i] * B.w[j] * std::exp(-a * b / p * ab2);
      for (int k = 0; k < C.n; ++k)
        for (int l = 0; l < D.n; ++l) {
          const double c = C.a[k], d = D.a[l], q = c + d;
          double Q[3], cd2 = 0, pq2 = 0;
          for (int x = 0; x < 3; ++x) {
            Q[x] = (c * C.c[x] + d * D.c[x]) / q;
            cd2 += (C.c[x] - D.c[x]) * (C.c[x] - D.c[x]);
            pq2 += (P[x] - Q[x]) * (P[x] - Q[x]);
          }
          const double Kcd = C.w[k] * D.w[l] * std::exp(-c * d / q * cd2);
          const double rho = p * q / (p + q);
          const double pref = two_pi_52 / (p * q * std::sqrt(p + q)) * Kab * Kcd;
          double F[G::L + 1];
          boys_host(G::L, rho * pq2, pref, F);
          const double rp = rho / p, rq = rho / q;
          double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
          G::template prim<double>(acc, F, P[0] - A.c[0], P[1] - A.c[1], P[2] - A.c[2], -rp * PQ[0], -rp * PQ[1],
                                   -rp * PQ[2], Q[0] - C.c[0], Q[1] - C.c[1], Q[2] - C.c[2], rq * PQ[0],
                                   rq * PQ[1], rq * PQ[2], 0.5 / p, 0.5 / q, 0.5 / (p + q), rp, rq, 2.0 * a,
                                   2.0 * b, 2.0 * c);
        }
    }
  const int nab = Q::NA * Q::NB, ncd = Q::NC * Q::ND, nac = Q::NA * Q::NC, nad = Q::NA * Q::ND,
            nbc = Q::NB * Q::NC;
  const double* Pab = pb;
  const double* Pcd = Pab + nab;
  const double* Pac = Pcd + ncd;
  const double* Pad = Pac + nac;
  const double* Pbc = Pad + nad;
  const double* Pbd = Pbc + nbc;
  G::template finish<double>(acc, A.c[0] - B.c[0], A.c[1] - B.c[1], A.c[2] - B.c[2], C.c[0] - D.c[0],
                             C.c[1] - D.c[1], C.c[2] - D.c[2], Pab, Pcd, Pac, Pad, Pbc, Pbd, f, g);
}

// The per-lane gradient code (one pass, or one per centre) contracted with the
// density blocks pb in the order P_ab, P_cd, P_ac, P_ad, P_bc, P_bd.
template <int LA, int LB, int LC, int LD>
static int quartet_grad_lane(const Shell& A, const Shell& B, const Shell& C, const Shell& D, const double* pb,
                             double f, double* out) {
  double g[9] = {0};
  quartet_grad_lane_pass<LA, LB, LC, LD, 0>(A, B, C, D, pb, f, g);
  if constexpr (EriGradLane<LA, LB, LC, LD>::NPASS == 3) {
    quartet_grad_lane_pass<LA, LB, LC, LD, 1>(A, B, C, D, pb, f, g);
    quartet_grad_lane_pass<LA, LB, LC, LD, 2>(A, B, C, D, pb, f, g);
  }
  for (int j = 0; j < 9; ++j) out[j] = g[j];
  return 0;
}

extern "C" int host_grad_lane(const int* l, const double* ctr, const int* n, const double* a, const double* w,
                              const double* pb, double f, double* out) {
  Shell s[4];
  unpack(n, ctr, a, w, s);
#define GX_CASE(A, B, C, D) \
  if (l[0] == A && l[1] == B && l[2] == C && l[3] == D) return quartet_grad_lane<A, B, C, D>(s[0], s[1], s[2], s[3], pb, f, out);
  GX_FOR_EACH_GRAD_LANE_CLASS(GX_CASE)
#undef GX_CASE
  return -1;
}

// Every class through its cooperative program (returns < 0 on a malformed program).
extern "C" int host_quartet_prog(const int* l, const double* ctr, const int* n, const double* a,
                                 const double* w, double* out) {
  Shell s[4];
  unpack(n, ctr, a, w, s);
#define GX_CASE(A, B, C, D) \
  if (l[0] == A && l[1] == B && l[2] == C && l[3] == D) return quartet_prog<A, B, C, D>(s[0], s[1], s[2], s[3], out);
  GX_FOR_EACH_CLASS(GX_CASE)
#undef GX_CASE
  return -1;
}

// Every class through the path the device takes: per-lane class code or the
// cooperative program.
extern "C" int host_quartet(const int* l, const double* ctr, const int* n, const double* a,
                            const double* w, double* out) {
  Shell s[4];
  unpack(n, ctr, a, w, s);
#define GX_CASE(A, B, C, D)                                                         \
  if (l[0] == A && l[1] == B && l[2] == C && l[3] == D) {                           \
    quartet<A, B, C, D>(s[0], s[1], s[2], s[3], out);                               \
    return 0;                                                                       \
  }
  GX_FOR_EACH_LANE_CLASS(GX_CASE)
#undef GX_CASE
#define GX_CASE(A, B, C, D) \
  if (l[0] == A && l[1] == B && l[2] == C && l[3] == D) return quartet_prog<A, B, C, D>(s[0], s[1], s[2], s[3], out);
  GX_FOR_EACH_COOP_CLASS(GX_CASE)
#undef GX_CASE
  return -1;
}

// The Rys-quadrature path: per primitive quartet and root, the 2D integrals
// of each axis (RysClass::g2d) and their products weighted by the rule of
// gx_rys.h.
template <int LA, int LB, int LC, int LD>
static int quartet_rys(const Shell& A, const Shell& B, const Shell& C, const Shell& D, double* out) {
  using R = RysClass<LA, LB, LC, LD>;
  constexpr int N = R::NROOT;
  static double acc[R::NOUT];
  for (int t = 0; t < R::NOUT; ++t) acc[t] = 0.0;
  const double two_pi_52 = 34.98683665524972;
  double AB[3], CD[3];
  for (int x = 0; x < 3; ++x) {
    AB[x] = A.c[x] - B.c[x];
    CD[x] = C.c[x] - D.c[x];
  }
  for (int i = 0; i < A.n; ++i)
    for (int j = 0; j < B.n; ++j) {
      const double a = A.a[i], b = B.a[j], p = a + b;
      double P[3], ab2 = 0;
      for (int x = 0; x < 3; ++x) {
        P[x] = (a * A.c[x] + b * B.c[x]) / p;
        ab2 += AB[x] * AB[x];
      }
      const double Kab = A.w[i] * B.w[j] * std::exp(-a * b / p * ab2);
      for (int k = 0; k < C.n; ++k)
        for (int l = 0; l < D.n; ++l) {
          const double c = C.a[k], d = D.a[l], q = c + d;
          double Q[3], cd2 = 0, pq2 = 0;
          for (int x = 0; x < 3; ++x) {
            Q[x] = (c * C.c[x] + d * D.c[x]) / q;
            cd2 += CD[x] * CD[x];
            pq2 += (P[x] - Q[x]) * (P[x] - Q[x]);
          }
          const double Kcd = C.w[k] * D.w[l] * std::exp(-c * d / q * cd2);
          const double rho = p * q / (p + q), rp = rho / p, rq = rho / q;
          const double pref = two_pi_52 / (p * q * std::sqrt(p + q)) * Kab * Kcd;
          const double T = rho * pq2;
          double G[3][R::NG];
          for (int r = 0; r < N; ++r) {
            const double u = rys_quantity(N, T, r), wr = rys_quantity(N, T, N + r) * pref;
            for (int x = 0; x < 3; ++x) {
              const double PQ = P[x] - Q[x];
              R::g2d(P[x] - A.c[x] - rp * PQ * u, Q[x] - C.c[x] + rq * PQ * u, 0.5 / p * (1.0 - rp * u),
                     0.5 / q * (1.0 - rq * u), 0.5 * u / (p + q), AB[x], CD[x], G[x]);
            }
            for (int t = 0; t < R::NOUT; ++t) {
              const int o = R::off()[t];
              acc[t] += wr * G[0][o & 1023] * G[1][(o >> 10) & 1023] * G[2][o >> 20];
            }
          }
        }
    }
  for (int t = 0; t < R::NOUT; ++t) out[t] = acc[t];
  return 0;
}

extern "C" int host_quartet_rys(const int* l, const double* ctr, const int* n, const double* a,
                                const double* w, double* out) {
  Shell s[4];
  unpack(n, ctr, a, w, s);
#define GX_CASE(A, B, C, D) \
  if (l[0] == A && l[1] == B && l[2] == C && l[3] == D) return quartet_rys<A, B, C, D>(s[0], s[1], s[2], s[3], out);
  GX_FOR_EACH_RYS_CLASS(GX_CASE)
#undef GX_CASE
  return -1;
}

// The Rys gradient path: per primitive quartet and root (n at L + 1), the
// 2D integrals and their A, B, C derivatives (RysClass::g2d_grad), in the
// layout of quartet_grad (component-major, then centre A, B, C, then axis).
template <int LA, int LB, int LC, int LD>
static int quartet_grad_rys(const Shell& A, const Shell& B, const Shell& C, const Shell& D, double* out) {
  using R = RysClass<LA, LB, LC, LD>;
  constexpr int N = R::NROOT_GRAD, NG = R::NG;
  static double acc[9 * R::NOUT];
  for (int t = 0; t < 9 * R::NOUT; ++t) acc[t] = 0.0;
  const double two_pi_52 = 34.98683665524972;
  double AB[3], CD[3];
  for (int x = 0; x < 3; ++x) {
    AB[x] = A.c[x] - B.c[x];
    CD[x] = C.c[x] - D.c[x];
  }
  for (int i = 0; i < A.n; ++i)
    for (int j = 0; j < B.n; ++j) {
      const double a = A.a[i], b = B.a[j], p = a + b;
      double P[3], ab2 = 0;
      for (int x = 0; x < 3; ++x) {
        P[x] = (a * A.c[x] + b * B.c[x]) / p;
        ab2 += AB[x] * AB[x];
      }
      const double Kab = A.w[i] * B.w[j] * std::exp(-a * b / p * ab2);
      for (int k = 0; k < C.n; ++k)
        for (int l = 0; l < D.n; ++l) {
          const double c = C.a[k], d = D.a[l], q = c + d;
          double Q[3], cd2 = 0, pq2 = 0;
          for (int x = 0; x < 3; ++x) {
            Q[x] = (c * C.c[x] + d * D.c[x]) / q;
            cd2 += CD[x] * CD[x];
            pq2 += (P[x] - Q[x]) * (P[x] - Q[x]);
          }
          const double Kcd = C.w[k] * D.w[l] * std::exp(-c * d / q * cd2);
          const double rho = p * q / (p + q), rp = rho / p, rq = rho / q;
          const double pref = two_pi_52 / (p * q * std::sqrt(p + q)) * Kab * Kcd;
          const double T = rho * pq2;
          double G[3][NG], GA[3][NG], GB[3][NG], GC[3][NG];
          for (int r = 0; r < N; ++r) {
            const double u = rys_quantity(N, T, r), wr = rys_quantity(N, T, N + r) * pref;
            for (int x = 0; x < 3; ++x) {
              const double PQ = P[x] - Q[x];
              R::g2d_grad(P[x] - A.c[x] - rp * PQ * u, Q[x] - C.c[x] + rq * PQ * u, 0.5 / p * (1.0 - rp * u),
                          0.5 / q * (1.0 - rq * u), 0.5 * u / (p + q), AB[x], CD[x], a, b, c, G[x], GA[x], GB[x],
                          GC[x]);
            }
            for (int t = 0; t < R::NOUT; ++t) {
              const int o = R::off()[t];
              const int ox = o & 1023, oy = (o >> 10) & 1023, oz = o >> 20;
              const double* Dv[3][3] = {{GA[0], GA[1], GA[2]}, {GB[0], GB[1], GB[2]}, {GC[0], GC[1], GC[2]}};
              for (int X = 0; X < 3; ++X) {
                acc[9 * t + 3 * X + 0] += wr * Dv[X][0][ox] * G[1][oy] * G[2][oz];
                acc[9 * t + 3 * X + 1] += wr * G[0][ox] * Dv[X][1][oy] * G[2][oz];
                acc[9 * t + 3 * X + 2] += wr * G[0][ox] * G[1][oy] * Dv[X][2][oz];
              }
            }
          }
        }
    }
  for (int t = 0; t < 9 * R::NOUT; ++t) out[t] = acc[t];
  return 0;
}

extern "C" int host_quartet_grad_rys(const int* l, const double* ctr, const int* n, const double* a,
                                     const double* w, double* out) {
  Shell s[4];
  unpack(n, ctr, a, w, s);
#define GX_CASE(A, B, C, D)                                                                              \
  if constexpr ((A + B + C + D + 1) / 2 + 1 <= RYS_NMAX)                                                 \
    if (l[0] == A && l[1] == B && l[2] == C && l[3] == D) return quartet_grad_rys<A, B, C, D>(s[0], s[1], s[2], s[3], out);
  GX_FOR_EACH_RYS_CLASS(GX_CASE)
#undef GX_CASE
  return -1;
}
