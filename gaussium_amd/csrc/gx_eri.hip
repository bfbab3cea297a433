// ERI kernels for gfx950: primitive-pair build, Schwarz bounds, dense and
// quartet-list ERIs, and the fused direct J/K build.
//
// Reference operators replaced (src/gaussium/ of the reference):
//   ObaraSaika.integrate          integrals/twoelectronrepulsion/obara_saika_scheme.py:17-82
//   TwoElectronRepulsion.*        matrixelements/electron_repulsion_matrix.py:18-47
//   FockMatrixRestricted.create   hartreefock/fock_matrix.py:12-22
//
// One angular-momentum class (la lb|lc ld) per template instance; the
// recursion code comes from gen_eri.py (EriClass<...>).
#include "gx_common.h"
#include "_gen/eri_gen.h"

namespace gx {

// ---------------------------------------------------------------------------
// Primitive-pair records.  One thread per primitive pair; `pair_of` maps the
// primitive pair to its shell pair, (ia, ib) are the primitive indices.
__global__ void k_primpairs(int npp, const int* __restrict__ pair_of, const int* __restrict__ pia,
                            const int* __restrict__ pib, const int* __restrict__ pairA,
                            const int* __restrict__ pairB, const int* __restrict__ sh_poff,
                            const double* __restrict__ ctr, const double* __restrict__ exps,
                            const double* __restrict__ coefs, PrimPair64* __restrict__ out64,
                            PrimPair32* __restrict__ out32) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npp) return;
  const int ip = pair_of[t];
  const int A = pairA[ip], B = pairB[ip];
  const int ga = sh_poff[A] + pia[t], gb = sh_poff[B] + pib[t];
  const double a = exps[ga], b = exps[gb];
  const double p = a + b, ip_ = 1.0 / p;
  const double Ax = ctr[3 * A], Ay = ctr[3 * A + 1], Az = ctr[3 * A + 2];
  const double Bx = ctr[3 * B], By = ctr[3 * B + 1], Bz = ctr[3 * B + 2];
  const double dx = Ax - Bx, dy = Ay - By, dz = Az - Bz;
  const double r2 = dx * dx + dy * dy + dz * dz;
  // 2 pi^(5/2) = 34.98683665524972 split evenly over bra and ket: sqrt = 5.914967172795612
  const double K = 5.914967172795612 * coefs[ga] * coefs[gb] * exp(-a * b * ip_ * r2);
  PrimPair64 r;
  r.p = p;
  r.ip = ip_;
  r.Px = (a * Ax + b * Bx) * ip_;
  r.Py = (a * Ay + b * By) * ip_;
  r.Pz = (a * Az + b * Bz) * ip_;
  r.Kip = K * ip_;
  r.PAx = r.Px - Ax;
  r.PAy = r.Py - Ay;
  r.PAz = r.Pz - Az;
  r.ea = a;
  out64[t] = r;
  PrimPair32 s;
  s.p = (float)p;
  s.ip = (float)ip_;
  s.Px = (float)r.Px;
  s.Py = (float)r.Py;
  s.Pz = (float)r.Pz;
  s.Kip = (float)r.Kip;
  s.PAx = (float)r.PAx;
  s.PAy = (float)r.PAy;
  s.PAz = (float)r.PAz;
  s.ea = (float)a;
  s.pad1 = s.pad2 = 0.0f;
  out32[t] = s;
}

// ---------------------------------------------------------------------------
// Contracted shell quartet (AB|CD), all Cartesian components, unscaled.
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void quartet64(const ShellView& sh, const PairView& pr, int ip, int jp,
                                          const double* __restrict__ tab,
                                          double (&out)[EriClass<LA, LB, LC, LD>::NOUT]) {
  using C = EriClass<LA, LB, LC, LD>;
  const int A = pr.A[ip], B = pr.B[ip], Cs = pr.A[jp], Ds = pr.B[jp];
  const double ABx = sh.ctr[3 * A] - sh.ctr[3 * B], ABy = sh.ctr[3 * A + 1] - sh.ctr[3 * B + 1],
               ABz = sh.ctr[3 * A + 2] - sh.ctr[3 * B + 2];
  const double CDx = sh.ctr[3 * Cs] - sh.ctr[3 * Ds], CDy = sh.ctr[3 * Cs + 1] - sh.ctr[3 * Ds + 1],
               CDz = sh.ctr[3 * Cs + 2] - sh.ctr[3 * Ds + 2];
  const PrimPair64* __restrict__ bp = pr.pp64 + pr.poff[ip];
  const PrimPair64* __restrict__ kp = pr.pp64 + pr.poff[jp];
  const int nb = pr.np[ip], nk = pr.np[jp];
  double acc[C::NACC];
#pragma unroll
  for (int i = 0; i < C::NACC; ++i) acc[i] = 0.0;
  for (int ib = 0; ib < nb; ++ib) {
    const PrimPair64 b = bp[ib];
    for (int ik = 0; ik < nk; ++ik) {
      const PrimPair64 k = kp[ik];
      const double rs = gx_rsqrt(b.p + k.p);
      const double is = rs * rs;
      const double rho = b.p * k.p * is;
      const double PQx = b.Px - k.Px, PQy = b.Py - k.Py, PQz = b.Pz - k.Pz;
      const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
      const double pref = b.Kip * k.Kip * rs;
      double F[C::L + 1];
      boys64<C::L>(T, pref, tab, F);
      const double rho_p = rho * b.ip, rho_q = rho * k.ip;
      C::template prim<double>(acc, F, b.PAx, b.PAy, b.PAz, -rho_p * PQx, -rho_p * PQy, -rho_p * PQz,
                               k.PAx, k.PAy, k.PAz, rho_q * PQx, rho_q * PQy, rho_q * PQz, 0.5 * b.ip,
                               0.5 * k.ip, 0.5 * is, rho_p, rho_q);
    }
  }
  C::template hrr<double>(acc, ABx, ABy, ABz, CDx, CDy, CDz, out);
}

template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void quartet32(const ShellView& sh, const PairView& pr, int ip, int jp,
                                          const float* __restrict__ tab,
                                          float (&out)[EriClass<LA, LB, LC, LD>::NOUT]) {
  using C = EriClass<LA, LB, LC, LD>;
  const int A = pr.A[ip], B = pr.B[ip], Cs = pr.A[jp], Ds = pr.B[jp];
  const float ABx = (float)(sh.ctr[3 * A] - sh.ctr[3 * B]);
  const float ABy = (float)(sh.ctr[3 * A + 1] - sh.ctr[3 * B + 1]);
  const float ABz = (float)(sh.ctr[3 * A + 2] - sh.ctr[3 * B + 2]);
  const float CDx = (float)(sh.ctr[3 * Cs] - sh.ctr[3 * Ds]);
  const float CDy = (float)(sh.ctr[3 * Cs + 1] - sh.ctr[3 * Ds + 1]);
  const float CDz = (float)(sh.ctr[3 * Cs + 2] - sh.ctr[3 * Ds + 2]);
  const PrimPair32* __restrict__ bp = pr.pp32 + pr.poff[ip];
  const PrimPair32* __restrict__ kp = pr.pp32 + pr.poff[jp];
  const int nb = pr.np[ip], nk = pr.np[jp];
  float acc[C::NACC];
#pragma unroll
  for (int i = 0; i < C::NACC; ++i) acc[i] = 0.0f;
  for (int ib = 0; ib < nb; ++ib) {
    const PrimPair32 b = bp[ib];
    for (int ik = 0; ik < nk; ++ik) {
      const PrimPair32 k = kp[ik];
      const float rs = gx_rsqrt(b.p + k.p);
      const float is = rs * rs;
      const float rho = b.p * k.p * is;
      const float PQx = b.Px - k.Px, PQy = b.Py - k.Py, PQz = b.Pz - k.Pz;
      const float T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
      const float pref = b.Kip * k.Kip * rs;
      float F[C::L + 1];
      boys32<C::L>(T, pref, tab, F);
      const float rho_p = rho * b.ip, rho_q = rho * k.ip;
      C::template prim<float>(acc, F, b.PAx, b.PAy, b.PAz, -rho_p * PQx, -rho_p * PQy, -rho_p * PQz,
                              k.PAx, k.PAy, k.PAz, rho_q * PQx, rho_q * PQy, rho_q * PQz, 0.5f * b.ip,
                              0.5f * k.ip, 0.5f * is, rho_p, rho_q);
    }
  }
  C::template hrr<float>(acc, ABx, ABy, ABz, CDx, CDy, CDz, out);
}

// ---------------------------------------------------------------------------
// J/K evaluators.  The bra shell pair is uniform across a wave, so its
// primitive pairs come through the scalar data path (kernel-argument
// pointers marked __restrict__ let the compiler use scalar loads); each lane
// loads its own ket primitive pairs once, in the outer loop.
// Bra primitive pairs are consumed in groups of JKG (FP64) / JKG32 (FP32):
// the group's scalar record loads and table gathers are issued together and
// the branch-free Boys evaluation keeps the group one basic block.
#ifndef GX_JKG
#define GX_JKG 2
#endif
#ifndef GX_JKG32
#define GX_JKG32 2
#endif
#ifndef GX_MIXED_LDS32
#define GX_MIXED_LDS32 1  // mixed builds: FP32 Taylor table in LDS (else read through L1)
#endif
#ifndef GX_DYN_CHUNK
#define GX_DYN_CHUNK 1
#endif
// FP64 group inside mixed builds: 1, 2 for (ss|ss) (measured on MI355X at
// the 220-function cluster: (ss|ss) 125 -> 117 us with groups of 2, (ps|ss)
// 92 -> 98 us).
template <int LA, int LB, int LC, int LD>
constexpr int jk_group_mixed64() {
  return LA + LB + LC + LD == 0 ? 2 : 1;
}

template <class C, typename R, typename PP, int G, class BG>
__device__ __forceinline__ void prim_group(R (&acc)[C::NACC], const PP* __restrict__ b, const PP& k,
                                           const R* __restrict__ tab) {
  R rs[G], rho[G], PQx[G], PQy[G], PQz[G], T[G], pref[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    rs[g] = gx_rsqrt(b[g].p + k.p);
    const R is = rs[g] * rs[g];
    rho[g] = b[g].p * k.p * is;
    PQx[g] = b[g].Px - k.Px;
    PQy[g] = b[g].Py - k.Py;
    PQz[g] = b[g].Pz - k.Pz;
    T[g] = rho[g] * (PQx[g] * PQx[g] + PQy[g] * PQy[g] + PQz[g] * PQz[g]);
    pref[g] = b[g].Kip * k.Kip * rs[g];
  }
  R F[G][C::L + 1];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if constexpr (sizeof(R) == 8)
      boys64_nb<C::L, BG>(T[g], pref[g], tab, F[g]);
    else
      boys32_nb<C::L, BG>(T[g], pref[g], tab, F[g]);
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const R is = rs[g] * rs[g];
    const R rho_p = rho[g] * b[g].ip, rho_q = rho[g] * k.ip;
    C::template prim<R>(acc, F[g], b[g].PAx, b[g].PAy, b[g].PAz, -rho_p * PQx[g], -rho_p * PQy[g],
                        -rho_p * PQz[g], k.PAx, k.PAy, k.PAz, rho_q * PQx[g], rho_q * PQy[g],
                        rho_q * PQz[g], R(0.5) * b[g].ip, R(0.5) * k.ip, R(0.5) * is, rho_p, rho_q);
  }
}

template <class C, typename R, typename PP, int G, class BG>
__device__ __forceinline__ void quartet_jk(const PP* __restrict__ bra, int nb, const PP* __restrict__ kp,
                                           int nk, R ABx, R ABy, R ABz, R CDx, R CDy, R CDz,
                                           const R* __restrict__ tab, R (&out)[C::NOUT]) {
  R acc[C::NACC];
#pragma unroll
  for (int i = 0; i < C::NACC; ++i) acc[i] = R(0);
  for (int ik = 0; ik < nk; ++ik) {
    const PP k = kp[ik];
    int ib = 0;
    for (; ib + G <= nb; ib += G) {
      PP b[G];
#pragma unroll
      for (int g = 0; g < G; ++g) b[g] = bra[ib + g];
      prim_group<C, R, PP, G, BG>(acc, b, k, tab);
    }
    for (; ib < nb; ++ib) {
      PP b[1] = {bra[ib]};
      prim_group<C, R, PP, 1, BG>(acc, b, k, tab);
    }
  }
  C::template hrr<R>(acc, ABx, ABy, ABz, CDx, CDy, CDz, out);
}

// The class code of one J/K pass: a class evaluated in one pass (every ket
// component, cidx the identity), or pass P of a class evaluated in passes
// over its ket components (EriJKPass, gen_eri.py).
template <class C>
struct JKOnePass : C {
  static __host__ __device__ constexpr int cidx(int i) { return i; }
};
template <class T>
struct JKPassTag {
  using type = T;
};
template <int LA, int LB, int LC, int LD, int P>
using JKPassClass = std::conditional_t<(EriJKLane<LA, LB, LC, LD>::NPASS > 1), EriJKPass<LA, LB, LC, LD, P>,
                                       JKOnePass<EriClass<LA, LB, LC, LD>>>;

// fn(JKPassTag<pass class>) for the passes P, P + 1, .. of the class.
template <int LA, int LB, int LC, int LD, int P = 0, class F>
__device__ __forceinline__ void jk_passes(F& fn) {
  if constexpr (P < EriJKLane<LA, LB, LC, LD>::NPASS) {
    fn(JKPassTag<JKPassClass<LA, LB, LC, LD, P>>{});
    jk_passes<LA, LB, LC, LD, P + 1>(fn);
  }
}

// ---------------------------------------------------------------------------
// Schwarz bounds: Q_AB = sqrt(max_ab |(ab|ab)| s_a^2 s_b^2), one thread per pair.
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_schwarz(ShellView sh, PairView pr, int p0, int np,
                                                const double* __restrict__ master,
                                                const double* __restrict__ scale,
                                                double* __restrict__ Q) {
  using C = EriClass<LA, LB, LA, LB>;
  __shared__ double tab[BOYS_NT * BOYS_ROW64];
  boys_stage64(C::L, master, tab);
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np) return;
  const int ip = p0 + t;
  double v[C::NOUT];
  quartet64<LA, LB, LA, LB>(sh, pr, ip, ip, tab, v);
  const int ao = sh.bf[pr.A[ip]], bo = sh.bf[pr.B[ip]];
  double m = 0.0;
#pragma unroll
  for (int ia = 0; ia < C::NA; ++ia)
#pragma unroll
    for (int ib = 0; ib < C::NB; ++ib) {
      const double s = scale[ao + ia] * scale[bo + ib];
      const double d = fabs(v[((ia * C::NB + ib) * C::NA + ia) * C::NB + ib]) * s * s;
      m = d > m ? d : m;
    }
  Q[ip] = sqrt(m);
}

// ---------------------------------------------------------------------------
// Dense ERI tensor for small N: one thread per (bra, ket) shell quartet,
// written to all eight permutational images.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_dense(ShellView sh, PairView pr, int bra0, int nbra,
                                              int ket0, int nket, int same,
                                              const double* __restrict__ master,
                                              const double* __restrict__ scale,
                                              double* __restrict__ out, long N) {
  using C = EriClass<LA, LB, LC, LD>;
  __shared__ double tab[BOYS_NT * BOYS_ROW64];
  boys_stage64(C::L, master, tab);
  __syncthreads();
  const int bi = blockIdx.y;
  const int kj = blockIdx.x * blockDim.x + threadIdx.x;
  if (bi >= nbra || kj >= nket || (same && kj > bi)) return;
  const int ip = bra0 + bi, jp = ket0 + kj;
  double v[C::NOUT];
  quartet64<LA, LB, LC, LD>(sh, pr, ip, jp, tab, v);
  const long ao = sh.bf[pr.A[ip]], bo = sh.bf[pr.B[ip]];
  const long co = sh.bf[pr.A[jp]], dof = sh.bf[pr.B[jp]];
  int k = 0;
  for (int ia = 0; ia < C::NA; ++ia)
    for (int ib = 0; ib < C::NB; ++ib)
      for (int ic = 0; ic < C::NC; ++ic)
        for (int id = 0; id < C::ND; ++id, ++k) {
          const long i = ao + ia, j = bo + ib, kk = co + ic, l = dof + id;
          const double x = v[k] * scale[i] * scale[j] * scale[kk] * scale[l];
          out[((i * N + j) * N + kk) * N + l] = x;
          out[((j * N + i) * N + kk) * N + l] = x;
          out[((i * N + j) * N + l) * N + kk] = x;
          out[((j * N + i) * N + l) * N + kk] = x;
          out[((kk * N + l) * N + i) * N + j] = x;
          out[((l * N + kk) * N + i) * N + j] = x;
          out[((kk * N + l) * N + j) * N + i] = x;
          out[((l * N + kk) * N + j) * N + i] = x;
        }
}

// ---------------------------------------------------------------------------
// Quartet list: task t evaluates shell quartet (bra[t]|ket[t]) and stores
// component comp[t] (scaled by sc[t]) at out[pos[t]].
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_quartets(ShellView sh, PairView pr, int ntask,
                                                 const int* __restrict__ bra,
                                                 const int* __restrict__ ket,
                                                 const int* __restrict__ comp,
                                                 const double* __restrict__ sc,
                                                 const long* __restrict__ pos,
                                                 const double* __restrict__ master,
                                                 double* __restrict__ out) {
  using C = EriClass<LA, LB, LC, LD>;
  __shared__ double tab[BOYS_NT * BOYS_ROW64];
  boys_stage64(C::L, master, tab);
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntask) return;
  double v[C::NOUT];
  quartet64<LA, LB, LC, LD>(sh, pr, bra[t], ket[t], tab, v);
  const int c = comp[t];
  double x = 0.0;
#pragma unroll
  for (int k = 0; k < C::NOUT; ++k) x = (k == c) ? v[k] : x;
  out[pos[t]] = x * sc[t];
}

// ---------------------------------------------------------------------------
// Fused direct J/K.
//
// Persistent workgroups take work items of one class combination (in
// decreasing-work order); all waves of a workgroup share one bra at a time.
// For each bra the ket class is split into segments of equal primitive-pair
// count, each sorted by decreasing Schwarz bound, so the kets surviving the
// cutoff -- and, in mixed mode, those that stay in FP64 -- are prefixes of
// the segments.  The host writes each bra's surviving kets as one list (FP64
// kets, then FP32 kets from the next multiple of 64); an item is a range of
// that list, which the waves sweep 64 at a time (one ket pair per lane).  The
// density rows of the bra's shells sit in LDS.  Digestion:
//   Jt_ab += 2 f v P_cd      (register sums, wave-reduced, one atomic per element)
//   J_cd  += 2 f v P_ab      (global FP64 atomics into the pair-indexed Jkp)
//   Kt_ac += f v P_bd, Kt_ad += f v P_bc, Kt_bc += f v P_ad, Kt_bd += f v P_ac
//                            (LDS rows of A and B, flushed once per item: the
//                            whole wave would otherwise hit the same two rows)
// with f the 8-fold degeneracy weight; J = Jt + Jt^T, K = Kt + Kt^T.
// JK_MAXSEG bounds the segments per (class, system) (host bookkeeping).
#ifndef JK_MAXSEG
#define JK_MAXSEG 16
#endif
#ifndef JK_TEAM
#define JK_TEAM 256  // threads working on one work item
#endif

// Boys grid of the per-lane J/K kernel of class (la lb|lc ld): the coarse
// grid (half the LDS of the fine one, two more terms per Taylor series) for
// the classes with at least six density rows in LDS and at most 54
// components, whose occupancy LDS rather than registers bounds (measured on
// MI355X: (ds|ds) 107 -> 82 us at 255 functions; the classes with fewer rows
// or more components run slower on it).
__host__ __device__ constexpr int jk_ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr bool jk_coarse_grid(int la, int lb, int lc, int ld) {
  return jk_ncart(la) + jk_ncart(lb) >= 6 && jk_ncart(la) * jk_ncart(lb) * jk_ncart(lc) * jk_ncart(ld) <= 54;
}

struct JKArgs {
  ShellView sh;
  PairView pr;
  SysView sy;
  int bra0;
  const int4* items;     // work items: bra position, ket-list range [v0, v1), p0 | p1 << 16; largest first
  int nitem;
  const int* kets;       // ket lists of the bras (GX_KET_* encoding), indexed by the items' [v0, v1)
  const double* tab64;   // per-L FP64 Taylor tables, fine grid (BoysFine::NT x ROW64 each)
  const float* tab32;    // per-L FP32 Taylor tables, fine grid
  const double* tab64c;  // the same, coarse grid (BoysCoarse)
  const float* tab32c;
  const double* P;       // packed densities, pre-scaled by s_k s_l (the K density)
  const double* P2;      // second K density (two-density builds), pre-scaled
  const double* Pj;      // the J density (P, or P + P2 in two-density builds), pre-scaled
  double* Jt;            // packed accumulator of the J_ab parts (one copy)
  double* Jkp;           // pair-indexed J_cd accumulator: GX_JT_COPIES copies of jstride doubles
  double* Kt;
  double* Kt2;           // K accumulator of the second density (single copy)
  int* item_ctr;         // [next item, workgroups done]: zero before a launch, reset by its last workgroup
  long jstride;          // doubles per copy of Jkp (sum over pairs of n_A n_B)
  int jkcs;              // Jkp stride between the components of a ket pair (pairs of the ket class)
  int mixed;
};

// J_cd = sum_ab (ab|cd) P_ab arrives from every bra at once, one value per
// lane and ket pair.  It is accumulated per ket pair (component k of pair p
// at jko + k * jkcs, jkcs = pairs of the ket class) rather than at (c, d) of
// the matrix: the kets of a chunk are mostly consecutive pairs, so a wave's
// atomics hit consecutive addresses (a few cache lines) instead of 64
// scattered ones.  Workgroups are dispatched
// to the eight XCDs round-robin, so copy blockIdx % GX_JT_COPIES collects the
// atomics of (mostly) one XCD: the same address is then not contended from
// every XCD at once.  k_symmetrize adds the copies into J.
__device__ __forceinline__ double* jkp_copy(const JKArgs& a) {
  return a.Jkp + (long)(blockIdx.x % GX_JT_COPIES) * a.jstride;
}

// MIXED = false: pure FP64 build, the FP32 path is compiled out and the FP64
// Taylor table is staged in LDS.  MIXED = true: the quartets below the FP32
// bound are FP32, the others FP64; both tables are staged in LDS, on the
// coarse grid (13.9 + 4.6 KB for every class).  Measured on MI355X at the
// 220-function cluster: with the FP64 table read through L1 instead, the
// FP64 quartets of a mixed (ss|ss) launch made it slower than the all-FP64
// kernel (211 vs 129 us) and the autotuner kept it in FP64; staged, (ss|ss)
// runs in 125 us and the mixed build takes 0.379 ms (0.430 before).
// NDEN = 2 (FP64 only): one evaluation of the integrals serves J of
// P + P2 and K of P and of P2 (unrestricted Fock builds).
// Classes whose per-lane J/K kernel is held to 256 VGPRs (two waves per
// SIMD): they need a few more without the bound and run one wave per SIMD
// then (measured on MI355X at the 255-function cluster: (ds|ds) 105 -> 81 us,
// (ds|pp) 65 -> 57 us; (dp|ps) and (dd|ss) were slower bounded).
template <int LA, int LB, int LC, int LD>
constexpr int jk_min_waves() {
  return ((LA == 2 && LB == 0 && LC == 2 && LD == 0) || (LA == 2 && LB == 0 && LC == 1 && LD == 1)) ? 2 : 1;
}

template <int LA, int LB, int LC, int LD, bool MIXED, int NDEN>
__global__ __launch_bounds__(JK_TEAM) __attribute__((amdgpu_waves_per_eu((jk_min_waves<LA, LB, LC, LD>())))) void k_jk(JKArgs a, const PrimPair64* __restrict__ pp64,
                                                const PrimPair32* __restrict__ pp32) {
  using C = EriClass<LA, LB, LC, LD>;
  constexpr int NA = C::NA, NB = C::NB, NC = C::NC, ND = C::ND;
  extern __shared__ double smem[];
  __shared__ double jab_s[NA * NB];
  __shared__ int next_chunk;
  const int tid = threadIdx.x;
  // dynamic LDS: Boys FP64 (pure FP64 build) | P rows | K rows (rows of the
  // item's system)
  constexpr bool T64 = true;                      // FP64 table staged in LDS
  constexpr bool T32 = MIXED && GX_MIXED_LDS32;   // FP32 table staged in LDS
  constexpr bool COARSE = jk_coarse_grid(LA, LB, LC, LD) || MIXED;
  using BG = std::conditional_t<COARSE, BoysCoarse, BoysFine>;
  constexpr int off32 = T64 ? BG::NT * BG::ROW64 : 0;              // in doubles
  constexpr int offrow = off32 + (T32 ? BG::NT * BG::ROW32 / 2 : 0);
  const double* __restrict__ gtab64 =
      (COARSE ? a.tab64c : a.tab64) + C::L * BG::NT * BG::ROW64;
  const float* __restrict__ gtab32 =
      (COARSE ? a.tab32c : a.tab32) + C::L * BG::NT * BG::ROW32;
  const double* __restrict__ tab64 = T64 ? smem : gtab64;
  const float* __restrict__ tab32 = T32 ? reinterpret_cast<const float*>(smem + off32) : gtab32;
  double* prow = smem + offrow;
  if (T64) boys_copy64<BG>(gtab64, smem);
  if (T32) boys_copy32<BG>(gtab32, reinterpret_cast<float*>(smem + off32));

  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int wave = tid / GX_WAVE;
  constexpr int nwave = JK_TEAM / GX_WAVE;

  // Items are sorted by decreasing cost.  A workgroup starts on item
  // blockIdx.x and then claims the next unclaimed one from a counter, so the
  // cheap items at the end even out the workgroups' finishing times.
  __shared__ int next_item;
  for (int w = blockIdx.x; w < a.nitem;) {
    const int4 it = a.items[w];
    const int ip = a.bra0 + it.x;
    const PairRec br = a.pr.rec[ip];
    const int A = br.A, B = br.B;
    const int ao = br.fa, bo = br.fb;
    const int rowB = (A == B) ? 0 : NA;
    const int sys = br.sys;
    const int N = a.sy.n[sys];
    const double* __restrict__ Pm = a.P + a.sy.mb[sys];
    double* __restrict__ Jm = a.Jt + a.sy.mb[sys];
    double* __restrict__ Jkp = jkp_copy(a);
    double* __restrict__ Km = a.Kt + a.sy.mb[sys];
    double* __restrict__ Km2 = (NDEN == 2 ? a.Kt2 : a.Kt) + a.sy.mb[sys];
    const double* __restrict__ Pm2 = (NDEN == 2 ? a.P2 : a.P) + a.sy.mb[sys];
    const double* __restrict__ Pjm = a.Pj + a.sy.mb[sys];
    const int nrk = (NA + NB) * N;  // one density's rows
    double* krow = prow + NDEN * nrk;
    // wave-uniform: kept in scalar registers so that the bra loop reads go
    // through the scalar data path
    const int p0 = __builtin_amdgcn_readfirstlane(it.w & 0xffff);
    const int nbp = __builtin_amdgcn_readfirstlane((it.w >> 16) - (it.w & 0xffff));
    const int bpoff = __builtin_amdgcn_readfirstlane(br.poff) + p0;
    const PrimPair64* __restrict__ bra64 = pp64 + bpoff;
    const PrimPair32* __restrict__ bra32 = pp32 + bpoff;
    const double ABx = br.ABx, ABy = br.ABy, ABz = br.ABz;
    __syncthreads();  // the previous item's flush is complete
    for (int i = tid; i < nrk; i += JK_TEAM) {
      const int r = i / N, c = i - r * N;
      const int g = r < NA ? ao + r : bo + (r - NA);
      prow[i] = Pm[(long)g * N + c];
      krow[i] = 0.0;
      if (NDEN == 2) {
        prow[nrk + i] = Pm2[(long)g * N + c];
        krow[nrk + i] = 0.0;
      }
    }
    if (tid < NA * NB) jab_s[tid] = 0.0;
    if (tid == 0) next_chunk = nwave;
    __syncthreads();
    const int v1 = it.z;
    const double fA = (A == B) ? 0.5 : 1.0;
    double jab[NA * NB];
#pragma unroll
    for (int i = 0; i < NA * NB; ++i) jab[i] = 0.0;

#if GX_DYN_CHUNK
    // waves take the item's chunks of 64 kets in turn from an LDS counter
    for (int v0 = it.y + wave * GX_WAVE; v0 < v1;) {
#else
    for (int v0 = it.y + wave * GX_WAVE; v0 < v1; v0 += nwave * GX_WAVE) {
#endif
      const int v = v0 + lane;
      const int ke = v < v1 ? a.kets[v] : GX_KET_NONE;
      const bool act = ke >= 0;
      const int jp = act ? (ke & GX_KET_ID) : 0;
      const bool lo = act && (ke & GX_KET_FP32);
      const bool hi = act && !(ke & GX_KET_FP32);
      const PairRec kq = a.pr.rec[jp];
      const double CDx = kq.ABx, CDy = kq.ABy, CDz = kq.ABz;
      // Evaluation and digestion of the ket components PC::cidx(0 .. PC::NC)
      // of the quartet: all of them, or those of one pass.
      auto pass = [&](auto tag) {
        using PC = typename decltype(tag)::type;
        constexpr int NCP = PC::NC;
        double val[PC::NOUT];
        if (hi)
          quartet_jk<PC, double, PrimPair64, MIXED ? jk_group_mixed64<LA, LB, LC, LD>() : GX_JKG, BG>(bra64, nbp, pp64 + kq.poff, kq.np, ABx,
                                                                          ABy, ABz, CDx, CDy, CDz, tab64, val);
        if constexpr (MIXED) {
          if (lo) {
            float vf[PC::NOUT];
            quartet_jk<PC, float, PrimPair32, GX_JKG32, BG>(bra32, nbp, pp32 + kq.poff, kq.np, (float)ABx,
                                                             (float)ABy, (float)ABz, (float)CDx, (float)CDy,
                                                             (float)CDz, tab32, vf);
#pragma unroll
            for (int k = 0; k < PC::NOUT; ++k) val[k] = (double)vf[k];
          }
        }
        if (!act) return;
        const int co = kq.fa, dof = kq.fb;
        double f = fA * ((kq.A == kq.B) ? 0.5 : 1.0);
        if (jp == ip) f *= 0.5;
#pragma unroll
        for (int k = 0; k < PC::NOUT; ++k) val[k] *= f;
        double Pcd[NCP * ND];
#pragma unroll
        for (int ic = 0; ic < NCP; ++ic)
#pragma unroll
          for (int id = 0; id < ND; ++id) Pcd[ic * ND + id] = Pjm[(long)(co + PC::cidx(ic)) * N + dof + id];
        double jcd[NCP * ND];
#pragma unroll
        for (int i = 0; i < NCP * ND; ++i) jcd[i] = 0.0;
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
          for (int ib = 0; ib < NB; ++ib) {
            double pab = prow[ia * N + bo + ib];
            if (NDEN == 2) pab += prow[nrk + ia * N + bo + ib];
#pragma unroll
            for (int ic = 0; ic < NCP; ++ic)
#pragma unroll
              for (int id = 0; id < ND; ++id) {
                const double x = val[((ia * NB + ib) * NCP + ic) * ND + id];
                jab[ia * NB + ib] = fma(x, Pcd[ic * ND + id], jab[ia * NB + ib]);
                jcd[ic * ND + id] = fma(x, pab, jcd[ic * ND + id]);
              }
          }
#pragma unroll
        for (int ic = 0; ic < NCP; ++ic)
#pragma unroll
          for (int id = 0; id < ND; ++id)
            atomicAdd(&Jkp[kq.jko + (long)(PC::cidx(ic) * ND + id) * a.jkcs], 2.0 * jcd[ic * ND + id]);
#pragma unroll
        for (int d = 0; d < NDEN; ++d) {
          const double* __restrict__ pr = prow + d * nrk;
          double* kr = krow + d * nrk;
          double kac[NA * NCP], kad[NA * ND], kbc[NB * NCP], kbd[NB * ND];
#pragma unroll
          for (int i = 0; i < NA * NCP; ++i) kac[i] = 0.0;
#pragma unroll
          for (int i = 0; i < NA * ND; ++i) kad[i] = 0.0;
#pragma unroll
          for (int i = 0; i < NB * NCP; ++i) kbc[i] = 0.0;
#pragma unroll
          for (int i = 0; i < NB * ND; ++i) kbd[i] = 0.0;
#pragma unroll
          for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
              for (int ic = 0; ic < NCP; ++ic) {
                const double pac = pr[ia * N + co + PC::cidx(ic)];
                const double pbc = pr[(rowB + ib) * N + co + PC::cidx(ic)];
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                  const double x = val[((ia * NB + ib) * NCP + ic) * ND + id];
                  const double pad = pr[ia * N + dof + id];
                  const double pbd = pr[(rowB + ib) * N + dof + id];
                  kac[ia * NCP + ic] = fma(x, pbd, kac[ia * NCP + ic]);
                  kad[ia * ND + id] = fma(x, pbc, kad[ia * ND + id]);
                  kbc[ib * NCP + ic] = fma(x, pad, kbc[ib * NCP + ic]);
                  kbd[ib * ND + id] = fma(x, pac, kbd[ib * ND + id]);
                }
              }
#pragma unroll
          for (int ia = 0; ia < NA; ++ia) {
#pragma unroll
            for (int ic = 0; ic < NCP; ++ic) atomicAdd(&kr[ia * N + co + PC::cidx(ic)], kac[ia * NCP + ic]);
#pragma unroll
            for (int id = 0; id < ND; ++id) atomicAdd(&kr[ia * N + dof + id], kad[ia * ND + id]);
          }
#pragma unroll
          for (int ib = 0; ib < NB; ++ib) {
#pragma unroll
            for (int ic = 0; ic < NCP; ++ic)
              atomicAdd(&kr[(rowB + ib) * N + co + PC::cidx(ic)], kbc[ib * NCP + ic]);
#pragma unroll
            for (int id = 0; id < ND; ++id) atomicAdd(&kr[(rowB + ib) * N + dof + id], kbd[ib * ND + id]);
          }
        }
      };
      jk_passes<LA, LB, LC, LD>(pass);
#if GX_DYN_CHUNK
      int ck = 0;
      if (lane == 0) ck = atomicAdd(&next_chunk, 1);
      v0 = it.y + __builtin_amdgcn_readfirstlane(ck) * GX_WAVE;
#endif
    }

    // J_ab: wave reduction, then one LDS atomic per wave and element.
#pragma unroll
    for (int i = 0; i < NA * NB; ++i) {
      double x = jab[i];
#pragma unroll
      for (int off = GX_WAVE / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, GX_WAVE);
      if (lane == 0 && x != 0.0) atomicAdd(&jab_s[i], x);
    }
    __syncthreads();
    if (tid < NA * NB) {
      const double x = jab_s[tid];
      if (x != 0.0) {
        const int ia = tid / NB, ib = tid - ia * NB;
        atomicAdd(&Jm[(long)(ao + ia) * N + bo + ib], 2.0 * x);
      }
    }
    const int nrow = (A == B) ? NA : NA + NB;
    for (int i = tid; i < nrow * N; i += JK_TEAM) {
      const int r = i / N, c = i - r * N;
      const int g = r < NA ? ao + r : bo + (r - NA);
      const double x = krow[i];
      if (x != 0.0) atomicAdd(&Km[(long)g * N + c], x);
      if (NDEN == 2) {
        const double y = krow[nrk + i];
        if (y != 0.0) atomicAdd(&Km2[(long)g * N + c], y);
      }
    }
    if (tid == 0) next_item = atomicAdd(&a.item_ctr[0], 1) + (int)gridDim.x;
    __syncthreads();
    w = next_item;
  }
  // the last workgroup to finish zeroes the counters for the next launch
  // (every other workgroup has made its last claim before counting itself)
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(&a.item_ctr[1], 1) == (int)gridDim.x - 1) {
      atomicExch(&a.item_ctr[0], 0);
      atomicExch(&a.item_ctr[1], 0);
    }
  }
}

// Element t of the packed per-system matrices -> (system, row, column).
__device__ __forceinline__ void packed_index(long t, int nsys, const long* __restrict__ mb,
                                             const int* __restrict__ n, int& s, int& i, int& j) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (mb[m] <= t) lo = m;
    else hi = m - 1;
  }
  s = lo;
  const long u = t - mb[s];
  i = (int)(u / n[s]);
  j = (int)(u - (long)i * n[s]);
}

// J = s (Jt + Jt^T) s, K = s (Kt + Kt^T) s per system (scale by global
// index), where Jt also takes the pair-indexed J_cd sums: element t of the
// packed matrices is entry jkp_of[t] of every Jkp copy (-1: t lies in the
// transposed half of a shell-pair block, whose sums sit at the transpose).
__global__ void k_symmetrize(long total, int nsys, const long* __restrict__ mb, const int* __restrict__ n,
                             const int* __restrict__ bfb, const double* __restrict__ scale,
                             const double* __restrict__ Jt, const double* __restrict__ Jkp, long jstride,
                             const int* __restrict__ jkp_of, const double* __restrict__ Kt,
                             double* __restrict__ J, double* __restrict__ K) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int s, i, j;
  packed_index(t, nsys, mb, n, s, i, j);
  const double f = scale[bfb[s] + i] * scale[bfb[s] + j];
  const long tt = mb[s] + (long)j * n[s] + i;
  double x = Jt[t] + Jt[tt];
  const int e1 = jkp_of[t], e2 = jkp_of[tt];
  if (e1 >= 0) {
#pragma unroll
    for (int k = 0; k < GX_JT_COPIES; ++k) x += Jkp[k * jstride + e1];
  }
  if (e2 >= 0) {
#pragma unroll
    for (int k = 0; k < GX_JT_COPIES; ++k) x += Jkp[k * jstride + e2];
  }
  J[t] = f * x;
  K[t] = f * (Kt[t] + Kt[tt]);
}

// K = s (Kt + Kt^T) s per system from a single-copy accumulator.
__global__ void k_symmetrize1(long total, int nsys, const long* __restrict__ mb, const int* __restrict__ n,
                              const int* __restrict__ bfb, const double* __restrict__ scale,
                              const double* __restrict__ Kt, double* __restrict__ K) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int s, i, j;
  packed_index(t, nsys, mb, n, s, i, j);
  const long tt = mb[s] + (long)j * n[s] + i;
  K[t] = scale[bfb[s] + i] * scale[bfb[s] + j] * (Kt[t] + Kt[tt]);
}

// P' = s P s per system.
__global__ void k_scale_density(long total, int nsys, const long* __restrict__ mb, const int* __restrict__ n,
                                const int* __restrict__ bfb, const double* __restrict__ scale,
                                const double* __restrict__ P, double* __restrict__ Ps) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int s, i, j;
  packed_index(t, nsys, mb, n, s, i, j);
  Ps[t] = P[t] * scale[bfb[s] + i] * scale[bfb[s] + j];
}

// Start of a J/K build in one launch: Ps = s P s (packed, `total`
// elements), Kt = 0 over the same elements, Jt = 0 over its `njt` doubles.
__global__ void k_jk_prepare(long total, int nsys, const long* __restrict__ mb, const int* __restrict__ n,
                             const int* __restrict__ bfb, const double* __restrict__ scale,
                             const double* __restrict__ P, double* __restrict__ Ps, double* __restrict__ Kt,
                             long njt, double* __restrict__ Jt) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < njt) Jt[t] = 0.0;
  if (t >= total) return;
  int s, i, j;
  packed_index(t, nsys, mb, n, s, i, j);
  Ps[t] = P[t] * scale[bfb[s] + i] * scale[bfb[s] + j];
  Kt[t] = 0.0;
}

// z = x + y, elementwise.
__global__ void k_add(long n, const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ z) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) z[t] = x[t] + y[t];
}

// F_0..F_L at T[i] by the branch-free evaluation of grid BG, in FP64 or
// FP32 (test entry point gx_boys_values).
template <int L, class BG, bool F32>
__global__ void k_boys_values(int n, const double* __restrict__ T, const double* __restrict__ t64,
                              const float* __restrict__ t32, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if constexpr (F32) {
    float F[L + 1];
    boys32_nb<L, BG>((float)T[i], 1.0f, t32 + L * BG::NT * BG::ROW32, F);
#pragma unroll
    for (int m = 0; m <= L; ++m) out[(long)i * (L + 1) + m] = (double)F[m];
  } else {
    double F[L + 1];
    boys64_nb<L, BG>(T[i], 1.0, t64 + L * BG::NT * BG::ROW64, F);
#pragma unroll
    for (int m = 0; m <= L; ++m) out[(long)i * (L + 1) + m] = F[m];
  }
}

// Per-L Taylor tables of grid BG in HBM (from the master table of that
// grid), the same layout the kernels stage into LDS.
template <class BG>
__global__ void k_boys_tables(int lmax, const double* __restrict__ master, double* __restrict__ t64,
                              float* __restrict__ t32) {
  const int L = blockIdx.x;
  if (L > lmax) return;
  boys_stage64<BG>(L, master, t64 + (long)L * BG::NT * BG::ROW64);
  boys_stage32<BG>(L, master, t32 + (long)L * BG::NT * BG::ROW32);
}

__global__ void k_fill(long n, double* __restrict__ x, double v) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = v;
}

}  // namespace gx
