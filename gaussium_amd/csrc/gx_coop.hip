// Wave-cooperative ERI classes for gfx950.
//
// Classes whose contracted [e0|f0] intermediates do not fit in one lane's
// registers -- (dp|pp), (dp|dp) and the dd classes with a ket HRR -- are
// evaluated by a whole wave on one shell quartet: for the Schwarz bounds,
// the dense and listed ERIs and the nuclear gradient (the J/K build of these
// classes uses the Rys kernel of gx_rys.hip).  The VRR table [e0|f0]^(m), the
// contracted accumulators and the HRR table sit in LDS, and the 64 lanes
// share the entries of each recursion level.  The recursions are the same as
// the per-lane class code (obara_saika_scheme.py:131-146 of the reference, in
// the VRR + HRR arrangement of head_gordon_pople.py:111-181); gen_eri.py
// writes them out as level-ordered programs (EriProg<...>, EriGradProg<...>).
//
// One workgroup is one wave, so __syncthreads() is the level barrier.
// LDS per wave, in doubles: max(NACC + NT, NH) for the quartet.

#ifndef GX_COOP_U
#define GX_COOP_U 4
#endif
#ifndef GX_COOP_U_MIN_NT
#define GX_COOP_U_MIN_NT 800
#endif

namespace gx {

template <class G>
constexpr int coop_lds() {
  return G::NACC + G::NT > G::NH ? G::NACC + G::NT : G::NH;
}

// Contracted shell quartet into H: on return H[G::out()[k]] is component k
// (reference order), unscaled.  Every lane of the wave calls it.
template <class G>
__device__ __forceinline__ void coop_quartet(const PrimPair64* __restrict__ bra, int nb,
                                             const PrimPair64* __restrict__ ket, int nk, double ABx,
                                             double ABy, double ABz, double CDx, double CDy, double CDz,
                                             const double* __restrict__ tab, double* __restrict__ H) {
  constexpr int W = GX_WAVE;
  // program entries per lane in flight: measured to pay off only for the
  // large programs (VRR tables of 800 doubles and more)
  constexpr int U = G::NT >= GX_COOP_U_MIN_NT ? GX_COOP_U : 1;
  const int sub = threadIdx.x & (GX_WAVE - 1);
  const int lane = sub;
  double* __restrict__ Hg = H;
  double* __restrict__ T = Hg + G::NACC;
  const int4* __restrict__ vp = G::vrr();
  const int* __restrict__ vl = G::vlev();
  const auto* __restrict__ ap = G::acc();
  __syncthreads();  // the previous quartet's readers of H are done
  for (int t = sub; t < G::NACC; t += W) Hg[t] = 0.0;
  if (sub == 0) T[0] = 0.0;
  const int nq = nk * nb;
  for (int q = 0; q < nq; ++q) {
    const int ik = q / nb, ib = q - ik * nb;
    const PrimPair64 k = ket[ik];
    const PrimPair64 b = bra[ib];
    const double rs = gx_rsqrt(b.p + k.p);
    const double is = rs * rs;
    const double rho = b.p * k.p * is;
    const double PQx = b.Px - k.Px, PQy = b.Py - k.Py, PQz = b.Pz - k.Pz;
    const double Tb = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
    const double pref = b.Kip * k.Kip * rs;
    double F[G::L + 1];
    boys64_nb<G::L>(Tb, pref, tab, F);
    const double rho_p = rho * b.ip, rho_q = rho * k.ip;
    const double WPx = -rho_p * PQx, WPy = -rho_p * PQy, WPz = -rho_p * PQz;
    const double WQx = rho_q * PQx, WQy = rho_q * PQy, WQz = rho_q * PQz;
    const double o2p = 0.5 * b.ip, o2q = 0.5 * k.ip, o2pq = 0.5 * is;
#pragma unroll
    for (int m = 0; m <= G::L; ++m)
      if (sub == m) T[1 + m] = F[m];
    __syncthreads();
    // Entries of one level read only earlier levels, so each lane takes
    // GX_COOP_U of them at a time: all their descriptor loads, then all their
    // table reads, then the writes, which keeps U dependency chains in flight
    // instead of one.
    for (int lev = 1; lev <= G::L; ++lev) {
      const int e0 = vl[lev], e1 = vl[lev + 1];
      for (int t0 = e0 + sub; t0 < e1; t0 += U * W) {
        int4 en[U];
#pragma unroll
        for (int u = 0; u < U; ++u) en[u] = vp[t0 + u * W < e1 ? t0 + u * W : e0];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const unsigned x = (unsigned)en[u].x, y = (unsigned)en[u].y, z = (unsigned)en[u].z;
          const bool ket_step = (x >> 16) & 1u;
          const unsigned ax = (x >> 17) & 3u;
          const double k1 = (double)((x >> 20) & 15u), k2 = (double)((x >> 24) & 15u);
          const double pa = ax == 0 ? b.PAx : (ax == 1 ? b.PAy : b.PAz);
          const double qc = ax == 0 ? k.PAx : (ax == 1 ? k.PAy : k.PAz);
          const double wp = ax == 0 ? WPx : (ax == 1 ? WPy : WPz);
          const double wq = ax == 0 ? WQx : (ax == 1 ? WQy : WQz);
          const double cA = ket_step ? qc : pa, cB = ket_step ? wq : wp;
          const double o = ket_step ? o2q : o2p, r = ket_step ? rho_q : rho_p;
          double w = cA * T[y & 0xffffu];
          w = fma(cB, T[y >> 16], w);
          w = fma(k1 * o, T[z & 0xffffu] - r * T[z >> 16], w);
          v[u] = fma(k2 * o2pq, T[(unsigned)en[u].w], w);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (t0 + u * W < e1) T[(unsigned)en[u].x & 0xffffu] = v[u];
      }
      __syncthreads();
    }
    for (int t0 = sub; t0 < G::NACC; t0 += U * W) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = T[ap[t0 + u * W < G::NACC ? t0 + u * W : 0]];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (t0 + u * W < G::NACC) Hg[t0 + u * W] += v[u];
    }
    __syncthreads();  // T is rewritten by the next primitive quartet
  }
  const int2* __restrict__ hp = G::hrr();
  const int* __restrict__ hl = G::hlev();
  for (int lev = 0; lev < G::NHLEV; ++lev) {
    const int e0 = hl[lev], e1 = hl[lev + 1];
    for (int t0 = e0 + lane; t0 < e1; t0 += U * GX_WAVE) {
      int2 en[U];
#pragma unroll
      for (int u = 0; u < U; ++u) en[u] = hp[t0 + u * GX_WAVE < e1 ? t0 + u * GX_WAVE : e0];
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const unsigned x = (unsigned)en[u].x, y = (unsigned)en[u].y;
        const bool ket_step = (x >> 16) & 1u;
        const unsigned ax = (x >> 17) & 3u;
        const double ab = ax == 0 ? ABx : (ax == 1 ? ABy : ABz);
        const double cd = ax == 0 ? CDx : (ax == 1 ? CDy : CDz);
        v[u] = fma(ket_step ? cd : ab, H[y >> 16], H[y & 0xffffu]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (t0 + u * GX_WAVE < e1) H[(unsigned)en[u].x & 0xffffu] = v[u];
    }
    __syncthreads();
  }
}

template <class G>
__device__ __forceinline__ void coop_quartet_pairs(const ShellView& sh, const PairView& pr, int ip, int jp,
                                                   const double* __restrict__ tab, double* __restrict__ H) {
  const int A = pr.A[ip], B = pr.B[ip], Cs = pr.A[jp], Ds = pr.B[jp];
  coop_quartet<G>(pr.pp64 + pr.poff[ip], pr.np[ip], pr.pp64 + pr.poff[jp], pr.np[jp],
                  sh.ctr[3 * A] - sh.ctr[3 * B], sh.ctr[3 * A + 1] - sh.ctr[3 * B + 1],
                  sh.ctr[3 * A + 2] - sh.ctr[3 * B + 2], sh.ctr[3 * Cs] - sh.ctr[3 * Ds],
                  sh.ctr[3 * Cs + 1] - sh.ctr[3 * Ds + 1], sh.ctr[3 * Cs + 2] - sh.ctr[3 * Ds + 2], tab, H);
}

// Schwarz bounds of one pair class, one wave per pair (as k_schwarz).
template <int LA, int LB>
__global__ __launch_bounds__(GX_WAVE) void k_schwarz_coop(ShellView sh, PairView pr, int p0,
                                                          const double* __restrict__ tab,
                                                          const double* __restrict__ scale,
                                                          double* __restrict__ Q) {
  using G = EriProg<LA, LB, LA, LB>;
  extern __shared__ double H[];
  const int ip = p0 + blockIdx.x;
  coop_quartet_pairs<G>(sh, pr, ip, ip, tab + G::L * BOYS_NT * BOYS_ROW64, H);
  if (threadIdx.x == 0) {
    const int ao = sh.bf[pr.A[ip]], bo = sh.bf[pr.B[ip]];
    const unsigned short* __restrict__ out = G::out();
    double m = 0.0;
    for (int ia = 0; ia < G::NA; ++ia)
      for (int ib = 0; ib < G::NB; ++ib) {
        const double s = scale[ao + ia] * scale[bo + ib];
        const double d = fabs(H[out[((ia * G::NB + ib) * G::NA + ia) * G::NB + ib]]) * s * s;
        m = d > m ? d : m;
      }
    Q[ip] = sqrt(m);
  }
}

// Dense ERI tensor, one wave per shell quartet (as k_dense).
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(GX_WAVE) void k_dense_coop(ShellView sh, PairView pr, int bra0, int ket0,
                                                        int same, const double* __restrict__ tab,
                                                        const double* __restrict__ scale,
                                                        double* __restrict__ out, long N) {
  using G = EriProg<LA, LB, LC, LD>;
  extern __shared__ double H[];
  const int bi = blockIdx.y, kj = blockIdx.x;
  if (same && kj > bi) return;
  const int ip = bra0 + bi, jp = ket0 + kj;
  coop_quartet_pairs<G>(sh, pr, ip, jp, tab + G::L * BOYS_NT * BOYS_ROW64, H);
  const long ao = sh.bf[pr.A[ip]], bo = sh.bf[pr.B[ip]];
  const long co = sh.bf[pr.A[jp]], dof = sh.bf[pr.B[jp]];
  const unsigned short* __restrict__ os = G::out();
  // Components that the quartet's own symmetries map onto each other (a == b
  // shells, c == d shells, bra pair == ket pair) share their eight images;
  // the lanes write them concurrently, so only the smallest component index
  // of each such orbit writes (its images cover the others' addresses), and
  // every address gets one value.
  const bool sab = pr.A[ip] == pr.B[ip], scd = pr.A[jp] == pr.B[jp], sbk = same && ip == jp;
  for (int k = threadIdx.x; k < G::NOUT; k += GX_WAVE) {
    const int id = k % G::ND, ic = (k / G::ND) % G::NC, ib = (k / (G::ND * G::NC)) % G::NB;
    const int ia = k / (G::ND * G::NC * G::NB);
    bool owner = true;
#pragma unroll
    for (int g = 1; g < 8; ++g) {
      const bool swab = g & 1, swcd = g & 2, swbk = g & 4;
      if ((swab && !sab) || (swcd && !scd) || (swbk && !sbk)) continue;
      int a2 = swab ? ib : ia, b2 = swab ? ia : ib, c2 = swcd ? id : ic, d2 = swcd ? ic : id;
      if (swbk) {
        const int ta = a2, tb = b2;
        a2 = c2;
        b2 = d2;
        c2 = ta;
        d2 = tb;
      }
      if (((a2 * G::NB + b2) * G::NC + c2) * G::ND + d2 < k) owner = false;
    }
    if (!owner) continue;
    const long i = ao + ia, j = bo + ib, kk = co + ic, l = dof + id;
    const double x = H[os[k]] * scale[i] * scale[j] * scale[kk] * scale[l];
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

// Quartet list, one wave per task (as k_quartets).
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(GX_WAVE) void k_quartets_coop(ShellView sh, PairView pr,
                                                           const int* __restrict__ bra,
                                                           const int* __restrict__ ket,
                                                           const int* __restrict__ comp,
                                                           const double* __restrict__ sc,
                                                           const long* __restrict__ pos,
                                                           const double* __restrict__ tab,
                                                           double* __restrict__ out) {
  using G = EriProg<LA, LB, LC, LD>;
  extern __shared__ double H[];
  const int t = blockIdx.x;
  coop_quartet_pairs<G>(sh, pr, bra[t], ket[t], tab + G::L * BOYS_NT * BOYS_ROW64, H);
  if (threadIdx.x == 0) out[pos[t]] = H[G::out()[comp[t]]] * sc[t];
}

}  // namespace gx
