// Per-lane two-electron gradient kernels for the smaller classes: the
// quartets of a bra are swept 64 kets at a time, one ket shell pair per lane,
// as in the fused J/K kernel; the lane accumulates the weighted [e0|f0] sets
// of EriGradLanePass, then its HRR and the contraction with
//   D_k = f (4 P_ab P_cd - P_ac P_bd - P_ad P_bc)
// give the nine derivatives on A, B, C (D by translational invariance).
// Classes whose sets fit the registers at once run one pass; those whose
// per-centre sets fit run three passes over the primitives (EriGradLane::NPASS).
// The larger classes use k_grad_coop (gx_coop.hip).
//
// MIXED = true: the kets whose Schwarz product is below the FP32 bound (the
// second range of every ket segment, as in k_jk) run their primitive loop in
// FP32 (FP32 pair records and Taylor table); the weighted sets are then
// converted and the HRR and contraction run in FP64.

namespace gx {

template <int L>
__device__ __forceinline__ void boys_nb(double T, double scale, const double* __restrict__ tab, double (&F)[L + 1]) {
  boys64_nb<L>(T, scale, tab, F);
}
template <int L>
__device__ __forceinline__ void boys_nb(float T, float scale, const float* __restrict__ tab, float (&F)[L + 1]) {
  boys32_nb<L>(T, scale, tab, F);
}

// One pass of the per-lane gradient: the weighted sets of pass X over the
// bra x ket primitive pairs in precision R (records PP, table tab), then its
// HRR and contraction into g in FP64.  The density blocks are read (from the
// LDS rows of the bra's shells and P_cd from HBM) only after the primitive
// loop, so they hold no registers there.
template <int LA, int LB, int LC, int LD, int X, typename R, typename PP>
__device__ __forceinline__ void grad_lane_pass(const PP* __restrict__ bra, int nbp, const PP* __restrict__ kp, int nk,
                                               const R* __restrict__ tab, double ABx, double ABy, double ABz,
                                               double CDx, double CDy, double CDz,
                                               const double* __restrict__ prow, const double* __restrict__ Pm,
                                               int N, int rowB, int bo, int co, int dof, double f, double (&g)[9]) {
  using P = EriGradLanePass<LA, LB, LC, LD, X>;
  using C = EriGradLane<LA, LB, LC, LD>;
  constexpr int NA = C::NA, NB = C::NB, NC = C::NC, ND = C::ND;
  R acc[P::NACC];
#pragma unroll
  for (int i = 0; i < P::NACC; ++i) acc[i] = R(0);
  for (int ik = 0; ik < nk; ++ik) {
    const PP k = kp[ik];
    const R w3 = R(2) * k.ea;
    for (int ib = 0; ib < nbp; ++ib) {
      const PP b = bra[ib];
      const R rs = gx_rsqrt(b.p + k.p);
      const R is = rs * rs;
      const R rho = b.p * k.p * is;
      const R PQx = b.Px - k.Px, PQy = b.Py - k.Py, PQz = b.Pz - k.Pz;
      const R T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
      R F[P::L + 1];
      boys_nb<P::L>(T, b.Kip * k.Kip * rs, tab, F);
      const R rho_p = rho * b.ip, rho_q = rho * k.ip;
      P::template prim<R>(acc, F, b.PAx, b.PAy, b.PAz, -rho_p * PQx, -rho_p * PQy, -rho_p * PQz, k.PAx, k.PAy,
                          k.PAz, rho_q * PQx, rho_q * PQy, rho_q * PQz, R(0.5) * b.ip, R(0.5) * k.ip, R(0.5) * is,
                          rho_p, rho_q, R(2) * b.ea, R(2) * (b.p - b.ea), w3);
    }
  }
  double accd[P::NACC];
#pragma unroll
  for (int i = 0; i < P::NACC; ++i) accd[i] = (double)acc[i];
  double Pab[NA * NB], Pcd[NC * ND], Pac[NA * NC], Pad[NA * ND], Pbc[NB * NC], Pbd[NB * ND];
#pragma unroll
  for (int ia = 0; ia < NA; ++ia) {
#pragma unroll
    for (int ib = 0; ib < NB; ++ib) Pab[ia * NB + ib] = prow[ia * N + bo + ib];
#pragma unroll
    for (int ic = 0; ic < NC; ++ic) Pac[ia * NC + ic] = prow[ia * N + co + ic];
#pragma unroll
    for (int id = 0; id < ND; ++id) Pad[ia * ND + id] = prow[ia * N + dof + id];
  }
#pragma unroll
  for (int ib = 0; ib < NB; ++ib) {
#pragma unroll
    for (int ic = 0; ic < NC; ++ic) Pbc[ib * NC + ic] = prow[(rowB + ib) * N + co + ic];
#pragma unroll
    for (int id = 0; id < ND; ++id) Pbd[ib * ND + id] = prow[(rowB + ib) * N + dof + id];
  }
#pragma unroll
  for (int ic = 0; ic < NC; ++ic)
#pragma unroll
    for (int id = 0; id < ND; ++id) Pcd[ic * ND + id] = Pm[(long)(co + ic) * N + dof + id];
  P::template finish<double>(accd, ABx, ABy, ABz, CDx, CDy, CDz, Pab, Pcd, Pac, Pad, Pbc, Pbd, f, g);
}

// All passes of one quartet in precision R.
template <int LA, int LB, int LC, int LD, typename R, typename PP>
__device__ __forceinline__ void grad_lane_quartet(const PP* __restrict__ bra, int nbp, const PP* __restrict__ kp,
                                                  int nk, const R* __restrict__ tab, double ABx, double ABy,
                                                  double ABz, double CDx, double CDy, double CDz,
                                                  const double* __restrict__ prow, const double* __restrict__ Pm,
                                                  int N, int rowB, int bo, int co, int dof, double f, double (&g)[9]) {
  grad_lane_pass<LA, LB, LC, LD, 0, R, PP>(bra, nbp, kp, nk, tab, ABx, ABy, ABz, CDx, CDy, CDz, prow, Pm, N, rowB,
                                           bo, co, dof, f, g);
  if constexpr (EriGradLane<LA, LB, LC, LD>::NPASS == 3) {
    grad_lane_pass<LA, LB, LC, LD, 1, R, PP>(bra, nbp, kp, nk, tab, ABx, ABy, ABz, CDx, CDy, CDz, prow, Pm, N,
                                             rowB, bo, co, dof, f, g);
    grad_lane_pass<LA, LB, LC, LD, 2, R, PP>(bra, nbp, kp, nk, tab, ABx, ABy, ABz, CDx, CDy, CDz, prow, Pm, N,
                                             rowB, bo, co, dof, f, g);
  }
}

template <int LA, int LB, int LC, int LD, bool MIXED>
__global__ __launch_bounds__(JK_TEAM) void k_grad_lane(GradArgs a) {
  using C = EriGradLane<LA, LB, LC, LD>;
  constexpr int NA = C::NA, NB = C::NB, NC = C::NC, ND = C::ND;
  extern __shared__ double prow[];  // P rows of the bra's shells
  __shared__ int r_start[2 * JK_MAXSEG], r_pre[2 * JK_MAXSEG + 1], r_len[2 * JK_MAXSEG];
  __shared__ int next_chunk;
  __shared__ double gsh[3 * GX_GRAD_MAXATOM];
  const int tid = threadIdx.x;
  const int lane = tid & (GX_WAVE - 1);
  const int wave = tid / GX_WAVE;
  constexpr int nwave = JK_TEAM / GX_WAVE;
  const double* __restrict__ tab = a.tab64 + C::L * BOYS_NT * BOYS_ROW64;
  const float* __restrict__ tab32 = a.tab32 + C::L * BOYS_NT * BOYS_ROW32;
  // nuclear sums of the current system, kept in LDS across the items of
  // the workgroup and flushed when the system changes and at the end
  int cur = -1, cur3 = 0;
  // items in decreasing cost: the first is blockIdx.x, the rest are claimed
  // from a counter as workgroups finish
  __shared__ int next_item;
  for (int w = blockIdx.x; w < a.nitem;) {
    const int4 it = a.items[w];
    const int ip = a.bra0 + it.x;
    const int A = a.pr.A[ip], B = a.pr.B[ip];
    const int ao = a.sh.lbf[A], bo = a.sh.lbf[B];
    const int rowB = (A == B) ? 0 : NA;
    const int sys = a.sh.sys[A];
    const int N = a.sy.n[sys];
    const double* __restrict__ Pm = a.P + a.sy.mb[sys];
    const int s0 = a.cseg[sys];
    const int nseg = a.cseg[sys + 1] - s0;
    const int at0 = a.sys_atom[sys], nat3 = 3 * (a.sys_atom[sys + 1] - at0);
    const int p0 = __builtin_amdgcn_readfirstlane(it.w & 0xffff);
    const int nbp = __builtin_amdgcn_readfirstlane((it.w >> 16) - (it.w & 0xffff));
    const int bpoff = __builtin_amdgcn_readfirstlane(a.pr.poff[ip]) + p0;
    const PrimPair64* __restrict__ bra = a.pr.pp64 + bpoff;
    const double ABx = a.sh.ctr[3 * A] - a.sh.ctr[3 * B], ABy = a.sh.ctr[3 * A + 1] - a.sh.ctr[3 * B + 1],
                 ABz = a.sh.ctr[3 * A + 2] - a.sh.ctr[3 * B + 2];
    __syncthreads();  // the previous item's readers of prow / r_* are done
    for (int i = tid; i < (NA + NB) * N; i += JK_TEAM) {
      const int r = i / N, c = i - r * N;
      prow[i] = Pm[(long)(r < NA ? ao + r : bo + (r - NA)) * N + c];
    }
    if (tid == 0) next_chunk = nwave;
    if (at0 != cur) {  // another system: flush the sums held for the previous one
      for (int t = tid; t < cur3; t += JK_TEAM)
        if (gsh[t] != 0.0) atomicAdd(&a.grad[3 * cur + t], gsh[t]);
      for (int t = tid; t < nat3; t += JK_TEAM) gsh[t] = 0.0;
      cur = at0;
      cur3 = nat3;
    }
    // ranges r < nseg: FP64 part of segment r; r >= nseg (MIXED): FP32 part
    // of segment r - nseg, starting at the next multiple of 64 so that no
    // chunk mixes precisions (the layout of the host's work items)
    const int nr = MIXED ? 2 * nseg : nseg;
    if (tid < nseg) {
      const int* rr = a.rng + (long)it.x * 2 * JK_MAXSEG;
      r_start[tid] = a.seg_lo[s0 + tid];
      r_pre[tid + 1] = rr[tid];
      if (MIXED) {
        r_start[nseg + tid] = a.seg_lo[s0 + tid] + rr[tid];
        r_pre[nseg + tid + 1] = rr[JK_MAXSEG + tid];
      }
    }
    __syncthreads();
    if (tid == 0) {
      int acc = 0;
      for (int r = 0; r < nr; ++r) {
        const int len = r_pre[r + 1];
        if (r == nseg) acc = (acc + GX_WAVE - 1) & ~(GX_WAVE - 1);
        r_pre[r + 1] = acc + len;
        r_len[r] = len;
        acc += len;
      }
      for (int r = 0; r < nr; ++r) r_pre[r] = r_pre[r + 1] - r_len[r];
      r_pre[nr] = acc;
    }
    __syncthreads();
    const int total = min(r_pre[nr], it.z);
    const double fA = (A == B) ? 0.5 : 1.0;
    double gab[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) gab[j] = 0.0;
    for (int v0 = it.y + wave * GX_WAVE; v0 < total;) {
      const int v = v0 + lane;
      int r = 0;
      if (v < total)
        while (r + 1 < nr && v >= r_pre[r + 1]) ++r;
      const bool act = v < total && v < r_pre[r] + r_len[r];
      if (act) {
        const int jp = r_start[r] + (v - r_pre[r]);
        const int Cs = a.pr.A[jp], Ds = a.pr.B[jp];
        const int co = a.sh.lbf[Cs], dof = a.sh.lbf[Ds];
        const PrimPair64* __restrict__ kp = a.pr.pp64 + a.pr.poff[jp];
        const int nk = a.pr.np[jp];
        double f = fA * ((Cs == Ds) ? 0.5 : 1.0);
        if (jp == ip) f *= 0.5;
        const double CDx = a.sh.ctr[3 * Cs] - a.sh.ctr[3 * Ds], CDy = a.sh.ctr[3 * Cs + 1] - a.sh.ctr[3 * Ds + 1],
                     CDz = a.sh.ctr[3 * Cs + 2] - a.sh.ctr[3 * Ds + 2];
        double g[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) g[j] = 0.0;
        if (!MIXED || r < nseg) {
          grad_lane_quartet<LA, LB, LC, LD, double, PrimPair64>(bra, nbp, kp, nk, tab, ABx, ABy, ABz, CDx, CDy, CDz,
                                                                prow, Pm, N, rowB, bo, co, dof, f, g);
        } else if constexpr (MIXED) {
          grad_lane_quartet<LA, LB, LC, LD, float, PrimPair32>(a.pr.pp32 + bpoff, nbp, a.pr.pp32 + a.pr.poff[jp],
                                                               nk, tab32, ABx, ABy, ABz, CDx, CDy, CDz, prow, Pm,
                                                               N, rowB, bo, co, dof, f, g);
        }
#pragma unroll
        for (int j = 0; j < 6; ++j) gab[j] += g[j];
        const int atC = a.atom[Cs] - at0, atD = a.atom[Ds] - at0;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          atomicAdd(&gsh[3 * atC + x], g[6 + x]);
          atomicAdd(&gsh[3 * atD + x], -(g[x] + g[3 + x] + g[6 + x]));
        }
      }
      int ck = 0;
      if (lane == 0) ck = atomicAdd(&next_chunk, 1);
      v0 = it.y + __builtin_amdgcn_readfirstlane(ck) * GX_WAVE;
    }
    // A and B are the item's bra: one wave reduction, one atomic per wave
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double x = gab[j];
#pragma unroll
      for (int off = GX_WAVE / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, GX_WAVE);
      gab[j] = x;
    }
    if (lane == 0) {
      const int atA = a.atom[A] - at0, atB = a.atom[B] - at0;
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        atomicAdd(&gsh[3 * atA + x], gab[x]);
        atomicAdd(&gsh[3 * atB + x], gab[3 + x]);
      }
    }
    if (tid == 0) next_item = atomicAdd(&a.item_ctr[0], 1) + (int)gridDim.x;
    __syncthreads();
    w = next_item;
  }
  __syncthreads();
  for (int t = tid; t < cur3; t += JK_TEAM)
    if (gsh[t] != 0.0) atomicAdd(&a.grad[3 * cur + t], gsh[t]);
  // the last workgroup to finish zeroes the counters for the next launch
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(&a.item_ctr[1], 1) == (int)gridDim.x - 1) {
      atomicExch(&a.item_ctr[0], 0);
      atomicExch(&a.item_ctr[1], 0);
    }
  }
}

}  // namespace gx
