// Per-lane two-electron gradient kernels for the smaller classes: the
// quartets of a bra are swept 64 kets at a time, one ket shell pair per lane,
// as in the fused J/K kernel; the lane accumulates the weighted [e0|f0] sets
// of EriGradLanePass, then its HRR and the contraction with
//   D_k = f (4 P_ab P_cd - P_ac P_bd - P_ad P_bc)
// give the nine derivatives on A, B, C (D by translational invariance).
// Classes whose sets fit the registers at once run one pass; those whose
// per-centre sets fit run three passes over the primitives (EriGradLane::NPASS).
// The larger classes use k_grad_rys (gx_rys.hip).
//
// MIXED = true: the kets whose Schwarz product is below the FP32 bound (the
// second range of every ket segment, as in k_jk) run their primitive loop in
// FP32 (FP32 pair records and Taylor table); the weighted sets are then
// converted and the HRR and contraction run in FP64.

namespace gx {

// Two-electron part of the RHF nuclear gradient over the canonical shell
// quartets (the same work items and screening as the J/K build):
//   g_X += sum_k D_k d(ab|cd)_k / dX,  D_k = f (4 P_ab P_cd - P_ac P_bd - P_ad P_bc)
// with f the 8-fold degeneracy weight (the energy of one canonical quartet is
// f v D_k / f = f v (4 P_ab P_cd - P_ac P_bd - P_ad P_bc), from
// E_2e = 1/2 tr(P J) - 1/4 tr(P K)).  Derivatives on A, B, C come from the
// recursions (per-lane class code, or Rys quadrature in gx_rys.hip); D from
// translational invariance.
struct GradArgs {
  ShellView sh;
  PairView pr;
  SysView sy;
  int bra0;
  const int4* items;   // as JKArgs::items
  int nitem;
  const int* kets;     // as JKArgs::kets
  const double* tab64;
  const float* tab32;  // FP32 Taylor tables (mixed-precision per-lane kernels)
  const double* tab64c;  // the same on the coarse grid (staged by the per-lane kernels)
  const float* tab32c;
  const double* P;     // packed densities, pre-scaled by s_k s_l
  const int* atom;     // shell -> atom (over all systems)
  const int* sys_atom; // per system: first nucleus (nsys + 1)
  double* grad;        // 3 natom
  int* item_ctr;       // [next item, workgroups done]: zero before a launch, reset by its last workgroup
};

// Gradient contributions collect in LDS per nucleus of the item's system (at
// most GX_GRAD_MAXATOM) over all the items a workgroup takes from that
// system, and go to HBM when the system changes and at the end: the few
// nuclei of a system would otherwise take every quartet's (or every item's)
// atomics, all on the same addresses.
#define GX_GRAD_MAXATOM 256

// The per-lane gradient kernels stage the coarse-grid Taylor tables of
// their class (FP64, and FP32 in mixed kernels) in LDS.  Measured on MI355X
// against the fine tables read through L1: 66-function gradient 0.352 ->
// 0.34 ms, 255-function 6.0 -> 5.8 ms ((ss|ss) 414 -> 319 us).
using GradGrid = BoysCoarse;
constexpr int GRAD_TAB64 = GradGrid::NT * GradGrid::ROW64;      // doubles
constexpr int GRAD_TAB_DOUBLES = GRAD_TAB64 + GradGrid::NT * GradGrid::ROW32 / 2;

template <int L>
__device__ __forceinline__ void boys_nb(double T, double scale, const double* __restrict__ tab, double (&F)[L + 1]) {
  boys64_nb<L, GradGrid>(T, scale, tab, F);
}
template <int L>
__device__ __forceinline__ void boys_nb(float T, float scale, const float* __restrict__ tab, float (&F)[L + 1]) {
  boys32_nb<L, GradGrid>(T, scale, tab, F);
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
  extern __shared__ double smem[];  // Taylor tables (FP64, FP32), then the P rows of the bra's shells
  double* __restrict__ prow = smem + GRAD_TAB_DOUBLES;
  __shared__ int next_chunk;
  __shared__ double gsh[3 * GX_GRAD_MAXATOM];
  const int tid = threadIdx.x;
  const int lane = tid & (GX_WAVE - 1);
  const int wave = tid / GX_WAVE;
  constexpr int nwave = JK_TEAM / GX_WAVE;
  const double* __restrict__ tab = smem;
  const float* __restrict__ tab32 = reinterpret_cast<const float*>(smem + GRAD_TAB64);
  boys_copy64<GradGrid>(a.tab64c + C::L * GradGrid::NT * GradGrid::ROW64, smem);
  if (MIXED)
    boys_copy32<GradGrid>(a.tab32c + C::L * GradGrid::NT * GradGrid::ROW32, reinterpret_cast<float*>(smem + GRAD_TAB64));
  // nuclear sums of the current system, kept in LDS across the items of
  // the workgroup and flushed when the system changes and at the end
  int cur = -1, cur3 = 0;
  // items in decreasing cost: the first is blockIdx.x, the rest are claimed
  // from a counter as workgroups finish
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
    const int at0 = a.sys_atom[sys], nat3 = 3 * (a.sys_atom[sys + 1] - at0);
    const int p0 = __builtin_amdgcn_readfirstlane(it.w & 0xffff);
    const int nbp = __builtin_amdgcn_readfirstlane((it.w >> 16) - (it.w & 0xffff));
    const int bpoff = __builtin_amdgcn_readfirstlane(br.poff) + p0;
    const PrimPair64* __restrict__ bra = a.pr.pp64 + bpoff;
    const double ABx = br.ABx, ABy = br.ABy, ABz = br.ABz;
    __syncthreads();  // the previous item's readers of prow are done
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
    __syncthreads();
    const int v1 = it.z;
    const double fA = (A == B) ? 0.5 : 1.0;
    double gab[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) gab[j] = 0.0;
    for (int v0 = it.y + wave * GX_WAVE; v0 < v1;) {
      const int v = v0 + lane;
      const int ke = v < v1 ? a.kets[v] : GX_KET_NONE;
      if (ke >= 0) {
        const int jp = ke & GX_KET_ID;
        const PairRec kq = a.pr.rec[jp];
        const int Cs = kq.A, Ds = kq.B;
        const int co = kq.fa, dof = kq.fb;
        const PrimPair64* __restrict__ kp = a.pr.pp64 + kq.poff;
        const int nk = kq.np;
        double f = fA * ((Cs == Ds) ? 0.5 : 1.0);
        if (jp == ip) f *= 0.5;
        const double CDx = kq.ABx, CDy = kq.ABy, CDz = kq.ABz;
        double g[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) g[j] = 0.0;
        if (!MIXED || !(ke & GX_KET_FP32)) {
          grad_lane_quartet<LA, LB, LC, LD, double, PrimPair64>(bra, nbp, kp, nk, tab, ABx, ABy, ABz, CDx, CDy, CDz,
                                                                prow, Pm, N, rowB, bo, co, dof, f, g);
        } else if constexpr (MIXED) {
          grad_lane_quartet<LA, LB, LC, LD, float, PrimPair32>(a.pr.pp32 + bpoff, nbp, a.pr.pp32 + kq.poff,
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
