// Rys-quadrature J/K kernel for the classes with many contracted
// intermediates ((dp|pp), (dp|dp), (dd|pp), (dd|dp), (dd|dd)).
//
// The Obara-Saika programs of these classes run level by level through
// tables of 400-3000 doubles per quartet, with few lanes busy per level and
// one quartet per wave in the LDS.  Rys quadrature factorises each primitive
// quartet into n = L/2 + 1 roots times three 2D integrals (RysClass::g2d,
// from gen_rys.py) of (la+1)(lb+1)(lc+1)(ld+1) values, after which every
// output component (ab|cd) = sum_i w_i Gx Gy Gz is independent: the lanes
// own output components and accumulate them over the primitive quartets in
// registers.  Work items of one ket per chunk (as for the cooperative
// classes), the screening and the digestion of k_jk.
//
// Reference operators replaced (src/gaussium/ of the reference): the same
// integrals as ObaraSaika.integrate (integrals/twoelectronrepulsion/
// obara_saika_scheme.py:17-82), here by quadrature over the Boys weight.
#include "gx_rys.h"

namespace gx {

// Primitive quartets evaluated side by side in one round: as many as the
// root (2n lanes each) and 2D-integral (3n lanes each) phases fit in a wave,
// while their 2D tables stay within 2048 doubles of LDS.
template <int LA, int LB, int LC, int LD>
constexpr int rys_pq() {
  using R = RysClass<LA, LB, LC, LD>;
  constexpr int by_lanes = GX_WAVE / (3 * R::NROOT);
  constexpr int by_lds = 2048 / (3 * R::NROOT * R::NG);
  constexpr int p = by_lanes < by_lds ? by_lanes : by_lds;
  return p < 1 ? 1 : p;
}

template <int LA, int LB, int LC, int LD>
constexpr int rys_lds() {  // doubles: rules (2n per primitive quartet), 2D tables (3 n NG each)
  using R = RysClass<LA, LB, LC, LD>;
  return rys_pq<LA, LB, LC, LD>() * (2 * R::NROOT + 3 * R::NROOT * R::NG);
}

// Component of output slot sl: (a, b) = sl % (NA NB), and for
// t = sl / (NA NB), c = (a + t / ND) % NC, d = (b + t % ND) % ND.
template <int NA, int NB, int NC, int ND>
__device__ __forceinline__ void rys_slot(int sl, int& ia, int& ib, int& ic, int& id) {
  const int l = sl % (NA * NB), t = sl / (NA * NB);
  ia = l / NB;
  ib = l - ia * NB;
  ic = (ia + t / ND) % NC;
  id = (ib + t % ND) % ND;
}

#ifndef RYS_RECOMPUTE_NPL
#define RYS_RECOMPUTE_NPL 16
#endif

// Threads per workgroup (one quartet at a time): two waves for the classes
// with more than 16 output components per lane of one wave ((dd|dd):
// 21 -> 11 per lane, 470 -> 256 VGPRs).
template <int LA, int LB, int LC, int LD>
constexpr int rys_threads() {
  return RysClass<LA, LB, LC, LD>::NOUT > 16 * GX_WAVE ? 2 * GX_WAVE : GX_WAVE;
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__((rys_threads<LA, LB, LC, LD>())) void k_jk_rys(JKArgs a) {
  using R = RysClass<LA, LB, LC, LD>;
  constexpr int NR = R::NROOT, NG = R::NG, NOUT = R::NOUT;
  constexpr int NA = R::NA, NB = R::NB, NC = R::NC, ND = R::ND;
  constexpr int SAB = NA * NB, SCD = NC * ND, SAC = NA * NC, SAD = NA * ND, SBC = NB * NC, SBD = NB * ND;
  constexpr int NBLK = SAB + SCD + SAC + SAD + SBC + SBD;
  constexpr int NT = rys_threads<LA, LB, LC, LD>();
  constexpr int NPL = (NOUT + NT - 1) / NT;              // output components per thread
  constexpr int PQ = rys_pq<LA, LB, LC, LD>();          // primitive quartets per round
  extern __shared__ double smem[];
  double* __restrict__ RW = smem;                      // [pq][roots u_i, then weights w_i times the prefactor]
  double* __restrict__ Gt = smem + 2 * NR * PQ;        // [pq][root][axis][NG]
  double* __restrict__ pblk = smem + rys_lds<LA, LB, LC, LD>();  // P_ab P_cd P_ac P_ad P_bc P_bd
  double* __restrict__ jblk = pblk + NBLK;             // J_ab J_cd K_ac K_ad K_bc K_bd partials
  const int lane = threadIdx.x;
  // Output slot s = lane + NT j holds component (a, b, c, d) with (a, b) =
  // s % (NA NB) and, for t = s / (NA NB), c = (a + t / ND) % NC and
  // d = (b + t % ND) % ND: every component once, and the 64 lanes of a slot
  // row meet each J_ab, J_cd, K_ad and K_bc entry at most twice in the
  // digestion's LDS atomics (consecutive components would put 36 lanes on one
  // J_ab entry).  off: the slot's x, y, z 2D offsets, packed.
  //   With NPL <= RYS_RECOMPUTE_NPL the slot's indices and the decoded
  // offsets are recomputed where they are used, from copies the compiler
  // cannot see through, instead of being kept in registers across the
  // primitive loop: measured to lift (dd|dp) from one to two waves per SIMD
  // (281 -> 208 VGPRs, 115 -> 98 us at the 255-function cluster); (dd|dd),
  // one wave per SIMD either way, runs faster with them kept.
  constexpr bool RECOMPUTE = NPL <= RYS_RECOMPUTE_NPL;
  int off[NPL];
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    int ia, ib, ic, id;
    rys_slot<NA, NB, NC, ND>(lane + j * NT, ia, ib, ic, id);
    off[j] = lane + j * NT < NOUT ? R::off()[((ia * NB + ib) * NC + ic) * ND + id] : 0;
  }
  // static striding over the items (their claiming from a counter measured
  // slower for the one-quartet-per-wave kernels)
  for (int w = blockIdx.x; w < a.nitem; w += gridDim.x) {
    const int4 it = a.items[w];
    const int ip = a.bra0 + it.x;
    const PairRec br = a.pr.rec[ip];
    const int ao = br.fa, bo = br.fb;
    const int sys = br.sys;
    const int N = a.sy.n[sys];
    const long mb = a.sy.mb[sys];
    const int p0 = it.w & 0xffff, nbp = (it.w >> 16) - p0;
    const PrimPair64* __restrict__ bra = a.pr.pp64 + br.poff + p0;
    const double AB[3] = {br.ABx, br.ABy, br.ABz};
    const double fA = (br.A == br.B) ? 0.5 : 1.0;
    for (int v = it.y; v < it.z; ++v) {
      const int jp = __builtin_amdgcn_readfirstlane(a.kets[v]);  // FP64 kets only
      const PairRec kq = a.pr.rec[jp];
      const int co = kq.fa, dof = kq.fb;
      const double CD[3] = {kq.ABx, kq.ABy, kq.ABz};
      const PrimPair64* __restrict__ ket = a.pr.pp64 + kq.poff;
      double f = fA * ((kq.A == kq.B) ? 0.5 : 1.0);
      if (jp == ip) f *= 0.5;
      // density blocks of this quartet, loaded into registers now (in flight
      // during the primitive loop) and stored to LDS after it
      constexpr int NPB = (NBLK + NT - 1) / NT;
      double pv[NPB];
#pragma unroll
      for (int i = 0; i < NPB; ++i) {
        const int t = lane + i * NT;
        int r0, c0, s;
        int u = t < NBLK ? t : 0;
        if (u < SAB) { r0 = ao; c0 = bo; s = NB; }
        else if ((u -= SAB) < SCD) { r0 = co; c0 = dof; s = ND; }
        else if ((u -= SCD) < SAC) { r0 = ao; c0 = co; s = NC; }
        else if ((u -= SAC) < SAD) { r0 = ao; c0 = dof; s = ND; }
        else if ((u -= SAD) < SBC) { r0 = bo; c0 = co; s = NC; }
        else { u -= SBC; r0 = bo; c0 = dof; s = ND; }
        pv[i] = a.P[mb + (long)(r0 + u / s) * N + c0 + u % s];
      }
      double acc[NPL];
#pragma unroll
      for (int j = 0; j < NPL; ++j) acc[j] = 0.0;
      const int nq = kq.np * nbp;
      // lane roles: roots phase (pq, quantity) for lane < 2 NR PQ; 2D phase
      // (pq, root, axis) for lane < 3 NR PQ
      const int pr = lane / (2 * NR), qr = lane - pr * (2 * NR);
      const int pg = lane / (3 * NR), rg = (lane - pg * (3 * NR)) / 3, xg = lane - pg * (3 * NR) - 3 * rg;
      for (int q0 = 0; q0 < nq; q0 += PQ) {
        const int npq = nq - q0 < PQ ? nq - q0 : PQ;
        __syncthreads();  // the previous round's readers of RW and Gt are done
        if (pr < npq) {
          const int q = q0 + pr, ik = q / nbp, ib = q - ik * nbp;
          const PrimPair64 b = bra[ib];
          const PrimPair64 k = ket[ik];
          const double rs = gx_rsqrt(b.p + k.p);
          const double rho = b.p * k.p * rs * rs;
          const double PQx = b.Px - k.Px, PQy = b.Py - k.Py, PQz = b.Pz - k.Pz;
          const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
          RW[lane] = rys_quantity(NR, T, qr) * (qr < NR ? 1.0 : b.Kip * k.Kip * rs);
        }
        __syncthreads();
        if (pg < npq) {
          const int q = q0 + pg, ik = q / nbp, ib = q - ik * nbp;
          const PrimPair64 b = bra[ib];
          const PrimPair64 k = ket[ik];
          const double rs = gx_rsqrt(b.p + k.p);
          const double is = rs * rs;
          const double rho = b.p * k.p * is;
          const double rp = rho * b.ip, rq = rho * k.ip;
          const double u = RW[pg * 2 * NR + rg];
          const double PA = xg == 0 ? b.PAx : (xg == 1 ? b.PAy : b.PAz);
          const double QC = xg == 0 ? k.PAx : (xg == 1 ? k.PAy : k.PAz);
          const double pq = xg == 0 ? b.Px - k.Px : (xg == 1 ? b.Py - k.Py : b.Pz - k.Pz);
          const double ab = xg == 0 ? AB[0] : (xg == 1 ? AB[1] : AB[2]);
          const double cd = xg == 0 ? CD[0] : (xg == 1 ? CD[1] : CD[2]);
          R::g2d(fma(-rp * pq, u, PA), fma(rq * pq, u, QC), 0.5 * b.ip * fma(-rp, u, 1.0),
                 0.5 * k.ip * fma(-rq, u, 1.0), 0.5 * is * u, ab, cd, Gt + lane * NG);
        }
        __syncthreads();
        for (int pp = 0; pp < npq; ++pp) {
          const double* __restrict__ W = RW + pp * 2 * NR + NR;
          const double* __restrict__ Gq = Gt + pp * 3 * NR * NG;
#pragma unroll
          for (int j = 0; j < NPL; ++j) {
            if (lane + j * NT < NOUT) {
              int o = off[j];
              if constexpr (RECOMPUTE) asm volatile("" : "+v"(o));
              const int ox = o & 1023, oy = (o >> 10) & 1023, oz = o >> 20;
              double s = 0.0;
#pragma unroll
              for (int r = 0; r < NR; ++r) {
                const double* g = Gq + 3 * r * NG;
                s = fma(W[r] * g[ox], g[NG + oy] * g[2 * NG + oz], s);
              }
              acc[j] += s;
            }
          }
        }
      }
#pragma unroll
      for (int i = 0; i < NPB; ++i) {
        const int t = lane + i * NT;
        if (t < NBLK) {
          pblk[t] = pv[i];
          jblk[t] = 0.0;
        }
      }
      __syncthreads();
      const double* pab = pblk;
      const double* pcd = pab + SAB;
      const double* pac = pcd + SCD;
      const double* pad = pac + SAC;
      const double* pbc = pad + SAD;
      const double* pbd = pbc + SBC;
      double* jab = jblk;
      double* jcd = jab + SAB;
      double* kac = jcd + SCD;
      double* kad = kac + SAC;
      double* kbc = kad + SAD;
      double* kbd = kbc + SBC;
      int ln = lane;
      if constexpr (RECOMPUTE) asm volatile("" : "+v"(ln));
#pragma unroll
      for (int j = 0; j < NPL; ++j) {
        if (lane + j * NT < NOUT) {
          int ia, ib, ic, id;
          rys_slot<NA, NB, NC, ND>(ln + j * NT, ia, ib, ic, id);
          const double x = f * acc[j];
          atomicAdd(&jab[ia * NB + ib], x * pcd[ic * ND + id]);
          atomicAdd(&jcd[ic * ND + id], x * pab[ia * NB + ib]);
          atomicAdd(&kac[ia * NC + ic], x * pbd[ib * ND + id]);
          atomicAdd(&kad[ia * ND + id], x * pbc[ib * NC + ic]);
          atomicAdd(&kbc[ib * NC + ic], x * pad[ia * ND + id]);
          atomicAdd(&kbd[ib * ND + id], x * pac[ia * NC + ic]);
        }
      }
      __syncthreads();
      for (int t = lane; t < NBLK; t += NT) {
        const double x = jblk[t];
        if (x == 0.0) continue;
        int r0, c0, s;
        int u = t;
        double* dst = a.Kt;
        double m = 1.0;
        if (u < SAB) { r0 = ao; c0 = bo; s = NB; dst = a.Jt; m = 2.0; }
        else if ((u -= SAB) < SCD) { r0 = co; c0 = dof; s = ND; dst = a.Jt; m = 2.0; }
        else if ((u -= SCD) < SAC) { r0 = ao; c0 = co; s = NC; }
        else if ((u -= SAC) < SAD) { r0 = ao; c0 = dof; s = ND; }
        else if ((u -= SAD) < SBC) { r0 = bo; c0 = co; s = NC; }
        else { u -= SBC; r0 = bo; c0 = dof; s = ND; }
        atomicAdd(&dst[mb + (long)(r0 + u / s) * N + c0 + u % s], m * x);
      }
    }
  }
}

}  // namespace gx

namespace gx {

// LDS of the Rys gradient kernel (doubles): the rule (2n) and, per root and
// axis, the 2D integrals and their A, B, C derivatives (4 NG each).
template <int LA, int LB, int LC, int LD>
constexpr int rys_grad_lds() {
  using R = RysClass<LA, LB, LC, LD>;
  return 2 * R::NROOT_GRAD + 12 * R::NROOT_GRAD * R::NG;
}

// Two-electron gradient of the classes evaluated one quartet per wave, by
// Rys quadrature at L + 1: d/dX_i (ab|cd) for X = A, B, C replaces the X
// factor of axis i by 2 alpha_X E(x_X + 1) - x_X E(x_X - 1) (RysClass::
// g2d_grad); contracted with D_k = f (4 P_ab P_cd - P_ac P_bd - P_ad P_bc)
// (f the 8-fold degeneracy weight), D from translational invariance.  Lanes
// own output components and accumulate the nine sums in registers; items in
// decreasing cost, the first blockIdx.x and the rest claimed from a counter;
// per-nucleus sums in LDS as in k_grad_lane.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(GX_WAVE) void k_grad_rys(GradArgs a) {
  using R = RysClass<LA, LB, LC, LD>;
  constexpr int NR = R::NROOT_GRAD, NG = R::NG, NOUT = R::NOUT;
  constexpr int NA = R::NA, NB = R::NB, NC = R::NC, ND = R::ND;
  constexpr int SAB = NA * NB, SCD = NC * ND, SAC = NA * NC, SAD = NA * ND, SBC = NB * NC, SBD = NB * ND;
  constexpr int NBLK = SAB + SCD + SAC + SAD + SBC + SBD;
  constexpr int NPL = (NOUT + GX_WAVE - 1) / GX_WAVE;
  extern __shared__ double smem[];
  double* __restrict__ RW = smem;          // roots, then weights times the prefactor
  double* __restrict__ Gt = smem + 2 * NR; // [root][axis][G, GA, GB, GC][NG]
  double* __restrict__ pblk = smem + rys_grad_lds<LA, LB, LC, LD>();
  __shared__ double gsh[3 * GX_GRAD_MAXATOM];
  const int lane = threadIdx.x;
  int off[NPL];
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    const int k = lane + j * GX_WAVE;
    off[j] = k < NOUT ? R::off()[k] : 0;
  }
  // nuclear sums of the current system, kept in LDS across the items of
  // the workgroup and flushed when the system changes and at the end
  int cur = -1, cur3 = 0;
  for (int w = blockIdx.x; w < a.nitem;) {
    const int4 it = a.items[w];
    const int ip = a.bra0 + it.x;
    const PairRec br = a.pr.rec[ip];
    const int A = br.A, B = br.B;
    const int ao = br.fa, bo = br.fb;
    const int sys = br.sys;
    const int N = a.sy.n[sys];
    const long mb = a.sy.mb[sys];
    const int at0 = a.sys_atom[sys], nat3 = 3 * (a.sys_atom[sys + 1] - at0);
    const int p0 = it.w & 0xffff, nbp = (it.w >> 16) - p0;
    const PrimPair64* __restrict__ bra = a.pr.pp64 + br.poff + p0;
    const double AB[3] = {br.ABx, br.ABy, br.ABz};
    const double fA = (A == B) ? 0.5 : 1.0;
    if (at0 != cur) {  // another system: flush the sums held for the previous one
      __syncthreads();
      for (int t = lane; t < cur3; t += GX_WAVE)
        if (gsh[t] != 0.0) atomicAdd(&a.grad[3 * cur + t], gsh[t]);
      for (int t = lane; t < nat3; t += GX_WAVE) gsh[t] = 0.0;
      cur = at0;
      cur3 = nat3;
      __syncthreads();
    }
    for (int v = it.y; v < it.z; ++v) {
      const int jp = __builtin_amdgcn_readfirstlane(a.kets[v]);  // FP64 kets only
      const PairRec kq = a.pr.rec[jp];
      const int Cs = kq.A, Ds = kq.B;
      const int co = kq.fa, dof = kq.fb;
      const double CD[3] = {kq.ABx, kq.ABy, kq.ABz};
      const PrimPair64* __restrict__ ket = a.pr.pp64 + kq.poff;
      double f = fA * ((Cs == Ds) ? 0.5 : 1.0);
      if (jp == ip) f *= 0.5;
      __syncthreads();  // the previous quartet's readers of pblk are done
      for (int t = lane; t < NBLK; t += GX_WAVE) {
        int r0, c0, s;
        int u = t;
        if (u < SAB) { r0 = ao; c0 = bo; s = NB; }
        else if ((u -= SAB) < SCD) { r0 = co; c0 = dof; s = ND; }
        else if ((u -= SCD) < SAC) { r0 = ao; c0 = co; s = NC; }
        else if ((u -= SAC) < SAD) { r0 = ao; c0 = dof; s = ND; }
        else if ((u -= SAD) < SBC) { r0 = bo; c0 = co; s = NC; }
        else { u -= SBC; r0 = bo; c0 = dof; s = ND; }
        pblk[t] = a.P[mb + (long)(r0 + u / s) * N + c0 + u % s];
      }
      __syncthreads();
      // density factor of each owned component
      double Dk[NPL];
#pragma unroll
      for (int j = 0; j < NPL; ++j) {
        const int k = lane + j * GX_WAVE;
        Dk[j] = 0.0;
        if (k < NOUT) {
          const int id = k % ND, ic = (k / ND) % NC, ib = (k / (ND * NC)) % NB, ia = k / (ND * NC * NB);
          const double* pab = pblk;
          const double* pcd = pab + SAB;
          const double* pac = pcd + SCD;
          const double* pad = pac + SAC;
          const double* pbc = pad + SAD;
          const double* pbd = pbc + SBC;
          Dk[j] = f * (4.0 * pab[ia * NB + ib] * pcd[ic * ND + id] - pac[ia * NC + ic] * pbd[ib * ND + id] -
                       pad[ia * ND + id] * pbc[ib * NC + ic]);
        }
      }
      double g[9];
#pragma unroll
      for (int t = 0; t < 9; ++t) g[t] = 0.0;
      const int nq = kq.np * nbp;
      for (int q = 0; q < nq; ++q) {
        const int ik = q / nbp, ib = q - ik * nbp;
        const PrimPair64 b = bra[ib];
        const PrimPair64 k = ket[ik];
        const double rs = gx_rsqrt(b.p + k.p);
        const double is = rs * rs;
        const double rho = b.p * k.p * is;
        const double PQx = b.Px - k.Px, PQy = b.Py - k.Py, PQz = b.Pz - k.Pz;
        const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
        const double rp = rho * b.ip, rq = rho * k.ip;
        __syncthreads();  // the previous primitive quartet's readers of RW and Gt are done
        if (lane < 2 * NR) RW[lane] = rys_quantity(NR, T, lane) * (lane < NR ? 1.0 : b.Kip * k.Kip * rs);
        __syncthreads();
        if (lane < 3 * NR) {
          const int r = lane / 3, x = lane - 3 * r;
          const double u = RW[r];
          const double PA = x == 0 ? b.PAx : (x == 1 ? b.PAy : b.PAz);
          const double QC = x == 0 ? k.PAx : (x == 1 ? k.PAy : k.PAz);
          const double pq = x == 0 ? PQx : (x == 1 ? PQy : PQz);
          const double ab = x == 0 ? AB[0] : (x == 1 ? AB[1] : AB[2]);
          const double cd = x == 0 ? CD[0] : (x == 1 ? CD[1] : CD[2]);
          double* Gl = Gt + lane * 4 * NG;
          R::g2d_grad(fma(-rp * pq, u, PA), fma(rq * pq, u, QC), 0.5 * b.ip * fma(-rp, u, 1.0),
                      0.5 * k.ip * fma(-rq, u, 1.0), 0.5 * is * u, ab, cd, b.ea, b.p - b.ea, k.ea, Gl, Gl + NG,
                      Gl + 2 * NG, Gl + 3 * NG);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NPL; ++j) {
          if (lane + j * GX_WAVE < NOUT) {
            const int o3[3] = {off[j] & 1023, (off[j] >> 10) & 1023, off[j] >> 20};
            double t9[9];
#pragma unroll
            for (int t = 0; t < 9; ++t) t9[t] = 0.0;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              const double wr = RW[NR + r];
              double G0[3], GX[3][3];  // G0[axis]; GX[centre][axis]
#pragma unroll
              for (int x = 0; x < 3; ++x) {
                const double* Gl = Gt + (r * 3 + x) * 4 * NG + o3[x];
                G0[x] = Gl[0];
                GX[0][x] = Gl[NG];
                GX[1][x] = Gl[2 * NG];
                GX[2][x] = Gl[3 * NG];
              }
              const double yz = wr * G0[1] * G0[2], xz = wr * G0[0] * G0[2], xy = wr * G0[0] * G0[1];
#pragma unroll
              for (int X = 0; X < 3; ++X) {
                t9[3 * X] = fma(GX[X][0], yz, t9[3 * X]);
                t9[3 * X + 1] = fma(GX[X][1], xz, t9[3 * X + 1]);
                t9[3 * X + 2] = fma(GX[X][2], xy, t9[3 * X + 2]);
              }
            }
#pragma unroll
            for (int t = 0; t < 9; ++t) g[t] = fma(Dk[j], t9[t], g[t]);
          }
        }
      }
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        double x = g[t];
#pragma unroll
        for (int sh = GX_WAVE / 2; sh > 0; sh >>= 1) x += __shfl_xor(x, sh, GX_WAVE);
        g[t] = x;
      }
      if (lane == 0) {
        const int at[3] = {a.atom[A] - at0, a.atom[B] - at0, a.atom[Cs] - at0};
        const int atD = a.atom[Ds] - at0;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          gsh[3 * at[0] + x] += g[x];
          gsh[3 * at[1] + x] += g[3 + x];
          gsh[3 * at[2] + x] += g[6 + x];
          gsh[3 * atD + x] -= g[x] + g[3 + x] + g[6 + x];
        }
      }
    }
    int nx = 0;
    if (lane == 0) nx = atomicAdd(&a.item_ctr[0], 1) + (int)gridDim.x;
    w = __shfl(nx, 0, GX_WAVE);
  }
  __syncthreads();
  for (int t = lane; t < cur3; t += GX_WAVE)
    if (gsh[t] != 0.0) atomicAdd(&a.grad[3 * cur + t], gsh[t]);
  // the last workgroup to finish zeroes the counters for the next launch
  if (lane == 0) {
    __threadfence();
    if (atomicAdd(&a.item_ctr[1], 1) == (int)gridDim.x - 1) {
      atomicExch(&a.item_ctr[0], 0);
      atomicExch(&a.item_ctr[1], 0);
    }
  }
}

}  // namespace gx
