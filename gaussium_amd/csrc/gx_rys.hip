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

template <int LA, int LB, int LC, int LD>
constexpr int rys_lds() {  // doubles: rule (2n), 2D tables (3 n NG)
  using R = RysClass<LA, LB, LC, LD>;
  return 2 * R::NROOT + 3 * R::NROOT * R::NG;
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(GX_WAVE) void k_jk_rys(JKArgs a) {
  using R = RysClass<LA, LB, LC, LD>;
  constexpr int NR = R::NROOT, NG = R::NG, NOUT = R::NOUT;
  constexpr int NA = R::NA, NB = R::NB, NC = R::NC, ND = R::ND;
  constexpr int SAB = NA * NB, SCD = NC * ND, SAC = NA * NC, SAD = NA * ND, SBC = NB * NC, SBD = NB * ND;
  constexpr int NBLK = SAB + SCD + SAC + SAD + SBC + SBD;
  constexpr int NPL = (NOUT + GX_WAVE - 1) / GX_WAVE;  // output components per lane
  extern __shared__ double smem[];
  double* __restrict__ RW = smem;                      // roots u_i, then weights w_i (times the prefactor)
  double* __restrict__ Gt = smem + 2 * NR;             // [root][axis][NG]
  double* __restrict__ pblk = smem + rys_lds<LA, LB, LC, LD>();  // P_ab P_cd P_ac P_ad P_bc P_bd
  double* __restrict__ jblk = pblk + NBLK;             // J_ab J_cd K_ac K_ad K_bc K_bd partials
  const int lane = threadIdx.x;
  // Output slot s = lane + 64 j holds component (a, b, c, d) with (a, b) =
  // s % (NA NB) and, for t = s / (NA NB), c = (a + t / ND) % NC and
  // d = (b + t % ND) % ND: every component once, and the 64 lanes of a slot
  // row meet each J_ab, J_cd, K_ad and K_bc entry at most twice in the
  // digestion's LDS atomics (consecutive components would put 36 lanes on one
  // J_ab entry).  kidx: the component index; off: its x, y, z 2D offsets.
  int off[NPL], kidx[NPL];
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    const int sl = lane + j * GX_WAVE;
    const int l = sl % (NA * NB), t = sl / (NA * NB);
    const int ia = l / NB, ib = l - ia * NB;
    const int ic = (ia + t / ND) % NC, id = (ib + t % ND) % ND;
    kidx[j] = ((ia * NB + ib) * NC + ic) * ND + id;
    off[j] = sl < NOUT ? R::off()[kidx[j]] : 0;
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
      constexpr int NPB = (NBLK + GX_WAVE - 1) / GX_WAVE;
      double pv[NPB];
#pragma unroll
      for (int i = 0; i < NPB; ++i) {
        const int t = lane + i * GX_WAVE;
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
      for (int q = 0; q < nq; ++q) {
        const int ik = q / nbp, ib = q - ik * nbp;
        const PrimPair64 b = bra[ib];
        const PrimPair64 k = ket[ik];
        const double rs = gx_rsqrt(b.p + k.p);
        const double is = rs * rs;
        const double rho = b.p * k.p * is;
        const double PQ[3] = {b.Px - k.Px, b.Py - k.Py, b.Pz - k.Pz};
        const double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
        const double pref = b.Kip * k.Kip * rs;
        const double rp = rho * b.ip, rq = rho * k.ip;
        __syncthreads();  // the previous primitive quartet's readers of RW and Gt are done
        if (lane < 2 * NR) RW[lane] = rys_quantity(NR, T, lane) * (lane < NR ? 1.0 : pref);
        __syncthreads();
        if (lane < 3 * NR) {
          const int r = lane / 3, x = lane - 3 * r;
          const double u = RW[r];
          const double PA = x == 0 ? b.PAx : (x == 1 ? b.PAy : b.PAz);
          const double QC = x == 0 ? k.PAx : (x == 1 ? k.PAy : k.PAz);
          const double pq = x == 0 ? PQ[0] : (x == 1 ? PQ[1] : PQ[2]);
          const double ab = x == 0 ? AB[0] : (x == 1 ? AB[1] : AB[2]);
          const double cd = x == 0 ? CD[0] : (x == 1 ? CD[1] : CD[2]);
          R::g2d(fma(-rp * pq, u, PA), fma(rq * pq, u, QC), 0.5 * b.ip * fma(-rp, u, 1.0),
                 0.5 * k.ip * fma(-rq, u, 1.0), 0.5 * is * u, ab, cd, Gt + lane * NG);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NPL; ++j) {
          if (lane + j * GX_WAVE < NOUT) {
            const int ox = off[j] & 1023, oy = (off[j] >> 10) & 1023, oz = off[j] >> 20;
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              const double* g = Gt + 3 * r * NG;
              s = fma(RW[NR + r] * g[ox], g[NG + oy] * g[2 * NG + oz], s);
            }
            acc[j] += s;
          }
        }
      }
#pragma unroll
      for (int i = 0; i < NPB; ++i) {
        const int t = lane + i * GX_WAVE;
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
#pragma unroll
      for (int j = 0; j < NPL; ++j) {
        if (lane + j * GX_WAVE < NOUT) {
          const int k = kidx[j];
          const int id = k % ND, ic = (k / ND) % NC, ib = (k / (ND * NC)) % NB, ia = k / (ND * NC * NB);
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
      for (int t = lane; t < NBLK; t += GX_WAVE) {
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
