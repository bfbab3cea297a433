// One-electron integrals S, T, V on gfx950 (McMurchie-Davidson).
//
// Replaces (src/gaussium/ of the reference):
//   orbital_overlap          integrals/orbital_overlap_integral.py:16-36
//   kinetic (del operator)   matrixelements/kinetic_energy_matrix.py:16-33,
//                            factory/primitive_basis_factory.py:89-122
//   nuclear_attraction       integrals/nuclear_attraction_integral.py:19-57
// The reference evaluates the Taketa-Huzinaga-O-ohata explicit sums; this
// kernel uses Hermite expansion coefficients E^{ij}_t and Hermite Coulomb
// integrals R_tuv, which give the same integrals.  Kinetic energy uses the
// reference's expansion  -1/2 nabla^2 g_b = b(2l_b+3) g_b - 2b^2 sum_i g_{b+2i}
//                        - 1/2 sum_i l_i(l_i-1) g_{b-2i}.
// One thread per shell pair (i >= j); all component blocks are written to
// both triangles.
#include "gx_common.h"

namespace gx {

#define ONE_LMAX 2            // shell l
#define ONE_EMAX (ONE_LMAX + 2)  // l + 2 for the kinetic expansion
#define ONE_RMAX (2 * ONE_LMAX)  // Hermite order for V
#define ONE_RDIM (2 * ONE_LMAX + 2)  // R_tuv index bound (the gradient raises L by one)
#define ONE_TMAX (ONE_LMAX + ONE_EMAX)  // Hermite index bound of E
// A primitive pair whose Gaussian-product factor |c_a c_b| exp(-ab/(a+b) R^2)
// is below this is skipped: its contribution to any one-electron integral or
// derivative is that factor times bounded polynomial and (2 pi / p) terms,
// far below 1e-12.
#define ONE_PRIM_EPS 1e-20
#define ONE_GRAD_NAT 128  // nuclei per block summed in LDS by the gradient kernels

// Boys F_0..F_n(T) from the master table in global memory (Taylor + downward).
__device__ inline void boys_master(int n, double T, const double* __restrict__ master, double* F) {
  if (T < BOYS_TMAX) {
    const int k0 = (int)(T * BOYS_INV_DT + 0.5);
    const double x = k0 * BOYS_DT - T;
    const double* row = master + k0 * (BOYS_MCOL + 1);
    const double invf[8] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040};
    double f = row[n + 7] * invf[7];
    for (int k = 6; k >= 0; --k) f = fma(f, x, row[n + k] * invf[k]);
    double e = 1.0 / 5040.0;
    e = fma(e, x, 1.0 / 720.0);
    e = fma(e, x, 1.0 / 120.0);
    e = fma(e, x, 1.0 / 24.0);
    e = fma(e, x, 1.0 / 6.0);
    e = fma(e, x, 0.5);
    e = fma(e, x, 1.0);
    e = fma(e, x, 1.0);
    e *= row[BOYS_MCOL];
    F[n] = f;
    for (int m = n; m >= 1; --m) {
      f = fma(2.0 * T, f, e) / (2 * m - 1);
      F[m - 1] = f;
    }
  } else {
    double f = 0.88622692545275801365 * sqrt(1.0 / T);
    F[0] = f;
    for (int m = 0; m < n; ++m) {
      f *= (m + 0.5) / T;
      F[m + 1] = f;
    }
  }
}

// Hermite coefficients E[i][j][t] (without the exp(-mu X^2) factor) for
// i <= imax, j <= jmax along one axis.
__device__ inline void hermite_e(int imax, int jmax, double PA, double PB, double o2p,
                                 double (*E)[ONE_EMAX + 1][ONE_TMAX + 1]) {
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j)
      for (int t = 0; t <= ONE_TMAX; ++t) E[i][j][t] = 0.0;
  E[0][0][0] = 1.0;
  for (int i = 0; i <= imax; ++i) {
    if (i > 0) {
      for (int t = 0; t <= i; ++t) {
        double v = PA * E[i - 1][0][t];
        if (t > 0) v += o2p * E[i - 1][0][t - 1];
        if (t + 1 <= i - 1) v += (t + 1) * E[i - 1][0][t + 1];
        E[i][0][t] = v;
      }
    }
    for (int j = 1; j <= jmax; ++j) {
      for (int t = 0; t <= i + j; ++t) {
        double v = (t <= i + j - 1) ? PB * E[i][j - 1][t] : 0.0;
        if (t > 0) v += o2p * E[i][j - 1][t - 1];
        if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
        E[i][j][t] = v;
      }
    }
  }
}

// Hermite Coulomb integrals R^{(0)}_{tuv} for t + u + v <= L (McMurchie-
// Davidson): R^{(n)}_{000} = (-2p)^n F_n(p |PC|^2), then
// R^{(n)}_{t+1,u,v} = t R^{(n+1)}_{t-1,u,v} + X_PC R^{(n+1)}_{t,u,v} (likewise u, v),
// from n = L down to 0 with two buffers.
__device__ inline void hermite_r(int L, double p, const double* PC, const double* F,
                                 double (*R)[ONE_RDIM][ONE_RDIM], double (*S)[ONE_RDIM][ONE_RDIM]) {
  double m2p[ONE_RDIM + 1];
  m2p[0] = 1.0;
  for (int n = 1; n <= L; ++n) m2p[n] = m2p[n - 1] * (-2.0 * p);
  // the buffer holding order n is R when (L - n) is even
  for (int n = L; n >= 0; --n) {
    double (*cur)[ONE_RDIM][ONE_RDIM] = ((L - n) & 1) ? S : R;
    double (*prv)[ONE_RDIM][ONE_RDIM] = ((L - n) & 1) ? R : S;
    const int top = L - n;
    for (int tt = 0; tt <= top; ++tt)
      for (int u = 0; u <= top - tt; ++u)
        for (int v = 0; v <= top - tt - u; ++v) {
          double r;
          if (tt + u + v == 0) {
            r = m2p[n] * F[n];
          } else if (tt > 0) {
            r = PC[0] * prv[tt - 1][u][v];
            if (tt > 1) r += (tt - 1) * prv[tt - 2][u][v];
          } else if (u > 0) {
            r = PC[1] * prv[tt][u - 1][v];
            if (u > 1) r += (u - 1) * prv[tt][u - 2][v];
          } else {
            r = PC[2] * prv[tt][u][v - 1];
            if (v > 1) r += (v - 1) * prv[tt][u][v - 2];
          }
          cur[tt][u][v] = r;
        }
  }
  if (L & 1)  // order 0 ended in S
    for (int tt = 0; tt <= L; ++tt)
      for (int u = 0; u <= L - tt; ++u)
        for (int v = 0; v <= L - tt - u; ++v) R[tt][u][v] = S[tt][u][v];
}

__device__ inline int ncart_d(int l) { return (l + 1) * (l + 2) / 2; }

__device__ inline void cart_d(int l, int k, int* e) {
  // reference component order for l <= 2
  const int c1[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const int c2[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
  if (l == 0) {
    e[0] = e[1] = e[2] = 0;
  } else if (l == 1) {
    e[0] = c1[k][0]; e[1] = c1[k][1]; e[2] = c1[k][2];
  } else {
    e[0] = c2[k][0]; e[1] = c2[k][1]; e[2] = c2[k][2];
  }
}

// Shell pairs of one system each (the ERI pair list: every unordered pair of
// shells of a system once); nuclei of the pair's system only.
struct OneArgs {
  int npair;
  const int* pA;
  const int* pB;
  const int* sl;
  const int* snp;
  const int* spoff;
  const int* sbf;     // global first function (scale index)
  const int* slbf;    // first function inside the system
  const int* ssys;
  const double* ctr1;
  const double* exps;
  const double* coefs;
  const double* scale;
  const int* sys_atom;  // nsys + 1: nuclei of system s are [sys_atom[s], sys_atom[s + 1])
  const double* nxyz;
  const double* nZ;
  const double* master;
  SysView sy;
};

__global__ __launch_bounds__(64) void k_one_electron(OneArgs a, double* __restrict__ S, double* __restrict__ T,
                                                     double* __restrict__ V) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.npair) return;
  const int si = a.pA[t], sj = a.pB[t];
  const int* __restrict__ sl = a.sl;
  const int* __restrict__ snp = a.snp;
  const int* __restrict__ spoff = a.spoff;
  const double* __restrict__ ctr1 = a.ctr1;
  const double* __restrict__ exps = a.exps;
  const double* __restrict__ coefs = a.coefs;
  const double* __restrict__ nxyz = a.nxyz;
  const double* __restrict__ nZ = a.nZ;
  const double* __restrict__ master = a.master;
  const int sys = a.ssys[si];
  const int c0 = a.sys_atom ? a.sys_atom[sys] : 0, c1 = a.sys_atom ? a.sys_atom[sys + 1] : 0;
  const int la = sl[si], lb = sl[sj];
  const int na = ncart_d(la), nb = ncart_d(lb);
  const double A[3] = {ctr1[3 * si], ctr1[3 * si + 1], ctr1[3 * si + 2]};
  const double B[3] = {ctr1[3 * sj], ctr1[3 * sj + 1], ctr1[3 * sj + 2]};
  double s_acc[36], t_acc[36], v_acc[36];
  for (int k = 0; k < 36; ++k) s_acc[k] = t_acc[k] = v_acc[k] = 0.0;
  double E[3][ONE_LMAX + 1][ONE_EMAX + 1][ONE_TMAX + 1];
  double R[ONE_RDIM][ONE_RDIM][ONE_RDIM], Rs[ONE_RDIM][ONE_RDIM][ONE_RDIM];
  double F[ONE_RDIM + 1];
  for (int ia = 0; ia < snp[si]; ++ia) {
    const double a = exps[spoff[si] + ia], ca = coefs[spoff[si] + ia];
    for (int ib = 0; ib < snp[sj]; ++ib) {
      const double b = exps[spoff[sj] + ib], cb = coefs[spoff[sj] + ib];
      const double p = a + b, o2p = 0.5 / p, mu = a * b / p;
      double P[3], r2 = 0.0;
      for (int x = 0; x < 3; ++x) r2 += (A[x] - B[x]) * (A[x] - B[x]);
      const double Kab = exp(-mu * r2) * ca * cb;
      if (fabs(Kab) < ONE_PRIM_EPS) continue;
      for (int x = 0; x < 3; ++x) {
        P[x] = (a * A[x] + b * B[x]) / p;
        hermite_e(la, lb + 2, P[x] - A[x], P[x] - B[x], o2p, E[x]);
      }
      const double s00 = Kab * pow(M_PI / p, 1.5);
      // overlap and kinetic
      for (int ka = 0; ka < na; ++ka) {
        int ea[3];
        cart_d(la, ka, ea);
        for (int kb = 0; kb < nb; ++kb) {
          int eb[3];
          cart_d(lb, kb, eb);
          const double sx = E[0][ea[0]][eb[0]][0], sy = E[1][ea[1]][eb[1]][0],
                       sz = E[2][ea[2]][eb[2]][0];
          const double sv = sx * sy * sz;
          const int lbt = eb[0] + eb[1] + eb[2];
          double tv = b * (2 * lbt + 3) * sv;
          tv -= 2.0 * b * b * (E[0][ea[0]][eb[0] + 2][0] * sy * sz + sx * E[1][ea[1]][eb[1] + 2][0] * sz +
                               sx * sy * E[2][ea[2]][eb[2] + 2][0]);
          if (eb[0] >= 2) tv -= 0.5 * eb[0] * (eb[0] - 1) * E[0][ea[0]][eb[0] - 2][0] * sy * sz;
          if (eb[1] >= 2) tv -= 0.5 * eb[1] * (eb[1] - 1) * sx * E[1][ea[1]][eb[1] - 2][0] * sz;
          if (eb[2] >= 2) tv -= 0.5 * eb[2] * (eb[2] - 1) * sx * sy * E[2][ea[2]][eb[2] - 2][0];
          s_acc[ka * nb + kb] += s00 * sv;
          t_acc[ka * nb + kb] += s00 * tv;
        }
      }
      // nuclear attraction
      const int L = la + lb;
      for (int c = c0; c < c1; ++c) {
        const double PC[3] = {P[0] - nxyz[3 * c], P[1] - nxyz[3 * c + 1], P[2] - nxyz[3 * c + 2]};
        const double Tb = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
        boys_master(L, Tb, master, F);
        hermite_r(L, p, PC, F, R, Rs);
        const double pref = -nZ[c] * 2.0 * M_PI / p * Kab;
        for (int ka = 0; ka < na; ++ka) {
          int ea[3];
          cart_d(la, ka, ea);
          for (int kb = 0; kb < nb; ++kb) {
            int eb[3];
            cart_d(lb, kb, eb);
            double sum = 0.0;
            for (int tt = 0; tt <= ea[0] + eb[0]; ++tt)
              for (int u = 0; u <= ea[1] + eb[1]; ++u)
                for (int v = 0; v <= ea[2] + eb[2]; ++v)
                  sum += E[0][ea[0]][eb[0]][tt] * E[1][ea[1]][eb[1]][u] * E[2][ea[2]][eb[2]][v] * R[tt][u][v];
            v_acc[ka * nb + kb] += pref * sum;
          }
        }
      }
    }
  }
  const int ao = a.sbf[si], bo = a.sbf[sj];
  const int la_ = a.slbf[si], lb_ = a.slbf[sj];
  const long N = a.sy.n[sys], mb = a.sy.mb[sys];
  for (int ka = 0; ka < na; ++ka)
    for (int kb = 0; kb < nb; ++kb) {
      const double s = a.scale[ao + ka] * a.scale[bo + kb];
      const long i = la_ + ka, j = lb_ + kb;
      const double sv = s * s_acc[ka * nb + kb], tv = s * t_acc[ka * nb + kb],
                   vv = s * v_acc[ka * nb + kb];
      S[mb + i * N + j] = sv;
      S[mb + j * N + i] = sv;
      T[mb + i * N + j] = tv;
      T[mb + j * N + i] = tv;
      if (V) {
        V[mb + i * N + j] = vv;
        V[mb + j * N + i] = vv;
      }
    }
}

}  // namespace gx

namespace gx {

// ---------------------------------------------------------------------------
// One-electron gradient, k_one_grad_t<LA, LB, KSPLIT, VPART>, for the shell
// pairs of class (LA, LB) (positions [p0, p0 + np) of the pair list), with the
// angular momenta as template parameters so that the Hermite tables live in
// registers:
//   g_X += sum_ab P_ab d(T + V)_ab / dX  -  W_ab dS_ab / dX
// with d/dA_i <a|O|b> = 2 alpha <a+1i|O|b> - a_i <a-1i|O|b> (likewise B); S
// and T depend on A - B only (d/dB = -d/dA); the nuclear-attraction term of
// nucleus C moves with C: d/dC = -(d/dA + d/dB).  The kinetic integrals use
// the reference's del-operator expansion on b (factory/primitive_basis_factory.py:89-122).
// One thread per (pair, primitive part): VPART = false computes the overlap
// and kinetic terms, VPART = true the nuclear attraction of the system's
// blockIdx.y-th nucleus (two kernels, each with its own register budget);
// with KSPLIT > 1 the primitives of shell A are dealt to KSPLIT threads
// (more threads for small systems).  Atomics into grad[3 * atom + axis].

// Cartesian exponent of component k of a shell of momentum l along axis x,
// in the reference's order s | x, y, z | xx, yy, zz, xy, xz, yz.
__host__ __device__ constexpr int cart_e(int l, int k, int x) {
  return l == 0 ? 0
         : l == 1 ? (k == x ? 1 : 0)
         : k < 3 ? (k == x ? 2 : 0)
         : k == 3 ? (x < 2 ? 1 : 0)
         : k == 4 ? (x != 1 ? 1 : 0)
         : (x > 0 ? 1 : 0);
}

// 1D Hermite coefficients E[i][j][t], i <= I, j <= J (no exp factor).
template <int I, int J>
__device__ __forceinline__ void herm_e_t(double PA, double PB, double o2p, double (&E)[I + 1][J + 1][I + J + 1]) {
#pragma unroll
  for (int i = 0; i <= I; ++i)
#pragma unroll
    for (int j = 0; j <= J; ++j)
#pragma unroll
      for (int t = 0; t <= I + J; ++t) {
        double v = 0.0;
        if (i == 0 && j == 0) {
          v = (t == 0) ? 1.0 : 0.0;
        } else if (j == 0) {
          if (t <= i - 1) v = PA * E[i - 1][0][t];
          if (t > 0 && t - 1 <= i - 1) v += o2p * E[i - 1][0][t - 1];
          if (t + 1 <= i - 1) v += (t + 1) * E[i - 1][0][t + 1];
        } else {
          if (t <= i + j - 1) v = PB * E[i][j - 1][t];
          if (t > 0 && t - 1 <= i + j - 1) v += o2p * E[i][j - 1][t - 1];
          if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
        }
        E[i][j][t] = v;
      }
}

// Hermite Coulomb integrals R_{tuv} = R^{(0)}_{tuv}, t + u + v <= L.
template <int L>
__device__ __forceinline__ void herm_r_t(double p, const double (&PC)[3], const double (&F)[L + 1],
                                         double (&R)[L + 1][L + 1][L + 1]) {
  double Rn[L + 1][L + 1][L + 1][L + 1];
  double m2p = 1.0;
#pragma unroll
  for (int n = 0; n <= L; ++n) {
    Rn[n][0][0][0] = m2p * F[n];
    m2p *= -2.0 * p;
  }
#pragma unroll
  for (int n = L - 1; n >= 0; --n)
#pragma unroll
    for (int t = 0; t <= L - n; ++t)
#pragma unroll
      for (int u = 0; u <= L - n - t; ++u)
#pragma unroll
        for (int v = 0; v <= L - n - t - u; ++v) {
          if (t + u + v == 0) continue;
          double r;
          if (t > 0) {
            r = PC[0] * Rn[n + 1][t - 1][u][v];
            if (t > 1) r += (t - 1) * Rn[n + 1][t - 2][u][v];
          } else if (u > 0) {
            r = PC[1] * Rn[n + 1][t][u - 1][v];
            if (u > 1) r += (u - 1) * Rn[n + 1][t][u - 2][v];
          } else {
            r = PC[2] * Rn[n + 1][t][u][v - 1];
            if (v > 1) r += (v - 1) * Rn[n + 1][t][u][v - 2];
          }
          Rn[n][t][u][v] = r;
        }
#pragma unroll
  for (int t = 0; t <= L; ++t)
#pragma unroll
    for (int u = 0; u <= L - t; ++u)
#pragma unroll
      for (int v = 0; v <= L - t - u; ++v) R[t][u][v] = Rn[0][t][u][v];
}

// Boys F_0..F_L(T) from the master table (as boys_master, fixed L).
template <int L>
__device__ __forceinline__ void boys_master_t(double T, const double* __restrict__ master, double (&F)[L + 1]) {
  if (T < BOYS_TMAX) {
    const int k0 = (int)(T * BOYS_INV_DT + 0.5);
    const double x = k0 * BOYS_DT - T;
    const double* row = master + k0 * (BOYS_MCOL + 1);
    const double invf[8] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040};
    double f = row[L + 7] * invf[7];
#pragma unroll
    for (int k = 6; k >= 0; --k) f = fma(f, x, row[L + k] * invf[k]);
    double e = 1.0 / 5040.0;
    e = fma(e, x, 1.0 / 720.0);
    e = fma(e, x, 1.0 / 120.0);
    e = fma(e, x, 1.0 / 24.0);
    e = fma(e, x, 1.0 / 6.0);
    e = fma(e, x, 0.5);
    e = fma(e, x, 1.0);
    e = fma(e, x, 1.0);
    e *= row[BOYS_MCOL];
    F[L] = f;
#pragma unroll
    for (int m = L; m >= 1; --m) {
      f = fma(2.0 * T, f, e) * (1.0 / (2 * m - 1));
      F[m - 1] = f;
    }
  } else {
    double f = 0.88622692545275801365 * sqrt(1.0 / T);
    F[0] = f;
#pragma unroll
    for (int m = 0; m < L; ++m) {
      f *= (m + 0.5) / T;
      F[m + 1] = f;
    }
  }
}

template <int LA, int LB, int KSPLIT, bool VPART>
__global__ __launch_bounds__(64) void k_one_grad_t(OneArgs a, int p0, int np, const int* __restrict__ satom,
                                                   const double* __restrict__ P, const double* __restrict__ W,
                                                   double* __restrict__ grad) {
  constexpr int NA = (LA + 1) * (LA + 2) / 2, NB = (LB + 1) * (LB + 2) / 2;
  // per-block sums of the nuclear contributions in LDS (atoms from the first
  // nucleus of the block's first pair's system; others go straight to grad)
  __shared__ double gs[3 * ONE_GRAD_NAT];
  for (int i = threadIdx.x; i < 3 * ONE_GRAD_NAT; i += blockDim.x) gs[i] = 0.0;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int tc = min(t, np * KSPLIT - 1);
  const int pos = p0 + tc / KSPLIT, part = tc % KSPLIT;
  const int slot = VPART ? blockIdx.y + 1 : 0;
  const int si = a.pA[pos], sj = a.pB[pos];
  const int sys = a.ssys[si];
  const int c0 = a.sys_atom ? a.sys_atom[sys] : 0, c1 = a.sys_atom ? a.sys_atom[sys + 1] : 0;
  const int first = p0 + min((int)(blockIdx.x * blockDim.x), np * KSPLIT - 1) / KSPLIT;
  const int base = a.sys_atom ? a.sys_atom[a.ssys[a.pA[first]]] : 0;
  const int npa = a.snp[si], npb = a.snp[sj];
  const bool active = t < np * KSPLIT && !(slot > 0 && c0 + slot - 1 >= c1) && part < npa;
  __syncthreads();
  const double A[3] = {a.ctr1[3 * si], a.ctr1[3 * si + 1], a.ctr1[3 * si + 2]};
  const double B[3] = {a.ctr1[3 * sj], a.ctr1[3 * sj + 1], a.ctr1[3 * sj + 2]};
  const int ao = a.sbf[si], bo = a.sbf[sj];
  const int la_ = a.slbf[si], lb_ = a.slbf[sj];
  const long N = a.sy.n[sys], mb = a.sy.mb[sys];
  const double wpair = (si == sj) ? 1.0 : 2.0;  // both triangles of an off-diagonal block
  double pab[NA * NB], wab[NA * NB];
#pragma unroll
  for (int ka = 0; ka < NA; ++ka)
#pragma unroll
    for (int kb = 0; kb < NB; ++kb) {
      const long i = la_ + ka, j = lb_ + kb;
      const double sc = wpair * a.scale[ao + ka] * a.scale[bo + kb];
      pab[ka * NB + kb] = sc * P[mb + i * N + j];
      wab[ka * NB + kb] = slot == 0 ? sc * W[mb + i * N + j] : 0.0;
    }
  double r2 = 0.0;
#pragma unroll
  for (int x = 0; x < 3; ++x) r2 += (A[x] - B[x]) * (A[x] - B[x]);
  double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0}, gC[3] = {0.0, 0.0, 0.0};
  const int c = c0 + slot - 1;
  for (int ia = part; active && ia < npa; ia += KSPLIT) {
    const double al = a.exps[a.spoff[si] + ia], ca = a.coefs[a.spoff[si] + ia];
    for (int ib = 0; ib < npb; ++ib) {
      const double be = a.exps[a.spoff[sj] + ib], cb = a.coefs[a.spoff[sj] + ib];
      const double p = al + be, o2p = 0.5 / p, mu = al * be / p;
      const double kab = exp(-mu * r2) * ca * cb;
      if (fabs(kab) < ONE_PRIM_EPS) continue;
      double Pc[3], PA[3], PB[3];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        Pc[x] = (al * A[x] + be * B[x]) / p;
        PA[x] = Pc[x] - A[x];
        PB[x] = Pc[x] - B[x];
      }
      if constexpr (!VPART) {
        // overlap and kinetic (depend on A - B only): E_0 of i <= LA + 1, j <= LB + 2
        double Ex[LA + 2][LB + 3][LA + LB + 4], Ey[LA + 2][LB + 3][LA + LB + 4], Ez[LA + 2][LB + 3][LA + LB + 4];
        herm_e_t<LA + 1, LB + 2>(PA[0], PB[0], o2p, Ex);
        herm_e_t<LA + 1, LB + 2>(PA[1], PB[1], o2p, Ey);
        herm_e_t<LA + 1, LB + 2>(PA[2], PB[2], o2p, Ez);
        const double s00 = kab * pow(M_PI / p, 1.5);
        auto S1 = [&](int x, int i, int j) -> double {
          return x == 0 ? Ex[i][j][0] : x == 1 ? Ey[i][j][0] : Ez[i][j][0];
        };
#pragma unroll
        for (int ka = 0; ka < NA; ++ka)
#pragma unroll
          for (int kb = 0; kb < NB; ++kb) {
            const double pw = pab[ka * NB + kb], ww = wab[ka * NB + kb];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
              // <a +- 1_x | S, T | b> for the derivative on A
              double dS = 0.0, dT = 0.0;
#pragma unroll
              for (int sgn = 0; sgn < 2; ++sgn) {
                const int dx = sgn == 0 ? 1 : -1;
                const int eax = cart_e(LA, ka, x) + dx;
                if (eax < 0) continue;
                const double w = sgn == 0 ? 2.0 * al : -(double)cart_e(LA, ka, x);
                int ea[3], eb[3];
#pragma unroll
                for (int y = 0; y < 3; ++y) {
                  ea[y] = cart_e(LA, ka, y) + (y == x ? dx : 0);
                  eb[y] = cart_e(LB, kb, y);
                }
                const double sx = S1(0, ea[0], eb[0]), sy = S1(1, ea[1], eb[1]), sz = S1(2, ea[2], eb[2]);
                const int lbt = eb[0] + eb[1] + eb[2];
                double tv = be * (2 * lbt + 3) * sx * sy * sz;
                tv -= 2.0 * be * be * (S1(0, ea[0], eb[0] + 2) * sy * sz + sx * S1(1, ea[1], eb[1] + 2) * sz +
                                       sx * sy * S1(2, ea[2], eb[2] + 2));
                if (eb[0] >= 2) tv -= 0.5 * eb[0] * (eb[0] - 1) * S1(0, ea[0], eb[0] - 2) * sy * sz;
                if (eb[1] >= 2) tv -= 0.5 * eb[1] * (eb[1] - 1) * sx * S1(1, ea[1], eb[1] - 2) * sz;
                if (eb[2] >= 2) tv -= 0.5 * eb[2] * (eb[2] - 1) * sx * sy * S1(2, ea[2], eb[2] - 2);
                dS += w * sx * sy * sz;
                dT += w * tv;
              }
              const double g = s00 * (pw * dT - ww * dS);
              gA[x] += g;
              gB[x] -= g;
            }
          }
      } else {
        // nuclear attraction of nucleus c: E of i <= LA + 1, j <= LB + 1, R up to LA + LB + 1
        constexpr int L = LA + LB + 1;
        double Ex[LA + 2][LB + 2][LA + LB + 3], Ey[LA + 2][LB + 2][LA + LB + 3], Ez[LA + 2][LB + 2][LA + LB + 3];
        herm_e_t<LA + 1, LB + 1>(PA[0], PB[0], o2p, Ex);
        herm_e_t<LA + 1, LB + 1>(PA[1], PB[1], o2p, Ey);
        herm_e_t<LA + 1, LB + 1>(PA[2], PB[2], o2p, Ez);
        const double PC[3] = {Pc[0] - a.nxyz[3 * c], Pc[1] - a.nxyz[3 * c + 1], Pc[2] - a.nxyz[3 * c + 2]};
        double F[L + 1];
        boys_master_t<L>(p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), a.master, F);
        double R[L + 1][L + 1][L + 1];
        herm_r_t<L>(p, PC, F, R);
        const double pref = -a.nZ[c] * 2.0 * M_PI / p * kab;
        auto Vval = [&](const int (&ea)[3], const int (&eb)[3]) -> double {
          double sum = 0.0;
#pragma unroll
          for (int tt = 0; tt <= LA + LB + 2; ++tt) {
            if (tt > ea[0] + eb[0]) break;
#pragma unroll
            for (int u = 0; u <= LA + LB + 2; ++u) {
              if (u > ea[1] + eb[1]) break;
#pragma unroll
              for (int v = 0; v <= LA + LB + 2; ++v) {
                if (v > ea[2] + eb[2]) break;
                sum += Ex[ea[0]][eb[0]][tt] * Ey[ea[1]][eb[1]][u] * Ez[ea[2]][eb[2]][v] * R[tt][u][v];
              }
            }
          }
          return sum;
        };
#pragma unroll
        for (int ka = 0; ka < NA; ++ka)
#pragma unroll
          for (int kb = 0; kb < NB; ++kb) {
            const double pw = pref * pab[ka * NB + kb];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
              int e0[3], f0[3];
#pragma unroll
              for (int y = 0; y < 3; ++y) {
                e0[y] = cart_e(LA, ka, y);
                f0[y] = cart_e(LB, kb, y);
              }
              int ua[3] = {e0[0], e0[1], e0[2]}, ub[3] = {f0[0], f0[1], f0[2]};
              ua[x] += 1;
              ub[x] += 1;
              double dA = 2.0 * al * Vval(ua, f0), dB = 2.0 * be * Vval(e0, ub);
              if (e0[x] > 0) {
                int da[3] = {e0[0], e0[1], e0[2]};
                da[x] -= 1;
                dA -= e0[x] * Vval(da, f0);
              }
              if (f0[x] > 0) {
                int db[3] = {f0[0], f0[1], f0[2]};
                db[x] -= 1;
                dB -= f0[x] * Vval(e0, db);
              }
              gA[x] += pw * dA;
              gB[x] += pw * dB;
              gC[x] -= pw * (dA + dB);
            }
          }
      }
    }
  }
  auto add = [&](int atom, int x, double v) {
    if (v == 0.0) return;
    const int i = atom - base;
    if (i >= 0 && i < ONE_GRAD_NAT) atomicAdd(&gs[3 * i + x], v);
    else atomicAdd(&grad[3 * atom + x], v);
  };
  if (active) {
    const int atA = satom[si], atB = satom[sj];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      if (VPART) add(c, x, gC[x]);
      add(atA, x, gA[x]);
      add(atB, x, gB[x]);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * ONE_GRAD_NAT; i += blockDim.x)
    if (gs[i] != 0.0) atomicAdd(&grad[3 * base + i], gs[i]);
}

// Nuclear repulsion gradient  dE_nuc/dR_A = -sum_B Z_A Z_B (R_A - R_B) / |R_A - R_B|^3
// over the nuclei B of A's system (common/nuclear_repulsion.py:5-24 of the
// reference gives E_nuc), added to grad.
__global__ void k_nuc_grad(int natom, const int* __restrict__ atom_sys, const int* __restrict__ sys_atom,
                           const double* __restrict__ nxyz, const double* __restrict__ nZ,
                           double* __restrict__ grad) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= natom) return;
  const int s = atom_sys[a];
  double g[3] = {0.0, 0.0, 0.0};
  for (int b = sys_atom[s]; b < sys_atom[s + 1]; ++b) {
    if (b == a) continue;
    const double d[3] = {nxyz[3 * a] - nxyz[3 * b], nxyz[3 * a + 1] - nxyz[3 * b + 1], nxyz[3 * a + 2] - nxyz[3 * b + 2]};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double f = nZ[a] * nZ[b] / (r2 * sqrt(r2));
    for (int x = 0; x < 3; ++x) g[x] -= f * d[x];
  }
  for (int x = 0; x < 3; ++x) grad[3 * a + x] += g[x];
}

}  // namespace gx
