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
// One-electron part of the RHF nuclear gradient (the reference has no
// gradient; its optimiser differentiates energies numerically,
// geometryoptimization/nelder_mead.py:126-144):
//   g_X += sum_ab P_ab d(T + V)_ab / dX  -  W_ab dS_ab / dX
// with d/dA_i <a|O|b> = 2 alpha <a+1i|O|b> - a_i <a-1i|O|b> (likewise B);
// S and T depend on A - B only (d/dB = -d/dA); the nuclear-attraction term of
// nucleus C moves with C:  d/dC = -(d/dA + d/dB).
// One thread per shell pair (i >= j); atomics into grad[3 * atom + axis].
#define ONE_GRAD_I (ONE_LMAX + 1)      // bra raised by 1
#define ONE_GRAD_J (ONE_LMAX + 2)      // ket raised by 2 (kinetic) or 1
#define ONE_GRAD_T (ONE_GRAD_I + ONE_GRAD_J)

__device__ inline void hermite_e_g(int imax, int jmax, double PA, double PB, double o2p,
                                   double (*E)[ONE_GRAD_J + 1][ONE_GRAD_T + 1]) {
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j)
      for (int t = 0; t <= i + j; ++t) E[i][j][t] = 0.0;
  E[0][0][0] = 1.0;
  for (int i = 0; i <= imax; ++i) {
    if (i > 0) {
      for (int t = 0; t <= i; ++t) {
        double v = (t <= i - 1) ? PA * E[i - 1][0][t] : 0.0;
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

__global__ __launch_bounds__(64) void k_one_grad(OneArgs a, const int* __restrict__ satom,
                                                 const double* __restrict__ P, const double* __restrict__ W,
                                                 double* __restrict__ grad) {
  // grid: x over shell pairs, y over the parts of one pair -- 0: overlap and
  // kinetic, y >= 1: the nuclear attraction of the system's (y - 1)-th nucleus
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.npair) return;
  const int slot = blockIdx.y;
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
  if (slot > 0 && c0 + slot - 1 >= c1) return;
  const int la = sl[si], lb = sl[sj];
  const int na = ncart_d(la), nb = ncart_d(lb);
  const double A[3] = {ctr1[3 * si], ctr1[3 * si + 1], ctr1[3 * si + 2]};
  const double B[3] = {ctr1[3 * sj], ctr1[3 * sj + 1], ctr1[3 * sj + 2]};
  const int ao = a.sbf[si], bo = a.sbf[sj];
  const int la_ = a.slbf[si], lb_ = a.slbf[sj];
  const long N = a.sy.n[sys], mb = a.sy.mb[sys];
  const double wpair = (si == sj) ? 1.0 : 2.0;  // both triangles of an off-diagonal block
  double pab[36], wab[36];
  for (int ka = 0; ka < na; ++ka)
    for (int kb = 0; kb < nb; ++kb) {
      const long i = la_ + ka, j = lb_ + kb;
      const double s = wpair * a.scale[ao + ka] * a.scale[bo + kb];
      pab[ka * nb + kb] = s * P[mb + i * N + j];
      wab[ka * nb + kb] = s * W[mb + i * N + j];
    }
  double E[3][ONE_GRAD_I + 1][ONE_GRAD_J + 1][ONE_GRAD_T + 1];
  double R[ONE_RDIM][ONE_RDIM][ONE_RDIM], Rs[ONE_RDIM][ONE_RDIM][ONE_RDIM];
  double F[ONE_RDIM + 1];
  double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
  // overlap and kinetic: depend on A - B only
  for (int ia = 0; slot == 0 && ia < snp[si]; ++ia) {
    const double a = exps[spoff[si] + ia], ca = coefs[spoff[si] + ia];
    for (int ib = 0; ib < snp[sj]; ++ib) {
      const double b = exps[spoff[sj] + ib], cb = coefs[spoff[sj] + ib];
      const double p = a + b, o2p = 0.5 / p, mu = a * b / p;
      double Pc[3], r2 = 0.0;
      for (int x = 0; x < 3; ++x) r2 += (A[x] - B[x]) * (A[x] - B[x]);
      const double kab = exp(-mu * r2) * ca * cb;
      if (fabs(kab) < ONE_PRIM_EPS) continue;
      for (int x = 0; x < 3; ++x) {
        Pc[x] = (a * A[x] + b * B[x]) / p;
        hermite_e_g(la + 1, lb + 2, Pc[x] - A[x], Pc[x] - B[x], o2p, E[x]);
      }
      const double s00 = kab * pow(M_PI / p, 1.5);
      auto S1 = [&](int x, int i, int j) { return E[x][i][j][0]; };
      // <a|T|b> for exponent triples ea, eb (the reference's del-operator expansion on b)
      auto Tval = [&](const int* ea, const int* eb) {
        const double sx = S1(0, ea[0], eb[0]), sy = S1(1, ea[1], eb[1]), sz = S1(2, ea[2], eb[2]);
        const int lbt = eb[0] + eb[1] + eb[2];
        double tv = b * (2 * lbt + 3) * sx * sy * sz;
        tv -= 2.0 * b * b * (S1(0, ea[0], eb[0] + 2) * sy * sz + sx * S1(1, ea[1], eb[1] + 2) * sz +
                             sx * sy * S1(2, ea[2], eb[2] + 2));
        if (eb[0] >= 2) tv -= 0.5 * eb[0] * (eb[0] - 1) * S1(0, ea[0], eb[0] - 2) * sy * sz;
        if (eb[1] >= 2) tv -= 0.5 * eb[1] * (eb[1] - 1) * sx * S1(1, ea[1], eb[1] - 2) * sz;
        if (eb[2] >= 2) tv -= 0.5 * eb[2] * (eb[2] - 1) * sx * sy * S1(2, ea[2], eb[2] - 2);
        return tv;
      };
      auto Sval = [&](const int* ea, const int* eb) {
        return S1(0, ea[0], eb[0]) * S1(1, ea[1], eb[1]) * S1(2, ea[2], eb[2]);
      };
      for (int ka = 0; ka < na; ++ka) {
        int ea[3];
        cart_d(la, ka, ea);
        for (int kb = 0; kb < nb; ++kb) {
          int eb[3];
          cart_d(lb, kb, eb);
          const double pw = pab[ka * nb + kb], ww = wab[ka * nb + kb];
          for (int x = 0; x < 3; ++x) {
            int up[3] = {ea[0], ea[1], ea[2]};
            up[x] += 1;
            double dS = 2.0 * a * Sval(up, eb), dT = 2.0 * a * Tval(up, eb);
            if (ea[x] > 0) {
              int dn[3] = {ea[0], ea[1], ea[2]};
              dn[x] -= 1;
              dS -= ea[x] * Sval(dn, eb);
              dT -= ea[x] * Tval(dn, eb);
            }
            const double g = s00 * (pw * dT - ww * dS);
            gA[x] += g;
            gB[x] -= g;
          }
        }
      }
    }
  }
  // nuclear attraction of one nucleus
  const int L = la + lb + 1;
  for (int c = c0 + slot - 1; slot > 0 && c == c0 + slot - 1; ++c) {
    double gC[3] = {0.0, 0.0, 0.0};
    for (int ia = 0; ia < snp[si]; ++ia) {
      const double a = exps[spoff[si] + ia], ca = coefs[spoff[si] + ia];
      for (int ib = 0; ib < snp[sj]; ++ib) {
        const double b = exps[spoff[sj] + ib], cb = coefs[spoff[sj] + ib];
        const double p = a + b, o2p = 0.5 / p, mu = a * b / p;
        double Pc[3], r2 = 0.0;
        for (int x = 0; x < 3; ++x) r2 += (A[x] - B[x]) * (A[x] - B[x]);
        const double kab = exp(-mu * r2) * ca * cb;
        if (fabs(kab) < ONE_PRIM_EPS) continue;
        for (int x = 0; x < 3; ++x) {
          Pc[x] = (a * A[x] + b * B[x]) / p;
          hermite_e_g(la + 1, lb + 1, Pc[x] - A[x], Pc[x] - B[x], o2p, E[x]);
        }
        const double PC[3] = {Pc[0] - nxyz[3 * c], Pc[1] - nxyz[3 * c + 1], Pc[2] - nxyz[3 * c + 2]};
        const double Tb = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
        boys_master(L, Tb, master, F);
        hermite_r(L, p, PC, F, R, Rs);
        const double pref = -nZ[c] * 2.0 * M_PI / p * kab;
        auto Vval = [&](const int* ea, const int* eb) {
          double sum = 0.0;
          for (int tt = 0; tt <= ea[0] + eb[0]; ++tt)
            for (int u = 0; u <= ea[1] + eb[1]; ++u)
              for (int v = 0; v <= ea[2] + eb[2]; ++v)
                sum += E[0][ea[0]][eb[0]][tt] * E[1][ea[1]][eb[1]][u] * E[2][ea[2]][eb[2]][v] * R[tt][u][v];
          return sum;
        };
        for (int ka = 0; ka < na; ++ka) {
          int ea[3];
          cart_d(la, ka, ea);
          for (int kb = 0; kb < nb; ++kb) {
            int eb[3];
            cart_d(lb, kb, eb);
            const double pw = pref * pab[ka * nb + kb];
            for (int x = 0; x < 3; ++x) {
              int ua[3] = {ea[0], ea[1], ea[2]}, ub[3] = {eb[0], eb[1], eb[2]};
              ua[x] += 1;
              ub[x] += 1;
              double dA = 2.0 * a * Vval(ua, eb), dB = 2.0 * b * Vval(ea, ub);
              if (ea[x] > 0) {
                int da[3] = {ea[0], ea[1], ea[2]};
                da[x] -= 1;
                dA -= ea[x] * Vval(da, eb);
              }
              if (eb[x] > 0) {
                int db[3] = {eb[0], eb[1], eb[2]};
                db[x] -= 1;
                dB -= eb[x] * Vval(ea, db);
              }
              gA[x] += pw * dA;
              gB[x] += pw * dB;
              gC[x] -= pw * (dA + dB);
            }
          }
        }
      }
    }
    for (int x = 0; x < 3; ++x)
      if (gC[x] != 0.0) atomicAdd(&grad[3 * c + x], gC[x]);
  }
  const int atA = satom[si], atB = satom[sj];
  for (int x = 0; x < 3; ++x) {
    atomicAdd(&grad[3 * atA + x], gA[x]);
    atomicAdd(&grad[3 * atB + x], gB[x]);
  }
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
