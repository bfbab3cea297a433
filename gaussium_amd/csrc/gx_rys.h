// Rys quadrature rule of a primitive quartet (tables from gen_rys.py):
// F_m(T) = sum_i w_i u_i^m for m < 2n.  Shared by the device kernels and the
// host test harness (which defines __device__ away).
#pragma once
#include "rys_tab.h"  // generated by gen_rys.py (include path: the generated-code folder)

// Quantity q of the n-point rule at T: the root u_q for q < n, the weight
// w_{q-n} for n <= q < 2n (roots ascending).  T < RYS_TMAX: Chebyshev series
// on the unit interval [k, k + 1) holding T (Clenshaw); T >= RYS_TMAX: the
// large-T rule u = s / T, w = l / (2 sqrt(T)).
__device__ __forceinline__ double rys_quantity(int n, double T, int q) {
  if (T < (double)RYS_TMAX) {
    const int k = (int)T;
    const double x = 2.0 * (T - k) - 1.0;
    const double* c = gx_rys_cheb + gx_rys_cheb_off[n - 1] + ((long)k * 2 * n + q) * RYS_NCHEB;
    double b1 = 0.0, b2 = 0.0;
    for (int m = RYS_NCHEB - 1; m >= 1; --m) {
      const double t = fma(2.0 * x, b1, c[m] - b2);
      b2 = b1;
      b1 = t;
    }
    return fma(x, b1, c[0] - b2);
  }
  const int a = n * (n - 1) / 2;
  if (q < n) return gx_rys_asym_s[a + q] / T;
  return gx_rys_asym_l[a + q - n] * 0.5 / sqrt(T);
}
