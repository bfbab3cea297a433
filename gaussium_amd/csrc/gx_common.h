// Shared device definitions for the gx ERI / J-K engine (gfx950, wave64).
//
// Layout in HBM (all arrays owned by a gx context):
//   shells   : per shell  l, nprim, prim offset, first basis-function index,
//              centre (ERI) and centre (one-electron)
//   pairs    : per shell pair (A,B) with l_A >= l_B: A, B, primitive-pair
//              offset/count, Schwarz bound Q_AB; pairs are grouped by class
//              (l_A,l_B) and, inside a class, sorted by primitive-pair count
//              then by Q (descending)
//   primpairs: per primitive pair one 64-byte FP64 record and one 32-byte
//              FP32 record (PrimPair64 / PrimPair32)
//   boys     : master tables F_m(T_n), m <= BOYS_MCOL - 1, plus exp(-T_n),
//              on the fine (T_n = n / 10) and coarse (T_n = n / 4) grids;
//              per-L Taylor tables F_{L+k}(T_n)/k! of both (BoysGrid),
//              which kernels read or stage into LDS
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GX_WAVE 64

// Copies of the pair-indexed J_cd accumulator (see jkp_copy in gx_eri.hip).
#ifndef GX_JT_COPIES
#define GX_JT_COPIES 8
#endif

// Boys-function grids.  Master tables F_m(T_n), m < BOYS_MCOL, plus
// exp(-T_n), on two grids: T_n = n * BOYS_DT (the one-electron kernels read
// it directly; the per-L Taylor tables of the fine grid below come from it)
// and T_n = n / 4 (the coarse grid's Taylor tables).
#define BOYS_DT 0.1
#define BOYS_INV_DT 10.0
#define BOYS_NT 361          // rows T_n = 0, 0.1, .., 36.0 (nearest row of any T < 36)
#define BOYS_TMAX 36.0
#define BOYS_MCOL 20         // F_0 .. F_19 stored (L + Taylor terms <= 20)
#define BOYS_TAYLOR64 8      // FP64: 8 Taylor terms, |dT| <= 0.05 -> < 1e-15 rel.
#define BOYS_TAYLOR32 5      // FP32: 5 Taylor terms -> < 3e-9 rel.
#define BOYS_ROW64 10        // LDS row: 8 coefficients, exp(-T_n), pad
#define BOYS_ROW32 8         // LDS row: 5 coefficients, pad

namespace gx {

// Taylor-table grid: rows T_n = n / TPU (n < NT, T_n up to BOYS_TMAX), each
// F_{L+k}(T_n) / k! for k < K64 (then exp(-T_n), padded to 16 bytes) in
// FP64 and k < K32 in FP32.  A Taylor remainder of F_L is at most
// e^h h^K / K! relative (h = 1 / (2 TPU) the largest |T - T_n|), as
// |F_{L+K}(T')| <= e^h F_L(T) for |T' - T| <= h; the same bound holds for
// exp(-T) = exp(-T_n) exp(T_n - T).
//   BoysFine   (TPU 10, K 8 / 5): < 1e-15 / 3e-9, 28.9 KB per L in FP64
//   BoysCoarse (TPU 4, K 10 / 7): < 3e-16 / 1.1e-10, 13.9 KB per L
// The coarse grid costs more arithmetic per Boys value and takes less
// LDS: the J/K kernels of the classes whose LDS (table plus density rows)
// rather than registers bounds the occupancy use it (jk_coarse_grid).
template <int TPU_, int K64_, int K32_>
struct BoysGrid {
  static constexpr int TPU = TPU_;
  static constexpr double DT = 1.0 / TPU_;
  static constexpr int NT = 36 * TPU_ + 1;
  static constexpr int K64 = K64_, K32 = K32_;
  static constexpr int ROW64 = (K64_ + 2) & ~1;
  static constexpr int ROW32 = (K32_ + 3) & ~3;
};
using BoysFine = BoysGrid<10, BOYS_TAYLOR64, BOYS_TAYLOR32>;
using BoysCoarse = BoysGrid<4, 10, 7>;
static_assert(BoysFine::NT == BOYS_NT && BoysFine::ROW64 == BOYS_ROW64 && BoysFine::ROW32 == BOYS_ROW32,
              "fine grid = master grid");


// One primitive pair (a on A, b on B) of an oriented shell pair (A, B):
//   p = a + b, ip = 1/p, P the Gaussian product centre, PA = P - A,
//   Kip = sqrt(2 pi^(5/2)) c_a c_b exp(-a b / p |AB|^2) / p,  ea = a
// so that a primitive quartet's prefactor is Kip_bra Kip_ket / sqrt(p + q);
// ea (and p - ea) weight the raised classes of the nuclear gradient.
struct __align__(16) PrimPair64 {
  double p, ip, Px, Py, Pz, Kip, PAx, PAy, PAz, ea;
};

struct __align__(16) PrimPair32 {
  float p, ip, Px, Py, Pz, Kip, PAx, PAy, PAz, ea, pad1, pad2;
};

// Device-side views passed to kernels by value.
struct ShellView {
  const int* l;
  const int* bf;        // first basis function (index over all systems)
  const double* ctr;    // ERI centre, 3 per shell
  const int* lbf;       // first basis function inside its system
  const int* sys;       // system of the shell
};

// Systems of a batch: each has its own nbf x nbf matrices, packed one after
// the other (P, J, K, S, T, V ...); quartets never mix systems.
struct SysView {
  const int* n;         // basis functions per system
  const long* mb;       // offset of the system's matrix in the packed arrays
};

// One shell pair (A, B) in a single 64-byte record, so that a kernel gets
// everything it needs about a ket (or bra) pair with one dependent load:
// A - B, the first basis functions of A and B inside their system, the
// primitive-pair range, the shells, the system and the pair's first entry
// in the pair-indexed J accumulator.
struct __align__(16) PairRec {
  double ABx, ABy, ABz;
  int fa, fb;    // first basis function of A and of B inside their system
  int poff, np;  // primitive pairs [poff, poff + np)
  int A, B;      // shells
  int sys, jko;  // system; entry of component 0 in the pair-indexed J accumulator
};

struct PairView {
  const int* A;
  const int* B;
  const int* poff;      // primitive-pair offset
  const int* np;        // primitive-pair count
  const double* Q;      // Schwarz bound (scaled)
  const PrimPair64* pp64;
  const PrimPair32* pp32;
  const PairRec* rec;   // per pair (J/K and gradient kernels)
};

// Ket lists of the work items: entry = shell-pair id, GX_KET_FP32 set when the
// quartet runs in FP32, GX_KET_NONE for the idle lanes between a bra's FP64
// and FP32 kets.
#define GX_KET_FP32 0x40000000
#define GX_KET_ID 0x3fffffff
#define GX_KET_NONE (-1)

// Reciprocal and reciprocal square root from the hardware estimates (about
// 2^-24 relative on gfx950, measured) plus Newton steps (quadratic
// convergence): two steps reach full FP64 precision for the reciprocal; one
// step leaves the reciprocal square root within ~2^-47 (4e-15) relative,
// far below the 1e-10 absolute bar of the integrals it scales.
__device__ __forceinline__ double gx_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// Side streams for the class-combination kernels of one build: with the
// caller's stream, the four hardware queues a process gets by default.
#ifndef GX_NSIDE
#define GX_NSIDE 3
#endif

#ifndef GX_RSQ_NR
#define GX_RSQ_NR 1  // Newton steps after the hardware estimate
#endif

__device__ __forceinline__ double gx_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  // y <- y (3 - x y^2) / 2
  const double h = 0.5 * x;
#pragma unroll
  for (int k = 0; k < GX_RSQ_NR; ++k) {
    const double t = fma(-h * y, y, 0.5);
    y = fma(y, t, y);
  }
  return y;
}

__device__ __forceinline__ float gx_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float gx_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

// 1/k! (k <= 13), folded to constants in unrolled loops.
__host__ __device__ constexpr double gx_inv_fact(int k) {
  double f = 1.0;
  for (int i = 2; i <= k; ++i) f *= i;
  return 1.0 / f;
}

// exp(x), |x| <= 1 / (2 BG::TPU), by its Taylor series with BG::K64 terms.
template <class BG>
__device__ __forceinline__ double boys_exp_taylor(double x) {
  double e = gx_inv_fact(BG::K64 - 1);
#pragma unroll
  for (int k = BG::K64 - 2; k >= 0; --k) e = fma(e, x, gx_inv_fact(k));
  return e;
}

// ---------------------------------------------------------------------------
// Boys function F_0..F_L at T from a Taylor table of the fine grid.
//   T < BOYS_TMAX : F_L by Taylor about the nearest grid point, exp(-T) from
//                   exp(-T_n) times a Taylor series, then the downward
//                   recursion F_{m-1} = (2T F_m + e^-T) / (2m - 1)
//   T >= BOYS_TMAX: F_0 = sqrt(pi / T) / 2 (relative error erfc(sqrt T) < 3e-17)
//                   and the upward recursion F_{m+1} = (2m + 1) F_m / (2T); the
//                   dropped e^-T / (2T) term leaves every F_m within 4e-18
//                   absolute of the exact value
// Every F_m is multiplied by `scale` (the quartet prefactor).
template <int L>
__device__ __forceinline__ void boys64(double T, double scale, const double* __restrict__ tab,
                                       double (&F)[L + 1]) {
  if (T < BOYS_TMAX) {
    const int n = (int)(T * BOYS_INV_DT + 0.5);
    const double x = n * BOYS_DT - T;  // -(T - T_n)
    const double* row = tab + n * BOYS_ROW64;
    double f = row[BOYS_TAYLOR64 - 1];
#pragma unroll
    for (int k = BOYS_TAYLOR64 - 2; k >= 0; --k) f = fma(f, x, row[k]);
    if (L == 0) {
      F[0] = f * scale;
      return;
    }
    // exp(-T) = exp(-T_n) * exp(x)
    const double e = boys_exp_taylor<BoysFine>(x) * row[BOYS_TAYLOR64];
    F[L] = f * scale;
    const double t2 = 2.0 * T;
#pragma unroll
    for (int m = L; m >= 1; --m) {
      f = fma(t2, f, e) * (1.0 / (2 * m - 1));
      F[m - 1] = f * scale;
    }
  } else {
    const double rt = gx_rsqrt(T);
    const double it = rt * rt;
    double f = 0.88622692545275801365 * rt;  // sqrt(pi)/2 / sqrt(T)
    F[0] = f * scale;
#pragma unroll
    for (int m = 0; m < L; ++m) {
      f = f * ((m + 0.5) * it);
      F[m + 1] = f * scale;
    }
  }
}

template <int L>
__device__ __forceinline__ void boys32(float T, float scale, const float* __restrict__ tab,
                                       float (&F)[L + 1]) {
  if (T < (float)BOYS_TMAX) {
    const int n = (int)(T * (float)BOYS_INV_DT + 0.5f);
    const float x = n * (float)BOYS_DT - T;
    const float* row = tab + n * BOYS_ROW32;
    float f = row[BOYS_TAYLOR32 - 1];
#pragma unroll
    for (int k = BOYS_TAYLOR32 - 2; k >= 0; --k) f = fmaf(f, x, row[k]);
    F[L] = f * scale;
    if (L > 0) {
      const float e = __expf(-T);
      const float t2 = 2.0f * T;
#pragma unroll
      for (int m = L; m >= 1; --m) {
        f = fmaf(t2, f, e) * (1.0f / (2 * m - 1));
        F[m - 1] = f * scale;
      }
    }
  } else {
    const float rt = gx_rsqrt(T);
    const float it = rt * rt;
    float f = 0.886226925f * rt;
    F[0] = f * scale;
#pragma unroll
    for (int m = 0; m < L; ++m) {
      f = f * ((m + 0.5f) * it);
      F[m + 1] = f * scale;
    }
  }
}

// Branch-free variants on grid BG: both the table and the asymptotic forms
// are evaluated (index clamped to the table) and one is selected, so that a
// loop over several primitive quartets stays a single basic block and the
// table gathers of all of them can be in flight together.  The nearest grid
// point comes from adding 2^52 (2^23 in FP32): the sum's low mantissa bits
// are rint(T TPU), with no float-integer conversions; the row offset is a
// 24-bit multiply.
template <int L, class BG = BoysFine>
__device__ __forceinline__ void boys64_nb(double T, double scale, const double* __restrict__ tab,
                                          double (&F)[L + 1]) {
  const bool small = T < BOYS_TMAX;
  const double Tc = fmin(T, BOYS_TMAX);  // row BG::NT - 1 for T >= BOYS_TMAX (value unused)
  const double tn = fma(Tc, (double)BG::TPU, 4503599627370496.0);  // 2^52 + rint(Tc TPU)
  const int n = __double2loint(tn);
  const double x = fma(tn - 4503599627370496.0, BG::DT, -Tc);  // -(T - T_n)
  const double* row = tab + __mul24(n, BG::ROW64);
  double f = row[BG::K64 - 1];
#pragma unroll
  for (int k = BG::K64 - 2; k >= 0; --k) f = fma(f, x, row[k]);
  const double rt = gx_rsqrt(fmax(T, BOYS_TMAX));  // used only for T >= BOYS_TMAX
  const double it = rt * rt;
  double g = 0.88622692545275801365 * rt;  // asymptotic F_0
  if (L == 0) {
    F[0] = (small ? f : g) * scale;
    return;
  }
  const double e = boys_exp_taylor<BG>(x) * row[BG::K64];  // exp(-T_n) exp(x)
  double Fs[L + 1], Fa[L + 1];
  Fs[L] = f;
  const double t2 = 2.0 * Tc;
#pragma unroll
  for (int m = L; m >= 1; --m) {
    f = fma(t2, f, e) * (1.0 / (2 * m - 1));
    Fs[m - 1] = f;
  }
  Fa[0] = g;
#pragma unroll
  for (int m = 0; m < L; ++m) {
    g = g * ((m + 0.5) * it);
    Fa[m + 1] = g;
  }
#pragma unroll
  for (int m = 0; m <= L; ++m) F[m] = (small ? Fs[m] : Fa[m]) * scale;
}

template <int L, class BG = BoysFine>
__device__ __forceinline__ void boys32_nb(float T, float scale, const float* __restrict__ tab,
                                          float (&F)[L + 1]) {
  const bool small = T < (float)BOYS_TMAX;
  const float Tc = fminf(T, (float)BOYS_TMAX);
  const float tn = fmaf(Tc, (float)BG::TPU, 8388608.0f);  // 2^23 + rint(Tc TPU)
  const int n = __float_as_int(tn) - 0x4b000000;
  const float x = fmaf(tn - 8388608.0f, (float)BG::DT, -Tc);
  const float* row = tab + __mul24(n, BG::ROW32);
  float f = row[BG::K32 - 1];
#pragma unroll
  for (int k = BG::K32 - 2; k >= 0; --k) f = fmaf(f, x, row[k]);
  const float rt = gx_rsqrt(fmaxf(T, (float)BOYS_TMAX));
  const float it = rt * rt;
  float g = 0.886226925f * rt;
  if (L == 0) {
    F[0] = (small ? f : g) * scale;
    return;
  }
  const float e = __expf(-Tc);
  float Fs[L + 1], Fa[L + 1];
  Fs[L] = f;
  const float t2 = 2.0f * Tc;
#pragma unroll
  for (int m = L; m >= 1; --m) {
    f = fmaf(t2, f, e) * (1.0f / (2 * m - 1));
    Fs[m - 1] = f;
  }
  Fa[0] = g;
#pragma unroll
  for (int m = 0; m < L; ++m) {
    g = g * ((m + 0.5f) * it);
    Fa[m + 1] = g;
  }
#pragma unroll
  for (int m = 0; m <= L; ++m) F[m] = (small ? Fs[m] : Fa[m]) * scale;
}

// Per-L Taylor table of grid BG from the master table of that grid (layout
// master[n * (BOYS_MCOL + 1) + m] = F_m(T_n), column BOYS_MCOL = exp(-T_n)):
// row n holds F_{L+k}(T_n) / k!, k < BG::K64, then exp(-T_n) (FP64); or
// F_{L+k}(T_n) / k!, k < BG::K32 (FP32).
template <class BG = BoysFine>
__device__ __forceinline__ void boys_stage64(int L, const double* __restrict__ master,
                                             double* __restrict__ out) {
  for (int i = threadIdx.x; i < BG::NT * BG::ROW64; i += blockDim.x) {
    const int n = i / BG::ROW64, k = i - n * BG::ROW64;
    const double* src = master + n * (BOYS_MCOL + 1);
    double v = 0.0;
    if (k < BG::K64) v = src[L + k] * gx_inv_fact(k);
    else if (k == BG::K64) v = src[BOYS_MCOL];
    out[i] = v;
  }
}

template <class BG = BoysFine>
__device__ __forceinline__ void boys_stage32(int L, const double* __restrict__ master,
                                             float* __restrict__ out) {
  for (int i = threadIdx.x; i < BG::NT * BG::ROW32; i += blockDim.x) {
    const int n = i / BG::ROW32, k = i - n * BG::ROW32;
    const double* src = master + n * (BOYS_MCOL + 1);
    out[i] = k < BG::K32 ? (float)(src[L + k] * gx_inv_fact(k)) : 0.0f;
  }
}

// Copy a per-L table of grid BG from HBM into LDS.
template <class BG>
__device__ __forceinline__ void boys_copy64(const double* __restrict__ src, double* __restrict__ lds) {
  const double2* s2 = reinterpret_cast<const double2*>(src);
  double2* d2 = reinterpret_cast<double2*>(lds);
  for (int i = threadIdx.x; i < BG::NT * BG::ROW64 / 2; i += blockDim.x) d2[i] = s2[i];
}
template <class BG>
__device__ __forceinline__ void boys_copy32(const float* __restrict__ src, float* __restrict__ lds) {
  const float4* s4 = reinterpret_cast<const float4*>(src);
  float4* d4 = reinterpret_cast<float4*>(lds);
  for (int i = threadIdx.x; i < BG::NT * BG::ROW32 / 4; i += blockDim.x) d4[i] = s4[i];
}

}  // namespace gx
