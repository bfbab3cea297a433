// C ABI of the gx engine (declared in include/gx.h): context, basis set-up,
// pair lists, launches.  Host bookkeeping is O(nshell^2); all floating-point
// integral work runs on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/gx.h"
#include "gx_common.h"
#include "gx_plan.h"
#include "_gen/eri_gen.h"
// Single translation unit: the kernels are templates instantiated below.
#include "gx_eri.hip"
#include "gx_coop.hip"
#include "gx_grad.hip"
#include "gx_rys.hip"
#include "gx_one.hip"

// Device scratch of gx_jk_noalloc: the scaled density, Kt, Jt and the
// GX_JT_COPIES copies of the pair-indexed J_cd accumulator of one build,
// 3 sum_k n_k^2 + GX_JT_COPIES sum_pairs n_A n_B doubles
// at most GX_NOALLOC_MAX_DOUBLES (5 MiB: one system of up to 256 functions).
// Every device the library runs on holds its own copy until the process
// ends; its address is taken once per device by gx_create.
#define GX_NOALLOC_MAX_DOUBLES (10L * 256 * 256)
static __device__ __align__(256) double scratch[GX_NOALLOC_MAX_DOUBLES];

namespace {

thread_local std::string g_err;

// Per-device address of `scratch` (hipGetSymbolAddress with the device
// current); host code uses these, never the array's name.
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;

hipError_t scratch_setup(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch_ptr[device] = static_cast<double*>(p);
  return e;
}

double* scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define GX_HIP(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return fail(GX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// As GX_HIP, inside a fork onto the side streams: on failure the caller's
// stream is joined to the side streams before returning (so that it still
// orders after everything already launched).
#define GX_HIP_FORKED(x, n)                                                              \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      (void)join_side(c, (n));                                                            \
      return fail(GX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));           \
    }                                                                                     \
  } while (0)

#define GX_TRY(x)            \
  do {                       \
    int r_ = (x);            \
    if (r_ != GX_OK) return r_; \
  } while (0)

constexpr int NCLS = (GX_ERI_LMAX + 1) * (GX_ERI_LMAX + 2) / 2;

inline int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }
inline int class_la(int c) {
  int la = 0;
  while ((la + 1) * (la + 2) / 2 <= c) ++la;
  return la;
}
inline int class_lb(int c) { return c - class_la(c) * (class_la(c) + 1) / 2; }
inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
inline int class_L(int X, int Y) { return class_la(X) + class_lb(X) + class_la(Y) + class_lb(Y); }

template <class T>
int dalloc(T** p, size_t n) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(GX_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  return GX_OK;
}

template <class T>
void dfree(T** p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

template <class T>
int upload(T** p, const std::vector<T>& v, hipStream_t s) {
  GX_TRY(dalloc(p, v.size()));
  if (!v.empty()) GX_HIP(hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return GX_OK;
}

// Boys master table on the grid T_n = n * dt, n < nt: F_m(T_n) by the
// everywhere-convergent series
// F_M(T) = e^-T sum_k (2T)^k / ((2M+1)(2M+3)...(2M+2k+1)) in long double,
// then downward recursion; last column exp(-T_n).
void boys_master_table(std::vector<double>& tab, int nt, long double dt) {
  const int M = BOYS_MCOL - 1;
  tab.assign((size_t)nt * (BOYS_MCOL + 1), 0.0);
  for (int n = 0; n < nt; ++n) {
    const long double T = (long double)n * dt;
    long double term = 1.0L / (2 * M + 1), sum = term;
    for (int k = 1; k < 4000; ++k) {
      term *= 2.0L * T / (2 * M + 2 * k + 1);
      sum += term;
      if (term < sum * 1e-22L) break;
    }
    const long double e = expl(-T);
    long double f = e * sum;
    double* row = &tab[(size_t)n * (BOYS_MCOL + 1)];
    row[M] = (double)f;
    for (int m = M; m >= 1; --m) {
      f = (2.0L * T * f + e) / (2 * m - 1);
      row[m - 1] = (double)f;
    }
    row[BOYS_MCOL] = (double)e;
  }
}

}  // namespace

struct gx_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  int max_lds = 65536;
  double* d_boys = nullptr;   // master table, fine grid (also the one-electron kernels')
  double* d_boyst = nullptr;  // master table, coarse grid
  double* d_tab64 = nullptr;  // per-L Taylor tables, fine grid (L <= BOYS_MCOL - BOYS_TAYLOR64)
  float* d_tab32 = nullptr;
  double* d_tab64c = nullptr;  // per-L Taylor tables, coarse grid (L <= BOYS_MCOL - BoysCoarse::K64)
  float* d_tab32c = nullptr;
  // basis
  int nshell = 0, nbf = 0, nprim = 0;
  std::vector<int> l, np, poff, bf, atom;
  std::vector<double> ctr, ctr1, exps, coefs, scale;
  int *d_l = nullptr, *d_np = nullptr, *d_poff = nullptr, *d_bf = nullptr;
  double *d_ctr = nullptr, *d_ctr1 = nullptr, *d_exps = nullptr, *d_coefs = nullptr,
         *d_scale = nullptr;
  // pairs
  int npair = 0, npp = 0;
  std::vector<int> pA, pB, pcls, ppoff, pnp;
  // Primitive-pair screening: pair p keeps the primitive pairs
  // kept_ab[kept_off[p] .. kept_off[p + 1]) (ia << 16 | ib); empty = all.
  double prim_eps = GX_PRIM_EPS_DEFAULT;
  std::vector<int> kept_off, kept_ab;
  std::vector<double> Q;
  std::vector<int> cls_start;  // NCLS + 1
  std::vector<int> pairmap;    // nshell*nshell -> pair id
  int *d_pA = nullptr, *d_pB = nullptr, *d_ppoff = nullptr, *d_pnp = nullptr;
  double* d_Q = nullptr;
  gx::PrimPair64* d_pp64 = nullptr;
  gx::PrimPair32* d_pp32 = nullptr;
  gx::PairRec* d_prec = nullptr;  // per pair, for the J/K and gradient kernels
  // Work sets per class combo: work items (4 ints: bra position, range
  // [v0, v1) of the combo's ket list, bra primitive-pair range p0 | p1 << 16),
  // largest first; and the ket list: per bra its surviving kets (FP64 ones,
  // then FP32 ones from the next multiple of 64, GX_KET_* encoding).  `jk` for
  // gx_jk, `gr` for gx_grad.
  struct WorkSet {
    std::vector<int*> d_items, d_kets;
    std::vector<int> n_items;
    std::vector<double> ipw;  // items per resident workgroup, per combo and mode (FP64, mixed; autotuned)
    std::vector<double> tms;  // launch time (ms) the autotuner measured, per combo and mode; 0 = not timed
    double cutoff = -1.0, fp32 = -1.0;
    int mixed = -1;
    void release() {
      for (auto& p : d_items) dfree(&p);
      for (auto& p : d_kets) dfree(&p);
      d_items.clear();
      d_kets.clear();
      n_items.clear();
      cutoff = fp32 = -1.0;
      mixed = -1;
    }
  } jk, gr;
  // per class combination: 0 when a mixed build evaluates it all in FP64
  // (chosen by gx_jk_autotune where the FP32 kernel is not faster)
  std::vector<char> jk_fp32;
  std::vector<int> wg_per_cu;  // [combo][FP64, mixed, two-density] resident workgroups per CU, 0 = not queried
  std::vector<int> wg_per_cu_grad;
  int* d_atom = nullptr;
  int* d_item_ctr = nullptr;  // per class combination: [next item, workgroups done] of the J/K, then the gradient kernels
  // side streams for the class-combination kernels of one build (created on
  // first use, default priority), fork and join events
  hipStream_t side[GX_NSIDE > 0 ? GX_NSIDE : 1] = {};
  hipEvent_t ev_fork = nullptr, ev_join[GX_NSIDE > 0 ? GX_NSIDE : 1] = {};
  bool side_ready = false;
  // ket segments: per class, runs of pairs with the same primitive-count key
  std::vector<int> seg_lo, seg_hi;
  std::vector<int> cseg;  // [class][system] first segment, NCLS x (nsys + 1)
  // systems of a batch (one system unless gx_set_basis_batch): shells
  // [sys_shell[s], sys_shell[s + 1]), nuclei [sys_atom[s], sys_atom[s + 1]),
  // nbf sn[s], first function sbfb[s], packed-matrix offset smb[s]
  int nsys = 1;
  std::vector<int> sys_shell, sys_atom, ssys, lbf, sn, sbfb;
  std::vector<long> smb;
  long mtot = 0;
  int nmax = 0;
  int *d_ssys = nullptr, *d_lbf = nullptr, *d_sn = nullptr, *d_sbfb = nullptr, *d_sys_atom = nullptr,
      *d_atom_sys = nullptr;
  long* d_smb = nullptr;
  bool sys_atom_ok = false;  // d_sys_atom / d_atom_sys match the current basis and nuclei
  std::vector<int64_t> pre_pnp, pre_nab;     // prefix sums over pair ids
  std::vector<int> nab;                       // basis-function pairs per shell pair
  int num_cu = 256;
  // scratch: d_Jt holds Jt (mtot doubles) then GX_JT_COPIES copies of the
  // pair-indexed J_cd accumulator (jkp_n doubles each; component k of pair p
  // at jko + k * pairs of its class); jkp_of maps a packed matrix element to
  // its Jkp entry (-1: none)
  double *d_Ps = nullptr, *d_Jt = nullptr, *d_Kt = nullptr;
  long jkp_n = 0;
  int* d_jkp_of = nullptr;
  long jt_doubles() const { return mtot + GX_JT_COPIES * jkp_n; }
  // two-density (unrestricted) builds, allocated on first use
  double *d_Ps2 = nullptr, *d_Psj = nullptr, *d_Kt2 = nullptr;
  // nuclei
  int natom = 0;
  std::vector<double> nxyz, nZ;
  double *d_nxyz = nullptr, *d_nZ = nullptr;
  // launch events of gx_jk calls made with GX_JK_TIMING since the last read:
  // pairs (ev[2i], ev[2i+1]) around the launch of class combo ev_combo[i]
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_combo;
  // the same for gx_grad_ex calls made with GX_JK_TIMING
  std::vector<hipEvent_t> gev;
  std::vector<int> gev_combo;

  gx::ShellView shells() const { return gx::ShellView{d_l, d_bf, d_ctr, d_lbf, d_ssys}; }
  gx::SysView systems() const { return gx::SysView{d_sn, d_smb}; }
  int cseg_at(int cls, int sys) const { return cseg[(size_t)cls * (nsys + 1) + sys]; }
  gx::PairView pairs() const { return gx::PairView{d_pA, d_pB, d_ppoff, d_pnp, d_Q, d_pp64, d_pp32, d_prec}; }

  void free_pairs() {
    dfree(&d_pA);
    dfree(&d_pB);
    dfree(&d_ppoff);
    dfree(&d_pnp);
    dfree(&d_Q);
    dfree(&d_pp64);
    dfree(&d_pp32);
    dfree(&d_prec);
    jk.release();
    gr.release();
    jk.ipw.clear();
    gr.ipw.clear();
    jk.tms.clear();
    gr.tms.clear();
    jk_fp32.clear();
    wg_per_cu.clear();
    wg_per_cu_grad.clear();
  }
  void free_all() {
    free_pairs();
    dfree(&d_atom);
    dfree(&d_ssys);
    dfree(&d_lbf);
    dfree(&d_sn);
    dfree(&d_sbfb);
    dfree(&d_smb);
    dfree(&d_sys_atom);
    dfree(&d_atom_sys);
    dfree(&d_l);
    dfree(&d_np);
    dfree(&d_poff);
    dfree(&d_bf);
    dfree(&d_ctr);
    dfree(&d_ctr1);
    dfree(&d_exps);
    dfree(&d_coefs);
    dfree(&d_scale);
    dfree(&d_Ps);
    dfree(&d_Jt);
    dfree(&d_jkp_of);
    dfree(&d_Kt);
    dfree(&d_Ps2);
    dfree(&d_Psj);
    dfree(&d_Kt2);
    dfree(&d_item_ctr);
    release_side();
    dfree(&d_nxyz);
    dfree(&d_nZ);
    dfree(&d_boys);
    dfree(&d_boyst);
    dfree(&d_tab64);
    dfree(&d_tab32);
    dfree(&d_tab64c);
    dfree(&d_tab32c);
    for (auto& e : ev) (void)hipEventDestroy(e);
    ev.clear();
    ev_combo.clear();
    for (auto& e : gev) (void)hipEventDestroy(e);
    gev.clear();
    gev_combo.clear();
  }
  // Destroys whichever side streams and events exist (also after a
  // set-up that failed part way).
  void release_side() {
    for (int k = 0; k < (GX_NSIDE > 0 ? GX_NSIDE : 1); ++k) {
      if (side[k]) (void)hipStreamDestroy(side[k]);
      if (ev_join[k]) (void)hipEventDestroy(ev_join[k]);
      side[k] = nullptr;
      ev_join[k] = nullptr;
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    ev_fork = nullptr;
    side_ready = false;
  }
};

namespace {

int combo_count() { return NCLS * (NCLS + 1) / 2; }

// ---------------------------------------------------------------------------
// Kernel dispatch by class.  Per-lane classes (EriClass) run one quartet per
// lane; cooperative classes (EriProg, gx_coop.hip) one quartet per wave with
// the recursion tables in dynamic LDS.
template <class G>
size_t coop_bytes() {
  return (size_t)gx::coop_lds<G>() * sizeof(double);
}

hipError_t launch_schwarz(int la, int lb, hipStream_t s, gx::ShellView sh, gx::PairView pr, int p0, int np,
                          const double* boys, const double* tab64, const double* scale, double* Q) {
#define GX_CASE(A, B)                                                                           \
  if (la == A && lb == B) {                                                                     \
    hipLaunchKernelGGL((gx::k_schwarz<A, B>), dim3((np + 63) / 64), dim3(64), 0, s, sh, pr, p0, np, \
                       boys, scale, Q);                                                         \
    return hipGetLastError();                                                                   \
  }
  GX_FOR_EACH_LANE_PAIRCLASS(GX_CASE)
#undef GX_CASE
#define GX_CASE(A, B)                                                                          \
  if (la == A && lb == B) {                                                                    \
    const size_t lds = coop_bytes<EriProg<A, B, A, B>>();                                      \
    hipError_t e = hipFuncSetAttribute((const void*)gx::k_schwarz_coop<A, B>,                   \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);  \
    if (e != hipSuccess) return e;                                                             \
    hipLaunchKernelGGL((gx::k_schwarz_coop<A, B>), dim3(np), dim3(GX_WAVE), lds, s, sh, pr, p0, \
                       tab64, scale, Q);                                                       \
    return hipGetLastError();                                                                  \
  }
  GX_FOR_EACH_COOP_PAIRCLASS(GX_CASE)
#undef GX_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_dense(int la, int lb, int lc, int ld, hipStream_t s, gx::ShellView sh, gx::PairView pr,
                        int b0, int nb, int k0, int nk, int same, const double* boys, const double* tab64,
                        const double* scale, double* out, long N) {
#define GX_CASE(A, B, C, D)                                                                   \
  if (la == A && lb == B && lc == C && ld == D) {                                             \
    hipLaunchKernelGGL((gx::k_dense<A, B, C, D>), dim3((nk + 63) / 64, nb), dim3(64), 0, s, sh, pr, \
                       b0, nb, k0, nk, same, boys, scale, out, N);                           \
    return hipGetLastError();                                                                 \
  }
  GX_FOR_EACH_LANE_CLASS(GX_CASE)
#undef GX_CASE
#define GX_CASE(A, B, C, D)                                                                   \
  if (la == A && lb == B && lc == C && ld == D) {                                             \
    const size_t lds = coop_bytes<EriProg<A, B, C, D>>();                                     \
    hipError_t e = hipFuncSetAttribute((const void*)gx::k_dense_coop<A, B, C, D>,              \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    if (e != hipSuccess) return e;                                                            \
    hipLaunchKernelGGL((gx::k_dense_coop<A, B, C, D>), dim3(nk, nb), dim3(GX_WAVE), lds, s, sh, pr, \
                       b0, k0, same, tab64, scale, out, N);                                   \
    return hipGetLastError();                                                                 \
  }
  GX_FOR_EACH_COOP_CLASS(GX_CASE)
#undef GX_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_quartets(int la, int lb, int lc, int ld, hipStream_t s, gx::ShellView sh, gx::PairView pr,
                           int n, const int* bra, const int* ket, const int* comp, const double* sc,
                           const long* pos, const double* boys, const double* tab64, double* out) {
#define GX_CASE(A, B, C, D)                                                                       \
  if (la == A && lb == B && lc == C && ld == D) {                                                 \
    hipLaunchKernelGGL((gx::k_quartets<A, B, C, D>), dim3((n + 63) / 64), dim3(64), 0, s, sh, pr, n, \
                       bra, ket, comp, sc, pos, boys, out);                                       \
    return hipGetLastError();                                                                     \
  }
  GX_FOR_EACH_LANE_CLASS(GX_CASE)
#undef GX_CASE
#define GX_CASE(A, B, C, D)                                                                       \
  if (la == A && lb == B && lc == C && ld == D) {                                                 \
    const size_t lds = coop_bytes<EriProg<A, B, C, D>>();                                         \
    hipError_t e = hipFuncSetAttribute((const void*)gx::k_quartets_coop<A, B, C, D>,               \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);     \
    if (e != hipSuccess) return e;                                                                \
    hipLaunchKernelGGL((gx::k_quartets_coop<A, B, C, D>), dim3(n), dim3(GX_WAVE), lds, s, sh, pr, bra, \
                       ket, comp, sc, pos, tab64, out);                                           \
    return hipGetLastError();                                                                     \
  }
  GX_FOR_EACH_COOP_CLASS(GX_CASE)
#undef GX_CASE
  return hipErrorInvalidValue;
}

// Persistent J/K launch: as many workgroups as stay resident (queried once
// per kernel and basis, cached in *per_cu), at most one per work item.
hipError_t launch_jk(int la, int lb, int lc, int ld, int num_cu, size_t lds, hipStream_t s,
                     const gx::JKArgs& a, int* per_cu, int nden = 1) {
#define GX_CASE(A, B, C, D)                                                                   \
  if (la == A && lb == B && lc == C && ld == D) {                                             \
    const void* fn = nden == 2 ? (const void*)gx::k_jk<A, B, C, D, false, 2>                 \
                     : a.mixed ? (const void*)gx::k_jk<A, B, C, D, true, 1>                   \
                               : (const void*)gx::k_jk<A, B, C, D, false, 1>;                 \
    /* the attribute is per function, shared by every context: set it each time */          \
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    if (e != hipSuccess) return e;                                                            \
    if (*per_cu <= 0) {                                                                       \
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, fn, JK_TEAM, lds);             \
      if (e != hipSuccess) return e;                                                          \
      if (*per_cu < 1) *per_cu = 1;                                                           \
    }                                                                                         \
    const int grid = std::min(a.nitem, *per_cu * num_cu);                                    \
    if (nden == 2)                                                                            \
      hipLaunchKernelGGL((gx::k_jk<A, B, C, D, false, 2>), dim3(grid), dim3(JK_TEAM), lds, s, a, \
                         a.pr.pp64, a.pr.pp32);                                               \
    else if (a.mixed)                                                                         \
      hipLaunchKernelGGL((gx::k_jk<A, B, C, D, true, 1>), dim3(grid), dim3(JK_TEAM), lds, s, a, \
                         a.pr.pp64, a.pr.pp32);                                               \
    else                                                                                      \
      hipLaunchKernelGGL((gx::k_jk<A, B, C, D, false, 1>), dim3(grid), dim3(JK_TEAM), lds, s, a, \
                         a.pr.pp64, a.pr.pp32);                                               \
    return hipGetLastError();                                                                 \
  }
  GX_FOR_EACH_LANE_CLASS(GX_CASE)
#undef GX_CASE
#define GX_CASE(A, B, C, D)                                                                   \
  if (la == A && lb == B && lc == C && ld == D) {                                             \
    const void* fn = (const void*)gx::k_jk_rys<A, B, C, D>;                                   \
    constexpr int nt = gx::rys_threads<A, B, C, D>();                                         \
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    if (e != hipSuccess) return e;                                                            \
    if (*per_cu <= 0) {                                                                       \
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, fn, nt, lds);                  \
      if (e != hipSuccess) return e;                                                          \
      if (*per_cu < 1) *per_cu = 1;                                                           \
    }                                                                                         \
    const int grid = std::min(a.nitem, *per_cu * num_cu);                                    \
    hipLaunchKernelGGL((gx::k_jk_rys<A, B, C, D>), dim3(grid), dim3(nt), lds, s, a);         \
    return hipGetLastError();                                                                 \
  }
  GX_FOR_EACH_COOP_CLASS(GX_CASE)
#undef GX_CASE
  return hipErrorInvalidValue;
}

// Two-electron gradient of one class combination (gx_grad.hip, k_grad_lane;
// gx_rys.hip, k_grad_rys).
size_t grad_lds_bytes(int la, int lb, int lc, int ld, int N) {
  if (gx_grad_is_lane(la, lb, lc, ld))
    return ((size_t)gx::GRAD_TAB_DOUBLES + (size_t)(ncart(la) + ncart(lb)) * N) * sizeof(double);
#define GX_CASE(A, B, C, D)                                                                     \
  if (la == A && lb == B && lc == C && ld == D) {                                               \
    using G = EriGradProg<A, B, C, D>;                                                          \
    const int nblk = (G::NA + G::NB) * (G::NC + G::ND) + G::NA * G::NB + G::NC * G::ND;          \
    return ((size_t)gx::rys_grad_lds<A, B, C, D>() + nblk) * sizeof(double);                    \
  }
  GX_FOR_EACH_CLASS(GX_CASE)
#undef GX_CASE
  return 0;
}

// mixed = 1 (per-lane classes only): kets below the FP32 bound run their
// primitive loops in FP32 (k_grad_lane<..., true>).
hipError_t launch_grad(int la, int lb, int lc, int ld, int num_cu, hipStream_t s, const gx::GradArgs& a,
                       int* per_cu, int N, int mixed) {
  const size_t lds = grad_lds_bytes(la, lb, lc, ld, N);
#define GX_LANE(A, B, C, D, M)                                                                \
  {                                                                                           \
    const void* fn = (const void*)gx::k_grad_lane<A, B, C, D, M>;                            \
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    if (e != hipSuccess) return e;                                                            \
    if (*per_cu <= 0) {                                                                       \
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, fn, JK_TEAM, lds);             \
      if (e != hipSuccess) return e;                                                          \
      if (*per_cu < 1) *per_cu = 1;                                                           \
    }                                                                                         \
    const int grid = std::min(a.nitem, *per_cu * num_cu);                                    \
    hipLaunchKernelGGL((gx::k_grad_lane<A, B, C, D, M>), dim3(grid), dim3(JK_TEAM), lds, s, a); \
    return hipGetLastError();                                                                 \
  }
#define GX_CASE(A, B, C, D)                                   \
  if (la == A && lb == B && lc == C && ld == D) {             \
    if (mixed) GX_LANE(A, B, C, D, true) else GX_LANE(A, B, C, D, false) \
  }
  GX_FOR_EACH_GRAD_LANE_CLASS(GX_CASE)
#undef GX_CASE
#undef GX_LANE
#define GX_CASE(A, B, C, D)                                                                   \
  if (la == A && lb == B && lc == C && ld == D) {                                             \
    const void* fn = (const void*)gx::k_grad_rys<A, B, C, D>;                                \
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    if (e != hipSuccess) return e;                                                            \
    if (*per_cu <= 0) {                                                                       \
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, fn, GX_WAVE, lds);             \
      if (e != hipSuccess) return e;                                                          \
      if (*per_cu < 1) *per_cu = 1;                                                           \
    }                                                                                         \
    const int grid = std::min(a.nitem, *per_cu * num_cu);                                    \
    hipLaunchKernelGGL((gx::k_grad_rys<A, B, C, D>), dim3(grid), dim3(GX_WAVE), lds, s, a);  \
    return hipGetLastError();                                                                 \
  }
  GX_FOR_EACH_GRAD_COOP_CLASS(GX_CASE)
#undef GX_CASE
  return hipErrorInvalidValue;
}

// Whether a class combination runs one quartet per wave (the Rys J/K kernel,
// the cooperative gradient, dense and Schwarz kernels; always FP64).
bool combo_is_coop(int X, int Y) { return gx_class_is_coop(class_la(X), class_lb(X), class_la(Y), class_lb(Y)); }

// Must match the layout at the top of gx::k_jk (per-lane classes) or
// gx::k_jk_rys (the classes evaluated one quartet per wave).
size_t jk_lds_bytes(int la, int lb, int lc, int ld, int N, int mixed, int nden = 1) {
#define GX_CASE(A, B, C, D)                                                                     \
  if (la == A && lb == B && lc == C && ld == D) {                                               \
    using G = EriProg<A, B, C, D>;                                                              \
    const int nblk = (G::NA + G::NB) * (G::NC + G::ND) + G::NA * G::NB + G::NC * G::ND;          \
    return (size_t)(gx::rys_lds<A, B, C, D>() + 2 * nblk) * sizeof(double);                    \
  }
  GX_FOR_EACH_COOP_CLASS(GX_CASE)
#undef GX_CASE
  const int na = ncart(la), nb = ncart(lb);
  const bool t64 = true, t32 = mixed && GX_MIXED_LDS32;
  const bool coarse = gx::jk_coarse_grid(la, lb, lc, ld) || mixed;
  const size_t nt = coarse ? gx::BoysCoarse::NT : gx::BoysFine::NT;
  const size_t r64 = coarse ? gx::BoysCoarse::ROW64 : gx::BoysFine::ROW64;
  const size_t r32 = coarse ? gx::BoysCoarse::ROW32 : gx::BoysFine::ROW32;
  const size_t b = (t64 ? nt * r64 * sizeof(double) : 0) + (t32 ? nt * r32 * sizeof(float) : 0);
  return b + 2 * (size_t)nden * (na + nb) * N * sizeof(double);
}

// Build the primitive-pair records of the current pair list on the device.
int build_primpairs(gx_ctx* c) {
  std::vector<int> pair_of, pia, pib;
  c->ppoff.assign(c->npair, 0);
  c->pnp.assign(c->npair, 0);
  const bool all = c->kept_off.empty();
  int off = 0;
  for (int p = 0; p < c->npair; ++p) {
    const int A = c->pA[p], B = c->pB[p];
    c->ppoff[p] = off;
    if (all) {
      c->pnp[p] = c->np[A] * c->np[B];
      for (int i = 0; i < c->np[A]; ++i)
        for (int j = 0; j < c->np[B]; ++j) {
          pair_of.push_back(p);
          pia.push_back(i);
          pib.push_back(j);
        }
    } else {
      c->pnp[p] = c->kept_off[p + 1] - c->kept_off[p];
      for (int k = c->kept_off[p]; k < c->kept_off[p + 1]; ++k) {
        pair_of.push_back(p);
        pia.push_back(c->kept_ab[k] >> 16);
        pib.push_back(c->kept_ab[k] & 0xffff);
      }
    }
    off += c->pnp[p];
  }
  c->npp = off;
  int *d_of = nullptr, *d_ia = nullptr, *d_ib = nullptr;
  int rc = GX_OK;
  if ((rc = upload(&d_of, pair_of, c->stream)) != GX_OK || (rc = upload(&d_ia, pia, c->stream)) != GX_OK ||
      (rc = upload(&d_ib, pib, c->stream)) != GX_OK) {
    dfree(&d_of);
    dfree(&d_ia);
    dfree(&d_ib);
    return rc;
  }
  GX_TRY(upload(&c->d_pA, c->pA, c->stream));
  GX_TRY(upload(&c->d_pB, c->pB, c->stream));
  GX_TRY(upload(&c->d_ppoff, c->ppoff, c->stream));
  GX_TRY(upload(&c->d_pnp, c->pnp, c->stream));
  GX_TRY(dalloc(&c->d_pp64, (size_t)c->npp));
  GX_TRY(dalloc(&c->d_pp32, (size_t)c->npp));
  if (c->npp > 0) {
    hipLaunchKernelGGL(gx::k_primpairs, dim3((c->npp + 255) / 256), dim3(256), 0, c->stream, c->npp,
                       d_of, d_ia, d_ib, c->d_pA, c->d_pB, c->d_poff, c->d_ctr, c->d_exps,
                       c->d_coefs, c->d_pp64, c->d_pp32);
  }
  hipError_t e = hipGetLastError();
  hipError_t e2 = hipStreamSynchronize(c->stream);
  dfree(&d_of);
  dfree(&d_ia);
  dfree(&d_ib);
  if (e != hipSuccess || e2 != hipSuccess)
    return fail(GX_ERR_HIP, std::string("k_primpairs: ") + hipGetErrorString(e != hipSuccess ? e : e2));
  return GX_OK;
}

// Primitive-pair screening.  For every primitive pair (a, b) of every shell
// pair, S_ab = sqrt(max over components |(ab|ab)|) of the scaled primitive
// product (the Schwarz kernels run on a view in which each primitive pair is
// its own shell pair).  By the Schwarz inequality a primitive quartet is
// bounded by S_ab S_cd <= S_ab S_max, so a primitive pair with
// S_ab S_max < prim_eps changes no contracted integral by more than
// prim_eps times its number of primitive quartets; it is dropped.  Each shell
// pair keeps at least its largest primitive pair.  Expects the records of
// every primitive pair (kept_off empty); leaves kept_off / kept_ab and the
// screened records.
int screen_primpairs(gx_ctx* c) {
  const int npp = c->npp;
  std::vector<int> vA(npp), vB(npp), voff(npp), vnp(npp, 1);
  for (int p = 0; p < c->npair; ++p)
    for (int k = 0; k < c->pnp[p]; ++k) {
      vA[c->ppoff[p] + k] = c->pA[p];
      vB[c->ppoff[p] + k] = c->pB[p];
    }
  std::iota(voff.begin(), voff.end(), 0);
  int *d_vA = nullptr, *d_vB = nullptr, *d_voff = nullptr, *d_vnp = nullptr;
  double* d_S = nullptr;
  std::vector<double> S(npp, 0.0);
  int rc = GX_OK;
  auto run = [&]() -> int {
    GX_TRY(upload(&d_vA, vA, c->stream));
    GX_TRY(upload(&d_vB, vB, c->stream));
    GX_TRY(upload(&d_voff, voff, c->stream));
    GX_TRY(upload(&d_vnp, vnp, c->stream));
    GX_TRY(dalloc(&d_S, (size_t)npp));
    const gx::PairView pv{d_vA, d_vB, d_voff, d_vnp, nullptr, c->d_pp64, c->d_pp32, nullptr};
    for (int k = 0; k < NCLS; ++k) {
      const int p0 = c->cls_start[k], p1 = c->cls_start[k + 1];
      if (p1 == p0) continue;
      const int q0 = c->ppoff[p0], q1 = c->ppoff[p1 - 1] + c->pnp[p1 - 1];
      if (q1 == q0) continue;
      GX_HIP(launch_schwarz(class_la(k), class_lb(k), c->stream, c->shells(), pv, q0, q1 - q0, c->d_boys,
                            c->d_tab64, c->d_scale, d_S));
    }
    GX_HIP(hipMemcpyAsync(S.data(), d_S, npp * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GX_HIP(hipStreamSynchronize(c->stream));
    return GX_OK;
  };
  rc = run();
  dfree(&d_vA);
  dfree(&d_vB);
  dfree(&d_voff);
  dfree(&d_vnp);
  dfree(&d_S);
  if (rc != GX_OK) return rc;
  double smax = 0.0;
  for (double v : S) smax = std::max(smax, v);
  c->kept_off.assign(c->npair + 1, 0);
  c->kept_ab.clear();
  for (int p = 0; p < c->npair; ++p) {
    const int nb = c->np[c->pB[p]];
    int best = 0;
    const size_t first = c->kept_ab.size();
    for (int k = 0; k < c->pnp[p]; ++k) {
      const double s = S[c->ppoff[p] + k];
      if (s > S[c->ppoff[p] + best]) best = k;
      if (s * smax >= c->prim_eps) c->kept_ab.push_back((k / nb) << 16 | (k % nb));
    }
    if (c->kept_ab.size() == first) c->kept_ab.push_back((best / nb) << 16 | (best % nb));
    c->kept_off[p + 1] = (int)c->kept_ab.size();
  }
  return build_primpairs(c);
}

int compute_schwarz(gx_ctx* c) {
  GX_TRY(dalloc(&c->d_Q, (size_t)c->npair));
  for (int k = 0; k < NCLS; ++k) {
    const int p0 = c->cls_start[k], n = c->cls_start[k + 1] - p0;
    if (n == 0) continue;
    GX_HIP(launch_schwarz(class_la(k), class_lb(k), c->stream, c->shells(), c->pairs(), p0, n, c->d_boys,
                          c->d_tab64, c->d_scale, c->d_Q));
  }
  c->Q.assign(c->npair, 0.0);
  GX_HIP(hipMemcpyAsync(c->Q.data(), c->d_Q, c->npair * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  GX_HIP(hipStreamSynchronize(c->stream));
  return GX_OK;
}

// Surviving kets of bra pair ip in ket segment sg: count with Q_ip Q_j >= thr
// (segments are sorted by decreasing Q).
int host_count(const gx_ctx* c, int lo, int hi, double QA, double thr) {
  if (thr <= 0.0) return hi - lo;
  int a = lo, b = hi;
  while (a < b) {
    const int m = (a + b) / 2;
    if (QA * c->Q[m] >= thr) a = m + 1;
    else b = m;
  }
  return a - lo;
}

// For every class combo and bra: [lo, hi, n64, n32] per ket segment.
template <class F>
void for_each_bra_segment(const gx_ctx* c, int X, int Y, double cutoff, double fp32, int mixed, F&& f,
                          bool jk_rules = true) {
  // J/K: the cooperative classes are always FP64, and so are the
  // combinations gx_jk_autotune put in FP64; the gradient's caller decides
  if (jk_rules && combo_is_coop(X, Y)) mixed = 0;
  const size_t combo = (size_t)X * (X + 1) / 2 + Y;
  if (jk_rules && combo < c->jk_fp32.size() && !c->jk_fp32[combo]) mixed = 0;
  const int b0 = c->cls_start[X], nb = c->cls_start[X + 1] - b0;
  for (int i = 0; i < nb; ++i) {
    const int ip = b0 + i;
    const double QA = c->Q[ip];
    const int sy = c->ssys[c->pA[ip]];
    for (int sg = c->cseg_at(Y, sy); sg < c->cseg_at(Y, sy + 1); ++sg) {
      const int lo = c->seg_lo[sg];
      int hi = c->seg_hi[sg];
      if (X == Y && hi > ip + 1) hi = ip + 1;
      if (hi <= lo) continue;
      const int ns = host_count(c, lo, hi, QA, cutoff);
      const int n64 = mixed ? std::min(ns, host_count(c, lo, hi, QA, fp32)) : ns;
      f(i, ip, lo, hi, n64, ns - n64);
    }
  }
}

// Work items of the J/K kernels.  A bra's surviving kets form a virtual
// index space (FP64 ranges, then FP32 ranges from the next multiple of 64,
// the same layout the kernel rebuilds); waves take 64 consecutive indices at
// a time.  Heavy bras are split into items over contiguous runs of those
// chunks and, if one chunk alone is too heavy, over their primitive pairs
// too: J and K are linear in the integrals, so each part is digested on its
// own.  Costs are in primitive-quartet lane iterations; items are sorted by
// decreasing cost and the persistent workgroups stride over them.
struct JKItem {
  int bpos, v0, v1, prims;
};

#ifndef GX_ITEM_OVERHEAD
#define GX_ITEM_OVERHEAD 256.0
#endif
#ifndef GX_ITEMS_PER_WG
#define GX_ITEMS_PER_WG 4.0
#endif

// Work items of one class combination (X|Y) into ws (see build_orders).
int build_combo(gx_ctx* c, gx_ctx::WorkSet& ws, int X, int Y, int combo, double cutoff, double fp32, int mixed,
                bool grad, double ipw) {
  constexpr double kItemOverhead = GX_ITEM_OVERHEAD;  // setup + flush of one item
  constexpr double kQuartetCost = 8.0;                // digestion of one shell quartet per lane
  dfree(&ws.d_items[combo]);
  dfree(&ws.d_kets[combo]);
  ws.n_items[combo] = 0;
  const int b0 = c->cls_start[X], nb = c->cls_start[X + 1] - b0;
  const int nk = c->cls_start[Y + 1] - c->cls_start[Y];
  if (nb == 0 || nk == 0) return GX_OK;
  // surviving kets per bra and ket segment: FP64 part, then FP32 part
  std::vector<int> rng((size_t)nb * 2 * JK_MAXSEG, 0);
  for_each_bra_segment(c, X, Y, cutoff, fp32, mixed, [&](int i, int ip, int lo, int, int n64, int n32) {
    const int s0 = c->cseg_at(Y, c->ssys[c->pA[ip]]);
    int sg = 0;
    while (c->seg_lo[s0 + sg] != lo) ++sg;
    rng[(size_t)i * 2 * JK_MAXSEG + sg] = n64;
    rng[(size_t)i * 2 * JK_MAXSEG + JK_MAXSEG + sg] = n32;
  }, !grad);
  // per bra: its ket list (appended to `kets`, from kofs[i]) and per-chunk
  // cost (widest ket of the chunk).  Per-lane classes: chunks of 64 kets,
  // four waves per item; cooperative classes: one ket per chunk, one wave
  // per item.
  const bool coop = grad ? !gx_grad_is_lane(class_la(X), class_lb(X), class_la(Y), class_lb(Y))
                         : combo_is_coop(X, Y);
  const int csz = coop ? 1 : GX_WAVE;
  const double waves = coop ? 1.0 : (double)(JK_TEAM / GX_WAVE);
  std::vector<std::vector<double>> chunk(nb);
  std::vector<int> kets, kofs(nb + 1, 0);
  double total = 0.0;
  std::vector<int> vk;
  for (int i = 0; i < nb; ++i) {
    const int* r = &rng[(size_t)i * 2 * JK_MAXSEG];
    const int sy = c->ssys[c->pA[b0 + i]];
    const int s0 = c->cseg_at(Y, sy), nseg = c->cseg_at(Y, sy + 1) - s0;
    vk.clear();
    for (int h = 0; h < 2; ++h) {
      // the FP32 kets start at the next multiple of 64, so no chunk mixes precisions
      if (h == 1) vk.resize((vk.size() + GX_WAVE - 1) / GX_WAVE * GX_WAVE, GX_KET_NONE);
      for (int sg = 0; sg < nseg; ++sg) {
        const int start = c->seg_lo[s0 + sg] + (h ? r[sg] : 0);
        for (int k = 0; k < r[h * JK_MAXSEG + sg]; ++k) vk.push_back((start + k) | (h ? GX_KET_FP32 : 0));
      }
    }
    while (!vk.empty() && vk.back() < 0) vk.pop_back();
    kofs[i] = (int)kets.size();
    kets.insert(kets.end(), vk.begin(), vk.end());
    const double nbp = c->pnp[b0 + i];
    for (size_t v0 = 0; v0 < vk.size(); v0 += csz) {
      int widest = 0;
      for (size_t v = v0; v < std::min(vk.size(), v0 + csz); ++v)
        if (vk[v] >= 0) widest = std::max(widest, c->pnp[vk[v] & GX_KET_ID]);
      chunk[i].push_back(nbp * widest + kQuartetCost);
      total += chunk[i].back();
    }
  }
  kofs[nb] = (int)kets.size();
  // item size: about `ipw` items per resident workgroup (four per CU
  // assumed), each spread over the waves of its workgroup, and not below the
  // per-item overhead scaled by GX_ITEMS_PER_WG / ipw (so that the autotuner's
  // larger ipw candidates reach the small items that light combinations with
  // few bras need to fill the device)
  const double slots = (double)c->num_cu * 4.0;
  const double target = std::max(total / (slots * waves * ipw), kItemOverhead * GX_ITEMS_PER_WG / ipw);
  std::vector<JKItem> items;
  std::vector<double> icost;
  for (int i = 0; i < nb; ++i) {
    const auto& ch = chunk[i];
    if (ch.empty()) continue;
    const int nbp = c->pnp[b0 + i];
    const int len = kofs[i + 1] - kofs[i];
    double widest = 0.0;
    for (double x : ch) widest = std::max(widest, x);
    // primitive split so that one chunk of each part fits the target
    const int nsplit = std::max(1, std::min(nbp, (int)std::ceil(widest / target)));
    for (int sp = 0; sp < nsplit; ++sp) {
      const int p0 = (int)((long)nbp * sp / nsplit), p1 = (int)((long)nbp * (sp + 1) / nsplit);
      const double frac = (double)(p1 - p0) / nbp;
      // runs of chunks of about target x waves
      size_t k0 = 0;
      double acc = 0.0;
      for (size_t k = 0; k < ch.size(); ++k) {
        acc += ch[k] * frac;
        const bool last = k + 1 == ch.size();
        if (last || (acc >= waves * target && k + 1 - k0 >= (size_t)waves)) {
          JKItem it;
          it.bpos = i;
          it.v0 = kofs[i] + (int)(k0 * csz);
          it.v1 = kofs[i] + std::min(len, (int)((k + 1) * csz));
          it.prims = p0 | (p1 << 16);
          items.push_back(it);
          icost.push_back(acc / waves + kItemOverhead);
          k0 = k + 1;
          acc = 0.0;
        }
      }
    }
  }
  std::vector<int> order(items.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return icost[a] > icost[b]; });
  std::vector<int> flat(4 * items.size());
  for (size_t k = 0; k < order.size(); ++k) {
    const JKItem& it = items[order[k]];
    flat[4 * k] = it.bpos;
    flat[4 * k + 1] = it.v0;
    flat[4 * k + 2] = it.v1;
    flat[4 * k + 3] = it.prims;
  }
  ws.n_items[combo] = (int)items.size();
  if (items.empty()) return GX_OK;
  GX_TRY(upload(&ws.d_items[combo], flat, c->stream));
  GX_TRY(upload(&ws.d_kets[combo], kets, c->stream));
  GX_HIP(hipStreamSynchronize(c->stream));  // before the host vectors go out of scope
  return GX_OK;
}

// Per-pair records (PairRec) of the current pair list on the device.
int build_pairrec(gx_ctx* c) {
  std::vector<gx::PairRec> rec(c->npair);
  std::vector<int> jkp_of((size_t)c->mtot, -1);
  std::vector<long> cbase(NCLS + 1, 0);
  for (int k = 0; k < NCLS; ++k)
    cbase[k + 1] = cbase[k] + (long)(c->cls_start[k + 1] - c->cls_start[k]) * ncart(class_la(k)) * ncart(class_lb(k));
  c->jkp_n = cbase[NCLS];
  for (int p = 0; p < c->npair; ++p) {
    const int A = c->pA[p], B = c->pB[p];
    gx::PairRec& r = rec[p];
    r.ABx = c->ctr[3 * A] - c->ctr[3 * B];
    r.ABy = c->ctr[3 * A + 1] - c->ctr[3 * B + 1];
    r.ABz = c->ctr[3 * A + 2] - c->ctr[3 * B + 2];
    r.fa = c->lbf[A];
    r.fb = c->lbf[B];
    r.poff = c->ppoff[p];
    r.np = c->pnp[p];
    r.A = A;
    r.B = B;
    r.sys = c->ssys[A];
    // class-major, then component-major: component k of the pair sits at
    // jko + k * (pairs of its class), so that consecutive pairs of a class
    // are consecutive for every component
    const int cls = c->pcls[p], ncls = c->cls_start[cls + 1] - c->cls_start[cls];
    const int na = ncart(c->l[A]), nb = ncart(c->l[B]);
    r.jko = (int)(cbase[cls] + (p - c->cls_start[cls]));
    for (int ia = 0; ia < na; ++ia)
      for (int ib = 0; ib < nb; ++ib)
        jkp_of[c->smb[r.sys] + (long)(r.fa + ia) * c->sn[r.sys] + r.fb + ib] = r.jko + (ia * nb + ib) * ncls;
  }
  GX_TRY(upload(&c->d_prec, rec, c->stream));
  GX_TRY(upload(&c->d_jkp_of, jkp_of, c->stream));
  GX_HIP(hipStreamSynchronize(c->stream));  // before the host vectors go out of scope
  return GX_OK;
}

int build_orders(gx_ctx* c, double cutoff, double fp32, int mixed, bool grad = false) {
  gx_ctx::WorkSet& ws = grad ? c->gr : c->jk;
  if (!mixed) fp32 = 0.0;
  if (cutoff == ws.cutoff && fp32 == ws.fp32 && mixed == ws.mixed && !ws.d_items.empty()) return GX_OK;
  ws.release();
  ws.d_items.assign(combo_count(), nullptr);
  ws.d_kets.assign(combo_count(), nullptr);
  ws.n_items.assign(combo_count(), 0);
  if ((int)ws.ipw.size() != 2 * combo_count()) ws.ipw.assign(2 * combo_count(), GX_ITEMS_PER_WG);
  if ((int)ws.tms.size() != 2 * combo_count()) ws.tms.assign(2 * combo_count(), 0.0);
  if ((int)c->jk_fp32.size() != combo_count()) c->jk_fp32.assign(combo_count(), 1);
  int combo = 0;
  for (int X = 0; X < NCLS; ++X)
    for (int Y = 0; Y <= X; ++Y, ++combo) {
      // gradients: the cooperative classes stay FP64; every mode uses the
      // item size tuned by gx_grad_autotune
      const int m = grad ? (mixed && gx_grad_is_lane(class_la(X), class_lb(X), class_la(Y), class_lb(Y))) : mixed;
      GX_TRY(build_combo(c, ws, X, Y, combo, cutoff, m ? fp32 : 0.0, m, grad,
                         ws.ipw[combo + (grad ? 0 : mixed) * combo_count()]));
    }
  ws.cutoff = cutoff;
  ws.fp32 = fp32;
  ws.mixed = mixed;
  return GX_OK;
}

// Nucleus ranges of the systems on the device (the single-system default
// covers every nucleus) and each nucleus's system.
int upload_sys_atoms(gx_ctx* c) {
  if (c->sys_atom_ok) return GX_OK;  // unchanged since the last upload (keeps graph capture legal)
  std::vector<int> sa(c->sys_atom);
  if (sa.size() == 2 && sa[1] < 0) sa[1] = c->natom;
  if ((int)sa.size() != c->nsys + 1 || sa[0] != 0 || sa[c->nsys] != c->natom)
    return fail(GX_ERR_STATE, "system nucleus ranges do not cover the " + std::to_string(c->natom) + " nuclei");
  std::vector<int> as(c->natom, 0);
  for (int k = 0; k < c->nsys; ++k)
    for (int a = sa[k]; a < sa[k + 1]; ++a) as[a] = k;
  for (int sh = 0; sh < c->nshell; ++sh) {
    const int at = c->atom[sh], k = c->ssys[sh];
    if (at >= 0 && (at < sa[k] || at >= sa[k + 1]))
      return fail(GX_ERR_STATE, "shell " + std::to_string(sh) + " sits on a nucleus of another system");
  }
  GX_TRY(upload(&c->d_sys_atom, sa, c->stream));
  GX_TRY(upload(&c->d_atom_sys, as, c->stream));
  GX_HIP(hipStreamSynchronize(c->stream));
  c->sys_atom_ok = true;
  return GX_OK;
}

gx::OneArgs one_args(const gx_ctx* c) {
  gx::OneArgs a;
  a.npair = c->npair;
  a.pA = c->d_pA;
  a.pB = c->d_pB;
  a.sl = c->d_l;
  a.snp = c->d_np;
  a.spoff = c->d_poff;
  a.sbf = c->d_bf;
  a.slbf = c->d_lbf;
  a.ssys = c->d_ssys;
  a.ctr1 = c->d_ctr1;
  a.exps = c->d_exps;
  a.coefs = c->d_coefs;
  a.scale = c->d_scale;
  a.sys_atom = c->d_sys_atom;
  a.nxyz = c->d_nxyz;
  a.nZ = c->d_nZ;
  a.master = c->d_boys;
  a.sy = c->systems();
  return a;
}

// The J/K kernel of class combination (X|Y) over its current work items.
// nden = 2: a two-density (unrestricted) build in FP64: J of Ps + Ps2 into
// Jt, K of Ps into Kt and of Ps2 into Kt2.
// Ps, Jt, Kt: the scaled density and accumulators (nullptr: the context's).
hipError_t launch_combo(gx_ctx* c, int X, int Y, int combo, int mixed, int nden = 1, hipStream_t st = nullptr,
                        double* Ps = nullptr, double* Jt = nullptr, double* Kt = nullptr) {
  if (!st) st = c->stream;
  if (!Ps) Ps = c->d_Ps;
  if (!Jt) Jt = c->d_Jt;
  if (!Kt) Kt = c->d_Kt;
  if ((size_t)combo < c->jk_fp32.size() && !c->jk_fp32[combo]) mixed = 0;  // this combination runs in FP64
  if (nden == 2) mixed = 0;
  gx::JKArgs a;
  a.sh = c->shells();
  a.pr = c->pairs();
  a.sy = c->systems();
  a.bra0 = c->cls_start[X];
  a.items = reinterpret_cast<const int4*>(c->jk.d_items[combo]);
  a.nitem = c->jk.n_items[combo];
  a.kets = c->jk.d_kets[combo];
  a.tab64 = c->d_tab64;
  a.tab32 = c->d_tab32;
  a.tab64c = c->d_tab64c;
  a.tab32c = c->d_tab32c;
  a.P = Ps;
  a.P2 = nden == 2 ? c->d_Ps2 : Ps;
  a.Pj = nden == 2 ? c->d_Psj : Ps;
  a.Jt = Jt;
  a.Jkp = Jt + c->mtot;
  a.Kt = Kt;
  a.Kt2 = nden == 2 ? c->d_Kt2 : Kt;
  a.item_ctr = c->d_item_ctr + 2 * combo;
  a.jstride = c->jkp_n;
  a.jkcs = c->cls_start[Y + 1] - c->cls_start[Y];
  a.mixed = mixed;
  if (c->wg_per_cu.empty()) c->wg_per_cu.assign(3 * combo_count(), 0);
  const int la = class_la(X), lb = class_lb(X), lc = class_la(Y), ld = class_lb(Y);
  int* per_cu = &c->wg_per_cu[3 * combo + (nden == 2 ? 2 : mixed)];
  if (nden == 2 && combo_is_coop(X, Y)) {
    // the cooperative kernels digest one density: run them once per density
    // (J is linear, so both launches add their part of J(P + P2) to Jt)
    const size_t lds = jk_lds_bytes(la, lb, lc, ld, c->nmax, 0);
    hipError_t e = launch_jk(la, lb, lc, ld, c->num_cu, lds, st, a, per_cu);
    if (e != hipSuccess) return e;
    a.P = a.Pj = c->d_Ps2;
    a.Kt = c->d_Kt2;
    return launch_jk(la, lb, lc, ld, c->num_cu, lds, st, a, per_cu);
  }
  return launch_jk(la, lb, lc, ld, c->num_cu, jk_lds_bytes(la, lb, lc, ld, c->nmax, mixed, nden), st, a, per_cu,
                   nden);
}

// One-electron gradient of the shell pairs of class (la, lb): one thread per
// (pair, primitive part) and system part (grid y: overlap/kinetic, then each
// nucleus of the pair's system).  Small jobs deal the primitives of shell A
// to four threads.
hipError_t launch_one_grad(gx_ctx* c, int la, int lb, const gx::OneArgs& oa, int p0, int np, int max_atoms,
                           const double* P, const double* W, double* grad) {
  const bool split = (long)np * (1 + max_atoms) < 200000;
#define GX_ONE_GRAD(A, B, KS)                                                                          \
  {                                                                                                    \
    hipLaunchKernelGGL((gx::k_one_grad_t<A, B, KS, false>), dim3((np * KS + 63) / 64, 1), dim3(64), 0, \
                       c->stream, oa, p0, np, c->d_atom, P, W, grad);                                 \
    if (max_atoms > 0)                                                                                 \
      hipLaunchKernelGGL((gx::k_one_grad_t<A, B, KS, true>), dim3((np * KS + 63) / 64, max_atoms),     \
                         dim3(64), 0, c->stream, oa, p0, np, c->d_atom, P, W, grad);                  \
  }
#define GX_CASE(A, B)                 \
  if (la == A && lb == B) {           \
    if (split) GX_ONE_GRAD(A, B, 4)   \
    else GX_ONE_GRAD(A, B, 1)         \
    return hipGetLastError();         \
  }
  GX_FOR_EACH_PAIRCLASS(GX_CASE)
#undef GX_CASE
#undef GX_ONE_GRAD
  return hipErrorInvalidValue;
}

// Two-electron gradient launch of one class combination (items of c->gr).
hipError_t launch_grad_combo(gx_ctx* c, int X, int Y, int combo, double* grad, hipStream_t st = nullptr) {
  if (!st) st = c->stream;
  if (c->gr.n_items[combo] == 0) return hipSuccess;
  if (c->wg_per_cu_grad.empty()) c->wg_per_cu_grad.assign(2 * combo_count(), 0);
  const int mixed = (c->gr.mixed && gx_grad_is_lane(class_la(X), class_lb(X), class_la(Y), class_lb(Y))) ? 1 : 0;
  gx::GradArgs a;
  a.sh = c->shells();
  a.pr = c->pairs();
  a.sy = c->systems();
  a.bra0 = c->cls_start[X];
  a.items = reinterpret_cast<const int4*>(c->gr.d_items[combo]);
  a.nitem = c->gr.n_items[combo];
  a.kets = c->gr.d_kets[combo];
  a.tab64 = c->d_tab64;
  a.tab32 = c->d_tab32;
  a.tab64c = c->d_tab64c;
  a.tab32c = c->d_tab32c;
  a.P = c->d_Ps;
  a.atom = c->d_atom;
  a.sys_atom = c->d_sys_atom;
  a.grad = grad;
  a.item_ctr = c->d_item_ctr + 2 * combo_count() + 2 * combo;
  return launch_grad(class_la(X), class_lb(X), class_la(Y), class_lb(Y), c->num_cu, st, a,
                     &c->wg_per_cu_grad[combo + mixed * combo_count()], c->nmax, mixed);
}

// Fork the context stream into the side streams (returns how many to use:
// 0 = launch everything on the context stream, as under graph capture).
int fork_side(gx_ctx* c) {
  if (GX_NSIDE <= 0) return 0;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return 0;
  if (!c->side_ready) {
    bool ok = true;
    for (int k = 0; k < GX_NSIDE && ok; ++k)
      ok = hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) == hipSuccess;
    if (!ok) {  // nothing half made stays behind; the next call tries again
      c->release_side();
      return -1;
    }
    c->side_ready = true;
  }
  if (hipEventRecord(c->ev_fork, c->stream) != hipSuccess) return -1;
  for (int k = 0; k < GX_NSIDE; ++k)
    if (hipStreamWaitEvent(c->side[k], c->ev_fork, 0) != hipSuccess) return -1;
  return GX_NSIDE;
}

// The context stream waits for the n side streams.
hipError_t join_side(gx_ctx* c, int n) {
  for (int k = 0; k < n; ++k) {
    hipError_t e = hipEventRecord(c->ev_join[k], c->side[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, c->ev_join[k], 0);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

extern "C" {

const char* gx_last_error(void) { return g_err.c_str(); }

int gx_create(gx_ctx** out, int device) {
  if (!out) return fail(GX_ERR_ARG, "gx_create: null output pointer");
  *out = nullptr;
  int ndev = 0;
  GX_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(GX_ERR_ARG, "gx_create: no such device");
  GX_HIP(hipSetDevice(device));
  // the no-alloc entry point's scratch address on this device, once
  GX_HIP(scratch_setup(device));
  gx_ctx* c = new gx_ctx();
  c->device = device;
  // the null stream until gx_set_stream; the GX_NSIDE side streams of
  // gx_jk / gx_jk_uhf / gx_grad are created on their first use (each context
  // holds its own set until gx_destroy)
  c->own = nullptr;
  c->stream = nullptr;
  int lds = 0, ncu = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
      lds > 0)
    c->max_lds = lds;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    c->num_cu = ncu;
  std::vector<double> tab, tabc;
  boys_master_table(tab, BOYS_NT, (long double)BOYS_DT);
  boys_master_table(tabc, gx::BoysCoarse::NT, 1.0L / gx::BoysCoarse::TPU);
  int rc = upload(&c->d_boys, tab, c->stream);
  if (rc == GX_OK) rc = upload(&c->d_boyst, tabc, c->stream);
  // highest L whose Taylor columns exist
  const int ltab = BOYS_MCOL - gx::BoysFine::K64, ltabc = BOYS_MCOL - gx::BoysCoarse::K64;
  if (rc == GX_OK) rc = dalloc(&c->d_tab64, (size_t)(ltab + 1) * BOYS_NT * BOYS_ROW64);
  if (rc == GX_OK) rc = dalloc(&c->d_tab32, (size_t)(ltab + 1) * BOYS_NT * BOYS_ROW32);
  if (rc == GX_OK) rc = dalloc(&c->d_tab64c, (size_t)(ltabc + 1) * gx::BoysCoarse::NT * gx::BoysCoarse::ROW64);
  if (rc == GX_OK) rc = dalloc(&c->d_tab32c, (size_t)(ltabc + 1) * gx::BoysCoarse::NT * gx::BoysCoarse::ROW32);
  if (rc == GX_OK) {
    hipLaunchKernelGGL(gx::k_boys_tables<gx::BoysFine>, dim3(ltab + 1), dim3(256), 0, c->stream, ltab, c->d_boys,
                       c->d_tab64, c->d_tab32);
    hipLaunchKernelGGL(gx::k_boys_tables<gx::BoysCoarse>, dim3(ltabc + 1), dim3(256), 0, c->stream, ltabc,
                       c->d_boyst, c->d_tab64c, c->d_tab32c);
    if (hipGetLastError() != hipSuccess) rc = fail(GX_ERR_HIP, "k_boys_tables launch");
  }
  // item counters start at zero; each J/K launch leaves them at zero
  if (rc == GX_OK) rc = dalloc(&c->d_item_ctr, (size_t)4 * combo_count());
  if (rc == GX_OK && hipMemsetAsync(c->d_item_ctr, 0, sizeof(int) * 4 * combo_count(), c->stream) != hipSuccess)
    rc = fail(GX_ERR_HIP, "gx_create: counter initialisation");
  if (rc == GX_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(GX_ERR_HIP, "boys upload");
  if (rc != GX_OK) {
    c->free_all();
    delete c;
    return rc;
  }
  *out = c;
  return GX_OK;
}

int gx_destroy(gx_ctx* c) {
  if (!c) return fail(GX_ERR_ARG, "gx_destroy: null context");
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->free_all();
  delete c;
  return GX_OK;
}

int gx_set_stream(gx_ctx* c, void* stream) {
  if (!c) return fail(GX_ERR_ARG, "gx_set_stream: null context");
  c->stream = stream ? (hipStream_t)stream : c->own;
  return GX_OK;
}

int gx_set_basis_batch(gx_ctx* c, int nsys, const int* sys_shell, const int* sys_atom, int nshell,
                       const int* l, const int* nprim, const int* prim_off, const int* bf_off, const double* ctr,
                       const double* ctr1, const int* atom, int nprim_total, const double* exps,
                       const double* coefs, int nbf, const double* bf_scale) {
  if (!c) return fail(GX_ERR_ARG, "gx_set_basis: null context");
  if (nshell <= 0 || nbf <= 0 || nprim_total <= 0 || !l || !nprim || !prim_off || !bf_off || !ctr ||
      !ctr1 || !atom || !exps || !coefs || !bf_scale)
    return fail(GX_ERR_ARG, "gx_set_basis: null array or non-positive size");
  if (nsys <= 0 || !sys_shell || sys_shell[0] != 0 || sys_shell[nsys] != nshell)
    return fail(GX_ERR_ARG, "gx_set_basis_batch: system shell ranges must start at 0 and end at nshell");
  for (int k = 0; k < nsys; ++k)
    if (sys_shell[k + 1] <= sys_shell[k])
      return fail(GX_ERR_ARG, "gx_set_basis_batch: system " + std::to_string(k) + " has no shells");
  if (sys_atom)
    for (int k = 0; k < nsys; ++k)
      if (sys_atom[k] < 0 || sys_atom[k + 1] < sys_atom[k])
        return fail(GX_ERR_ARG, "gx_set_basis_batch: system nucleus ranges must be non-decreasing");
  GX_HIP(hipSetDevice(c->device));
  int covered = 0;
  for (int s = 0; s < nshell; ++s) {
    if (l[s] < 0 || l[s] > GX_ERI_LMAX)
      return fail(GX_ERR_ARG, "gx_set_basis: shell angular momentum " + std::to_string(l[s]) +
                                  " not supported (max " + std::to_string(GX_ERI_LMAX) + ")");
    if (nprim[s] <= 0 || prim_off[s] < 0 || prim_off[s] + nprim[s] > nprim_total)
      return fail(GX_ERR_ARG, "gx_set_basis: primitive range of shell " + std::to_string(s));
    if (nprim[s] > 65535)
      return fail(GX_ERR_ARG, "gx_set_basis: too many primitives in shell " + std::to_string(s));
    if (bf_off[s] < 0 || bf_off[s] + ncart(l[s]) > nbf)
      return fail(GX_ERR_ARG, "gx_set_basis: basis-function range of shell " + std::to_string(s));
    covered += ncart(l[s]);
  }
  if (covered != nbf) return fail(GX_ERR_ARG, "gx_set_basis: shells do not cover nbf functions");
  for (int p = 0; p < nprim_total; ++p)
    if (!(exps[p] > 0.0)) return fail(GX_ERR_ARG, "gx_set_basis: non-positive exponent");
  // systems: contiguous function ranges in system order
  std::vector<int> sn(nsys, 0), sbfb(nsys, 0), ssys(nshell), lbf(nshell);
  std::vector<long> smb(nsys + 1, 0);
  int next_bf = 0, nmax = 0;
  for (int k = 0; k < nsys; ++k) {
    int lo = nbf, hi = 0;
    for (int sh = sys_shell[k]; sh < sys_shell[k + 1]; ++sh) {
      lo = std::min(lo, bf_off[sh]);
      hi = std::max(hi, bf_off[sh] + ncart(l[sh]));
      sn[k] += ncart(l[sh]);
      ssys[sh] = k;
    }
    if (lo != next_bf || hi - lo != sn[k])
      return fail(GX_ERR_ARG, "gx_set_basis_batch: functions of system " + std::to_string(k) +
                                  " are not one contiguous range after the previous system's");
    sbfb[k] = lo;
    next_bf = hi;
    smb[k + 1] = smb[k] + (long)sn[k] * sn[k];
    nmax = std::max(nmax, sn[k]);
    for (int sh = sys_shell[k]; sh < sys_shell[k + 1]; ++sh) lbf[sh] = bf_off[sh] - lo;
  }
  GX_HIP(hipStreamSynchronize(c->stream));
  c->free_pairs();
  c->nshell = nshell;
  c->nbf = nbf;
  c->nprim = nprim_total;
  c->nsys = nsys;
  c->sys_atom_ok = false;
  c->sys_shell.assign(sys_shell, sys_shell + nsys + 1);
  if (sys_atom) c->sys_atom.assign(sys_atom, sys_atom + nsys + 1);
  else c->sys_atom = {0, -1};  // one system with every nucleus
  c->ssys.swap(ssys);
  c->lbf.swap(lbf);
  c->sn.swap(sn);
  c->sbfb.swap(sbfb);
  c->smb.assign(smb.begin(), smb.end() - 1);
  c->mtot = smb[nsys];
  c->nmax = nmax;
  c->l.assign(l, l + nshell);
  c->np.assign(nprim, nprim + nshell);
  c->poff.assign(prim_off, prim_off + nshell);
  c->bf.assign(bf_off, bf_off + nshell);
  c->atom.assign(atom, atom + nshell);
  c->ctr.assign(ctr, ctr + 3 * nshell);
  c->ctr1.assign(ctr1, ctr1 + 3 * nshell);
  c->exps.assign(exps, exps + nprim_total);
  c->coefs.assign(coefs, coefs + nprim_total);
  c->scale.assign(bf_scale, bf_scale + nbf);
  GX_TRY(upload(&c->d_l, c->l, c->stream));
  GX_TRY(upload(&c->d_np, c->np, c->stream));
  GX_TRY(upload(&c->d_poff, c->poff, c->stream));
  GX_TRY(upload(&c->d_bf, c->bf, c->stream));
  GX_TRY(upload(&c->d_ctr, c->ctr, c->stream));
  GX_TRY(upload(&c->d_ctr1, c->ctr1, c->stream));
  GX_TRY(upload(&c->d_exps, c->exps, c->stream));
  GX_TRY(upload(&c->d_coefs, c->coefs, c->stream));
  GX_TRY(upload(&c->d_scale, c->scale, c->stream));
  GX_TRY(upload(&c->d_atom, c->atom, c->stream));
  GX_TRY(upload(&c->d_ssys, c->ssys, c->stream));
  GX_TRY(upload(&c->d_lbf, c->lbf, c->stream));
  GX_TRY(upload(&c->d_sn, c->sn, c->stream));
  GX_TRY(upload(&c->d_sbfb, c->sbfb, c->stream));
  GX_TRY(upload(&c->d_smb, c->smb, c->stream));
  GX_TRY(dalloc(&c->d_Ps, (size_t)c->mtot));
  GX_TRY(dalloc(&c->d_Kt, (size_t)c->mtot));
  // two-density buffers: reallocated at the new size on their next use
  dfree(&c->d_Ps2);
  dfree(&c->d_Psj);
  dfree(&c->d_Kt2);

  // Shell pairs inside each system, oriented l_A >= l_B, sorted by class,
  // system, then primitive-pair count.
  struct P {
    int A, B, cls, sys, kp;
  };
  std::vector<P> pl;
  for (int k = 0; k < nsys; ++k)
    for (int i = sys_shell[k]; i < sys_shell[k + 1]; ++i)
      for (int j = sys_shell[k]; j <= i; ++j) {
        int A = i, B = j;
        if (c->l[j] > c->l[i]) std::swap(A, B);
        pl.push_back(P{A, B, pair_class(c->l[A], c->l[B]), k, c->np[A] * c->np[B]});
      }
  std::stable_sort(pl.begin(), pl.end(), [](const P& a, const P& b) {
    if (a.cls != b.cls) return a.cls < b.cls;
    if (a.sys != b.sys) return a.sys < b.sys;
    return a.kp > b.kp;
  });
  c->npair = (int)pl.size();
  c->pA.resize(c->npair);
  c->pB.resize(c->npair);
  c->pcls.resize(c->npair);
  c->cls_start.assign(NCLS + 1, 0);
  for (int p = 0; p < c->npair; ++p) {
    c->pA[p] = pl[p].A;
    c->pB[p] = pl[p].B;
    c->pcls[p] = pl[p].cls;
  }
  for (int k = 0, p = 0; k <= NCLS; ++k) {
    while (p < c->npair && c->pcls[p] < k) ++p;
    c->cls_start[k] = p;
  }
  c->kept_off.clear();
  c->kept_ab.clear();
  GX_TRY(build_primpairs(c));
  if (c->prim_eps > 0.0) GX_TRY(screen_primpairs(c));
  GX_TRY(compute_schwarz(c));

  // Segment key per pair: its primitive-pair count, or floor(log2) of it when
  // a (class, system) has more than JK_MAXSEG distinct counts.  Within
  // (class, system, key) pairs are ordered by decreasing Schwarz bound.
  std::vector<int> key(c->npair);
  for (int p0 = 0; p0 < c->npair;) {
    int p1 = p0;
    const int cls = c->pcls[p0], sy = c->ssys[c->pA[p0]];
    while (p1 < c->npair && c->pcls[p1] == cls && c->ssys[c->pA[p1]] == sy) ++p1;
    std::vector<int> distinct;
    for (int p = p0; p < p1; ++p) distinct.push_back(c->pnp[p]);
    std::sort(distinct.begin(), distinct.end());
    distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
    const bool coarse = (int)distinct.size() > JK_MAXSEG;
    for (int p = p0; p < p1; ++p) {
      int kk = c->pnp[p];
      if (coarse) {
        int lg = 0;
        while ((2 << lg) <= kk) ++lg;
        kk = lg;
      }
      key[p] = kk;
    }
    p0 = p1;
  }
  std::vector<int> perm(c->npair);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
    if (c->pcls[a] != c->pcls[b]) return c->pcls[a] < c->pcls[b];
    const int sa = c->ssys[c->pA[a]], sb = c->ssys[c->pA[b]];
    if (sa != sb) return sa < sb;
    if (key[a] != key[b]) return key[a] > key[b];
    return c->Q[a] > c->Q[b];
  });
  {
    std::vector<int> A2(c->npair), B2(c->npair), C2(c->npair), K2(c->npair);
    std::vector<double> Q2(c->npair);
    for (int p = 0; p < c->npair; ++p) {
      A2[p] = c->pA[perm[p]];
      B2[p] = c->pB[perm[p]];
      C2[p] = c->pcls[perm[p]];
      Q2[p] = c->Q[perm[p]];
      K2[p] = key[perm[p]];
    }
    c->pA.swap(A2);
    c->pB.swap(B2);
    c->pcls.swap(C2);
    c->Q.swap(Q2);
    key.swap(K2);
    if (!c->kept_off.empty()) {
      std::vector<int> off2(c->npair + 1, 0), ab2;
      ab2.reserve(c->kept_ab.size());
      for (int p = 0; p < c->npair; ++p) {
        const int q = perm[p];
        ab2.insert(ab2.end(), c->kept_ab.begin() + c->kept_off[q], c->kept_ab.begin() + c->kept_off[q + 1]);
        off2[p + 1] = (int)ab2.size();
      }
      c->kept_off.swap(off2);
      c->kept_ab.swap(ab2);
    }
  }
  GX_TRY(build_primpairs(c));
  GX_TRY(upload(&c->d_Q, c->Q, c->stream));
  c->seg_lo.clear();
  c->seg_hi.clear();
  c->cseg.assign((size_t)NCLS * (nsys + 1), 0);
  {
    int p = 0;
    for (int k = 0; k < NCLS; ++k)
      for (int sy = 0; sy <= nsys; ++sy) {
        c->cseg[(size_t)k * (nsys + 1) + sy] = (int)c->seg_lo.size();
        if (sy == nsys) break;
        const int first = (int)c->seg_lo.size();
        while (p < c->cls_start[k + 1] && c->ssys[c->pA[p]] == sy) {
          int q = p;
          while (q < c->cls_start[k + 1] && c->ssys[c->pA[q]] == sy && key[q] == key[p]) ++q;
          c->seg_lo.push_back(p);
          c->seg_hi.push_back(q);
          p = q;
        }
        if ((int)c->seg_lo.size() - first > JK_MAXSEG)
          return fail(GX_ERR_STATE, "gx_set_basis: too many primitive-count segments");
      }
  }
  GX_TRY(build_pairrec(c));
  GX_TRY(dalloc(&c->d_Jt, (size_t)c->jt_doubles()));
  c->nab.assign(c->npair, 0);
  c->pre_pnp.assign(c->npair + 1, 0);
  c->pre_nab.assign(c->npair + 1, 0);
  for (int p = 0; p < c->npair; ++p) {
    const int na = ncart(c->l[c->pA[p]]), nb = ncart(c->l[c->pB[p]]);
    c->nab[p] = (c->pA[p] == c->pB[p]) ? na * (na + 1) / 2 : na * nb;
    c->pre_pnp[p + 1] = c->pre_pnp[p] + c->pnp[p];
    c->pre_nab[p + 1] = c->pre_nab[p] + c->nab[p];
  }
  c->pairmap.clear();
  if (nsys == 1) {
    c->pairmap.assign((size_t)nshell * nshell, -1);
    for (int p = 0; p < c->npair; ++p) {
      c->pairmap[(size_t)c->pA[p] * nshell + c->pB[p]] = p;
      c->pairmap[(size_t)c->pB[p] * nshell + c->pA[p]] = p;
    }
  }
  GX_HIP(hipStreamSynchronize(c->stream));
  return GX_OK;
}

int gx_set_basis(gx_ctx* c, int nshell, const int* l, const int* nprim, const int* prim_off,
                 const int* bf_off, const double* ctr, const double* ctr1, const int* atom,
                 int nprim_total, const double* exps, const double* coefs, int nbf,
                 const double* bf_scale) {
  const int sys_shell[2] = {0, nshell};
  return gx_set_basis_batch(c, 1, sys_shell, nullptr, nshell, l, nprim, prim_off, bf_off, ctr, ctr1, atom,
                            nprim_total, exps, coefs, nbf, bf_scale);
}

int gx_set_nuclei(gx_ctx* c, int natom, const double* xyz, const double* charge) {
  if (!c) return fail(GX_ERR_ARG, "gx_set_nuclei: null context");
  if (natom <= 0 || !xyz || !charge) return fail(GX_ERR_ARG, "gx_set_nuclei: bad arguments");
  GX_HIP(hipSetDevice(c->device));
  GX_HIP(hipStreamSynchronize(c->stream));
  c->natom = natom;
  c->sys_atom_ok = false;
  c->nxyz.assign(xyz, xyz + 3 * natom);
  c->nZ.assign(charge, charge + natom);
  GX_TRY(upload(&c->d_nxyz, c->nxyz, c->stream));
  GX_TRY(upload(&c->d_nZ, c->nZ, c->stream));
  GX_HIP(hipStreamSynchronize(c->stream));
  return GX_OK;
}

int gx_nbf(gx_ctx* c) { return c ? c->nbf : fail(GX_ERR_ARG, "gx_nbf: null context"); }
int gx_npair(gx_ctx* c) { return c ? c->npair : fail(GX_ERR_ARG, "gx_npair: null context"); }

int gx_set_prim_screen(gx_ctx* c, double eps) {
  if (!c) return fail(GX_ERR_ARG, "gx_set_prim_screen: null context");
  if (!(eps >= 0.0)) return fail(GX_ERR_ARG, "gx_set_prim_screen: threshold must be >= 0");
  c->prim_eps = eps;
  return GX_OK;
}

int gx_nprimpair(gx_ctx* c) { return c ? c->npp : fail(GX_ERR_ARG, "gx_nprimpair: null context"); }

int gx_schwarz(gx_ctx* c, double* Q, int* shellA, int* shellB) {
  if (!c) return fail(GX_ERR_ARG, "gx_schwarz: null context");
  if (c->npair == 0) return fail(GX_ERR_STATE, "gx_schwarz: no basis set");
  for (int p = 0; p < c->npair; ++p) {
    if (Q) Q[p] = c->Q[p];
    if (shellA) shellA[p] = c->pA[p];
    if (shellB) shellB[p] = c->pB[p];
  }
  return GX_OK;
}

int gx_eri_dense(gx_ctx* c, double* out) {
  if (!c || !out) return fail(GX_ERR_ARG, "gx_eri_dense: null argument");
  if (c->npair == 0) return fail(GX_ERR_STATE, "gx_eri_dense: no basis set");
  if (c->nsys != 1) return fail(GX_ERR_STATE, "gx_eri_dense: one system only");
  GX_HIP(hipSetDevice(c->device));
  const long N = c->nbf;
  const long n4 = N * N * N * N;
  hipLaunchKernelGGL(gx::k_fill, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, c->stream, n4, out,
                     0.0);
  GX_HIP(hipGetLastError());
  for (int X = 0; X < NCLS; ++X)
    for (int Y = 0; Y <= X; ++Y) {
      const int b0 = c->cls_start[X], nb = c->cls_start[X + 1] - b0;
      const int k0 = c->cls_start[Y], nk = c->cls_start[Y + 1] - k0;
      if (nb == 0 || nk == 0) continue;
      GX_HIP(launch_dense(class_la(X), class_lb(X), class_la(Y), class_lb(Y), c->stream, c->shells(),
                          c->pairs(), b0, nb, k0, nk, X == Y, c->d_boys, c->d_tab64, c->d_scale, out, N));
    }
  GX_HIP(hipStreamSynchronize(c->stream));
  return GX_OK;
}

int gx_eri_quartets(gx_ctx* c, int64_t n, const int32_t* ijkl, double* out) {
  if (!c || (n > 0 && (!ijkl || !out))) return fail(GX_ERR_ARG, "gx_eri_quartets: null argument");
  if (c->npair == 0) return fail(GX_ERR_STATE, "gx_eri_quartets: no basis set");
  if (c->nsys != 1) return fail(GX_ERR_STATE, "gx_eri_quartets: one system only");
  if (n == 0) return GX_OK;
  GX_HIP(hipSetDevice(c->device));
  std::vector<int> shell_of(c->nbf);
  for (int s = 0; s < c->nshell; ++s)
    for (int k = 0; k < ncart(c->l[s]); ++k) shell_of[c->bf[s] + k] = s;
  const int ncombo = combo_count();
  std::vector<std::vector<int>> tb(ncombo), tk(ncombo), tc(ncombo);
  std::vector<std::vector<double>> ts(ncombo);
  std::vector<std::vector<long>> tp(ncombo);
  auto combo_id = [](int X, int Y) { return X * (X + 1) / 2 + Y; };
  for (int64_t t = 0; t < n; ++t) {
    int f[4];
    for (int k = 0; k < 4; ++k) {
      f[k] = ijkl[4 * t + k];
      if (f[k] < 0 || f[k] >= c->nbf) return fail(GX_ERR_ARG, "gx_eri_quartets: index out of range");
    }
    int s[4], cc[4];
    for (int k = 0; k < 4; ++k) {
      s[k] = shell_of[f[k]];
      cc[k] = f[k] - c->bf[s[k]];
    }
    int bra = c->pairmap[(size_t)s[0] * c->nshell + s[1]];
    int ket = c->pairmap[(size_t)s[2] * c->nshell + s[3]];
    int ca = cc[0], cb = cc[1], ccc = cc[2], cd = cc[3];
    if (c->pA[bra] != s[0]) std::swap(ca, cb);
    if (c->pA[ket] != s[2]) std::swap(ccc, cd);
    if (c->pcls[bra] < c->pcls[ket]) {
      std::swap(bra, ket);
      std::swap(ca, ccc);
      std::swap(cb, cd);
    }
    const int X = c->pcls[bra], Y = c->pcls[ket];
    const int NB = ncart(class_lb(X)), NC = ncart(class_la(Y)), ND = ncart(class_lb(Y));
    const int id = combo_id(X, Y);
    tb[id].push_back(bra);
    tk[id].push_back(ket);
    tc[id].push_back(((ca * NB + cb) * NC + ccc) * ND + cd);
    ts[id].push_back(c->scale[f[0]] * c->scale[f[1]] * c->scale[f[2]] * c->scale[f[3]]);
    tp[id].push_back((long)t);
  }
  double* d_out = nullptr;
  GX_TRY(dalloc(&d_out, (size_t)n));
  int rc = GX_OK;
  std::vector<int*> keep_i;
  std::vector<double*> keep_d;
  std::vector<long*> keep_l;
  for (int X = 0; X < NCLS && rc == GX_OK; ++X)
    for (int Y = 0; Y <= X && rc == GX_OK; ++Y) {
      const int id = combo_id(X, Y);
      const int m = (int)tb[id].size();
      if (m == 0) continue;
      int *db = nullptr, *dk = nullptr, *dc = nullptr;
      double* ds = nullptr;
      long* dp = nullptr;
      rc = upload(&db, tb[id], c->stream);
      if (rc == GX_OK) rc = upload(&dk, tk[id], c->stream);
      if (rc == GX_OK) rc = upload(&dc, tc[id], c->stream);
      if (rc == GX_OK) rc = upload(&ds, ts[id], c->stream);
      if (rc == GX_OK) rc = upload(&dp, tp[id], c->stream);
      keep_i.push_back(db);
      keep_i.push_back(dk);
      keep_i.push_back(dc);
      keep_d.push_back(ds);
      keep_l.push_back(dp);
      if (rc != GX_OK) break;
      hipError_t e = launch_quartets(class_la(X), class_lb(X), class_la(Y), class_lb(Y), c->stream,
                                     c->shells(), c->pairs(), m, db, dk, dc, ds, dp, c->d_boys, c->d_tab64,
                                     d_out);
      if (e != hipSuccess) rc = fail(GX_ERR_HIP, std::string("k_quartets: ") + hipGetErrorString(e));
    }
  if (rc == GX_OK) {
    hipError_t e = hipMemcpyAsync(out, d_out, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(GX_ERR_HIP, std::string("gx_eri_quartets: ") + hipGetErrorString(e));
  } else {
    (void)hipStreamSynchronize(c->stream);
  }
  for (auto p : keep_i) dfree(&p);
  for (auto p : keep_d) dfree(&p);
  for (auto p : keep_l) dfree(&p);
  dfree(&d_out);
  return rc;
}

int gx_jk(gx_ctx* c, const double* P, double* J, double* K, double cutoff, double fp32_bound,
          int flags) {
  if (!c || !P || !J || !K) return fail(GX_ERR_ARG, "gx_jk: null argument");
  if (c->npair == 0) return fail(GX_ERR_STATE, "gx_jk: no basis set");
  if (!(cutoff >= 0.0) || !(fp32_bound >= 0.0)) return fail(GX_ERR_ARG, "gx_jk: negative threshold");
  GX_HIP(hipSetDevice(c->device));
  const int N = c->nmax;  // LDS rows are sized for the largest system
  const long nn = c->mtot;
  const int mixed = (flags & GX_JK_MIXED) ? 1 : 0;
  for (int X = 0; X < NCLS; ++X)
    for (int Y = 0; Y <= X; ++Y) {
      if (c->cls_start[X + 1] == c->cls_start[X] || c->cls_start[Y + 1] == c->cls_start[Y]) continue;
      const size_t lds = jk_lds_bytes(class_la(X), class_lb(X), class_la(Y), class_lb(Y), N, mixed);
      if (lds > (size_t)c->max_lds)
        return fail(GX_ERR_ARG, "gx_jk: basis too large for the LDS row cache (nbf=" + std::to_string(N) +
                                    ", needs " + std::to_string(lds) + " bytes)");
    }
  const unsigned g = (unsigned)((nn + 255) / 256);
  const long njt = c->jt_doubles();
  hipLaunchKernelGGL(gx::k_jk_prepare, dim3((unsigned)((std::max(nn, njt) + 255) / 256)), dim3(256), 0, c->stream,
                     nn, c->nsys, c->d_smb, c->d_sn, c->d_sbfb, c->d_scale, P, c->d_Ps, c->d_Kt, njt, c->d_Jt);
  GX_HIP(hipGetLastError());
  GX_TRY(build_orders(c, cutoff, fp32_bound, mixed));
  const bool timing = (flags & GX_JK_TIMING) != 0;
  // one event pair per launch; the pool grows on first use and is reused
  // after every gx_jk_launch_times read
  auto next_events = [&](int combo_id) -> hipError_t {
    const size_t i = c->ev_combo.size();
    while (c->ev.size() < 2 * i + 2) {
      hipEvent_t e;
      hipError_t err = hipEventCreate(&e);
      if (err != hipSuccess) return err;
      c->ev.push_back(e);
    }
    c->ev_combo.push_back(combo_id);
    return hipSuccess;
  };
  // independent class combinations run on the side streams (their tails and
  // small grids overlap); timed builds keep one stream so that each launch's
  // events bracket it alone
  const int nside = timing ? 0 : fork_side(c);
  if (nside < 0) return fail(GX_ERR_HIP, "gx_jk: side streams");
  for (const gx::PlanStep& p : gx::launch_plan(NCLS, c->jk.n_items, c->jk.tms, mixed, nside)) {
    if (timing) {
      GX_HIP(next_events(p.combo));
      GX_HIP(hipEventRecord(c->ev[2 * (c->ev_combo.size() - 1)], c->stream));
    }
    GX_HIP_FORKED(launch_combo(c, p.X, p.Y, p.combo, mixed, 1, p.side >= 0 ? c->side[p.side] : c->stream), nside);
    if (timing) GX_HIP(hipEventRecord(c->ev[2 * (c->ev_combo.size() - 1) + 1], c->stream));
  }
  GX_HIP(join_side(c, nside));
  hipLaunchKernelGGL(gx::k_symmetrize, dim3(g), dim3(256), 0, c->stream, nn, c->nsys, c->d_smb, c->d_sn,
                     c->d_sbfb, c->d_scale, c->d_Jt, c->d_Jt + nn, c->jkp_n, c->d_jkp_of, c->d_Kt, J, K);
  GX_HIP(hipGetLastError());
  return GX_OK;
}

int gx_jk_noalloc(gx_ctx* c, const double* P, double* J, double* K, double cutoff, double fp32_bound,
                  int flags) {
  if (!c || !P || !J || !K) return fail(GX_ERR_ARG, "gx_jk_noalloc: null argument");
  if (c->npair == 0) return fail(GX_ERR_STATE, "gx_jk_noalloc: no basis set");
  if (!(cutoff >= 0.0) || !(fp32_bound >= 0.0)) return fail(GX_ERR_ARG, "gx_jk_noalloc: negative threshold");
  if (flags & ~GX_JK_MIXED) return fail(GX_ERR_ARG, "gx_jk_noalloc: only GX_JK_MIXED is accepted");
  const int mixed = (flags & GX_JK_MIXED) ? 1 : 0;
  const long nn = c->mtot;
  const long need = 2L * nn + c->jt_doubles();
  if (need > GX_NOALLOC_MAX_DOUBLES)
    return fail(GX_ERR_ARG, "gx_jk_noalloc: the build needs " + std::to_string(need) +
                                " doubles of scratch (3 sum_k n_k^2 + " + std::to_string(GX_JT_COPIES) +
                                " x the pair blocks), at most " + std::to_string(GX_NOALLOC_MAX_DOUBLES));
  double* base = scratch_of(c->device);
  if (!base) return fail(GX_ERR_STATE, "gx_jk_noalloc: no scratch address for this device");
  // the work items are built (and uploaded) by gx_jk / gx_jk_autotune; this
  // call only uses them
  const double fp32 = mixed ? fp32_bound : 0.0;
  if (c->jk.d_items.empty() || c->jk.cutoff != cutoff || c->jk.fp32 != fp32 || c->jk.mixed != mixed)
    return fail(GX_ERR_STATE, "gx_jk_noalloc: no work items for these thresholds (call gx_jk with the same "
                              "cutoff, fp32_bound and flags once first)");
  const int N = c->nmax;
  for (int X = 0; X < NCLS; ++X)
    for (int Y = 0; Y <= X; ++Y) {
      if (c->cls_start[X + 1] == c->cls_start[X] || c->cls_start[Y + 1] == c->cls_start[Y]) continue;
      const size_t lds = jk_lds_bytes(class_la(X), class_lb(X), class_la(Y), class_lb(Y), N, mixed);
      if (lds > (size_t)c->max_lds)
        return fail(GX_ERR_ARG, "gx_jk_noalloc: basis too large for the LDS row cache (nbf=" +
                                    std::to_string(N) + ")");
    }
  GX_HIP(hipSetDevice(c->device));
  double* Ps = base;
  double* Kt = base + nn;
  double* Jt = base + 2 * nn;
  const unsigned g = (unsigned)((nn + 255) / 256);
  // the scratch keeps the previous call's contents: the accumulators are
  // cleared by kernels (no memset node, so a captured call stays replayable)
  const long njt = c->jt_doubles();
  hipLaunchKernelGGL(gx::k_jk_prepare, dim3((unsigned)((std::max(nn, njt) + 255) / 256)), dim3(256), 0, c->stream,
                     nn, c->nsys, c->d_smb, c->d_sn, c->d_sbfb, c->d_scale, P, Ps, Kt, njt, Jt);
  GX_HIP(hipGetLastError());
  int combo = 0;
  for (int X = 0; X < NCLS; ++X)
    for (int Y = 0; Y <= X; ++Y, ++combo) {
      if (c->jk.n_items[combo] == 0) continue;
      GX_HIP(launch_combo(c, X, Y, combo, mixed, 1, c->stream, Ps, Jt, Kt));
    }
  hipLaunchKernelGGL(gx::k_symmetrize, dim3(g), dim3(256), 0, c->stream, nn, c->nsys, c->d_smb, c->d_sn,
                     c->d_sbfb, c->d_scale, Jt, Jt + nn, c->jkp_n, c->d_jkp_of, Kt, J, K);
  GX_HIP(hipGetLastError());
  return GX_OK;
}

int gx_jk_uhf(gx_ctx* c, const double* Pa, const double* Pb, double* J, double* Ka, double* Kb, double cutoff,
              int flags) {
  if (!c || !Pa || !Pb || !J || !Ka || !Kb) return fail(GX_ERR_ARG, "gx_jk_uhf: null argument");
  if (c->npair == 0) return fail(GX_ERR_STATE, "gx_jk_uhf: no basis set");
  if (!(cutoff >= 0.0)) return fail(GX_ERR_ARG, "gx_jk_uhf: negative threshold");
  if (flags & GX_JK_MIXED) return fail(GX_ERR_ARG, "gx_jk_uhf: two-density builds are FP64 only");
  GX_HIP(hipSetDevice(c->device));
  const int N = c->nmax;
  const long nn = c->mtot;
  for (int X = 0; X < NCLS; ++X)
    for (int Y = 0; Y <= X; ++Y) {
      if (c->cls_start[X + 1] == c->cls_start[X] || c->cls_start[Y + 1] == c->cls_start[Y]) continue;
      const size_t lds = jk_lds_bytes(class_la(X), class_lb(X), class_la(Y), class_lb(Y), N, 0, 2);
      if (lds > (size_t)c->max_lds)
        return fail(GX_ERR_ARG, "gx_jk_uhf: basis too large for the LDS row cache (nbf=" + std::to_string(N) +
                                    ", needs " + std::to_string(lds) + " bytes)");
    }
  if (!c->d_Ps2) {
    GX_TRY(dalloc(&c->d_Ps2, (size_t)nn));
    GX_TRY(dalloc(&c->d_Psj, (size_t)nn));
    GX_TRY(dalloc(&c->d_Kt2, (size_t)nn));
  }
  const unsigned g = (unsigned)((nn + 255) / 256);
  hipLaunchKernelGGL(gx::k_scale_density, dim3(g), dim3(256), 0, c->stream, nn, c->nsys, c->d_smb, c->d_sn,
                     c->d_sbfb, c->d_scale, Pa, c->d_Ps);
  hipLaunchKernelGGL(gx::k_scale_density, dim3(g), dim3(256), 0, c->stream, nn, c->nsys, c->d_smb, c->d_sn,
                     c->d_sbfb, c->d_scale, Pb, c->d_Ps2);
  hipLaunchKernelGGL(gx::k_add, dim3(g), dim3(256), 0, c->stream, nn, c->d_Ps, c->d_Ps2, c->d_Psj);
  hipLaunchKernelGGL(gx::k_fill, dim3((unsigned)((c->jt_doubles() + 255) / 256)), dim3(256), 0, c->stream,
                     c->jt_doubles(), c->d_Jt, 0.0);
  // Jt and the Jkp copies; Kt and Kt2 one each
  hipLaunchKernelGGL(gx::k_fill, dim3(g), dim3(256), 0, c->stream, nn, c->d_Kt, 0.0);
  hipLaunchKernelGGL(gx::k_fill, dim3(g), dim3(256), 0, c->stream, nn, c->d_Kt2, 0.0);
  GX_HIP(hipGetLastError());
  GX_TRY(build_orders(c, cutoff, 0.0, 0));
  const int nside = fork_side(c);
  if (nside < 0) return fail(GX_ERR_HIP, "gx_jk_uhf: side streams");
  for (const gx::PlanStep& p : gx::launch_plan(NCLS, c->jk.n_items, c->jk.tms, 0, nside))
    GX_HIP_FORKED(launch_combo(c, p.X, p.Y, p.combo, 0, 2, p.side >= 0 ? c->side[p.side] : c->stream), nside);
  GX_HIP(join_side(c, nside));
  hipLaunchKernelGGL(gx::k_symmetrize, dim3(g), dim3(256), 0, c->stream, nn, c->nsys, c->d_smb, c->d_sn,
                     c->d_sbfb, c->d_scale, c->d_Jt, c->d_Jt + nn, c->jkp_n, c->d_jkp_of, c->d_Kt, J, Ka);
  hipLaunchKernelGGL(gx::k_symmetrize1, dim3(g), dim3(256), 0, c->stream, nn, c->nsys, c->d_smb, c->d_sn,
                     c->d_sbfb, c->d_scale, c->d_Kt2, Kb);
  GX_HIP(hipGetLastError());
  return GX_OK;
}

int gx_jk_autotune(gx_ctx* c, const double* P, double cutoff, double fp32_bound, int flags,
                   const double* cand, int ncand, int reps, double* out, int cap) {
  if (!c || !P || !cand || !out || ncand <= 0 || reps <= 0) return fail(GX_ERR_ARG, "gx_jk_autotune: bad argument");
  if (c->npair == 0) return fail(GX_ERR_STATE, "gx_jk_autotune: no basis set");
  for (int k = 0; k < ncand; ++k)
    if (!(cand[k] > 0.0)) return fail(GX_ERR_ARG, "gx_jk_autotune: candidates must be positive");
  const int mixed = (flags & GX_JK_MIXED) ? 1 : 0;
  // one untimed build: fills Ps, builds the default items and checks sizes
  double *dJ = nullptr, *dK = nullptr;
  GX_TRY(dalloc(&dJ, (size_t)c->mtot));
  int rc = dalloc(&dK, (size_t)c->mtot);
  if (rc == GX_OK) rc = gx_jk(c, P, dJ, dK, cutoff, fp32_bound, mixed);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == GX_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(GX_ERR_HIP, "gx_jk_autotune: hipEventCreate");
  int row = 0, combo = 0;
  for (int X = 0; X < NCLS && rc == GX_OK; ++X)
    for (int Y = 0; Y <= X && rc == GX_OK; ++Y, ++combo) {
      if (c->jk.n_items[combo] == 0) continue;
      if (row >= cap) {
        rc = fail(GX_ERR_ARG, "gx_jk_autotune: output too small");
        break;
      }
      double& ipw = c->jk.ipw[combo + mixed * combo_count()];
      double best = ipw, best_ms = 1e30;
      int best_fp32 = 1;
      // mixed builds also try the combination all in FP64 (kept where it is faster)
      for (int use32 = mixed; use32 >= 0 && rc == GX_OK; --use32) {
        if (mixed) c->jk_fp32[combo] = (char)use32;
        for (int k = 0; k < ncand && rc == GX_OK; ++k) {
          rc = build_combo(c, c->jk, X, Y, combo, cutoff, fp32_bound, mixed, false, cand[k]);
          if (rc != GX_OK) break;
          double ms_min = 1e30;
          for (int r = 0; r < reps; ++r) {
            float ms = 0.f;
            if (hipEventRecord(e0, c->stream) != hipSuccess || launch_combo(c, X, Y, combo, mixed) != hipSuccess ||
                hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
              rc = fail(GX_ERR_HIP, "gx_jk_autotune: timed launch");
              break;
            }
            ms_min = std::min(ms_min, (double)ms);
          }
          if (ms_min < best_ms) best_ms = ms_min, best = cand[k], best_fp32 = use32;
        }
      }
      if (rc != GX_OK) break;
      ipw = best;
      c->jk.tms[combo + mixed * combo_count()] = best_ms;
      if (mixed) c->jk_fp32[combo] = (char)best_fp32;
      rc = build_combo(c, c->jk, X, Y, combo, cutoff, fp32_bound, mixed, false, best);
      double* r = out + 6 * row++;
      r[0] = class_la(X);
      r[1] = class_lb(X);
      r[2] = class_la(Y);
      r[3] = class_lb(Y);
      r[4] = best;
      r[5] = best_ms;
      if (mixed && !best_fp32) r[4] = -best;  // negative: the combination runs in FP64
    }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dfree(&dJ);
  dfree(&dK);
  return rc == GX_OK ? row : rc;
}

int gx_grad(gx_ctx* c, const double* P, const double* W, double* grad, double cutoff) {
  return gx_grad_ex(c, P, W, grad, cutoff, 0.0, 0);
}

int gx_grad_ex(gx_ctx* c, const double* P, const double* W, double* grad, double cutoff, double fp32_bound,
               int flags) {
  if (!c || !P || !W || !grad) return fail(GX_ERR_ARG, "gx_grad: null argument");
  if (c->npair == 0) return fail(GX_ERR_STATE, "gx_grad: no basis set");
  if (c->natom == 0) return fail(GX_ERR_STATE, "gx_grad: no nuclei set");
  if (!(cutoff >= 0.0) || !(fp32_bound >= 0.0)) return fail(GX_ERR_ARG, "gx_grad: negative threshold");
  const int mixed = (flags & GX_JK_MIXED) ? 1 : 0;
  const bool timing = (flags & GX_JK_TIMING) != 0;
  for (int sh = 0; sh < c->nshell; ++sh)
    if (c->atom[sh] < 0 || c->atom[sh] >= c->natom)
      return fail(GX_ERR_STATE, "gx_grad: shell " + std::to_string(sh) + " has no nucleus (atom index " +
                                    std::to_string(c->atom[sh]) + ")");
  for (int X = 0; X < NCLS; ++X)
    for (int Y = 0; Y <= X; ++Y) {
      if (c->cls_start[X + 1] == c->cls_start[X] || c->cls_start[Y + 1] == c->cls_start[Y]) continue;
      const size_t lds = grad_lds_bytes(class_la(X), class_lb(X), class_la(Y), class_lb(Y), c->nmax);
      if (lds > (size_t)c->max_lds)
        return fail(GX_ERR_ARG, "gx_grad: class (" + std::to_string(class_la(X)) + std::to_string(class_lb(X)) +
                                    "|" + std::to_string(class_la(Y)) + std::to_string(class_lb(Y)) +
                                    ") needs " + std::to_string(lds) + " bytes of LDS, the device has " +
                                    std::to_string(c->max_lds));
    }
  GX_TRY(upload_sys_atoms(c));
  GX_HIP(hipSetDevice(c->device));
  const long nn = c->mtot;
  const unsigned g = (unsigned)((nn + 255) / 256);
  hipLaunchKernelGGL(gx::k_scale_density, dim3(g), dim3(256), 0, c->stream, nn, c->nsys, c->d_smb, c->d_sn,
                     c->d_sbfb, c->d_scale, P, c->d_Ps);
  hipLaunchKernelGGL(gx::k_fill, dim3((3 * c->natom + 255) / 256), dim3(256), 0, c->stream, (long)3 * c->natom,
                     grad, 0.0);
  const gx::OneArgs oa = one_args(c);
  int max_atoms = 0;
  for (int k = 0; k < c->nsys; ++k) {
    const int hi = (c->sys_atom.size() == 2 && c->sys_atom[1] < 0) ? c->natom : c->sys_atom[k + 1];
    max_atoms = std::max(max_atoms, hi - c->sys_atom[k]);
  }
  if (max_atoms > GX_GRAD_MAXATOM)
    return fail(GX_ERR_ARG, "gx_grad: a system has " + std::to_string(max_atoms) + " nuclei (at most " +
                                std::to_string(GX_GRAD_MAXATOM) + ")");
  GX_HIP(hipGetLastError());
  GX_TRY(build_orders(c, cutoff, fp32_bound, mixed, true));
  // timed calls keep every kernel on the context stream, bracketed by its events
  const int nside = timing ? 0 : fork_side(c);
  if (nside < 0) return fail(GX_ERR_HIP, "gx_grad: side streams");
  // the one-electron terms (atomics into grad) on the context stream, beside
  // the two-electron kernels on the side streams
  for (int k = 0; k < NCLS; ++k) {
    const int p0 = c->cls_start[k], np = c->cls_start[k + 1] - p0;
    if (np > 0) GX_HIP_FORKED(launch_one_grad(c, class_la(k), class_lb(k), oa, p0, np, max_atoms, P, W, grad), nside);
  }
  for (const gx::PlanStep& p : gx::launch_plan(NCLS, c->gr.n_items, c->gr.tms, 0, nside)) {
    if (timing) {
      const size_t i = c->gev_combo.size();
      while (c->gev.size() < 2 * i + 2) {
        hipEvent_t e;
        GX_HIP(hipEventCreate(&e));
        c->gev.push_back(e);
      }
      c->gev_combo.push_back(p.combo);
      GX_HIP(hipEventRecord(c->gev[2 * i], c->stream));
    }
    GX_HIP_FORKED(launch_grad_combo(c, p.X, p.Y, p.combo, grad, p.side >= 0 ? c->side[p.side] : c->stream), nside);
    if (timing) GX_HIP(hipEventRecord(c->gev[2 * (c->gev_combo.size() - 1) + 1], c->stream));
  }
  GX_HIP(join_side(c, nside));
  // the nuclear repulsion term adds without atomics: after every other term
  hipLaunchKernelGGL(gx::k_nuc_grad, dim3((c->natom + 63) / 64), dim3(64), 0, c->stream, c->natom, c->d_atom_sys,
                     c->d_sys_atom, c->d_nxyz, c->d_nZ, grad);
  GX_HIP(hipGetLastError());
  return GX_OK;
}

int gx_grad_launch_times(gx_ctx* c, double* out, int cap) {
  if (!c || !out) return fail(GX_ERR_ARG, "gx_grad_launch_times: null argument");
  GX_HIP(hipSetDevice(c->device));
  const int nc = combo_count();
  std::vector<double> ms(nc, 0.0), cnt(nc, 0.0);
  if (!c->gev_combo.empty()) GX_HIP(hipEventSynchronize(c->gev[2 * c->gev_combo.size() - 1]));
  for (size_t i = 0; i < c->gev_combo.size(); ++i) {
    float t = 0.f;
    GX_HIP(hipEventElapsedTime(&t, c->gev[2 * i], c->gev[2 * i + 1]));
    ms[c->gev_combo[i]] += t;
    cnt[c->gev_combo[i]] += 1.0;
  }
  c->gev_combo.clear();
  int row = 0, combo = 0;
  for (int X = 0; X < NCLS; ++X)
    for (int Y = 0; Y <= X; ++Y, ++combo) {
      if (cnt[combo] == 0.0) continue;
      if (row >= cap) return fail(GX_ERR_ARG, "gx_grad_launch_times: output too small");
      double* r = out + 6 * row++;
      r[0] = class_la(X);
      r[1] = class_lb(X);
      r[2] = class_la(Y);
      r[3] = class_lb(Y);
      r[4] = ms[combo];
      r[5] = cnt[combo];
    }
  return row;
}

int gx_grad_autotune(gx_ctx* c, const double* P, const double* W, double cutoff, const double* cand, int ncand,
                     int reps, double* out, int cap) {
  if (!c || !P || !W || !cand || !out || ncand <= 0 || reps <= 0)
    return fail(GX_ERR_ARG, "gx_grad_autotune: bad argument");
  for (int k = 0; k < ncand; ++k)
    if (!(cand[k] > 0.0)) return fail(GX_ERR_ARG, "gx_grad_autotune: candidates must be positive");
  double* dg = nullptr;
  GX_TRY(dalloc(&dg, (size_t)3 * std::max(c->natom, 1)));
  // one untimed gradient: scaled density, default items, sizes checked
  int rc = gx_grad(c, P, W, dg, cutoff);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == GX_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(GX_ERR_HIP, "gx_grad_autotune: hipEventCreate");
  int row = 0, combo = 0;
  for (int X = 0; X < NCLS && rc == GX_OK; ++X)
    for (int Y = 0; Y <= X && rc == GX_OK; ++Y, ++combo) {
      if (c->gr.n_items[combo] == 0) continue;
      if (row >= cap) {
        rc = fail(GX_ERR_ARG, "gx_grad_autotune: output too small");
        break;
      }
      double& ipw = c->gr.ipw[combo];
      double best = ipw, best_ms = 1e30;
      for (int k = 0; k < ncand && rc == GX_OK; ++k) {
        rc = build_combo(c, c->gr, X, Y, combo, cutoff, 0.0, 0, true, cand[k]);
        if (rc != GX_OK) break;
        double ms_min = 1e30;
        for (int r = 0; r < reps; ++r) {
          float ms = 0.f;
          if (hipEventRecord(e0, c->stream) != hipSuccess || launch_grad_combo(c, X, Y, combo, dg) != hipSuccess ||
              hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(GX_ERR_HIP, "gx_grad_autotune: timed launch");
            break;
          }
          ms_min = std::min(ms_min, (double)ms);
        }
        if (ms_min < best_ms) best_ms = ms_min, best = cand[k];
      }
      if (rc != GX_OK) break;
      ipw = best;
      c->gr.tms[combo] = best_ms;
      rc = build_combo(c, c->gr, X, Y, combo, cutoff, 0.0, 0, true, best);
      double* r = out + 6 * row++;
      r[0] = class_la(X);
      r[1] = class_lb(X);
      r[2] = class_la(Y);
      r[3] = class_lb(Y);
      r[4] = best;
      r[5] = best_ms;
    }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(c->stream);
  dfree(&dg);
  return rc == GX_OK ? row : rc;
}

int gx_jk_stats(gx_ctx* c, double cutoff, double fp32_bound, int flags, gx_stats* out) {
  if (!c || !out) return fail(GX_ERR_ARG, "gx_jk_stats: null argument");
  if (c->npair == 0) return fail(GX_ERR_STATE, "gx_jk_stats: no basis set");
  std::memset(out, 0, sizeof(*out));
  const int mixed = (flags & GX_JK_MIXED) ? 1 : 0;
  auto bfq = [&](int ip, int lo, int hi) -> int64_t {
    // basis-function quartets of bra ip with kets [lo, hi); the diagonal
    // quartet (ip|ip) counts nab (nab + 1) / 2 instead of nab^2
    int64_t n = (int64_t)c->nab[ip] * (c->pre_nab[hi] - c->pre_nab[lo]);
    if (ip >= lo && ip < hi) n -= (int64_t)c->nab[ip] * c->nab[ip] - (int64_t)c->nab[ip] * (c->nab[ip] + 1) / 2;
    return n;
  };
  for (int X = 0; X < NCLS; ++X)
    for (int Y = 0; Y <= X; ++Y) {
      for_each_bra_segment(c, X, Y, 0.0, 0.0, 0, [&](int, int ip, int lo, int hi, int, int) {
        out->shell_quartets += hi - lo;
        out->bf_quartets += bfq(ip, lo, hi);
      });
      for_each_bra_segment(c, X, Y, cutoff, fp32_bound, mixed, [&](int, int ip, int lo, int, int n64, int n32) {
        out->shell_quartets_fp64 += n64;
        out->shell_quartets_fp32 += n32;
        out->bf_quartets_fp64 += bfq(ip, lo, lo + n64);
        out->bf_quartets_fp32 += bfq(ip, lo + n64, lo + n64 + n32);
      });
    }
  return GX_OK;
}

int gx_jk_class_counts(gx_ctx* c, double cutoff, double fp32_bound, int flags, int64_t* out, int cap) {
  if (!c || !out) return fail(GX_ERR_ARG, "gx_jk_class_counts: null argument");
  if (c->npair == 0) return fail(GX_ERR_STATE, "gx_jk_class_counts: no basis set");
  const int mixed = (flags & GX_JK_MIXED) ? 1 : 0;
  int row = 0;
  for (int X = 0; X < NCLS; ++X)
    for (int Y = 0; Y <= X; ++Y) {
      if (row >= cap) return fail(GX_ERR_ARG, "gx_jk_class_counts: output too small");
      int64_t* r = out + 10 * row;
      r[0] = class_la(X);
      r[1] = class_lb(X);
      r[2] = class_la(Y);
      r[3] = class_lb(Y);
      for (int k = 4; k < 10; ++k) r[k] = 0;
      for_each_bra_segment(c, X, Y, cutoff, fp32_bound, mixed, [&](int, int ip, int lo, int, int n64, int n32) {
        r[4] += n64;
        r[5] += n32;
        r[6] += (int64_t)c->pnp[ip] * (c->pre_pnp[lo + n64] - c->pre_pnp[lo]);
        r[7] += (int64_t)c->pnp[ip] * (c->pre_pnp[lo + n64 + n32] - c->pre_pnp[lo + n64]);
        int64_t b64 = (int64_t)c->nab[ip] * (c->pre_nab[lo + n64] - c->pre_nab[lo]);
        if (ip >= lo && ip < lo + n64) b64 -= (int64_t)c->nab[ip] * (c->nab[ip] - 1) / 2;
        int64_t b32 = (int64_t)c->nab[ip] * (c->pre_nab[lo + n64 + n32] - c->pre_nab[lo + n64]);
        if (ip >= lo + n64 && ip < lo + n64 + n32) b32 -= (int64_t)c->nab[ip] * (c->nab[ip] - 1) / 2;
        r[8] += b64;
        r[9] += b32;
      });
      ++row;
    }
  return row;
}

int gx_jk_launch_times(gx_ctx* c, double* out, int cap) {
  if (!c || !out) return fail(GX_ERR_ARG, "gx_jk_launch_times: null argument");
  GX_HIP(hipSetDevice(c->device));
  const int nc = combo_count();
  std::vector<double> ms(nc, 0.0), cnt(nc, 0.0);
  if (!c->ev_combo.empty()) GX_HIP(hipEventSynchronize(c->ev[2 * c->ev_combo.size() - 1]));
  for (size_t i = 0; i < c->ev_combo.size(); ++i) {
    float t = 0.f;
    GX_HIP(hipEventElapsedTime(&t, c->ev[2 * i], c->ev[2 * i + 1]));
    ms[c->ev_combo[i]] += t;
    cnt[c->ev_combo[i]] += 1.0;
  }
  c->ev_combo.clear();
  int row = 0, combo = 0;
  for (int X = 0; X < NCLS; ++X)
    for (int Y = 0; Y <= X; ++Y, ++combo) {
      if (cnt[combo] == 0.0) continue;
      if (row >= cap) return fail(GX_ERR_ARG, "gx_jk_launch_times: output too small");
      double* r = out + 6 * row++;
      r[0] = class_la(X);
      r[1] = class_lb(X);
      r[2] = class_la(Y);
      r[3] = class_lb(Y);
      r[4] = ms[combo];
      r[5] = cnt[combo];
    }
  return row;
}

int gx_kernel_kind(int la, int lb, int lc, int ld, int gradient) {
  if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > GX_ERI_LMAX || lb > la || lc > GX_ERI_LMAX || ld > lc ||
      pair_class(lc, ld) > pair_class(la, lb))
    return fail(GX_ERR_ARG, "gx_kernel_kind: not a canonical class combination");
  if (gradient) return gx_grad_is_lane(la, lb, lc, ld) ? 0 : 1;
  return gx_class_is_coop(la, lb, lc, ld) ? 1 : 0;
}

int gx_boys_values(gx_ctx* c, int grid, int fp32, int L, const double* T, int n, double* out) {
  if (!c || (n > 0 && (!T || !out))) return fail(GX_ERR_ARG, "gx_boys_values: null argument");
  if (grid < 0 || grid > 1 || fp32 < 0 || fp32 > 1 || L < 0 || L > 10 || n < 0)
    return fail(GX_ERR_ARG, "gx_boys_values: grid 0 or 1, fp32 0 or 1, 0 <= L <= 10, n >= 0");
  if (n == 0) return GX_OK;
  GX_HIP(hipSetDevice(c->device));
  const unsigned g = (unsigned)((n + 255) / 256);
#define GX_BOYS_CASE(LL)                                                                                  \
  case LL:                                                                                               \
    if (grid == 0 && !fp32)                                                                              \
      hipLaunchKernelGGL((gx::k_boys_values<LL, gx::BoysFine, false>), dim3(g), dim3(256), 0, c->stream, n, \
                         T, c->d_tab64, c->d_tab32, out);                                               \
    else if (grid == 0)                                                                                  \
      hipLaunchKernelGGL((gx::k_boys_values<LL, gx::BoysFine, true>), dim3(g), dim3(256), 0, c->stream, n,  \
                         T, c->d_tab64, c->d_tab32, out);                                               \
    else if (!fp32)                                                                                      \
      hipLaunchKernelGGL((gx::k_boys_values<LL, gx::BoysCoarse, false>), dim3(g), dim3(256), 0, c->stream, \
                         n, T, c->d_tab64c, c->d_tab32c, out);                                           \
    else                                                                                                 \
      hipLaunchKernelGGL((gx::k_boys_values<LL, gx::BoysCoarse, true>), dim3(g), dim3(256), 0, c->stream,  \
                         n, T, c->d_tab64c, c->d_tab32c, out);                                           \
    break;
  switch (L) {
    GX_BOYS_CASE(0) GX_BOYS_CASE(1) GX_BOYS_CASE(2) GX_BOYS_CASE(3) GX_BOYS_CASE(4) GX_BOYS_CASE(5)
    GX_BOYS_CASE(6) GX_BOYS_CASE(7) GX_BOYS_CASE(8) GX_BOYS_CASE(9) GX_BOYS_CASE(10)
  }
#undef GX_BOYS_CASE
  GX_HIP(hipGetLastError());
  return GX_OK;
}

int gx_one_electron(gx_ctx* c, double* S, double* T, double* V) {
  if (!c || !S || !T) return fail(GX_ERR_ARG, "gx_one_electron: null argument");
  if (c->npair == 0) return fail(GX_ERR_STATE, "gx_one_electron: no basis set");
  if (V && c->natom == 0) return fail(GX_ERR_STATE, "gx_one_electron: no nuclei set");
  if (c->natom > 0) GX_TRY(upload_sys_atoms(c));
  GX_HIP(hipSetDevice(c->device));
  gx::OneArgs oa = one_args(c);
  if (c->natom == 0) oa.sys_atom = nullptr;
  hipLaunchKernelGGL(gx::k_one_electron, dim3((c->npair + 63) / 64), dim3(64), 0, c->stream, oa, S, T, V);
  GX_HIP(hipGetLastError());
  return GX_OK;
}

}  // extern "C"
