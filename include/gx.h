/*
 * gx -- MI355X (gfx950) two-electron-integral and J/K Fock-build engine.
 *
 * C ABI: plain pointers and sizes only.  Every call returns GX_OK (0) or a
 * negative error code and then leaves a message for gx_last_error() (per
 * host thread).  A context is bound to one device and one HIP stream; it is
 * not thread-safe.  Unless stated otherwise a call returns after its GPU
 * work has been enqueued on the context stream and, for host outputs, after
 * that stream has drained.
 *
 * Reference interfaces replaced (paths relative to src/gaussium/ of the
 * reference, ChiCheng45/Gaussium):
 *   gx_set_basis      Basis / PrimitiveBasis data model,
 *                     objects/basis_function.py:6-53,
 *                     objects/primitive_basis_function.py:6-42 (normalisation
 *                     is folded into coefs and bf_scale by the caller)
 *   gx_eri_quartets   TwoElectronRepulsion.calculate_integral
 *                     matrixelements/electron_repulsion_matrix.py:18-26 and
 *                     ObaraSaika.integrate
 *                     integrals/twoelectronrepulsion/obara_saika_scheme.py:17-82
 *   gx_eri_dense      TwoElectronRepulsion.create_repulsion_matrix
 *                     matrixelements/electron_repulsion_matrix.py:28-47
 *   gx_jk             FockMatrixRestricted.create (the G = J - K/2 part)
 *                     hartreefock/fock_matrix.py:12-22
 *   gx_one_electron   OrbitalOverlapMatrix / KineticEnergyMatrix /
 *                     NuclearAttractionMatrix .create
 *                     matrixelements/orbital_overlap_matrix.py:15-27,
 *                     matrixelements/kinetic_energy_matrix.py:16-33,
 *                     matrixelements/nuclear_attraction_matrix.py:16-29
 *   gx_grad*          no counterpart (the reference has no gradients; they
 *                     replace the function-count-driven Nelder-Mead search,
 *                     geometryoptimization/nelder_mead.py:28-144)
 */
#ifndef GX_H
#define GX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GX_OK 0
#define GX_ERR_ARG (-1)      /* null pointer, bad size, unsupported input   */
#define GX_ERR_HIP (-2)      /* a HIP runtime call failed                    */
#define GX_ERR_STATE (-3)    /* call out of order (e.g. no basis set yet)    */
#define GX_ERR_NOMEM (-4)    /* device allocation failed                     */

/* gx_jk flags */
#define GX_JK_MIXED 1        /* quartets with Schwarz bound < fp32_bound in FP32 */
#define GX_JK_TIMING 2       /* record HIP events around each class-combination launch */

/* Default primitive-pair screening threshold (gx_set_prim_screen). */
#define GX_PRIM_EPS_DEFAULT 1e-18

typedef struct gx_ctx gx_ctx;

typedef struct gx_stats {
  int64_t shell_quartets;       /* canonical shell quartets                       */
  int64_t bf_quartets;          /* canonical basis-function quartets               */
  int64_t shell_quartets_fp64;  /* surviving the cutoff, evaluated in FP64         */
  int64_t shell_quartets_fp32;  /* surviving the cutoff, evaluated in FP32         */
  int64_t bf_quartets_fp64;
  int64_t bf_quartets_fp32;
  double flops_fp64;            /* model flops of the evaluated FP64 quartets      */
  double flops_fp32;            /* model flops of the evaluated FP32 quartets      */
} gx_stats;

/* Last error message of this host thread ("" if none). */
const char* gx_last_error(void);

/* Create a context on HIP device `device`.  It launches on the device's null
 * stream until gx_set_stream names another. */
int gx_create(gx_ctx** ctx, int device);
int gx_destroy(gx_ctx* ctx);

/* Launch on `stream` (a hipStream_t, e.g. torch's current stream); NULL
 * restores the null stream. */
int gx_set_stream(gx_ctx* ctx, void* stream);

/* Basis set as shells of Cartesian functions (all host arrays).
 *   l[s], nprim[s], prim_off[s]  angular momentum, primitive count/offset
 *   bf_off[s]                    index of the shell's first basis function;
 *                                components follow in the order
 *                                s | x,y,z | xx,yy,zz,xy,xz,yz
 *   ctr[3s]                      centre used by the two-electron integrals
 *   ctr1[3s]                     centre used by the one-electron integrals
 *   atom[s]                      atom index of the shell (gradients)
 *   exps[p], coefs[p]            primitive exponent and radial coefficient
 *                                (contraction x primitive norm x contracted norm
 *                                of the shell's first component)
 *   bf_scale[nbf]                per-function factor relative to that component
 * Builds primitive pairs and Schwarz bounds on the device.  l <= 2 (s, p,
 * d Cartesian shells). */
int gx_set_basis(gx_ctx* ctx, int nshell, const int* l, const int* nprim, const int* prim_off,
                 const int* bf_off, const double* ctr, const double* ctr1, const int* atom,
                 int nprim_total, const double* exps, const double* coefs, int nbf,
                 const double* bf_scale);

/* A batch of independent molecules (systems) in one context, so that one
 * launch per class combination builds J/K (or gradients) for all of them.
 * Shells [sys_shell[k], sys_shell[k + 1]) form system k (sys_shell[0] = 0,
 * sys_shell[nsys] = nshell); their functions must be one contiguous range
 * following the previous system's.  Nuclei [sys_atom[k], sys_atom[k + 1])
 * of the gx_set_nuclei array belong to system k (NULL: one system with all
 * nuclei).  Quartets never mix systems.  Every nbf x nbf matrix of the other
 * calls (P, J, K, W, S, T, V) is then the systems' matrices packed one after
 * the other, row-major, sum_k n_k^2 doubles.  gx_set_basis is the one-system
 * case.  No counterpart in the reference, which builds one molecule at a
 * time (geometryoptimization/nelder_mead.py:126-144 evaluates simplex
 * vertices one after the other). */
int gx_set_basis_batch(gx_ctx* ctx, int nsys, const int* sys_shell, const int* sys_atom, int nshell,
                       const int* l, const int* nprim, const int* prim_off, const int* bf_off, const double* ctr,
                       const double* ctr1, const int* atom, int nprim_total, const double* exps,
                       const double* coefs, int nbf, const double* bf_scale);

/* Primitive-pair screening threshold for the following gx_set_basis /
 * gx_set_basis_batch.  A primitive pair (a, b) of a shell pair is dropped
 * when S_ab * S_max < eps, with S_ab = sqrt(max over components |(ab|ab)|)
 * of the scaled primitive product and S_max the largest S over the basis:
 * by the Schwarz inequality every primitive quartet left out is below eps,
 * so a contracted integral moves by less than eps times its number of
 * primitive quartets.  Each shell pair keeps at least its largest primitive
 * pair.  eps = 0 keeps every primitive pair.  Default GX_PRIM_EPS_DEFAULT.
 * No counterpart in the reference, which sums every primitive quartet
 * (integrals/twoelectronrepulsion/obara_saika_scheme.py:41-80). */
int gx_set_prim_screen(gx_ctx* ctx, double eps);

/* Primitive pairs kept for the current basis (summed over shell pairs). */
int gx_nprimpair(gx_ctx* ctx);

/* Nuclei (host arrays): xyz[3*natom] in bohr, charge[natom]. */
int gx_set_nuclei(gx_ctx* ctx, int natom, const double* xyz, const double* charge);

/* Number of basis functions / shell pairs of the current basis. */
int gx_nbf(gx_ctx* ctx);
int gx_npair(gx_ctx* ctx);

/* Schwarz bounds sqrt(max |(ab|ab)|) of all shell pairs, and the pair's
 * shells, copied to host arrays of length gx_npair(). */
int gx_schwarz(gx_ctx* ctx, double* Q, int* shellA, int* shellB);

/* Values of n basis-function quartets ijkl[4n] (host) -> out[n] (host). */
int gx_eri_quartets(gx_ctx* ctx, int64_t n, const int32_t* ijkl, double* out);

/* Dense (ij|kl) tensor into a device buffer of nbf^4 doubles (row-major). */
int gx_eri_dense(gx_ctx* ctx, double* out_dev);

/* Direct J and K from the density P (all nbf x nbf FP64 device buffers):
 *   J_ij = sum_kl (ij|kl) P_kl,  K_ij = sum_kl (ik|jl) P_kl.
 * Shell quartets with Q_ab Q_cd < cutoff are skipped; with GX_JK_MIXED,
 * those with Q_ab Q_cd < fp32_bound are evaluated in FP32 and accumulated
 * into the FP64 J and K. */
int gx_jk(gx_ctx* ctx, const double* P_dev, double* J_dev, double* K_dev, double cutoff,
          double fp32_bound, int flags);

/* gx_jk that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  The scaled density and the J/K accumulators live in
 * one file-scope device array of the library (5 MiB per device: at most
 * 655360 doubles for the scaled density, Kt and Jt, 3 sum_k n_k^2, and 8
 * copies of the pair-indexed J accumulator, about 8 n^2 / 2 -- one system
 * of up to 256 functions fits); larger inputs are rejected with GX_ERR_ARG.  The work items
 * must already exist for these thresholds (a gx_jk call with the same
 * cutoff, fp32_bound and GX_JK_MIXED flag builds them; otherwise
 * GX_ERR_STATE).  Only GX_JK_MIXED is accepted; every kernel runs on the
 * context stream.  All calls on one device share that array: calls that
 * overlap in time on one device (other streams, other contexts, other
 * threads) race; consecutive calls do not (each clears what it uses). */
int gx_jk_noalloc(gx_ctx* ctx, const double* P_dev, double* J_dev, double* K_dev, double cutoff,
                  double fp32_bound, int flags);

/* Unrestricted (two-density) direct build, FP64: J of P_a + P_b, K of P_a
 * and K of P_b from one evaluation of the integrals (the cooperative d-shell
 * classes evaluate theirs once per density), same thresholds and work items
 * as gx_jk; all nbf x nbf device buffers.  Replaces the Coulomb and exchange
 * sums of the reference's unrestricted Fock builder
 * (hartreefock/fock_matrix.py:25-55): F_a = H + J - K_a, F_b = H + J - K_b. */
int gx_jk_uhf(gx_ctx* ctx, const double* Pa_dev, const double* Pb_dev, double* J_dev, double* Ka_dev,
              double* Kb_dev, double cutoff, int flags);

/* Autotuner of the J/K work-item size ("tile") per class combination: for
 * each combination, the kernel is timed (best of `reps` launches, HIP
 * events on the context stream) with the items sized for each of the
 * `ncand` candidate values of "items per resident workgroup", and the
 * fastest is kept for later gx_jk calls with these thresholds (until the
 * next gx_set_basis).  With GX_JK_MIXED each combination is also timed with
 * all its quartets in FP64, and runs so in later mixed builds where that is
 * faster (never less accurate).  Writes one row of 6 doubles per
 * combination: la, lb, lc, ld, chosen value (negated when the combination
 * runs in FP64 in mixed builds), its milliseconds.  The measured times
 * also order later builds on the side streams: once every combination of a
 * build has one, the longest is launched first, each onto the stream (the
 * side streams or the caller's) with the least time queued.  Returns the
 * number of rows or an error code.  No counterpart in the reference. */
int gx_jk_autotune(gx_ctx* ctx, const double* P_dev, double cutoff, double fp32_bound, int flags,
                   const double* candidates, int ncand, int reps, double* out, int cap);

/* Nuclear gradient of the closed-shell RHF energy, dE/dR (3 natom doubles,
 * device buffer, atom-major), from the converged density P and the
 * energy-weighted density W = 2 sum_occ e_i C_i C_i^T (nbf x nbf device
 * buffers):
 *   dE/dX = sum P d(T + V)/dX - sum W dS/dX
 *         + sum over canonical quartets f (4 P_ab P_cd - P_ac P_bd - P_ad P_bc) d(ab|cd)/dX
 *         + dE_nuc/dX
 * Shell quartets with Q_ab Q_cd < cutoff are skipped.  Needs gx_set_nuclei;
 * every shell's atom index must name a nucleus.  The reference has no
 * analytic gradient: it replaces the numerical differentiation of energies
 * by its optimiser (geometryoptimization/nelder_mead.py:126-144). */
int gx_grad(gx_ctx* ctx, const double* P_dev, const double* W_dev, double* grad_dev, double cutoff);

/* gx_grad with precision and timing flags.  GX_JK_MIXED: in the per-lane
 * classes (up to (pp|pp) and the smaller d classes) the shell quartets with
 * Q_ab Q_cd < fp32_bound run their primitive loops in FP32 (FP32 pair
 * records and Boys table), the HRR, the contraction with the densities and
 * the accumulation stay FP64; the wave-per-quartet classes stay FP64.
 * GX_JK_TIMING: every class-combination kernel runs on the context stream
 * between two HIP events (read with gx_grad_launch_times).  gx_grad is
 * gx_grad_ex(..., 0.0, 0).  No counterpart in the reference (see gx_grad). */
int gx_grad_ex(gx_ctx* ctx, const double* P_dev, const double* W_dev, double* grad_dev, double cutoff,
               double fp32_bound, int flags);

/* Kernel durations of the two-electron gradient kernels of the gx_grad_ex
 * calls made with GX_JK_TIMING since the previous call of this function, as
 * gx_jk_launch_times: rows of la, lb, lc, ld, total milliseconds, launches.
 * No counterpart in the reference (measurement). */
int gx_grad_launch_times(gx_ctx* ctx, double* out, int cap);

/* Autotuner of the gradient's work-item size per class combination, as
 * gx_jk_autotune (rows of la, lb, lc, ld, chosen value, milliseconds); the
 * choice is kept for later gx_grad calls with this cutoff (until the next
 * gx_set_basis), and its times order the gradient's side-stream launches
 * as for gx_jk.  Returns the number of rows or an error code.  No
 * counterpart in the reference. */
int gx_grad_autotune(gx_ctx* ctx, const double* P_dev, const double* W_dev, double cutoff,
                     const double* candidates, int ncand, int reps, double* out, int cap);

/* Shell-quartet and flop counts that gx_jk would evaluate (host-side count). */
int gx_jk_stats(gx_ctx* ctx, double cutoff, double fp32_bound, int flags, gx_stats* out);

/* Per class combination (la lb|lc ld) of the last gx_jk-style sweep with these
 * thresholds, one row of 10 int64 in `out` (at most `cap` rows):
 *   la, lb, lc, ld, shell quartets FP64, shell quartets FP32,
 *   primitive quartets FP64, primitive quartets FP32,
 *   basis-function quartets FP64, basis-function quartets FP32.
 * Returns the number of rows (>= 0) or an error code. */
int gx_jk_class_counts(gx_ctx* ctx, double cutoff, double fp32_bound, int flags, int64_t* out, int cap);

/* Kernel durations of the gx_jk calls made with GX_JK_TIMING since the
 * previous call of this function, from HIP events recorded on the context's
 * stream around each class-combination launch.  One row of 6 doubles per
 * class combination that ran: la, lb, lc, ld, total milliseconds, launches;
 * at most `cap` rows.  Waits for the last event, then forgets the recorded
 * launches.  Returns the number of rows or an error code.  No counterpart in
 * the reference (measurement). */
int gx_jk_launch_times(gx_ctx* ctx, double* out, int cap);

/* Kernel family of class combination (la lb|lc ld) (canonical: la >= lb,
 * lc >= ld, (lc, ld) not after (la, lb)): 0 = one shell quartet per lane
 * (k_jk, k_grad_lane), 1 = one quartet per wave (k_jk_rys, k_grad_rys),
 * for the J/K (gradient = 0) or gradient (1) kernels; negative on bad
 * arguments.  No counterpart in the reference (measurement). */
int gx_kernel_kind(int la, int lb, int lc, int ld, int gradient);

/* Boys function values of the device code, for tests: out[i (L + 1) + m] =
 * F_m(T[i]), m <= L <= 10, from the per-L Taylor tables of the fine
 * (grid = 0) or coarse (grid = 1) grid by the branch-free evaluation the
 * J/K kernels use, in FP64 (fp32 = 0) or FP32 (fp32 = 1).  T and out are
 * device buffers of n and n (L + 1) doubles; runs on the context's stream.
 * Replaces no call of the reference: it exposes the device counterpart of
 * integrals/boys_function.py:4-76 to the tests. */
int gx_boys_values(gx_ctx* ctx, int grid, int fp32, int L, const double* T, int n, double* out);

/* One-electron matrices S, T, V (nbf x nbf FP64 device buffers); V needs
 * gx_set_nuclei first. */
int gx_one_electron(gx_ctx* ctx, double* S_dev, double* T_dev, double* V_dev);

#ifdef __cplusplus
}
#endif

#endif /* GX_H */
