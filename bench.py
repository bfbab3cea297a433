#!/usr/bin/env python3
"""Benchmark: direct J/K build of a 40-atom, 220-function cluster on one MI355X.

One step = one direct J/K build (all ERIs screened by Schwarz bounds,
evaluated and digested into J and K) at the converged SCF density; inputs
are resident in HBM.  The mixed-precision build (quartets with Schwarz bound
< 1e-4 in FP32) is the headline when its converged RHF energy is within
1e-6 Eh of the FP64 one; the plain FP64 build is timed in the same run.  The
SCF runs to convergence in both precisions in the untimed set-up with the
opt-in second-order SCF (the reference's DIIS loop cycles on this cluster);
its energies, iteration counts, wall times and the smallest overlap
eigenvalue are reported.

Prints one JSON line.  ``--gpus`` above 1 is not supported (one GPU per run).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector (= FP64 matrix) peak
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (packed)
PEAK_HBM_GBS = 8000.0
CUTOFF = 1e-11
FP32_BOUND = 1e-4
PRIM_EPS = 1e-18  # primitive-pair screening threshold (the library default)
COPY_BYTES = 256 << 20
SCF_MAX_ITER = 5000
COMPUTE_ROOF = ('FP64 arithmetic issued on the vector ALUs (the recurrences are not GEMM-shaped); MI355X '
                'peaks at 78.6 TFLOP/s FP64 on its vector and its matrix units alike; FP32 parts of a mixed '
                'kernel are priced at 157.3')


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--cpu-seconds', type=float, default=12.0,
                    help='wall seconds of the sampled CPU baseline (0 to skip)')
    ap.add_argument('--seed', type=int, default=2)
    ap.add_argument('--workload', default='cluster_220bf',
                    choices=['cluster_220bf', 'gradient_66bf', 'batch_256', 'cluster_255bf_d', 'uhf_220bf'],
                    help='cluster_220bf: the headline J/K line; the others print one line each for the '
                         'gradient (66 functions), the 256-geometry batch, the d-shell cluster and the '
                         'unrestricted (two-density) build')
    ap.add_argument('--kernels-only', action='store_true',
                    help='counter runs: no SCF; time the J/K kernels on the core-guess density and run '
                         'one %d MiB device copy as the HBM-counter calibration' % (COPY_BYTES >> 20))
    return ap.parse_args()


# --------------------------------------------------------------------------- helpers
def kernel_name(combo, gradient=False, mixed=False):
    """rocprofv3 name of the kernel that runs class combination `combo`."""
    from gaussium_amd import _lib
    la, lb, lc, ld = combo
    coop = _lib.load().gx_kernel_kind(la, lb, lc, ld, 1 if gradient else 0) == 1
    if gradient:
        if coop:
            return 'k_grad_rys<%d, %d, %d, %d>' % combo
        return 'k_grad_lane<%d, %d, %d, %d, %s>' % (la, lb, lc, ld, 'true' if mixed else 'false')
    if coop:
        return 'k_jk_rys<%d, %d, %d, %d>' % combo
    return 'k_jk<%d, %d, %d, %d, %s, 1>' % (la, lb, lc, ld, 'true' if mixed else 'false')


def grad_rows(eng, mixed):
    """Class counts of the gradient's quartets: the J/K counts with these
    thresholds, with the FP32 part of the wave-per-quartet gradient classes
    (always FP64) moved to FP64."""
    from gaussium_amd import _lib
    rows = []
    for r in eng.jk_class_counts(cutoff=CUTOFF, fp32_bound=FP32_BOUND if mixed else 0.0, mixed=mixed):
        r = list(r)
        if _lib.load().gx_kernel_kind(r[0], r[1], r[2], r[3], 1) == 1:
            r[4], r[5], r[6], r[7], r[8], r[9] = r[4] + r[5], 0, r[6] + r[7], 0, r[8] + r[9], 0
        rows.append(tuple(r))
    return rows


def roofline(kernel_ms, rows, model, name, share_of=None, traffic=None, traffic_source=None):
    """Roofline of the dominant kernel: the class combination with the largest
    event-timed share, its model flops per launch (`model(rows) -> (f64, f32)`)
    over its mean duration, against the peak its FP64/FP32 split implies."""
    combo = max(kernel_ms, key=lambda k: kernel_ms[k])
    row = next(r for r in rows if tuple(r[:4]) == combo)
    k64, k32 = model([row])
    k_s = kernel_ms[combo] * 1e-3
    floor = k64 / (PEAK_FP64_TFLOPS * 1e12) + k32 / (PEAK_FP32_TFLOPS * 1e12)
    peak = (k64 + k32) / floor / 1e12
    achieved = (k64 + k32) / k_s / 1e12
    out = {'bound': 'valu_fp64', 'achieved': achieved, 'peak': peak, 'unit': 'TFLOP/s', 'frac': achieved / peak,
           'traffic': traffic, 'compute_roof': COMPUTE_ROOF, 'kernel': name(combo),
           'kernel_ms': kernel_ms[combo], 'flops_per_launch': [k64, k32],
           'class_combination': '(%d%d|%d%d)' % combo}
    if share_of:
        out['kernel_share'] = kernel_ms[combo] / share_of
    if traffic_source:
        out['traffic_source'] = traffic_source
    return out, combo


def per_kernel_flops(kernel_ms, rows, model, name):
    """{kernel: {ms, model flops (FP64, FP32), frac of its peak}} for every timed kernel."""
    out = {}
    for combo, ms in kernel_ms.items():
        row = next(r for r in rows if tuple(r[:4]) == combo)
        f64, f32 = model([row])
        floor = f64 / (PEAK_FP64_TFLOPS * 1e12) + f32 / (PEAK_FP32_TFLOPS * 1e12)
        out[name(combo)] = {'ms': ms, 'flops': [f64, f32], 'frac': floor / (ms * 1e-3) if ms > 0 else None}
    return out


def timed_builds(eng, P, steps, warmup, **kw):
    """Median build time (torch events around each gx_jk call; the class
    combinations run on the library's side streams), then `steps` more builds
    with per-launch HIP events (gx_jk then launches every kernel on its own
    stream, one after the other) for the per-kernel durations."""
    import torch
    for _ in range(warmup):
        eng.jk(P, **kw)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record()
        eng.jk(P, **kw)
        ends[i].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = sorted(s.elapsed_time(e) for s, e in zip(starts, ends))
    eng.jk_launch_times()  # forget earlier timed launches
    for i in range(steps):
        starts[i].record()
        eng.jk(P, timing=True, **kw)
        ends[i].record()
    torch.cuda.synchronize()
    serial = sorted(s.elapsed_time(e) for s, e in zip(starts, ends))
    kernels = {k: v[0] / v[1] for k, v in eng.jk_launch_times().items()}
    return {'median_ms': ms[len(ms) // 2], 'mean_ms': sum(ms) / len(ms), 'min_ms': ms[0], 'wall_s': wall,
            'kernel_ms': kernels, 'serial_median_ms': serial[len(serial) // 2]}


def timed_grads(eng, P, W, steps, warmup, **kw):
    """As timed_builds for gx_grad_ex: median call time, then per-kernel
    event times of the two-electron gradient kernels."""
    import torch
    ms = _median_ms(lambda: eng.grad(P, W, **kw), steps, warmup)
    eng.grad_launch_times()
    for _ in range(steps):
        eng.grad(P, W, timing=True, **kw)
    torch.cuda.synchronize()
    return ms, {k: v[0] / v[1] for k, v in eng.grad_launch_times().items()}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (None when the summary does not cover it)."""
    path = os.path.join(ROOT, 'profiles', 'jk_220bf_traffic.json')
    try:
        with open(path) as f:
            return json.load(f)['hbm_bytes_per_launch'].get(kernel)
    except (OSError, ValueError, KeyError):
        return None


def cpu_baseline(basis, seconds):
    from oracle.ref_faithful import baseline
    b = baseline(basis, seconds=seconds)
    return {'value': b['quartets_per_s'], 'unit': 'bf-quartets/s', 'cores': b['procs'], 'kind': 'port',
            'sample': ('pure-Python recursive Obara-Saika (reference algorithm, no memo) on %d random '
                       'canonical quartets of the 220-function cluster over a %d-process Pool for %.0f s; '
                       'ERI build %.3g s + dense fill %.3g s + one Fock loop %.3g s, all extrapolated '
                       'to the %d canonical quartets'
                       % (b['sampled_quartets'], b['procs'], b['sample_wall_s'], b['eri_s'], b['fill_s'],
                          b['fock_s'], b['canonical_quartets'])),
            'extrapolated_total_s': b['total_s'], 'eri_quartets_per_s': b['eri_quartets_per_s']}


def _median_ms(fn, steps, warmup):
    """Median of `steps` calls of fn() timed with torch events on the current stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    en = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    for i in range(steps):
        st[i].record()
        fn()
        en[i].record()
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in zip(st, en))
    return ms[len(ms) // 2]


def core_guess_density(eng, ne):
    """Density of the core-Hamiltonian guess (the work of every kernel depends
    on the Schwarz bounds only, so tuning needs no SCF)."""
    from gaussium_amd.scf import LinearAlgebra, density_matrix_restricted
    S, T, V = eng.one_electron()
    _, C = LinearAlgebra(S).diagonalize(T + V)
    return density_matrix_restricted(C, ne)


def converged_scf(eng, nuc, basis, ne, precision='fp64', initial_density=None):
    """RHF to convergence with the second-order SCF; returns the report and
    the converged density and energy-weighted density."""
    import torch
    from gaussium_amd.basis import nuclear_repulsion
    from gaussium_amd.hartree_fock import RestrictedHF
    kw = dict(precision='mixed', fp32_bound=FP32_BOUND) if precision == 'mixed' else {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # basis None: the engine keeps its basis, pairs and tuned work items
    hf = RestrictedHF(nuc, None, ne, engine=eng, schwarz_cutoff=CUTOFF, stabilizer='second_order',
                      max_iter=SCF_MAX_ITER, initial_density=initial_density, **kw)
    e = hf.energies()[0] + nuclear_repulsion(nuc)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    so = hf.scf_method.stabilizer
    no = ne // 2
    c = hf.scf_method.orbital_coefficients[:, :no]
    W = 2.0 * (c * hf.scf_method.orbital_energies[:no]) @ c.T
    report = {'energy': e, 'converged': bool(hf.scf_method.converged), 'wall_s': wall,
              'energy_evaluations': hf.scf_method.iterations, 'hessian_products': so.hessian_products,
              'stage_iterations': dict(so.stage_iterations), 'fock_builds': hf.fock_builder.builds,
              'min_overlap_eigenvalue': hf.smallest_overlap_eigenvalue,
              'method': 'second-order SCF (optimal damping, L-BFGS, trust-region Newton), |dE| < 1e-12, '
                        'max |orbital gradient| < 1e-7'}
    return report, hf.scf_method.density, W, hf


# --------------------------------------------------------------------------- modes
def kernels_only(args):
    """The J/K kernels of both modes (220-function cluster, or the 255-function
    one with --workload cluster_255bf_d) on the core-Hamiltonian guess density (the
    work per kernel depends on the Schwarz bounds only, not on P), plus one
    plain device copy of COPY_BYTES whose FETCH_SIZE / WRITE_SIZE calibrate the
    HBM counters."""
    import torch
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    if args.workload == 'cluster_255bf_d':
        nuc, basis, ne, info = workloads.cluster_255bf()
    else:
        nuc, basis, ne, info = workloads.cluster_220bf(seed=args.seed)
    eng = Engine(basis, nuc)
    P = core_guess_density(eng, ne)
    out = {'workload': info['workload'], 'copy_bytes': COPY_BYTES}
    for name, kw in (('fp64_build', {}), ('mixed_build', dict(fp32_bound=FP32_BOUND, mixed=True))):
        t = eng.autotune(P, cutoff=CUTOFF, reps=3, **kw)  # as in the full bench (its launches come first)
        f64c = {k for k, v in t.items() if v[0] < 0}
        r = timed_builds(eng, P, args.steps, args.warmup, cutoff=CUTOFF, **kw)
        out[name] = {kernel_name(k, mixed=name == 'mixed_build' and k not in f64c): v
                     for k, v in r['kernel_ms'].items()}
        out[name]['median_build_ms'] = r['median_ms']
    src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps(out), flush=True)


def bench_gradient(args):
    """Analytic gradient of the 12-atom, 66-function cluster at the converged
    SCF density: gx_grad_ex in FP64 and mixed precision (one-electron +
    two-electron derivative kernels); the value is the mixed gradient's time
    when it is within 1e-6 Eh/bohr of the FP64 one."""
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    from gaussium_amd.flops import grad_model_flops
    nuc, basis, ne, info = workloads.cluster_66bf()
    eng = Engine(basis, nuc)
    P0 = core_guess_density(eng, ne)
    eng.autotune(P0, cutoff=CUTOFF)  # untimed set-up (the J/K items of the SCF)
    scf, P, W, hf = converged_scf(eng, nuc, basis, ne)
    tuned = eng.grad_autotune(P, W, cutoff=CUTOFF)
    ms64, k64 = timed_grads(eng, P, W, args.steps, args.warmup, cutoff=CUTOFF)
    msmx, kmx = timed_grads(eng, P, W, args.steps, args.warmup, cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True)
    ms_jk = _median_ms(lambda: eng.jk(P, cutoff=CUTOFF), args.steps, args.warmup)
    g64 = eng.grad(P, W, cutoff=CUTOFF)
    gmx = eng.grad(P, W, cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True)
    err = float((gmx - g64).abs().max())
    ok = err < 1e-6
    rows64, rowsmx = grad_rows(eng, False), grad_rows(eng, True)
    rows, kms = (rowsmx, kmx) if ok else (rows64, k64)
    name = lambda c: kernel_name(c, gradient=True, mixed=ok)
    roof, _ = roofline(kms, rows, grad_model_flops, name, share_of=sum(kms.values()))
    f64, f32 = grad_model_flops(rows)
    return {'metric': 'RHF nuclear gradient ms (66 bf), mixed vs FP64 at |dg| < 1e-6 Eh/bohr',
            'value': msmx if ok else ms64, 'unit': 'ms', 'higher_is_better': False, 'n_gpus': 1,
            'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': msmx if ok else ms64,
            'dtype': 'mixed-f32/f64' if ok else 'f64', 'data': 'synthetic (seeded random cluster and basis)',
            'config': {'workload': info['workload'], 'nbf': info['nbf'], 'atoms': info['n_atoms'],
                       'electrons': ne, 'schwarz_cutoff': CUTOFF, 'fp32_bound': FP32_BOUND},
            'grad_ms_fp64': ms64, 'grad_ms_mixed': msmx, 'mixed_speedup': ms64 / msmx,
            'max_abs_dg_mixed_vs_fp64': err, 'jk_ms_same_density': ms_jk,
            'max_abs_gradient': float(g64.abs().max()), 'translation_residual': float(g64.sum(0).abs().max()),
            'min_overlap_eigenvalue': scf['min_overlap_eigenvalue'], 'scf': scf,
            'roofline': roof,
            'two_electron_model_flops': {'fp64': f64, 'fp32': f32,
                                         'achieved_TFLOPs_all_kernels': (f64 + f32) / (sum(kms.values()) * 1e-3)
                                         / 1e12},
            'kernel_ms': {'fp64': {kernel_name(k, True): v for k, v in k64.items()},
                          'mixed': {kernel_name(k, True, True): v for k, v in kmx.items()}},
            'per_kernel': per_kernel_flops(kms, rows, grad_model_flops, name),
            'autotuned_grad_items_per_wg': {'%d%d%d%d' % k: v[0] for k, v in tuned.items()}}


def bench_batch(args):
    """256 geometries of 4-40 atoms (22-220 functions) in one context: J/K and
    gradient for all of them, one launch per class combination, captured once
    in a HIP graph and replayed; work items autotuned per class first."""
    import numpy as np
    import torch
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    from gaussium_amd.flops import grad_model_flops, jk_model_flops
    batch = workloads.batch(256)
    systems = [(b, n) for n, b, ne, info in batch]
    eng = Engine()
    eng.set_systems(systems)
    rng = np.random.default_rng(0)
    mats = []
    for b, _ in systems:
        a = rng.standard_normal((len(b), len(b)))
        mats.append((a + a.T) / len(b))
    P = eng.pack(mats)
    print('batch: %d systems set up' % len(systems), file=sys.stderr, flush=True)
    tuned = eng.autotune(P, cutoff=CUTOFF, reps=2)
    tuned_grad = eng.grad_autotune(P, P, cutoff=CUTOFF, reps=1)
    print('batch: autotuned', file=sys.stderr, flush=True)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        eng.jk(P, cutoff=CUTOFF)
        eng.grad(P, P, cutoff=CUTOFF)
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        eng.jk(P, cutoff=CUTOFF)
        eng.grad(P, P, cutoff=CUTOFF)
    print('batch: graph captured', file=sys.stderr, flush=True)
    ms = _median_ms(graph.replay, args.steps, args.warmup)
    ms_jk = _median_ms(lambda: eng.jk(P, cutoff=CUTOFF), args.steps, args.warmup)
    ms_grad = _median_ms(lambda: eng.grad(P, P, cutoff=CUTOFF), args.steps, args.warmup)
    # per-kernel event times (one stream) and model flops of J/K and gradient kernels
    eng.jk_launch_times()
    eng.grad_launch_times()
    reps = max(2, args.steps // 5)
    for _ in range(reps):
        eng.jk(P, cutoff=CUTOFF, timing=True)
        eng.grad(P, P, cutoff=CUTOFF, timing=True)
    torch.cuda.synchronize()
    kjk = {k: v[0] / v[1] for k, v in eng.jk_launch_times().items()}
    kgr = {k: v[0] / v[1] for k, v in eng.grad_launch_times().items()}
    rows_jk = eng.jk_class_counts(cutoff=CUTOFF)
    rows_gr = grad_rows(eng, False)
    jk_name = lambda c: kernel_name(c)
    gr_name = lambda c: kernel_name(c, gradient=True)
    roof_jk, _ = roofline(kjk, rows_jk, jk_model_flops, jk_name)
    roof_gr, _ = roofline(kgr, rows_gr, grad_model_flops, gr_name)
    total = sum(kjk.values()) + sum(kgr.values())
    roof = roof_gr if max(kgr.values()) >= max(kjk.values()) else roof_jk
    roof['kernel_share'] = roof['kernel_ms'] / total
    st = eng.jk_stats(cutoff=CUTOFF)
    sizes = [len(b) for b, _ in systems]
    fj = sum(jk_model_flops(rows_jk))
    fg = sum(grad_model_flops(rows_gr))
    return {'metric': 'batched J/K + gradient, 256 geometries, ms per graph replay', 'value': ms, 'unit': 'ms',
            'higher_is_better': False, 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
            'ms_per_step': ms, 'dtype': 'f64', 'data': 'synthetic (seeded random clusters and basis)',
            'config': {'workload': 'batch_256', 'geometries': len(systems), 'nbf_min': min(sizes),
                       'nbf_max': max(sizes), 'nbf_total': sum(sizes), 'schwarz_cutoff': CUTOFF},
            'geometries_per_s': len(systems) / (ms * 1e-3), 'jk_ms_eager': ms_jk, 'grad_ms_eager': ms_grad,
            'bf_quartets_per_s_jk': st['bf_quartets'] / (ms_jk * 1e-3),
            'roofline': roof, 'roofline_jk_dominant': roof_jk,
            'model_flops': {'jk': fj, 'grad_two_electron': fg,
                            'replay_TFLOPs': (fj + fg) / (ms * 1e-3) / 1e12},
            'per_kernel': {'jk': per_kernel_flops(kjk, rows_jk, jk_model_flops, jk_name),
                           'grad': per_kernel_flops(kgr, rows_gr, grad_model_flops, gr_name)},
            'launch_structure': 'one persistent launch per class combination (all systems), captured as one '
                                'graph chain',
            'autotuned_items_per_wg': {'%d%d%d%d' % k: v[0] for k, v in tuned.items()},
            'autotuned_grad_items_per_wg': {'%d%d%d%d' % k: v[0] for k, v in tuned_grad.items()}}


def bench_d(args):
    """30-atom polarised cluster (255 functions, d shells): converged SCF in
    FP64 and mixed precision, then FP64 and mixed J/K and the FP64 gradient at
    the converged density, with per-kernel rooflines."""
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    from gaussium_amd.flops import grad_model_flops, jk_model_flops
    nuc, basis, ne, info = workloads.cluster_255bf()
    eng = Engine(basis, nuc)
    P0 = core_guess_density(eng, ne)
    eng.autotune(P0, cutoff=CUTOFF)  # untimed set-up
    tuned_mx = eng.autotune(P0, cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True)
    f64c = {k for k, v in tuned_mx.items() if v[0] < 0}
    scf64, P, W, hf = converged_scf(eng, nuc, basis, ne)
    scfmx, _, _, _ = converged_scf(eng, nuc, basis, ne, 'mixed')
    eng.grad_autotune(P, W, cutoff=CUTOFF)
    fp64 = timed_builds(eng, P, args.steps, args.warmup, cutoff=CUTOFF)
    mixed = timed_builds(eng, P, args.steps, args.warmup, cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True)
    msg, kgr = timed_grads(eng, P, W, max(3, args.steps // 4), 1, cutoff=CUTOFF)
    rows64 = eng.jk_class_counts(cutoff=CUTOFF)
    ok = abs(scfmx['energy'] - scf64['energy']) < 1e-6
    head = mixed if ok else fp64
    rows = eng.jk_class_counts(cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True) if ok else rows64
    name = lambda c: kernel_name(c, mixed=ok and c not in f64c)
    roof, _ = roofline(head['kernel_ms'], rows, jk_model_flops, name, share_of=head['serial_median_ms'])
    gr_name = lambda c: kernel_name(c, gradient=True)
    roof_gr, _ = roofline(kgr, grad_rows(eng, False), grad_model_flops, gr_name, share_of=sum(kgr.values()))
    dd = next((r for r in rows64 if tuple(r[:4]) == (2, 2, 2, 2)), None)
    st = eng.jk_stats(cutoff=CUTOFF)
    out = {'metric': 'J/K build ms (255 bf with d shells), FP64 and mixed', 'value': head['median_ms'],
           'unit': 'ms', 'higher_is_better': False, 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
           'ms_per_step': head['median_ms'], 'dtype': 'mixed-f32/f64' if ok else 'f64',
           'data': 'synthetic (seeded random cluster and basis)',
           'config': {'workload': info['workload'], 'nbf': info['nbf'], 'atoms': info['n_atoms'],
                      'electrons': ne, 'schwarz_cutoff': CUTOFF, 'fp32_bound': FP32_BOUND},
           'jk_ms_fp64': fp64['median_ms'], 'jk_ms_mixed': mixed['median_ms'], 'grad_ms_fp64': msg,
           'dE_mixed_Eh': abs(scfmx['energy'] - scf64['energy']),
           'dE_definition': 'converged RHF energies of the mixed and the FP64 SCF',
           'min_overlap_eigenvalue': scf64['min_overlap_eigenvalue'],
           'scf': {'fp64': scf64, 'mixed': scfmx},
           'bf_quartets_per_s_fp64': st['bf_quartets'] / (fp64['median_ms'] * 1e-3),
           'roofline': roof, 'roofline_gradient': roof_gr,
           'per_kernel': {'jk': per_kernel_flops(head['kernel_ms'], rows, jk_model_flops, name),
                          'grad': per_kernel_flops(kgr, grad_rows(eng, False), grad_model_flops, gr_name)}}
    if dd is not None and (2, 2, 2, 2) in fp64['kernel_ms']:
        f = sum(jk_model_flops([dd]))
        out['dddd'] = {'kernel': kernel_name((2, 2, 2, 2)), 'shell_quartets': dd[4], 'model_flops': f,
                       'kernel_ms': fp64['kernel_ms'][(2, 2, 2, 2)],
                       'frac': f / (fp64['kernel_ms'][(2, 2, 2, 2)] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}
    return out


def bench_uhf(args):
    """Unrestricted J/K of the 220-function cluster (FP64): one two-density
    build (gx_jk_uhf) against two single-density builds, at alpha/beta
    densities made from the converged RHF orbitals with one electron moved
    from beta to alpha (a triplet occupancy)."""
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    nuc, basis, ne, info = workloads.cluster_220bf()
    eng = Engine(basis, nuc)
    P0 = core_guess_density(eng, ne)
    eng.autotune(P0, cutoff=CUTOFF)  # untimed set-up
    scf, P, W, hf = converged_scf(eng, nuc, basis, ne)
    c = hf.scf_method.orbital_coefficients
    na, nb = ne // 2 + 1, ne // 2 - 1
    Pa = (c[:, :na] @ c[:, :na].T).contiguous()
    Pb = (c[:, :nb] @ c[:, :nb].T).contiguous()
    ms_fused = _median_ms(lambda: eng.jk_uhf(Pa, Pb, cutoff=CUTOFF), args.steps, args.warmup)

    def two_builds():
        eng.jk(Pa, cutoff=CUTOFF)
        eng.jk(Pb, cutoff=CUTOFF)
    ms_two = _median_ms(two_builds, args.steps, args.warmup)
    J, Ka, Kb = eng.jk_uhf(Pa, Pb, cutoff=CUTOFF)
    Ja, Ka1 = eng.jk(Pa, cutoff=CUTOFF)
    Jb, Kb1 = eng.jk(Pb, cutoff=CUTOFF)
    diff = max(float((J - Ja - Jb).abs().max()), float((Ka - Ka1).abs().max()), float((Kb - Kb1).abs().max()))
    return {'metric': 'unrestricted J/K build ms (220 bf, FP64), one two-density build', 'value': ms_fused,
            'unit': 'ms', 'higher_is_better': False, 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
            'ms_per_step': ms_fused, 'dtype': 'f64', 'data': 'synthetic (seeded random cluster and basis)',
            'config': {'workload': info['workload'] + '_uhf', 'nbf': info['nbf'], 'atoms': info['n_atoms'],
                       'n_alpha': na, 'n_beta': nb, 'schwarz_cutoff': CUTOFF},
            'two_single_density_builds_ms': ms_two, 'speedup_vs_two_builds': ms_two / ms_fused,
            'max_abs_diff_vs_two_builds': diff, 'min_overlap_eigenvalue': scf['min_overlap_eigenvalue']}


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py measures one GPU; --gpus %d is not supported (run with --gpus 1)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.kernels_only:
        return kernels_only(args)
    extra = {'gradient_66bf': bench_gradient, 'batch_256': bench_batch, 'cluster_255bf_d': bench_d,
             'uhf_220bf': bench_uhf}
    if args.workload in extra:
        print(json.dumps(extra[args.workload](args)), flush=True)
        return
    import torch
    from gaussium_amd import workloads
    from gaussium_amd.basis import nuclear_repulsion
    from gaussium_amd.engine import Engine
    from gaussium_amd.flops import jk_model_flops

    nuc, basis, ne, info = workloads.cluster_220bf(seed=args.seed)
    eng = Engine(basis, nuc, prim_eps=PRIM_EPS)
    e_nuc = nuclear_repulsion(nuc)

    # work-item sizes (and, in mixed builds, FP32 or FP64 per class
    # combination) tuned once per mode on the core-guess density (untimed
    # set-up; the work of every kernel depends on the Schwarz bounds only)
    P0 = core_guess_density(eng, ne)
    tuned = {'fp64': eng.autotune(P0, cutoff=CUTOFF, reps=3),
             'mixed': eng.autotune(P0, cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True, reps=3)}
    fp64_combos = {k for k, v in tuned['mixed'].items() if v[0] < 0}
    print('bench: autotuned', file=sys.stderr, flush=True)
    # converged SCF in both precisions (untimed set-up; wall times reported)
    scf64, P, W, hf = converged_scf(eng, nuc, basis, ne)
    print('bench: FP64 SCF %.10f (%s)' % (scf64['energy'], scf64['converged']), file=sys.stderr, flush=True)
    scfmx, Pmx, _, _ = converged_scf(eng, nuc, basis, ne, 'mixed')
    print('bench: mixed SCF %.10f (%s)' % (scfmx['energy'], scfmx['converged']), file=sys.stderr, flush=True)
    dE = abs(scfmx['energy'] - scf64['energy'])
    H = hf.core_hamiltonian
    energy_at = {}
    for mode, kw in (('fp64', {}), ('mixed', dict(fp32_bound=FP32_BOUND, mixed=True))):
        J, K = eng.jk(P, cutoff=CUTOFF, **kw)
        energy_at[mode] = 0.5 * float(torch.sum(P * (2.0 * H + J - 0.5 * K))) + e_nuc
    # the same FP64 build over every primitive pair (no primitive-pair
    # screening): energy change and time, reported beside the headline
    full = Engine(basis, nuc, prim_eps=0.0)
    J, K = full.jk(P, cutoff=CUTOFF)
    e_full = 0.5 * float(torch.sum(P * (2.0 * H + J - 0.5 * K))) + e_nuc
    fp64_full = timed_builds(full, P, args.steps, args.warmup, cutoff=CUTOFF)
    prim = {'kept': eng.n_primitive_pairs, 'all': full.n_primitive_pairs, 'threshold': PRIM_EPS,
            'dE_fp64_vs_unscreened_Eh': abs(energy_at['fp64'] - e_full),
            'jk_ms_fp64_unscreened': fp64_full['median_ms']}
    full.close()
    fp64 = timed_builds(eng, P, args.steps, args.warmup, cutoff=CUTOFF)
    mixed = timed_builds(eng, P, args.steps, args.warmup, cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True)
    st64 = eng.jk_stats(cutoff=CUTOFF)
    stmx = eng.jk_stats(cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True)
    rows64 = eng.jk_class_counts(cutoff=CUTOFF)
    rowsmx = eng.jk_class_counts(cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True)
    f64_a, _ = jk_model_flops(rows64)
    fm64, fm32 = jk_model_flops(rowsmx)

    ok = dE < 1e-6 and scf64['converged'] and scfmx['converged']
    head = mixed if ok else fp64
    head_stats = stmx if ok else st64
    head_rows = rowsmx if ok else rows64
    t = head['median_ms'] * 1e-3
    name = lambda c: kernel_name(c, mixed=ok and c not in fp64_combos)
    roof, combo = roofline(head['kernel_ms'], head_rows, jk_model_flops, name, share_of=head['serial_median_ms'],
                           traffic_source='profiles/jk_220bf_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE per '
                                          'launch, scaled by a 256 MiB copy)')
    roof['traffic'] = measured_traffic(roof['kernel'])
    if ok:
        b_flops = fm64 + fm32
        b_peak = b_flops / (fm64 / (PEAK_FP64_TFLOPS * 1e12) + fm32 / (PEAK_FP32_TFLOPS * 1e12)) / 1e12
    else:
        b_flops, b_peak = f64_a, PEAK_FP64_TFLOPS
    roof['build'] = {'flops': b_flops, 'achieved': b_flops / t / 1e12, 'peak': b_peak,
                     'frac': b_flops / t / 1e12 / b_peak}
    roof['fp64_build'] = {'flops': f64_a, 'achieved': f64_a / (fp64['median_ms'] * 1e-3) / 1e12,
                          'frac': f64_a / (fp64['median_ms'] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}
    line = {
        'metric': 'ERI quartets/s + J/K build ms (220 bf), mixed vs FP64 at dE<1e-6 Eh',
        'value': head_stats['bf_quartets'] / t,
        'unit': 'bf-quartets/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': head['median_ms'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'mixed-f32/f64' if ok else 'f64',
        'data': 'synthetic (seeded random cluster and basis)',
        'config': {'workload': info['workload'], 'nbf': info['nbf'], 'atoms': info['n_atoms'],
                   'electrons': ne, 'schwarz_cutoff': CUTOFF, 'fp32_bound': FP32_BOUND, 'seed': args.seed},
        'jk_ms_fp64': fp64['median_ms'],
        'jk_ms_mixed': mixed['median_ms'],
        'mixed_speedup': fp64['median_ms'] / mixed['median_ms'],
        'jk_ms_one_stream': {'fp64': fp64['serial_median_ms'], 'mixed': mixed['serial_median_ms'],
                             'note': 'every kernel on the caller stream, with per-launch events (the kernel_ms '
                                     'runs); the headline builds overlap class combinations on side streams'},
        'dE_mixed_Eh': dE,
        'dE_definition': 'converged RHF total energies of the mixed and the FP64 SCF (both with the bench '
                         'screening), |E_mixed - E_fp64|',
        'dE_mixed_Eh_fixed_density': abs(energy_at['mixed'] - energy_at['fp64']),
        'energy_at_density': energy_at,
        'min_overlap_eigenvalue': scf64['min_overlap_eigenvalue'],
        'scf': {'fp64': scf64, 'mixed': scfmx},
        'primitive_pairs': prim,
        'autotuned_items_per_wg': {m: {'%d%d%d%d' % k: abs(v[0]) for k, v in t_.items()} for m, t_ in tuned.items()},
        'mixed_build_fp64_combinations': ['%d%d%d%d' % k for k in sorted(fp64_combos)],
        'quartets': {'canonical_bf': st64['bf_quartets'], 'canonical_shell': st64['shell_quartets'],
                     'evaluated_bf_fp64_build': st64['bf_quartets_fp64'],
                     'evaluated_bf_mixed_build': [stmx['bf_quartets_fp64'], stmx['bf_quartets_fp32']],
                     'evaluated_bf_per_s': (stmx['bf_quartets_fp64'] + stmx['bf_quartets_fp32']) / t
                     if ok else st64['bf_quartets_fp64'] / t},
        'roofline': roof,
        'kernel_ms': {'fp64_build': {kernel_name(k): v for k, v in fp64['kernel_ms'].items()},
                      'mixed_build': {name(k): v for k, v in mixed['kernel_ms'].items()}},
        'per_kernel': per_kernel_flops(head['kernel_ms'], head_rows, jk_model_flops, name),
    }
    line['class_counts'] = {'columns': ['la', 'lb', 'lc', 'ld', 'shq64', 'shq32', 'primq64', 'primq32', 'bfq64',
                                        'bfq32'], 'fp64_build': rows64, 'mixed_build': rowsmx}
    if args.cpu_seconds > 0:
        cb = cpu_baseline(basis, args.cpu_seconds)
        line['cpu_baseline'] = cb
        line['gpu_vs_cpu'] = line['value'] / cb['value']
    else:
        line['cpu_baseline'] = None
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
