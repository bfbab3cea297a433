#!/usr/bin/env python3
"""Benchmark: direct J/K build of a 40-atom, 220-function cluster on one MI355X.

One step = one direct J/K build (all ERIs screened by Schwarz bounds,
evaluated and digested into J and K) from an SCF density; inputs are
resident in HBM.  The mixed-precision build (quartets with Schwarz bound
< 1e-4 in FP32) is the headline when its energy error is below 1e-6 Eh; the
plain FP64 build is timed in the same run.  The energy error is the RHF
energy functional at one density with the mixed and the FP64 J/K (the
seeded random cluster does not converge under the reference's DIIS, so
SCF end points are not comparable; they are reported as they are).

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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--cpu-seconds', type=float, default=12.0,
                    help='wall seconds of the sampled CPU baseline (0 to skip)')
    ap.add_argument('--seed', type=int, default=2)
    ap.add_argument('--scf-iter', type=int, default=30, help='SCF iterations of the untimed set-up')
    ap.add_argument('--workload', default='cluster_220bf',
                    choices=['cluster_220bf', 'gradient_66bf', 'batch_256', 'cluster_255bf_d', 'uhf_220bf'],
                    help='cluster_220bf: the headline J/K line; the others print one line each for the '
                         'gradient (66 functions), the 256-geometry batch, the d-shell cluster and the '
                         'unrestricted (two-density) build')
    ap.add_argument('--kernels-only', action='store_true',
                    help='counter runs: no SCF; time the J/K kernels on the core-guess density and run '
                         'one %d MiB device copy as the HBM-counter calibration' % (COPY_BYTES >> 20))
    return ap.parse_args()


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
    serial = []
    for i in range(steps):
        starts[i].record()
        eng.jk(P, timing=True, **kw)
        ends[i].record()
    torch.cuda.synchronize()
    serial = sorted(s.elapsed_time(e) for s, e in zip(starts, ends))
    kernels = {k: v[0] / v[1] for k, v in eng.jk_launch_times().items()}
    return {'median_ms': ms[len(ms) // 2], 'mean_ms': sum(ms) / len(ms), 'min_ms': ms[0], 'wall_s': wall,
            'kernel_ms': kernels, 'serial_median_ms': serial[len(serial) // 2]}


def kernel_name(combo, mixed, fp64_combos=()):
    """Kernel of class combination `combo`; in a mixed build the combinations
    the autotuner put in FP64 (`fp64_combos`) run the FP64 kernel."""
    la, lb, lc, ld = combo
    m = mixed and tuple(combo) not in fp64_combos
    return 'k_jk<%d, %d, %d, %d, %s, 1>' % (la, lb, lc, ld, 'true' if m else 'false')


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (None when the summary does not cover it)."""
    path = os.path.join(ROOT, 'profiles', 'jk_220bf_traffic.json')
    try:
        with open(path) as f:
            return json.load(f)['hbm_bytes_per_launch'].get(kernel)
    except (OSError, ValueError, KeyError):
        return None


def class_rows(eng, mixed):
    return eng.jk_class_counts(cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=mixed)


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


def kernels_only(args):
    """The J/K kernels of both modes on the core-Hamiltonian guess density (the
    work per kernel depends on the Schwarz bounds only, not on P), plus one
    plain device copy of COPY_BYTES whose FETCH_SIZE / WRITE_SIZE calibrate the
    HBM counters."""
    import torch
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    from gaussium_amd.scf import LinearAlgebra, density_matrix_restricted
    nuc, basis, ne, info = workloads.cluster_220bf(seed=args.seed)
    eng = Engine(basis, nuc)
    S, T, V = eng.one_electron()
    _, C = LinearAlgebra(S).diagonalize(T + V)
    P = density_matrix_restricted(C, ne)
    out = {'workload': info['workload'], 'copy_bytes': COPY_BYTES}
    for name, kw in (('fp64_build', {}), ('mixed_build', dict(fp32_bound=FP32_BOUND, mixed=True))):
        t = eng.autotune(P, cutoff=CUTOFF, reps=3, **kw)  # as in the full bench (its launches come first)
        f64c = {k for k, v in t.items() if v[0] < 0}
        r = timed_builds(eng, P, args.steps, args.warmup, cutoff=CUTOFF, **kw)
        out[name] = {kernel_name(k, name == 'mixed_build', f64c): v for k, v in r['kernel_ms'].items()}
        out[name]['median_build_ms'] = r['median_ms']
    src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps(out), flush=True)


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


def _scf_density(eng, nuc, basis, ne, iters):
    """P and W = 2 sum_occ e C C^T after `iters` SCF iterations (untimed)."""
    from gaussium_amd.hartree_fock import RestrictedHF
    hf = RestrictedHF(nuc, basis, ne, engine=eng, schwarz_cutoff=CUTOFF, max_iter=iters, threshold=1e-9)
    hf.begin_scf()
    c = hf.scf_method.orbital_coefficients[:, :ne // 2]
    eps = hf.scf_method.orbital_energies[:ne // 2]
    return hf.scf_method.density, 2.0 * (c * eps) @ c.T, hf


def bench_gradient(args):
    """Analytic gradient of the 12-atom, 66-function cluster: one gx_grad call
    (one-electron + two-electron derivative kernels, FP64) at an SCF density."""
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    nuc, basis, ne, info = workloads.cluster_66bf()
    eng = Engine(basis, nuc)
    P, W, hf = _scf_density(eng, nuc, basis, ne, args.scf_iter)
    tuned = eng.grad_autotune(P, W, cutoff=CUTOFF)  # untimed set-up
    eng.autotune(P, cutoff=CUTOFF)
    ms = _median_ms(lambda: eng.grad(P, W, cutoff=CUTOFF), args.steps, args.warmup)
    ms_jk = _median_ms(lambda: eng.jk(P, cutoff=CUTOFF), args.steps, args.warmup)
    g = eng.grad(P, W, cutoff=CUTOFF)
    return {'metric': 'RHF nuclear gradient ms (66 bf, FP64)', 'value': ms, 'unit': 'ms', 'higher_is_better': False,
            'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'dtype': 'f64',
            'data': 'synthetic (seeded random cluster and basis)',
            'config': {'workload': info['workload'], 'nbf': info['nbf'], 'atoms': info['n_atoms'],
                       'schwarz_cutoff': CUTOFF},
            'jk_ms_same_density': ms_jk, 'max_abs_gradient': float(g.abs().max()),
            'translation_residual': float(g.sum(0).abs().max()),
            'autotuned_grad_items_per_wg': {'%d%d%d%d' % k: v[0] for k, v in tuned.items()}}


def bench_batch(args):
    """256 geometries of 4-40 atoms (22-220 functions) in one context: J/K and
    gradient for all of them, one launch per class combination, captured once
    in a HIP graph and replayed; work items autotuned per class first."""
    import torch
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    batch = workloads.batch(256)
    systems = [(b, n) for n, b, ne, info in batch]
    eng = Engine()
    eng.set_systems(systems)
    import numpy as np
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
    st = eng.jk_stats(cutoff=CUTOFF)
    sizes = [len(b) for b, _ in systems]
    return {'metric': 'batched J/K + gradient, 256 geometries, ms per graph replay', 'value': ms, 'unit': 'ms',
            'higher_is_better': False, 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'dtype': 'f64',
            'data': 'synthetic (seeded random clusters and basis)',
            'config': {'workload': 'batch_256', 'geometries': len(systems), 'nbf_min': min(sizes),
                       'nbf_max': max(sizes), 'nbf_total': sum(sizes), 'schwarz_cutoff': CUTOFF},
            'geometries_per_s': len(systems) / (ms * 1e-3), 'jk_ms_eager': ms_jk, 'grad_ms_eager': ms_grad,
            'bf_quartets_per_s_jk': st['bf_quartets'] / (ms_jk * 1e-3),
            'autotuned_items_per_wg': {'%d%d%d%d' % k: v[0] for k, v in tuned.items()},
            'autotuned_grad_items_per_wg': {'%d%d%d%d' % k: v[0] for k, v in tuned_grad.items()}}


def bench_d(args):
    """30-atom polarised cluster (255 functions, d shells): FP64 and mixed J/K
    and the FP64 gradient at an SCF density."""
    import torch
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    nuc, basis, ne, info = workloads.cluster_255bf()
    eng = Engine(basis, nuc)
    P, W, hf = _scf_density(eng, nuc, basis, ne, args.scf_iter)
    H = hf.core_hamiltonian
    eng.autotune(P, cutoff=CUTOFF)  # untimed set-up
    eng.autotune(P, cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True)
    eng.grad_autotune(P, W, cutoff=CUTOFF)
    ms64 = _median_ms(lambda: eng.jk(P, cutoff=CUTOFF), args.steps, args.warmup)
    msmx = _median_ms(lambda: eng.jk(P, cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True), args.steps, args.warmup)
    msg = _median_ms(lambda: eng.grad(P, W, cutoff=CUTOFF), args.steps, args.warmup)
    e = {}
    for mode, kw in (('fp64', {}), ('mixed', dict(fp32_bound=FP32_BOUND, mixed=True))):
        J, K = eng.jk(P, cutoff=CUTOFF, **kw)
        e[mode] = 0.5 * float(torch.sum(P * (2.0 * H + J - 0.5 * K)))
    st = eng.jk_stats(cutoff=CUTOFF)
    return {'metric': 'J/K build ms (255 bf with d shells), FP64 and mixed', 'value': msmx, 'unit': 'ms',
            'higher_is_better': False, 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
            'dtype': 'mixed-f32/f64', 'data': 'synthetic (seeded random cluster and basis)',
            'config': {'workload': info['workload'], 'nbf': info['nbf'], 'atoms': info['n_atoms'],
                       'electrons': ne, 'schwarz_cutoff': CUTOFF, 'fp32_bound': FP32_BOUND},
            'jk_ms_fp64': ms64, 'jk_ms_mixed': msmx, 'grad_ms_fp64': msg,
            'dE_mixed_Eh_fixed_density': abs(e['mixed'] - e['fp64']),
            'bf_quartets_per_s_fp64': st['bf_quartets'] / (ms64 * 1e-3)}


def bench_uhf(args):
    """Unrestricted J/K of the 220-function cluster (FP64): one two-density
    build (gx_jk_uhf) against two single-density builds, at alpha/beta
    densities made from the SCF orbitals with one electron moved from beta to
    alpha (a triplet occupancy)."""
    import torch
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    nuc, basis, ne, info = workloads.cluster_220bf()
    eng = Engine(basis, nuc)
    P, W, hf = _scf_density(eng, nuc, basis, ne, args.scf_iter)
    c = hf.scf_method.orbital_coefficients
    na, nb = ne // 2 + 1, ne // 2 - 1
    Pa = (c[:, :na] @ c[:, :na].T).contiguous()
    Pb = (c[:, :nb] @ c[:, :nb].T).contiguous()
    eng.autotune(P, cutoff=CUTOFF)  # untimed set-up
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
            'dtype': 'f64', 'data': 'synthetic (seeded random cluster and basis)',
            'config': {'workload': info['workload'] + '_uhf', 'nbf': info['nbf'], 'atoms': info['n_atoms'],
                       'n_alpha': na, 'n_beta': nb, 'schwarz_cutoff': CUTOFF},
            'two_single_density_builds_ms': ms_two, 'speedup_vs_two_builds': ms_two / ms_fused,
            'max_abs_diff_vs_two_builds': diff}


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
    from gaussium_amd.hartree_fock import RestrictedHF

    nuc, basis, ne, info = workloads.cluster_220bf(seed=args.seed)
    eng = Engine(basis, nuc, prim_eps=PRIM_EPS)
    e_nuc = nuclear_repulsion(nuc)

    # SCF in both modes (untimed set-up).  The reference's DIIS does not
    # converge on the seeded random clusters (the energy enters a cycle), so the
    # precision criterion is taken at a fixed density: the RHF energy
    # functional E[P] = 1/2 tr(P (H + F)) with F = H + J - K/2 from the FP64 and
    # from the mixed build, at the density the FP64 SCF ends with.
    scf = {}
    for mode, kw in (('fp64', dict(precision='fp64')), ('mixed', dict(precision='mixed', fp32_bound=FP32_BOUND))):
        t0 = time.perf_counter()
        hf = RestrictedHF(nuc, basis, ne, engine=eng, schwarz_cutoff=CUTOFF, max_iter=args.scf_iter,
                          threshold=1e-9, **kw)
        e = hf.energies()[0] + e_nuc
        scf[mode] = {'energy': e, 'iterations': hf.scf_method.iterations, 'wall_s': time.perf_counter() - t0,
                     'density': hf.scf_method.density, 'converged': bool(hf.scf_method.converged),
                     'core_hamiltonian': hf.core_hamiltonian}
    P = scf['fp64']['density']
    H = scf['fp64']['core_hamiltonian']
    # work-item sizes (and, in mixed builds, FP32 or FP64 per class
    # combination) tuned once per mode (untimed set-up), before the energies
    tuned = {'fp64': eng.autotune(P, cutoff=CUTOFF, reps=3),
             'mixed': eng.autotune(P, cutoff=CUTOFF, fp32_bound=FP32_BOUND, mixed=True, reps=3)}
    fp64_combos = {k for k, v in tuned['mixed'].items() if v[0] < 0}
    energy_at = {}
    for mode, kw in (('fp64', {}), ('mixed', dict(fp32_bound=FP32_BOUND, mixed=True))):
        J, K = eng.jk(P, cutoff=CUTOFF, **kw)
        energy_at[mode] = 0.5 * float(torch.sum(P * (2.0 * H + J - 0.5 * K))) + e_nuc
    dE = abs(energy_at['mixed'] - energy_at['fp64'])
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
    rows64, rowsmx = class_rows(eng, False), class_rows(eng, True)
    f64_a, _ = jk_model_flops(rows64)
    fm64, fm32 = jk_model_flops(rowsmx)

    ok = dE < 1e-6
    head = mixed if ok else fp64
    head_stats = stmx if ok else st64
    head_rows = rowsmx if ok else rows64
    t = head['median_ms'] * 1e-3

    # Roofline of the dominant kernel: the class-combination kernel with the
    # largest share of the build, its model flops per launch over its mean
    # event-timed duration.  Mixed kernels are priced against the blend of the
    # FP64 and FP32 peaks that their own flop split implies.
    combo = max(head['kernel_ms'], key=lambda k: head['kernel_ms'][k])
    row = next(r for r in head_rows if tuple(r[:4]) == combo)
    k64, k32 = jk_model_flops([row])
    k_s = head['kernel_ms'][combo] * 1e-3
    k_floor = k64 / (PEAK_FP64_TFLOPS * 1e12) + k32 / (PEAK_FP32_TFLOPS * 1e12)
    k_peak = (k64 + k32) / k_floor / 1e12
    k_achieved = (k64 + k32) / k_s / 1e12
    if ok:
        b_flops = fm64 + fm32
        b_peak = b_flops / (fm64 / (PEAK_FP64_TFLOPS * 1e12) + fm32 / (PEAK_FP32_TFLOPS * 1e12)) / 1e12
    else:
        b_flops, b_peak = f64_a, PEAK_FP64_TFLOPS
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
        'dE_definition': 'fixed density: |E[P] mixed J/K - E[P] FP64 J/K|, P = final FP64 SCF density',
        'energy_at_density': energy_at,
        'primitive_pairs': prim,
        'autotuned_items_per_wg': {m: {'%d%d%d%d' % k: abs(v[0]) for k, v in t.items()} for m, t in tuned.items()},
        'mixed_build_fp64_combinations': ['%d%d%d%d' % k for k, v in tuned['mixed'].items() if v[0] < 0],
        'scf': {m: {'energy': scf[m]['energy'], 'iterations': scf[m]['iterations'],
                    'converged': scf[m]['converged'], 'wall_s': scf[m]['wall_s']} for m in scf},
        'quartets': {'canonical_bf': st64['bf_quartets'], 'canonical_shell': st64['shell_quartets'],
                     'evaluated_bf_fp64_build': st64['bf_quartets_fp64'],
                     'evaluated_bf_mixed_build': [stmx['bf_quartets_fp64'], stmx['bf_quartets_fp32']],
                     'evaluated_bf_per_s': (stmx['bf_quartets_fp64'] + stmx['bf_quartets_fp32']) / t
                     if ok else st64['bf_quartets_fp64'] / t},
        'roofline': {'bound': 'mfma', 'achieved': k_achieved, 'peak': k_peak, 'unit': 'TFLOP/s',
                     'frac': k_achieved / k_peak, 'traffic': measured_traffic(kernel_name(combo, ok, fp64_combos)),
                     'compute_roof': 'FP64 arithmetic, issued on the vector ALUs: MI355X peaks at 78.6 TFLOP/s '
                                     'FP64 on its vector and its matrix units alike (FP32 parts of a mixed '
                                     'build priced at 157.3)',
                     'traffic_source': 'profiles/jk_220bf_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE per '
                                       'launch, scaled by a 256 MiB copy)',
                     'kernel': kernel_name(combo, ok, fp64_combos), 'kernel_ms': head['kernel_ms'][combo],
                     'kernel_share': head['kernel_ms'][combo] / head['serial_median_ms'],
                     'flops_per_launch': [k64, k32],
                     'build': {'flops': b_flops, 'achieved': b_flops / t / 1e12, 'peak': b_peak,
                               'frac': b_flops / t / 1e12 / b_peak},
                     'fp64_build': {'flops': f64_a, 'achieved': f64_a / (fp64['median_ms'] * 1e-3) / 1e12,
                                    'frac': f64_a / (fp64['median_ms'] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}},
        'kernel_ms': {'fp64_build': {kernel_name(k, False): v for k, v in fp64['kernel_ms'].items()},
                      'mixed_build': {kernel_name(k, True, fp64_combos): v for k, v in mixed['kernel_ms'].items()}},
    }
    line['class_counts'] = {'columns': ['la', 'lb', 'lc', 'ld', 'shq64', 'shq32', 'primq64', 'primq32', 'bfq64',
                                        'bfq32'], 'fp64_build': rows64, 'mixed_build': rowsmx}
    traffic_file = os.path.join(ROOT, 'profiles', 'jk_220bf_traffic.json')
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        per_kernel = tr.get('hbm_bytes_per_launch', {})
        line['roofline']['traffic'] = per_kernel.get(line['roofline']['kernel'])
    if args.cpu_seconds > 0:
        cb = cpu_baseline(basis, args.cpu_seconds)
        line['cpu_baseline'] = cb
        line['gpu_vs_cpu'] = line['value'] / cb['value']
    else:
        line['cpu_baseline'] = None
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
