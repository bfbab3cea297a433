"""gx_jk_noalloc (the J/K build that allocates nothing, its scratch in a
static device array) against gx_jk, through the C ABI: two calls in a row,
the second on a smaller system, FP64 and mixed; and its error returns."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _call(eng, P, cutoff, fp32_bound, flags):
    n = P.shape[0]
    J = torch.full((n, n), np.nan, dtype=torch.float64, device=eng.device)
    K = torch.full((n, n), np.nan, dtype=torch.float64, device=eng.device)
    eng._bind_stream()
    rc = eng.lib.gx_jk_noalloc(eng.ctx, ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(J.data_ptr()),
                               ctypes.c_void_p(K.data_ptr()), float(cutoff), float(fp32_bound), int(flags))
    return rc, J, K


def _density(n, ne, seed):
    c = np.random.default_rng(seed).standard_normal((n, ne // 2))
    return torch.from_numpy(c @ c.T / n).cuda()


@pytest.mark.parametrize('mixed', [False, True])
def test_noalloc_matches_gx_jk_two_calls(mixed):
    from gaussium_amd import _lib, workloads
    from gaussium_amd.engine import Engine
    flags = _lib.GX_JK_MIXED if mixed else 0
    kw = dict(cutoff=1e-11, fp32_bound=1e-4 if mixed else 0.0, mixed=mixed)
    for wl, seed in ((workloads.cluster_220bf, 1), (workloads.cluster_66bf, 2)):  # larger, then smaller
        nuc, basis, ne, info = wl()
        eng = Engine(basis, nuc)
        P = _density(len(basis), ne, seed)
        J0, K0 = eng.jk(P, **kw)  # builds the work items for these thresholds
        rc, J1, K1 = _call(eng, P, kw['cutoff'], kw['fp32_bound'], flags)
        assert rc == 0, eng.lib.gx_last_error()
        torch.cuda.synchronize()
        dj = float((J1 - J0).abs().max())
        dk = float((K1 - K0).abs().max())
        print('%s mixed=%s: max |noalloc - gx_jk| J %.1e K %.1e' % (info['workload'], mixed, dj, dk))
        assert dj < 1e-10 and dk < 1e-10
        eng.close()


def test_noalloc_errors():
    from gaussium_amd import _lib, workloads
    from gaussium_amd.engine import Engine
    nuc, basis, ne, info = workloads.cluster_66bf()
    eng = Engine(basis, nuc)
    P = _density(len(basis), ne, 3)
    rc, _, _ = _call(eng, P, 1e-11, 0.0, 0)  # no work items for these thresholds yet
    assert rc == -3  # GX_ERR_STATE
    rc, _, _ = _call(eng, P, 1e-11, 0.0, _lib.GX_JK_TIMING)
    assert rc == -1
    eng.close()
    # 330 functions: 3 * 330^2 + 8 * (about 330^2 / 2) doubles exceed the scratch
    nuc, basis, ne, info = workloads.cluster(30, 30, 14.0, seed=7)
    assert len(basis) > 256
    eng = Engine(basis, nuc)
    P = _density(len(basis), ne, 4)
    rc, _, _ = _call(eng, P, 1e-11, 0.0, 0)
    assert rc == -1 and b'scratch' in eng.lib.gx_last_error()
    eng.close()
