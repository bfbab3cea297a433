"""The device Boys function (gx_boys_values: the Taylor tables and the
branch-free evaluation the J/K kernels use) against the oracle's exact
F_m(T) (convergent series / incomplete-gamma form, pinned to mpmath in
test_oracle.py), on both grids, FP64 and FP32, below and above the table's
end (T = 36, where the asymptotic form takes over), at grid points and
half-way between them (the largest Taylor step).  Tolerances: FP64 4e-15
relative on the tables; on the asymptotic form 1.2e-14 relative (its rsqrt
is the hardware estimate plus one Newton step) plus 5e-18 absolute (the
e^-T / (2T) terms its upward recursion drops); FP32 2e-6 relative."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _T():
    fine = np.arange(0, 361) * 0.1
    coarse = np.arange(0, 145) * 0.25
    mids = np.concatenate([fine[:-1] + 0.05, coarse[:-1] + 0.125])
    rng = np.random.default_rng(3)
    return np.unique(np.concatenate([fine, coarse, mids, rng.uniform(0, 36, 600), rng.uniform(36, 80, 100),
                                     [35.999, 36.0, 36.001, 1e-12, 1e-6]]))


@pytest.mark.parametrize('fp32', [0, 1], ids=['fp64', 'fp32'])
@pytest.mark.parametrize('grid', [0, 1], ids=['fine', 'coarse'])
def test_device_boys_matches_exact(grid, fp32):
    import torch
    from gaussium_amd import _lib, workloads
    from gaussium_amd.engine import Engine
    from oracle.boys import boys_exact
    nuc, basis, ne, info = workloads.triatomic_7bf()
    eng = Engine(basis, nuc)
    T = _T()
    Td = torch.from_numpy(T).cuda()
    for L in (0, 1, 4, 8, 10):
        out = torch.full((len(T), L + 1), np.nan, dtype=torch.float64, device='cuda')
        eng._bind_stream()
        rc = eng.lib.gx_boys_values(eng.ctx, grid, fp32, L, ctypes.c_void_p(Td.data_ptr()), len(T),
                                    ctypes.c_void_p(out.data_ptr()))
        assert rc == 0, eng.lib.gx_last_error()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        ref = np.array([[boys_exact(m, t) for m in range(L + 1)] for t in T])
        err = np.abs(got - ref)
        # tables (T < 36): relative; asymptotic form (T >= 36): relative (the
        # rsqrt) plus the absolute error of the dropped e^-T / (2T) terms of
        # the upward recursion (4e-18, large against F_m of high m near 36)
        for sel, rtol, atol in ((T < 36.0, 2e-6 if fp32 else 4e-15, 0.0),
                                (T >= 36.0, 2e-6 if fp32 else 1.2e-14, 5e-18)):
            bad = err[sel] - (rtol * ref[sel] + atol)
            worst = np.unravel_index(np.argmax(bad), bad.shape)
            assert bad.max() <= 0, (grid, fp32, L, T[sel][worst[0]], worst[1], err[sel][worst], ref[sel][worst])
    eng.close()


def test_boys_values_rejects_bad_arguments():
    from gaussium_amd import workloads
    from gaussium_amd.engine import Engine
    nuc, basis, ne, info = workloads.triatomic_7bf()
    eng = Engine(basis, nuc)
    assert eng.lib.gx_boys_values(eng.ctx, 2, 0, 0, None, 0, None) == -1
    assert eng.lib.gx_boys_values(eng.ctx, 0, 0, 11, None, 0, None) == -1
    assert eng.lib.gx_boys_values(eng.ctx, 0, 0, 4, None, 0, None) == 0  # nothing to do
    eng.close()
