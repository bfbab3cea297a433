"""ctypes binding of libgx.so (the C ABI in include/gx.h).

The library must have been built in-tree (``gaussium_amd.build.build()``);
there is no fallback: every entry point raises if it is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libgx.so')

GX_JK_MIXED = 1
GX_JK_TIMING = 2

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_double_p = ctypes.POINTER(ctypes.c_double)


class GxStats(ctypes.Structure):
    _fields_ = [('shell_quartets', ctypes.c_int64), ('bf_quartets', ctypes.c_int64),
                ('shell_quartets_fp64', ctypes.c_int64), ('shell_quartets_fp32', ctypes.c_int64),
                ('bf_quartets_fp64', ctypes.c_int64), ('bf_quartets_fp32', ctypes.c_int64),
                ('flops_fp64', ctypes.c_double), ('flops_fp32', ctypes.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    'gx_last_error': (ctypes.c_char_p, []),
    'gx_create': (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    'gx_destroy': (ctypes.c_int, [ctypes.c_void_p]),
    'gx_set_stream': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    'gx_set_basis': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_void_p]),
    'gx_set_basis_batch': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_void_p]),
    'gx_set_nuclei': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    'gx_nbf': (ctypes.c_int, [ctypes.c_void_p]),
    'gx_npair': (ctypes.c_int, [ctypes.c_void_p]),
    'gx_set_prim_screen': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    'gx_nprimpair': (ctypes.c_int, [ctypes.c_void_p]),
    'gx_schwarz': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    'gx_eri_quartets': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    'gx_eri_dense': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    'gx_jk': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    'gx_jk_noalloc': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    'gx_jk_uhf': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int]),
    'gx_jk_autotune': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_int]),
    'gx_grad': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_double]),
    'gx_grad_ex': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    'gx_grad_launch_times': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    'gx_grad_autotune': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]),
    'gx_jk_stats': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.POINTER(GxStats)]),
    'gx_jk_class_counts': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_int]),
    'gx_jk_launch_times': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    'gx_kernel_kind': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    'gx_boys_values': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_int, ctypes.c_void_p]),
    'gx_one_electron': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
}

_lib = None


class GxError(RuntimeError):
    """A gx C-ABI call returned an error code."""


def load():
    """Load libgx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GxError('libgx.so not found at %s: run gaussium_amd.build.build() '
                          '(python -m gaussium_amd.build)' % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc, what):
    if rc != 0:
        msg = load().gx_last_error()
        raise GxError('%s failed (%d): %s' % (what, rc, msg.decode() if msg else ''))
    return rc
