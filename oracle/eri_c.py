"""ctypes front end of the C oracle ``oracle/c/os_eri.c`` (test infrastructure only).

The C file restates ``integrals/twoelectronrepulsion/obara_saika_scheme.py``
and ``integrals/boys_function.py`` of the reference; this module flattens a
list of duck-typed ``Basis`` objects into the arrays it takes.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, '_build', 'liboracle_eri.so')
_lib = None


def build():
    subprocess.run(['make', '-s', '-C', _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_boys.restype = ctypes.c_double
        _lib.oracle_boys.argtypes = [ctypes.c_int, ctypes.c_double]
    return _lib


def flatten(basis):
    """Per-function centre, (lx,ly,lz), primitive range; per-primitive exponent and weight n*c*N."""
    ctr, lxyz, poff, npr, pexp, pw = [], [], [], [], [], []
    for b in basis:
        ctr += list(b.coordinates)
        lxyz += list(b.integral_exponents)
        poff.append(len(pexp))
        npr.append(len(b.primitive_gaussian_array))
        for g in b.primitive_gaussian_array:
            pexp.append(g.exponent)
            pw.append(b.normalisation * g.contraction * g.normalisation)
    return (np.ascontiguousarray(ctr, dtype=np.float64), np.ascontiguousarray(lxyz, dtype=np.int32),
            np.ascontiguousarray(poff, dtype=np.int32), np.ascontiguousarray(npr, dtype=np.int32),
            np.ascontiguousarray(pexp, dtype=np.float64), np.ascontiguousarray(pw, dtype=np.float64))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_exact_boys(on):
    """Select the Boys evaluation: False = the reference's truncated series
    (integrals/boys_function.py:4-54), True = the convergent series in long double."""
    _load().oracle_set_exact_boys(1 if on else 0)


class exact_boys:
    """Context manager: evaluate with the exact Boys function inside the block."""

    def __enter__(self):
        set_exact_boys(True)

    def __exit__(self, *exc):
        set_exact_boys(False)


def eri_keys(basis, keys):
    """Values of the (i,j,k,l) quartets in ``keys`` (int array [n,4])."""
    lib = _load()
    flat = flatten(basis)
    keys = np.ascontiguousarray(keys, dtype=np.int32).reshape(-1, 4)
    out = np.zeros(len(keys))
    lib.oracle_eri_keys(ctypes.c_int(len(basis)), *[_ptr(a) for a in flat],
                        ctypes.c_int64(len(keys)), _ptr(keys), _ptr(out))
    return out


def eri_dense(basis):
    """Dense [N,N,N,N] tensor, every canonical quartet evaluated once."""
    lib = _load()
    flat = flatten(basis)
    n = len(basis)
    out = np.zeros((n, n, n, n))
    lib.oracle_eri_dense(ctypes.c_int(n), *[_ptr(a) for a in flat], _ptr(out))
    return out


def boys(v, x):
    return _load().oracle_boys(int(v), float(x))
