"""Boys function F_v(x), restated from the reference.

Follows ``integrals/boys_function.py:4-54`` (series for x <= 25, asymptotic
series for x > 25, each truncated at the first term below 1e-10) and
``integrals/boys_function.py:57-76`` (downward recursion).

Test infrastructure only; see ``oracle/__init__.py``.
"""
from math import exp, gamma, sqrt, pi, erf, fsum


_EXACT = [False]


class exact:
    """Context manager: ``boys_function`` evaluates the convergent series
    F_v(x) = e^-x sum_k (2x)^k / ((2v+1)(2v+3)...(2v+2k+1)) (no truncation at
    1e-10) inside the block when ``on`` is true."""

    def __init__(self, on=True):
        self.on = on

    def __enter__(self):
        self.prev = _EXACT[0]
        _EXACT[0] = bool(self.on)

    def __exit__(self, *exc):
        _EXACT[0] = self.prev


def boys_exact(v, x):
    """Convergent series for x < 60; Gamma(v+1/2) / (2 x^(v+1/2)) above (the
    neglected upper incomplete gamma is < 1e-16 relative for v <= 12)."""
    if x >= 60.0:
        return gamma(v + 0.5) / (2.0 * x ** (v + 0.5))
    term = 1.0 / (2 * v + 1)
    terms = [term]
    k = 1
    while True:
        term *= 2.0 * x / (2 * v + 2 * k + 1)
        terms.append(term)
        if term < terms[0] * 1e-20 and 2.0 * x < 2 * v + 2 * k + 1:
            break
        k += 1
    return exp(-x) * fsum(terms)


def boys_function(v, x):
    """F_v(x) exactly as the reference evaluates it (boys_function.py:4-54)."""
    if _EXACT[0]:
        return boys_exact(v, x)
    if x <= 25:
        i = 0
        ans = 0.0
        g_v = gamma(v + 0.5)
        while True:
            seq = (g_v / gamma(v + i + 1.5)) * x ** i
            if seq < 1e-10:
                break
            ans += seq
            i += 1
        return ans * 0.5 * exp(-x)
    i = 0
    ans = 0.0
    g_v = gamma(v + 0.5)
    while True:
        seq = (g_v / gamma(v - i + 1.5)) * x ** (-i)
        if seq < 1e-10:
            break
        ans += seq
        i += 1
    ans *= 0.5 * exp(-x)
    return (g_v / (2 * x ** (v + 0.5))) - ans


def boys_function_recursion(v, x, f_v):
    """F_{v-1}(x) from F_v(x) (boys_function.py:57-76)."""
    return (exp(-x) + 2 * x * f_v) / (2 * v - 1)


def boys_f0_closed_form(x):
    """F_0(x) = 1/2 sqrt(pi/x) erf(sqrt x): the closed form used to pin F_0."""
    if x < 1e-15:
        return 1.0 - x / 3.0
    return 0.5 * sqrt(pi / x) * erf(sqrt(x))


def boys_mpmath(v, x, dps=40):
    """F_v(x) = gamma_lower(v+1/2, x) / (2 x^(v+1/2)) at high precision."""
    import mpmath
    with mpmath.workdps(dps):
        x = mpmath.mpf(x)
        if x == 0:
            return mpmath.mpf(1) / (2 * v + 1)
        return mpmath.gammainc(v + mpmath.mpf(0.5), 0, x) / (2 * x ** (v + mpmath.mpf(0.5)))
