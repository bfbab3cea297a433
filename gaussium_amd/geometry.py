"""Gradient-driven geometry optimisation for ``start(..., geometry_optimization=...)``.

The reference minimises the energy with a Nelder-Mead simplex over internal
coordinates, one full SCF per vertex (geometryoptimization/nelder_mead.py:11-187,
called from main.py:21-23).  Here the analytic RHF gradient (``gx_grad``)
drives a quasi-Newton (BFGS) minimisation in Cartesian coordinates with a
backtracking line search; the result is the same kind of number -- the total
energy at the minimum -- reached with one energy and gradient per step
instead of (3n - 6) + 1 energies per simplex.
"""
import numpy as np

from .basis import read_basis_set_file


def bfgs_minimize(fun, x, gtol=3e-5, etol=1e-9, max_steps=200, max_step=0.3):
    """BFGS with a backtracking (Armijo) line search on ``fun(x) -> (e, g)``.

    Returns (x, e, history of (e, max |g|)).  A step is only taken when it
    lowers the energy sufficiently; when the line search fails the inverse
    Hessian restarts from steepest descent, and a second failure in a row
    ends the search at the best point so far.
    """
    x = np.asarray(x, dtype=np.float64).copy()
    e, g = fun(x)
    g = np.asarray(g, dtype=np.float64).ravel()
    n = x.size
    hinv = np.eye(n) * 0.5  # inverse Hessian guess (bohr^2 / Eh)
    history = [(e, float(np.abs(g).max()))]
    steepest = False
    for _ in range(max_steps):
        if np.abs(g).max() < gtol:
            break
        d = -hinv @ g
        step = np.abs(d).max()
        if step > max_step:
            d *= max_step / step
        if g @ d >= 0.0:  # not a descent direction: restart from steepest descent
            hinv = np.eye(n) * 0.5
            d = -hinv @ g
        t = 1.0
        accepted = False
        while t >= 1e-3:
            x_new = x + t * d
            e_new, g_new = fun(x_new)
            if e_new <= e + 1e-4 * t * (g @ d):
                accepted = True
                break
            t *= 0.5
        if not accepted:
            if steepest:
                break
            hinv = np.eye(n) * 0.5
            steepest = True
            continue
        steepest = False
        g_new = np.asarray(g_new, dtype=np.float64).ravel()
        s, y = x_new - x, g_new - g
        sy = s @ y
        if sy > 1e-12:
            rho = 1.0 / sy
            v = np.eye(n) - rho * np.outer(s, y)
            hinv = v @ hinv @ v.T + rho * np.outer(s, s)
        de = e - e_new
        x, e, g = x_new, e_new, g_new
        history.append((e, float(np.abs(g).max())))
        if abs(de) < etol and np.abs(g).max() < 10 * gtol:
            break
    return x, e, history


def optimize(basis_file, energy_object, nuclei, method=None, gtol=3e-5, etol=1e-9, max_steps=200,
             max_step_bohr=0.3):
    """Minimise the total energy over the nuclear coordinates; returns it.

    ``method`` is accepted for signature parity (the reference passes
    whatever ``geometry_optimization`` was given); the minimiser is BFGS on
    the analytic gradient.  Stops when max |dE/dR| < ``gtol`` (Eh/bohr) or
    the energy changes by less than ``etol`` in a step.
    """
    from .gradient import RHFSurface
    if energy_object.method != 'RHF':
        raise NotImplementedError('geometry optimisation needs the RHF gradient (method %r)'
                                  % (energy_object.method,))
    basis = read_basis_set_file(basis_file, nuclei)
    opts = {k: v for k, v in energy_object.options.items()
            if k in ('schwarz_cutoff', 'max_iter', 'stabilizer', 'precision', 'fp32_bound')}
    surface = RHFSurface(nuclei, basis, energy_object.electrons, **opts)
    x, e, history = bfgs_minimize(surface.energy_and_gradient, surface.xyz0.ravel().copy(), gtol=gtol,
                                  etol=etol, max_steps=max_steps, max_step=max_step_bohr)
    energy_object.optimisation = {'steps': len(history) - 1, 'history': history,
                                  'coordinates': x.reshape(-1, 3), 'energy_evaluations': surface.evaluations}
    return e
