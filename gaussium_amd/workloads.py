"""Seeded synthetic molecules and basis sets for the benchmark workloads.

Each builder returns ``(nuclei, basis, electrons, info)`` with the data model
of ``gaussium_amd.basis``.  Basis sets are per element (every heavy atom
carries the same shells, as in a ``.gbs`` file); exponents are log-uniform,
raw contraction coefficients uniform in (0.1, 1) and one-primitive shells
have coefficient 1.0 (the contracted normalisation of ``Basis`` is applied as
in the reference).

Workloads:
  triatomic_7bf      bent triatomic, heavy Z=8 at the origin, two Z=1 at
                     1.81 bohr, 104.5 deg; heavy s,s,p, light s; 3 primitives
                     each, exponents in [0.1, 130]; 7 functions, 10 electrons
  cluster(n_heavy, n_light, box, d_min, polarised)
                     random cluster of Z=6 / Z=1 atoms; heavy s[6] sp[2] sp[1]
                     (+ d[1] with exponent 0.8 if polarised), light s[2] s[1];
                     exponents in [0.08, 3000]; closed shell (cation if the
                     neutral electron count is odd)
  cluster_220bf      20 + 20 atoms in a 12-bohr box, d_min 1.8: 220 functions
  cluster_66bf       6 + 6 atoms (box 12 * (12/40)^(1/3) bohr): 66 functions
  cluster_255bf      15 + 15 atoms, polarised: 255 functions, 104 electrons
  batch(256)         uneven 4..40-atom clusters, each a 0.1-bohr Gaussian
                     perturbation of its base cluster
"""
import math

import numpy as np

from .basis import Nuclei, expand_shells

HEAVY = ('CARBON', 6.0, 12.0)
LIGHT = ('HYDROGEN', 1.0, 1.0)


def _exps(rng, n, lo, hi):
    return sorted(np.exp(rng.uniform(math.log(lo), math.log(hi), n)).tolist(), reverse=True)


def _coefs(rng, n):
    return [1.0] if n == 1 else rng.uniform(0.1, 1.0, n).tolist()


def _shell(rng, letter, n, lo, hi):
    e = _exps(rng, n, lo, hi)
    c = _coefs(rng, n)
    if letter == 'L':
        cp = _coefs(rng, n)
        return ['L'] + [[c[i], e[i], cp[i]] for i in range(n)]
    return [letter] + [[c[i], e[i]] for i in range(n)]


def split_valence_basis(rng, polarised=False, lo=0.08, hi=3000.0):
    heavy = [_shell(rng, 'S', 6, lo, hi), _shell(rng, 'L', 2, lo, hi), _shell(rng, 'L', 1, lo, hi)]
    if polarised:
        heavy.append(['D', [1.0, 0.8]])
    light = [_shell(rng, 'S', 2, lo, hi), _shell(rng, 'S', 1, lo, hi)]
    return heavy, light


def random_positions(rng, n, box, d_min, max_tries=200000):
    pts = []
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError('could not place %d atoms in a %.1f bohr box with d_min %.2f' % (n, box, d_min))
        x = rng.uniform(0.0, box, 3)
        if all(np.linalg.norm(x - p) >= d_min for p in pts):
            pts.append(x)
    return np.array(pts)


def build(elements, xyz, shells_by_element):
    nuclei = [Nuclei(el[0], el[1], el[2], tuple(float(v) for v in x)) for el, x in zip(elements, xyz)]
    basis = []
    for nuc, el in zip(nuclei, elements):
        basis += expand_shells(shells_by_element[el[0]], nuc.coordinates)
    return nuclei, basis


def triatomic_7bf(seed=1):
    rng = np.random.default_rng(seed)
    heavy = [_shell(rng, 'S', 3, 0.1, 130.0), _shell(rng, 'S', 3, 0.1, 130.0), _shell(rng, 'P', 3, 0.1, 130.0)]
    light = [_shell(rng, 'S', 3, 0.1, 130.0)]
    r, th = 1.81, math.radians(104.5)
    xyz = np.array([[0.0, 0.0, 0.0],
                    [r * math.sin(th / 2), 0.0, r * math.cos(th / 2)],
                    [-r * math.sin(th / 2), 0.0, r * math.cos(th / 2)]])
    elements = [('OXYGEN', 8.0, 16.0), LIGHT, LIGHT]
    nuclei, basis = build(elements, xyz, {'OXYGEN': heavy, 'HYDROGEN': light})
    return nuclei, basis, 10, {'workload': 'triatomic_7bf', 'seed': seed, 'nbf': len(basis)}


def cluster(n_heavy, n_light, box, d_min=1.8, polarised=False, seed=2, xyz=None):
    rng = np.random.default_rng(seed)
    heavy, light = split_valence_basis(rng, polarised)
    n = n_heavy + n_light
    if xyz is None:
        xyz = random_positions(rng, n, box, d_min)
    elements = [HEAVY] * n_heavy + [LIGHT] * n_light
    nuclei, basis = build(elements, xyz, {'CARBON': heavy, 'HYDROGEN': light})
    electrons = 6 * n_heavy + n_light
    charge = electrons % 2
    electrons -= charge
    info = {'n_atoms': n, 'n_heavy': n_heavy, 'n_light': n_light, 'box_bohr': box, 'd_min_bohr': d_min,
            'polarised': polarised, 'seed': seed, 'nbf': len(basis), 'electrons': electrons, 'charge': charge}
    return nuclei, basis, electrons, info


def cluster_220bf(seed=2):
    nuc, bas, ne, info = cluster(20, 20, 12.0, seed=seed)
    info['workload'] = 'cluster_40atom_220bf'
    return nuc, bas, ne, info


def cluster_66bf(seed=3):
    nuc, bas, ne, info = cluster(6, 6, 12.0 * (12.0 / 40.0) ** (1.0 / 3.0), seed=seed)
    info['workload'] = 'cluster_12atom_66bf'
    return nuc, bas, ne, info


def cluster_255bf(seed=5):
    nuc, bas, ne, info = cluster(15, 15, 12.0 * (30.0 / 40.0) ** (1.0 / 3.0), polarised=True, seed=seed)
    info['workload'] = 'cluster_30atom_255bf_d'
    return nuc, bas, ne, info


def batch(n_geom=256, seed=4, sigma=0.1):
    """Uneven clusters of 4..40 atoms (heavy = ceil(n/2)), perturbed by N(0, sigma)."""
    rng = np.random.default_rng(seed)
    out = []
    for g in range(n_geom):
        n = int(rng.integers(4, 41))
        nh = (n + 1) // 2
        box = 12.0 * (n / 40.0) ** (1.0 / 3.0)
        sub_seed = int(rng.integers(0, 2 ** 31))
        base = random_positions(np.random.default_rng(sub_seed), n, box, 1.8)
        xyz = base + rng.normal(0.0, sigma, base.shape)
        nuc, bas, ne, info = cluster(nh, n - nh, box, seed=seed, xyz=xyz)
        info['workload'] = 'batch_%d' % g
        out.append((nuc, bas, ne, info))
    return out
