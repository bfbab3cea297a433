"""Basis-set data model, file readers and shell packing.

The duck types keep the reference's attribute names so that objects of
either package can be handed to the engine:

* ``PrimitiveBasis``  (contraction, exponent, coordinates, integral_exponents,
  normalisation) -- objects/primitive_basis_function.py:6-42 of the reference
* ``Basis``           (primitive_gaussian_array, coordinates,
  integral_exponents, normalisation) -- objects/basis_function.py:6-53
* ``Nuclei``          (element, charge, mass, coordinates) -- objects/nuclei.py

Normalisation uses (-1)!! = 1 (the convention the reference's published
energies assume), a single-primitive function keeps contracted
normalisation 1 (basis_function.py:35-36).  ``read_basis_set_file`` /
``read_mol_file`` keep the reference's file semantics (common/file_input.py:6-88):
GAMESS-UK ``.gbs`` with coefficient-first rows, shell letters S/L/P/D,
``.mol`` in bohr with a ``charge multiplicity`` header.

``pack_shells`` turns a list of contracted functions into the shell arrays
of the C ABI (``gx_set_basis``).
"""
import itertools
import re
from math import pi, sqrt

import numpy as np


def double_factorial(n):
    if n <= 0:
        return 1
    out = 1
    while n > 1:
        out *= n
        n -= 2
    return out


class Nuclei:
    def __init__(self, element, charge, mass, coordinates):
        self.element = element
        self.charge = charge
        self.mass = mass
        self.coordinates = tuple(float(x) for x in coordinates)


class PrimitiveBasis:
    def __init__(self, contraction, exponent, coordinates, integral_exponents):
        self.contraction = contraction
        self.exponent = exponent
        self.coordinates = tuple(coordinates)
        self.integral_exponents = tuple(integral_exponents)
        self._norm = None

    @property
    def normalisation(self):
        if self._norm is None:
            l, m, n = self.integral_exponents
            df = double_factorial(2 * l - 1) * double_factorial(2 * m - 1) * double_factorial(2 * n - 1)
            self._norm = 1.0 / sqrt(df * (pi / (2 * self.exponent)) ** 1.5 / (4 * self.exponent) ** (l + m + n))
        return self._norm


class Basis:
    def __init__(self, primitive_gaussian_array, coordinates, integral_exponents):
        self.primitive_gaussian_array = primitive_gaussian_array
        self.coordinates = tuple(coordinates)
        self.integral_exponents = tuple(integral_exponents)
        self._norm = None

    @property
    def normalisation(self):
        if self._norm is None:
            prims = self.primitive_gaussian_array
            if len(prims) == 1:
                self._norm = 1
            else:
                l, m, n = self.integral_exponents
                df = double_factorial(2 * l - 1) * double_factorial(2 * m - 1) * double_factorial(2 * n - 1)
                s = 0.0
                for a, b in itertools.product(prims, repeat=2):
                    g = a.exponent + b.exponent
                    s += (a.contraction * b.contraction * a.normalisation * b.normalisation * df
                          * (pi / g) ** 1.5 / (2 * g) ** (l + m + n))
                self._norm = 1.0 / sqrt(s)
        return self._norm


# Cartesian component order of the reference (primitive_basis_factory.py:30-85).
COMPONENTS = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
}
_LETTER_L = {'S': 0, 'P': 1, 'D': 2}


def expand_shells(shells, coordinates):
    """Shell list ``[[letter, [c, a(, c_p)], ...], ...]`` -> list of ``Basis``."""
    out = []
    for shell in shells:
        letter, rows = shell[0], shell[1:]
        if letter == 'L':
            out.append(Basis([PrimitiveBasis(r[0], r[1], coordinates, (0, 0, 0)) for r in rows],
                             coordinates, (0, 0, 0)))
            for comp in COMPONENTS[1]:
                out.append(Basis([PrimitiveBasis(r[2], r[1], coordinates, comp) for r in rows],
                                 coordinates, comp))
        elif letter in _LETTER_L:
            for comp in COMPONENTS[_LETTER_L[letter]]:
                out.append(Basis([PrimitiveBasis(r[0], r[1], coordinates, comp) for r in rows],
                                 coordinates, comp))
        else:
            raise ValueError('unsupported shell letter %r' % letter)
    return out


def parse_gbs(text, element):
    """Shells of ``element`` in GAMESS-UK text (same section lookup as file_input.py:27-54)."""
    flat = ' '.join(text.replace('\n', ':').split()) + '#'
    m = re.search(element + '.*?#', flat)
    if m is None:
        raise ValueError('element %s not found in basis file' % element)
    lines = re.split(':', m.group(0).replace(': ', ':'))[1:-2]
    shells = []
    for line in lines:
        if any(letter in line for letter in ('S', 'L', 'P', 'D')):
            shells.append([line[0]])
        elif line.strip():
            shells[-1].append([float(x) for x in line.split()])
    return shells


def read_basis_set_file(path, nuclei_array):
    with open(path) as f:
        text = f.read()
    out = []
    for nuc in nuclei_array:
        out += expand_shells(parse_gbs(text, nuc.element), nuc.coordinates)
    return out


def read_mol_file(path):
    """(nuclei, electron count, multiplicity) as file_input.py:59-88."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines()]
    charge, mult = (int(x) for x in lines[0].split()[:2])
    nuclei, z = [], 0
    for line in lines[1:]:
        arr = line.split()
        if not arr:
            continue
        z += int(arr[1])
        nuclei.append(Nuclei(arr[0], float(arr[1]), float(arr[2]), (float(arr[3]), float(arr[4]), float(arr[5]))))
    return nuclei, z - charge, mult


def nuclear_repulsion(nuclei):
    """sum_{A<B} Z_A Z_B / R_AB (common/nuclear_repulsion.py:5-24, energy/energy.py:27-28)."""
    xyz = np.array([n.coordinates for n in nuclei], dtype=np.float64)
    z = np.array([float(n.charge) for n in nuclei])
    e = 0.0
    for i in range(len(nuclei)):
        d = np.sqrt(((xyz[i + 1:] - xyz[i]) ** 2).sum(axis=1))
        e += float((z[i] * z[i + 1:] / d).sum())
    return e


class ShellSet:
    """Shell arrays of the C ABI plus the map back to basis functions."""

    def __init__(self, l, nprim, prim_off, bf_off, ctr, ctr1, atom, exps, coefs, bf_scale):
        self.l = np.ascontiguousarray(l, dtype=np.int32)
        self.nprim = np.ascontiguousarray(nprim, dtype=np.int32)
        self.prim_off = np.ascontiguousarray(prim_off, dtype=np.int32)
        self.bf_off = np.ascontiguousarray(bf_off, dtype=np.int32)
        self.ctr = np.ascontiguousarray(ctr, dtype=np.float64).reshape(-1, 3)
        self.ctr1 = np.ascontiguousarray(ctr1, dtype=np.float64).reshape(-1, 3)
        self.atom = np.ascontiguousarray(atom, dtype=np.int32)
        self.exps = np.ascontiguousarray(exps, dtype=np.float64)
        self.coefs = np.ascontiguousarray(coefs, dtype=np.float64)
        self.bf_scale = np.ascontiguousarray(bf_scale, dtype=np.float64)

    @property
    def nshell(self):
        return len(self.l)

    def moved(self, old_xyz, new_xyz):
        """The same shells with every shell moved as rigidly as its atom
        (``atom`` indices into ``old_xyz`` / ``new_xyz``, arrays [natom, 3])."""
        if np.any(self.atom < 0):
            raise ValueError('every shell needs an atom to move with the geometry')
        d = np.asarray(new_xyz, dtype=np.float64) - np.asarray(old_xyz, dtype=np.float64)
        out = ShellSet(self.l, self.nprim, self.prim_off, self.bf_off, self.ctr + d[self.atom],
                       self.ctr1 + d[self.atom], self.atom, self.exps, self.coefs, self.bf_scale)
        return out

    @property
    def nbf(self):
        return len(self.bf_scale)


def _effective(b):
    """Per-primitive weights n_bf * c_i * N_i and exponents of one function."""
    n = b.normalisation
    return ([n * g.contraction * g.normalisation for g in b.primitive_gaussian_array],
            [g.exponent for g in b.primitive_gaussian_array])


def pack_shells(basis, nuclei=None, rtol=1e-10):
    """Group consecutive Cartesian functions into shells.

    A shell is a run of functions on one centre with identical exponents whose
    components are exactly the reference order for l = sum(integral_exponents)
    and whose primitive weights are proportional (per-component normalisation).
    ``nuclei`` (optional) assigns each shell the index of the nucleus at its
    centre (-1 if none).
    """
    l_out, np_out, poff, bfo, ctr, ctr1, atom, exps, coefs = [], [], [], [], [], [], [], [], []
    scale = np.ones(len(basis))
    nuc_xyz = None if nuclei is None else np.array([n.coordinates for n in nuclei], dtype=np.float64)
    i = 0
    while i < len(basis):
        b0 = basis[i]
        lq = int(sum(b0.integral_exponents))
        if lq not in COMPONENTS:
            raise ValueError('basis function %d: angular momentum %d not supported' % (i, lq))
        comps = COMPONENTS[lq]
        group = basis[i:i + len(comps)]
        if len(group) != len(comps) or [tuple(g.integral_exponents) for g in group] != comps:
            raise ValueError('basis functions %d.. do not form a complete l=%d shell in reference '
                             'component order' % (i, lq))
        w0, e0 = _effective(b0)
        c0 = tuple(b0.coordinates)
        for k, g in enumerate(group):
            w, e = _effective(g)
            if tuple(g.coordinates) != c0 or len(e) != len(e0) or any(x != y for x, y in zip(e, e0)):
                raise ValueError('basis function %d does not share centre/exponents with its shell' % (i + k))
            ratio = w[0] / w0[0]
            if any(abs(wi - ratio * w0i) > rtol * abs(wi) for wi, w0i in zip(w, w0)):
                raise ValueError('basis function %d: contraction not proportional to its shell' % (i + k))
            scale[i + k] = ratio
        pc = tuple(b0.primitive_gaussian_array[0].coordinates)
        l_out.append(lq)
        np_out.append(len(e0))
        poff.append(len(exps))
        bfo.append(i)
        ctr.append(c0)
        ctr1.append(pc)
        if nuc_xyz is not None:
            d = np.abs(nuc_xyz - np.array(c0)).sum(axis=1)
            k = int(np.argmin(d))
            atom.append(k if d[k] < 1e-12 else -1)
        else:
            atom.append(-1)
        exps += e0
        coefs += w0
        i += len(comps)
    return ShellSet(l_out, np_out, poff, bfo, ctr, ctr1, atom, exps, coefs, scale)
