"""Basis data model and file readers, restated from the reference.

* ``PrimitiveBasis.normalisation`` follows ``objects/primitive_basis_function.py:26-42``.
* ``Basis.normalisation`` follows ``objects/basis_function.py:22-53``
  (a single primitive keeps normalisation 1, ``:35-36``).
* ``expand_basis_set`` follows ``factory/primitive_basis_factory.py:5-86``
  (component order S | s,px,py,pz | px,py,pz | xx,yy,zz,xy,xz,yz).
* ``read_basis_set_file`` / ``read_mol_file`` follow ``common/file_input.py:6-88``.

The double factorial uses (-1)!! = 1, the convention the reference's
published energies were produced with (older scipy ``factorial2``).

Test infrastructure only; see ``oracle/__init__.py``.
"""
import re
from math import pi, sqrt
import itertools


def factorial2(n):
    """n!! with (-1)!! = 1 and (-3)!! treated as 0 (never reached)."""
    if n <= 0:
        return 1 if n >= -1 else 0
    out = 1
    while n > 1:
        out *= n
        n -= 2
    return out


class Nuclei:
    """Nucleus: element name, charge Z, mass, coordinates in bohr (objects/nuclei.py:1-16)."""

    def __init__(self, element, charge, mass, coordinates):
        self.element = element
        self.charge = charge
        self.mass = mass
        self.coordinates = tuple(coordinates)


class PrimitiveBasis:
    """Primitive Cartesian Gaussian (objects/primitive_basis_function.py:6-42)."""

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
            out1 = factorial2(2 * l - 1) * factorial2(2 * m - 1) * factorial2(2 * n - 1)
            out2 = (pi / (2 * self.exponent)) ** 1.5
            out3 = (4 * self.exponent) ** (l + m + n)
            self._norm = 1 / sqrt((out1 * out2) / out3)
        return self._norm


class Basis:
    """Contracted Cartesian Gaussian (objects/basis_function.py:6-53)."""

    def __init__(self, primitive_gaussian_array, coordinates, integral_exponents):
        self.primitive_gaussian_array = primitive_gaussian_array
        self.coordinates = tuple(coordinates)
        self.integral_exponents = tuple(integral_exponents)
        self._norm = None

    @property
    def normalisation(self):
        if self._norm is None:
            l, m, n = self.integral_exponents
            if len(self.primitive_gaussian_array) == 1:
                self._norm = 1
            else:
                ans = 0.0
                for pa, pb in itertools.product(self.primitive_gaussian_array, repeat=2):
                    a1, a2 = pa.exponent, pb.exponent
                    out1 = factorial2(2 * l - 1) * factorial2(2 * m - 1) * factorial2(2 * n - 1)
                    out2 = (pi / (a1 + a2)) ** 1.5
                    out3 = (2 * (a1 + a2)) ** (l + m + n)
                    ans += (pa.contraction * pb.contraction * pa.normalisation * pb.normalisation
                            * out1 * out2) / out3
                self._norm = 1 / sqrt(ans)
        return self._norm


_COMPONENTS = {
    'S': [(0, 0, 0)],
    'P': [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    'D': [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
}


def expand_basis_set(file_list, coordinates):
    """Shell list -> contracted Cartesian functions (primitive_basis_factory.py:5-86).

    ``file_list`` is a list of shells ``[letter, [c, a], [c, a], ...]``; an
    'L' shell rows are ``[c_s, a, c_p]``.
    """
    out = []
    for shell in file_list:
        letter, rows = shell[0], shell[1:]
        if letter == 'L':
            out.append(Basis([PrimitiveBasis(r[0], r[1], coordinates, (0, 0, 0)) for r in rows],
                             coordinates, (0, 0, 0)))
            for comp in _COMPONENTS['P']:
                out.append(Basis([PrimitiveBasis(r[2], r[1], coordinates, comp) for r in rows],
                                 coordinates, comp))
        elif letter in _COMPONENTS:
            for comp in _COMPONENTS[letter]:
                out.append(Basis([PrimitiveBasis(r[0], r[1], coordinates, comp) for r in rows],
                                 coordinates, comp))
    return out


def parse_basis_text(text, element):
    """Shell list of ``element`` from GAMESS-UK ``.gbs`` text (file_input.py:27-54)."""
    lines = text.replace('\n', ':')
    lines = ' '.join(lines.split()) + '#'
    lines = re.search(element + '.*?#', lines).group(0)
    lines = re.split(':', lines.replace(': ', ':'))
    lines = lines[1:len(lines) - 2]
    shells = []
    cur = None
    for line in lines:
        if any(letter in line for letter in ('S', 'L', 'P', 'D')):
            if cur is not None:
                shells.append(cur)
            cur = [line[0]]
        else:
            cur.append([float(b) for b in line.split()])
    if cur is not None:
        shells.append(cur)
    return shells


def read_basis_set_file(path, nuclei_array):
    """List of ``Basis`` for every nucleus in file order (file_input.py:6-56)."""
    with open(path) as f:
        text = f.read()
    basis = []
    for nuc in nuclei_array:
        basis += expand_basis_set(parse_basis_text(text, nuc.element), nuc.coordinates)
    return basis


def read_mol_file(path):
    """(nuclei, electrons, multiplicity) from a ``.mol`` file (file_input.py:59-88)."""
    with open(path) as f:
        lines = f.readlines()
    charge = int(lines[0].split()[0])
    mult = int(lines[0].split()[1])
    nuclei = []
    z_total = 0
    for line in lines[1:]:
        arr = line.split()
        if not arr:
            continue
        z_total += int(arr[1])
        nuclei.append(Nuclei(arr[0], float(arr[1]), float(arr[2]),
                             (float(arr[3]), float(arr[4]), float(arr[5]))))
    return nuclei, z_total - charge, mult


def nuclear_repulsion(nuclei):
    """E_nuc = sum_{A<B} Z_A Z_B / R_AB (common/nuclear_repulsion.py:5-42, energy/energy.py:27-28)."""
    e = 0.0
    for i, a in enumerate(nuclei):
        for b in nuclei[i + 1:]:
            r = sqrt(sum((x - y) ** 2 for x, y in zip(a.coordinates, b.coordinates)))
            e += a.charge * b.charge / r
    return e


def sort_index(a, b, c, d):
    """Canonical 8-fold representative of (ab|cd) (common/symmetry.py:186-197)."""
    if a > b:
        a, b = b, a
    if c > d:
        c, d = d, c
    if a > c:
        a, c = c, a
        b, d = d, b
    if a == c and b > d:
        a, c = c, a
        b, d = d, b
    return a, b, c, d


def canonical_keys(n):
    """Canonical quartet keys in the reference's order (electron_repulsion_matrix.py:30-33)."""
    keys = []
    for a, b, c, d in itertools.product(range(n), repeat=4):
        if not (a > b or c > d or a > c or (a == c and b > d)):
            keys.append((a, b, c, d))
    return keys
