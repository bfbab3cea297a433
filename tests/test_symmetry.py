"""Point groups and the integrals they force to zero (gaussium_amd.symmetry).

* The reference's molecule-factory fixtures (tests/golden/point_groups.json:
  nuclei and the labels and element counts its tests assert) give the same
  labels and counts.
* The zero pattern is checked against the C oracle's integrals: on exactly
  symmetric molecules every integral flagged zero is zero to 1e-14 of the
  largest, and every integral below 1e-13 of the largest is flagged (the
  test is exact, not a heuristic); C_1 flags nothing.
"""
import json
import os

import numpy as np
import pytest

from tests.conftest import ROOT

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def _factory():
    from gaussium_amd.symmetry import MoleculeFactory
    return MoleculeFactory(symmetry=True)


@pytest.mark.parametrize('name', sorted(json.load(open(os.path.join(GOLDEN, 'point_groups.json')))['molecules']))
def test_point_group_labels_match_reference_fixtures(name):
    from gaussium_amd.basis import Nuclei
    m = json.load(open(os.path.join(GOLDEN, 'point_groups.json')))['molecules'][name]
    nuclei = [Nuclei('X', z, 1, c) for z, c in m['atoms']]
    out, pg = _factory().create(nuclei)
    assert pg.label == m['label']
    if 'rotations' in m:
        assert len(pg.elements.rotations) == m['rotations']
    if 'reflections' in m:
        assert len(pg.elements.reflections) == m['reflections']
    # the nuclei are centred and every operation maps the framework onto itself
    X = np.array([n.coordinates for n in out])
    Z = np.array([n.charge for n in out])
    assert np.abs(X.mean(axis=0)).max() < 1e-9
    for R in pg.operations:
        Y = X @ np.asarray(R).T
        for i in range(len(Z)):
            assert np.any((np.linalg.norm(X - Y[i], axis=1) < 1e-3) & (Z == Z[i]))


def _methane(a=1.18, shift=(0.3, 0.1, -0.2)):
    from gaussium_amd.basis import Nuclei
    s = np.asarray(shift)
    return [Nuclei('CARBON', 6, 12, tuple(s))] + [
        Nuclei('HYDROGEN', 1, 1, tuple(s + a * np.array([sx, sy, sx * sy]))) for sx in (1, -1) for sy in (1, -1)]


def _molecule(mol):
    from gaussium_amd.basis import read_mol_file
    if mol == 'CH4':
        return _methane()
    return read_mol_file(os.path.join(GOLDEN, mol + '.mol'))[0]


@pytest.mark.parametrize('mol,gbs,label', [('H2O', 'STO-3G', 'C_{2v}'), ('C2H4', 'STO-3G', 'D_{2h}'),
                                           ('CO', 'STO-3G', 'C_{4v}'), ('H2O', '3-21G', 'C_{2v}'),
                                           ('CH4', 'STO-3G', 'T_{d}')])
def test_zero_pattern_matches_oracle_integrals(mol, gbs, label):
    from gaussium_amd.basis import read_basis_set_file
    from gaussium_amd.symmetry import Symmetry
    from oracle import eri_c
    nuclei, pg = _factory().create(_molecule(mol))
    assert pg.label == label
    basis = read_basis_set_file(os.path.join(GOLDEN, gbs + '.gbs'), nuclei)
    sym = Symmetry(pg, basis)
    n = len(basis)
    mask = sym.zero_mask()
    vals = eri_c.eri_keys(basis, np.indices((n,) * 4).reshape(4, -1).T).reshape((n,) * 4)
    scale = np.abs(vals).max()
    assert mask.sum() > 0
    assert np.abs(vals[mask]).max() <= 1e-14 * scale
    small = np.abs(vals) < 1e-13 * scale
    assert np.array_equal(small, mask), (int(small.sum()), int(mask.sum()))
    # the per-index query agrees with the mask
    rng = np.random.default_rng(5)
    for q in rng.integers(0, n, (200, 4)):
        assert sym.none_zero_integral(tuple(int(x) for x in q)) == (not mask[tuple(q)])


def test_c1_flags_nothing():
    from gaussium_amd.basis import Nuclei, read_basis_set_file
    from gaussium_amd.symmetry import MoleculeFactory, Symmetry
    # four nuclei out of one plane (three are always planar: C_s)
    nuclei = [Nuclei('OXYGEN', 8, 16, (0.1, 0.2, 0.3)), Nuclei('HYDROGEN', 1, 1, (1.5, 0.3, 0.2)),
              Nuclei('HYDROGEN', 1, 1, (-0.3, 1.7, 0.9)), Nuclei('HYDROGEN', 1, 1, (0.4, -0.6, 1.8))]
    nuclei, pg = MoleculeFactory(symmetry=True).create(nuclei)
    assert pg.label == 'C_{1}'
    sym = Symmetry(pg, read_basis_set_file(os.path.join(GOLDEN, 'STO-3G.gbs'), nuclei))
    assert not sym.zero_mask().any() and sym.none_zero_integral((0, 1, 2, 3))
    # symmetry=False leaves the nuclei alone and reports C_1
    same, pg0 = MoleculeFactory(symmetry=False).create(nuclei)
    assert pg0.label == 'C_{1}' and same is nuclei
