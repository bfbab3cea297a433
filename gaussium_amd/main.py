"""``start()`` and ``Energy`` with the reference's signatures.

main.py:11-34 and energy/energy.py:17-81 of the reference.  The 'RHF',
'UHF', 'GUHF' and 'MP2' methods run on this engine (the other branches of
``Energy.calculate_energy`` -- DFT, CCSD, TDHF, CIS -- are outside the
ERI + J/K path and raise ``NotImplementedError``).
``symmetry=True`` centres and orients the molecule and finds its point
group (``symmetry.MoleculeFactory``, main.py:25-27 of the reference); the
``Symmetry`` object then zeroes the integrals the group forbids in the dense
repulsion matrix (``TwoElectronRepulsionMatrixOS``).  The direct J/K build
of the SCF evaluates whole shell quartets and is unaffected: the energy is
the same as without symmetry.
"""
import time

from .basis import nuclear_repulsion, read_basis_set_file, read_mol_file
from .hartree_fock import BlockedHartreeFock, RestrictedHF, UnrestrictedHF


class Energy:
    def __init__(self, electrons, multiplicity, processors, method, **options):
        self.electrons = electrons
        self.multiplicity = multiplicity
        self.processors = processors
        self.method = method
        self.options = options
        self.symmetry_object = None
        self.point_group = None
        self.last = None

    def calculate_energy(self, nuclei_array, basis_set):
        e_nuc = nuclear_repulsion(nuclei_array)
        if self.method == 'RHF':
            calc = RestrictedHF(nuclei_array, basis_set, self.electrons, self.symmetry_object, self.processors,
                                **self.options)
        elif self.method == 'UHF':
            calc = UnrestrictedHF(nuclei_array, basis_set, self.electrons, self.multiplicity, self.symmetry_object,
                                  self.processors, **self.options)
        elif self.method == 'GUHF':
            calc = BlockedHartreeFock(nuclei_array, basis_set, self.electrons, self.multiplicity,
                                      self.symmetry_object, self.processors, **self.options)
        elif self.method == 'MP2':
            from .mp2 import MoellerPlesset
            calc = MoellerPlesset(RestrictedHF(nuclei_array, basis_set, self.electrons, self.symmetry_object,
                                               self.processors, **self.options))
        else:
            raise NotImplementedError('method %r is not provided by gaussium_amd (RHF, UHF, GUHF, MP2)'
                                      % (self.method,))
        electron_energy, correlation = calc.energies()
        self.last = calc
        return electron_energy + e_nuc + correlation


def start(mol_file, basis_file, method, processors, symmetry=False, geometry_optimization=None, **options):
    """Total energy of the molecule in ``mol_file`` with ``basis_file`` (main.py:11-34)."""
    t0 = time.perf_counter()
    nuclei, electrons, multiplicity = read_mol_file(mol_file)
    energy_object = Energy(electrons, multiplicity, processors, method, **options)
    if symmetry and geometry_optimization is None:
        from .symmetry import MoleculeFactory, Symmetry
        nuclei, point_group = MoleculeFactory(symmetry=True).create(nuclei)
        energy_object.point_group = point_group
        energy_object.symmetry_object = Symmetry(point_group, read_basis_set_file(basis_file, nuclei))
    if geometry_optimization is not None:
        from .geometry import optimize
        energy = optimize(basis_file, energy_object, nuclei, method=geometry_optimization)
    else:
        basis = read_basis_set_file(basis_file, nuclei)
        energy = energy_object.calculate_energy(nuclei, basis)
    energy_object.wall_time = time.perf_counter() - t0
    return energy
