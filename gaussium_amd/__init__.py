"""gaussium_amd -- MI355X-native ERI and J/K Fock-build engine.

Drop-in for the ERI evaluation and restricted J/K Fock construction of the
Gaussium quantum-chemistry code: ``start(mol, basis, method, procs)`` and the
integral / SCF call sites keep their signatures; the work runs in
hand-written HIP kernels (libgx.so, C ABI in include/gx.h).
"""
from .basis import (Basis, Nuclei, PrimitiveBasis, nuclear_repulsion, pack_shells, read_basis_set_file,
                    read_mol_file)

__all__ = ['Basis', 'Nuclei', 'PrimitiveBasis', 'nuclear_repulsion', 'pack_shells', 'read_basis_set_file',
           'read_mol_file', 'start', 'Energy', 'Engine', 'RestrictedHF', 'FockMatrixRestricted',
           'TwoElectronRepulsionMatrixOS']


def __getattr__(name):
    # torch-dependent parts load lazily so that the data model imports without a GPU stack.
    if name in ('start', 'Energy'):
        from . import main
        return getattr(main, name)
    if name == 'Engine':
        from .engine import Engine
        return Engine
    if name in ('RestrictedHF', 'FockMatrixRestricted', 'TwoElectronRepulsionMatrixOS'):
        from . import hartree_fock
        return getattr(hartree_fock, name)
    raise AttributeError(name)
