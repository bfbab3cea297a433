"""Analytic RHF nuclear gradients behind ``torch.autograd``.

The reference has no analytic gradient: its geometry optimiser evaluates
full energies at simplex vertices (geometryoptimization/nelder_mead.py:126-144,
``main.py:21-23``).  Here ``RHFSurface`` holds a molecule and its basis as
shells attached to atoms; ``energy_and_gradient(xyz)`` moves the shells with
the atoms, runs the SCF on the device and evaluates dE/dR with ``gx_grad``
(one-electron derivatives, the two-electron derivative quartets contracted
with the density, nuclear repulsion).  ``RHFEnergy.apply(xyz, surface)`` is
the same as a differentiable function of the coordinates.
"""
import numpy as np
import torch

from .basis import Nuclei, nuclear_repulsion, pack_shells
from .engine import Engine
from .hartree_fock import RestrictedHF


class RHFSurface:
    """Closed-shell RHF energy as a function of the nuclear coordinates (bohr)."""

    def __init__(self, nuclei, basis, electrons, engine=None, schwarz_cutoff=0.0, threshold=1e-12,
                 density_threshold=1e-9, max_iter=300, stabilizer=None, precision='fp64', fp32_bound=0.0):
        self.nuclei = list(nuclei)
        self.electrons = electrons
        self.xyz0 = np.array([n.coordinates for n in self.nuclei], dtype=np.float64)
        self.shells0 = pack_shells(basis, self.nuclei)
        if np.any(self.shells0.atom < 0):
            raise ValueError('every basis function must sit on a nucleus')
        self.engine = engine if engine is not None else Engine()
        self.options = dict(schwarz_cutoff=schwarz_cutoff, threshold=threshold,
                            density_threshold=density_threshold, max_iter=max_iter, stabilizer=stabilizer,
                            precision=precision, fp32_bound=fp32_bound)
        self.evaluations = 0
        self.last = None

    def _nuclei_at(self, xyz):
        return [Nuclei(n.element, n.charge, n.mass, tuple(x)) for n, x in zip(self.nuclei, xyz)]

    def _hf(self, xyz, initial_density=None):
        xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
        nuc = self._nuclei_at(xyz)
        self.engine.set_basis(self.shells0.moved(self.xyz0, xyz))
        self.engine.set_nuclei(nuc)
        hf = RestrictedHF(nuc, None, self.electrons, engine=self.engine, initial_density=initial_density,
                          **self.options)
        e = hf.energies()[0] + nuclear_repulsion(nuc)
        if not hf.scf_method.converged:
            raise RuntimeError('SCF did not converge in %d iterations (|dE| < %g, max |dP| < %g); the '
                               'gradient needs a converged density' % (hf.scf_method.iterations,
                                                                       self.options['threshold'],
                                                                       self.options['density_threshold']))
        self.evaluations += 1
        self.last = hf
        return hf, e

    def energy(self, xyz, initial_density=None):
        """Total energy at ``xyz``.  ``initial_density`` (e.g. the converged
        density of a nearby geometry, ``last.scf_method.density``) starts the
        SCF from the orbitals of its Fock matrix instead of the core-H guess,
        so that displaced geometries stay on the same SCF solution."""
        return self._hf(xyz, initial_density)[1]

    def energy_and_gradient(self, xyz, initial_density=None):
        hf, e = self._hf(xyz, initial_density)
        return e, hf.gradient().cpu().numpy()


class RHFEnergy(torch.autograd.Function):
    """E(xyz) with dE/dxyz from the analytic gradient kernels.

    ``xyz``: [natom, 3] float64 tensor (any device); returns a 0-d float64
    tensor on the same device."""

    @staticmethod
    def forward(ctx, xyz, surface):
        e, g = surface.energy_and_gradient(xyz.detach().cpu().numpy())
        ctx.save_for_backward(torch.as_tensor(g, dtype=torch.float64, device=xyz.device))
        return torch.tensor(e, dtype=torch.float64, device=xyz.device)

    @staticmethod
    def backward(ctx, grad_output):
        (g,) = ctx.saved_tensors
        return grad_output * g, None
