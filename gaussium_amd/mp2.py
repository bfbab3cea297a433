"""Second-order Moller-Plesset energy on the device.

moellerplesset/moeller_plesset.py:7-26 of the reference: a converged RHF,
the AO->MO transformation of the repulsion integrals
(matrixelements/transformations.py:5-26) and

    E_MP2 = sum_{ijab} [2 (ia|jb)^2 - (ia|jb)(ib|ja)] / (e_i + e_j - e_a - e_b)

over occupied i, j and virtual a, b.  The reference transforms the whole
N^4 tensor in place, one N x N slice at a time; here the dense AO tensor
comes from the class kernels (``gx_eri_dense``) and only the (ia|jb) block is
formed, by four FP64 GEMMs on the device (one index at a time: occupied,
virtual, occupied, virtual), which rocBLAS runs on the FP64 matrix cores.
"""
import torch


def transform_ovov(eri, c_occ, c_vir):
    """(ia|jb) [o, v, o, v] from the AO tensor eri [N, N, N, N] (device, FP64)."""
    n = eri.shape[0]
    o, v = c_occ.shape[1], c_vir.shape[1]
    t = (c_occ.T @ eri.reshape(n, n * n * n)).reshape(o, n, n, n)             # [i, nu, la, si]
    t = t.permute(0, 2, 3, 1).contiguous().reshape(o * n * n, n) @ c_vir       # [i, la, si, a]
    t = t.reshape(o, n, n, v).permute(0, 3, 2, 1).contiguous().reshape(o * v * n, n) @ c_occ  # [i, a, si, j]
    t = t.reshape(o, v, n, o).permute(0, 1, 3, 2).contiguous().reshape(o * v * o, n) @ c_vir  # [i, a, j, b]
    return t.reshape(o, v, o, v)


def mp2_correlation(ovov, eps_occ, eps_vir):
    """Closed-shell MP2 correlation energy from (ia|jb) and orbital energies."""
    denom = (eps_occ[:, None, None, None] - eps_vir[None, :, None, None]
             + eps_occ[None, None, :, None] - eps_vir[None, None, None, :])
    exch = ovov.permute(0, 3, 2, 1)  # (ib|ja)
    return float(torch.sum((2.0 * ovov * ovov - ovov * exch) / denom))


class MoellerPlesset:
    """MP2 on top of a ``RestrictedHF`` (moeller_plesset.py:7-26)."""

    def __init__(self, hartree_fock):
        self.hartree_fock = hartree_fock
        self.hartree_fock_energy, eps, c = hartree_fock.begin_scf()
        nocc = hartree_fock.electrons // 2
        self.orbital_energies = eps
        eri = hartree_fock.engine.eri_dense()
        self.ovov = transform_ovov(eri, c[:, :nocc].contiguous(), c[:, nocc:].contiguous())
        del eri
        self.nocc = nocc

    def second_order(self):
        eps = self.orbital_energies
        return mp2_correlation(self.ovov, eps[:self.nocc], eps[self.nocc:])

    def energies(self):
        return self.hartree_fock_energy, self.second_order()
