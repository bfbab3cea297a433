"""Python handle of a gx context: basis upload, ERIs, direct J/K, one-electron.

Device buffers are torch tensors (FP64) whose ``data_ptr()`` is handed to the
C ABI; the context launches on torch's current stream so that the ordering
with surrounding torch work needs no extra synchronisation.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .basis import ShellSet, pack_shells


def _p(a):
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(a.data_ptr())


class Engine:
    """One gx context bound to a CUDA/HIP device.

    ``Engine(basis, nuclei)`` packs the contracted functions into shells,
    builds primitive pairs and Schwarz bounds on the device and keeps them for
    every later call.  ``prim_eps`` is the primitive-pair screening
    threshold (``gx_set_prim_screen``; None keeps the library default, 0
    keeps every primitive pair).
    """

    def __init__(self, basis=None, nuclei=None, device=None, prim_eps=None):
        if not torch.cuda.is_available():
            raise _lib.GxError('gaussium_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
        self.device = torch.device('cuda', torch.cuda.current_device() if device is None else device)
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.gx_create(ctypes.byref(h), self.device.index), 'gx_create')
        self.ctx = h
        self.shells = None
        self.nuclei = None
        if prim_eps is not None:
            self.set_prim_screen(prim_eps)
        if basis is not None:
            self.set_basis(basis, nuclei)
        if nuclei is not None:
            self.set_nuclei(nuclei)

    def close(self):
        if getattr(self, 'ctx', None) is not None and self.ctx.value:
            self.lib.gx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up -------------------------------------------------------------
    def _bind_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self.lib.gx_set_stream(self.ctx, ctypes.c_void_p(s)), 'gx_set_stream')

    def set_prim_screen(self, eps):
        """Primitive-pair screening threshold for the next set_basis / set_systems."""
        _lib.check(self.lib.gx_set_prim_screen(self.ctx, ctypes.c_double(eps)), 'gx_set_prim_screen')

    @property
    def n_primitive_pairs(self):
        """Primitive pairs kept for the current basis."""
        return int(self.lib.gx_nprimpair(self.ctx))

    def set_basis(self, basis, nuclei=None):
        sh = basis if isinstance(basis, ShellSet) else pack_shells(basis, nuclei)
        self._bind_stream()
        _lib.check(self.lib.gx_set_basis(
            self.ctx, sh.nshell, _p(sh.l), _p(sh.nprim), _p(sh.prim_off), _p(sh.bf_off), _p(sh.ctr),
            _p(sh.ctr1), _p(sh.atom), len(sh.exps), _p(sh.exps), _p(sh.coefs), sh.nbf, _p(sh.bf_scale)),
            'gx_set_basis')
        self.shells = sh
        self.sizes = [sh.nbf]
        self.offsets = np.array([0, sh.nbf * sh.nbf], dtype=np.int64)
        return sh

    def set_systems(self, systems):
        """A batch of independent molecules, ``[(basis, nuclei), ...]``: one
        context, one launch per class combination for all of them.  Matrices
        are then packed (``pack`` / ``unpack``)."""
        shells = [pack_shells(b, n) for b, n in systems]
        l, nprim, poff, bfo, ctr, ctr1, atom, exps, coefs, scale = ([] for _ in range(10))
        sys_shell, sys_atom, sizes = [0], [0], []
        nbf = npr = nat = 0
        nuclei = []
        for sh, (b, nuc) in zip(shells, systems):
            if np.any(sh.atom < 0):
                raise ValueError('every shell of a batched system must sit on one of its nuclei')
            l.append(sh.l)
            nprim.append(sh.nprim)
            poff.append(sh.prim_off + npr)
            bfo.append(sh.bf_off + nbf)
            ctr.append(sh.ctr)
            ctr1.append(sh.ctr1)
            atom.append(sh.atom + nat)
            exps.append(sh.exps)
            coefs.append(sh.coefs)
            scale.append(sh.bf_scale)
            nbf += sh.nbf
            npr += len(sh.exps)
            nat += len(nuc)
            nuclei += list(nuc)
            sys_shell.append(sys_shell[-1] + sh.nshell)
            sys_atom.append(nat)
            sizes.append(sh.nbf)
        cat = lambda xs, t: np.ascontiguousarray(np.concatenate(xs), dtype=t)
        allsh = ShellSet(cat(l, np.int32), cat(nprim, np.int32), cat(poff, np.int32), cat(bfo, np.int32),
                         cat(ctr, np.float64), cat(ctr1, np.float64), cat(atom, np.int32),
                         cat(exps, np.float64), cat(coefs, np.float64), cat(scale, np.float64))
        ss = np.ascontiguousarray(sys_shell, dtype=np.int32)
        sa = np.ascontiguousarray(sys_atom, dtype=np.int32)
        self._bind_stream()
        _lib.check(self.lib.gx_set_basis_batch(
            self.ctx, len(systems), _p(ss), _p(sa), allsh.nshell, _p(allsh.l), _p(allsh.nprim),
            _p(allsh.prim_off), _p(allsh.bf_off), _p(allsh.ctr), _p(allsh.ctr1), _p(allsh.atom), len(allsh.exps),
            _p(allsh.exps), _p(allsh.coefs), allsh.nbf, _p(allsh.bf_scale)), 'gx_set_basis_batch')
        self.shells = allsh
        self.sizes = sizes
        self.offsets = np.concatenate([[0], np.cumsum([n * n for n in sizes])]).astype(np.int64)
        self.set_nuclei(nuclei)
        return allsh

    def pack(self, mats):
        """List of [n_k, n_k] matrices -> one packed device vector."""
        return torch.cat([self._dev(m).reshape(-1) for m in mats])

    def unpack(self, flat):
        """Packed vector -> list of [n_k, n_k] views."""
        return [flat[self.offsets[k]:self.offsets[k + 1]].view(n, n) for k, n in enumerate(self.sizes)]

    @property
    def packed_size(self):
        return int(self.offsets[-1])

    def set_nuclei(self, nuclei):
        xyz = np.ascontiguousarray([n.coordinates for n in nuclei], dtype=np.float64)
        z = np.ascontiguousarray([float(n.charge) for n in nuclei], dtype=np.float64)
        _lib.check(self.lib.gx_set_nuclei(self.ctx, len(nuclei), _p(xyz), _p(z)), 'gx_set_nuclei')
        self.nuclei = nuclei

    @property
    def nbf(self):
        return self.shells.nbf

    def _matrix_shape(self, M, what):
        """(n, n) for one system, the packed length for a batch."""
        if len(self.sizes) == 1:
            n = self.nbf
            if M.shape != (n, n):
                raise ValueError('%s has shape %s, expected (%d, %d)' % (what, tuple(M.shape), n, n))
            return (n, n)
        if M.shape != (self.packed_size,):
            raise ValueError('%s must be the packed vector of %d elements' % (what, self.packed_size))
        return (self.packed_size,)

    def _empty(self, *shape):
        return torch.empty(*shape, dtype=torch.float64, device=self.device)

    def _dev(self, x):
        t = torch.as_tensor(x, dtype=torch.float64, device=self.device)
        return t.contiguous()

    # -- integrals -------------------------------------------------------------
    def schwarz(self):
        n = self.lib.gx_npair(self.ctx)
        q = np.zeros(n)
        a = np.zeros(n, dtype=np.int32)
        b = np.zeros(n, dtype=np.int32)
        _lib.check(self.lib.gx_schwarz(self.ctx, _p(q), _p(a), _p(b)), 'gx_schwarz')
        return q, a, b

    def eri_quartets(self, keys):
        """Values of basis-function quartets ``keys`` (int array [n, 4]) as a numpy array."""
        k = np.ascontiguousarray(np.asarray(keys, dtype=np.int32).reshape(-1, 4))
        out = np.zeros(len(k))
        self._bind_stream()
        _lib.check(self.lib.gx_eri_quartets(self.ctx, len(k), _p(k), _p(out)), 'gx_eri_quartets')
        return out

    def eri_dense(self):
        """Dense (ij|kl) tensor [N,N,N,N] on the device."""
        n = self.nbf
        out = self._empty(n, n, n, n)
        self._bind_stream()
        _lib.check(self.lib.gx_eri_dense(self.ctx, _p(out)), 'gx_eri_dense')
        return out

    def jk(self, P, cutoff=0.0, fp32_bound=0.0, mixed=False, out=None, timing=False):
        """Direct (J, K) from density ``P`` (FP64 device tensors).

        ``timing=True`` records HIP events around each class-combination
        kernel on the stream it runs on; read them with ``jk_launch_times``.
        """
        P = self._dev(P)
        shape = self._matrix_shape(P, 'density')
        J, K = (self._empty(*shape), self._empty(*shape)) if out is None else out
        self._bind_stream()
        flags = (_lib.GX_JK_MIXED if mixed else 0) | (_lib.GX_JK_TIMING if timing else 0)
        _lib.check(self.lib.gx_jk(self.ctx, _p(P), _p(J), _p(K), float(cutoff), float(fp32_bound), flags),
                   'gx_jk')
        return J, K

    def jk_uhf(self, Pa, Pb, cutoff=0.0):
        """Unrestricted direct build (gx_jk_uhf, FP64): (J of Pa + Pb, K of Pa,
        K of Pb) from one evaluation of the integrals."""
        Pa, Pb = self._dev(Pa), self._dev(Pb)
        shape = self._matrix_shape(Pa, 'alpha density')
        self._matrix_shape(Pb, 'beta density')
        J, Ka, Kb = self._empty(*shape), self._empty(*shape), self._empty(*shape)
        self._bind_stream()
        _lib.check(self.lib.gx_jk_uhf(self.ctx, _p(Pa), _p(Pb), _p(J), _p(Ka), _p(Kb), float(cutoff), 0),
                   'gx_jk_uhf')
        return J, Ka, Kb

    def autotune(self, P, cutoff=0.0, fp32_bound=0.0, mixed=False,
                 candidates=(1.0, 2.0, 4.0, 8.0, 16.0, 32.0, 64.0), reps=3):
        """Pick the work-item size of each class combination by timing (see
        gx_jk_autotune); returns {(la, lb, lc, ld): (items per workgroup, ms)}.
        In mixed mode a negative value marks a combination that runs all in
        FP64 because that was faster."""
        P = self._dev(P)
        self._matrix_shape(P, 'density')
        cand = np.ascontiguousarray(candidates, dtype=np.float64)
        out = np.zeros((64, 6))
        self._bind_stream()
        n = self.lib.gx_jk_autotune(self.ctx, _p(P), float(cutoff), float(fp32_bound),
                                    _lib.GX_JK_MIXED if mixed else 0, _p(cand), len(cand), int(reps), _p(out), 64)
        if n < 0:
            _lib.check(n, 'gx_jk_autotune')
        return {tuple(int(x) for x in r[:4]): (float(r[4]), float(r[5])) for r in out[:n]}

    def grad_autotune(self, P, W, cutoff=0.0, candidates=(1.0, 2.0, 4.0, 8.0, 16.0, 32.0, 64.0), reps=2):
        """Pick the gradient work-item size of each class combination by timing
        (gx_grad_autotune); returns {(la, lb, lc, ld): (items per workgroup, ms)}."""
        P, W = self._dev(P), self._dev(W)
        self._matrix_shape(P, 'P')
        cand = np.ascontiguousarray(candidates, dtype=np.float64)
        out = np.zeros((64, 6))
        self._bind_stream()
        n = self.lib.gx_grad_autotune(self.ctx, _p(P), _p(W), float(cutoff), _p(cand), len(cand), int(reps),
                                      _p(out), 64)
        if n < 0:
            _lib.check(n, 'gx_grad_autotune')
        return {tuple(int(x) for x in r[:4]): (float(r[4]), float(r[5])) for r in out[:n]}

    def jk_class_counts(self, cutoff=0.0, fp32_bound=0.0, mixed=False):
        """Per class combination: (la, lb, lc, ld, shell quartets FP64, FP32,
        primitive quartets FP64, FP32, bf quartets FP64, FP32) that gx_jk
        evaluates with these thresholds (gx_jk_class_counts)."""
        out = np.zeros((64, 10), dtype=np.int64)
        n = self.lib.gx_jk_class_counts(self.ctx, float(cutoff), float(fp32_bound),
                                        _lib.GX_JK_MIXED if mixed else 0, _p(out), 64)
        if n < 0:
            _lib.check(n, 'gx_jk_class_counts')
        return [tuple(int(x) for x in r) for r in out[:n]]

    def jk_launch_times(self):
        """{(la, lb, lc, ld): (total ms, launches)} over the ``jk(..., timing=True)``
        calls since the previous read."""
        out = np.zeros((64, 6))
        n = self.lib.gx_jk_launch_times(self.ctx, _p(out), 64)
        if n < 0:
            _lib.check(n, 'gx_jk_launch_times')
        return {tuple(int(x) for x in r[:4]): (float(r[4]), int(r[5])) for r in out[:n]}

    def grad(self, P, W, cutoff=0.0, fp32_bound=0.0, mixed=False, timing=False):
        """Nuclear gradient dE/dR [natom, 3] of the closed-shell RHF energy from
        the density ``P`` and the energy-weighted density ``W`` (device tensors).
        ``mixed``: quartets with Schwarz product below ``fp32_bound`` run their
        primitive loops in FP32 (gx_grad_ex); ``timing`` records HIP events
        around each two-electron kernel (``grad_launch_times``)."""
        if self.nuclei is None:
            raise _lib.GxError('grad needs nuclei (set_nuclei)')
        P, W = self._dev(P), self._dev(W)
        self._matrix_shape(P, 'P')
        self._matrix_shape(W, 'W')
        g = self._empty(len(self.nuclei), 3)
        self._bind_stream()
        flags = (_lib.GX_JK_MIXED if mixed else 0) | (_lib.GX_JK_TIMING if timing else 0)
        _lib.check(self.lib.gx_grad_ex(self.ctx, _p(P), _p(W), _p(g), float(cutoff), float(fp32_bound), flags),
                   'gx_grad_ex')
        return g

    def grad_launch_times(self):
        """{(la, lb, lc, ld): (total ms, launches)} of the two-electron gradient
        kernels over the ``grad(..., timing=True)`` calls since the previous read."""
        out = np.zeros((64, 6))
        n = self.lib.gx_grad_launch_times(self.ctx, _p(out), 64)
        if n < 0:
            _lib.check(n, 'gx_grad_launch_times')
        return {tuple(int(x) for x in r[:4]): (float(r[4]), int(r[5])) for r in out[:n]}

    def jk_stats(self, cutoff=0.0, fp32_bound=0.0, mixed=False):
        st = _lib.GxStats()
        _lib.check(self.lib.gx_jk_stats(self.ctx, float(cutoff), float(fp32_bound),
                                        _lib.GX_JK_MIXED if mixed else 0, ctypes.byref(st)), 'gx_jk_stats')
        return {name: getattr(st, name) for name, _ in st._fields_}

    def one_electron(self):
        """(S, T, V) on the device; V needs nuclei."""
        shape = (self.nbf, self.nbf) if len(self.sizes) == 1 else (self.packed_size,)
        S, T = self._empty(*shape), self._empty(*shape)
        V = self._empty(*shape) if self.nuclei is not None else None
        self._bind_stream()
        _lib.check(self.lib.gx_one_electron(self.ctx, _p(S), _p(T), _p(V) if V is not None else None),
                   'gx_one_electron')
        return S, T, V
