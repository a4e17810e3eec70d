"""Gradients with respect to the initial field over the C ABI
(gt_adj_forward / gt_adj_backward of include/gturbo.h), for initial-
condition inversion.

The reference has no adjoint.  An :class:`Adjoint` differentiates the run of
src/main.f90:89-136 -- fft_dir and trunc of the input frame, ``nsteps`` x
(step or step4, forcing), fft_inv -- with the library's hand-written
transposed passes and checkpoint-and-recompute every ``kchk`` steps.
:func:`run` wraps it as a ``torch.autograd.Function`` on device tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check
from .params import Params
from .solver import Solver, _addr


class Adjoint:
    def __init__(self, params: Params, nsteps: int, kchk: int = 20):
        self.params = params
        self.n = int(params.n)
        self.nsteps = int(nsteps)
        self.kchk = int(kchk)
        self.solver = Solver(params)
        self._bufs = [self.solver.field() for _ in range(4)]  # w0, wT, gT, g0
        for a in self._bufs:
            self.solver.enter_data(a, copyin=False)

    def close(self):
        if self.solver is not None:
            self.solver.close()
            self.solver = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _pad(self, buf, w):
        w = np.asarray(w, np.float64)
        buf[:] = 0.0
        buf[:, : w.shape[1]] = w
        self.solver.update_device(buf)

    def forward(self, w0: np.ndarray, seed: int = 481929):
        """Final field (N, N) of the run from the physical field w0 and the
        forcing seed after it; keeps the checkpoints for backward."""
        b0, bT = self._bufs[0], self._bufs[1]
        self._pad(b0, w0)
        s = C.c_int32(int(seed))
        check(self.solver.L.gt_adj_forward(self.solver.h, _addr(b0), _addr(bT), self.nsteps,
                                           self.kchk, C.byref(s)))
        self.solver.update_host(bT)
        return bT[:, : self.n].copy(), int(s.value)

    def backward(self, gT: np.ndarray) -> np.ndarray:
        """dL/dw0 (N, N) from dL/dwT for the run of the last forward."""
        bg, b0 = self._bufs[2], self._bufs[3]
        self._pad(bg, gT)
        check(self.solver.L.gt_adj_backward(self.solver.h, _addr(bg), _addr(b0)))
        self.solver.update_host(b0)
        return b0[:, : self.n].copy()

    # device tensors (torch): the tensors are attached as the device copies
    def forward_device(self, w0_pad, wT_pad, seed: int) -> int:
        sv = self.solver
        for t in (w0_pad, wT_pad):
            sv.map_data(t)
        try:
            s = C.c_int32(int(seed))
            check(sv.L.gt_adj_forward(sv.h, w0_pad.data_ptr(), wT_pad.data_ptr(), self.nsteps,
                                      self.kchk, C.byref(s)))
        finally:
            for t in (w0_pad, wT_pad):
                sv.unmap_data(t)
        return int(s.value)

    def backward_device(self, gT_pad, g0_pad):
        sv = self.solver
        for t in (gT_pad, g0_pad):
            sv.map_data(t)
        try:
            check(sv.L.gt_adj_backward(sv.h, gT_pad.data_ptr(), g0_pad.data_ptr()))
        finally:
            for t in (gT_pad, g0_pad):
                sv.unmap_data(t)


def run(w0, adj: Adjoint, seed: int = 481929):
    """torch.autograd-aware final field of the run from w0, an (N, N)
    float64 tensor.  Backward goes through the library's adjoint
    (gt_adj_backward); one Adjoint serves one pending backward.

    A GPU tensor is handed to the library in place (gt_map_data); that needs
    torch imported before gturbo_amd loaded its library, so that both use
    torch's HIP runtime.  A CPU tensor goes through host copies."""
    import torch
    from . import _lib

    n = adj.n
    if tuple(w0.shape) != (n, n) or w0.dtype != torch.float64:
        raise ValueError(f"w0 must be a ({n}, {n}) float64 tensor")
    if w0.is_cuda and not _lib.TORCH_FIRST:
        raise _lib.GturboError(
            "GPU tensors need torch imported before gturbo_amd's library is loaded "
            "(one HIP runtime per process); import torch first, or pass CPU tensors")

    class _Run(torch.autograd.Function):
        @staticmethod
        def forward(ctx, w):
            if not w.is_cuda:
                wT, _ = adj.forward(w.detach().numpy(), seed)
                return torch.from_numpy(wT)
            b0 = torch.zeros((n, n + 2), dtype=torch.float64, device=w.device)
            b0[:, :n] = w
            bT = torch.empty_like(b0)
            torch.cuda.synchronize(w.device)  # the library runs on its own stream
            adj.forward_device(b0, bT, seed)
            return bT[:, :n].clone()

        @staticmethod
        def backward(ctx, g):
            if not g.is_cuda:
                return torch.from_numpy(adj.backward(g.detach().numpy()))
            bg = torch.zeros((n, n + 2), dtype=torch.float64, device=g.device)
            bg[:, :n] = g
            b0 = torch.empty_like(bg)
            torch.cuda.synchronize(g.device)
            adj.backward_device(bg, b0)
            return b0[:, :n].clone()

    return _Run.apply(w0)
