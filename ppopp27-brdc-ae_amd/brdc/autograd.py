"""Differentiable eigenvalues of symmetric tridiagonals (torch.autograd).

``eigvalsh_tridiagonal(d, e)`` returns the ascending eigenvalues of
tridiag(e, d, e) through the BR solve; its backward is the hand-written
gfx950 kernel behind ``brdc_dstebr_grad_device`` (dL/dD_j = sum_i g_i
v_ij^2, dL/dE_j = 2 sum_i g_i v_ij v_i,j+1 with every eigenvector recomputed
by twisted factorization at its eigenvalue; nothing is stored by the forward
pass beyond D, E and the eigenvalues).  ``eigvalsh_tridiagonal_batched``
does the same for a batch's concatenated rows through one merge forest.
The reference solver is values-only and has no backward; this layer is new.

Repeated eigenvalues: eigenvalues within 2^-44 ||T|| of a neighbour (a
matrix split into identical blocks, exact multiplicities) get their vectors
from inverse iteration with Gram-Schmidt inside the cluster, so the cluster's
vectors are orthonormal; only losses that give a cluster's members equal
weight have a gradient there, and for them the result is exact to rounding.
Pairs farther apart than that but closer than about 1e-8 ||T|| whose vectors
overlap (Wilkinson's W21+ below its top pair) keep twisted-factorization
vectors, accurate to about eps ||T|| / gap.
"""
from __future__ import annotations

import ctypes

import torch

from . import BatchSolver, Solver, _lib


def _scratch(nbytes: int, device):
    buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
    if buf.data_ptr() % 256:
        raise _lib.BrdcError("scratch allocation is not 256-byte aligned")
    return buf


def grad_single(d, e, w, g, stream=None):
    """(dL/dD [n], dL/dE [n-1]) of one matrix; all arguments float64 device tensors."""
    L = _lib.load()
    n = d.numel()
    dd = torch.empty_like(d)
    de = torch.empty(max(n, 1), dtype=torch.float64, device=d.device)
    if n == 0:
        return dd, de[:0]
    nbytes = ctypes.c_size_t(0)
    _lib.check(L.brdc_dstebr_grad_workspace(n, ctypes.byref(nbytes)), "grad workspace")
    scratch = _scratch(nbytes.value, d.device)
    e_in = e if e.numel() >= max(n - 1, 1) else torch.zeros(1, dtype=torch.float64,
                                                             device=d.device)
    st = stream if stream is not None else torch.cuda.current_stream(d.device)
    _lib.check(L.brdc_dstebr_grad_device(n, d.data_ptr(), e_in.data_ptr(), w.data_ptr(),
                                         g.data_ptr(), dd.data_ptr(), de.data_ptr(),
                                         scratch.data_ptr(), scratch.numel(), st.cuda_stream),
               "brdc_dstebr_grad_device")
    return dd, de[: n - 1]


def grad_batched(solver: BatchSolver, d, e, w, g, stream=None):
    """(dL/dD, dL/dE) over a batch's concatenated rows (dE is 0 at each
    matrix's last row)."""
    L = _lib.load()
    dd = torch.empty_like(d)
    de = torch.empty_like(d)
    if solver.rows == 0:
        return dd, de
    nbytes = ctypes.c_size_t(0)
    _lib.check(L.brdc_batch_grad_workspace(solver._plan, ctypes.byref(nbytes)), "grad workspace")
    scratch = _scratch(nbytes.value, d.device)
    st = stream if stream is not None else torch.cuda.current_stream(d.device)
    _lib.check(L.brdc_batch_grad(solver._plan, d.data_ptr(), e.data_ptr(), w.data_ptr(),
                                 g.data_ptr(), dd.data_ptr(), de.data_ptr(), solver.ws.data_ptr(),
                                 solver.bytes, scratch.data_ptr(), scratch.numel(),
                                 st.cuda_stream),
               "brdc_batch_grad")
    return dd, de


def _check_info(info, what):
    bad = int(info.abs().sum().item())
    if bad:
        raise _lib.BrdcError(f"{what}: nonzero INFO from the eigenvalue solve")


class _Eigvals(torch.autograd.Function):
    @staticmethod
    def forward(ctx, d, e, leaf_max):
        d0 = d.detach().contiguous()
        e0 = e.detach().contiguous()
        solver = Solver(d0.numel(), device=d0.device, leaf_max=leaf_max)
        w = d0.clone()
        solver.solve(w, e0 if e0.numel() else d0.new_zeros(1))
        _check_info(solver.info, "eigvalsh_tridiagonal")
        ctx.save_for_backward(d0, e0, w)
        return w

    @staticmethod
    def backward(ctx, g):
        d0, e0, w = ctx.saved_tensors
        dd, de = grad_single(d0, e0, w, g.contiguous())
        return dd, de, None


class _EigvalsBatched(torch.autograd.Function):
    @staticmethod
    def forward(ctx, d, e, solver):
        d0 = d.detach().contiguous()
        e0 = e.detach().contiguous()
        w = d0.clone()
        solver.solve(w, e0)
        _check_info(solver.info, "eigvalsh_tridiagonal_batched")
        ctx.solver = solver
        ctx.save_for_backward(d0, e0, w)
        return w

    @staticmethod
    def backward(ctx, g):
        d0, e0, w = ctx.saved_tensors
        dd, de = grad_batched(ctx.solver, d0, e0, w, g.contiguous())
        return dd, de, None


def eigvalsh_tridiagonal(d, e, leaf_max: int = 0):
    """Ascending eigenvalues of tridiag(e, d, e) (d: n, e: n-1 float64 CUDA
    tensors), differentiable in d and e."""
    return _Eigvals.apply(d, e, leaf_max)


def eigvalsh_tridiagonal_batched(d, e, solver: BatchSolver):
    """Per-matrix ascending eigenvalues over a batch's concatenated rows (the
    layout of :func:`brdc.generators.pack_batch`), differentiable in d and e;
    ``solver`` fixes the batch's sizes."""
    return _EigvalsBatched.apply(d, e, solver)
