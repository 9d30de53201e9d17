"""Differentiable sparse solve for PyTorch: x = A^{-1} b with A SPD, by PCG (src/is.cpp:39-121)
preconditioned by a spaND factorization of A, with gradients for b and for A's stored
values. There is no counterpart in the reference; the gradients are the adjoint ones:

  lambda = A^{-T} dL/dx            (PCG with the same factorization; A^T = A)
  dL/db  = lambda
  dL/dA_ij = -sum_r lambda_i,r x_j,r  on A's stored entries (spand_matrix_sampled_outer)

Values are given in the CSC order of a fixed pattern, so that a caller can build them from
its own parameters with torch operations (e.g. face coefficients of a finite-volume
operator, ``fv_values``) and let autograd carry the gradient further.
"""
from __future__ import annotations

import numpy as np
import torch

from . import FP32, FP64, Matrix, Tree, cg_device
from .problems import CSC, DiffusionFV, linspace_nd, symmetric_graph


class SpdSolver:
    """Factorization and operator state for one SPD pattern. ``update(vals)`` assembles and
    factorizes A for new values (the partition is computed once); the solves below use the
    last factorization."""

    def __init__(self, colptr, rowval, X, nlevels: int, tol: float = 1e-2, precision=FP64, pcg_tol: float = 1e-12,
                 maxit: int = 500):
        self.colptr = np.ascontiguousarray(colptr, dtype=np.int32)
        self.rowval = np.ascontiguousarray(rowval, dtype=np.int32)
        self.N = len(self.colptr) - 1
        self.pcg_tol, self.maxit = pcg_tol, maxit
        self.tree = Tree(nlevels)
        self.tree.set_use_geo(True)
        self.tree.set_Xcoo(X)
        self.tree.set_tol(tol)
        self.tree.set_precision(precision)
        ones = CSC(self.N, self.colptr, self.rowval, np.ones(len(self.rowval)))
        self.tree.partition(symmetric_graph(ones))
        self.matrix = None
        self.iters = []

    def update(self, vals: np.ndarray):
        A = CSC(self.N, self.colptr, self.rowval, np.asarray(vals, dtype=np.float64))
        self.tree.assemble(A)
        self.tree.factorize()
        self.matrix = Matrix(A)

    def pcg(self, b: torch.Tensor) -> torch.Tensor:
        """A^{-1} b for b of shape (N, nrhs), float64 on the GPU."""
        bt = b.detach().t().contiguous()  # column-major N x nrhs, leading dimension N
        xt = torch.zeros_like(bt)
        its, info = cg_device(self.matrix, self.tree, bt.data_ptr(), xt.data_ptr(), bt.shape[0], self.N,
                              iters=self.maxit, tol=self.pcg_tol)
        if info["status"] != 0:
            raise RuntimeError(f"PCG did not converge in {self.maxit} iterations: {its}")
        self.iters.append(its.tolist())
        return xt.t()

    def sampled_outer(self, u: torch.Tensor, v: torch.Tensor, alpha: float) -> torch.Tensor:
        ut, vt = u.detach().t().contiguous(), v.detach().t().contiguous()
        out = torch.zeros(self.matrix.nnz, dtype=torch.float64, device=u.device)
        self.matrix.sampled_outer_device(ut.data_ptr(), vt.data_ptr(), ut.shape[0], self.N, alpha, out.data_ptr())
        return out


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, b, solver):
        solver.update(vals.detach().cpu().numpy())
        x = solver.pcg(b)
        ctx.solver = solver
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, gx):
        (x,) = ctx.saved_tensors
        solver = ctx.solver
        lam = solver.pcg(gx.contiguous())  # A^{-T} gx, A symmetric
        gvals = solver.sampled_outer(lam, x, -1.0)
        return gvals, lam, None


def spd_solve(vals: torch.Tensor, b: torch.Tensor, solver: SpdSolver) -> torch.Tensor:
    """x = A(vals)^{-1} b, differentiable in vals (stored entries, CSC order) and b (N, nrhs)."""
    return _Solve.apply(vals, b, solver)


def fv_values(fv: DiffusionFV, kappa: torch.Tensor) -> torch.Tensor:
    """Stored entries of ``fv``'s matrix for cell-wise kappa, as torch operations (so that
    gradients reach kappa): harmonic-mean interior faces, kappa_a on boundary faces."""
    dev = kappa.device
    fa = torch.as_tensor(fv.face_a, device=dev)
    fb = torch.as_tensor(fv.face_b, device=dev)
    bc = torch.as_tensor(fv.bnd_cell, device=dev)
    ka, kb = kappa[fa], kappa[fb]
    t = torch.cat([2.0 * ka * kb / (ka + kb), kappa[bc]])
    ek = torch.as_tensor(fv.entry_index, device=dev)
    ef = torch.as_tensor(fv.entry_face, device=dev)
    es = torch.as_tensor(fv.entry_sign, device=dev, dtype=kappa.dtype)
    return torch.zeros(fv.nnz, dtype=kappa.dtype, device=dev).index_add(0, ek, es * t[ef])


def fv_solver(fv: DiffusionFV, nlevels: int, **kw) -> SpdSolver:
    """An SpdSolver on ``fv``'s pattern with grid coordinates for the geometric partition."""
    return SpdSolver(fv.colptr, fv.rowval, linspace_nd(fv.n, fv.d), nlevels, **kw)
