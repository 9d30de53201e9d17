"""Differentiable sparse solve for PyTorch: x = A^{-1} b by a Krylov method preconditioned by
a spaND factorization of A (PCG for SPD A, src/is.cpp:39-121; GMRES for a general A,
src/is.cpp:123-300), with gradients for b and for A's stored values. There is no
counterpart in the reference; the gradients are the adjoint ones:

  lambda = A^{-T} dL/dx            (the adjoint solve, reusing the factorization: PCG with
                                    M^{-1} for SPD A, GMRES on A^T with M^{-T} otherwise)
  dL/db  = lambda
  dL/dA_ij = -sum_r lambda_i,r x_j,r  on A's stored entries (spand_matrix_sampled_outer)

Values are given in the CSC order of a fixed pattern, so that a caller can build them from
its own parameters with torch operations (e.g. face coefficients of a finite-volume
operator, ``fv_values``) and let autograd carry the gradient further. The values never
leave the GPU: they go to the tree (spand_assemble_values_device) and to the Krylov
operator (spand_matrix_set_values_device) as device arrays.
"""
from __future__ import annotations

import numpy as np
import torch

from . import FP64, Matrix, Tree, cg_device, gmres_device
from .problems import CSC, DiffusionFV, linspace_nd, symmetric_graph


def check_values(vals: torch.Tensor, nnz: int) -> torch.Tensor:
    """The stored values as a contiguous float64 CUDA vector of nnz entries (ValueError
    otherwise: the library reads them as device memory)."""
    if not isinstance(vals, torch.Tensor) or not vals.is_cuda:
        raise ValueError("values must be a CUDA tensor")
    if vals.dtype != torch.float64:
        raise ValueError("values must be float64")
    if vals.numel() != nnz:
        raise ValueError(f"values: expected {nnz} entries, got {vals.numel()}")
    return vals.detach().contiguous().reshape(-1)


def check_rhs(b: torch.Tensor, N: int) -> torch.Tensor:
    """b as an (N, nrhs) float64 CUDA tensor, 1 <= nrhs <= 64 (a vector (N,) becomes (N, 1));
    ValueError otherwise."""
    if not isinstance(b, torch.Tensor) or not b.is_cuda:
        raise ValueError("right-hand side must be a CUDA tensor")
    if b.dtype != torch.float64:
        raise ValueError("right-hand side must be float64")
    if b.dim() == 1:
        b = b.reshape(-1, 1)
    if b.dim() != 2 or b.shape[0] != N:
        raise ValueError(f"right-hand side must have shape ({N},) or ({N}, nrhs)")
    if not 1 <= b.shape[1] <= 64:
        raise ValueError("right-hand side: 1 to 64 columns")
    return b


class SparseSolver:
    """Factorization and operator state for one sparsity pattern. ``update(vals)`` re-assembles
    and factorizes A for new values (the partition is computed once); the solves use the last
    factorization. ``version`` counts the updates, so that a backward pass can tell whether
    the factorization still belongs to its forward pass.

    symm: "SPD" (LLT factor, PCG) or "GEN" (PLU factor, GMRES(restart))."""

    def __init__(self, colptr, rowval, X, nlevels: int, tol: float = 1e-2, precision=FP64, symm: str = "SPD",
                 krylov_tol: float = 1e-12, maxit: int = 500, restart: int = 30, pcg_tol: float | None = None):
        if symm not in ("SPD", "GEN"):
            raise ValueError("symm must be 'SPD' or 'GEN'")
        self.colptr = np.ascontiguousarray(colptr, dtype=np.int32)
        self.rowval = np.ascontiguousarray(rowval, dtype=np.int32)
        self.N = len(self.colptr) - 1
        self.nnz = int(self.colptr[-1])
        self.symm = symm
        self.krylov_tol = krylov_tol if pcg_tol is None else pcg_tol
        self.maxit, self.restart = maxit, restart
        self.tree = Tree(nlevels)
        self.tree.set_use_geo(True)
        self.tree.set_Xcoo(X)
        self.tree.set_tol(tol)
        self.tree.set_precision(precision)
        if symm == "GEN":
            self.tree.set_symm_kind("GEN")
        ones = CSC(self.N, self.colptr, self.rowval, np.ones(self.nnz))
        self.tree.partition(symmetric_graph(ones))
        # the pattern is assembled once from the host; later values arrive on the device
        self.tree.assemble(ones)
        self.matrix = Matrix(ones)
        self.version = 0
        self.iters = []

    def _check_vals(self, vals: torch.Tensor) -> torch.Tensor:
        return check_values(vals, self.nnz)

    def _check_rhs(self, b: torch.Tensor) -> torch.Tensor:
        return check_rhs(b, self.N)

    def update(self, vals: torch.Tensor):
        """Re-assembles and factorizes A for new stored values (device tensor, CSC order)."""
        v = self._check_vals(vals)
        torch.cuda.synchronize(v.device)  # the library's calls are ordered on the null stream
        self.tree.assemble_values_device(v.data_ptr(), self.nnz)
        self.matrix.set_values_device(v.data_ptr(), self.nnz)
        self.tree.factorize()
        self.version += 1

    def _krylov(self, b: torch.Tensor, transpose: bool) -> torch.Tensor:
        squeeze = b.dim() == 1
        b = self._check_rhs(b)
        bt = b.detach().t().contiguous()  # column-major N x nrhs, leading dimension N
        xt = torch.zeros_like(bt)
        torch.cuda.synchronize(b.device)
        if self.symm == "SPD":
            its, info = cg_device(self.matrix, self.tree, bt.data_ptr(), xt.data_ptr(), bt.shape[0], self.N,
                                  iters=self.maxit, tol=self.krylov_tol)
        else:
            its, info = gmres_device(self.matrix, self.tree, bt.data_ptr(), xt.data_ptr(), bt.shape[0], self.N,
                                     iters=self.maxit, restart=self.restart, tol=self.krylov_tol,
                                     transpose=transpose)
        if info["status"] != 0:
            raise RuntimeError(f"Krylov solve did not converge in {self.maxit} iterations: {its}")
        self.iters.append(its.tolist())
        x = xt.t()
        return x.reshape(-1) if squeeze else x

    def solve(self, b: torch.Tensor) -> torch.Tensor:
        """A^{-1} b for b of shape (N,) or (N, nrhs), float64 on the GPU."""
        return self._krylov(b, transpose=False)

    def solve_adjoint(self, g: torch.Tensor) -> torch.Tensor:
        """A^{-T} g with the same factorization (A^T = A for SPD)."""
        return self._krylov(g, transpose=True)

    # kept for callers of the SPD-only interface
    def pcg(self, b: torch.Tensor) -> torch.Tensor:
        return self.solve(b)

    def sampled_outer(self, u: torch.Tensor, v: torch.Tensor, alpha: float) -> torch.Tensor:
        u, v = self._check_rhs(u), self._check_rhs(v)
        if u.shape != v.shape:
            raise ValueError("sampled_outer: u and v differ in shape")
        ut, vt = u.detach().t().contiguous(), v.detach().t().contiguous()
        out = torch.zeros(self.nnz, dtype=torch.float64, device=u.device)
        torch.cuda.synchronize(u.device)
        self.matrix.sampled_outer_device(ut.data_ptr(), vt.data_ptr(), ut.shape[0], self.N, alpha, out.data_ptr())
        return out


SpdSolver = SparseSolver


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, b, solver):
        solver.update(vals)
        x = solver.solve(b)
        ctx.solver = solver
        ctx.version = solver.version
        ctx.save_for_backward(vals.detach(), x)
        return x

    @staticmethod
    def backward(ctx, gx):
        vals, x = ctx.saved_tensors
        solver = ctx.solver
        if solver.version != ctx.version:
            # another solve on this solver came in between: restore this forward's factorization
            solver.update(vals)
            ctx.version = solver.version
        lam = solver.solve_adjoint(gx.contiguous())
        x2, lam2 = (x.reshape(-1, 1), lam.reshape(-1, 1)) if x.dim() == 1 else (x, lam)
        gvals = solver.sampled_outer(lam2, x2, -1.0)
        return gvals, lam, None


def sparse_solve(vals: torch.Tensor, b: torch.Tensor, solver: SparseSolver) -> torch.Tensor:
    """x = A(vals)^{-1} b, differentiable in vals (stored entries, CSC order) and b ((N,) or
    (N, nrhs))."""
    return _Solve.apply(vals, b, solver)


spd_solve = sparse_solve


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


def fv_solver(fv: DiffusionFV, nlevels: int, **kw) -> SparseSolver:
    """A solver on ``fv``'s pattern with grid coordinates for the geometric partition."""
    return SparseSolver(fv.colptr, fv.rowval, linspace_nd(fv.n, fv.d), nlevels, **kw)
