"""Differentiable sparse solve x = A(theta)^{-1} b on the GPU.

``SparseSolve`` is a ``torch.autograd.Function``:
  forward   x = A^{-1} b with the library's FP64 Krylov solver (Cg for a
            symmetric positive definite A, Gmres otherwise);
  backward  the adjoint solve A^T lam = dL/dx, then
            dL/db = lam and dL/dA_ij = -lam_i x_j on the stored pattern
            (shk_sddmm, a fused SDDMM-style kernel).
For RowSymSparse storage a stored upper value a_ij (i != j) stands for both
a_ij and a_ji, so its gradient is -(lam_i x_j + lam_j x_i).

``DiffusionOperator`` assembles the 7-point matrix of -div(exp(theta) grad u)
on an n^3 grid (harmonic-mean face coefficients, Dirichlet boundary faces use
the cell value -- the layout of seldon_amd.workloads.diffusion3d) from theta
with torch ops, so dL/dtheta follows from dL/dA by autograd.
"""
from __future__ import annotations

import numpy as np
import torch

from . import krylov as K
from . import workloads as wl


class SparseSolve(torch.autograd.Function):
    """x = A(values)^{-1} b.  A and AT are device handles whose values are
    overwritten by every solve on them, so backward sets them again from the
    saved ``values`` (and ``perm_T``, AT's values as a permutation of A's)
    before the adjoint solve: several solves on one operator before one
    backward each get their own matrix."""

    @staticmethod
    def forward(ctx, values, b, A: "K.Matrix", AT: "K.Matrix | None", spd: bool, tol: float, max_iter: int,
                restart: int, perm_T: "torch.Tensor | None" = None):
        A.set_values(values.detach().contiguous())
        x = torch.zeros_like(b)
        it = K.Iteration(max_iter, tol)
        it.HideMessages()
        if spd:
            code = K.Cg(A, x, b.detach().contiguous(), K.JacobiPreconditioner(), it)
        else:
            it.SetRestart(restart)
            code = K.Gmres(A, x, b.detach().contiguous(), K.JacobiPreconditioner(), it)
        if code not in (0,):
            raise RuntimeError(f"forward solve did not converge (code {code}, {it.nb_iter} iterations)")
        ctx.save_for_backward(values, x)
        ctx.A, ctx.AT, ctx.spd, ctx.tol, ctx.max_iter, ctx.restart = A, AT, spd, tol, max_iter, restart
        ctx.perm_T = perm_T
        ctx.forward_iters = it.nb_iter
        return x

    @staticmethod
    def backward(ctx, gx):
        values, x = ctx.saved_tensors
        A = ctx.A
        lam = torch.zeros_like(x)
        it = K.Iteration(ctx.max_iter, ctx.tol)
        it.HideMessages()
        gx = gx.contiguous()
        if ctx.spd:
            A.set_values(values.detach().contiguous())
            code = K.Cg(A, lam, gx, K.JacobiPreconditioner(), it)
        else:
            AT = ctx.AT
            AT.set_values(values.detach()[ctx.perm_T].contiguous())
            it.SetRestart(ctx.restart)
            code = K.Gmres(AT, lam, gx, K.JacobiPreconditioner(), it)
        if code != 0:
            raise RuntimeError(f"adjoint solve did not converge (code {code})")
        gvals = torch.empty_like(values)
        from . import _lib as L
        L.check(L.load().shk_sddmm(A.ctx.h, A.h, K._ptr(lam), K._ptr(x), K._ptr(gvals)))
        return gvals, lam, None, None, None, None, None, None, None


class SparseOperator:
    """A fixed sparsity pattern whose values are a differentiable tensor."""

    def __init__(self, pattern: "wl.CSR", spd: bool, storage: str = "row", ctx=None):
        self.pattern = pattern
        self.spd = spd
        self.A = K.Matrix(pattern, storage=storage, dtype=torch.float64, ctx=ctx)
        self.AT = None
        self.perm_T = None
        if not spd:
            # A^T on its own handle; its values are a permutation of A's
            S = pattern.to_scipy().tocsr()
            idx = S.copy()
            idx.data = np.arange(S.nnz, dtype=np.float64)
            T = idx.T.tocsr()
            T.sort_indices()
            self.perm_T = torch.tensor(T.data.astype(np.int64), device="cuda")
            PT = wl.CSR(T.shape[0], T.shape[1], T.indptr.astype(np.int32), T.indices.astype(np.int32),
                        pattern.val[T.data.astype(np.int64)].copy())
            self.AT = K.Matrix(PT, storage="row", dtype=torch.float64, ctx=ctx)

    def solve(self, values, b, tol=1e-12, max_iter=20000, restart=50):
        return SparseSolve.apply(values, b, self.A, self.AT, self.spd, tol, max_iter, restart, self.perm_T)


class DiffusionOperator(SparseOperator):
    """-div(exp(theta) grad u) on an n^3 grid, 7 points, SPD."""

    def __init__(self, n: int, ctx=None):
        self.n = n
        N = n ** 3
        pattern = wl.diffusion3d(n, np.zeros(N))
        super().__init__(pattern, spd=True, storage="row", ctx=ctx)
        ptr, ind = pattern.ptr.astype(np.int64), pattern.ind.astype(np.int64)
        rows = np.repeat(np.arange(N), np.diff(ptr))
        # face table in the generator's order: for each cell the six faces
        # -z,-y,-x,+x,+y,+z; neighbour index or -1 at the boundary
        r = np.arange(N)
        ix, iy, iz = r % n, (r // n) % n, r // (n * n)
        nb = np.full((N, 6), -1, np.int64)
        nb[:, 0] = np.where(iz > 0, r - n * n, -1)
        nb[:, 1] = np.where(iy > 0, r - n, -1)
        nb[:, 2] = np.where(ix > 0, r - 1, -1)
        nb[:, 3] = np.where(ix < n - 1, r + 1, -1)
        nb[:, 4] = np.where(iy < n - 1, r + n, -1)
        nb[:, 5] = np.where(iz < n - 1, r + n * n, -1)
        # CSR position of each (cell, face) off-diagonal and of each diagonal
        key = rows * N + ind
        order = np.argsort(key)
        def pos(i, j):
            k = np.searchsorted(key[order], i * N + j)
            return order[k]
        self.nb = torch.tensor(nb, device="cuda")
        offpos = np.full((N, 6), -1, np.int64)
        for f in range(6):
            m = nb[:, f] >= 0
            offpos[m, f] = pos(r[m], nb[m, f])
        self.offpos = torch.tensor(offpos, device="cuda")
        self.diagpos = torch.tensor(pos(r, r), device="cuda")
        self.nnz = pattern.nnz

    def values(self, theta: torch.Tensor) -> torch.Tensor:
        """CSR values of A(theta) (same arithmetic as the C generator)."""
        kc = torch.exp(theta)
        nbc = self.nb.clamp(min=0)
        kn = kc[nbc]                               # [N, 6]
        kf = 2.0 * kc[:, None] * kn / (kc[:, None] + kn)
        kf = torch.where(self.nb >= 0, kf, kc[:, None].expand_as(kf))
        diag = kf[:, 0]
        for f in range(1, 6):
            diag = diag + kf[:, f]
        vals = torch.zeros(self.nnz, dtype=theta.dtype, device=theta.device)
        m = self.nb >= 0
        vals = vals.index_put((self.offpos[m],), -kf[m])
        vals = vals.index_put((self.diagpos,), diag)
        return vals
