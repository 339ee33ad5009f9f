"""Seeded synthetic workloads for the sparse Krylov path.

Thin ctypes front-end over ``libshk_workloads.so`` (built from
``csrc/shk_workloads.c``), so the HIP path, the CPU oracle and the benchmark
all consume bit-identical matrices and vectors.

Matrices are returned as :class:`CSR` (0-based ``int32`` ``ptr``/``ind``,
``float64`` values), the layout of Seldon's ``Matrix<T,General,RowSparse>``
(``matrix_sparse/Matrix_Sparse.hxx:42-142``).

Named workloads (sizes are the full ones; every builder takes smaller sizes
for tests):

* ``laplace2d``     -- 2-D 5-point Laplacian, 256x256 -> CG, identity.
* ``convdiff3d``    -- 3-D 7-point upwind convection-diffusion, 128^3,
  v = (1, 0.5, 0.25), global Peclet |v| L / nu = 10 with L = 1 ->
  BiCGSTAB / GMRES(30), Jacobi.
* ``powerlaw``      -- unstructured-FEM-like rows, 4,194,304 rows, row length
  P(k) ~ k^-1.7452 on [4, 512] (mean ~27), columns within +-20,000 of the
  diagonal, N(0,1) off-diagonals, diagonal 1.1 x row abs-sum -> FP32 GMRES(50)
  with FP64 iterative refinement.
* ``diffusion3d``   -- -div(exp(theta) grad u) on 64^3, harmonic-mean faces,
  theta ~ N(0, 0.5^2) -> differentiable CG.
* ``aniso2d``       -- -eps u_xx - u_yy on n x n -> the batched CG systems.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libshk_workloads.so")
_lib = None

_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(
            f"{_LIB_PATH} is missing: build the package first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(_LIB_PATH)
    c_i64, c_u64, c_int, c_dbl = ctypes.c_int64, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    lib.shk_wl_uniform.argtypes = [c_i64, c_u64, c_dbl, c_dbl, _F64P]
    lib.shk_wl_normal.argtypes = [c_i64, c_u64, c_dbl, c_dbl, _F64P]
    lib.shk_wl_aniso2d_nnz.argtypes = [c_int, c_int]
    lib.shk_wl_aniso2d_nnz.restype = c_i64
    lib.shk_wl_aniso2d.argtypes = [c_int, c_int, c_dbl, _I32P, _I32P, _F64P]
    lib.shk_wl_stencil3d_nnz.argtypes = [c_int]
    lib.shk_wl_stencil3d_nnz.restype = c_i64
    lib.shk_wl_convdiff3d.argtypes = [c_int, c_dbl, c_dbl, c_dbl, c_dbl, _I32P, _I32P, _F64P]
    lib.shk_wl_diffusion3d.argtypes = [c_int, _F64P, _I32P, _I32P, _F64P]
    lib.shk_wl_powerlaw_ptr.argtypes = [c_i64, c_u64, c_int, c_int, c_dbl, c_int, _I32P]
    lib.shk_wl_powerlaw_ptr.restype = c_i64
    lib.shk_wl_powerlaw_fill.argtypes = [c_i64, c_u64, c_int, c_int, c_dbl, c_int, c_dbl,
                                         _I32P, _I32P, _F64P]
    lib.shk_wl_upper.argtypes = [c_int, _I32P, _I32P, _F64P, _I32P,
                                 ctypes.c_void_p, ctypes.c_void_p]
    lib.shk_wl_upper.restype = c_i64
    lib.shk_wl_csr_apply.argtypes = [c_i64, _I32P, _I32P, _F64P, _F64P, _F64P]
    _lib = lib
    return lib


@dataclass
class CSR:
    """0-based CSR (Seldon RowSparse layout).  ``sym`` marks an upper-triangle
    RowSymSparse matrix (diagonal included)."""
    m: int
    n: int
    ptr: np.ndarray
    ind: np.ndarray
    val: np.ndarray
    sym: bool = False

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1])

    def diagonal(self) -> np.ndarray:
        d = np.zeros(self.m)
        rows = np.repeat(np.arange(self.m, dtype=np.int64), np.diff(self.ptr))
        on = self.ind == rows
        d[rows[on]] = self.val[on]
        return d

    def to_scipy(self):
        import scipy.sparse as sp
        A = sp.csr_matrix((self.val, self.ind, self.ptr), shape=(self.m, self.n))
        if self.sym:
            A = A + sp.triu(A, 1).T
        return A.tocsr()

    def astype(self, dtype) -> "CSR":
        return CSR(self.m, self.n, self.ptr, self.ind,
                   np.ascontiguousarray(self.val.astype(dtype)), self.sym)


def uniform(n: int, seed: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    _load().shk_wl_uniform(n, seed, lo, hi, out)
    return out


def normal(n: int, seed: int, mean: float = 0.0, sd: float = 1.0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    _load().shk_wl_normal(n, seed, mean, sd, out)
    return out


def aniso2d(nx: int, ny: int, eps: float) -> CSR:
    lib = _load()
    nnz = lib.shk_wl_aniso2d_nnz(nx, ny)
    m = nx * ny
    ptr = np.empty(m + 1, np.int32)
    ind = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    lib.shk_wl_aniso2d(nx, ny, eps, ptr, ind, val)
    return CSR(m, m, ptr, ind, val)


def laplace2d(nx: int, ny: int | None = None) -> CSR:
    return aniso2d(nx, nx if ny is None else ny, 1.0)


def upwind2d(g: int, vx: float = 20.0, vy: float = 10.0) -> CSR:
    """Upwind convection-diffusion on a g x g grid (natural order, sorted
    columns): -lap u + h (vx d/dx + vy d/dy) u with rows scaled by h^2,
    h = 1 / (g + 1); nonsymmetric.  The C++ overlay program
    (tests/cpp/seldon_overlay_check.cpp) builds the same matrix inside Seldon,
    entry for entry."""
    n, h = g * g, 1.0 / (g + 1)
    ptr = np.zeros(n + 1, np.int32)
    ind, val = [], []
    for iy in range(g):
        for ix in range(g):
            r = iy * g + ix
            if iy > 0:
                ind.append(r - g); val.append(-1 - h * vy)
            if ix > 0:
                ind.append(r - 1); val.append(-1 - h * vx)
            ind.append(r); val.append(4 + h * (vx + vy))
            if ix < g - 1:
                ind.append(r + 1); val.append(-1.0)
            if iy < g - 1:
                ind.append(r + g); val.append(-1.0)
            ptr[r + 1] = len(ind)
    return CSR(n, n, ptr, np.array(ind, np.int32), np.array(val, np.float64))


def convdiff3d(n: int, v=(1.0, 0.5, 0.25), peclet: float = 10.0) -> CSR:
    lib = _load()
    nu = math.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2) / peclet
    nnz = lib.shk_wl_stencil3d_nnz(n)
    m = n ** 3
    ptr = np.empty(m + 1, np.int32)
    ind = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    lib.shk_wl_convdiff3d(n, v[0], v[1], v[2], nu, ptr, ind, val)
    return CSR(m, m, ptr, ind, val)


def diffusion3d(n: int, theta: np.ndarray) -> CSR:
    lib = _load()
    kappa = np.ascontiguousarray(np.exp(np.asarray(theta, dtype=np.float64)))
    assert kappa.size == n ** 3
    nnz = lib.shk_wl_stencil3d_nnz(n)
    m = n ** 3
    ptr = np.empty(m + 1, np.int32)
    ind = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    lib.shk_wl_diffusion3d(n, kappa, ptr, ind, val)
    return CSR(m, m, ptr, ind, val)


def powerlaw(n: int, seed: int, kmin: int = 4, kmax: int = 512,
             expo: float = 1.7452, half_width: int = 20000, dd: float = 1.1) -> CSR:
    lib = _load()
    ptr = np.empty(n + 1, np.int32)
    nnz = lib.shk_wl_powerlaw_ptr(n, seed, kmin, kmax, expo, half_width, ptr)
    ind = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    lib.shk_wl_powerlaw_fill(n, seed, kmin, kmax, expo, half_width, dd, ptr, ind, val)
    return CSR(n, n, ptr, ind, val)


def upper(A: CSR) -> CSR:
    """Upper triangle incl. diagonal (what Seldon's RowSymSparse stores)."""
    lib = _load()
    ptr_u = np.empty(A.m + 1, np.int32)
    nnz_u = lib.shk_wl_upper(A.m, A.ptr, A.ind, A.val, ptr_u, None, None)
    ind_u = np.empty(nnz_u, np.int32)
    val_u = np.empty(nnz_u, np.float64)
    lib.shk_wl_upper(A.m, A.ptr, A.ind, A.val, ptr_u,
                     ind_u.ctypes.data_as(ctypes.c_void_p),
                     val_u.ctypes.data_as(ctypes.c_void_p))
    return CSR(A.m, A.n, ptr_u, ind_u, val_u, sym=True)


def first_columns(A: CSR, ncols: int) -> CSR:
    """The m x ncols matrix of A's entries in columns [0, ncols), stored order
    kept (a rectangular case for the transposed products)."""
    keep = A.ind < ncols
    rows = np.repeat(np.arange(A.m), np.diff(A.ptr))[keep]
    ptr = np.zeros(A.m + 1, np.int32)
    np.add.at(ptr, rows + 1, 1)
    return CSR(A.m, ncols, np.cumsum(ptr).astype(np.int32), A.ind[keep].astype(np.int32),
               A.val[keep].astype(np.float64))


def apply(A: CSR, x: np.ndarray) -> np.ndarray:
    """y = A x (full CSR only), sequential per-row sums in double."""
    assert not A.sym
    y = np.empty(A.m, np.float64)
    _load().shk_wl_csr_apply(A.m, A.ptr, A.ind, A.val, np.ascontiguousarray(x, np.float64), y)
    return y


# ----------------------------------------------------------- named configs --

SEED_XTRUE = 12345


def cg_laplace(grid: int = 256, seed: int = SEED_XTRUE):
    """2-D 5-point Laplacian; x_true ~ U[-1,1]; b = A x_true."""
    A = laplace2d(grid)
    xt = uniform(A.m, seed)
    return A, apply(A, xt), xt


def convdiff(grid: int = 128, seed: int = SEED_XTRUE + 1):
    A = convdiff3d(grid)
    xt = uniform(A.m, seed)
    return A, apply(A, xt), xt


def upwind(grid: int = 24):
    """2-D upwind convection-diffusion (vx 20, vy 10); x_true_i = cos(0.21 i)
    - 0.3 sin(0.7 i); b = A x_true (the overlay program's nonsymmetric case)."""
    A = upwind2d(grid)
    xt = np.array([math.cos(0.21 * k) - 0.3 * math.sin(0.7 * k) for k in range(A.m)])
    return A, apply(A, xt), xt


def fem_powerlaw(n: int = 1 << 22, seed: int = 2024, xseed: int = SEED_XTRUE + 2):
    A = powerlaw(n, seed)
    xt = uniform(A.m, xseed)
    return A, apply(A, xt), xt


def batched_systems(count: int = 512, seed: int = 77, nmin: int = 32, nmax: int = 256):
    """Grid sides n_s ~ U{nmin..nmax}, anisotropy eps_s ~ logU[1,100]."""
    u = uniform(2 * count, seed, 0.0, 1.0)
    sides = (nmin + np.floor(u[:count] * (nmax - nmin + 1))).astype(int)
    sides = np.minimum(sides, nmax)
    eps = np.exp(u[count:] * math.log(100.0))
    return sides, eps
