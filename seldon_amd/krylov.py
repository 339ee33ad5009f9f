"""Seldon-named Python front-end over the HIP library.

Mirrors the names and argument meaning of Seldon's sparse Krylov API
(computation/solver/iterative/Iterative.hxx, Cg.cxx, BiCgStab.cxx,
Gmres.cxx, computation/basic_functions/Functions_Vector.cxx,
Functions_MatVect.cxx) so the parity tests read like Seldon code:

    A = Matrix(csr, storage="row")          # Matrix<double,General,RowSparse>
    it = Iteration(max_iter, tol)            # Iteration<double>
    code = Cg(A, x, b, Preconditioner_Base(), it)

Vectors are torch tensors on the GPU (torch is used for device memory and
streams only); every operation runs in libshk.so.  The calls are ordered on
torch's current stream.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L

_DT = {torch.float64: L.SHK_F64, torch.float32: L.SHK_F32}
_NP = {torch.float64: np.float64, torch.float32: np.float32}
_STORAGE = {"row": L.SHK_ROWSPARSE, "sym": L.SHK_ROWSYMSPARSE, "array": L.SHK_ARRAYROWSPARSE}


class Context:
    """One device, one stream (the current torch stream), its workspace."""

    _default = None

    def __init__(self, device: int = 0):
        lib = L.load()
        h = ctypes.c_void_p()
        L.check(lib.shk_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device
        self.bind_stream()

    def bind_stream(self, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        L.check(L.load().shk_ctx_set_stream(self.h, ctypes.c_void_p(s.cuda_stream)))

    def synchronize(self):
        L.check(L.load().shk_ctx_synchronize(self.h))

    def info(self):
        cus, xcds = ctypes.c_int(), ctypes.c_int()
        arch = ctypes.create_string_buffer(64)
        L.check(L.load().shk_ctx_device_info(self.h, ctypes.byref(cus), ctypes.byref(xcds), arch, 64))
        return dict(cus=cus.value, xcds=xcds.value, arch=arch.value.decode())

    def close(self):
        if self.h:
            L.load().shk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = Context(torch.cuda.current_device())
        return cls._default


def _ptr(t: torch.Tensor):
    assert t.is_cuda and t.is_contiguous(), "device vectors must be contiguous CUDA tensors"
    return ctypes.c_void_p(t.data_ptr())


@dataclass
class DeviceCSR:
    """CSR arrays already on the device (torch tensors), e.g. from
    ConvertMatrix_from_Coordinates; Matrix(DeviceCSR(...)) takes them."""
    m: int
    n: int
    ptr: torch.Tensor
    ind: torch.Tensor
    val: torch.Tensor


class Matrix:
    """A sparse matrix resident on the GPU.

    storage: "row" (RowSparse), "sym" (RowSymSparse: upper triangle given),
    "array" (ArrayRowSparse, rows laid end to end)."""

    LAYOUTS = {"auto": 0, "rows": 1, "panels": 2, "sliced": 3}

    def __init__(self, csr, storage: str | None = None, dtype=torch.float64, ctx: Context | None = None,
                 layout: str = "auto"):
        self.ctx = ctx or Context.default()
        if storage is None:
            storage = "sym" if getattr(csr, "sym", False) else "row"
        self.storage = storage
        self.dtype = dtype
        self.m, self.n = int(csr.m), int(csr.n)
        if isinstance(csr.ptr, torch.Tensor):  # CSR arrays already on the device
            assert csr.ptr.is_cuda and csr.val.dtype == dtype
            ptr, ind, val = (t.contiguous() for t in (csr.ptr, csr.ind, csr.val))
            self.nnz = int(ind.numel())
            args, on_dev = (_ptr(ptr), _ptr(ind) if self.nnz else None, _ptr(val) if self.nnz else None), 1
        else:
            ptr = np.ascontiguousarray(csr.ptr, np.int32)
            ind = np.ascontiguousarray(csr.ind, np.int32)
            val = np.ascontiguousarray(np.asarray(csr.val), _NP[dtype])
            self.nnz = int(ptr[-1])
            args, on_dev = tuple(a.ctypes.data_as(ctypes.c_void_p) for a in (ptr, ind, val)), 0
        h = ctypes.c_void_p()
        L.check(L.load().shk_csr_create_layout(self.ctx.h, _STORAGE[storage], _DT[dtype], self.m, self.n,
                                               self.nnz, *args, on_dev, self.LAYOUTS[layout], ctypes.byref(h)))
        self.h = h

    def Write(self, path: str):
        """Seldon's binary CSR file of the stored arrays (Matrix_Sparse.cxx:1532-1556)."""
        L.check(L.load().shk_csr_write(self.ctx.h, self.h, str(path).encode()))

    @classmethod
    def _wrap(cls, h, ctx, storage, dtype) -> "Matrix":
        self = cls.__new__(cls)
        self.ctx, self.storage, self.dtype, self.h = ctx, storage, dtype, h
        m, n, a, b, d, st = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64(),
                             ctypes.c_int(), ctypes.c_int())
        L.check(L.load().shk_csr_info(h, ctypes.byref(m), ctypes.byref(n), ctypes.byref(a), ctypes.byref(b),
                                      ctypes.byref(d), ctypes.byref(st)))
        self.m, self.n, self.nnz = m.value, n.value, int(a.value)
        return self

    @classmethod
    def Read(cls, path: str, storage: str = "row", dtype=torch.float64, layout: str = "auto",
             ctx: Context | None = None) -> "Matrix":
        """A matrix from Seldon's binary CSR file (Matrix_Sparse.cxx:1649-1682)."""
        ctx = ctx or Context.default()
        h = ctypes.c_void_p()
        L.check(L.load().shk_csr_read(ctx.h, str(path).encode(), _STORAGE[storage], _DT[dtype],
                                      cls.LAYOUTS[layout], ctypes.byref(h)))
        return cls._wrap(h, ctx, storage, dtype)

    @property
    def layout(self) -> str:
        """The device layout chosen at creation: "rows", "panels" or "sliced"."""
        v = ctypes.c_int()
        L.check(L.load().shk_csr_layout(self.h, ctypes.byref(v)))
        return {1: "rows", 2: "panels", 3: "sliced"}[v.value]

    def GetM(self):
        return self.m

    def GetN(self):
        return self.n

    def set_values(self, val):
        if isinstance(val, torch.Tensor) and val.is_cuda:
            assert val.dtype == self.dtype and val.numel() == self.nnz
            L.check(L.load().shk_csr_set_values(self.ctx.h, self.h, _ptr(val.contiguous()), 1))
        else:
            v = np.ascontiguousarray(np.asarray(val), _NP[self.dtype])
            L.check(L.load().shk_csr_set_values(self.ctx.h, self.h, v.ctypes.data_as(ctypes.c_void_p), 0))

    def spmv_bytes(self, beta_nonzero=False):
        return L.load().shk_csr_spmv_bytes(self.h, int(beta_nonzero))

    def device_nnz(self):
        m, n, a, b, d, s = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64(),
                            ctypes.c_int(), ctypes.c_int())
        L.check(L.load().shk_csr_info(self.h, ctypes.byref(m), ctypes.byref(n), ctypes.byref(a),
                                      ctypes.byref(b), ctypes.byref(d), ctypes.byref(s)))
        return b.value

    def close(self):
        if getattr(self, "h", None):
            L.load().shk_csr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ctx_of(*objs):
    for o in objs:
        if isinstance(o, Matrix):
            return o.ctx
    return Context.default()


# ----------------------------------------------------------- MltAdd / Mlt --

class SeldonTranspose:
    """Seldon's transpose tags (share/MatrixFlag.hxx:32-80): pass
    SeldonTrans / SeldonNoTrans / SeldonConjTrans as the second argument of
    MltAdd or the first of Mlt.  ConjTrans is Trans for real values."""

    def __init__(self, status: int, name: str):
        self.status = status  # 0: Trans, 1: NoTrans, 2: ConjTrans (MatrixFlag.hxx:38)
        self.name = name

    def Trans(self) -> bool:
        return self.status == 0

    def NoTrans(self) -> bool:
        return self.status == 1

    def ConjTrans(self) -> bool:
        return self.status == 2

    def __repr__(self):
        return self.name


SeldonTrans = SeldonTranspose(0, "SeldonTrans")
SeldonNoTrans = SeldonTranspose(1, "SeldonNoTrans")
SeldonConjTrans = SeldonTranspose(2, "SeldonConjTrans")


def MltAdd(alpha, *args):
    """MltAdd(alpha, A, X, beta, Y): Y <- beta Y + alpha A X
    (Functions_MatVect.cxx:398-431, :526-564; Functions_MatrixArray.cxx:534-575).
    MltAdd(alpha, SeldonTrans, A, X, beta, Y): Y <- beta Y + alpha A^T X
    (Functions_MatVect.cxx:729-758; Functions_MatrixArray.cxx:591-632;
    RowSymSparse: the plain product, Functions_MatVect.cxx:918-927)."""
    if isinstance(args[0], SeldonTranspose):
        trans, A, X, beta, Y = args
    else:
        trans, (A, X, beta, Y) = SeldonNoTrans, args
    fn = L.load().shk_spmv if trans.NoTrans() else L.load().shk_spmv_trans
    L.check(fn(A.ctx.h, A.h, float(alpha), _ptr(X), float(beta), _ptr(Y)))


def Mlt(*args):
    """Mlt(A, X, Y): Y = A X (Functions_MatVect.cxx:72-81);
    Mlt(SeldonTrans, A, X, Y): Y = A^T X (Functions_MatVect.cxx:123-130);
    Mlt(alpha, X): X *= alpha (Functions_Vector.cxx:69-74)."""
    if len(args) == 4:
        trans, A, X, Y = args
        fn = L.load().shk_mlt if trans.NoTrans() else L.load().shk_mlt_trans
        L.check(fn(A.ctx.h, A.h, _ptr(X), _ptr(Y)))
    elif len(args) == 3:
        A, X, Y = args
        L.check(L.load().shk_mlt(A.ctx.h, A.h, _ptr(X), _ptr(Y)))
    else:
        alpha, X = args
        L.check(L.load().shk_scal(Context.default().h, _DT[X.dtype], X.numel(), float(alpha), _ptr(X)))


def MltDot(A: Matrix, X, Y) -> float:
    r = ctypes.c_double()
    L.check(L.load().shk_mlt_dot(A.ctx.h, A.h, _ptr(X), _ptr(Y), ctypes.byref(r)))
    return r.value


def ConvertMatrix_from_Coordinates(IndRow: torch.Tensor, IndCol: torch.Tensor, Val: torch.Tensor, index: int = 0,
                                   m: int = 0, n: int = 0):
    """CSR arrays (device tensors ptr, ind, val) and (m, n) from coordinates,
    as Seldon's ConvertMatrix_from_Coordinates builds a RowSparse matrix
    (Matrix_Conversions.cxx:1288-1335): entries ordered by (row, column),
    duplicates kept in input order; m, n at least the given ones (the
    matrix's size before the call).  Matrix(DeviceCSR(m, n, ptr, ind, val))
    uploads nothing: it plans from the device arrays."""
    dt = _DT[Val.dtype]
    row = IndRow.to(torch.int32).contiguous()
    col = IndCol.to(torch.int32).contiguous()
    val = Val.contiguous()
    nnz = row.numel()
    ctx = Context.default()
    mm, nn = ctypes.c_int(), ctypes.c_int()
    null = ctypes.c_void_p(0)
    L.check(L.load().shk_coo_to_csr(ctx.h, dt, nnz, _ptr(row) if nnz else null, _ptr(col) if nnz else null,
                                    _ptr(val) if nnz else null, int(index), int(m), int(n), 0,
                                    ctypes.byref(mm), ctypes.byref(nn), null, null, null))
    ptr = torch.empty(mm.value + 1, dtype=torch.int32, device=row.device)
    ind = torch.empty(max(nnz, 1), dtype=torch.int32, device=row.device)
    out = torch.empty(max(nnz, 1), dtype=val.dtype, device=row.device)
    L.check(L.load().shk_coo_to_csr(ctx.h, dt, nnz, _ptr(row) if nnz else null, _ptr(col) if nnz else null,
                                    _ptr(val) if nnz else null, int(index), int(m), int(n), mm.value,
                                    ctypes.byref(mm), ctypes.byref(nn), _ptr(ptr), _ptr(ind), _ptr(out)))
    return mm.value, nn.value, ptr, ind[:nnz], out[:nnz]


def WriteVector(X: torch.Tensor, path: str, with_size: bool = True):
    """Vector::Write(path, with_size) (Vector.cxx:846-874) of a device vector:
    the length as an int, then the elements."""
    L.check(L.load().shk_vec_write(Context.default().h, _DT[X.dtype], X.numel(), _ptr(X), str(path).encode(),
                                   int(with_size)))


def ReadVector(path: str, dtype=torch.float64, with_size: bool = True, n: int | None = None) -> torch.Tensor:
    """Vector::Read(path, with_size) (Vector.cxx:1060-1091) into a new device
    vector: the file's length, or n elements when the file has none."""
    ctx = Context.default()
    got = ctypes.c_int64()
    if with_size:
        L.check(L.load().shk_vec_read(ctx.h, _DT[dtype], str(path).encode(), 1, 0, None, ctypes.byref(got)))
        n = got.value
    elif n is None:
        raise ValueError("ReadVector without the length needs n")
    X = torch.empty(n, dtype=dtype, device="cuda")
    L.check(L.load().shk_vec_read(ctx.h, _DT[dtype], str(path).encode(), int(with_size), n,
                                  _ptr(X) if n else None, ctypes.byref(got)))
    return X


def WriteHarwellBoeing(A: Matrix, path: str):
    """WriteHarwellBoeing(A, path) (IOMatrixMarket.cxx:397-510): the file
    Seldon writes for the same storage and arrays, byte for byte."""
    L.check(L.load().shk_csr_write_hb(A.ctx.h, A.h, str(path).encode()))


def ReadHarwellBoeing(path: str, dtype=torch.float64, layout: str = "auto", ctx: Context | None = None) -> Matrix:
    """ReadHarwellBoeing(path, A) (IOMatrixMarket.cxx:80-360) into a ColSparse
    matrix, then to RowSparse through coordinates, as a Seldon program does:
    the RowSparse matrix, assembled on the device."""
    ctx = ctx or Context.default()
    h = ctypes.c_void_p()
    L.check(L.load().shk_csr_read_hb(ctx.h, str(path).encode(), _DT[dtype], Matrix.LAYOUTS[layout],
                                     ctypes.byref(h)))
    return Matrix._wrap(h, ctx, "row", dtype)


def SOR(A: Matrix, X: torch.Tensor, B: torch.Tensor, omega, iter: int, type_ssor: int = 2, omega_dtype=None):
    """SOR(A, X, B, omega, iter, type_ssor): 2 forward sweeps, 3 backward, 0
    both (Relaxation_MatVect.cxx:41-95, :106-167, :179-300); X updated in
    place, bitwise as Seldon's sequential sweeps on all three storages.
    omega_dtype is the C++ type of omega: by default the value type (as
    SorPreconditioner passes it); torch.float64 on float32 values makes the
    relaxation step double, as C++ does for a double omega."""
    wide = omega_dtype is not None and torch.empty(0, dtype=omega_dtype).element_size() == 8
    L.check(L.load().shk_sor(A.ctx.h, A.h, _ptr(X), _ptr(B), float(omega),
                             L.SHK_F64 if (wide or X.dtype == torch.float64) else L.SHK_F32, int(iter), int(type_ssor)))


def DotProd(X, Y) -> float:
    r = ctypes.c_double()
    L.check(L.load().shk_dot(Context.default().h, _DT[X.dtype], X.numel(), _ptr(X), _ptr(Y), ctypes.byref(r)))
    return r.value


DotProdConj = DotProd


def Norm2(X) -> float:
    r = ctypes.c_double()
    L.check(L.load().shk_nrm2(Context.default().h, _DT[X.dtype], X.numel(), _ptr(X), ctypes.byref(r)))
    return r.value


def Add(alpha, X, Y):
    L.check(L.load().shk_axpy(Context.default().h, _DT[X.dtype], X.numel(), float(alpha), _ptr(X), _ptr(Y)))


def Copy(X, Y):
    L.check(L.load().shk_copy(Context.default().h, _DT[X.dtype], X.numel(), _ptr(X), _ptr(Y)))


def Fill(X, value):
    L.check(L.load().shk_fill(Context.default().h, _DT[X.dtype], X.numel(), float(value), _ptr(X)))


def AddDot(alpha, X, Y, Z) -> float:
    r = ctypes.c_double()
    L.check(L.load().shk_axpy_dot(Context.default().h, _DT[X.dtype], X.numel(), float(alpha), _ptr(X), _ptr(Y),
                                  _ptr(Z), ctypes.byref(r)))
    return r.value


# ------------------------------------------------------- preconditioners --

class Preconditioner_Base:
    """Identity preconditioner (Iterative.cxx:56-77)."""
    kind = L.SHK_PREC_IDENTITY

    def Solve(self, A: Matrix, r, z):
        Copy(r, z)

    TransSolve = Solve


class JacobiPreconditioner(Preconditioner_Base):
    """z_i = r_i / a_ii (new; division form of SOR's /a_jj,
    matrix_sparse/Relaxation_MatVect.cxx:79)."""
    kind = L.SHK_PREC_JACOBI

    def Solve(self, A: Matrix, r, z):
        L.check(L.load().shk_jacobi(A.ctx.h, A.h, _ptr(r), _ptr(z)))

    TransSolve = Solve


# --------------------------------------------------------------- Iteration --

@dataclass
class Iteration:
    """Iteration<double> (Iterative.hxx:47-108, Iterative.cxx:81-354)."""
    max_iter: int = 100
    tolerance: float = 1e-6
    nb_iter: int = 0
    facteur_reste: float = 0.0
    error_code: int = 0
    fail_convergence: bool = False
    print_level: int = 1
    init_guess_null: bool = True
    parameter_restart: int = 10
    history: np.ndarray = field(default_factory=lambda: np.zeros(0))
    history_it: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    hist_len: int = 0

    def __init__(self, max_iter: int = 100, tol: float = 1e-6):
        self.max_iter, self.tolerance = int(max_iter), float(tol)
        self.nb_iter, self.facteur_reste, self.error_code = 0, 0.0, 0
        self.fail_convergence, self.print_level = False, 1
        self.init_guess_null, self.parameter_restart = True, 10
        self.history = np.zeros(0)
        self.history_it = np.zeros(0, np.int32)
        self.hist_len = 0

    def SetRestart(self, m):
        self.parameter_restart = int(m)

    def GetRestart(self):
        return self.parameter_restart

    def SetTolerance(self, t):
        self.tolerance = float(t)

    def SetMaxNumberIteration(self, m):
        self.max_iter = int(m)

    def SetNumberIteration(self, nb):
        self.nb_iter = int(nb)

    def GetNumberIteration(self):
        return self.nb_iter

    def SetInitGuess(self, null: bool):
        self.init_guess_null = bool(null)

    def IsInitGuess_Null(self):
        return self.init_guess_null

    def ShowMessages(self):
        self.print_level = 1

    def ShowFullHistory(self):
        self.print_level = 6

    def HideMessages(self):
        self.print_level = 0

    def ErrorCode(self):
        if self.nb_iter >= self.max_iter:
            return -2
        if self.fail_convergence:
            return self.error_code
        return 0

    def _to_c(self, cap):
        self._hist = np.zeros(max(cap, 1), np.float64)
        self._hist_it = np.zeros(max(cap, 1), np.int32)
        c = L.ShkIteration()
        c.tolerance, c.facteur_reste = self.tolerance, self.facteur_reste
        c.max_iter, c.nb_iter = self.max_iter, self.nb_iter
        c.error_code, c.fail_convergence = self.error_code, int(self.fail_convergence)
        c.print_level, c.init_guess_null = self.print_level, int(self.init_guess_null)
        c.parameter_restart = self.parameter_restart
        c.hist_cap = cap
        c.hist = self._hist.ctypes.data_as(ctypes.c_void_p)
        c.hist_it = self._hist_it.ctypes.data_as(ctypes.c_void_p)
        return c

    def _from_c(self, c):
        self.facteur_reste, self.nb_iter = c.facteur_reste, c.nb_iter
        self.error_code, self.fail_convergence = c.error_code, bool(c.fail_convergence)
        self.hist_len = c.hist_len
        n = min(c.hist_len, c.hist_cap)
        self.history = self._hist[:n].copy()
        self.history_it = self._hist_it[:n].copy()
        if self.print_level >= 1:
            for k, (it, res) in enumerate(zip(self.history_it, self.history)):
                if (self.print_level >= 6) or it % 100 == 0:
                    print(f"Residu at iteration number {it}  {res:.17g}")


def _solve(fn, A: Matrix, x, b, M, it: Iteration, hist_cap=None):
    assert x.dtype == A.dtype and b.dtype == A.dtype
    cap = hist_cap if hist_cap is not None else 2 * it.max_iter + 4 * (it.max_iter // max(it.parameter_restart, 1)) + 16
    c = it._to_c(cap)
    code = ctypes.c_int()
    L.check(fn(A.ctx.h, A.h, _ptr(x), _ptr(b), M.kind, ctypes.byref(c), ctypes.byref(code)))
    it._from_c(c)
    return code.value


def Cg(A, x, b, M, it):
    """Cg (computation/solver/iterative/Cg.cxx:45-116) on the GPU."""
    return _solve(L.load().shk_cg, A, x, b, M, it)


def BiCgStab(A, x, b, M, it):
    """BiCgStab (computation/solver/iterative/BiCgStab.cxx:44-145) on the GPU."""
    return _solve(L.load().shk_bicgstab, A, x, b, M, it)


def Gmres(A, x, b, M, it):
    """Gmres(m) (computation/solver/iterative/Gmres.cxx:45-182) on the GPU."""
    return _solve(L.load().shk_gmres, A, x, b, M, it)


def Cgs(A, x, b, M, it):
    """Cgs (computation/solver/iterative/Cgs.cxx:46-137) on the GPU, fused."""
    return _solve(L.load().shk_cgs, A, x, b, M, it)


def BiCg(A, x, b, M, it):
    """BiCg (computation/solver/iterative/BiCg.cxx:47-131) on the GPU, fused;
    its A^T products run on the library's device-built transpose."""
    return _solve(L.load().shk_bicg, A, x, b, M, it)


def TfQmr(A, x, b, M, it):
    """TfQmr (computation/solver/iterative/TfQmr.cxx:45-240) on the GPU, fused."""
    return _solve(L.load().shk_tfqmr, A, x, b, M, it)


def Qmr(A, x, b, M, it):
    """Qmr (computation/solver/iterative/Qmr.cxx:42-190) on the GPU, fused;
    its A^T products run on the library's device-built transpose."""
    return _solve(L.load().shk_qmr, A, x, b, M, it)


def CoCg(A, x, b, M, it):
    """CoCg (computation/solver/iterative/CoCg.cxx:44-115): for real values
    its loop is Cg's operation for operation, so it runs the fused Cg."""
    return _solve(L.load().shk_cg, A, x, b, M, it)


def BiCgcr(A, x, b, M, it):
    """BiCgcr (computation/solver/iterative/BiCgcr.cxx:45-117) on the GPU, fused."""
    return _solve(L.load().shk_bicgcr, A, x, b, M, it)


def QmrSym(A, x, b, M, it):
    """QmrSym (computation/solver/iterative/QmrSym.cxx:42-166) on the GPU, fused."""
    return _solve(L.load().shk_qmrsym, A, x, b, M, it)


def MinRes(A, x, b, M, it):
    """MinRes (computation/solver/iterative/MinRes.cxx:47-151) on the GPU, fused."""
    return _solve(L.load().shk_minres, A, x, b, M, it)


def Symmlq(A, x, b, M, it):
    """Symmlq (computation/solver/iterative/Symmlq.cxx:45-161) on the GPU, fused."""
    return _solve(L.load().shk_symmlq, A, x, b, M, it)


def QCgs(A, x, b, M, it):
    """QCgs (computation/solver/iterative/QCgs.cxx:47-164) on the GPU, fused."""
    return _solve(L.load().shk_qcgs, A, x, b, M, it)


def Gcr(A, x, b, M, it):
    """Gcr (computation/solver/iterative/Gcr.cxx:45-139) on the GPU, fused;
    restarted every it.parameter_restart iterations."""
    return _solve(L.load().shk_gcr, A, x, b, M, it)


def BiCgStabl(A, x, b, M, it):
    """BiCgStabl (computation/solver/iterative/BiCgStabl.cxx:43-180) on the
    GPU, fused; l = it.parameter_restart BiCG steps per cycle."""
    return _solve(L.load().shk_bicgstabl, A, x, b, M, it)


def Cgne(A, x, b, M, it):
    """Cgne (computation/solver/iterative/Cgne.cxx:41-124) on the GPU, fused."""
    return _solve(L.load().shk_cgne, A, x, b, M, it)


def Lsqr(A, x, b, M, it):
    """Lsqr (computation/solver/iterative/Lsqr.cxx:41-150) on the GPU, fused;
    like the template it applies no preconditioner."""
    return _solve(L.load().shk_lsqr, A, x, b, M, it)


def solve_ir(A64: Matrix, A32: Matrix, x, b, M, tol=1e-10, inner_restart=50, inner_tol=1e-6,
             inner_max_iter=2000, max_refinements=50, inner_method="gmres"):
    """FP32 inner Krylov + FP64 iterative refinement (no Seldon counterpart)."""
    p = L.ShkIrParams()
    p.tolerance, p.max_refinements = tol, max_refinements
    p.inner_restart, p.inner_max_iter, p.inner_tolerance = inner_restart, inner_max_iter, inner_tol
    p.inner_method = 0 if inner_method == "gmres" else 1
    hist = np.zeros(max_refinements + 1)
    p.relres_hist = hist.ctypes.data_as(ctypes.c_void_p)
    code = ctypes.c_int()
    L.check(L.load().shk_solve_ir(A64.ctx.h, A64.h, A32.h, _ptr(x), _ptr(b), M.kind, ctypes.byref(p),
                                  ctypes.byref(code)))
    return dict(code=code.value, refinements=p.refinements, inner_iters=p.total_inner_iter,
                relres=p.final_relres, relres_hist=hist[:p.refinements + 1].copy())
