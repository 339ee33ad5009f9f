"""TEST INFRASTRUCTURE ONLY -- the CPU oracle for the sparse Krylov path.

ctypes front-end over ``oracle/liboracle.so`` (built from ``seldon_oracle.c``
by ``oracle/Makefile``), a plain-C restatement of Seldon's generic templates:
SpMV on RowSparse / RowSymSparse / ArrayRowSparse and its transposed forms,
DotProd / Norm2 / Add / Mlt / Copy, Iteration, Cg, BiCgStab, Gmres, Cgs, BiCg, TfQmr, Qmr, Cgne,
Lsqr, BiCgcr, QmrSym, MinRes, Symmlq, QCgs, Gcr and BiCgStabl, SOR,
and ConvertMatrix_from_Coordinates (file:line citations are in
``seldon_oracle_body.h``); plus, in Python below, the Harwell-Boeing writer
and reader.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` legs of ``bench.py`` may use it, and only as the reference
the GPU results are compared with.

Parity pins (see ``tests/test_oracle.py``):
  * bitwise equality with drivers compiled from the real Seldon headers
    (``oracle/_ref``) on SpMV (plain and transposed), Cg, BiCgStab, SOR on
    the three storages, the coordinate conversion and the Harwell-Boeing
    files, and on committed fixtures (``tests/golden``) made by them;
  * Cgs, QCgs, BiCg, TfQmr, Qmr, Cgne, Lsqr, Gcr, BiCgStabl, BiCgcr, QmrSym, MinRes and Symmlq
    (and Cg as Seldon's CoCg):
    bitwise against Seldon's own
    runs of them (the overlay program's CPU legs,
    ``tests/golden/overlay_cpu.npz``), closed cases (A = 2.5 I; diagonal A
    under Jacobi; Lsqr's exact breakdown on A = 2 I, b = e_0), BiCg
    following Cg on an SPD matrix (history at 2k = Cg's at k within 1e-9),
    Cgne following Cg on the explicit A^T A, TfQmr's quasi-residual bound on
    the true residual, Qmr's carried residual and Lsqr's |phi_bar| equal to
    the true one, FP32 within kappa eps_32; BiCgStab(l) on a diagonal A
    with four distinct values stopping after ceil(4 / l) cycles with
    x = A^-1 b, and BiCgStab(1) under Jacobi solving it in one cycle through
    the final change of coordinates (both also in float);
  * Gmres: iteration counts and residual histories against the MKL-linked
    real Seldon build (tolerance, MKL's reductions differ in order);
  * closed forms (5-point Laplacian eigenpairs, SOR as forward substitution
    and its convergence to A^-1 b), dense numpy solves, scipy/numpy
    triangular solves, numpy's stable lexsort and the textbook Givens
    rotation;
  * the FP32 family: ``orc_s_cg`` and ``orc_s_bicgstab`` bitwise against
    Seldon's own float build (Iteration<double>, generic templates) on the
    three storages (``tests/golden/solve_f32_*``); ``orc_s_gmres`` keeps its
    Hessenberg, rotations and triangular solve in double (the FP32 inner
    solver of the mixed-precision path), so it is pinned by the iteration
    counts and error codes of Seldon's float Gmres (Iteration<float>, float
    Hessenberg), by its residual history within FP32 rounding of that build,
    by A = I (one step, happy breakdown), by diagonal A with k distinct
    values (at most k steps), by the Arnoldi relation A M^-1 V_k = V_{k+1}
    Hbar_k with V orthonormal to FP32 rounding, and by its solution within
    kappa * eps_32 of the FP64 oracle's.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ROWSPARSE, ROWSYMSPARSE, ARRAYROWSPARSE = 0, 1, 2
PREC = {"id": 0, "identity": 0, "jacobi": 1}


class OrcMat(ctypes.Structure):
    _fields_ = [("storage", ctypes.c_int), ("m", ctypes.c_int), ("n", ctypes.c_int),
                ("nnz", ctypes.c_int64), ("ptr", ctypes.c_void_p),
                ("ind", ctypes.c_void_p), ("val", ctypes.c_void_p)]


class OrcIter(ctypes.Structure):
    _fields_ = [("tolerance", ctypes.c_double), ("facteur_reste", ctypes.c_double),
                ("max_iter", ctypes.c_int), ("nb_iter", ctypes.c_int),
                ("error_code", ctypes.c_int), ("fail_convergence", ctypes.c_int),
                ("print_level", ctypes.c_int), ("init_guess_null", ctypes.c_int),
                ("parameter_restart", ctypes.c_int),
                ("hist", ctypes.c_void_p), ("hist_it", ctypes.c_void_p),
                ("hist_len", ctypes.c_int), ("hist_cap", ctypes.c_int),
                ("arnoldi_h", ctypes.c_void_p), ("arnoldi_v", ctypes.c_void_p),
                ("arnoldi_steps", ctypes.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        for sfx, T in (("d", ctypes.c_double), ("s", ctypes.c_float)):
            getattr(L, f"orc_{sfx}_mltadd").argtypes = [T, ctypes.POINTER(OrcMat), vp, T, vp]
            getattr(L, f"orc_{sfx}_mlt").argtypes = [ctypes.POINTER(OrcMat), vp, vp]
            getattr(L, f"orc_{sfx}_mltadd_trans").argtypes = [T, ctypes.POINTER(OrcMat), vp, T, vp]
            getattr(L, f"orc_{sfx}_mlt_trans").argtypes = [ctypes.POINTER(OrcMat), vp, vp]
            getattr(L, f"orc_{sfx}_sor").argtypes = [ctypes.POINTER(OrcMat), vp, vp, ctypes.c_double,
                                                      ctypes.c_int, ctypes.c_int, ctypes.c_int]
            getattr(L, f"orc_{sfx}_coo_to_csr").argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int,
                                                            ctypes.c_int, vp, vp, vp]
            getattr(L, f"orc_{sfx}_dot").argtypes = [ctypes.c_int, vp, vp]
            getattr(L, f"orc_{sfx}_dot").restype = T
            getattr(L, f"orc_{sfx}_norm2").argtypes = [ctypes.c_int, vp]
            getattr(L, f"orc_{sfx}_norm2").restype = T
            for solver in ("cg", "bicgstab", "gmres", "cgs", "bicg", "tfqmr", "qmr", "cgne", "lsqr", "bicgcr", "qmrsym", "minres", "symmlq", "qcgs", "gcr",
                           "bicgstabl"):
                f = getattr(L, f"orc_{sfx}_{solver}")
                f.argtypes = [ctypes.POINTER(OrcMat), vp, vp, ctypes.c_int, vp,
                              ctypes.POINTER(OrcIter)]
                f.restype = ctypes.c_int
            getattr(L, f"orc_{sfx}_jacobi_diag").argtypes = [ctypes.POINTER(OrcMat), vp]
            getattr(L, f"orc_{sfx}_sddmm").argtypes = [ctypes.c_int, vp, vp, vp, vp, vp]
        L.orc_iter_init.argtypes = [ctypes.POINTER(OrcIter), ctypes.c_int, ctypes.c_double]
        L.orc_coo_dims.argtypes = [ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.orc_trsv_upper.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp]
        L.orc_genrot.argtypes = [ctypes.POINTER(ctypes.c_double)] * 4
        L.orc_applyrot.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                   ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype):
    return "s" if np.dtype(dtype) == np.float32 else "d"


class _Mat:
    """Keeps the numpy arrays alive alongside the C struct."""

    def __init__(self, A, storage, dtype):
        self.ptr = np.ascontiguousarray(A.ptr, np.int32)
        self.ind = np.ascontiguousarray(A.ind, np.int32)
        self.val = np.ascontiguousarray(A.val, dtype)
        self.s = OrcMat(storage, A.m, A.n, int(self.ptr[-1]), _p(self.ptr), _p(self.ind), _p(self.val))


def storage_of(A, storage=None):
    if storage is not None:
        return {"row": ROWSPARSE, "sym": ROWSYMSPARSE, "array": ARRAYROWSPARSE}.get(storage, storage)
    return ROWSYMSPARSE if getattr(A, "sym", False) else ROWSPARSE


def mltadd(alpha, A, x, beta, y, storage=None, dtype=np.float64):
    M = _Mat(A, storage_of(A, storage), dtype)
    x = np.ascontiguousarray(x, dtype)
    y = np.array(y, dtype=dtype, copy=True)
    getattr(lib(), f"orc_{_sfx(dtype)}_mltadd")(alpha, ctypes.byref(M.s), _p(x), beta, _p(y))
    return y


def mlt(A, x, storage=None, dtype=np.float64):
    M = _Mat(A, storage_of(A, storage), dtype)
    x = np.ascontiguousarray(x, dtype)
    y = np.empty(A.m, dtype)
    getattr(lib(), f"orc_{_sfx(dtype)}_mlt")(ctypes.byref(M.s), _p(x), _p(y))
    return y


def mltadd_trans(alpha, A, x, beta, y, storage=None, dtype=np.float64):
    """MltAdd(alpha, SeldonTrans, A, x, beta, y): x has A.m entries, y A.n
    (A.m for RowSymSparse)."""
    M = _Mat(A, storage_of(A, storage), dtype)
    x = np.ascontiguousarray(x, dtype)
    y = np.array(y, dtype=dtype, copy=True)
    getattr(lib(), f"orc_{_sfx(dtype)}_mltadd_trans")(alpha, ctypes.byref(M.s), _p(x), beta, _p(y))
    return y


def mlt_trans(A, x, storage=None, dtype=np.float64):
    """Mlt(SeldonTrans, A, x, y): Seldon's generic loop, which reads A(j, i)
    for every (j, i): O(m n), for small and medium test sizes."""
    st = storage_of(A, storage)
    M = _Mat(A, st, dtype)
    x = np.ascontiguousarray(x, dtype)
    y = np.empty(A.m if st == ROWSYMSPARSE else A.n, dtype)
    getattr(lib(), f"orc_{_sfx(dtype)}_mlt_trans")(ctypes.byref(M.s), _p(x), _p(y))
    return y


def sor(A, x, b, omega, iter, type_ssor=2, storage=None, dtype=np.float64, omega_dtype=None):
    """SOR(A, X, B, omega, iter, type_ssor) on a copy of x; returns it.
    omega_dtype: the C++ type of omega (default: the value type); float64
    with float32 values makes the relaxation step double, as in C++."""
    M = _Mat(A, storage_of(A, storage), dtype)
    x = np.array(x, dtype=dtype, copy=True)
    b = np.ascontiguousarray(b, dtype)
    wide = int(np.dtype(dtype) == np.float32 and omega_dtype is not None and np.dtype(omega_dtype) == np.float64)
    getattr(lib(), f"orc_{_sfx(dtype)}_sor")(ctypes.byref(M.s), _p(x), _p(b), float(omega), wide, iter, type_ssor)
    return x


def coo_to_csr(row, col, val, index=0, m_min=0, n_min=0, dtype=np.float64):
    """ConvertMatrix_from_Coordinates (RowSparse): returns (m, n, ptr, ind, val)."""
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, dtype)
    m, n = ctypes.c_int(), ctypes.c_int()
    lib().orc_coo_dims(row.size, _p(row), _p(col), index, m_min, n_min, ctypes.byref(m), ctypes.byref(n))
    ptr = np.empty(m.value + 1, np.int32)
    ind = np.empty(row.size, np.int32)
    out = np.empty(row.size, dtype)
    getattr(lib(), f"orc_{_sfx(dtype)}_coo_to_csr")(row.size, _p(row), _p(col), _p(val), index, m.value,
                                                    _p(ptr), _p(ind), _p(out))
    return m.value, n.value, ptr, ind, out


def dot(x, y):
    x = np.ascontiguousarray(x); y = np.ascontiguousarray(y, x.dtype)
    return getattr(lib(), f"orc_{_sfx(x.dtype)}_dot")(x.size, _p(x), _p(y))


def norm2(x):
    x = np.ascontiguousarray(x)
    return getattr(lib(), f"orc_{_sfx(x.dtype)}_norm2")(x.size, _p(x))


def jacobi_diag(A, storage=None, dtype=np.float64):
    M = _Mat(A, storage_of(A, storage), dtype)
    d = np.empty(A.m, dtype)
    getattr(lib(), f"orc_{_sfx(dtype)}_jacobi_diag")(ctypes.byref(M.s), _p(d))
    return d


@dataclass
class SolveResult:
    x: np.ndarray
    code: int
    nb_iter: int
    hist: np.ndarray      # every relative residual Finished() compared
    hist_it: np.ndarray   # iteration number at each of those calls
    arnoldi: tuple | None = None  # Gmres, first cycle: (V (k+1) x n, Hbar (k+1) x k)


def solve(method, A, b, precond="id", tol=1e-10, max_iter=10000, restart=10,
          x0=None, storage=None, dtype=np.float64, hist_cap=None, arnoldi=False) -> SolveResult:
    st = storage_of(A, storage)
    M = _Mat(A, st, dtype)
    b = np.ascontiguousarray(b, dtype)
    x = np.zeros(A.m, dtype) if x0 is None else np.array(x0, dtype=dtype, copy=True)
    diag = jacobi_diag(A, st, dtype)
    cap = hist_cap or (2 * max_iter + 4 * (max_iter // max(restart, 1)) + 16)
    hist = np.zeros(cap, np.float64)
    hist_it = np.zeros(cap, np.int32)
    it = OrcIter()
    lib().orc_iter_init(ctypes.byref(it), max_iter, tol)
    it.parameter_restart = restart
    it.init_guess_null = 1 if x0 is None else 0
    it.hist, it.hist_it, it.hist_cap = _p(hist), _p(hist_it), cap
    if arnoldi:
        mm = restart
        ah = np.zeros((mm, mm + 1), np.float64)          # column i = Hbar(:, i)
        av = np.zeros((mm + 1, A.m), dtype)
        it.arnoldi_h, it.arnoldi_v = _p(ah), _p(av)
    code = getattr(lib(), f"orc_{_sfx(dtype)}_{method}")(
        ctypes.byref(M.s), _p(x), _p(b), PREC[precond], _p(diag), ctypes.byref(it))
    n = min(it.hist_len, cap)
    arn = None
    if arnoldi:
        k = it.arnoldi_steps
        arn = (av[:k + 1].copy(), ah[:k, :k + 1].T.copy())
    return SolveResult(x, code, it.nb_iter, hist[:n].copy(), hist_it[:n].copy(), arn)


@dataclass
class IrResult:
    x: np.ndarray
    code: int
    refinements: int
    inner_iters: list      # inner iterations of each refinement step
    relres_hist: np.ndarray


def solve_ir(A, b, precond="jacobi", tol=1e-10, inner_method="gmres", inner_restart=50, inner_tol=1e-6,
             inner_max_iter=2000, max_refinements=50) -> IrResult:
    """FP32 inner solve + FP64 iterative refinement, the definition the
    library's shk_solve_ir implements (no Seldon counterpart): x0 = 0;
    while |b - A x| / |b| >= tol: r = b - A x (FP64 MltAdd), d = the FP32
    solver (``solve(..., dtype=float32)``, a null initial guess) on A d =
    float(r * (1/|r|)), x += |r| * double(d); stop with code 1 when the true
    residual does not decrease, -2 after max_refinements."""
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(A.m, np.float64)
    bn = float(norm2(b))
    if bn == 0.0:
        return IrResult(x, 0, 0, [], np.zeros(1))
    r = mltadd(-1.0, A, x, 1.0, b)
    rn = float(norm2(r))
    rel = rn / bn
    hist, inner, code, k = [rel], [], 0, 0
    while rel >= tol and k < max_refinements:
        r32 = (r * (1.0 / rn)).astype(np.float32)
        res = solve(inner_method, A, r32, precond=precond, tol=inner_tol, max_iter=inner_max_iter,
                    restart=inner_restart, dtype=np.float32)
        inner.append(res.nb_iter)
        x += rn * res.x.astype(np.float64)
        r = mltadd(-1.0, A, x, 1.0, b)
        rn = float(norm2(r))
        prev, rel = rel, rn / bn
        k += 1
        hist.append(rel)
        if not rel < prev:
            code = 1
            break
    if code == 0 and rel >= tol:
        code = -2
    return IrResult(x, code, k, inner, np.array(hist))


def sddmm(A, lam, x):
    dtype = np.asarray(lam).dtype
    lam = np.ascontiguousarray(lam); x = np.ascontiguousarray(x, dtype)
    g = np.empty(A.nnz, dtype)
    ptr = np.ascontiguousarray(A.ptr, np.int32); ind = np.ascontiguousarray(A.ind, np.int32)
    getattr(lib(), f"orc_{_sfx(dtype)}_sddmm")(A.m, _p(ptr), _p(ind), _p(lam), _p(x), _p(g))
    return g


def trsv_upper(H, s):
    H = np.asfortranarray(H, np.float64)
    s = np.array(s, np.float64, copy=True)
    n = s.size
    lib().orc_trsv_upper(n, H.ctypes.data_as(ctypes.c_void_p), H.shape[0], _p(s))
    return s


def genrot(a, b):
    a_, b_, c_, s_ = (ctypes.c_double(a), ctypes.c_double(b), ctypes.c_double(), ctypes.c_double())
    lib().orc_genrot(ctypes.byref(a_), ctypes.byref(b_), ctypes.byref(c_), ctypes.byref(s_))
    return a_.value, b_.value, c_.value, s_.value


# ------------------------------------------------------------ reference run --

REF_DIR = os.path.join(HERE, "_ref")


def ref_available(kind="generic"):
    return os.path.exists(os.path.join(REF_DIR, f"seldon_ref_{kind}"))


def write_seldon_csr(path, A, dtype=np.float64):
    """Seldon binary CSR: int m, n, nz; ptr[m+1]; ind[nz]; data[nz]
    (matrix_sparse/Matrix_Sparse.cxx:1532-1556)."""
    with open(path, "wb") as f:
        np.array([A.m, A.n, A.nnz], np.int32).tofile(f)
        np.ascontiguousarray(A.ptr, np.int32).tofile(f)
        np.ascontiguousarray(A.ind, np.int32).tofile(f)
        np.ascontiguousarray(A.val, dtype).tofile(f)


def write_seldon_vec(path, x):
    """Seldon binary vector: int m; T[m] (vector/Vector.cxx:846-874)."""
    with open(path, "wb") as f:
        np.array([x.size], np.int32).tofile(f)
        np.ascontiguousarray(x).tofile(f)


def read_seldon_csr(path, dtype=np.float64):
    """(m, n, ptr, ind, val) from a Seldon binary CSR."""
    with open(path, "rb") as f:
        m, n, nz = np.fromfile(f, np.int32, 3)
        ptr = np.fromfile(f, np.int32, m + 1)
        ind = np.fromfile(f, np.int32, nz)
        val = np.fromfile(f, dtype, nz)
    return int(m), int(n), ptr, ind, val


def read_seldon_vec(path, dtype=np.float64):
    with open(path, "rb") as f:
        m = int(np.fromfile(f, np.int32, 1)[0])
        return np.fromfile(f, dtype, m)


def run_ref(args, kind="generic", cwd=None):
    exe = os.path.join(REF_DIR, f"seldon_ref_{kind}")
    env = dict(os.environ, MKL_NUM_THREADS="1", MKL_CBWR="COMPATIBLE")
    out = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True,
                         cwd=cwd, env=env, check=True)
    return out.stdout


def parse_ref_history(stdout):
    hist, its, code, nb = [], [], None, None
    for line in stdout.splitlines():
        if line.startswith("Residu at iteration number"):
            parts = line.split()
            its.append(int(parts[4])); hist.append(float(parts[5]))
        elif line.startswith("ERRCODE"):
            parts = line.split()
            code, nb = int(parts[1]), int(parts[3])
    return np.array(hist), np.array(its, np.int32), code, nb


# ------------------------------------------------ adjoint reference (FP64) --

def sddmm_sym_fold(U, lam, x):
    """RowSymSparse gradient: stored upper (i, c) -> -(lam_i x_c) + -(lam_c x_i)
    (diagonal: first term only), from the plain per-entry products."""
    g_up = sddmm(U, lam, x)
    rows = np.repeat(np.arange(U.m), np.diff(U.ptr))
    g_lo = -(lam[U.ind] * x[rows])
    off = U.ind != rows
    g = g_up.copy()
    g[off] = g_up[off] + g_lo[off]
    return g


def diffusion_grad_reference(n, theta, values, pattern, b, x_obs, tol=1e-13, max_iter=20000):
    """dL/dtheta and dL/db for L = |x - x_obs|^2, x = A(theta)^{-1} b, with the
    oracle's FP64 Cg (Jacobi) for the forward and adjoint solves, the oracle's
    SDDMM for dL/dA, and the chain rule through the harmonic-mean faces written
    out in numpy (independent of the torch assembly)."""
    import copy
    A = copy.copy(pattern)
    A.val = np.ascontiguousarray(values, np.float64)
    fx = solve("cg", A, b, precond="jacobi", tol=tol, max_iter=max_iter)
    x = fx.x
    gx = 2.0 * (x - x_obs)
    ad = solve("cg", A, gx, precond="jacobi", tol=tol, max_iter=max_iter)
    lam = ad.x
    G = sddmm(A, lam, x)
    N = n ** 3
    ptr = A.ptr.astype(np.int64)
    rows = np.repeat(np.arange(N), np.diff(ptr))
    key = rows * N + A.ind.astype(np.int64)
    order = np.argsort(key)
    ks = key[order]

    def pos(i, j):
        return order[np.searchsorted(ks, i * N + j)]

    k = np.exp(theta)
    r = np.arange(N)
    ix, iy, iz = r % n, (r // n) % n, r // (n * n)
    dk = np.zeros(N)
    gd = G[pos(r, r)]
    for f, (cond, off) in enumerate(((iz > 0, -n * n), (iy > 0, -n), (ix > 0, -1),
                                     (ix < n - 1, 1), (iy < n - 1, n), (iz < n - 1, n * n))):
        c = r[cond]
        nb = c + off
        kc, kn = k[c], k[nb]
        dLdkf = gd[c] + gd[nb] - G[pos(c, nb)] - G[pos(nb, c)]
        dk[c] += dLdkf * 2.0 * kn * kn / ((kc + kn) ** 2)
        bnd = r[~cond]
        dk[bnd] += gd[bnd]  # boundary face: k_f = k_c
    return dict(x=x, lam=lam, grad_theta=k * dk, grad_b=lam, loss=float(np.sum((x - x_obs) ** 2)),
                its=(fx.nb_iter, ad.nb_iter))


# ---- Harwell-Boeing files (matrix_sparse/IOMatrixMarket.cxx) --------------
def _libc():
    c = ctypes.CDLL(None)
    c.strtof.restype = ctypes.c_float
    c.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    c.strtod.restype = ctypes.c_double
    c.strtod.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    return c


def hb_write_text(A, storage, dtype=np.float64) -> bytes:
    """WriteHarwellBoeing(A, name) (IOMatrixMarket.cxx:397-510) for the
    stored arrays of A: ConvertToCSC's order (row-major coordinates, the
    symmetric storage mirrored after each off-diagonal entry,
    Matrix_Conversions.cxx:113-160, then MergeSort by column -- stable;
    Functions_Arrays.cxx:412-480), then Seldon's left-aligned layout."""
    import math
    rows, cols, vals = [], [], []
    val = np.asarray(A.val, dtype)
    for i in range(A.m):
        for k in range(A.ptr[i], A.ptr[i + 1]):
            c = int(A.ind[k])
            rows.append(i); cols.append(c); vals.append(val[k])
            if storage == "sym" and c != i:
                rows.append(c); cols.append(i); vals.append(val[k])
    nnz, m, n = len(rows), A.m, A.n
    key = (lambda e: (cols[e], rows[e])) if storage == "sym" else (lambda e: cols[e])
    order = sorted(range(nnz), key=key)  # Python's sort is stable
    cptr = np.zeros(n + 1, np.int64)
    np.add.at(cptr, np.asarray(cols, np.int64) + 1, 1)
    cptr = np.cumsum(cptr)
    ptr_len = int(math.ceil(math.log10(0.1 + nnz + 1))) + 1
    ptr_np = min(80 // ptr_len, n + 1)
    ind_len = int(math.ceil(math.log10(0.1 + n + 1))) + 1
    ind_np = min(80 // ind_len, nnz)
    ptrcrd, indcrd, valcrd = n // ptr_np + 1, (nnz - 1) // ind_np + 1, (nnz - 1) // 3 + 1
    out = [f"{'Seldon':<72}{'S0000008':<8}\n",
           "".join(f"{x:<14d}" for x in (ptrcrd + indcrd + valcrd, ptrcrd, indcrd, valcrd, 0)) + "\n",
           ("RUA" if m == n else "RRA") + " " * 11 + "".join(f"{x:<14d}" for x in (m, n, nnz, 0)) + "\n",
           f"{'(%dI%d)' % (ptr_np, ptr_len):<16}{'(%dI%d)' % (ind_np, ind_len):<16}{'(3D23.15)':<20}"
           f"{'(3D23.15)':<20}\n"]
    for i in range(0, n + 1, ptr_np):
        out.append("".join(f"{int(cptr[j]) + 1:<{ptr_len}d}" for j in range(i, min(i + ptr_np, n + 1))) + "\n")
    for i in range(0, nnz, ind_np):
        out.append("".join(f"{rows[order[j]] + 1:<{ind_len}d}" for j in range(i, min(i + ind_np, nnz))) + "\n")
    for i in range(0, nnz, 3):
        out.append("".join(f"{float(vals[order[j]]):<23.15e}" for j in range(i, min(i + 3, nnz))) + "\n")
    return "".join(out).encode()


def _istream_int(s):
    t = s.lstrip(" \t\n\r\v\f")
    j = 1 if t[:1] in "+-" and t[:1] != "" else 0
    k = j
    while k < len(t) and t[k].isdigit():
        k += 1
    if k == j:
        raise ValueError(f"no integer in {s!r}")
    return int(t[:k])


def _istream_value(s, dtype):
    import re
    t = s.lstrip(" \t\n\r\v\f")
    mt = re.match(r"[+-]?(\d*)(\.\d*)?", t)
    if not mt or not (mt.group(1) or (mt.group(2) or "")[1:]):
        raise ValueError(f"no value in {s!r}")
    tok = mt.group(0)
    me = re.match(r"[eE][+-]?\d+", t[len(tok):])
    if t[len(tok):len(tok) + 1] in ("e", "E") and not me:
        raise ValueError(f"exponent without digits in {s!r}")
    if me:
        tok += me.group(0)
    c = _libc()
    return dtype((c.strtod if dtype == np.float64 else c.strtof)(tok.encode(), None))


def hb_read(text: bytes, dtype=np.float64):
    """ReadHarwellBoeing(name, A) (IOMatrixMarket.cxx:80-360) into a
    ColSparse matrix, then ConvertMatrix_to_Coordinates and
    ConvertMatrix_from_Coordinates into a RowSparse matrix of the same size:
    returns (m, n, ptr, ind, val) of the latter."""
    lines = text.decode().split("\n")
    if len(lines[0]) < 72:
        raise ValueError("title line shorter than 72 characters")
    pos = 1
    toks = []
    while len(toks) < 10:  # formatted extraction over lines 2 and 3
        toks += lines[pos].split()
        pos += 1
    nl_ptr, nl_ind, nl_val, nl_rhs = (int(x) for x in toks[1:5])
    typ, nrow, ncol, nnz, nelt = toks[5], int(toks[6]), int(toks[7]), int(toks[8]), int(toks[9])
    if len(typ) != 3 or typ[0] not in "RCP" or typ[0] == "C" or typ[1] in "HZS" or nelt != 0 or typ[2] == "E":
        raise ValueError(f"unsupported matrix type {typ}")
    fmt = lines[pos]
    pos += 1
    import re
    lw_ptr, ew_ptr = (int(x) for x in re.match(r"\((\d+)I(\d+)\)", fmt[0:16]).groups())
    lw_ind, ew_ind = (int(x) for x in re.match(r"\((\d+)I(\d+)\)", fmt[16:32]).groups())
    if typ[0] != "P":
        element = fmt[32:48]
        vect = []
        for tok in re.split(r"[ (ABEFGILOPZ*.)]+", element):
            if tok:
                try:
                    vect.append(_istream_int(tok))
                except ValueError:
                    vect.append(0)
        if len(vect) < 3:
            raise ValueError(f"value format {element!r} could not be parsed")
        lw_val, ew_val = vect[-3], vect[-2]
    if nl_rhs != 0:
        pos += 1

    def fields(count, per_line, width, parse):
        nonlocal pos
        out = []
        while len(out) < count:
            line = lines[pos]
            pos += 1
            k = 0
            for _ in range(per_line):
                if len(out) == count:
                    break
                if k > len(line):
                    raise ValueError("missing field")
                out.append(parse(line[k:k + width]))
                k += width
        return out

    cptr = np.array(fields(ncol + 1, lw_ptr, ew_ptr, _istream_int), np.int64) - 1
    rind = np.array(fields(nnz, lw_ind, ew_ind, _istream_int), np.int64) - 1
    if typ[0] == "P":
        val = np.ones(nnz, dtype)
    else:
        val = np.array(fields(nnz, lw_val, ew_val, lambda s: _istream_value(s, dtype)), dtype)
    ccol = np.repeat(np.arange(ncol), np.diff(cptr))
    m, n, ptr, ind, vv = coo_to_csr(rind.astype(np.int32), ccol.astype(np.int32), val, 0, nrow, ncol, dtype)
    return m, n, ptr, ind, vv
