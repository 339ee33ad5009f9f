"""GPU parity of the SpMV path (MltAdd / Mlt on RowSparse, RowSymSparse,
ArrayRowSparse) through the C ABI: bitwise against the real-Seldon fixtures
and against the oracle at larger sizes (row sums are sequential in stored
order on the GPU too, with no FMA contraction)."""
import numpy as np
import pytest
import torch

from seldon_amd import workloads as wl
from tests.conftest import load_fixture

pytestmark = pytest.mark.gpu

DT = {"d": (torch.float64, np.float64), "f": (torch.float32, np.float32)}


def _dev(a, npdt):
    return torch.tensor(np.ascontiguousarray(a, npdt), device="cuda")


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("layout", ["rows", "auto"])
@pytest.mark.parametrize("st", ["row", "sym", "array"])
@pytest.mark.parametrize("c", ["d", "f"])
def test_spmv_golden_fixture(gpu, st, c, layout):
    from seldon_amd import krylov as K
    fx = load_fixture("spmv_powerlaw300.npz")
    sp = fx["spec"]
    A = wl.powerlaw(sp["n"], sp["seed"], kmin=sp["kmin"], kmax=sp["kmax"], half_width=sp["half_width"])
    M = wl.upper(A) if st == "sym" else A
    tdt, ndt = DT[c]
    x = wl.uniform(A.m, sp["xseed"]); y = wl.uniform(A.m, sp["yseed"]); y[sp["nan_at"]] = np.nan
    Ag = K.Matrix(M, storage=st, dtype=tdt, layout=layout)
    xg = _dev(x, ndt)
    for k, (al, be) in enumerate(fx["alpha_beta"]):
        yg = _dev(y, ndt)
        K.MltAdd(al, Ag, xg, be, yg)
        got = yg.cpu().numpy()
        assert _same(got, fx[f"{st}_{c}_{k}"]), (st, c, al, be)
    yg = torch.full((A.m,), 7.0, dtype=tdt, device="cuda")
    K.Mlt(Ag, xg, yg)
    assert _same(yg.cpu().numpy(), fx[f"{st}_{c}_mlt"])


def _cases():
    yield "powerlaw20k", wl.powerlaw(20000, 3, half_width=3000)
    yield "convdiff24", wl.convdiff3d(24)
    yield "laplace64", wl.laplace2d(64)
    # rows longer than the LDS budget (2048 / 4096 entries) and empty rows
    A = wl.powerlaw(12000, 9, kmin=1, kmax=6000, expo=0.9, half_width=5000)
    yield "longrows", A
    ptr = A.ptr.copy()
    lens = np.diff(ptr)
    lens[::7] = 0
    ptr2 = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    keep = np.concatenate([np.arange(A.ptr[i], A.ptr[i] + lens[i]) for i in range(A.m)])
    yield "emptyrows", wl.CSR(A.m, A.n, ptr2, A.ind[keep].copy(), A.val[keep].copy())


@pytest.mark.parametrize("layout", ["rows", "auto"])
@pytest.mark.parametrize("st", ["row", "sym", "array"])
@pytest.mark.parametrize("c", ["d", "f"])
def test_spmv_bitwise_vs_oracle(gpu, st, c, layout):
    from oracle import oracle as orc
    from seldon_amd import krylov as K
    tdt, ndt = DT[c]
    for name, A in _cases():
        M = wl.upper(A) if st == "sym" else A
        x = wl.uniform(A.n, 21)
        y = wl.uniform(A.m, 22)
        Ag = K.Matrix(M, storage=st, dtype=tdt, layout=layout)
        if layout == "auto":
            # banded / stencil rows with sorted columns take the panel layout
            # (RowSymSparse: upper parts, then transposed parts)
            assert Ag.layout == "panels", name
        print(f"[spmv] {name} {st} {c} {Ag.layout}", flush=True)
        xg = _dev(x, ndt)
        for al, be in ((1.0, 0.0), (-1.0, 1.0), (0.7, -0.3)):
            ref = orc.mltadd(ndt(al), M, x.astype(ndt), ndt(be), y.astype(ndt), storage=st, dtype=ndt)
            yg = _dev(y, ndt)
            K.MltAdd(al, Ag, xg, be, yg)
            assert _same(yg.cpu().numpy(), ref), (name, st, c, al, be)
        ref = orc.mlt(M, x.astype(ndt), storage=st, dtype=ndt)
        yg = torch.empty(A.m, dtype=tdt, device="cuda")
        K.Mlt(Ag, xg, yg)
        assert _same(yg.cpu().numpy(), ref), (name, st, c, "mlt")


def test_spmv_closed_form_laplacian(gpu):
    """5-point Laplacian eigenpairs: A u_jk = lambda_jk u_jk with
    lambda = 4 - 2cos(j pi/(g+1)) - 2cos(k pi/(g+1)), u = sin x sin."""
    from seldon_amd import krylov as K
    g = 64
    A = wl.laplace2d(g)
    i = np.arange(1, g + 1)
    for (j, k) in ((1, 1), (3, 7), (g, g - 1)):
        u = np.outer(np.sin(k * i * np.pi / (g + 1)), np.sin(j * i * np.pi / (g + 1))).ravel()
        lam = 4 - 2 * np.cos(j * np.pi / (g + 1)) - 2 * np.cos(k * np.pi / (g + 1))
        Ag = K.Matrix(A)
        ug = _dev(u, np.float64)
        yg = torch.empty_like(ug)
        K.Mlt(Ag, ug, yg)
        assert np.allclose(yg.cpu().numpy(), lam * u, rtol=0, atol=1e-12)


def test_spmv_empty_matrix(gpu):
    from seldon_amd import krylov as K
    A = wl.CSR(0, 0, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    Ag = K.Matrix(A)
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    K.Mlt(Ag, x, x)


def test_mlt_dot_fused(gpu):
    from oracle import oracle as orc
    from seldon_amd import krylov as K
    A = wl.convdiff3d(20)
    x = wl.uniform(A.m, 5)
    Ag = K.Matrix(A)
    xg = _dev(x, np.float64)
    yg = torch.empty_like(xg)
    d = K.MltDot(Ag, xg, yg)
    y = orc.mlt(A, x)
    assert np.array_equal(yg.cpu().numpy(), y)
    assert abs(d - orc.dot(x, y)) <= 1e-13 * abs(orc.dot(np.abs(x), np.abs(y)))


def test_panel_layout_choice_and_fallback(gpu):
    """Rows with unsorted columns, or columns spread far from the diagonal,
    keep the row layout (and refuse an explicit panel request); results stay
    bitwise either way."""
    from oracle import oracle as orc
    from seldon_amd import _lib as L
    from seldon_amd import krylov as K
    A = wl.laplace2d(40)
    # reverse the columns of every third row: Seldon's row sum order follows
    # the stored order, which the panel walk could not keep
    ind, val = A.ind.copy(), A.val.copy()
    for r in range(0, A.m, 3):
        a, b = A.ptr[r], A.ptr[r + 1]
        ind[a:b] = ind[a:b][::-1].copy()
        val[a:b] = val[a:b][::-1].copy()
    U = wl.CSR(A.m, A.n, A.ptr.copy(), ind, val)
    Ug = K.Matrix(U)
    assert Ug.layout == "rows"
    with pytest.raises(L.ShkError):
        K.Matrix(U, layout="panels")
    x = wl.uniform(A.n, 4)
    yg = torch.empty(A.m, dtype=torch.float64, device="cuda")
    K.Mlt(Ug, _dev(x, np.float64), yg)
    assert _same(yg.cpu().numpy(), orc.mlt(U, x))
    # columns uniform over the whole range: x windows would cost more than
    # the entries, so AUTO keeps the row layout
    R = wl.powerlaw(100000, 5, kmin=4, kmax=40, half_width=100000)
    assert K.Matrix(R).layout == "rows"
    assert K.Matrix(R, dtype=torch.float32).layout == "rows"


@pytest.mark.parametrize("c", ["d", "f"])
def test_panel_rectangular_and_values_update(gpu, c):
    """Panels on a rectangular matrix, m not a multiple of the panel height,
    then new values through set_values (the device scatter into the panel
    arrays) -- bitwise against the oracle both times."""
    from oracle import oracle as orc
    from seldon_amd import krylov as K
    tdt, ndt = DT[c]
    A = wl.powerlaw(5000, 11, kmin=2, kmax=300, half_width=700)
    keep = A.ind < 4200
    lens = np.array([keep[A.ptr[i]:A.ptr[i + 1]].sum() for i in range(A.m)])
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    R = wl.CSR(A.m, 4200, ptr, A.ind[keep].copy(), A.val[keep].copy())
    Rg = K.Matrix(R, dtype=tdt, layout="panels")
    x = wl.uniform(R.n, 6)
    xg = _dev(x, ndt)
    yg = torch.empty(R.m, dtype=tdt, device="cuda")
    K.Mlt(Rg, xg, yg)
    assert _same(yg.cpu().numpy(), orc.mlt(R, x.astype(ndt), dtype=ndt))
    v2 = wl.normal(R.nnz, 8)
    R2 = wl.CSR(R.m, R.n, R.ptr, R.ind, v2)
    Rg.set_values(_dev(v2, ndt))
    K.Mlt(Rg, xg, yg)
    assert _same(yg.cpu().numpy(), orc.mlt(R2, x.astype(ndt), dtype=ndt))
    Rg.set_values(v2.astype(ndt))  # host source
    K.Mlt(Rg, xg, yg)
    assert _same(yg.cpu().numpy(), orc.mlt(R2, x.astype(ndt), dtype=ndt))


# ------------------------------------------------ transposed products --

@pytest.mark.parametrize("layout", ["rows", "auto"])
@pytest.mark.parametrize("st", ["row", "sym", "array"])
@pytest.mark.parametrize("c", ["d", "f"])
def test_spmv_trans_golden_fixture(gpu, st, c, layout):
    """MltAdd(alpha, SeldonTrans, ..) and Mlt(SeldonTrans, ..) bitwise against
    the real Seldon build, on a 300 x 220 matrix (RowSymSparse: square)."""
    from seldon_amd import krylov as K
    fx = load_fixture("spmvt_rect300x220.npz")
    sp = fx["spec"]
    A = wl.powerlaw(sp["n"], sp["seed"], kmin=sp["kmin"], kmax=sp["kmax"], half_width=sp["half_width"])
    M = wl.upper(A) if st == "sym" else wl.first_columns(A, sp["ncols"])
    tdt, ndt = DT[c]
    x = wl.uniform(sp["n"], sp["xseed"])
    y = wl.uniform(A.m if st == "sym" else sp["ncols"], sp["yseed_sym"] if st == "sym" else sp["yseed_rect"])
    y[sp["nan_at"]] = np.nan
    Ag = K.Matrix(M, storage=st, dtype=tdt, layout=layout)
    xg = _dev(x, ndt)
    for k, (al, be) in enumerate(fx["alpha_beta"]):
        yg = _dev(y, ndt)
        K.MltAdd(al, K.SeldonTrans, Ag, xg, be, yg)
        assert _same(yg.cpu().numpy(), fx[f"{st}_{c}_{k}"]), (st, c, al, be)
    yg = torch.full((len(y),), 7.0, dtype=tdt, device="cuda")
    K.Mlt(K.SeldonTrans, Ag, xg, yg)
    assert _same(yg.cpu().numpy(), fx[f"{st}_{c}_mlt"])


@pytest.mark.parametrize("layout", ["rows", "auto"])
@pytest.mark.parametrize("st", ["row", "sym", "array"])
def test_spmv_trans_bitwise_vs_oracle(gpu, st, layout):
    """Transposed products against the oracle on matrices with long and empty
    rows, a stencil, and a wide rectangular one (A^T is built on the first
    call, in the layout AUTO picks for it)."""
    from oracle import oracle as orc
    from seldon_amd import krylov as K
    cases = [(n, A) for n, A in _cases() if n in ("powerlaw20k", "convdiff24", "longrows", "emptyrows")]
    P = wl.powerlaw(9000, 4, kmin=1, kmax=300, half_width=2500)
    cases.append(("rect9000x6000", wl.first_columns(P, 6000)))
    for name, A in cases:
        if st == "sym" and A.m != A.n:
            continue
        M = wl.upper(A) if st == "sym" else A
        n_out = A.m if st == "sym" else A.n
        x = wl.uniform(A.m, 31)
        y = wl.uniform(n_out, 32)
        Ag = K.Matrix(M, storage=st, dtype=torch.float64, layout=layout)
        print(f"[spmvt] {name} {st} {Ag.layout}", flush=True)
        xg = _dev(x, np.float64)
        for al, be in ((1.0, 0.0), (-1.0, 1.0), (0.7, -0.3)):
            ref = orc.mltadd_trans(al, M, x, be, y, storage=st)
            yg = _dev(y, np.float64)
            K.MltAdd(al, K.SeldonTrans, Ag, xg, be, yg)
            assert _same(yg.cpu().numpy(), ref), (name, st, al, be)
        if A.m <= 12000:  # the oracle's Mlt(SeldonTrans) is Seldon's O(m n) loop
            yg = torch.empty(n_out, dtype=torch.float64, device="cuda")
            K.Mlt(K.SeldonTrans, Ag, xg, yg)
            assert _same(yg.cpu().numpy(), orc.mlt_trans(M, x, storage=st)), (name, st, "mlt")


def test_spmv_trans_follows_new_values(gpu):
    """A^T is rebuilt after new values; SeldonNoTrans is the plain product."""
    from oracle import oracle as orc
    from seldon_amd import krylov as K
    A = wl.first_columns(wl.powerlaw(5000, 6, kmin=2, kmax=50, half_width=800), 4000)
    x = wl.uniform(A.m, 41)
    Ag = K.Matrix(A)
    xg = _dev(x, np.float64)
    yg = torch.empty(A.n, dtype=torch.float64, device="cuda")
    K.Mlt(K.SeldonTrans, Ag, xg, yg)
    assert _same(yg.cpu().numpy(), orc.mltadd_trans(1.0, A, x, 0.0, np.zeros(A.n)))
    B = wl.CSR(A.m, A.n, A.ptr, A.ind, A.val * -1.5 + 0.25)
    Ag.set_values(B.val)
    K.MltAdd(2.0, K.SeldonTrans, Ag, xg, 0.0, yg)
    assert _same(yg.cpu().numpy(), orc.mltadd_trans(2.0, B, x, 0.0, np.zeros(A.n)))
    y2 = torch.empty(A.m, dtype=torch.float64, device="cuda")
    K.Mlt(K.SeldonNoTrans, Ag, _dev(wl.uniform(A.n, 42), np.float64), y2)
    assert _same(y2.cpu().numpy(), orc.mlt(B, wl.uniform(A.n, 42)))


# ---------------------------------------------------------------- SOR --

@pytest.mark.parametrize("layout", ["rows", "auto"])
def test_sor_golden_fixture(gpu, layout):
    """SOR / SSOR sweeps (level-scheduled) bitwise against the real Seldon:
    RowSparse, ArrayRowSparse, RowSymSparse; float values with a float and a
    double omega; scrambled rows, one hiding its diagonal from A(j, j)."""
    from seldon_amd import krylov as K
    from tests.golden.make_golden import sor_fixture_cases
    fx = load_fixture("sor.npz")
    mats = {}
    for key, A, st, ndt, wide, om, it, ty in sor_fixture_cases():
        tdt = torch.float64 if ndt == np.float64 else torch.float32
        mk = (id(A), st, tdt)
        if mk not in mats:
            mats[mk] = K.Matrix(A, storage=st, dtype=tdt, layout=layout)
        xg = _dev(wl.uniform(A.m, fx["spec"]["xseed"]), ndt)
        bg = _dev(wl.uniform(A.m, fx["spec"]["bseed"]), ndt)
        K.SOR(mats[mk], xg, bg, om, it, ty, omega_dtype=torch.float64 if wide else None)
        assert _same(xg.cpu().numpy(), fx[key]), key


@pytest.mark.parametrize("st", ["row", "array", "sym"])
def test_sor_bitwise_vs_oracle(gpu, st):
    """Larger matrices (a stencil with few long levels, a banded random one
    with many short ones), and new values dropping the level plan."""
    from oracle import oracle as orc
    from seldon_amd import krylov as K
    # levels of at most 1024 rows run in one work-group, larger ones get a
    # launch each: the 3-D matrices (middle levels of ~1700-3000 rows) mix both
    mats = (("laplace96", wl.laplace2d(96)), ("convdiff24", wl.convdiff3d(24)),
            ("powerlaw6k", wl.powerlaw(6000, 23, kmin=2, kmax=60, half_width=500)),
            ("convdiff64", wl.convdiff3d(64)))
    if st == "sym":
        mats = (("laplace96", wl.upper(wl.laplace2d(96))), ("aniso64", wl.upper(wl.aniso2d(64, 48, 0.01))),
                ("powerlaw6k", wl.upper(wl.powerlaw(6000, 23, kmin=2, kmax=60, half_width=500))),
                ("convdiff48", wl.upper(wl.convdiff3d(48))))
    for name, A in mats:
        Ag = K.Matrix(A, storage=st)
        x = wl.uniform(A.m, 61)
        b = wl.uniform(A.m, 62)
        print(f"[sor] {name} {st}", flush=True)
        xg, bg = _dev(x, np.float64), _dev(b, np.float64)
        K.SOR(Ag, xg, bg, 1.25, 2, 0)
        assert _same(xg.cpu().numpy(), orc.sor(A, x, b, 1.25, 2, 0, storage=st)), (name, st)
        B = wl.CSR(A.m, A.n, A.ptr, A.ind, A.val * 1.5, A.sym)
        Ag.set_values(B.val)
        xg = _dev(x, np.float64)
        K.SOR(Ag, xg, bg, 0.9, 1, 2)
        assert _same(xg.cpu().numpy(), orc.sor(B, x, b, 0.9, 1, 2, storage=st)), (name, st, "new values")
