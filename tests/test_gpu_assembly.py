"""GPU parity of the CSR assembly from coordinates (Seldon's
ConvertMatrix_from_Coordinates for RowSparse, Matrix_Conversions.cxx:
1288-1335) through the C ABI: bitwise against the real-Seldon fixtures and
against the oracle on larger inputs, then the assembled matrix used in place
(created from the device arrays) for a bitwise SpMV."""
import numpy as np
import pytest
import torch

from seldon_amd import workloads as wl
from tests.conftest import load_fixture
from tests.golden.make_golden import COO_CASES, coo_case

pytestmark = pytest.mark.gpu


def _run(row, col, val, index, m_min, n_min, tdt):
    from seldon_amd import krylov as K
    dev = lambda a, dt: torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")
    return K.ConvertMatrix_from_Coordinates(dev(row, torch.int32), dev(col, torch.int32), dev(val, tdt),
                                            index=index, m=m_min, n=n_min)


@pytest.mark.parametrize("name", sorted(COO_CASES))
@pytest.mark.parametrize("c", ["d", "f"])
def test_coo_to_csr_golden_fixture(gpu, name, c):
    fx = load_fixture("coo_to_csr.npz")
    tdt, ndt = (torch.float64, np.float64) if c == "d" else (torch.float32, np.float32)
    row, col, val, index, m_min, n_min = coo_case(name)
    m, n, ptr, ind, out = _run(row, col, val.astype(ndt), index, m_min, n_min, tdt)
    assert [m, n] == list(fx[f"{name}_{c}_mn"])
    assert np.array_equal(ptr.cpu().numpy(), fx[f"{name}_{c}_ptr"])
    assert np.array_equal(ind.cpu().numpy(), fx[f"{name}_{c}_ind"])
    assert np.array_equal(out.cpu().numpy(), fx[f"{name}_{c}_val"])


def test_coo_to_csr_large_vs_oracle_then_spmv(gpu):
    """A shuffled power-law matrix with 10% duplicated coordinates, assembled
    on the device, equals the oracle's conversion; the matrix built from the
    device arrays multiplies bitwise like the oracle's."""
    from oracle import oracle as orc
    from seldon_amd import krylov as K
    A = wl.powerlaw(200000, 12, half_width=20000)
    row = np.repeat(np.arange(A.m, dtype=np.int32), np.diff(A.ptr))
    col = A.ind.astype(np.int32)
    val = A.val.copy()
    rng = np.random.default_rng(5)
    k = len(row) // 10
    where = rng.integers(0, len(row), k)
    pick = rng.integers(0, len(row), k)
    row = np.concatenate([row, row[pick]])
    col = np.concatenate([col, col[pick]])
    val = np.concatenate([val, rng.standard_normal(k)])
    perm = rng.permutation(len(row))
    row, col, val = row[perm], col[perm], val[perm]
    print(f"[coo] {len(row)} entries", flush=True)
    m, n, ptr, ind, out = _run(row, col, val, 0, 0, 0, torch.float64)
    rm, rn, rptr, rind, rval = orc.coo_to_csr(row, col, val)
    assert (m, n) == (rm, rn)
    assert np.array_equal(ptr.cpu().numpy(), rptr)
    assert np.array_equal(ind.cpu().numpy(), rind)
    assert np.array_equal(out.cpu().numpy(), rval)
    Ag = K.Matrix(K.DeviceCSR(m, n, ptr, ind, out))
    x = wl.uniform(n, 3)
    xg = torch.tensor(x, device="cuda")
    yg = torch.empty(m, dtype=torch.float64, device="cuda")
    K.Mlt(Ag, xg, yg)
    ref = orc.mlt(wl.CSR(rm, rn, rptr, rind, rval), x)
    assert np.array_equal(yg.cpu().numpy(), ref)


def test_coo_to_csr_errors(gpu):
    from seldon_amd import _lib as L
    from seldon_amd import krylov as K
    row = torch.tensor([0, 1, 2], dtype=torch.int32, device="cuda")
    col = torch.tensor([0, 5, 1], dtype=torch.int32, device="cuda")
    val = torch.ones(3, dtype=torch.float64, device="cuda")
    with pytest.raises(L.ShkError):  # index 1 makes row 0 negative
        K.ConvertMatrix_from_Coordinates(row, col, val, index=1)
    import ctypes
    mm, nn = ctypes.c_int(), ctypes.c_int()
    ptr = torch.empty(2, dtype=torch.int32, device="cuda")
    ind = torch.empty(3, dtype=torch.int32, device="cuda")
    out = torch.empty(3, dtype=torch.float64, device="cuda")
    st = L.load().shk_coo_to_csr(K.Context.default().h, 0, 3, K._ptr(row), K._ptr(col), K._ptr(val), 0, 0, 0, 1,
                                 ctypes.byref(mm), ctypes.byref(nn), K._ptr(ptr), K._ptr(ind), K._ptr(out))
    assert st != 0 and (mm.value, nn.value) == (3, 6)  # the row pointer is too short: m = 3


@pytest.mark.parametrize("layout", ["rows", "auto"])
@pytest.mark.parametrize("st", ["row", "sym", "array"])
def test_csr_binary_file_roundtrip(gpu, tmp_path, st, layout):
    """Seldon's binary CSR file (Matrix_Sparse.cxx:1532-1556, 1649-1682):
    the device matrix writes the caller's arrays byte for byte as Seldon's
    layout has them, and a matrix read back multiplies bitwise the same."""
    from oracle import oracle as orc
    from seldon_amd import krylov as K
    A = wl.powerlaw(30000, 8, kmin=2, kmax=200, half_width=4000)
    M = wl.upper(A) if st == "sym" else A
    Ag = K.Matrix(M, storage=st, layout=layout)
    Ag.Write(tmp_path / "dev.bin")
    orc.write_seldon_csr(tmp_path / "ref.bin", M)
    assert (tmp_path / "dev.bin").read_bytes() == (tmp_path / "ref.bin").read_bytes()
    Bg = K.Matrix.Read(tmp_path / "ref.bin", storage=st, layout=layout)
    x = torch.tensor(wl.uniform(A.n, 4), device="cuda")
    y1 = torch.empty(A.m, dtype=torch.float64, device="cuda")
    y2 = torch.empty_like(y1)
    K.Mlt(Ag, x, y1)
    K.Mlt(Bg, x, y2)
    assert torch.equal(y1, y2)
    assert np.array_equal(y1.cpu().numpy(), orc.mlt(M, x.cpu().numpy(), storage=st))


def test_csr_read_file_written_by_seldon(gpu, tmp_path):
    """A CSR file written by the real Seldon (its ConvertMatrix_from_Coordinates
    then Matrix::Write, recorded in tests/golden/seldon_files.npz) reads into
    the same matrix the device assembles."""
    from seldon_amd import krylov as K
    fx = load_fixture("seldon_files.npz")
    row, col, val, index, m_min, n_min = coo_case("dup")
    (tmp_path / "A.bin").write_bytes(fx["csr_dup_d"].tobytes())
    Ag = K.Matrix.Read(tmp_path / "A.bin")
    m, n, ptr, ind, out = _run(row, col, val, index, m_min, n_min, torch.float64)
    assert (Ag.m, Ag.n) == (m, n)
    Bg = K.Matrix(K.DeviceCSR(m, n, ptr, ind, out))
    Ag.Write(tmp_path / "a2.bin")
    Bg.Write(tmp_path / "b2.bin")
    assert (tmp_path / "a2.bin").read_bytes() == (tmp_path / "b2.bin").read_bytes() == (tmp_path / "A.bin").read_bytes()


@pytest.mark.parametrize("c", ["d", "f"])
def test_vector_file_roundtrip(gpu, c, tmp_path):
    """Seldon's binary vector file (Vector.cxx:846-874, :1060-1091) from a
    device vector: the same bytes as the oracle's writer (whose layout the
    real-Seldon fixtures pin), read back exactly, with and without the
    length; a vector the real Seldon wrote reads back the same."""
    import ctypes

    from oracle import oracle as orc
    from seldon_amd import _lib as L
    from seldon_amd import krylov as K
    tdt, ndt = (torch.float64, np.float64) if c == "d" else (torch.float32, np.float32)
    v = wl.uniform(12345, 17).astype(ndt)
    X = torch.tensor(v, device="cuda")
    K.WriteVector(X, tmp_path / "v.bin")
    orc.write_seldon_vec(tmp_path / "o.bin", v)
    assert (tmp_path / "v.bin").read_bytes() == (tmp_path / "o.bin").read_bytes()
    Y = K.ReadVector(tmp_path / "v.bin", dtype=tdt)
    assert Y.cpu().numpy().tobytes() == v.tobytes()
    K.WriteVector(X, tmp_path / "n.bin", with_size=False)
    assert (tmp_path / "n.bin").read_bytes() == v.tobytes()
    Z = K.ReadVector(tmp_path / "n.bin", dtype=tdt, with_size=False, n=v.size)
    assert Z.cpu().numpy().tobytes() == v.tobytes()
    K.WriteVector(torch.empty(0, dtype=tdt, device="cuda"), tmp_path / "e.bin")
    assert (tmp_path / "e.bin").read_bytes() == np.int32(0).tobytes()
    assert K.ReadVector(tmp_path / "e.bin", dtype=tdt).numel() == 0
    small = torch.empty(10, dtype=tdt, device="cuda")
    got = ctypes.c_int64()
    with pytest.raises(RuntimeError, match="longer than the destination"):
        L.check(L.load().shk_vec_read(K.Context.default().h, K._DT[tdt], str(tmp_path / "v.bin").encode(), 1, 10,
                                      K._ptr(small), ctypes.byref(got)))
    # a file the real Seldon's Vector::Write wrote (tests/golden/seldon_files.npz:
    # uniform(1000, 17) / 2, through its SOR driver)
    fx = load_fixture("seldon_files.npz")
    raw = fx[f"vec_{c}"].tobytes()
    (tmp_path / "s.bin").write_bytes(raw)
    S = K.ReadVector(tmp_path / "s.bin", dtype=tdt)
    assert np.int32(S.numel()).tobytes() + S.cpu().numpy().tobytes() == raw
    n1 = 1000
    D = wl.CSR(n1, n1, np.arange(n1 + 1, dtype=np.int32), np.arange(n1, dtype=np.int32), np.full(n1, 2.0))
    want = orc.sor(D, np.zeros(n1, ndt), wl.uniform(n1, 17).astype(ndt), 1.0, 1, 2, dtype=ndt)
    assert np.array_equal(S.cpu().numpy(), want)
    K.WriteVector(S, tmp_path / "s2.bin")
    assert (tmp_path / "s2.bin").read_bytes() == raw
