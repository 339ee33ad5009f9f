"""GPU parity of Harwell-Boeing files (matrix_sparse/IOMatrixMarket.cxx)
through the C ABI: WriteHarwellBoeing of a device matrix byte for byte
against the text the real Seldon writes for the same arrays, and
ReadHarwellBoeing (into ColSparse, then to RowSparse through coordinates, as
a Seldon program does) against the arrays the real Seldon reads, then larger
matrices against the oracle."""
import os
import tempfile

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from seldon_amd import workloads as wl
from tests.conftest import load_fixture
from tests.golden.make_golden import hb_cases, hb_text_e

pytestmark = pytest.mark.gpu

DT = {"d": (torch.float64, np.float64), "f": (torch.float32, np.float32)}


def _written(A, st, tdt, layout):
    from seldon_amd import krylov as K
    with tempfile.TemporaryDirectory() as tmp:
        K.WriteHarwellBoeing(K.Matrix(A, storage=st, dtype=tdt, layout=layout), os.path.join(tmp, "o.rua"))
        return open(os.path.join(tmp, "o.rua"), "rb").read()


def _read(text, tdt, ndt, layout="auto"):
    """The device matrix read from text, as (m, n, ptr, ind, val) of its
    stored arrays (written back through Seldon's binary CSR layout)."""
    from seldon_amd import krylov as K
    with tempfile.TemporaryDirectory() as tmp:
        with open(os.path.join(tmp, "i.rua"), "wb") as f:
            f.write(text)
        M = K.ReadHarwellBoeing(os.path.join(tmp, "i.rua"), dtype=tdt, layout=layout)
        M.Write(os.path.join(tmp, "o.bin"))
        return M, orc.read_seldon_csr(os.path.join(tmp, "o.bin"), ndt)


@pytest.mark.parametrize("layout", ["rows", "auto"])
@pytest.mark.parametrize("c", ["d", "f"])
def test_hb_write_golden_fixture(gpu, c, layout):
    fx = load_fixture("harwell_boeing.npz")
    tdt, _ = DT[c]
    writers, _ = hb_cases()
    for name, (A, st) in writers.items():
        print(f"[hb write] {name} {c} {layout}", flush=True)
        assert _written(A, st, tdt, layout) == fx[f"w_{name}_{c}"].tobytes(), (name, c, layout)


@pytest.mark.parametrize("c", ["d", "f"])
def test_hb_read_golden_fixture(gpu, c):
    fx = load_fixture("harwell_boeing.npz")
    tdt, ndt = DT[c]
    _, readers = hb_cases()
    for name, text in readers.items():
        print(f"[hb read] {name} {c}", flush=True)
        M, (m, n, ptr, ind, val) = _read(text, tdt, ndt)
        assert [m, n] == list(fx[f"r_{name}_{c}_dims"]) and [M.m, M.n] == [m, n], name
        assert np.array_equal(ptr, fx[f"r_{name}_{c}_ptr"]) and np.array_equal(ind, fx[f"r_{name}_{c}_ind"]), name
        assert val.tobytes() == fx[f"r_{name}_{c}_val"].tobytes(), (name, c)


def test_hb_larger_vs_oracle_and_spmv(gpu):
    """A 40K-row power-law matrix: written text equal to the oracle's; the
    file read back (E values at 17 digits) equals the oracle's reading and
    multiplies bitwise like the original."""
    from seldon_amd import krylov as K
    A = wl.powerlaw(40000, 92, kmin=2, kmax=40, half_width=2000)
    assert _written(A, "row", torch.float64, "auto") == orc.hb_write_text(A, "row")
    S = A.to_scipy().tocsc()
    text = hb_text_e(S.shape[0], S.shape[1], S.indptr, S.indices, S.data, "R", (8, 10), (8, 10), (3, 26, 16))
    M, got = _read(text, torch.float64, np.float64)
    ref = orc.hb_read(text)
    assert got[:2] == ref[:2] and all(np.array_equal(x, y) for x, y in zip(got[2:], ref[2:]))
    assert np.array_equal(got[4], A.val)  # 17 significant digits: the values come back exactly
    x = wl.uniform(A.n, 93)
    y = torch.empty(A.m, dtype=torch.float64, device="cuda")
    K.Mlt(M, torch.tensor(x, device="cuda"), y)
    assert y.cpu().numpy().tobytes() == orc.mlt(A, x).tobytes()


def test_hb_errors(gpu):
    """Seldon's reader refuses its writer's D-format value line, and symmetric
    or complex types; the library raises the same way."""
    fx = load_fixture("harwell_boeing.npz")
    with pytest.raises(RuntimeError, match="could not be parsed"):
        _read(fx["w_powerlaw150_row_d"].tobytes(), torch.float64, np.float64)
    _, readers = hb_cases()
    text = readers["e4x20"].decode().split("\n")
    for typ, msg in (("RSA", "symmetric"), ("CUA", "complex")):
        t = list(text)
        t[2] = typ + t[2][3:]
        with pytest.raises(RuntimeError, match=msg):
            _read("\n".join(t).encode(), torch.float64, np.float64)
    with pytest.raises(RuntimeError, match="without entries"):
        Z = wl.CSR(4, 4, np.zeros(5, np.int32), np.zeros(0, np.int32), np.zeros(0))
        _written(Z, "row", torch.float64, "rows")
