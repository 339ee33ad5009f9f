"""CPU checks of the native boundary: libshk.so loads and exports every entry
point include/shk.h declares; calls fail cleanly without a GPU; the seeded
workload generators are deterministic and have the stated shapes."""
import ctypes
import os
import re

import numpy as np

from seldon_amd import _lib as L
from seldon_amd import workloads as wl
from tests.conftest import ROOT


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = _declared("shk.h")
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(L.exported_symbols()) | {"shk_ctx_device_info"}


def test_binding_loads_and_reports_no_device_cleanly():
    import torch
    lib = L.load()
    assert lib.shk_version().startswith(b"shk")
    if torch.cuda.is_available():
        return
    h = ctypes.c_void_p()
    st = lib.shk_ctx_create(0, ctypes.byref(h))
    assert st == -5  # SHK_ERR_NODEV
    assert b"device" in lib.shk_last_error()
    assert lib.shk_cg(None, None, None, None, 0, None, None) == -1  # SHK_ERR_ARG


def test_generators_shapes_and_determinism():
    A = wl.laplace2d(16)
    assert A.m == 256 and A.nnz == 5 * 256 - 4 * 16
    C = wl.convdiff3d(8)
    assert C.nnz == 7 * 512 - 6 * 64
    assert np.all(np.diff(C.ptr) >= 4)
    P1 = wl.powerlaw(50000, 9, half_width=20000)
    P2 = wl.powerlaw(50000, 9, half_width=20000)
    assert np.array_equal(P1.ind, P2.ind) and np.array_equal(P1.val, P2.val)
    lens = np.diff(P1.ptr)
    assert lens.min() >= 4 and lens.max() <= 512
    assert 24 < lens.mean() < 30
    rows = np.repeat(np.arange(P1.m), lens)
    assert np.all(np.abs(P1.ind - rows) <= 20000)
    d = P1.diagonal()
    off = np.abs(P1.val).copy()
    off[P1.ind == rows] = 0
    asum = np.bincount(rows, weights=off, minlength=P1.m)
    assert np.allclose(d, 1.1 * asum, rtol=1e-12)
    for i in range(0, P1.m, 997):  # sorted, distinct columns
        seg = P1.ind[P1.ptr[i]:P1.ptr[i + 1]]
        assert np.all(np.diff(seg) > 0)


def test_upper_is_triu():
    A = wl.convdiff3d(5)
    U = wl.upper(A)
    import scipy.sparse as sp
    ref = sp.triu(A.to_scipy()).tocsr()
    assert np.array_equal(U.ptr, ref.indptr) and np.array_equal(U.ind, ref.indices)
    assert np.array_equal(U.val, ref.data)


def test_diffusion_is_symmetric_positive():
    th = wl.normal(6 ** 3, 3, 0.0, 0.5)
    A = wl.diffusion3d(6, th).to_scipy()
    assert abs(A - A.T).max() == 0
    assert np.linalg.eigvalsh(A.toarray()).min() > 0


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and a message,
    before touching a device."""
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2, r.stderr
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
