"""shk_cg_noalloc (no device allocation: work memory in the library's fixed
scratch array) against shk_cg through the C ABI: same code, iteration count,
residual history and solution, bit for bit, on two calls in a row (the second
on a smaller system, so stale scratch contents from the first would show),
with and without an initial guess, FP64 and FP32; and a system too large for
the scratch is refused with an error instead of being solved."""
import numpy as np
import pytest
import torch

from seldon_amd import _lib as L
from seldon_amd import krylov as K
from seldon_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _run(fn, A, b, dtype, jac, x0):
    Ag = K.Matrix(A, dtype=dtype)
    bg = torch.tensor(b, dtype=dtype, device="cuda")
    xg = torch.tensor(x0, dtype=dtype, device="cuda") if x0 is not None else torch.full_like(bg, 7.0)
    it = K.Iteration(3000, 1e-10 if dtype == torch.float64 else 1e-5)
    it.HideMessages()
    if x0 is not None:
        it.SetInitGuess(False)
    M = K.JacobiPreconditioner() if jac else K.Preconditioner_Base()
    code = K._solve(fn, Ag, xg, bg, M, it)
    return code, it.nb_iter, it.history.copy(), xg.cpu().numpy()


CASES = [
    # (grid, dtype, Jacobi, seeded initial guess): each pair of consecutive
    # cases is a larger system followed by a smaller one
    (256, torch.float64, True, False),
    (64, torch.float64, False, True),
    (200, torch.float32, False, False),
    (48, torch.float32, True, True),
]


def test_cg_noalloc_matches_cg(gpu):
    lib = L.load()
    for grid, dtype, jac, guess in CASES:
        print(f"[no-alloc] grid {grid} {dtype} jacobi={jac} guess={guess}", flush=True)
        A, b, _ = wl.cg_laplace(grid)
        x0 = wl.uniform(A.m, 99) if guess else None
        ref = _run(lib.shk_cg, A, b, dtype, jac, x0)
        got = _run(lib.shk_cg_noalloc, A, b, dtype, jac, x0)
        assert got[0] == ref[0] == 0
        assert got[1] == ref[1] > 0
        assert np.array_equal(got[2], ref[2])
        assert np.array_equal(got[3], ref[3])


def test_cg_noalloc_refuses_too_large(gpu):
    A, b, _ = wl.cg_laplace(1500)  # 2.25M FP64 rows: more than the 32 MiB scratch holds
    with pytest.raises(L.ShkError, match="fixed scratch"):
        _run(L.load().shk_cg_noalloc, A, b, torch.float64, True, None)
