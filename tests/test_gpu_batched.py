"""Batched persistent Cg: every system against the oracle's Cg (Seldon's
recurrences), graph and stream launches identical, and the FP32+FP64
refinement variant to the FP64 tolerance."""
import numpy as np
import pytest
import torch

from seldon_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _small_batch(count=24):
    from seldon_amd.batched import stack
    sides, eps = wl.batched_systems(count, 5, 8, 40)
    systems = [wl.aniso2d(int(n), int(n), float(e)) for n, e in zip(sides, eps)]
    A, row_off = stack(systems)
    xt = wl.uniform(A.m, 9)
    b = wl.apply(A, xt)
    b[row_off[3]:row_off[4]] = 0.0  # a zero right-hand side: Init fails, x untouched, code 0
    return A, row_off, b, systems


@pytest.mark.parametrize("group,rpp", [(64, 1), (256, 4), (512, 2), (1024, 8)])
def test_batched_matches_oracle_per_system(gpu, group, rpp):
    from oracle import oracle as orc
    from seldon_amd import krylov as K
    from seldon_amd.batched import cg_batched
    A, row_off, b, systems = _small_batch()
    Ag = K.Matrix(A, layout="rows")
    bg = torch.tensor(b, device="cuda")
    x = torch.full_like(bg, 7.0)
    its, codes = cg_batched(Ag, row_off, x, bg, tol=1e-8, max_iter=5000, group=group, use_graph=False,
                            rows_per_pass=rpp)
    xh = x.cpu().numpy()
    for s, S in enumerate(systems):
        sl = slice(row_off[s], row_off[s + 1])
        r = orc.solve("cg", S, b[sl], tol=1e-8, max_iter=5000)
        assert codes[s] == r.code
        assert abs(int(its[s]) - r.nb_iter) <= 1, (s, its[s], r.nb_iter)
        if s == 3:
            assert np.all(xh[sl] == 7.0) and its[s] == 0
        else:
            assert np.linalg.norm(xh[sl] - r.x) <= 1e-6 * np.linalg.norm(r.x)


def test_batched_graph_equals_stream_and_max_iter(gpu):
    from seldon_amd import krylov as K
    from seldon_amd.batched import cg_batched
    A, row_off, b, systems = _small_batch()
    Ag = K.Matrix(A, layout="rows")
    bg = torch.tensor(b, device="cuda")
    x1, x2 = torch.zeros_like(bg), torch.zeros_like(bg)
    i1, c1 = cg_batched(Ag, row_off, x1, bg, group=256, use_graph=False)
    i2, c2 = cg_batched(Ag, row_off, x2, bg, group=256, use_graph=True)
    assert np.array_equal(i1, i2) and np.array_equal(c1, c2)
    assert torch.equal(x1, x2)  # deterministic reductions
    for _ in range(3):          # graph replays (fresh graph per call) stay identical
        x3 = torch.zeros_like(bg)
        i3, c3 = cg_batched(Ag, row_off, x3, bg, group=256, use_graph=True)
        assert torch.equal(x1, x3) and np.array_equal(i1, i3)
    i4, c4 = cg_batched(Ag, row_off, torch.zeros_like(bg), bg, max_iter=5, group=256)
    live = [s for s in range(len(systems)) if s != 3]
    assert np.all(c4[live] == -2) and np.all(i4[live] == 5)


@pytest.mark.parametrize("inner_tol", [1e-3, 1e-4])
def test_batched_ir_reaches_fp64_tolerance(gpu, inner_tol):
    from seldon_amd import krylov as K
    from seldon_amd.batched import cg_batched, cg_batched_ir
    A, row_off, b, systems = _small_batch()
    A64 = K.Matrix(A, layout="rows")
    A32 = K.Matrix(A, dtype=torch.float32, layout="rows")
    bg = torch.tensor(b, device="cuda")
    x = torch.zeros_like(bg)
    its, codes = cg_batched_ir(A64, A32, row_off, x, bg, tol=1e-10, inner_tol=inner_tol, group=256)
    assert np.all(codes == 0), codes
    xref = torch.zeros_like(bg)
    cg_batched(A64, row_off, xref, bg, tol=1e-12, group=256)
    xh, xr = x.cpu().numpy(), xref.cpu().numpy()
    from oracle import oracle as orc
    for s, S in enumerate(systems):
        sl = slice(row_off[s], row_off[s + 1])
        if s == 3:
            continue
        res = np.linalg.norm(b[sl] - orc.mlt(S, xh[sl])) / np.linalg.norm(b[sl])
        assert res < 1e-10, (s, res)
        assert np.linalg.norm(xh[sl] - xr[sl]) <= 1e-7 * np.linalg.norm(xr[sl])


def test_autotune_picks_a_candidate(gpu):
    from seldon_amd import krylov as K
    from seldon_amd.batched import CANDIDATES, ROWS_PER_PASS, autotune
    A, row_off, b, systems = _small_batch()
    Ag = K.Matrix(A, layout="rows")
    bg = torch.tensor(b, device="cuda")
    (g, u), entry = autotune(Ag, row_off, bg, use_cache=False)
    assert g in CANDIDATES and u in ROWS_PER_PASS
    assert set(entry["seconds"]) == {f"{c}x{r}" for c in CANDIDATES for r in ROWS_PER_PASS}
    assert entry["seconds"][f"{g}x{u}"] == min(entry["seconds"].values())


def test_batched_rows_per_pass_results_identical(gpu):
    """Keeping more rows in flight per thread changes no arithmetic: the
    same iterations, codes and solutions, bit for bit, for every setting."""
    from seldon_amd import krylov as K
    from seldon_amd.batched import cg_batched, cg_batched_ir
    A, row_off, b, systems = _small_batch()
    A64 = K.Matrix(A, layout="rows")
    A32 = K.Matrix(A, dtype=torch.float32, layout="rows")
    bg = torch.tensor(b, device="cuda")
    ref = None
    for u in (1, 2, 4, 8):
        x = torch.zeros_like(bg)
        its, codes = cg_batched(A64, row_off, x, bg, group=256, rows_per_pass=u)
        xi = torch.zeros_like(bg)
        its_i, codes_i = cg_batched_ir(A64, A32, row_off, xi, bg, tol=1e-10, group=256, rows_per_pass=u)
        got = (its, codes, x.cpu().numpy(), its_i, codes_i, xi.cpu().numpy())
        if ref is None:
            ref = got
        for a, r in zip(got, ref):
            assert np.array_equal(a, r), u


def test_batched_rejects_a_column_in_another_system(gpu):
    """A stacked matrix whose system 2 has an entry in system 5's columns is
    refused (it would read another system's p while that one writes it);
    nothing is solved and x is left as it was."""
    from seldon_amd import _lib as L
    from seldon_amd import krylov as K
    from seldon_amd.batched import cg_batched
    A, row_off, b, systems = _small_batch()
    ind = A.ind.copy()
    ind[A.ptr[row_off[2]]] = row_off[5]  # first entry of system 2's first row
    B = wl.CSR(A.m, A.n, A.ptr, ind, A.val)
    Bg = K.Matrix(B, layout="rows")
    bg = torch.tensor(b, device="cuda")
    x = torch.full_like(bg, 3.0)
    with pytest.raises(L.ShkError, match="outside its own rows"):
        cg_batched(Bg, row_off, x, bg, group=256)
    assert torch.all(x == 3.0)


def test_batched_one_large_system_on_a_team_of_sixteen(gpu):
    """A batch of one 200 x 200 anisotropic Laplacian (40,000 rows): its share
    of the work is all of it, so 16 work-groups solve it together (members
    meeting in device memory, the dots added in member order) -- same
    iterations (within 1) and solution as the oracle's Cg, and the same
    result bit for bit on a second run (the team's sums have a fixed order)."""
    from oracle import oracle as orc
    from seldon_amd import krylov as K
    from seldon_amd.batched import cg_batched, cg_batched_ir, stack
    S = wl.aniso2d(200, 200, 30.0)
    A, row_off = stack([S])
    xt = wl.uniform(A.m, 17)
    b = wl.apply(A, xt)
    Ag = K.Matrix(A, layout="rows")
    A32 = K.Matrix(A, dtype=torch.float32, layout="rows")
    bg = torch.tensor(b, device="cuda")
    r = orc.solve("cg", S, b, tol=1e-8, max_iter=5000)
    xs = []
    for _ in range(2):
        x = torch.zeros_like(bg)
        its, codes = cg_batched(Ag, row_off, x, bg, tol=1e-8, max_iter=5000, group=256)
        assert codes[0] == r.code == 0
        assert abs(int(its[0]) - r.nb_iter) <= 1, (its[0], r.nb_iter)
        xs.append(x.cpu().numpy())
    assert np.array_equal(xs[0], xs[1])
    assert np.linalg.norm(xs[0] - r.x) <= 1e-6 * np.linalg.norm(r.x)
    x = torch.zeros_like(bg)
    its, codes = cg_batched_ir(Ag, A32, row_off, x, bg, tol=1e-10, inner_tol=1e-3, group=256)
    assert codes[0] == 0
    xh = x.cpu().numpy()
    assert np.linalg.norm(b - orc.mlt(S, xh)) <= 1e-10 * np.linalg.norm(b)
