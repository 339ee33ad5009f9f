"""Differentiable solve: SDDMM gradient kernel against the oracle, and the
adjoint gradients against an FP64 CPU reference and central finite
differences (bar: 1e-6 relative)."""
import numpy as np
import pytest
import torch

from seldon_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("st", ["row", "sym", "array"])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_sddmm_vs_oracle(gpu, st, dt):
    from oracle import oracle as orc
    from seldon_amd import _lib as L
    from seldon_amd import krylov as K
    ndt = np.float64 if dt == torch.float64 else np.float32
    A = wl.powerlaw(30000, 12, kmin=1, kmax=5000, expo=1.2, half_width=6000)  # includes rows > budget
    M = wl.upper(A) if st == "sym" else A
    lam = wl.uniform(A.m, 3).astype(ndt)
    x = wl.uniform(A.m, 4).astype(ndt)
    Mg = K.Matrix(M, storage=st, dtype=dt)
    g = torch.empty(M.nnz, dtype=dt, device="cuda")
    lam_d = torch.tensor(lam, device="cuda")  # keep the device buffers alive across the call
    x_d = torch.tensor(x, device="cuda")
    L.check(L.load().shk_sddmm(Mg.ctx.h, Mg.h, K._ptr(lam_d), K._ptr(x_d), K._ptr(g)))
    ref = orc.sddmm_sym_fold(M, lam, x) if st == "sym" else orc.sddmm(M, lam, x)
    assert np.array_equal(g.cpu().numpy(), ref.astype(ndt))


def _problem(n, seed=41):
    theta = wl.normal(n ** 3, seed, 0.0, 0.5)
    b = wl.uniform(n ** 3, seed + 1, 0.0, 1.0)
    dtheta = wl.normal(n ** 3, seed + 2, 0.0, 0.1)
    return theta, b, dtheta


def _loss_grad(op, theta_t, b_t, x_obs_t, tol=1e-13):
    theta_t = theta_t.clone().requires_grad_(True)
    b_t = b_t.clone().requires_grad_(True)
    vals = op.values(theta_t)
    x = op.solve(vals, b_t, tol=tol)
    loss = torch.sum((x - x_obs_t) ** 2)
    loss.backward()
    return float(loss), theta_t.grad.detach().cpu().numpy(), b_t.grad.detach().cpu().numpy(), vals.detach()


@pytest.mark.parametrize("n", [12, 24])
def test_adjoint_matches_fp64_cpu_reference(gpu, n):
    from oracle import oracle as orc
    from seldon_amd.adjoint import DiffusionOperator
    theta, b, dtheta = _problem(n)
    op = DiffusionOperator(n)
    th = torch.tensor(theta, device="cuda")
    bt = torch.tensor(b, device="cuda")
    with torch.no_grad():
        x_obs = op.solve(op.values(th + torch.tensor(dtheta, device="cuda")), bt, tol=1e-13)
    loss, gth, gb, vals = _loss_grad(op, th, bt, x_obs)
    # the torch assembly reproduces the C generator's matrix
    gen = wl.diffusion3d(n, theta)
    assert np.allclose(vals.cpu().numpy(), gen.val, rtol=1e-14, atol=0)  # exp() ulps differ GPU/CPU
    ref = orc.diffusion_grad_reference(n, theta, vals.cpu().numpy(), op.pattern, b, x_obs.cpu().numpy())
    assert abs(loss - ref["loss"]) <= 1e-9 * abs(ref["loss"])
    for got, want in ((gth, ref["grad_theta"]), (gb, ref["grad_b"])):
        rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert rel <= 1e-6, rel


def test_adjoint_central_finite_differences(gpu):
    from seldon_amd.adjoint import DiffusionOperator
    n = 16
    theta, b, dtheta = _problem(n, seed=77)
    op = DiffusionOperator(n)
    th = torch.tensor(theta, device="cuda")
    bt = torch.tensor(b, device="cuda")
    with torch.no_grad():
        x_obs = op.solve(op.values(th + torch.tensor(dtheta, device="cuda")), bt, tol=1e-14)
    _, gth, gb, _ = _loss_grad(op, th, bt, x_obs, tol=1e-14)
    h = 1e-6

    def L(th_, b_):
        with torch.no_grad():
            x = op.solve(op.values(th_), b_, tol=1e-14)
            return float(torch.sum((x - x_obs) ** 2))

    rng = np.random.default_rng(5)
    scale_t = np.max(np.abs(gth))
    for j in rng.choice(n ** 3, 8, replace=False):
        e = torch.zeros_like(th); e[j] = h
        fd = (L(th + e, bt) - L(th - e, bt)) / (2 * h)
        assert abs(fd - gth[j]) <= 1e-6 * scale_t + 1e-6 * abs(gth[j]), (j, fd, gth[j])
    scale_b = np.max(np.abs(gb))
    for j in rng.choice(n ** 3, 4, replace=False):
        e = torch.zeros_like(bt); e[j] = h
        fd = (L(th, bt + e) - L(th, bt - e)) / (2 * h)
        assert abs(fd - gb[j]) <= 1e-6 * scale_b + 1e-6 * abs(gb[j]), (j, fd, gb[j])


def test_adjoint_nonsymmetric_values_gradient(gpu):
    """Gmres forward, transposed Gmres adjoint; dL/dA against the dense
    formula -lam x^T on the pattern with lam from a dense solve."""
    from seldon_amd.adjoint import SparseOperator
    A = wl.convdiff3d(8)
    b = wl.uniform(A.m, 3)
    x_obs = wl.uniform(A.m, 4)
    op = SparseOperator(A, spd=False)
    v = torch.tensor(A.val, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda")
    x = op.solve(v, bt, tol=1e-13)
    loss = torch.sum((x - torch.tensor(x_obs, device="cuda")) ** 2)
    loss.backward()
    D = A.to_scipy().toarray()
    xd = np.linalg.solve(D, b)
    lam = np.linalg.solve(D.T, 2 * (xd - x_obs))
    rows = np.repeat(np.arange(A.m), np.diff(A.ptr))
    want = -(lam[rows] * xd[A.ind])
    got = v.grad.cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-8 * np.linalg.norm(want)


def test_adjoint_nonsymmetric_two_solves_one_backward(gpu):
    """Two solves on one nonsymmetric operator at different values, one
    backward through both: each adjoint Gmres must run on A^T of its own
    forward's values (the handle holds the second solve's values when the
    first backward runs).  Checked against the dense formula per solve."""
    from seldon_amd.adjoint import SparseOperator
    A = wl.convdiff3d(8)
    b = wl.uniform(A.m, 3)
    x_obs = wl.uniform(A.m, 4)
    op = SparseOperator(A, spd=False)
    v1 = torch.tensor(A.val, device="cuda", requires_grad=True)
    scale = np.where(np.arange(A.nnz) % 3 == 0, 1.3, 0.8)  # a different nonsymmetric matrix
    v2 = torch.tensor(A.val * scale, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda")
    xo = torch.tensor(x_obs, device="cuda")
    x1 = op.solve(v1, bt, tol=1e-13)
    x2 = op.solve(v2, bt, tol=1e-13)
    loss = torch.sum((x1 - xo) ** 2) + 0.5 * torch.sum((x2 - xo) ** 2)
    loss.backward()
    rows = np.repeat(np.arange(A.m), np.diff(A.ptr))
    for v, vals, wgt in ((v1, A.val, 2.0), (v2, A.val * scale, 1.0)):
        M = wl.CSR(A.m, A.n, A.ptr, A.ind, vals)
        D = M.to_scipy().toarray()
        xd = np.linalg.solve(D, b)
        lam = np.linalg.solve(D.T, wgt * (xd - x_obs))
        want = -(lam[rows] * xd[A.ind])
        got = v.grad.cpu().numpy()
        assert np.linalg.norm(got - want) <= 1e-8 * np.linalg.norm(want)


def test_adjoint_diffusion64_full_size(gpu):
    """The differentiable solve at its stated size: 64^3 cells (262,144
    rows), theta ~ N(0, 0.5^2), b ~ U[0, 1], x_obs from theta + N(0, 0.1^2).
    Gradients w.r.t. theta and b against the oracle's FP64 CPU gradient
    (its Cg forward and adjoint, SDDMM and chain rule) within 1e-6 relative,
    and against central finite differences (h = 1e-6) of the GPU loss at 32
    theta and 32 b components.  A central difference here carries the noise
    of the two FP64 solves it subtracts: the loss (~3e3) is known to about
    1e-14 of |x - x_obs| |x|, which over 2h = 2e-6 is ~1e-5 of the largest
    gradient component, so the FD bar is 1e-4 of max |grad| (at 16^3, where
    that noise is smaller, the FD test holds 1e-6)."""
    from oracle import oracle as orc
    from seldon_amd.adjoint import DiffusionOperator
    n = 64
    theta, b, dtheta = _problem(n, seed=2025)
    op = DiffusionOperator(n)
    th = torch.tensor(theta, device="cuda")
    bt = torch.tensor(b, device="cuda")
    with torch.no_grad():
        x_obs = op.solve(op.values(th + torch.tensor(dtheta, device="cuda")), bt, tol=1e-14)
    loss, gth, gb, vals = _loss_grad(op, th, bt, x_obs, tol=1e-14)
    ref = orc.diffusion_grad_reference(n, theta, vals.cpu().numpy(), op.pattern, b, x_obs.cpu().numpy(),
                                       tol=1e-14)
    print(f"[adjoint64] loss {loss:.6e} (oracle {ref['loss']:.6e}), oracle its {ref['its']}", flush=True)
    assert abs(loss - ref["loss"]) <= 1e-9 * abs(ref["loss"])
    for got, want in ((gth, ref["grad_theta"]), (gb, ref["grad_b"])):
        rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"[adjoint64] gradient rel. error vs oracle {rel:.2e}", flush=True)
        assert rel <= 1e-6, rel
    h = 1e-6

    def L(th_, b_):
        with torch.no_grad():
            x = op.solve(op.values(th_), b_, tol=1e-14)
            return float(torch.sum((x - x_obs) ** 2))

    rng = np.random.default_rng(64)
    worst = 0.0
    for grad, which in ((gth, "theta"), (gb, "b")):
        scale = np.max(np.abs(grad))
        for j in rng.choice(n ** 3, 32, replace=False):
            e = torch.zeros_like(th)
            e[j] = h
            if which == "theta":
                fd = (L(th + e, bt) - L(th - e, bt)) / (2 * h)
            else:
                fd = (L(th, bt + e) - L(th, bt - e)) / (2 * h)
            err = abs(fd - grad[j]) / scale
            worst = max(worst, err)
            assert abs(fd - grad[j]) <= 1e-4 * scale, (which, j, fd, grad[j])
    print(f"[adjoint64] worst central-FD deviation {worst:.2e} (of max |grad|)", flush=True)
