"""GPU BLAS-1 (DotProd, Norm2, Add, Mlt(alpha,X), Copy, Fill) against the
oracle: elementwise ops bitwise, reductions within the fixed-order tree's
rounding bound (|err| <= n eps sum|x_i y_i|)."""
import numpy as np
import pytest
import torch

from seldon_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1000, 262144, 1 << 22])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_blas1(gpu, n, dt):
    from oracle import oracle as orc
    from seldon_amd import krylov as K
    ndt = np.float64 if dt == torch.float64 else np.float32
    eps = np.finfo(ndt).eps
    x = wl.uniform(n, 1).astype(ndt); y = wl.uniform(n, 2).astype(ndt)
    xg = torch.tensor(x, device="cuda"); yg = torch.tensor(y, device="cuda")
    d = K.DotProd(xg, yg)
    ref = orc.dot(x, y) if n else 0.0
    bound = 2 * max(n, 1) * eps * float(np.sum(np.abs(x.astype(np.float64) * y)))
    assert abs(d - ref) <= bound + 1e-300
    nr = K.Norm2(xg)
    assert abs(nr - (orc.norm2(x) if n else 0.0)) <= 2 * max(n, 1) * eps * (float(np.linalg.norm(x)) + 1e-300)
    for al in (0.0, -1.5, 0.25):
        yy = yg.clone()
        K.Add(al, xg, yy)
        yr = y.copy()
        if al != 0:
            yr = (yr + ndt(al) * x).astype(ndt)
        assert np.array_equal(yy.cpu().numpy(), yr)
    z = xg.clone(); K.Mlt(ndt(0.3), z)
    assert np.array_equal(z.cpu().numpy(), (x * ndt(0.3)).astype(ndt))
    z = torch.empty_like(xg); K.Copy(xg, z)
    assert np.array_equal(z.cpu().numpy(), x)
    K.Fill(z, 2.5)
    assert np.all(z.cpu().numpy() == ndt(2.5))
