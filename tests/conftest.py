import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def load_fixture(name):
    import json

    import numpy as np
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["spec"] = json.loads(str(d["spec"]))
    return d


@pytest.fixture(scope="session")
def gpu():
    """The HIP path on cuda:0; fails loudly (no fallback) when unusable."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from seldon_amd import krylov
    ctx = krylov.Context.default()
    info = ctx.info()
    assert info["arch"].startswith("gfx950"), info
    return ctx
