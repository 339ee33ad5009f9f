"""GPU tests of the Seldon-compatible C++ layer (include/seldon_amd).

* tests/cpp/seldon_hip_driver (standalone mode) takes the reference
  harness's command line and prints Iteration::Finished lines; its solves are
  compared with the real-Seldon fixtures, its SpMV bitwise.
* oracle/_ref/seldon_overlay_check (compiled against the real Seldon
  headers where that tree exists; the binary embeds those headers) runs the
  fused GPU Cg and Seldon's generic Cg template on GPU vectors; both must
  agree with Seldon's CPU Cg, recorded in tests/golden/overlay_cpu.npz.  It
  also runs fourteen more of Seldon's Krylov templates (BiCg, Qmr, Cgne,
  Lsqr with Mlt(SeldonTrans, ...) inside; Cgs, QCgs, TfQmr, Gcr, BiCgStabl,
  MinRes, Symmlq, CoCg, BiCgcr, QmrSym) on the CPU and on the GPU storages;
  they must agree within the spread of Seldon's own builds.
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle as orc
from tests.conftest import ROOT, load_fixture
from tests.golden.make_golden import build_case

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "seldon_amd", "_build", "bin", "seldon_hip_driver")
OVERLAY = os.path.join(ROOT, "oracle", "_ref", "seldon_overlay_check")


def _run(args):
    r = subprocess.run([DRIVER] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("name", ["cg_laplace64", "cg_laplace64_sym", "cg_laplace64_array", "cg_laplace64_x0",
                                  "cg_diffusion16_jacobi", "gmres30_convdiff16_jacobi", "gmres5_diag_identity"])
def test_driver_solve_matches_seldon(gpu, name):
    assert os.path.exists(DRIVER), "C++ driver not built"
    fx = load_fixture(f"solve_{name}.npz")
    spec = fx["spec"]
    A, b = build_case(spec)
    with tempfile.TemporaryDirectory() as tmp:
        orc.write_seldon_csr(f"{tmp}/A.bin", A)
        orc.write_seldon_vec(f"{tmp}/b.bin", b)
        x0 = "-"
        if spec.get("x0seed") is not None:
            from seldon_amd import workloads as wl
            orc.write_seldon_vec(f"{tmp}/x0.bin", wl.uniform(A.m, spec["x0seed"]))
            x0 = f"{tmp}/x0.bin"
        out = _run(["solve", spec["method"], spec.get("storage", "row"), spec["precond"], f"{tmp}/A.bin",
                    f"{tmp}/b.bin", x0, spec["tol"], spec["max_iter"], spec.get("restart", 10), f"{tmp}/x.bin"])
        hist, its, code, nb = orc.parse_ref_history(out)
        x = orc.read_seldon_vec(f"{tmp}/x.bin")
    assert code == int(fx["code"]) and nb == int(fx["nb_iter"])
    assert np.array_equal(its, fx["hist_it"])
    spread = np.maximum.accumulate(fx["spread"]) if len(fx["spread"]) else fx["spread"]
    rel = np.abs(hist - fx["hist"]) / np.abs(fx["hist"])
    assert np.all(rel <= np.maximum(1e-12, 10 * spread))
    assert abs(np.linalg.norm(x) - float(fx["x_norm"])) <= 1e-6 * float(fx["x_norm"])


@pytest.mark.parametrize("st", ["row", "sym", "array"])
def test_driver_spmv_matches_seldon(gpu, st):
    from seldon_amd import workloads as wl
    fx = load_fixture("spmv_powerlaw300.npz")
    sp = fx["spec"]
    A = wl.powerlaw(sp["n"], sp["seed"], kmin=sp["kmin"], kmax=sp["kmax"], half_width=sp["half_width"])
    M = wl.upper(A) if st == "sym" else A
    x = wl.uniform(A.m, sp["xseed"])
    y = wl.uniform(A.m, sp["yseed"])
    y[sp["nan_at"]] = np.nan
    with tempfile.TemporaryDirectory() as tmp:
        orc.write_seldon_csr(f"{tmp}/A.bin", M)
        orc.write_seldon_vec(f"{tmp}/x.bin", x)
        orc.write_seldon_vec(f"{tmp}/y.bin", y)
        for k, (al, be) in enumerate(fx["alpha_beta"]):
            _run(["spmv", st, "d", f"{tmp}/A.bin", f"{tmp}/x.bin", f"{tmp}/y.bin", repr(float(al)),
                  repr(float(be)), f"{tmp}/o.bin"])
            got = orc.read_seldon_vec(f"{tmp}/o.bin")
            assert np.array_equal(got, fx[f"{st}_d_{k}"], equal_nan=True), (st, al, be)


def _sections(out):
    sec, cur = {}, None
    for line in out.splitlines():
        if line.startswith("== "):
            cur = line[3:].strip()
            sec[cur] = []
        elif cur:
            sec[cur].append(line)
    return {k: orc.parse_ref_history("\n".join(v)) for k, v in sec.items()}


def test_overlay_in_real_seldon(gpu):
    """The GPU legs of the overlay program (built against the real Seldon
    headers into oracle/_ref, which embeds those headers) against its CPU
    legs recorded in this repository (tests/golden/overlay_cpu.npz, written
    by tests/golden/make_golden.py where the reference tree exists): the
    reference itself does not run here."""
    import hashlib
    import tempfile
    assert os.path.exists(OVERLAY), "oracle/_ref/seldon_overlay_check missing: build it where the reference exists"
    fx = np.load(os.path.join(ROOT, "tests", "golden", "overlay_cpu.npz"), allow_pickle=False)
    tmp = tempfile.mkdtemp()
    with open(os.path.join(tmp, "b_cpu.bin"), "wb") as f:  # Seldon's Vector::Write of the right-hand side
        f.write(fx["b_cpu"].tobytes())
    r = subprocess.run([OVERLAY] + json.loads(str(fx["args"])) + [tmp, "gpu"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # Seldon's WriteHarwellBoeing on the device storages: the CPU matrix's file
    for name, key in (("gpu.rua", "hb_sha256"), ("gpu_sym.rua", "hb_sym_sha256")):
        assert hashlib.sha256(open(os.path.join(tmp, name), "rb").read()).hexdigest() == str(fx[key]), name
    # Vector::Write from the device vector, and Seldon's file read back
    for name in ("b_gpu.bin", "b_back.bin"):
        assert open(os.path.join(tmp, name), "rb").read() == fx["b_cpu"].tobytes(), name
    s = _sections(r.stdout)

    def cpu(label):
        return (fx[f"{label}_hist"], fx[f"{label}_its"], int(fx[f"{label}_code"]), int(fx[f"{label}_nb"]))

    h_cpu, it_cpu, c_cpu, n_cpu = cpu("cpu")
    for label in ("gpu_fused", "gpu_generic"):
        h, it, c, n = s[label]
        assert c == c_cpu == 0 and n == n_cpu, (label, c, n, n_cpu)
        assert np.array_equal(it, it_cpu)
        rel = np.abs(h - h_cpu) / np.abs(h_cpu)
        # GPU reductions round differently from the sequential CPU sums; on
        # Laplacian Cg runs Seldon's own generic and MKL builds differ by up
        # to ~2e-12 over 720 iterations (tests/golden spread), the same order
        assert rel.max() < 1e-10, (label, rel.max())
    # Seldon's other Krylov templates on both sides: on the 24 x 24
    # convection-diffusion system the transpose-using BiCg, Qmr, Cgne, Lsqr
    # and Cgs, QCgs, TfQmr, Gcr, BiCgStabl; on the 64 x 64 Laplacian MinRes,
    # Symmlq, CoCg, BiCgcr, QmrSym, and Cg with Seldon's SorPreconditioner
    # on RowSparse and on RowSymSparse (SSOR sweeps level-scheduled on the
    # GPU, bitwise; to 1e-10); the others
    # to 1e-8.  Where Seldon's own builds
    # agree bitwise (generic, FMA, MKL-linked: spread 0) the GPU run must take
    # the same number of iterations with histories within 1e-6.  The others
    # are rounding-sensitive: those builds follow each other bitwise for the
    # first `agree` Finished() calls, then drift (Cgne 812 / 814 / 814
    # iterations and histories up to 0.23 apart; Lsqr up to 0.07; Cgs 5e-4;
    # QCgs 1e-6; TfQmr 0.02; BiCgStabl 0.13); there the first `agree` entries
    # must match within 1e-6 and the counts within 1% (GPU sums round
    # differently from the first dot product on).
    agree = {"cgne": 100, "lsqr": 100, "cgs": 40, "qcgs": 40, "tfqmr": 70, "bicgstabl": 2}
    # BiCg, Qmr, Cgne, Lsqr, Cgs, QCgs, TfQmr, Gcr, BiCgStabl, CoCg (= Cg on reals), BiCgcr, QmrSym, MinRes and Symmlq with
    # Preconditioner_Base run the
    # library's fused solvers
    # (label _gpu), and Seldon's templates with a user preconditioner type
    # (label _gpu_generic); both against the same CPU run
    legs = [(solver, "gpu") for solver in ("bicg", "qmr", "cgne", "lsqr", "cgs", "qcgs", "tfqmr", "gcr",
                                           "bicgstabl", "minres", "symmlq", "cocg", "bicgcr", "qmrsym",
                                           "ssor_cg", "ssor_cg_sym")]
    legs += [(m, "gpu_generic") for m in ("bicg", "qmr", "cgne", "lsqr", "cgs", "qcgs", "tfqmr", "gcr", "bicgstabl",
                                          "cocg", "bicgcr", "qmrsym", "minres", "symmlq")]
    for solver, leg in legs:
        h_c, it_c, c_c, n_c = cpu(f"{solver}_cpu")
        h_g, it_g, c_g, n_g = s[f"{solver}_{leg}"]
        k = min(len(h_c), len(h_g))
        rel = np.abs(h_g[:k] - h_c[:k]) / np.abs(h_c[:k])
        head = agree.get(solver, k)
        print(f"[overlay] {solver} {leg}: cpu {c_c}/{n_c} gpu {c_g}/{n_g} max rel history deviation {rel.max():.2e}, "
              f"first {head}: {rel[:head].max():.2e}", flush=True)
        assert c_c == c_g == 0, (solver, c_c, c_g)
        if solver not in agree:
            assert n_c == n_g and np.array_equal(it_g, it_c), (solver, n_c, n_g)
            assert rel.max() < 1e-6, (solver, rel.max())
        else:
            assert abs(n_g - n_c) <= max(2, n_c // 100), (solver, n_c, n_g)
            assert rel[:head].max() < 1e-6, (solver, rel[:head].max())
