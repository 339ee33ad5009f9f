"""Write the golden fixtures under tests/golden/ from the real Seldon build.

Runs the drivers that oracle/Makefile compiles from the reference headers
(oracle/_ref/seldon_ref_generic: Seldon's generic templates;
oracle/_ref/seldon_ref_blas: the same with SELDON_WITH_BLAS linked to the
image's MKL, needed by Gmres) on seeded inputs from seldon_amd.workloads and
stores their outputs as small .npz files.  The inputs themselves are not
stored: each fixture names its generator call, and a checksum of the
generated arrays guards against generator drift.

Usage:  python tests/golden/make_golden.py      (needs /root/reference)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from seldon_amd import workloads as wl  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def csr_digest(A) -> str:
    h = hashlib.sha256()
    for a in (A.ptr, A.ind, A.val):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def build_case(spec):
    """Regenerate (A, b) from a fixture spec (also used by the tests)."""
    kind = spec["kind"]
    if kind == "laplace2d":
        A = wl.laplace2d(spec["grid"])
    elif kind == "convdiff3d":
        A = wl.convdiff3d(spec["grid"])
    elif kind == "powerlaw":
        A = wl.powerlaw(spec["n"], spec["seed"], half_width=spec["half_width"])
    elif kind == "diffusion3d":
        A = wl.diffusion3d(spec["grid"], wl.normal(spec["grid"] ** 3, spec["seed"], 0.0, 0.5))
    elif kind == "aniso2d":
        A = wl.aniso2d(spec["grid"], spec["grid"], spec["eps"])
    elif kind == "diag":
        m = spec["n"]
        vals = np.array(spec["values"], np.float64)
        A = wl.CSR(m, m, np.arange(m + 1, dtype=np.int32), np.arange(m, dtype=np.int32),
                   vals[np.arange(m) % vals.size].copy())
    else:
        raise ValueError(kind)
    xt = wl.uniform(A.m, spec.get("xseed", wl.SEED_XTRUE))
    b = wl.apply(A, xt)
    if spec.get("zero_rhs"):
        b = np.zeros_like(b)
    if spec.get("storage") == "sym":
        A = wl.upper(A)
    return A, b


def run_solve(tmp, spec, kind):
    A, b = build_case(spec)
    orc.write_seldon_csr(f"{tmp}/A.bin", A)
    orc.write_seldon_vec(f"{tmp}/b.bin", b)
    x0 = "-"
    if spec.get("x0seed") is not None:
        x0v = wl.uniform(A.m, spec["x0seed"])
        orc.write_seldon_vec(f"{tmp}/x0.bin", x0v)
        x0 = f"{tmp}/x0.bin"
    out = orc.run_ref(["solve", spec["method"], spec.get("storage", "row"), spec["precond"],
                       f"{tmp}/A.bin", f"{tmp}/b.bin", x0, spec["tol"], spec["max_iter"],
                       spec.get("restart", 10), f"{tmp}/x.bin"], kind=kind)
    hist, its, code, nb = orc.parse_ref_history(out)
    x = orc.read_seldon_vec(f"{tmp}/x.bin")
    return A, hist, its, code, nb, x


SOLVES = {
    "cg_laplace64": dict(kind="laplace2d", grid=64, method="cg", precond="id", tol=1e-10, max_iter=2000),
    "cg_laplace256": dict(kind="laplace2d", grid=256, method="cg", precond="id", tol=1e-10, max_iter=5000),
    "cg_laplace64_sym": dict(kind="laplace2d", grid=64, method="cg", precond="id", tol=1e-10,
                             max_iter=2000, storage="sym"),
    "cg_laplace64_array": dict(kind="laplace2d", grid=64, method="cg", precond="id", tol=1e-10,
                               max_iter=2000, storage="array"),
    "cg_laplace64_x0": dict(kind="laplace2d", grid=64, method="cg", precond="id", tol=1e-10,
                            max_iter=2000, x0seed=99),
    "cg_laplace64_maxit": dict(kind="laplace2d", grid=64, method="cg", precond="id", tol=1e-10, max_iter=7),
    "cg_laplace64_zero_rhs": dict(kind="laplace2d", grid=64, method="cg", precond="id", tol=1e-10,
                                  max_iter=100, zero_rhs=True),
    "cg_diffusion16_jacobi": dict(kind="diffusion3d", grid=16, seed=31, method="cg", precond="jacobi",
                                  tol=1e-12, max_iter=2000),
    "cg_aniso64_eps30": dict(kind="aniso2d", grid=64, eps=30.0, method="cg", precond="id", tol=1e-8,
                             max_iter=5000),
    "cg_diag5": dict(kind="diag", n=500, values=[1.0, 2.0, 3.5, 7.0, 11.0], method="cg", precond="id",
                     tol=1e-12, max_iter=100),
    "bicgstab_convdiff16_jacobi": dict(kind="convdiff3d", grid=16, method="bicgstab", precond="jacobi",
                                       tol=1e-10, max_iter=2000, xseed=wl.SEED_XTRUE + 1),
    "bicgstab_convdiff24_jacobi": dict(kind="convdiff3d", grid=24, method="bicgstab", precond="jacobi",
                                       tol=1e-10, max_iter=2000, xseed=wl.SEED_XTRUE + 1),
    "bicgstab_laplace64_id": dict(kind="laplace2d", grid=64, method="bicgstab", precond="id", tol=1e-10,
                                  max_iter=2000),
    "gmres30_convdiff16_jacobi": dict(kind="convdiff3d", grid=16, method="gmres", precond="jacobi",
                                      tol=1e-10, max_iter=2000, restart=30, xseed=wl.SEED_XTRUE + 1),
    "gmres30_convdiff24_jacobi": dict(kind="convdiff3d", grid=24, method="gmres", precond="jacobi",
                                      tol=1e-10, max_iter=2000, restart=30, xseed=wl.SEED_XTRUE + 1),
    "gmres50_powerlaw20k_jacobi": dict(kind="powerlaw", n=20000, seed=2024, half_width=2000,
                                       method="gmres", precond="jacobi", tol=1e-10, max_iter=2000,
                                       restart=50, xseed=wl.SEED_XTRUE + 2),
    "gmres10_laplace32_id": dict(kind="laplace2d", grid=32, method="gmres", precond="id", tol=1e-10,
                                 max_iter=3000, restart=10),
    "gmres5_diag_identity": dict(kind="diag", n=300, values=[1.0], method="gmres", precond="id",
                                 tol=1e-10, max_iter=50, restart=5),
}


def run_solve_f32(tmp, spec, kind):
    """Seldon's solver on float matrix and vectors (the harness's `solvef`:
    Iteration<double> for Cg / BiCgStab; Iteration<float> for Gmres, the only
    float Gmres that compiles)."""
    A, b = build_case(spec)
    orc.write_seldon_csr(f"{tmp}/A.bin", A, np.float32)
    orc.write_seldon_vec(f"{tmp}/b.bin", b.astype(np.float32))
    out = orc.run_ref(["solvef", spec["method"], spec.get("storage", "row"), spec["precond"],
                       f"{tmp}/A.bin", f"{tmp}/b.bin", "-", spec["tol"], spec["max_iter"],
                       spec.get("restart", 10), f"{tmp}/x.bin"], kind=kind)
    hist, its, code, nb = orc.parse_ref_history(out)
    x = orc.read_seldon_vec(f"{tmp}/x.bin", np.float32)
    return A, hist, its, code, nb, x


# FP32 solves: every vector, matrix value and recurrence scalar in float.
SOLVES_F32 = {
    "cg_laplace64": dict(kind="laplace2d", grid=64, method="cg", precond="id", tol=1e-6, max_iter=2000),
    "cg_laplace64_sym": dict(kind="laplace2d", grid=64, method="cg", precond="id", tol=1e-6,
                             max_iter=2000, storage="sym"),
    "cg_laplace64_array": dict(kind="laplace2d", grid=64, method="cg", precond="id", tol=1e-6,
                               max_iter=2000, storage="array"),
    "cg_diffusion16_jacobi": dict(kind="diffusion3d", grid=16, seed=31, method="cg", precond="jacobi",
                                  tol=1e-6, max_iter=2000),
    "cg_laplace128": dict(kind="laplace2d", grid=128, method="cg", precond="id", tol=1e-5, max_iter=4000),
    "bicgstab_convdiff16_jacobi": dict(kind="convdiff3d", grid=16, method="bicgstab", precond="jacobi",
                                       tol=1e-6, max_iter=2000, xseed=wl.SEED_XTRUE + 1),
    "bicgstab_laplace64_id": dict(kind="laplace2d", grid=64, method="bicgstab", precond="id", tol=1e-5,
                                  max_iter=2000),
    # Gmres tolerances stay a decade above FP32's attainable accuracy (about
    # 1e-6 relative on these systems), where the residual estimate stagnates
    "gmres30_convdiff16_jacobi": dict(kind="convdiff3d", grid=16, method="gmres", precond="jacobi",
                                      tol=1e-5, max_iter=2000, restart=30, xseed=wl.SEED_XTRUE + 1),
    "gmres50_powerlaw20k_jacobi": dict(kind="powerlaw", n=20000, seed=2024, half_width=2000,
                                       method="gmres", precond="jacobi", tol=1e-5, max_iter=2000,
                                       restart=50, xseed=wl.SEED_XTRUE + 2),
    "gmres10_laplace32_id": dict(kind="laplace2d", grid=32, method="gmres", precond="id", tol=1e-5,
                                 max_iter=3000, restart=10),
}


def make_solves_f32():
    """solve_f32_*.npz.  Cg / BiCgStab: `hist` etc. from Seldon's generic
    float build, which the oracle's FP32 Cg / BiCgStab equal bit for bit;
    `other_*` from the MKL build.  Gmres: `hist` from Seldon's float Gmres
    (Iteration<float>: float Hessenberg and rotations, MKL's srotg / strsv);
    `other_*` from the oracle's FP32 Gmres, which keeps the Hessenberg,
    rotations and triangular solve in double (the FP32 inner solver of the
    mixed-precision path)."""
    with tempfile.TemporaryDirectory() as tmp:
        for name, spec in SOLVES_F32.items():
            spec = dict(spec, dtype="float32")
            A, b = build_case(spec)
            if spec["method"] == "gmres":
                _, hist, its, code, nb, x = run_solve_f32(tmp, spec, "blas")
                g = orc.solve("gmres", A, b, precond=spec["precond"], tol=spec["tol"], max_iter=spec["max_iter"],
                              restart=spec.get("restart", 10), storage=spec.get("storage"), dtype=np.float32)
                other = (g.hist, g.hist_it, g.code, g.nb_iter, g.x)
            else:
                _, hist, its, code, nb, x = run_solve_f32(tmp, spec, "generic")
                _, oh, oi, oc, onb, ox = run_solve_f32(tmp, spec, "blas")
                other = (oh, oi, oc, onb, ox)
            n = min(len(other[0]), len(hist))
            spread = np.zeros(len(hist))
            if n:
                spread[:n] = np.abs(other[0][:n] - hist[:n]) / np.maximum(np.abs(hist[:n]), 1e-300)
            np.savez_compressed(
                os.path.join(OUT, f"solve_f32_{name}.npz"),
                spec=json.dumps(spec), digest=csr_digest(A),
                hist=hist, hist_it=its, code=code, nb_iter=nb, x=x,
                x_sum=np.sum(x.astype(np.float64)), x_norm=np.linalg.norm(x.astype(np.float64)),
                build="blas-float-iteration" if spec["method"] == "gmres" else "generic",
                other_hist=other[0], other_code=other[2], other_nb_iter=other[3],
                spread=spread)
            print(f"f32 {name}: code {code} its {nb} hist {len(hist)} "
                  f"(other: code {other[2]} its {other[3]}, max spread {spread.max() if spread.size else 0:.2e})")


def make_solves():
    with tempfile.TemporaryDirectory() as tmp:
        for name, spec in SOLVES.items():
            res = {}
            kinds = ["blas"] if spec["method"] == "gmres" else ["generic", "blas"]
            for kind in kinds:
                A, hist, its, code, nb, x = run_solve(tmp, spec, kind)
                res[kind] = (hist, its, code, nb, x)
            primary = res["blas"] if spec["method"] == "gmres" else res["generic"]
            hist, its, code, nb, x = primary
            other = res["blas"]
            if spec["method"] == "gmres":
                # envelope for Gmres: the generic-template restatement against
                # the MKL build (the only build of Seldon's Gmres there is)
                Ar, br = build_case(spec)
                x0 = None if spec.get("x0seed") is None else wl.uniform(Ar.m, spec["x0seed"])
                g = orc.solve("gmres", Ar, br, precond=spec["precond"], tol=spec["tol"],
                              max_iter=spec["max_iter"], restart=spec.get("restart", 10), x0=x0,
                              storage=spec.get("storage"))
                other = (g.hist, g.hist_it, g.code, g.nb_iter, g.x)
            n = min(len(other[0]), len(hist))
            spread = np.zeros(len(hist))
            if n:
                spread[:n] = np.abs(other[0][:n] - hist[:n]) / np.maximum(np.abs(hist[:n]), 1e-300)
            np.savez_compressed(
                os.path.join(OUT, f"solve_{name}.npz"),
                spec=json.dumps(spec), digest=csr_digest(A),
                hist=hist, hist_it=its, code=code, nb_iter=nb,
                x_sum=np.sum(x), x_norm=np.linalg.norm(x), x_head=x[:16],
                build="blas" if spec["method"] == "gmres" else "generic",
                other_hist=other[0], other_code=other[2], other_nb_iter=other[3],
                spread=spread)
            print(f"{name}: code {code} its {nb} hist {len(hist)} "
                  f"(MKL build: code {other[2]} its {other[3]}, max spread {spread.max() if spread.size else 0:.2e})")


SPMV_ALPHA_BETA = [(1.0, 0.0), (0.3, -1.0), (-1.0, 1.0), (0.0, 0.3), (1.0, 1.0), (2.5, 0.0)]


def make_spmv():
    A = wl.powerlaw(300, 5, kmin=2, kmax=40, half_width=40)
    U = wl.upper(A)
    x = wl.uniform(A.m, 11)
    y = wl.uniform(A.m, 12)
    y[3] = np.nan  # beta*NaN: RowSparse/RowSym propagate it, ArrayRowSparse with beta=0 does not
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for st, M in (("row", A), ("sym", U), ("array", A)):
            for dt, c in ((np.float64, "d"), (np.float32, "f")):
                orc.write_seldon_csr(f"{tmp}/M.bin", M, dt)
                orc.write_seldon_vec(f"{tmp}/x.bin", x.astype(dt))
                orc.write_seldon_vec(f"{tmp}/y.bin", y.astype(dt))
                for k, (al, be) in enumerate(SPMV_ALPHA_BETA):
                    orc.run_ref(["spmv", st, c, f"{tmp}/M.bin", f"{tmp}/x.bin", f"{tmp}/y.bin",
                                 repr(al), repr(be), f"{tmp}/o.bin"])
                    out[f"{st}_{c}_{k}"] = orc.read_seldon_vec(f"{tmp}/o.bin", dt)
                orc.run_ref(["mlt", st, c, f"{tmp}/M.bin", f"{tmp}/x.bin", f"{tmp}/o.bin"])
                out[f"{st}_{c}_mlt"] = orc.read_seldon_vec(f"{tmp}/o.bin", dt)
    np.savez_compressed(os.path.join(OUT, "spmv_powerlaw300.npz"),
                        spec=json.dumps(dict(n=300, seed=5, kmin=2, kmax=40, half_width=40,
                                             xseed=11, yseed=12, nan_at=3)),
                        alpha_beta=np.array(SPMV_ALPHA_BETA), digest=csr_digest(A), **out)
    print("spmv fixtures:", len(out))


def make_spmvt():
    """MltAdd(alpha, SeldonTrans, M, X, beta, Y) and Mlt(SeldonTrans, M, X, Y)
    from the reference build, on a 300 x 220 matrix (RowSparse and
    ArrayRowSparse) and on its square parent's upper triangle
    (RowSymSparse)."""
    A = wl.powerlaw(300, 7, kmin=2, kmax=40, half_width=60)
    R = wl.first_columns(A, 220)
    U = wl.upper(A)
    x = wl.uniform(300, 21)
    y_rect = wl.uniform(220, 22)
    y_rect[5] = np.nan
    y_sym = wl.uniform(300, 23)
    y_sym[5] = np.nan
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for st, M, y in (("row", R, y_rect), ("array", R, y_rect), ("sym", U, y_sym)):
            for dt, c in ((np.float64, "d"), (np.float32, "f")):
                orc.write_seldon_csr(f"{tmp}/M.bin", M, dt)
                orc.write_seldon_vec(f"{tmp}/x.bin", x.astype(dt))
                orc.write_seldon_vec(f"{tmp}/y.bin", y.astype(dt))
                for k, (al, be) in enumerate(SPMV_ALPHA_BETA):
                    orc.run_ref(["spmvt", st, c, f"{tmp}/M.bin", f"{tmp}/x.bin", f"{tmp}/y.bin",
                                 repr(al), repr(be), f"{tmp}/o.bin"])
                    out[f"{st}_{c}_{k}"] = orc.read_seldon_vec(f"{tmp}/o.bin", dt)
                orc.run_ref(["mltt", st, c, f"{tmp}/M.bin", f"{tmp}/x.bin", f"{tmp}/o.bin"])
                out[f"{st}_{c}_mlt"] = orc.read_seldon_vec(f"{tmp}/o.bin", dt)
    np.savez_compressed(os.path.join(OUT, "spmvt_rect300x220.npz"),
                        spec=json.dumps(dict(n=300, seed=7, kmin=2, kmax=40, half_width=60, ncols=220,
                                             xseed=21, yseed_rect=22, yseed_sym=23, nan_at=5)),
                        alpha_beta=np.array(SPMV_ALPHA_BETA), digest=csr_digest(R), **out)
    print("spmvt fixtures:", len(out))


SOR_CASES = [(1.0, 1, 2), (1.3, 2, 2), (1.3, 2, 3), (0.8, 1, 0), (1.3, 2, 0)]  # (omega, iter, type_ssor)
SOR_BACKWARD_ONCE = [(1.3, 1, 3), (1.0, 1, 3)]


def scramble_rows(A, seed, hide_first_diagonal):
    """Each row reordered: its columns below the diagonal in random order,
    the diagonal, then its larger columns in random order, so Seldon's
    A(j, j) (a scan that stops at the first column not below j) still finds
    the diagonal.  With hide_first_diagonal, row 0 starts with a larger
    column and that scan reads A(0, 0) as 0."""
    rng = np.random.default_rng(seed)
    ind, val = A.ind.copy(), A.val.copy()
    for j in range(A.m):
        a, b = A.ptr[j], A.ptr[j + 1]
        c = ind[a:b]
        lo, dg, hi = np.flatnonzero(c < j), np.flatnonzero(c == j), np.flatnonzero(c > j)
        order = np.concatenate([rng.permutation(lo), dg, rng.permutation(hi)])
        if j == 0 and hide_first_diagonal and len(hi):
            order = np.concatenate([hi[:1], dg, hi[1:]])
        ind[a:b], val[a:b] = c[order], A.val[a:b][order]
    return wl.CSR(A.m, A.n, A.ptr.copy(), ind, val)


def sor_case():
    """name -> (matrix, storages, cases): a 512-row 3-D convection-diffusion
    matrix (nonsymmetric), a 400-row power-law matrix given a dominant
    diagonal, the upper triangle of a 16^2 Laplacian (RowSymSparse), and the
    convection-diffusion matrix with scrambled rows whose row 0 hides its
    diagonal (swept backward once, so only x(0) is not finite)."""
    C = wl.convdiff3d(8)
    P = wl.powerlaw(400, 17, kmin=2, kmax=30, half_width=60)
    return {"convdiff8": (C, ("row", "array"), SOR_CASES),
            "powerlaw400": (P, ("row", "array"), SOR_CASES),
            "laplace16": (wl.upper(wl.laplace2d(16)), ("sym",), SOR_CASES),
            "convdiff8scrambled": (scramble_rows(C, 53, True), ("row",), SOR_BACKWARD_ONCE)}


SOR_TYPES = (("d", np.float64, False), ("f", np.float32, False), ("fw", np.float32, True))  # fw: double omega


def sor_fixture_cases():
    """(key, A, storage, dtype, wide, omega, iter, type_ssor) of every fixture."""
    for name, (A, storages, cases) in sor_case().items():
        for c, dt, wide in SOR_TYPES:
            for st in storages:
                for k, (om, it, ty) in enumerate(cases):
                    yield f"{name}_{st}_{c}_{k}", A, st, dt, wide, om, it, ty


def make_sor():
    """SOR(A, X, B, omega, iter, type_ssor) from the reference build."""
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for key, A, st, dt, wide, om, it, ty in sor_fixture_cases():
            orc.write_seldon_csr(f"{tmp}/A.bin", A, dt)
            orc.write_seldon_vec(f"{tmp}/x.bin", wl.uniform(A.m, 51).astype(dt))
            orc.write_seldon_vec(f"{tmp}/b.bin", wl.uniform(A.m, 52).astype(dt))
            args = ["sor", st, "d" if dt == np.float64 else "f", f"{tmp}/A.bin", f"{tmp}/x.bin", f"{tmp}/b.bin",
                    repr(om), it, ty, f"{tmp}/o.bin"] + (["wide"] if wide else [])
            orc.run_ref(args)
            out[key] = orc.read_seldon_vec(f"{tmp}/o.bin", dt)
    np.savez_compressed(os.path.join(OUT, "sor.npz"),
                        spec=json.dumps({"cases": SOR_CASES, "backward_once": SOR_BACKWARD_ONCE, "xseed": 51,
                                         "bseed": 52, "scramble_seed": 53}), **out)
    print("sor fixtures:", len(out))


COO_CASES = {
    # name: (seed, nnz, rows, cols, duplicates, index, m_min, n_min)
    "dup": (31, 4000, 300, 200, 400, 1, 0, 250),
    "gaps": (32, 2500, 900, 700, 50, 0, 1200, 0),
    "none": (33, 0, 1, 1, 0, 0, 7, 3),
}


def coo_case(name):
    """Seeded coordinates with duplicates (repeated (row, col) pairs with other
    values) and, for "gaps", only every third row used."""
    seed, nnz, rows, cols, dup, index, m_min, n_min = COO_CASES[name]
    rng = np.random.default_rng(seed)
    row = rng.integers(0, rows, nnz)
    if name == "gaps":
        row = 3 * (row // 3)
    col = rng.integers(0, cols, nnz)
    if dup:
        pick = rng.integers(0, nnz, dup)
        where = rng.integers(0, nnz, dup)
        row[where] = row[pick]
        col[where] = col[pick]
    val = rng.standard_normal(nnz)
    return (row + index).astype(np.int32), (col + index).astype(np.int32), val, index, m_min, n_min


def make_coo():
    """ConvertMatrix_from_Coordinates (RowSparse) from the reference build."""
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name in COO_CASES:
            row, col, val, index, m_min, n_min = coo_case(name)
            for dt, c in ((np.float64, "d"), (np.float32, "f")):
                orc.write_seldon_vec(f"{tmp}/r.bin", row)
                orc.write_seldon_vec(f"{tmp}/c.bin", col)
                orc.write_seldon_vec(f"{tmp}/v.bin", val.astype(dt))
                orc.run_ref(["coo2csr", c, f"{tmp}/r.bin", f"{tmp}/c.bin", f"{tmp}/v.bin", index, m_min, n_min,
                             f"{tmp}/o.bin"])
                m, n, ptr, ind, v = orc.read_seldon_csr(f"{tmp}/o.bin", dt)
                out[f"{name}_{c}_mn"] = np.array([m, n])
                out[f"{name}_{c}_ptr"], out[f"{name}_{c}_ind"], out[f"{name}_{c}_val"] = ptr, ind, v
    np.savez_compressed(os.path.join(OUT, "coo_to_csr.npz"), spec=json.dumps(COO_CASES), **out)
    print("coo fixtures:", len(out))


def hb_cases():
    """Writer: name -> (matrix, storage).  Reader: name -> HB text in the
    formats Seldon's reader parses (E values; an RHS descriptor line; a
    pattern matrix; a rectangular one; rows unsorted within columns with a
    duplicated entry)."""
    P = wl.powerlaw(150, 41, kmin=2, kmax=12, half_width=40)
    writers = {"powerlaw150_row": (P, "row"), "powerlaw150_array": (P, "array"),
               "scrambled150_row": (scramble_rows(P, 42, False), "row"),
               "rect120x90_row": (wl.first_columns(wl.powerlaw(120, 43, kmin=2, kmax=12, half_width=40), 90), "row"),
               "laplace12_sym": (wl.upper(wl.laplace2d(12)), "sym")}
    rng = np.random.default_rng(44)
    readers = {}
    Q = wl.powerlaw(200, 45, kmin=2, kmax=10, half_width=50)
    S = Q.to_scipy().tocsc()
    readers["e4x20"] = hb_text_e(S.shape[0], S.shape[1], S.indptr, S.indices, S.data, "R", (8, 10), (8, 10),
                                 (4, 20, 12))
    R = wl.first_columns(wl.powerlaw(160, 46, kmin=2, kmax=10, half_width=50), 130).to_scipy().tocsc()
    readers["rect_rhs"] = hb_text_e(R.shape[0], R.shape[1], R.indptr, R.indices, R.data, "R", (10, 7), (12, 6),
                                    (3, 25, 16), rhs=True)
    readers["pattern"] = hb_text_e(S.shape[0], S.shape[1], S.indptr, S.indices, None, "P", (16, 5), (16, 5), None)
    # rows unsorted inside each column, and one entry given twice
    ptr, ind, dat = S.indptr.copy(), S.indices.copy(), S.data.copy()
    for j in range(S.shape[1]):
        a, b = ptr[j], ptr[j + 1]
        p = rng.permutation(b - a)
        ind[a:b], dat[a:b] = ind[a:b][p], dat[a:b][p]
    ind[ptr[3] + 1] = ind[ptr[3]]
    readers["unsorted_dup"] = hb_text_e(S.shape[0], S.shape[1], ptr, ind, dat, "R", (5, 16), (5, 16), (2, 40, 30))
    return writers, readers


def hb_text_e(m, n, cptr, rind, vals, kind, pf, inf, vf, rhs=False):
    """A Harwell-Boeing file with right-aligned integer fields and E-format
    values (Fortran widths pf = (per line, width) etc.), 1-based."""
    nnz = int(cptr[-1])
    cell = lambda k, w: (k + w - 1) // w
    nl_ptr, nl_ind = cell(n + 1, pf[0]), cell(nnz, inf[0])
    nl_val = 0 if kind == "P" else cell(nnz, vf[0])
    lines = [f"{'generated test matrix':<72}{'TEST0001':<8}",
             "".join(f"{x:14d}" for x in (nl_ptr + nl_ind + nl_val, nl_ptr, nl_ind, nl_val, 1 if rhs else 0)),
             f"{kind + ('UA' if m == n else 'RA'):<14}" + "".join(f"{x:14d}" for x in (m, n, nnz, 0)),
             f"{'(%dI%d)' % pf:<16}{'(%dI%d)' % inf:<16}" + (f"{'(%dE%d.%d)' % vf:<20}{'':<20}" if vf else "")]
    if rhs:
        lines.append(f"{'F':<14}{1:14d}{0:14d}")

    def block(items, per, w, fmt):
        for i in range(0, len(items), per):
            lines.append("".join(fmt(x).rjust(w) for x in items[i:i + per]))
    block([int(x) + 1 for x in cptr], pf[0], pf[1], str)
    block([int(x) + 1 for x in rind], inf[0], inf[1], str)
    if kind != "P":
        block(list(vals), vf[0], vf[1], lambda x: f"{float(x):.{vf[2]}E}")
    return ("\n".join(lines) + "\n").encode()


def make_hb():
    """Seldon's WriteHarwellBoeing output (the text) and what its
    ReadHarwellBoeing + coordinate conversion gives (the arrays)."""
    writers, readers = hb_cases()
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (A, st) in writers.items():
            for dt, c in ((np.float64, "d"), (np.float32, "f")):
                orc.write_seldon_csr(f"{tmp}/A.bin", A, dt)
                orc.run_ref(["hbwrite", st, c, f"{tmp}/A.bin", f"{tmp}/o.rua"])
                out[f"w_{name}_{c}"] = np.frombuffer(open(f"{tmp}/o.rua", "rb").read(), np.uint8)
        for name, text in readers.items():
            out[f"r_{name}_text"] = np.frombuffer(text, np.uint8)
            with open(f"{tmp}/i.rua", "wb") as f:
                f.write(text)
            for dt, c in ((np.float64, "d"), (np.float32, "f")):
                orc.run_ref(["hbread", c, f"{tmp}/i.rua", f"{tmp}/o.bin"])
                m, n, ptr, ind, val = orc.read_seldon_csr(f"{tmp}/o.bin", dt)
                out[f"r_{name}_{c}_dims"] = np.array([m, n])
                out[f"r_{name}_{c}_ptr"], out[f"r_{name}_{c}_ind"], out[f"r_{name}_{c}_val"] = ptr, ind, val
    np.savez_compressed(os.path.join(OUT, "harwell_boeing.npz"), spec=json.dumps({"writers": sorted(writers)}),
                        **out)
    print("hb fixtures:", len(out))


OVERLAY_ARGS = ["64", "24"]


def parse_sections(text):
    """'== label' sections of the overlay program's output, each parsed as the
    harness's history lines: {label: (hist, its, code, nb_iter)}."""
    sec, cur = {}, None
    for line in text.splitlines():
        if line.startswith("== "):
            cur = line[3:].strip()
            sec[cur] = []
        elif cur:
            sec[cur].append(line)
    return {k: orc.parse_ref_history("\n".join(v)) for k, v in sec.items()}


def make_overlay():
    """overlay_cpu.npz: the CPU legs of tests/cpp/seldon_overlay_check (Seldon's
    own Cg and fourteen more Krylov templates on its CPU storages, Seldon's
    WriteHarwellBoeing and Vector::Write), run here from the program built
    against the reference headers; the GPU test runs only its GPU legs and
    compares them with these.  The Harwell-Boeing files are kept as SHA-256
    digests (600 KB each); Seldon's vector file whole (the GPU legs read it)."""
    import hashlib
    import subprocess
    exe = os.path.join(orc.REF_DIR, "seldon_overlay_check")
    with tempfile.TemporaryDirectory() as tmp:
        out = subprocess.run([exe] + OVERLAY_ARGS + [tmp, "cpu"], capture_output=True, text=True, check=True,
                             timeout=600).stdout
        files = {k: open(os.path.join(tmp, k), "rb").read() for k in ("cpu.rua", "cpu_sym.rua", "b_cpu.bin")}
    secs = parse_sections(out)
    arrays = {}
    for label, (hist, its, code, nb) in secs.items():
        if label == "hb":
            continue
        arrays[f"{label}_hist"] = hist
        arrays[f"{label}_its"] = its
        arrays[f"{label}_code"] = code
        arrays[f"{label}_nb"] = nb
    np.savez_compressed(os.path.join(OUT, "overlay_cpu.npz"), args=json.dumps(OVERLAY_ARGS),
                        labels=json.dumps(sorted(k for k in secs if k != "hb")),
                        hb_sha256=hashlib.sha256(files["cpu.rua"]).hexdigest(),
                        hb_sym_sha256=hashlib.sha256(files["cpu_sym.rua"]).hexdigest(),
                        hb_size=len(files["cpu.rua"]), hb_sym_size=len(files["cpu_sym.rua"]),
                        b_cpu=np.frombuffer(files["b_cpu.bin"], np.uint8), **arrays)
    print("overlay fixture:", sorted(k for k in secs if k != "hb"))


def make_files():
    """seldon_files.npz: files written by the real Seldon, as bytes: the CSR
    file of ConvertMatrix_from_Coordinates + Matrix::Write on the "dup"
    coordinates, and Vector::Write of a double and of a float vector (from
    its SOR driver: one forward sweep of 2 I leaves x = b / 2)."""
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        row, col, val, index, m_min, n_min = coo_case("dup")
        orc.write_seldon_vec(f"{tmp}/r.bin", row)
        orc.write_seldon_vec(f"{tmp}/c.bin", col)
        orc.write_seldon_vec(f"{tmp}/v.bin", val)
        orc.run_ref(["coo2csr", "d", f"{tmp}/r.bin", f"{tmp}/c.bin", f"{tmp}/v.bin", index, m_min, n_min,
                     f"{tmp}/A.bin"])
        out["csr_dup_d"] = np.frombuffer(open(f"{tmp}/A.bin", "rb").read(), np.uint8)
        for c, dt in (("d", np.float64), ("f", np.float32)):
            v = wl.uniform(1000, 17).astype(dt)
            n = v.size
            A = wl.CSR(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.full(n, 2.0))
            orc.write_seldon_csr(f"{tmp}/D.bin", A, dt)
            orc.write_seldon_vec(f"{tmp}/x.bin", np.zeros(n, dt))
            orc.write_seldon_vec(f"{tmp}/b.bin", v)
            orc.run_ref(["sor", "row", c, f"{tmp}/D.bin", f"{tmp}/x.bin", f"{tmp}/b.bin", "1.0", 1, 2,
                         f"{tmp}/s.bin"])
            out[f"vec_{c}"] = np.frombuffer(open(f"{tmp}/s.bin", "rb").read(), np.uint8)
    np.savez_compressed(os.path.join(OUT, "seldon_files.npz"),
                        spec=json.dumps(dict(coo="dup", vec_n=1000, vec_seed=17)), **out)
    print("seldon files:", {k: v.size for k, v in out.items()})


if __name__ == "__main__":
    if not orc.ref_available("generic") or not orc.ref_available("blas"):
        orc.build()
        import subprocess
        subprocess.run(["make", "-s", "-C", orc.HERE, "ref"], check=True)
    which = sys.argv[1:] or ["spmv", "spmvt", "coo", "sor", "hb", "solves", "solves_f32", "overlay", "files"]
    if "hb" in which:
        make_hb()
    if "coo" in which:
        make_coo()
    if "sor" in which:
        make_sor()
    if "spmv" in which:
        make_spmv()
    if "spmvt" in which:
        make_spmvt()
    if "solves" in which:
        make_solves()
    if "solves_f32" in which:
        make_solves_f32()
    if "overlay" in which:
        make_overlay()
    if "files" in which:
        make_files()
