"""Benchmark of the sparse Krylov hot path on one MI355X.

Workload (the matrix the SpMV %-of-peak figure is quoted on): the
unstructured-FEM-like power-law matrix, 4,194,304 rows, row lengths
P(k) ~ k^-1.7452 on [4, 512] (mean 27), columns within +-20,000 of the
diagonal, N(0,1) off-diagonals, diagonal = 1.1 x row abs-sum; ~113.2M stored
entries (1.44 GB of FP64 CSR, 0.96 GB FP32).  Synthetic, seeded.

A step = one CSR SpMV y = A x over the whole matrix (Seldon's
Mlt(A, X, Y), FP64), inputs resident in HBM.  `value` = algorithmic GB/s of
the timed steps; `roofline` = the same bytes over the kernel's own average
duration from HIP events on the stream it runs on.  Also reported: FP32 SpMV,
time to a 1e-10 relative true residual with plain FP64 GMRES(50) and with the
FP32 GMRES(50) + FP64 iterative refinement solver, and the oracle's CPU SpMV
on the same matrix as `cpu_baseline`.

Also: the 128^3 convection-diffusion BiCgStab / Gmres(30) solves and the
512-system batched Cg (FP64 and FP32+IR), each timed to its tolerance.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 10] [--quick] [--no-cpu] [--only convdiff128]
                       [--no-solve] [--no-extra] [--probe]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy kernel of known size), committed under
    profiles/ as JSON."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--quick", action="store_true", help="smaller matrix (1M rows)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--layout", default="auto", choices=["auto", "rows", "panels", "sliced"],
                    help="device layout of the benchmark matrix (auto: sliced for RowSparse matrices of 2M rows or more)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 128^3 convection-diffusion solves and the batched Cg")
    ap.add_argument("--only", choices=["convdiff128", "batched"],
                    help="run one extra section alone and print its JSON (for profiling its kernels)")
    ap.add_argument("--probe", action="store_true",
                    help="profiling run: one 1 GiB copy kernel (counter calibration) and 40 FP64 + 40 FP32 SpMVs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU only: the sparse Krylov path is "
                                   "measured on a single MI355X; multi-GPU sharding is design-only"}))
        sys.exit(2)

    import numpy as np
    import torch

    from seldon_amd import krylov as K
    from seldon_amd import workloads as wl

    if args.only == "convdiff128":
        torch.cuda.set_device(0)
        print(json.dumps({"convdiff128": convdiff_solves(K, wl, np, torch, cpu=not args.no_cpu, layout=args.layout)}),
              flush=True)
        return
    if args.only == "batched":
        torch.cuda.set_device(0)
        print(json.dumps({"batched": batched_cg(K, np, torch, cpu=not args.no_cpu)}), flush=True)
        return

    n_rows = (1 << 20) if args.quick else (1 << 22)
    t0 = time.time()
    A, b, xt = wl.fem_powerlaw(n_rows)
    log(f"matrix {A.m} rows, {A.nnz} nnz, built in {time.time() - t0:.1f}s")
    torch.cuda.set_device(0)
    ctx = K.Context.default()
    # upload + layout planning, timed apart from every solve (host arrays to
    # the device, the layout's plan built on the host)
    t0 = time.perf_counter()
    A64 = K.Matrix(A, dtype=torch.float64, layout=args.layout)
    torch.cuda.synchronize()
    up64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    A32 = K.Matrix(A, dtype=torch.float32, layout=args.layout)
    torch.cuda.synchronize()
    up32 = time.perf_counter() - t0
    log(f"device layout: fp64 {A64.layout} ({up64:.2f}s upload+plan), fp32 {A32.layout} ({up32:.2f}s)")
    x64 = torch.tensor(xt, device="cuda")
    y64 = torch.empty_like(x64)
    x32 = x64.float()
    y32 = torch.empty_like(x32)
    bytes64 = A64.spmv_bytes(False)
    bytes32 = A32.spmv_bytes(False)

    if args.probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        for _ in range(40):
            K.Mlt(A64, x64, y64)
        for _ in range(40):
            K.Mlt(A32, x32, y32)
        torch.cuda.synchronize()
        log("probe done")
        return

    for _ in range(args.warmup):
        K.Mlt(A64, x64, y64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        K.Mlt(A64, x64, y64)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    value = bytes64 / (ms_per_step * 1e-3) / 1e9
    log(f"fp64 spmv {ms_per_step:.3f} ms/step = {value:.0f} GB/s")

    # kernel duration from HIP events on the library's stream
    import ctypes

    from seldon_amd import _lib as L
    avg = ctypes.c_double()
    L.check(L.load().shk_time_spmv(ctx.h, A64.h, 0, K._ptr(x64), K._ptr(y64), max(args.steps, 5),
                                   ctypes.byref(avg)))
    k64_ms = avg.value
    achieved = bytes64 / (k64_ms * 1e-3) / 1e9
    L.check(L.load().shk_time_spmv(ctx.h, A32.h, 0, K._ptr(x32), K._ptr(y32), max(args.steps, 5),
                                   ctypes.byref(avg)))
    k32_ms = avg.value
    achieved32 = bytes32 / (k32_ms * 1e-3) / 1e9
    log(f"kernel fp64 {k64_ms:.3f} ms ({achieved:.0f} GB/s), fp32 {k32_ms:.3f} ms ({achieved32:.0f} GB/s)")

    result = {
        "metric": "SpMV HBM GB/s (%peak); time-to-1e-10 residual, FP32+IR vs FP64 Krylov",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded power-law FEM-like CSR)",
        "config": {"workload": "FEM-like power-law CSR SpMV, FP64, Mlt(A,X,Y)", "layout": A64.layout,
                   "rows": A.m, "nnz": A.nnz, "row_len": "P(k)~k^-1.7452 on [4,512], mean 27",
                   "band_half_width": 20000, "algorithmic_bytes_per_step": bytes64},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles("spmv_f64")},
        "spmv_f32": {"kernel_ms": round(k32_ms, 5), "achieved_GBs": round(achieved32, 2),
                     "frac": round(achieved32 / HBM_PEAK_GBS, 4), "bytes": bytes32},
        "spmv_f64_kernel_ms": round(k64_ms, 5),
        "upload_plan_seconds": {"fp64": round(up64, 3), "fp32": round(up32, 3)},
    }

    result["transposed"] = transposed_product(K, A64, x64, y64, bytes64, torch)

    if not args.no_solve:
        result["solve"] = solve_compare(K, A, A64, A32, b, xt, np, torch, cpu=not args.no_cpu)

    if not args.no_extra:
        del A64, A32, x64, y64, x32, y32
        torch.cuda.empty_cache()
        result["assembly"] = assembly(K, A, np, torch, cpu=not args.no_cpu)
        result["cg_laplace256"] = cg_config1(K, wl, np, torch, cpu=not args.no_cpu)
        torch.cuda.empty_cache()
        result["convdiff128"] = convdiff_solves(K, wl, np, torch, cpu=not args.no_cpu)
        torch.cuda.empty_cache()
        result["batched"] = batched_cg(K, np, torch, cpu=not args.no_cpu)
        torch.cuda.empty_cache()
        result["diffusion64_gradient"] = diffusion_gradient(wl, np, torch, cpu=not args.no_cpu)

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(A, xt, bytes64)
        result["host"] = host_info()
    print(json.dumps(result), flush=True)


def transposed_product(K, A64, x64, y64, nbytes, torch, reps=10):
    """Mlt(SeldonTrans, A, X, Y) on the power-law matrix: the first call
    builds A^T (device sort of the coordinates, layout plan on the host,
    upload), the later ones are one SpMV with it."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    K.Mlt(K.SeldonTrans, A64, x64, y64)
    torch.cuda.synchronize()
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(reps):
        K.Mlt(K.SeldonTrans, A64, x64, y64)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    log(f"transposed product: first call (builds A^T) {first:.3f}s, then {ms:.3f} ms per call")
    return {"first_call_seconds": round(first, 3), "ms_per_call": round(ms, 4),
            "GBs_per_call": round(nbytes / (ms * 1e-3) / 1e9, 1)}


def host_info():
    """The CPU the baselines ran on (lscpu's model name and counts)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "oracle_build": ORACLE_BUILD}


ORACLE_BUILD = "gcc -O3 -march=x86-64-v3 -ffp-contract=off -fno-fast-math (oracle/Makefile)"


def cpu_threads():
    """Worker threads for the CPU batch: the host cores this process may use,
    at most 16 (one GPU's share of the host)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def solve_compare(K, A, A64, A32, b, xt, np, torch, cpu=True):
    """Plain FP64 GMRES(50) vs FP32 GMRES(50) + FP64 refinement, both to a
    1e-10 relative true residual (Jacobi preconditioner).  With `cpu`, the
    oracle's FP64 Gmres(50) (Seldon's template restated in C, one core) runs
    the same solve: its time, and the error of both GPU solutions against
    its solution."""
    out = {}
    x_orc = None
    if cpu:
        from oracle import oracle as orc
        t0 = time.perf_counter()
        r = orc.solve("gmres", A, b, precond="jacobi", tol=1e-10, max_iter=5000, restart=50)
        ct = time.perf_counter() - t0
        x_orc = r.x
        out["cpu_oracle_fp64_gmres50"] = {"seconds": round(ct, 3), "iterations": r.nb_iter, "code": r.code,
                                          "cores": 1, "kind": "port"}
        log(f"oracle fp64 gmres(50) (1 core): {ct:.2f}s, {r.nb_iter} its")
    bg = torch.tensor(b, device="cuda")
    bn = float(np.linalg.norm(b))
    xg = torch.zeros_like(bg)
    it = K.Iteration(5000, 1e-10)
    it.SetRestart(50)
    it.HideMessages()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    code = K.Gmres(A64, xg, bg, K.JacobiPreconditioner(), it)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    r = torch.empty_like(bg)
    K.Copy(bg, r)
    K.MltAdd(-1.0, A64, xg, 1.0, r)
    rel64 = K.Norm2(r) / bn
    x64 = xg.clone()
    out["fp64_gmres50"] = {"seconds": round(t64, 4), "iterations": it.nb_iter, "code": code,
                           "true_relres": rel64}
    xtn = float(np.linalg.norm(xt))
    out["fp64_gmres50"]["rel_error_vs_xtrue"] = float(np.linalg.norm(x64.cpu().numpy() - xt)) / xtn
    if x_orc is not None:
        out["fp64_gmres50"]["rel_error_vs_fp64_oracle"] = float(
            np.linalg.norm(x64.cpu().numpy() - x_orc) / np.linalg.norm(x_orc))
    log(f"fp64 gmres(50): {t64:.3f}s, {it.nb_iter} its, true relres {rel64:.2e}")
    try:
        xg = torch.zeros_like(bg)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = K.solve_ir(A64, A32, xg, bg, K.JacobiPreconditioner(), tol=1e-10, inner_restart=50)
        torch.cuda.synchronize()
        tir = time.perf_counter() - t0
        K.Copy(bg, r)
        K.MltAdd(-1.0, A64, xg, 1.0, r)
        relir = K.Norm2(r) / bn
        err = float(torch.linalg.norm(xg - x64) / torch.linalg.norm(x64))
        xi = xg.cpu().numpy()
        out["fp32_ir_gmres50"] = {"seconds": round(tir, 4), "refinements": info["refinements"],
                                  "inner_iterations": info["inner_iters"], "true_relres": relir,
                                  "rel_error_vs_fp64_gpu": err, "speedup_vs_fp64": round(t64 / tir, 3),
                                  "rel_error_vs_xtrue": float(np.linalg.norm(xi - xt)) / xtn}
        if x_orc is not None:
            out["fp32_ir_gmres50"]["rel_error_vs_fp64_oracle"] = float(
                np.linalg.norm(xi - x_orc) / np.linalg.norm(x_orc))
        log(f"fp32+IR: {tir:.3f}s, {info['refinements']} refinements, {info['inner_iters']} inner its, "
            f"relres {relir:.2e}, err vs fp64 {err:.2e}")
    except Exception as e:  # reported, not hidden
        out["fp32_ir_gmres50"] = {"error": str(e)}
    return out


def _true_relres(K, A, x, b, bn):
    import torch
    r = torch.empty_like(b)
    K.Copy(b, r)
    K.MltAdd(-1.0, A, x, 1.0, r)
    return K.Norm2(r) / bn


def cg_config1(K, wl, np, torch, cpu=True):
    """Plain FP64 Cg, identity preconditioner, on the 256 x 256 5-point
    Laplacian (65,536 rows), b = A x_true, to a 1e-10 relative residual: the
    reference's own CPU-runnable case.  Time to solution (second of two runs),
    iterations, and the oracle's Cg (Seldon's generic template restated in C,
    one core) on the same system."""
    A, b, xt = wl.cg_laplace(256)
    Ad = K.Matrix(A, dtype=torch.float64)
    bg = torch.tensor(b, device="cuda")
    out = {"rows": A.m, "nnz": A.nnz}
    for _ in range(2):
        x = torch.zeros_like(bg)
        it = K.Iteration(10000, 1e-10)
        it.HideMessages()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        code = K.Cg(Ad, x, bg, K.Preconditioner_Base(), it)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    err = float(np.abs(x.cpu().numpy() - xt).max())
    out["gpu"] = {"seconds": round(dt, 4), "iterations": it.nb_iter, "code": code,
                  "us_per_iteration": round(dt / max(it.nb_iter, 1) * 1e6, 2), "max_abs_err_vs_xtrue": err}
    log(f"config-1 Cg: {dt * 1e3:.1f} ms, {it.nb_iter} its ({dt / max(it.nb_iter, 1) * 1e6:.1f} us/it), code {code}")
    if cpu:
        from oracle import oracle as orc
        t0 = time.perf_counter()
        r = orc.solve("cg", A, b, tol=1e-10, max_iter=10000)
        ct = time.perf_counter() - t0
        out["cpu_oracle"] = {"seconds": round(ct, 4), "iterations": r.nb_iter, "cores": 1, "kind": "port"}
        log(f"config-1 Cg on the oracle (1 core): {ct * 1e3:.1f} ms, {r.nb_iter} its")
    return out


def convdiff_solves(K, wl, np, torch, cpu=True, layout="auto"):
    """FP64 BiCgStab, Gmres(30), Cgs, BiCg, TfQmr, Qmr, Cgne, Lsqr and BiCgStab(10) with the Jacobi preconditioner
    on the 128^3 7-point upwind convection-diffusion matrix (2,097,152 rows),
    to a 1e-10 relative residual: time to solution and iterations/s, plus the
    SpMV bandwidth on this matrix and the time of one symmetric SOR sweep
    (with the oracle's sequential sweep on one core beside it)."""
    import ctypes

    from seldon_amd import _lib as L
    A, b, xt = wl.convdiff(128)
    Ad = K.Matrix(A, dtype=torch.float64, layout=layout)
    bg = torch.tensor(b, device="cuda")
    bn = float(np.linalg.norm(b))
    out = {"rows": A.m, "nnz": A.nnz, "layout": Ad.layout}
    y = torch.empty_like(bg)
    avg = ctypes.c_double()
    L.check(L.load().shk_time_spmv(Ad.ctx.h, Ad.h, 0, K._ptr(bg), K._ptr(y), 20, ctypes.byref(avg)))
    sb = Ad.spmv_bytes(False)
    out["spmv"] = {"kernel_ms": round(avg.value, 5), "GBs": round(sb / (avg.value * 1e-3) / 1e9, 2),
                   "frac": round(sb / (avg.value * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "bytes": sb}
    # one symmetric SOR sweep (forward then backward, omega 1), level-scheduled
    x = torch.zeros_like(bg)
    K.SOR(Ad, x, bg, 1.0, 1, 0)  # builds the level plan
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        K.SOR(Ad, x, bg, 1.0, 1, 0)
    torch.cuda.synchronize()
    st = (time.perf_counter() - t0) / 5
    from oracle import oracle as orc
    t0 = time.perf_counter()
    orc.sor(A, np.zeros(A.m), b, 1.0, 1, 0)
    ct = time.perf_counter() - t0
    out["ssor_sweep"] = {"seconds": round(st, 5), "cpu_oracle_seconds": round(ct, 4), "cores": 1}
    log(f"convdiff128 SSOR sweep: {st * 1e3:.2f} ms (oracle, 1 core: {ct * 1e3:.1f} ms)")
    # BiCg's A^T (built on the device on first use) is made before its timed
    # solve; the first call's time is recorded beside it
    t0 = time.perf_counter()
    it = K.Iteration(2, 1e-10)
    it.HideMessages()
    K.BiCg(Ad, torch.zeros_like(bg), bg, K.JacobiPreconditioner(), it)
    torch.cuda.synchronize()
    bicg_first = time.perf_counter() - t0
    for name, fn, restart in (("bicgstab", K.BiCgStab, None), ("gmres30", K.Gmres, 30), ("cgs", K.Cgs, None),
                              ("bicg", K.BiCg, None), ("tfqmr", K.TfQmr, None), ("qmr", K.Qmr, None),
                              ("cgne", K.Cgne, None), ("lsqr", K.Lsqr, None), ("bicgstabl", K.BiCgStabl, None)):
        x = torch.zeros_like(bg)
        it = K.Iteration(20000, 1e-10)
        if restart:
            it.SetRestart(restart)
        it.HideMessages()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        code = fn(Ad, x, bg, K.JacobiPreconditioner(), it)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        rel = _true_relres(K, Ad, x, bg, bn)
        out[name] = {"seconds": round(dt, 4), "iterations": it.nb_iter, "code": code,
                     "iterations_per_s": round(it.nb_iter / dt, 1), "true_relres": rel}
        if name == "bicg":
            out[name]["first_call_seconds_with_transpose_build"] = round(bicg_first, 4)
        log(f"convdiff128 {name}: {dt:.3f}s, {it.nb_iter} its, code {code}, true relres {rel:.2e}")
        if cpu and name in ("bicg", "qmr", "cgne", "lsqr"):
            # Seldon's BiCg, Qmr, Cgne and Lsqr take A^T products through the generic transposed
            # product (Functions_MatVect.cxx:1279-1316), a dense loop over
            # every (row, column) pair: 4.4e12 terms per product here
            out[name]["cpu_oracle"] = {"skipped": "Mlt(SeldonTrans) on RowSparse is the generic dense loop, "
                                                  "O(rows x columns) per product"}
        elif cpu:
            # the oracle's solver on one core, a bounded sample of its
            # iterations (a full solve takes minutes); time to solution is
            # projected from its iteration rate and the GPU's count
            cap = 40
            t0 = time.perf_counter()
            r = orc.solve("gmres" if name == "gmres30" else name, A, b, precond="jacobi", tol=1e-10,
                          max_iter=cap, restart=restart or 10)
            ct = time.perf_counter() - t0
            ips = r.nb_iter / ct
            out[name]["cpu_oracle"] = {"sample_iterations": r.nb_iter, "seconds": round(ct, 3),
                                       "iterations_per_s": round(ips, 2),
                                       "projected_seconds_to_solution": round(it.nb_iter / ips, 2),
                                       "cores": 1, "kind": "port"}
            log(f"convdiff128 {name} on the oracle (1 core): {r.nb_iter} its in {ct:.2f}s")
    return out


def batched_cg(K, np, torch, cpu=True):
    """512 anisotropic 2-D Laplacians (n x n, n in [32, 256], anisotropy
    log-uniform in [1, 100]) solved to 1e-8 by the persistent batched Cg,
    FP64 and FP32 inner + FP64 refinement, each under hipGraph with the
    autotuned work-group size and rows per pass (and, mixed, inner
    tolerance).  Bytes per Cg iteration of system s (FP64):
    (4+8) nnz_s + 4 (n_s+1) for the matrix and 11 n_s x 8 for the vectors
    (p = r + beta p: 3; q = A p: 2; x += alpha p, r -= alpha q, |r|^2: 6);
    the mixed variant's figure counts its FP32 inner iterations only (its
    FP64 residuals at each refinement are not in it)."""
    from seldon_amd import batched as B
    A, row_off, b, xt, sides, eps = B.aniso_batch()
    A64 = K.Matrix(A, dtype=torch.float64, layout="rows")
    A32 = K.Matrix(A, dtype=torch.float32, layout="rows")
    bg = torch.tensor(b, device="cuda")
    n_s = np.diff(row_off).astype(np.float64)
    nnz_s = (A.ptr[row_off[1:]] - A.ptr[row_off[:-1]]).astype(np.float64)
    out = {"systems": len(n_s), "rows": A.m, "nnz": A.nnz}
    for variant in ("fp64", "ir"):
        (group, rpp), entry = B.autotune(A64, row_off, bg, variant=variant, A32=A32, use_cache=False)
        x = torch.zeros_like(bg)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if variant == "fp64":
            its, codes = B.cg_batched(A64, row_off, x, bg, tol=1e-8, group=group, use_graph=True,
                                      rows_per_pass=rpp)
        else:
            its, codes = B.cg_batched_ir(A64, A32, row_off, x, bg, tol=1e-8, group=group, use_graph=True,
                                         rows_per_pass=rpp, inner_tol=entry["inner_tol"])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        V = 8.0 if variant == "fp64" else 4.0
        per_it = (4 + V) * nnz_s + 4 * (n_s + 1) + 11 * n_s * V
        gbs = float((per_it * its).sum()) / dt / 1e9
        err = float(torch.linalg.norm(x - torch.tensor(xt, device="cuda")) / np.linalg.norm(xt))
        out[variant] = {"seconds": round(dt, 4), "group": group, "rows_per_pass": rpp,
                        "tuned_seconds": entry["seconds"],
                        "iterations_total": int(its.sum()), "iterations_max": int(its.max()),
                        "iterations_per_s": round(float(its.sum()) / dt, 1),
                        "codes_nonzero": int((codes != 0).sum()),
                        ("GBs" if variant == "fp64" else "inner_GBs"): round(gbs, 2),
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "rel_error_vs_xtrue": err}
        if variant == "ir":
            out[variant]["inner_tol"] = entry["inner_tol"]
            out[variant]["tuned_inner_tol_seconds"] = entry["inner_tol_seconds"]
            out[variant]["speedup_vs_fp64"] = round(out["fp64"]["seconds"] / dt, 3)
        log(f"batched {variant}: {dt:.3f}s group {group}, {int(its.sum())} its, {gbs:.0f} GB/s, "
            f"err {err:.2e}, nonzero codes {int((codes != 0).sum())}")
    if cpu:
        out["cpu_oracle"] = cpu_batched(A, row_off, b, np)
    return out


def diffusion_gradient(wl, np, torch, cpu=True, n=64, fd_samples=4):
    """The differentiable solve on -div(exp(theta) grad u) over a 64^3 grid
    (262,144 rows, SPD): loss = |x - x_obs|^2 with x = A(theta)^{-1} b; FP64
    Cg forward (Jacobi, 1e-14) and the adjoint Cg plus the SDDMM gradient in
    backward, through torch.autograd.  Reported: forward and backward time,
    their iterations, and the gradients' relative error against the oracle's
    FP64 CPU gradient (its Cg forward and adjoint, SDDMM and chain rule;
    timed on one core) and against central finite differences (h = 1e-6) at
    a few theta and b components."""
    from seldon_amd.adjoint import DiffusionOperator
    N = n ** 3
    theta = wl.normal(N, 2025, 0.0, 0.5)
    b = wl.uniform(N, 2026, 0.0, 1.0)
    dtheta = wl.normal(N, 2027, 0.0, 0.1)
    op = DiffusionOperator(n)
    th = torch.tensor(theta, device="cuda")
    bt = torch.tensor(b, device="cuda")
    with torch.no_grad():
        x_obs = op.solve(op.values(th + torch.tensor(dtheta, device="cuda")), bt, tol=1e-14)
    out = {"rows": N, "nnz": op.nnz}
    for rep in range(2):  # the second run is timed (first-use costs excluded)
        tt = th.clone().requires_grad_(True)
        bb = bt.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vals = op.values(tt)
        x = op.solve(vals, bb, tol=1e-14)
        loss = torch.sum((x - x_obs) ** 2)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    gth = tt.grad.detach().cpu().numpy()
    gb = bb.grad.detach().cpu().numpy()
    out["gpu"] = {"forward_seconds": round(t1 - t0, 4), "backward_seconds": round(t2 - t1, 4),
                  "loss": float(loss)}
    log(f"diffusion64 gradient: forward {1e3 * (t1 - t0):.1f} ms, backward {1e3 * (t2 - t1):.1f} ms")
    if cpu:
        from oracle import oracle as orc
        t0 = time.perf_counter()
        ref = orc.diffusion_grad_reference(n, theta, vals.detach().cpu().numpy(), op.pattern, b,
                                           x_obs.cpu().numpy(), tol=1e-14)
        ct = time.perf_counter() - t0
        out["cpu_oracle"] = {"seconds": round(ct, 3), "cg_iterations": list(ref["its"]), "cores": 1, "kind": "port"}
        out["rel_error_vs_fp64_oracle"] = {
            "theta": float(np.linalg.norm(gth - ref["grad_theta"]) / np.linalg.norm(ref["grad_theta"])),
            "b": float(np.linalg.norm(gb - ref["grad_b"]) / np.linalg.norm(ref["grad_b"]))}
        log(f"diffusion64 gradient vs the oracle's: {out['rel_error_vs_fp64_oracle']}, oracle {ct:.2f}s")
    h = 1e-6

    def L(th_, b_):
        with torch.no_grad():
            xx = op.solve(op.values(th_), b_, tol=1e-14)
            return float(torch.sum((xx - x_obs) ** 2))

    rng = np.random.default_rng(64)
    worst = 0.0
    for grad, which in ((gth, "theta"), (gb, "b")):
        scale = float(np.max(np.abs(grad)))
        for j in rng.choice(N, fd_samples, replace=False):
            e = torch.zeros_like(th)
            e[j] = h
            if which == "theta":
                fd = (L(th + e, bt) - L(th - e, bt)) / (2 * h)
            else:
                fd = (L(th, bt + e) - L(th, bt - e)) / (2 * h)
            worst = max(worst, abs(fd - grad[j]) / scale)
    out["central_fd_max_deviation_of_max_grad"] = worst
    return out


def cpu_batched(A, row_off, b, np):
    """The oracle's FP64 Cg (Seldon's template restated in C) on every one of
    the systems, to 1e-8, spread over the host cores in a thread pool (each
    solve is single-threaded, as Seldon's)."""
    import concurrent.futures as cf

    from oracle import oracle as orc
    from seldon_amd import workloads as wl

    def system(k):
        r0, r1 = int(row_off[k]), int(row_off[k + 1])
        e0, e1 = int(A.ptr[r0]), int(A.ptr[r1])
        As = wl.CSR(r1 - r0, r1 - r0, (A.ptr[r0:r1 + 1] - e0).astype(np.int32),
                    (A.ind[e0:e1] - r0).astype(np.int32), A.val[e0:e1])
        return As, b[r0:r1]

    mats = [system(k) for k in range(len(row_off) - 1)]

    def run(k):
        r = orc.solve("cg", mats[k][0], mats[k][1], tol=1e-8, max_iter=20000)
        return r.nb_iter, r.code

    nt = cpu_threads()
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=nt) as ex:
        res = list(ex.map(run, range(len(mats))))
    ct = time.perf_counter() - t0
    its = int(sum(r[0] for r in res))
    log(f"batched fp64 on the oracle ({nt} threads): {ct:.2f}s, {its} its")
    return {"seconds": round(ct, 3), "iterations_total": its, "codes_nonzero": int(sum(r[1] != 0 for r in res)),
            "iterations_per_s": round(its / ct, 1), "cores": nt, "kind": "port"}


def assembly(K, A, np, torch, cpu=True, reps=5):
    """ConvertMatrix_from_Coordinates on the device: the power-law matrix's
    113.2M entries as shuffled FP64 coordinates, to a RowSparse CSR (stable
    (row, column) order).  Bytes counted: read row, col, val (16 B per entry),
    write ind and val (12 B) and ptr (4(m+1)); the sort's own passes are on
    top.  The oracle (Seldon's MergeSort conversion restated in C, one core)
    is timed on a 1M-entry slice of the same coordinates."""
    rng = np.random.default_rng(77)
    row = np.repeat(np.arange(A.m, dtype=np.int32), np.diff(A.ptr))
    perm = rng.permutation(row.size)
    rg = torch.from_numpy(row[perm]).cuda()
    cg = torch.from_numpy(A.ind[perm].astype(np.int32)).cuda()
    vg = torch.from_numpy(A.val[perm]).cuda()
    K.ConvertMatrix_from_Coordinates(rg, cg, vg)  # warm-up (workspace, rocPRIM setup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        m, n, ptr, ind, out = K.ConvertMatrix_from_Coordinates(rg, cg, vg)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    ok = bool(torch.equal(ptr.cpu(), torch.from_numpy(A.ptr)) and torch.equal(ind.cpu(), torch.from_numpy(A.ind)))
    nbytes = 28.0 * row.size + 4.0 * (A.m + 1)
    res = {"entries": int(row.size), "seconds": round(dt, 5), "entries_per_s": round(row.size / dt, 1),
           "GBs": round(nbytes / dt / 1e9, 2), "matches_original_csr": ok}
    log(f"assembly: {row.size} coordinates -> CSR in {dt * 1e3:.1f} ms ({row.size / dt / 1e9:.2f} G entries/s), "
        f"same CSR as generated: {ok}")
    if cpu:
        from oracle import oracle as orc
        k = 1 << 20
        t0 = time.perf_counter()
        orc.coo_to_csr(row[perm[:k]], A.ind[perm[:k]], A.val[perm[:k]])
        ct = time.perf_counter() - t0
        res["cpu_oracle"] = {"entries": k, "seconds": round(ct, 4), "entries_per_s": round(k / ct, 1), "cores": 1,
                             "kind": "port"}
    return res


def cpu_baseline(A, xt, bytes64, budget_s=10.0):
    """The oracle's FP64 SpMV (Seldon's generic RowSparse loop restated in C,
    one core) on the same matrix; a bounded sample of ~10 s."""
    from oracle import oracle as orc
    reps, t = 0, 0.0
    while t < budget_s and reps < 200:
        t0 = time.perf_counter()
        orc.mlt(A, xt)
        t += time.perf_counter() - t0
        reps += 1
    gbs = bytes64 * reps / t / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{reps} FP64 SpMVs of the same 4M-row matrix, {t:.1f}s, oracle C "
                      f"(-O3 -march=x86-64-v3 -ffp-contract=off, 1 thread)"}


if __name__ == "__main__":
    main()
