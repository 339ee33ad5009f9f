"""Batched Cg over many independent SPD systems (one persistent kernel).

``stack(systems)`` lays the systems out block-diagonally in one CSR;
``cg_batched`` runs Seldon's Cg (identity preconditioner) on every system
in one launch of the library's persistent kernel (work-group per system,
device work queue, largest systems first), optionally captured in a hipGraph;
``cg_batched_ir`` is the FP32-inner / FP64-refinement variant.

The stacked matrices must use the row layout: ``K.Matrix(A, layout="rows")``.

``autotune`` picks the work-group size ("group": threads per system) and the
rows each thread keeps in flight per pass ("rows_per_pass") by timing every
pair on the batch itself -- and for the mixed variant then the inner solves'
tolerance ("inner_tol": a looser one means more refinements of shorter FP32
solves) -- and caches the choice per batch signature (variant, number of
systems, rows, stored entries) in ``autotune.json`` next to this file.
"""
from __future__ import annotations

import ctypes
import json
import os
import time

import numpy as np
import torch

from . import _lib as L
from . import krylov as K
from . import workloads as wl

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune.json")
CANDIDATES = (256, 512, 1024)   # threads per system work-group
ROWS_PER_PASS = (1, 2, 4, 8)     # rows a thread keeps in flight per pass
INNER_TOLS = (1e-3, 1e-4, 1e-5)  # mixed variant: each FP32 inner solve's residual reduction


def stack(systems):
    """Block-diagonal CSR of a list of CSR systems; returns (A, row_off)."""
    m = sum(s.m for s in systems)
    nnz = sum(s.nnz for s in systems)
    ptr = np.empty(m + 1, np.int64)
    ind = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    row_off = np.zeros(len(systems) + 1, np.int32)
    r = e = 0
    ptr[0] = 0
    for k, s in enumerate(systems):
        ptr[r + 1:r + s.m + 1] = s.ptr[1:].astype(np.int64) + e
        ind[e:e + s.nnz] = s.ind + r
        val[e:e + s.nnz] = s.val
        r += s.m
        e += s.nnz
        row_off[k + 1] = r
    return wl.CSR(m, m, ptr.astype(np.int32), ind, val), row_off


def aniso_batch(count=512, seed=77, nmin=32, nmax=256, xseed=wl.SEED_XTRUE + 5):
    """The batched workload: 2-D anisotropic 5-point Laplacians of random
    side and anisotropy; b_s = A_s x_true,s with x_true ~ U[-1,1]."""
    sides, eps = wl.batched_systems(count, seed, nmin, nmax)
    systems = [wl.aniso2d(int(n), int(n), float(e)) for n, e in zip(sides, eps)]
    A, row_off = stack(systems)
    xt = wl.uniform(A.m, xseed)
    b = wl.apply(A, xt)
    return A, row_off, b, xt, sides, eps


def _run(A64, row_off, x, b, tol, max_iter, group, use_graph, A32=None, inner_tol=1e-4, rows_per_pass=0):
    count = len(row_off) - 1
    its = np.zeros(count, np.int32)
    codes = np.zeros(count, np.int32)
    ro = np.ascontiguousarray(row_off, np.int32)
    lib = L.load()
    if A32 is None:
        L.check(lib.shk_cg_batched(A64.ctx.h, A64.h, count, ro.ctypes.data_as(ctypes.c_void_p), K._ptr(x), K._ptr(b),
                                   tol, max_iter, group, rows_per_pass, int(use_graph),
                                   its.ctypes.data_as(ctypes.c_void_p), codes.ctypes.data_as(ctypes.c_void_p)))
    else:
        L.check(lib.shk_cg_batched_ir(A64.ctx.h, A64.h, A32.h, count, ro.ctypes.data_as(ctypes.c_void_p), K._ptr(x),
                                      K._ptr(b), tol, max_iter, inner_tol, group, rows_per_pass, int(use_graph),
                                      its.ctypes.data_as(ctypes.c_void_p), codes.ctypes.data_as(ctypes.c_void_p)))
    return its, codes


def cg_batched(A64: "K.Matrix", row_off, x, b, tol=1e-8, max_iter=100000, group=0, use_graph=True,
               rows_per_pass=0):
    return _run(A64, row_off, x, b, tol, max_iter, group, use_graph, rows_per_pass=rows_per_pass)


def cg_batched_ir(A64: "K.Matrix", A32: "K.Matrix", row_off, x, b, tol=1e-8, max_iter=100000, inner_tol=1e-4,
                  group=0, use_graph=True, rows_per_pass=0):
    return _run(A64, row_off, x, b, tol, max_iter, group, use_graph, A32=A32, inner_tol=inner_tol,
                rows_per_pass=rows_per_pass)


def _signature(variant, A, row_off):
    return f"{variant}:{len(row_off) - 1}:{int(row_off[-1])}:{A.nnz}"


def autotune(A64, row_off, b, variant="fp64", A32=None, tol=1e-8, reps=2, use_cache=True):
    """Fastest (group, rows_per_pass) for this batch: every pair of the
    candidate work-group sizes and rows kept in flight per thread is timed
    on the batch itself (best of `reps`); for the mixed variant the inner
    tolerance is then timed with that pair (entry["inner_tol"]).  The choice
    is cached by the batch's signature."""
    key = _signature(variant, A64, row_off)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            cache = json.load(open(CACHE))
        except ValueError:
            cache = {}
    if use_cache and key in cache:
        e = cache[key]
        return (e["group"], e["rows_per_pass"]), e
    times = {}
    x = torch.zeros_like(b)
    for g in CANDIDATES:
        for u in ROWS_PER_PASS:
            best = None
            for _ in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                _run(A64, row_off, x, b, tol, 100000, g, True, A32=A32 if variant == "ir" else None,
                     rows_per_pass=u)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            times[(g, u)] = best
    choice = min(times, key=times.get)
    entry = {"group": choice[0], "rows_per_pass": choice[1],
             "seconds": {f"{g}x{u}": v for (g, u), v in times.items()}}
    if variant == "ir":
        itimes = {}
        for itol in INNER_TOLS:
            best = None
            for _ in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                _run(A64, row_off, x, b, tol, 100000, choice[0], True, A32=A32, inner_tol=itol,
                     rows_per_pass=choice[1])
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            itimes[itol] = best
        entry["inner_tol"] = min(itimes, key=itimes.get)
        entry["inner_tol_seconds"] = {f"{t:g}": v for t, v in itimes.items()}
    if use_cache:
        cache[key] = entry
        try:
            json.dump(cache, open(CACHE, "w"), indent=1)
        except OSError:
            pass
    return choice, entry
