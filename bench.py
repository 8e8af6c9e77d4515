#!/usr/bin/env python3
"""Benchmark: Q2 hexahedral Laplace assembly on a perturbed 64^3 mesh.

Workload: unit cube, 64^3 Q2 hex cells (262,144 cells, 2,146,689 DOFs,
27 DOFs and 27 Gauss points per cell); interior vertices displaced by a
seeded uniform +-0.2h per component; kappa = 1.  One step = one numeric
assembly of the stiffness values (CSC nzval, FP64 cross-cell accumulation)
over the whole mesh, with mesh, plan and pattern already resident in HBM.

Both arithmetic forms are timed: the FP64 drop-in and the mixed form
(FP64 geometry and accumulation, FP32 per-cell contraction).  `value` is
the mixed form; the FP64 form and the mixed-vs-FP64 error are reported
beside it.  The CPU baseline is the oracle (a C restatement of the
reference CPU loop) on a bounded sample of the same workload.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def load_package():
    path = os.path.join(ROOT, "galerkintoolkit.jl_amd")
    spec = importlib.util.spec_from_file_location("gtk_amd", os.path.join(path, "__init__.py"),
                                                  submodule_search_locations=[path])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gtk_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def survey_bytes_per_cell_q2(n):
    """Algorithmic HBM bytes of one Q2-hex matrix assembly launch on an n^3
    mesh: coordinates + cell->node + cell->DOF + local-entry->nnz map (4 B
    per local entry) + FP64 values, each moved once."""
    nodes = (n + 1) ** 3
    cells = n ** 3
    nnz = (8 * n + 1) ** 3
    return 8 * 3 * nodes + 4 * 8 * cells + 4 * 27 * cells + 4 * 729 * cells + 8 * nnz


def measured_traffic(n):
    """HBM bytes per launch from the committed counter summary of this
    workload (rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by a plain copy), or
    None when no summary matches."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_q2hex64.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("cells_per_dir") != n:
        return None
    return d["bytes_per_launch"]


def copy_calibration(nbytes):
    """One plain device-to-device copy of nbytes (the HIP runtime's copy
    kernel) so hardware byte counters can be scaled against known traffic."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"copy calibration: {nbytes} bytes read, {nbytes} bytes written", file=sys.stderr)


def cpu_baseline(n_sample, seed):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker, timed here only as the CPU baseline
    GT = sys.modules["gtk_amd"]
    m = GT.cartesian_mesh((0.0, 1.0) * 3, (n_sample,) * 3)
    x = GT.perturb_interior_nodes(m, 0.2, seed)
    p = O.Problem([n_sample] * 3).space(2)
    p.set_coords(x)
    p.quadrature(4)
    t0 = time.perf_counter()
    p.assemble_matrix(O.LAPLACE)
    dt = time.perf_counter() - t0
    return {"value": n_sample ** 3 / dt, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": f"{n_sample}^3 Q2 hex perturbed mesh ({n_sample ** 3} cells), FP64 numeric assembly + "
                      f"COO->CSC in the reference loop order, single thread, {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=64, help="cells per direction")
    ap.add_argument("--cpu-sample", type=int, default=48, help="cells per direction of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one plain device-to-device copy of this many MiB (HIP runtime) so hardware "
                         "byte counters can be scaled against a copy of known size")
    ap.add_argument("--only", choices=["both", "fp64", "mixed"], default="both",
                    help="time only one arithmetic form (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; a multi-GPU run is not part of it (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    GT = load_package()
    n = args.cells
    t_setup = time.perf_counter()
    mesh = GT.cartesian_mesh((0.0, 1.0) * 3, (n, n, n))
    mesh.coords = GT.perturb_interior_nodes(mesh, 0.2, args.seed)
    Om = GT.interior(mesh)
    V = GT.lagrange_space(Om, 2)
    dO = GT.measure(Om, 4)
    form = GT.Laplace(dO)
    A64, cache = GT.assemble_matrix(form, float, V, V, precision=GT.FP64, reuse=True)
    ref = A64.nzval.copy()
    Amx = GT.assemble_matrix(form, float, V, V, precision=GT.MIXED)
    rel_err = float(np.linalg.norm(Amx.nzval - ref) / np.linalg.norm(ref))
    setup_s = time.perf_counter() - t_setup
    asm = cache[0]
    ctx = asm.ctx
    from gtk_amd import _lib as L

    def run(prec):
        for _ in range(args.warmup):
            ctx.assemble_matrix_async(L.GTK_LAPLACE, prec, L.GTK_FREE, L.GTK_FREE)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.assemble_matrix_async(L.GTK_LAPLACE, prec, L.GTK_FREE, L.GTK_FREE)
        ctx.sync()
        wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        kern_ms = ctx.time_matrix(L.GTK_LAPLACE, prec, L.GTK_FREE, L.GTK_FREE, args.steps)
        return wall_ms, kern_ms

    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib << 20)
    wall64 = kern64 = wallmx = kernmx = float("nan")
    if args.only in ("both", "fp64"):
        wall64, kern64 = run(L.GTK_FP64)
    if args.only in ("both", "mixed"):
        wallmx, kernmx = run(L.GTK_MIXED)
    cells = mesh.cells.shape[0]
    traffic = measured_traffic(n)
    bytes_alg = survey_bytes_per_cell_q2(n)
    flops_alg = cells * 27 * (2 * 3 * 27 * 27 + 2 * 9 * 27 + 2 * 8 * 9)
    achieved = bytes_alg / (kernmx * 1e-3) / 1e9
    out = {
        "metric": "cells/s + HBM GB/s, Q2 hex Laplace assembly; mixed vs FP64 at stated rel err",
        "value": cells / (wallmx * 1e-3),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wallmx,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 contraction, fp64 geometry + accumulation",
        "data": "synthetic",
        "config": {"workload": f"3D Laplace, {n}^3 Q2 hex cells on the unit cube, interior vertices perturbed "
                               f"U(-0.2h,0.2h) (seed {args.seed}), 27-point Gauss, kappa=1",
                   "cells": cells, "dofs": int(V.nfree), "nnz": int(A64.nnz), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (traffic or {}).get("mixed")},
        "fp64": {"cells_per_s": cells / (wall64 * 1e-3), "ms_per_step": wall64, "kernel_ms": kern64,
                 "traffic": (traffic or {}).get("fp64"),
                 "hbm_gbs_alg": bytes_alg / (kern64 * 1e-3) / 1e9,
                 "fp64_tflops_alg": flops_alg / (kern64 * 1e-3) / 1e12},
        "mixed": {"kernel_ms": kernmx, "rel_fro_err_vs_fp64": rel_err,
                  "speedup_vs_fp64": kern64 / kernmx},
        "dofs_per_s": V.nfree / (wallmx * 1e-3),
        "bytes_alg_per_launch": bytes_alg,
        "setup_s": setup_s,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample, args.seed)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
