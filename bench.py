#!/usr/bin/env python3
"""Benchmark: Q2 hexahedral Laplace assembly on a perturbed 64^3 mesh.

Workload: unit cube, 64^3 Q2 hex cells (262,144 cells, 2,146,689 DOFs,
27 DOFs and 27 Gauss points per cell); interior vertices displaced by a
seeded uniform +-0.2h per component; kappa = 1.  One step = one numeric
assembly of the stiffness values (CSC nzval, FP64 cross-cell accumulation)
over the whole mesh, with mesh, plan and pattern already resident in HBM.

Both arithmetic forms are timed: the FP64 drop-in and the mixed form
(FP64 geometry and accumulation, split-FP16 per-cell contraction).  `value`
is the mixed form; the FP64 form and the mixed-vs-FP64 error are reported
beside it.  The CPU baseline is the oracle (a C restatement of the
reference CPU loop), median of five runs, single thread on a bounded sample
and on all host threads over the whole workload.

The default run also measures the other workloads and reports them as
sub-results of the same line: `q1hex` (Q1 hex 128^3, lognormal kappa),
`p2tet` (P2 tet elasticity 48^3 x 6 with the mixed-precision refined solve)
and `batch` (256 uneven 2D Poisson problems in one graph launch).
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


def measured_traffic(name, **match):
    """HBM bytes per launch from the committed counter summary of a workload
    (rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by a plain copy), or None when
    no summary matches it."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{name}.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if any(d.get(k) != v for k, v in match.items()):
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


def host_threads():
    """Host threads for the threaded CPU baseline: OMP_NUM_THREADS when set,
    else the CPU count, at most 16."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)
    return max(1, min(16, n))


def cpu_baseline(n_sample, n_full, seed, runs=5):
    """The oracle on the Q2 hex workload: single thread on an n_sample^3
    sample (the reference's path is single-threaded; the whole 64^3 mesh
    takes about 10 s per run there), and on all host threads over the whole
    n_full^3 workload; median of `runs` runs each, with the phase split."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker, timed here only as the CPU baseline
    GT = sys.modules["gtk_amd"]

    def timed(n, threads):
        m = GT.cartesian_mesh((0.0, 1.0) * 3, (n,) * 3)
        x = GT.perturb_interior_nodes(m, 0.2, seed)
        p = O.Problem([n] * 3).space(2)
        p.set_coords(x)
        p.quadrature(4)
        O.set_threads(threads)
        rec = []
        for _ in range(runs):
            t0 = time.perf_counter()
            p.assemble_matrix(O.LAPLACE)
            rec.append((time.perf_counter() - t0, O.phase_ms()))
        O.set_threads(1)
        dt, ph = sorted(rec, key=lambda r: r[0])[len(rec) // 2]
        return {"cells_per_s": n ** 3 / dt, "s_median": dt, "s_all": [round(r[0], 4) for r in rec],
                "phase_ms": {k: round(v, 2) for k, v in ph.items()}, "cells": n ** 3, "threads": threads}

    one = timed(n_sample, 1)
    nt = host_threads()
    full = timed(n_full, nt)
    return {"value": one["cells_per_s"], "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": f"{n_sample}^3 Q2 hex perturbed mesh ({n_sample ** 3} cells), FP64 numeric assembly + COO->CSC "
                      f"in the reference loop order, single thread, median of {runs} runs "
                      f"({one['s_median']:.2f} s each; a 64^3 run takes about 10 s single-threaded)",
            "single_thread": one,
            "threaded": {**full, "note": f"the whole {n_full}^3 workload on {nt} host threads (cell loop and "
                                         f"per-column compress threaded; same COO order, same result)"}}


def alg_bytes(D, n_geo, cells, nv, nd, ncoef, nnz):
    """Algorithmic HBM bytes of one matrix launch: coordinates, cell->node,
    cell->DOF, per-cell coefficients, a 4-byte local-entry->nnz map and the
    FP64 values, each moved once."""
    return 8 * D * n_geo + 4 * nv * cells + 4 * nd * cells + 8 * ncoef * cells + 4 * nd * nd * cells + 8 * nnz


def time_form(ctx, L, form, prec, steps, warmup):
    """Wall time per launch over `steps` queued launches, and the kernel's own
    time from HIP events on the context stream."""
    for _ in range(warmup):
        ctx.assemble_matrix_async(form, prec, L.GTK_FREE, L.GTK_FREE)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.assemble_matrix_async(form, prec, L.GTK_FREE, L.GTK_FREE)
    ctx.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    return wall_ms, ctx.time_matrix(form, prec, L.GTK_FREE, L.GTK_FREE, steps)


def tuned_launch(GT, args, ctx, form, V, code):
    """Launch shape of both precisions: the autotuner's choice (tuned on the
    problem unless cached), or with --no-tune the library's defaults."""
    from gtk_amd import _lib as L
    if args.no_tune:
        return {name: {"tuned": False, "plan": ctx.plan_info(code, prec, L.GTK_FREE, L.GTK_FREE)}
                for name, prec in (("fp64", L.GTK_FP64), ("mixed", L.GTK_MIXED))}
    return {name: GT.autotune_matrix(form, V, precision=prec, cache=args.tune_cache or None)
            for name, prec in (("fp64", GT.FP64), ("mixed", GT.MIXED))}


def bench_laplace(GT, args, n, order, degree, perturb, lognormal_kappa, label):
    """Laplace assembly on an n^3 hex mesh: FP64 and mixed forms."""
    from gtk_amd import _lib as L
    t_setup = time.perf_counter()
    mesh = GT.cartesian_mesh((0.0, 1.0) * 3, (n, n, n))
    if perturb:
        mesh.coords = GT.perturb_interior_nodes(mesh, perturb, args.seed)
    Om = GT.interior(mesh)
    V = GT.lagrange_space(Om, order)
    dO = GT.measure(Om, degree)
    cells = mesh.cells.shape[0]
    kappa = GT.inputs.lognormal(args.seed, cells) if lognormal_kappa else None
    form = GT.Laplace(dO, GT.face_constant_field(kappa, Om) if lognormal_kappa else None)
    A64, cache = GT.assemble_matrix(form, float, V, V, precision=GT.FP64, reuse=True)
    ref = A64.nzval.copy()
    Amx, cache_mx = GT.assemble_matrix(form, float, V, V, precision=GT.MIXED, reuse=True)
    rel_err = float(np.linalg.norm(Amx.nzval - ref) / np.linalg.norm(ref))
    setup_s = time.perf_counter() - t_setup
    ctx = A64.device[0].ctx
    launch = tuned_launch(GT, args, ctx, form, V, L.GTK_LAPLACE)
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib << 20)
    wall64 = kern64 = wallmx = kernmx = float("nan")
    if args.only in ("both", "fp64"):
        wall64, kern64 = time_form(ctx, L, L.GTK_LAPLACE, L.GTK_FP64, args.steps, args.warmup)
    if args.only in ("both", "mixed"):
        wallmx, kernmx = time_form(ctx, L, L.GTK_LAPLACE, L.GTK_MIXED, args.steps, args.warmup)
    # symbolic phase of the same block on host and device (fresh contexts)
    from gtk_amd import assembly as AS
    sym = {}
    for where in ("host", "device"):
        a_sym = AS._Assembler(V, dO)
        t0 = time.perf_counter()
        if where == "host":
            a_sym.ctx.build_pattern(L.GTK_FREE, L.GTK_FREE)
        else:
            sym["device_kernels_ms"] = a_sym.ctx.build_pattern_device(L.GTK_FREE, L.GTK_FREE)[2]
        sym[f"{where}_ms"] = (time.perf_counter() - t0) * 1e3
        a_sym.ctx.close()
    # the same mixed assembly through the host boundary: launch + the values'
    # copy to host memory (PCIe), wall time of gtk_assemble_matrix
    pcie_ms = min(GT.update_matrix(Amx, cache_mx).stats.total_ms for _ in range(3))
    nd = V.fe.nnodes
    nq = len(dO.w)
    traffic = measured_traffic(f"q{order}hex{n}", cells_per_dir=n, order=order, degree=degree, perturb=perturb,
                               lognormal_kappa=lognormal_kappa)
    bytes_alg = alg_bytes(3, mesh.coords.shape[0], cells, 8, nd, 1 if lognormal_kappa else 0, A64.nnz)
    flops_alg = cells * nq * (2 * 3 * nd * nd + 2 * 9 * nd + 2 * 8 * 9)
    achieved = bytes_alg / (kernmx * 1e-3) / 1e9
    kdesc = "lognormal kappa per cell" if lognormal_kappa else "kappa=1"
    pdesc = f"interior vertices perturbed U(-{perturb}h,{perturb}h) (seed {args.seed}), " if perturb else ""
    return {
        "metric": f"cells/s + HBM GB/s, Q{order} hex Laplace assembly; mixed vs FP64 at stated rel err",
        "value": cells / (wallmx * 1e-3),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wallmx,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp16x2-split contraction (hi+lo parts, fp32 accumulate) on fp16 MFMA"
                  if order == 2 else "fp32 contraction") + ", fp64 geometry + accumulation",
        "data": "synthetic",
        "config": {"workload": f"3D Laplace, {n}^3 Q{order} hex cells on the unit cube, {pdesc}"
                               f"{nq}-point Gauss, {kdesc}",
                   "name": label, "cells": cells, "dofs": int(V.nfree), "nnz": int(A64.nnz),
                   "parallelism": "single GPU", "launch": launch},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (traffic or {}).get("mixed")},
        "fp64": {"cells_per_s": cells / (wall64 * 1e-3), "ms_per_step": wall64, "kernel_ms": kern64,
                 "traffic": (traffic or {}).get("fp64"),
                 "hbm_gbs_alg": bytes_alg / (kern64 * 1e-3) / 1e9,
                 "roofline_frac": bytes_alg / (kern64 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                 "fp64_tflops_alg": flops_alg / (kern64 * 1e-3) / 1e12},
        "mixed": {"kernel_ms": kernmx, "rel_fro_err_vs_fp64": rel_err,
                  "speedup_vs_fp64": kern64 / kernmx},
        "dofs_per_s": V.nfree / (wallmx * 1e-3),
        "symbolic": {**sym, "note": "pattern of the block: host C++ (threads) vs GPU kernels; device_ms includes "
                                   "the copy of colptr/rowval to host memory"},
        "host_values": {"ms": pcie_ms, "cells_per_s": cells / (pcie_ms * 1e-3),
                        "note": "mixed assembly with the values copied to host memory (not the value)"},
        "bytes_alg_per_launch": bytes_alg,
        "setup_s": setup_s,
        "refined_solve": (refined_solve(GT, mesh, order, degree, kappa)
                          if label == "q2hex" and not args.no_solve else None),
    }


def refined_solve(GT, mesh, order, degree, kappa):
    """The refined-solution check of the mixed-precision path on the same
    mesh with Dirichlet values on the whole boundary and f = 1: the FP64
    reference (FP64 assembly, FP64 Jacobi-PCG to 1e-11); the mixed
    refinement of the FP64 assembly (FP32 Jacobi-PCG on its FP32 copy inside
    FP64 residuals with the FP64 values, to 1e-10); the mixed refinement of
    the mixed assembly (its residuals use the mixed values)."""
    Om = GT.interior(mesh)
    V = GT.lagrange_space(Om, order, dirichlet_boundary=GT.boundary(mesh))
    dO = GT.measure(Om, degree)
    form = GT.Laplace(dO, GT.face_constant_field(kappa, Om) if kappa is not None else None)
    b = GT.assemble_vector(GT.Load(dO, 1.0), float, V)
    A64 = GT.assemble_matrix(form, float, V, V, precision=GT.FP64)
    x64, st64 = GT.solve(A64, b, precision=GT.FP64, rtol=1e-11)
    xr, str_ = GT.solve(A64, b, precision=GT.MIXED, rtol=1e-10)
    Amx = GT.assemble_matrix(form, float, V, V, precision=GT.MIXED)
    xm, stm = GT.solve(Amx, b, precision=GT.MIXED, rtol=1e-10)
    nx = np.linalg.norm(x64)

    def rep(x, st):
        return {"rel_err_vs_fp64_solve": float(np.linalg.norm(x - x64) / nx), "rel_residual": st.rel_residual,
                "outer_iterations": st.outer_iterations, "inner_iterations": st.inner_iterations, "ms": st.solve_ms}
    return {"dofs": int(V.nfree), "nnz": int(A64.nnz),
            "fp64_reference": {"rel_residual": st64.rel_residual, "inner_iterations": st64.inner_iterations,
                               "ms": st64.solve_ms},
            "mixed_refinement_on_fp64_assembly": rep(xr, str_),
            "mixed_refinement_on_mixed_assembly": rep(xm, stm),
            "note": "Dirichlet on the whole boundary, f = 1; the error of the second against the FP64 solve is the "
                    "assembly's FP32 contraction rounding amplified by the condition number"}


def bench_elasticity(GT, args, n):
    """P2-tet linear elasticity on n^3 cubes x 6: FP32 assembly, body force,
    clamped on x = 0, solved by FP32 Jacobi-PCG inside FP64 refinement to
    relative residual 1e-10."""
    from gtk_amd import _lib as L
    t_setup = time.perf_counter()
    mesh = GT.cartesian_mesh((0.0, 1.0) * 3, (n, n, n), simplexify=True)
    Om = GT.interior(mesh)
    V = GT.lagrange_space(Om, 2, tensor_size=(3,), dirichlet_boundary=GT.boundary(mesh, group_names=["2-face-5"]))
    dO = GT.measure(Om, 5)
    cells = mesh.cells.shape[0]
    lam = 0.5 + 1.5 * GT.inputs.uniform(args.seed, cells)
    mu = 0.5 + 1.5 * GT.inputs.uniform(args.seed + 1, cells)
    form = GT.Elasticity(dO, GT.face_constant_field(lam, Om), GT.face_constant_field(mu, Om))
    A64 = GT.assemble_matrix(form, float, V, V, precision=GT.FP64)
    ref = A64.nzval.copy()
    A = GT.assemble_matrix(form, float, V, V, precision=GT.MIXED)
    rel_err = float(np.linalg.norm(A.nzval - ref) / np.linalg.norm(ref))
    b = GT.assemble_vector(GT.Load(dO, (0.0, 0.0, -1.0)), float, V)
    setup_s = time.perf_counter() - t_setup
    ctx = A.device[0].ctx
    launch = tuned_launch(GT, args, ctx, form, V, L.GTK_ELASTICITY)
    wall64, kern64 = time_form(ctx, L, L.GTK_ELASTICITY, L.GTK_FP64, args.steps, args.warmup)
    wallmx, kernmx = time_form(ctx, L, L.GTK_ELASTICITY, L.GTK_MIXED, args.steps, args.warmup)
    if args.no_solve:
        return elasticity_result(args, n, cells, V, A, launch, wall64, kern64, wallmx, kernmx, rel_err, mesh,
                                 None, None, setup_s)
    # the FP32 assembly is the one solved (mixed refinement to 1e-10); the
    # FP64 solve of the FP64-assembled system to 1e-11 is the reference
    A = GT.assemble_matrix(form, float, V, V, precision=GT.MIXED)
    x, st = GT.solve(A, b, precision=GT.MIXED, rtol=1e-10)
    x64, st64 = GT.solve(A64, b, precision=GT.FP64, rtol=1e-11)
    # the same refinement on the FP64-assembled system: its error against
    # that system's FP64 solve is the solver's own
    xr, str_ = GT.solve(A64, b, precision=GT.MIXED, rtol=1e-10)
    r64 = b - A64.to_scipy() @ x
    sol = {"rel_err_vs_fp64_solve": float(np.linalg.norm(x - x64) / np.linalg.norm(x64)),
           "rel_residual_fp64_system": float(np.linalg.norm(r64) / np.linalg.norm(b)),
           "fp64_reference": {"rel_residual": st64.rel_residual, "inner_iterations": st64.inner_iterations,
                              "ms": st64.solve_ms},
           "mixed_refinement_on_fp64_assembly": {
               "rel_err_vs_fp64_solve": float(np.linalg.norm(xr - x64) / np.linalg.norm(x64)),
               "rel_residual": str_.rel_residual, "outer_iterations": str_.outer_iterations,
               "inner_iterations": str_.inner_iterations, "ms": str_.solve_ms},
           "note": "x: mixed assembly, FP32 Jacobi-PCG inside FP64 refinement to |b - A x| <= 1e-10 |b|; "
                   "reference: FP64 assembly, FP64 Jacobi-PCG to 1e-11; the difference includes the assembly's "
                   "FP32 rounding (rel_fro_err_vs_fp64 of the matrix) amplified by the condition number"}
    return elasticity_result(args, n, cells, V, A, launch, wall64, kern64, wallmx, kernmx, rel_err, mesh, st, sol,
                             setup_s)


def elasticity_result(args, n, cells, V, A, launch, wall64, kern64, wallmx, kernmx, rel_err, mesh, st, sol,
                      setup_s):
    """The p2tet line; st/sol None: assembly only (--no-solve)."""
    nd = 30
    bytes_alg = alg_bytes(3, mesh.coords.shape[0], cells, 4, nd, 2, A.nnz)
    achieved = bytes_alg / (kernmx * 1e-3) / 1e9
    return {
        "metric": "cells/s, P2 tet elasticity FP32 assembly + mixed-precision refined solve",
        "value": cells / (wallmx * 1e-3),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wallmx,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 contraction, fp64 geometry + accumulation; fp32 CG in fp64 refinement",
        "data": "synthetic",
        "config": {"workload": f"3D linear elasticity, {n}^3 cubes x 6 P2 tets, 14-point Strang, lambda and mu "
                               f"U[0.5,2] per cell (seed {args.seed}), clamped on x=0, body force (0,0,-1)",
                   "name": "p2tet", "cells": cells, "dofs": int(V.nfree), "nnz": int(A.nnz),
                   "parallelism": "single GPU", "launch": launch},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "note": "algorithmic bytes of the mixed assembly over its kernel time; the kernel is bound by the "
                             "latency of its per-contributor loads, not by HBM bandwidth"},
        "fp64": {"ms_per_step": wall64, "kernel_ms": kern64, "cells_per_s": cells / (wall64 * 1e-3)},
        "mixed": {"kernel_ms": kernmx, "rel_fro_err_vs_fp64": rel_err, "speedup_vs_fp64": kern64 / kernmx},
        "bytes_alg_per_launch": bytes_alg,
        "solve": None if st is None else {"rel_residual": st.rel_residual, "outer_iterations": st.outer_iterations,
                                          "inner_iterations": st.inner_iterations, "ms": st.solve_ms, **sol},
        "setup_s": setup_s,
    }


def bench_dg(GT, args, n):
    """Interior-penalty DG (GalerkinToolkitExamples/src/poisson.jl:65-107):
    Laplace + interior-penalty skeleton terms + Nitsche terms on the whole
    boundary, a discontinuous Q2 space on n^2 perturbed quadrilaterals,
    degree 4, gamma = degree (degree + 1) / 10; each part's kernel timed in
    both arithmetic forms, and the summed matrix's wall time."""
    from gtk_amd import _lib as L
    from gtk_amd import assembly as AS
    t_setup = time.perf_counter()
    mesh = GT.cartesian_mesh((0.0, 1.0) * 2, (n, n))
    mesh.coords = GT.perturb_interior_nodes(mesh, 0.2, args.seed)
    Om = GT.interior(mesh)
    V = GT.lagrange_space(Om, 2, continuous=False)
    degree = 4
    gamma = degree * (degree + 1) / 10.0
    dO, dL, dG = GT.measure(Om, degree), GT.measure(GT.skeleton(mesh), degree), GT.measure(GT.boundary(mesh), degree)
    form = GT.Laplace(dO) + GT.InteriorPenalty(dL, gamma) + GT.Nitsche(dG, gamma)
    A64 = GT.assemble_matrix(form, float, V, V, precision=GT.FP64)
    Amx = GT.assemble_matrix(form, float, V, V, precision=GT.MIXED)
    rel_err = float(np.linalg.norm(Amx.nzval - A64.nzval) / np.linalg.norm(A64.nzval))
    setup_s = time.perf_counter() - t_setup
    cells = mesh.cells.shape[0]
    parts = {}
    for name, f, code, kind in (("volume", form.terms[0], L.GTK_LAPLACE, None),
                                ("skeleton", form.terms[1], L.GTK_SKELETON, None),
                                ("nitsche", form.terms[2], L.GTK_NITSCHE, "nitsche")):
        a = AS._assembler(V, f.dOmega, 0, kind)
        _, k64 = time_form(a.ctx, L, code, L.GTK_FP64, args.steps, args.warmup)
        _, kmx = time_form(a.ctx, L, code, L.GTK_MIXED, args.steps, args.warmup)
        nnz = a.pattern(L.GTK_FREE, L.GTK_FREE)[0]
        items = getattr(a, "items", cells)
        per = 2 if name == "skeleton" else 1
        nd = V.fe.nnodes * per if name != "volume" else V.fe.nnodes
        nv = 4 * per
        b = alg_bytes(2, mesh.coords.shape[0], items, nv, nd, 0 if name == "volume" else 1, nnz)
        parts[name] = {"items": int(items), "nnz": int(nnz), "fp64_kernel_ms": k64, "mixed_kernel_ms": kmx,
                       "bytes_alg_per_launch": b, "roofline_frac_mixed": b / (kmx * 1e-3) / 1e9 / HBM_PEAK_GBS}
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        GT.assemble_matrix(form, float, V, V, precision=GT.MIXED)
        walls.append((time.perf_counter() - t0) * 1e3)
    wall = float(np.median(walls))
    return {
        "metric": "cells/s, interior-penalty DG Laplace + skeleton + Nitsche assembly (Q2 quads)",
        "value": cells / (wall * 1e-3),
        "unit": "cells/s",
        "ms_per_step": wall,
        "dtype": "fp32 contraction, fp64 geometry + accumulation (mixed); fp64",
        "config": {"workload": f"2D interior-penalty DG, {n}^2 Q2 quadrilaterals (discontinuous), interior "
                               f"vertices perturbed U(-0.2h,0.2h) (seed {args.seed}), degree {degree}, gamma {gamma}",
                   "name": "dgip", "cells": cells, "dofs": int(V.nfree), "nnz": int(A64.nnz)},
        "parts": parts,
        "mixed": {"rel_fro_err_vs_fp64": rel_err},
        "note": "value: wall time of the mixed summed assembly (three kernels, two block additions on the device, "
                "the values copied to host memory)",
        "setup_s": setup_s,
    }


BATCH_SHAPES = {2: (16, 256, 3), 3: (8, 40, 2)}  # n per side in [nmin, nmax], orders 1..kmax


def batch_problems(GT, count, seed, dim=2):
    """`count` unit-square (dim 2) or unit-cube (dim 3) Poisson problems: n
    per side uniform in [16, 256] (3D: [8, 40]), Lagrange order uniform in
    {1, 2, 3} (3D: {1, 2}), Gauss with order+1 points per direction,
    Dirichlet on the whole boundary, lognormal kappa per cell and Gaussian
    nodal f (seeded per problem)."""
    nmin, nmax, kmax = BATCH_SHAPES[dim]
    u = GT.inputs.uniform(seed, 2 * count)
    probs = []
    for i in range(count):
        n = nmin + int(u[2 * i] * (nmax - nmin + 1))
        order = 1 + int(u[2 * i + 1] * kmax)
        m = GT.cartesian_mesh((0.0, 1.0) * dim, (n,) * dim)
        Om = GT.interior(m)
        V = GT.lagrange_space(Om, order, dirichlet_boundary=GT.boundary(m))
        dO = GT.measure(Om, 2 * order)
        kappa = GT.inputs.lognormal(seed + 1 + i, m.cells.shape[0])
        f = GT.inputs.normal(seed + 100_000 + i, V.nfree + V.ndiri)
        probs.append(dict(n=n, order=order, m=m, Om=Om, V=V, dO=dO, kappa=kappa, f=f))
    return probs


def cpu_baseline_batch(probs, stride):
    """The oracle (single thread) on every `stride`-th problem: matrix and
    load vector, cells/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker, timed here only as the CPU baseline
    sample = probs[::stride]
    cells, dt = 0, 0.0
    for q in sample:
        p = O.Problem((q["n"],) * q["m"].D).space(q["order"], dirichlet=1)
        p.quadrature(2 * q["order"])
        t0 = time.perf_counter()
        p.assemble_matrix(O.LAPLACE, coef=q["kappa"])
        p.assemble_vector(2, q["f"])
        dt += time.perf_counter() - t0
        cells += q["m"].cells.shape[0]
    return {"value": cells / dt, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": f"every {stride}th problem of the batch ({len(sample)} problems, {cells} cells): FP64 matrix "
                      f"+ nodal load vector + COO->CSC in the reference loop order, single thread, {dt:.2f} s"}


def bench_batch(GT, args, count, dim=2):
    """Batched uneven 2D (or 3D) Poisson: every problem's matrix and load
    vector in one persistent launch replayed from a HIP graph."""
    t_setup = time.perf_counter()
    probs = batch_problems(GT, count, args.seed, dim)
    forms = [GT.Laplace(q["dO"], GT.face_constant_field(q["kappa"], q["Om"])) for q in probs]
    loads = [GT.Load(q["dO"], GT.DiscreteField(q["V"], q["f"][:q["V"].nfree], q["f"][q["V"].nfree:]))
             for q in probs]
    cache = args.tune_cache or None
    bm = GT.MatrixBatch(forms, [q["V"] for q in probs], precision=GT.MIXED, tuning="auto", cache=cache, loads=loads)
    b64 = bm.with_precision(GT.FP64)
    setup_s = time.perf_counter() - t_setup
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib << 20)
    cells = bm.cells
    nnz = 0
    nbytes = 0
    for q, (a, _, nc) in zip(probs, bm.problems):
        pnnz = a.pattern(GT.FREE, GT.FREE)[0]
        V = q["V"]
        nd = V.cell_dofs.shape[1]
        nnz += pnnz
        nbytes += alg_bytes(dim, q["m"].coords.shape[0], nc, 1 << dim, nd, 1, pnnz)
        nbytes += 8 * (V.nfree + V.ndiri) + 8 * V.nfree  # nodal f in, load vector out

    def timed(b):
        for _ in range(args.warmup):
            b.run()
        kern = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            kern.append(b.run().kernel_ms)
        return (time.perf_counter() - t0) * 1e3 / args.steps, float(np.mean(kern))

    wall64, kern64 = timed(b64)
    ref = [A.nzval.copy() for A in b64.matrices()]
    vref = b64.vectors()
    wallmx, kernmx = timed(bm)
    errs = np.array([np.linalg.norm(A.nzval - r) / np.linalg.norm(r) for A, r in zip(bm.matrices(), ref)])
    verrs = np.array([np.linalg.norm(v - r) / np.linalg.norm(r) for v, r in zip(bm.vectors(), vref)])
    orders = [q["order"] for q in probs]
    achieved = nbytes / (kernmx * 1e-3) / 1e9
    traffic = measured_traffic("batch256" if dim == 2 else "batch3d64", problems=count, seed=args.seed)
    nmin, nmax, kmax = BATCH_SHAPES[dim]
    shape = "unit-square" if dim == 2 else "unit-cube"
    out = {
        "metric": f"aggregate cells/s, batched uneven {dim}D Poisson (matrix + load vector per problem), one graph "
                  f"launch",
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
        "config": {"workload": f"{count} {shape} problems, n per side U[{nmin},{nmax}], Lagrange order "
                               f"U{{1..{kmax}}}, (order+1)^{dim} Gauss, lognormal kappa per cell, Gaussian nodal f, "
                               f"Dirichlet boundary (seed {args.seed})",
                   "name": "batch" if dim == 2 else "batch3d", "problems": count, "cells": int(cells),
                   "nnz": int(nnz), "orders": {str(o): orders.count(o) for o in range(1, kmax + 1)},
                   "parallelism": "single GPU",
                   "launch": {name: {"waves": b.waves,
                                     "tiles": {GT.batch.ELEM_NAMES[k]: v for k, v in b.tiles.items()}}
                              for name, b in (("mixed", bm), ("fp64", b64))}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic and traffic["mixed"],
                     "alg_bytes": int(nbytes)},
        "fp64": {"ms_per_step": wall64, "kernel_ms": kern64, "cells_per_s": cells / (wall64 * 1e-3),
                 "traffic": traffic and traffic["fp64"]},
        "mixed": {"kernel_ms": kernmx, "speedup_vs_fp64": kern64 / kernmx,
                  "rel_fro_err_vs_fp64": {"max": float(errs.max()), "median": float(np.median(errs))},
                  "load_vector_rel_err_vs_fp64_max": float(verrs.max())},
        "tuning": {"mixed": bm.tuning_report, "fp64": b64.tuning_report},
        "setup_s": setup_s,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(probs, args.batch_cpu_stride if dim == 2 else 8)
    return out


SUB_DROP = ("n_gpus", "steps", "warmup", "higher_is_better", "scaling", "vs_baseline", "data")


def sub_result(r):
    """A workload's result as a sub-result of the headline line (the fields
    the headline already states dropped)."""
    return {k: v for k, v in r.items() if k not in SUB_DROP}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["all", "q2hex", "q1hex", "p2tet", "batch", "batch3d", "dgip"],
                    default="all",
                    help="all (default): the Q2 hex headline with the others as sub-results of the same line; "
                         "q2hex: Q2 hex 64^3 perturbed Laplace (the headline) alone; q1hex: Q1 hex 128^3 with "
                         "lognormal kappa; p2tet: P2 tet elasticity 48^3 x 6 with the refined solve; batch: 256 "
                         "uneven 2D Poisson problems in one graph launch; batch3d: 64 uneven 3D ones (Q1/Q2 hexahedra); "
                         "dgip: interior-penalty DG on 512^2 "
                         "discontinuous Q2 quadrilaterals (volume, skeleton and Nitsche kernels)")
    ap.add_argument("--problems", type=int, default=256, help="problems of the batch workload")
    ap.add_argument("--batch-cpu-stride", type=int, default=2, help="CPU baseline: every k-th problem of the batch")
    ap.add_argument("--tune-cache", default="", help="autotuner cache file (default: the user cache)")
    ap.add_argument("--no-tune", action="store_true",
                    help="single problems keep the library's default launch shapes (no tuning launches, for "
                         "profiling runs)")
    ap.add_argument("--cells", type=int, default=0, help="cells per direction (default per workload)")
    ap.add_argument("--cpu-sample", type=int, default=48, help="cells per direction of the single-thread CPU sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true",
                    help="assembly only: no refined solves (q2hex, p2tet; profiling runs)")
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

    def step(name):
        print(f"bench: {name}", file=sys.stderr, flush=True)

    wl = args.workload
    if wl in ("all", "q2hex"):
        step("q2hex")
        out = bench_laplace(GT, args, args.cells or 64, 2, 4, 0.2, False, "q2hex")
        if not args.no_cpu_baseline:
            step("q2hex cpu baseline")
            out["cpu_baseline"] = cpu_baseline(args.cpu_sample, args.cells or 64, args.seed)
        if wl == "all":
            step("q1hex")
            out["q1hex"] = sub_result(bench_laplace(GT, args, 128, 1, 2, 0.0, True, "q1hex"))
            step("p2tet")
            out["p2tet"] = sub_result(bench_elasticity(GT, args, 48))
            step("batch")
            out["batch"] = sub_result(bench_batch(GT, args, args.problems))
            step("dgip")
            out["dgip"] = sub_result(bench_dg(GT, args, 512))
    elif wl == "q1hex":
        out = bench_laplace(GT, args, args.cells or 128, 1, 2, 0.0, True, "q1hex")
    elif wl == "p2tet":
        out = bench_elasticity(GT, args, args.cells or 48)
    elif wl == "dgip":
        out = bench_dg(GT, args, args.cells or 512)
    elif wl == "batch3d":
        out = bench_batch(GT, args, 64, dim=3)
    else:
        out = bench_batch(GT, args, args.problems)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
