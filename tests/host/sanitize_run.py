"""Run by tests/test_host_sanitizers.py in a process with the address and
undefined-behaviour sanitizer runtimes preloaded: the oracle (built with
them, GTK_ORACLE_LIB) assembles matrices and vectors of every element kind,
and the symbolic phase (gtk_plan.cpp, built with them into the plan
harness given as argv[1]) builds and replays patterns, tiles, row offsets
and batch records against a COO assembly.  Prints one line when all pass."""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import conftest  # noqa: E402
import oracle as O  # noqa: E402
from test_plan_host import run  # noqa: E402

GT = conftest.load_package()


def harness(path):
    L = C.CDLL(path)
    f64 = np.ctypeslib.ndpointer(np.float64, flags="C")
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C")
    L.harness_run.argtypes = [C.c_int, C.c_int64, f64, C.c_int64, C.c_int, i32, C.c_int, C.c_int64, C.c_int64, i32,
                              C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64), i32, i32, f64, C.c_int64,
                              C.POINTER(C.c_int), C.c_int, C.c_char_p, C.c_int]
    return L


def oracle_cases():
    for cells, order, simplex, ncomp in [((4, 3), 1, False, 1), ((3, 4), 2, False, 1), ((3, 3), 3, False, 1),
                                         ((2, 3, 2), 1, False, 1), ((2, 2, 2), 2, False, 1),
                                         ((2, 2, 2), 2, True, 1), ((2, 2, 2), 2, True, 3)]:
        D = len(cells)
        p = O.Problem(list(cells), simplexify=simplex).space(order, dirichlet=1, ncomp=ncomp)
        p.quadrature(2 * order if not simplex else 4)
        if ncomp == 1:
            cp, rv, nz = p.assemble_matrix(O.LAPLACE)
            assert np.all(np.diff(cp) >= 0) and np.isfinite(nz).all()
            b = p.assemble_vector(0, [1.0])
            assert np.isfinite(b).all()
        else:
            nc = p.cell_nodes().shape[0]
            coef = np.empty(2 * nc)
            coef[0::2], coef[1::2] = 1.0, 0.5
            cp, rv, nz = p.assemble_matrix(O.ELASTICITY, coef=coef)
            assert np.isfinite(nz).all()


def plan_cases(L):
    for cells, order, simplex, diri, max_cells, chunk in [((6, 5), 1, False, 1, 64, 8), ((5, 4), 2, False, 0, 27, 1),
                                                          ((4, 5), 3, False, 1, 200, 64), ((3, 4, 3), 2, False, 1, 64, 8),
                                                          ((3, 3, 2), 2, True, 0, 64, 8)]:
        for mode in (0, 1):
            run(L, GT, cells, order, simplex, diri, max_cells=max_cells, chunk=chunk, mode=mode)
    run(L, GT, (5, 6), 2, diri=1, rows=1, cols=2, mode=0)


if __name__ == "__main__":
    oracle_cases()
    plan_cases(harness(sys.argv[1]))
    print("sanitized run ok", flush=True)
