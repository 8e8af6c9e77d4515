"""Host logic of the symbolic phase: CSC pattern, column-owner tiles, row
offsets (VALU kernel) and batch records (matrix-core kernel), replayed on
the CPU with exact integer-valued entries and compared bit for bit with a
COO assembly in the reference's contribution order (cells ascending, trial
outer, test inner; src/assembly.jl:189-208)."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "galerkintoolkit.jl_amd", "csrc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = tmp_path_factory.mktemp("plan") / "libplan_harness.so"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-shared", "-fPIC", "-I", CSRC,
                    os.path.join(ROOT, "tests", "host", "plan_harness.cpp"), os.path.join(CSRC, "gtk_plan.cpp"),
                    "-o", str(out), "-lpthread"], check=True)
    L = C.CDLL(str(out))
    f64 = np.ctypeslib.ndpointer(np.float64, flags="C")
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C")
    L.harness_run.argtypes = [C.c_int, C.c_int64, f64, C.c_int64, C.c_int, i32, C.c_int, C.c_int64, C.c_int64, i32,
                              C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64), i32, i32, f64, C.c_int64,
                              C.POINTER(C.c_int), C.c_int, C.c_char_p, C.c_int]
    return L


def coo_reference(cell_dofs, rows, cols, nrows, ncols):
    """COO triplets in the reference order, duplicates summed, CSC."""
    import scipy.sparse as sp
    I, J, V = [], [], []
    nc, nd = cell_dofs.shape
    sel = (lambda d: d > 0) if True else None
    for c in range(nc):
        for b in range(nd):
            dj = cell_dofs[c, b]
            if not ((dj > 0) if cols == 1 else (dj < 0)):
                continue
            for a in range(nd):
                di = cell_dofs[c, a]
                if not ((di > 0) if rows == 1 else (di < 0)):
                    continue
                I.append(abs(di) - 1)
                J.append(abs(dj) - 1)
                V.append(1.0 + c * 7.0 + a * 131.0 + b * 977.0)
    A = sp.coo_matrix((V, (I, J)), shape=(nrows, ncols)).tocsc()
    A.sum_duplicates()
    A.sort_indices()
    return A


def run(harness, GT, cells, order, simplex=False, diri=0, rows=1, cols=1, max_cells=64, chunk=8, mode=0):
    D = len(cells)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells, simplexify=simplex)
    db = GT.boundary(m) if diri else None
    V = GT.lagrange_space(GT.interior(m), order, dirichlet_boundary=db)
    cd = np.ascontiguousarray(V.cell_dofs, np.int32)
    cn = np.ascontiguousarray(m.cells, np.int32)
    cap = 10_000_000
    colptr = np.zeros((V.nfree if cols == 1 else V.ndiri) + 1, np.int32)
    rowval = np.zeros(cap, np.int32)
    nz = np.zeros(cap)
    nnz = C.c_int64()
    nt = C.c_int()
    err = C.create_string_buffer(256)
    rc = harness.harness_run(D, m.coords.shape[0], np.ascontiguousarray(m.coords), cn.shape[0], cn.shape[1], cn,
                             cd.shape[1], V.nfree, V.ndiri, cd, rows, cols, max_cells, chunk, C.byref(nnz), colptr,
                             rowval, nz, cap, C.byref(nt), mode, err, 256)
    if rc == 2:
        return None  # batch records not applicable (a column with more than 16 contributors)
    assert rc == 0, err.value.decode()
    nrows = V.nfree if rows == 1 else V.ndiri
    ref = coo_reference(V.cell_dofs, rows, cols, nrows, len(colptr) - 1)
    n = nnz.value
    assert np.array_equal(colptr - 1, ref.indptr)
    assert np.array_equal(rowval[:n] - 1, ref.indices)
    assert np.array_equal(nz[:n], ref.data)
    return nt.value


@pytest.mark.parametrize("cells,order,simplex,diri", [
    ((6, 5), 1, False, 1), ((5, 4), 2, False, 0), ((4, 5), 3, False, 1), ((4, 3, 5), 1, False, 0),
    ((3, 4, 3), 2, False, 1), ((3, 3, 2), 2, True, 0), ((4, 4), 2, True, 1)])
@pytest.mark.parametrize("max_cells,chunk", [(64, 8), (27, 1), (200, 64)])
@pytest.mark.parametrize("mode", [0, 1], ids=["offsets", "batches"])
def test_plan_replay(harness, GT, cells, order, simplex, diri, max_cells, chunk, mode):
    run(harness, GT, cells, order, simplex, diri, max_cells=max_cells, chunk=chunk, mode=mode)


@pytest.mark.parametrize("mode", [0, 1], ids=["offsets", "batches"])
def test_plan_dirichlet_columns(harness, GT, mode):
    run(harness, GT, (5, 6), 2, diri=1, rows=1, cols=2, mode=mode)
    # P1 triangles and tetrahedra: the corner vertices' columns hold no free
    # row (every vertex of their cells is a Dirichlet vertex)
    run(harness, GT, (6, 5), 1, simplex=True, diri=1, rows=1, cols=2, mode=mode)
    run(harness, GT, (3, 3, 2), 1, simplex=True, diri=1, rows=1, cols=2, mode=mode)


def test_batches_cover_hex_q2(harness, GT):
    """Q2 hexahedra: every column has at most 8 contributors, so the batch
    records apply (the replay must not report them as inapplicable)."""
    assert run(harness, GT, (4, 3, 3), 2, diri=1, max_cells=180, chunk=32, mode=1) is not None


def test_transpose_perm(harness, GT):
    """Row-order permutation of a square CSC block (the solver's CSR view):
    values gathered through it equal scipy's CSR data, and a pattern that is
    not structurally symmetric is rejected."""
    import scipy.sparse as sp
    rng = np.random.default_rng(3)
    m = GT.cartesian_mesh((0.0, 1.0) * 2, (5, 4))
    V = GT.lagrange_space(GT.interior(m), 2, dirichlet_boundary=GT.boundary(m))
    ref = coo_reference(V.cell_dofs, 1, 1, V.nfree, V.nfree)
    ref.data = rng.standard_normal(ref.nnz)  # values need not be symmetric
    colptr = (ref.indptr + 1).astype(np.int32)
    rowval = (ref.indices + 1).astype(np.int32)
    perm = np.zeros(ref.nnz, np.int32)
    err = C.create_string_buffer(128)
    H = harness
    H.harness_transpose.argtypes = [C.c_int64, np.ctypeslib.ndpointer(np.int32, flags="C"),
                                    np.ctypeslib.ndpointer(np.int32, flags="C"),
                                    np.ctypeslib.ndpointer(np.int32, flags="C"), C.c_char_p, C.c_int]
    assert H.harness_transpose(V.nfree, colptr, rowval, perm, err, 128) == 0, err.value
    csr = ref.tocsr()
    csr.sort_indices()
    assert np.array_equal(ref.data[perm], csr.data)
    assert np.array_equal(csr.indices + 1, rowval) and np.array_equal(csr.indptr + 1, colptr)
    # drop one off-diagonal entry: no longer structurally symmetric
    A = ref.tolil()
    i, j = 0, int(ref.indices[ref.indptr[0] + 1])
    A[j, i] = 0.0
    B = sp.csc_matrix(A)
    B.eliminate_zeros()
    cp = (B.indptr + 1).astype(np.int32)
    rv = (B.indices + 1).astype(np.int32)
    assert H.harness_transpose(V.nfree, cp, rv, np.zeros(B.nnz + 1, np.int32), err, 128) != 0


@pytest.mark.parametrize("mode", [0, 2], ids=["offsets", "batches"])
def test_duplicate_cell_dof_rejected(harness, GT, mode):
    """A cell listing one DOF twice would make two lanes of one accumulation
    instruction add into the same nonzero (order no longer fixed): both the
    VALU offsets and the batch records refuse such a space."""
    m = GT.cartesian_mesh((0.0, 1.0) * 2, (3, 3))
    V = GT.lagrange_space(GT.interior(m), 1)
    cd = np.ascontiguousarray(V.cell_dofs, np.int32).copy()
    cd[4, 1] = cd[4, 0]
    cn = np.ascontiguousarray(m.cells, np.int32)
    cap = 100_000
    colptr = np.zeros(V.nfree + 1, np.int32)
    err = C.create_string_buffer(256)
    rc = harness.harness_run(2, m.coords.shape[0], np.ascontiguousarray(m.coords), cn.shape[0], cn.shape[1], cn,
                             cd.shape[1], V.nfree, V.ndiri, cd, 1, 1, 64, 8, C.byref(C.c_int64()), colptr,
                             np.zeros(cap, np.int32), np.zeros(cap), cap, C.byref(C.c_int()), mode, err, 256)
    assert rc == -1 and b"twice" in err.value, (rc, err.value)


def test_nnz_over_int32_rejected(harness, GT):
    """The pattern's nonzero count must fit SparseMatrixCSC{Float64,Int32}:
    the symbolic phase refuses a pattern past its limit (forced here with a
    limit one below this pattern's nnz) instead of wrapping colptr."""
    m = GT.cartesian_mesh((0.0, 1.0) * 2, (4, 3))
    V = GT.lagrange_space(GT.interior(m), 2)
    cd = np.ascontiguousarray(V.cell_dofs, np.int32)
    cn = np.ascontiguousarray(m.cells, np.int32)
    H = harness
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C")
    H.harness_pattern.argtypes = [C.c_int, C.c_int64, np.ctypeslib.ndpointer(np.float64, flags="C"), C.c_int64,
                                  C.c_int, i32, C.c_int, C.c_int64, C.c_int64, i32, C.c_int, C.c_int, C.c_int64,
                                  C.POINTER(C.c_int64), C.c_char_p, C.c_int]
    args = (2, m.coords.shape[0], np.ascontiguousarray(m.coords), cn.shape[0], cn.shape[1], cn, cd.shape[1], V.nfree,
            V.ndiri, cd, 1, 1)
    nnz, err = C.c_int64(), C.create_string_buffer(256)
    assert H.harness_pattern(*args, 2**31 - 2, C.byref(nnz), err, 256) == 0
    ref = coo_reference(V.cell_dofs, 1, 1, V.nfree, V.nfree)
    assert nnz.value == ref.nnz
    assert H.harness_pattern(*args, nnz.value, C.byref(C.c_int64()), err, 256) == 0
    assert H.harness_pattern(*args, nnz.value - 1, C.byref(C.c_int64()), err, 256) == -1
    assert b"Int32" in err.value
