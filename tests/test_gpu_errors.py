"""Failure handling of the C ABI, each condition forced from the host: bad
node ids, DOF ids outside the GT signed encoding, NULL arrays, calls out of
order, unknown forms and precisions, a device allocation refused by the
per-allocation cap (GTK_MEM_LIMIT_MB), and a context that keeps working
after each refused call.  No kernel runs on invalid input: every check
returns before a launch.  (The reference checks such input with asserts,
e.g. src/assembly.jl:474-476; the library returns GTK_E* codes instead.)"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EINVAL, ESTATE, EDEVICE, ENOMEM, EUNSUP = -1, -2, -3, -4, -5


def q1_square(GT, n=3, diri=True):
    m = GT.cartesian_mesh((0.0, 1.0) * 2, (n, n))
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 1, dirichlet_boundary=GT.boundary(m) if diri else None)
    return m, Om, V


def err(ctx):
    return ctx.L.gtk_last_error(ctx.h).decode()


def test_mesh_and_space_validation(GT):
    print("case errors: mesh and space", flush=True)
    L = GT._lib
    m, Om, V = q1_square(GT)
    ctx = L.Context(0)
    lib, h = ctx.L, ctx.h
    xyz = np.ascontiguousarray(m.coords)
    cn = np.ascontiguousarray(m.cell_nodes_1b(), np.int32)
    # space before mesh
    cd = np.ascontiguousarray(V.cell_dofs, np.int32)
    assert lib.gtk_set_space(h, 1, cd.shape[1], V.nfree, V.ndiri, cd) == ESTATE
    # node id 0 (ids are 1-based) and past the last node
    for bad in (0, xyz.shape[0] + 1):
        c2 = cn.copy()
        c2[2, 1] = bad
        assert lib.gtk_set_mesh(h, 2, xyz.shape[0], xyz, c2.shape[0], c2.shape[1], c2) == EINVAL
        assert "out of range" in err(ctx)
    # NULL coordinates, through a raw pointer argument
    f = lib.gtk_set_mesh
    assert C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int64, C.c_void_p, C.c_int64, C.c_int, C.c_void_p)(
        C.cast(f, C.c_void_p).value)(h, 2, xyz.shape[0], None, cn.shape[0], cn.shape[1], cn.ctypes.data) == EINVAL
    assert lib.gtk_set_mesh(h, 4, xyz.shape[0], xyz, cn.shape[0], cn.shape[1], cn) == EINVAL  # dim 4
    assert lib.gtk_set_mesh(h, 2, xyz.shape[0], xyz, cn.shape[0], cn.shape[1], cn) == 0
    # DOF ids outside the signed encoding: 0, > n_free, < -n_diri
    for bad in (0, V.nfree + 1, -(V.ndiri + 1)):
        c2 = cd.copy()
        c2[1, 2] = bad
        assert lib.gtk_set_space(h, 1, c2.shape[1], V.nfree, V.ndiri, c2) == EINVAL
        assert "GT signed encoding" in err(ctx)
    assert lib.gtk_set_space(h, 2, 3, V.nfree, V.ndiri, cd) == EINVAL  # dofs_per_cell not a multiple of ncomp
    assert lib.gtk_set_space(h, 1, cd.shape[1], V.nfree, V.ndiri, cd) == 0
    ctx.close()


def test_call_order_and_arguments(GT, O):
    print("case errors: call order, forms, precisions", flush=True)
    L = GT._lib
    m, Om, V = q1_square(GT, n=4)
    ctx = L.Context(0)
    lib, h = ctx.L, ctx.h
    ctx.set_mesh(2, m.coords, m.cell_nodes_1b())
    nnz = C.c_int64()
    assert lib.gtk_build_pattern(h, L.GTK_FREE, L.GTK_FREE, C.byref(nnz), None) == ESTATE  # no space yet
    ctx.set_space(1, V.cell_dofs, V.nfree, V.ndiri)
    out = np.zeros(V.nfree)
    # assembly before the pattern; vector before the reference tables
    assert lib.gtk_assemble_matrix(h, L.GTK_LAPLACE, L.GTK_FP64, L.GTK_FREE, L.GTK_FREE, None, None) == ESTATE
    assert lib.gtk_assemble_vector(h, L.GTK_F_CONST, np.ones(1), L.GTK_FREE, out.ctypes.data, None) == ESTATE
    assert lib.gtk_build_pattern(h, 3, L.GTK_FREE, C.byref(nnz), None) == EINVAL  # rows selector
    assert lib.gtk_build_pattern(h, L.GTK_FREE, L.GTK_FREE, C.byref(nnz), None) == 0
    dO = GT.measure(Om, 2)
    phi, gphi = V.fe.tabulate(dO.x)
    _, gN = GT.refcell.Lagrange(m.cellkind, 1).tabulate(dO.x)
    ctx.set_reference(dO.w, phi, gphi, gN)
    assert lib.gtk_assemble_matrix(h, 7, L.GTK_FP64, L.GTK_FREE, L.GTK_FREE, None, None) == EINVAL  # form
    assert lib.gtk_assemble_matrix(h, L.GTK_LAPLACE, 5, L.GTK_FREE, L.GTK_FREE, None, None) == EINVAL  # precision
    assert lib.gtk_assemble_matrix(h, L.GTK_ELASTICITY, L.GTK_FP64, L.GTK_FREE, L.GTK_FREE, None, None) == ESTATE
    assert lib.gtk_assemble_matrix(h, L.GTK_LAPLACE, L.GTK_FP64, L.GTK_FREE, L.GTK_DIRICHLET, None, None) == ESTATE
    assert lib.gtk_assemble_vector(h, 9, np.ones(1), L.GTK_FREE, out.ctypes.data, None) == EINVAL  # load kind
    assert lib.gtk_lift_dirichlet(h, L.GTK_FREE, np.zeros(V.ndiri), out) == ESTATE  # no (FREE, DIRICHLET) block
    x = np.zeros(V.nfree)
    st = L.GtkSolveStats()
    assert lib.gtk_solve(h, 3, out, x, 1e-10, 5, 100, 1e-4, C.byref(st)) == EINVAL  # precision
    # after the refused calls the context still assembles the oracle's matrix
    vals, _ = ctx.assemble_matrix(L.GTK_LAPLACE, L.GTK_FP64, L.GTK_FREE, L.GTK_FREE, nnz.value)
    p = O.Problem([4, 4]).space(1, dirichlet=1)
    p.quadrature(2)
    ref = p.assemble_matrix(O.LAPLACE)[2]
    assert np.linalg.norm(vals - ref) <= 1e-13 * np.linalg.norm(ref)
    ctx.close()


def test_allocation_cap(GT, O, monkeypatch):
    """A device allocation past GTK_MEM_LIMIT_MB fails with GTK_ENOMEM (the
    path of a failed hipMalloc); smaller problems on the same setting work."""
    print("case errors: allocation cap", flush=True)
    L = GT._lib
    monkeypatch.setenv("GTK_MEM_LIMIT_MB", "1")
    m, Om, V = q1_square(GT, n=200, diri=False)  # 40,401 nodes: 970 KB of coordinates
    ctx = L.Context(0)
    ctx.set_mesh(2, m.coords, m.cell_nodes_1b())
    ctx.set_space(1, V.cell_dofs, V.nfree, V.ndiri)
    nnz = C.c_int64()
    rc = ctx.L.gtk_build_pattern(ctx.h, L.GTK_FREE, L.GTK_FREE, C.byref(nnz), None)  # values: 2.8 MB
    assert rc == ENOMEM and "GTK_MEM_LIMIT_MB" in err(ctx)
    ctx.close()
    m2, Om2, V2 = q1_square(GT, n=6, diri=True)
    dO = GT.measure(Om2, 2)
    A = GT.assemble_matrix(GT.Laplace(dO), float, V2, V2)
    p = O.Problem([6, 6]).space(1, dirichlet=1)
    p.quadrature(2)
    ref = p.assemble_matrix(O.LAPLACE)[2]
    assert np.linalg.norm(A.nzval - ref) <= 1e-13 * np.linalg.norm(ref)
