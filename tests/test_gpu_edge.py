"""Edge cases on the GPU: a single Q2 cell with one free DOF (the
reference's test/assembly_tests.jl:72-73), a mesh whose DOFs are all
Dirichlet (empty FREE block, FREE x DIRICHLET block with zero rows), a
space with no Dirichlet DOFs (empty FREE x DIRICHLET block), for the matrix
kernels, the load vector, the device symbolic phase and the solver."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_single_q2_cell(GT):
    print("case single Q2 cell", flush=True)
    m = GT.cartesian_mesh((0.0, 1.0) * 2, (1, 1))
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 2, dirichlet_boundary=GT.boundary(m))
    assert (V.nfree, V.ndiri) == (1, 8)
    dO = GT.measure(Om, 4)
    for prec in (GT.FP64, GT.MIXED):
        A = GT.assemble_matrix(GT.Laplace(dO), float, V, V, precision=prec)
        p = O.Problem([1, 1]).space(2, dirichlet=1)
        p.quadrature(4)
        cp, rv, nz = p.assemble_matrix(O.LAPLACE)
        assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
        assert abs(A.nzval[0] - nz[0]) <= (1e-13 if prec == GT.FP64 else 1e-6) * abs(nz[0])
    b = GT.assemble_vector(GT.Load(dO, 1.0), float, V)
    assert abs(b[0] - p.assemble_vector(0, [1.0])[0]) <= 1e-13 * abs(b[0])
    Ad = GT.assemble_matrix(GT.Laplace(dO), float, V, V, (GT.FREE, GT.DIRICHLET))
    assert Ad.shape == (1, 8) and Ad.nnz == 8
    x, st = GT.solve(A, b, rtol=1e-12)
    assert abs(x[0] * A.nzval[0] - b[0]) <= 1e-12 * abs(b[0])


def test_all_dirichlet(GT):
    """Q1 on one cell with the whole boundary Dirichlet: no free DOF."""
    print("case all dirichlet", flush=True)
    m = GT.cartesian_mesh((0.0, 1.0) * 2, (1, 1))
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 1, dirichlet_boundary=GT.boundary(m))
    assert (V.nfree, V.ndiri) == (0, 4)
    dO = GT.measure(Om, 2)
    A = GT.assemble_matrix(GT.Laplace(dO), float, V, V)
    assert A.shape == (0, 0) and A.nnz == 0 and len(A.colptr) == 1
    Ad = GT.assemble_matrix(GT.Laplace(dO), float, V, V, (GT.FREE, GT.DIRICHLET))
    assert Ad.shape == (0, 4) and Ad.nnz == 0 and np.array_equal(Ad.colptr, np.ones(5, np.int32))
    b = GT.assemble_vector(GT.Load(dO, 1.0), float, V)
    assert b.shape == (0,)
    bd = GT.assemble_vector(GT.Load(dO, 1.0), float, V, GT.DIRICHLET)
    assert abs(bd.sum() - 1.0) < 1e-14
    from gtk_amd import assembly as AS
    a = AS._Assembler(V, dO)
    nnz, ncols, _ = a.ctx.build_pattern_device(GT.FREE, GT.FREE)
    assert (nnz, ncols) == (0, 0)
    nnz, ncols, _ = a.ctx.build_pattern_device(GT.FREE, GT.DIRICHLET)
    assert (nnz, ncols) == (0, 4)


def test_no_dirichlet_columns(GT):
    """No Dirichlet DOFs: the FREE x DIRICHLET block has no columns."""
    print("case no dirichlet", flush=True)
    m = GT.cartesian_mesh((0.0, 1.0) * 3, (2, 2, 2))
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 2)
    dO = GT.measure(Om, 4)
    Ad = GT.assemble_matrix(GT.Mass(dO), float, V, V, (GT.FREE, GT.DIRICHLET), precision=GT.MIXED)
    assert Ad.shape == (V.nfree, 0) and Ad.nnz == 0
    M = GT.assemble_matrix(GT.Mass(dO), float, V, V)
    assert abs(M.nzval.sum() - 1.0) < 1e-13  # sum of the mass matrix = |Omega|
