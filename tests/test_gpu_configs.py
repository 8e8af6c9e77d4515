"""Benchmark configurations and kernel paths at oracle-checkable sizes,
through the C ABI:
  - the 2D Poisson workload exactly as the reference's CPU example is
    quoted (64 x 64 Q1 quads, full-boundary Dirichlet, 2 x 2 Gauss,
    stiffness with coefficient 1 and the load f = 1), FP64 and mixed;
  - the persistent matrix-core kernel with more tiles than workgroups, so
    that a workgroup's second and later tiles and the cross-tile prefetch of
    cell ids, node ids and coefficients run (GTK_LDS_KB shrinks the tiles),
    under lognormal kappa on Q1 and Q2 hexahedra, FP64 and mixed;
  - the device symbolic phase against the oracle's pattern directly;
  - the launch shape of the VALU kernel's 4-wave fallback.
Pattern bit-exact, FP64 within 1e-13, mixed within 1e-6 (relative
Frobenius)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-13, "mixed": 1e-6}


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def problem(GT, O, cells, order, diri, perturb=0.0, seed=7, simplex=False, ncomp=1, group=0):
    D = len(cells)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells, simplexify=simplex)
    if perturb:
        m.coords = GT.perturb_interior_nodes(m, perturb, seed)
    Om = GT.interior(m)
    db = None if diri == 0 else (GT.boundary(m) if diri == 1 else GT.boundary(m, [f"{D - 1}-face-{group}"]))
    V = GT.lagrange_space(Om, order, tensor_size=(ncomp,) if ncomp > 1 else None, dirichlet_boundary=db)
    p = O.Problem(cells, simplexify=simplex).space(order, ncomp=ncomp, dirichlet=diri, group=group)
    p.set_coords(m.coords)
    return m, Om, V, p


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_poisson2d_q1_64x64(GT, O, precision):
    print(f"case poisson2d q1 64x64 {precision}", flush=True)
    m, Om, V, p = problem(GT, O, (64, 64), 1, diri=1)
    assert (V.nfree, V.ndiri, m.cells.shape[0]) == (3969, 256, 4096)
    dO = GT.measure(Om, 2)
    p.quadrature(2)
    prec = GT.FP64 if precision == "fp64" else GT.MIXED
    A, b = GT.assemble_matrix_and_vector(GT.Laplace(dO), GT.Load(dO, 1.0), float, V, V, precision=prec)
    cp, rv, nz = p.assemble_matrix(O.LAPLACE)
    assert A.nnz == 34969
    assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
    assert relfro(A.nzval, nz) <= TOL[precision]
    assert relfro(b, p.assemble_vector(0, [1.0])) <= 1e-13


@pytest.mark.parametrize("cells,order,degree", [((24, 22, 20), 1, 2), ((14, 12, 13), 2, 4)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_persistent_multi_tile(GT, O, monkeypatch, cells, order, degree, precision):
    """Small LDS budget: several tiles per workgroup of the persistent grid
    (checked through gtk_plan_info), lognormal kappa per cell, perturbed
    nodes; every nonzero against the oracle."""
    monkeypatch.setenv("GTK_LDS_KB", "16")
    print(f"case persistent multi-tile {cells} Q{order} {precision}", flush=True)
    m, Om, V, p = problem(GT, O, cells, order, diri=1, perturb=0.2, seed=21)
    dO = GT.measure(Om, degree)
    p.quadrature(degree)
    k = GT.inputs.lognormal(22, m.cells.shape[0])
    prec = GT.FP64 if precision == "fp64" else GT.MIXED
    A, cache = GT.assemble_matrix(GT.Laplace(dO, GT.face_constant_field(k, Om)), float, V, V, precision=prec,
                                  reuse=True)
    L = GT._lib
    info = A.device[0].ctx.plan_info(L.GTK_LAPLACE, L.GTK_FP64 if precision == "fp64" else L.GTK_MIXED,
                                     L.GTK_FREE, L.GTK_FREE)
    print(f"  plan {info}", flush=True)
    assert info["matrix_cores"] == 1
    assert info["tiles"] >= 2 * info["grid"], info  # every workgroup runs at least two tiles
    cp, rv, nz = p.assemble_matrix(O.LAPLACE, coef=k)
    assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
    assert relfro(A.nzval, nz) <= TOL[precision]
    # a second launch on the same plan (the prefetch registers start over)
    k2 = GT.inputs.lognormal(23, m.cells.shape[0])
    GT.update_matrix(A, cache, parameters={"kappa": GT.face_constant_field(k2, Om)})
    assert relfro(A.nzval, p.assemble_matrix(O.LAPLACE, coef=k2)[2]) <= TOL[precision]


@pytest.mark.parametrize("cells,order,simplex,ncomp,diri,group,rows,cols", [
    ((9, 7), 3, False, 1, 1, 0, 1, 1), ((5, 4, 6), 2, False, 1, 1, 0, 1, 1), ((6, 5, 4), 1, False, 1, 0, 0, 1, 1),
    ((3, 3, 2), 2, True, 3, 2, 5, 1, 1), ((6, 5), 2, False, 1, 1, 0, 1, 2)])
def test_device_pattern_vs_oracle(GT, O, cells, order, simplex, ncomp, diri, group, rows, cols):
    """colptr/rowval of the device symbolic phase equal the oracle's COO ->
    CSC compress bit for bit (Q1-Q3, hexahedra, P2 vector tetrahedra, a
    FREE x DIRICHLET block)."""
    print(f"case device pattern {cells} Q{order} simplex={simplex} rows={rows} cols={cols}", flush=True)
    m, Om, V, p = problem(GT, O, cells, order, diri=diri, simplex=simplex, ncomp=ncomp, group=group)
    dO = GT.measure(Om, 2 * order)
    a = GT.assembly._Assembler(V, dO)
    nnz, ncols, nrows, colptr, rowval = a.pattern(rows, cols, symbolic="device")
    p.quadrature(2 * order)
    cp, rv, _ = p.assemble_matrix(O.MASS if ncomp == 1 else O.ELASTICITY,
                                  coef=None if ncomp == 1 else np.ones(2 * m.cells.shape[0]), rows=rows, cols=cols)
    assert np.array_equal(colptr, cp) and np.array_equal(rowval, rv)


def test_valu_four_wave_fallback_shape(GT, monkeypatch):
    """GTK_VALU_WAVES=4 must give the 256-thread VALU workgroup (and the
    default the 768/512-thread one)."""
    print("case valu 4-wave launch shape", flush=True)
    L = GT._lib
    shapes = {}
    for env in (None, "4"):
        if env:
            monkeypatch.setenv("GTK_VALU_WAVES", env)
        m = GT.cartesian_mesh((0.0, 1.0) * 3, (3, 2, 3), simplexify=True)
        Om = GT.interior(m)
        V = GT.lagrange_space(Om, 2, tensor_size=(3,), dirichlet_boundary=GT.boundary(m, ["2-face-5"]))
        dO = GT.measure(Om, 5)
        a = GT.assembly._Assembler(V, dO)
        a.pattern(L.GTK_FREE, L.GTK_FREE)
        nc = m.cells.shape[0]
        a.ctx.set_cell_coefficients(np.ones(2 * nc), 2)
        shapes[env] = a.ctx.plan_info(L.GTK_ELASTICITY, L.GTK_MIXED, L.GTK_FREE, L.GTK_FREE)
    assert shapes[None]["matrix_cores"] == 0 and shapes[None]["threads"] == 768
    assert shapes["4"]["threads"] == 256
