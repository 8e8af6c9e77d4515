"""Reuse and orchestration paths of the GPU assembly against the CPU oracle
(through the C ABI): update_matrix / update_vector after new coefficients
and after moved nodes (gtk_update_coordinates), the
assemble_matrix_and_vector* families with their update twins
(src/problems.jl:276-285, :352-435), the Dirichlet lifting on the device
(src/problems.jl:439-452), and the solve of a matrix whose device values
were overwritten by another form on the same block."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-13
TOL_MIXED = 1e-6


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def setup(GT, O, cells, order, diri=1, perturb=0.15, seed=9):
    D = len(cells)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells)
    m.coords = GT.perturb_interior_nodes(m, perturb, seed)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, order, dirichlet_boundary=GT.boundary(m) if diri else None)
    p = O.Problem(cells).space(order, dirichlet=diri)
    p.set_coords(m.coords)
    return m, Om, V, p


@pytest.mark.parametrize("cells,order,degree,precision", [((7, 6), 2, 4, "fp64"), ((4, 3, 4), 2, 4, "fp64"),
                                                          ((4, 3, 4), 2, 4, "mixed"), ((5, 4, 3), 1, 2, "mixed")])
def test_update_matrix_coefficients_and_coordinates(GT, O, cells, order, degree, precision):
    """update_matrix with a new kappa, then with moved nodes, equals the
    oracle's assembly of the new problem; the pattern object is unchanged."""
    print(f"case update_matrix {cells} Q{order} {precision}", flush=True)
    tol = TOL_FP64 if precision == "fp64" else TOL_MIXED
    prec = GT.FP64 if precision == "fp64" else GT.MIXED
    m, Om, V, p = setup(GT, O, cells, order)
    dO = GT.measure(Om, degree)
    p.quadrature(degree)
    nc = m.cells.shape[0]
    k1, k2 = GT.inputs.lognormal(41, nc), GT.inputs.lognormal(42, nc)
    A, cache = GT.assemble_matrix(GT.Laplace(dO, GT.face_constant_field(k1, Om)), float, V, V, precision=prec,
                                  reuse=True)
    nz_id = id(A.nzval)
    assert relfro(A.nzval, p.assemble_matrix(O.LAPLACE, coef=k1)[2]) <= tol
    GT.update_matrix(A, cache, parameters={"kappa": GT.face_constant_field(k2, Om)})
    assert id(A.nzval) == nz_id  # refilled in place, like update_matrix!
    assert relfro(A.nzval, p.assemble_matrix(O.LAPLACE, coef=k2)[2]) <= tol
    x2 = GT.perturb_interior_nodes(m, 0.1, 77)
    GT.update_matrix(A, cache, coords=x2)
    p.set_coords(x2)
    cp, rv, nz = p.assemble_matrix(O.LAPLACE, coef=k2)
    assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
    assert relfro(A.nzval, nz) <= tol


def test_update_vector_and_matrix_and_vector(GT, O):
    """assemble_matrix_and_vector / update_matrix_and_vector and
    update_vector with a new nodal load, against the oracle."""
    print("case update_vector / matrix_and_vector", flush=True)
    m, Om, V, p = setup(GT, O, (6, 5), 3)
    dO = GT.measure(Om, 7)
    p.quadrature(7)
    f1 = GT.inputs.normal(51, V.nfree + V.ndiri)
    f2 = GT.inputs.normal(52, V.nfree + V.ndiri)
    nc = m.cells.shape[0]
    k = GT.inputs.lognormal(53, nc)
    a = GT.Mass(dO, GT.face_constant_field(k, Om))
    l = GT.Load(dO, GT.DiscreteField(V, f1[:V.nfree], f1[V.nfree:]))
    A, b, cache = GT.assemble_matrix_and_vector(a, l, float, V, V, reuse=True)
    assert relfro(A.nzval, p.assemble_matrix(O.MASS, coef=k)[2]) <= TOL_FP64
    assert relfro(b, p.assemble_vector(2, f1)) <= TOL_FP64
    b_id = id(b)
    GT.update_matrix_and_vector(A, b, cache, parameters={"f": GT.DiscreteField(V, f2[:V.nfree], f2[V.nfree:]),
                                                         "kappa": None})
    assert id(b) == b_id  # refilled in place, like update_vector!
    assert relfro(A.nzval, p.assemble_matrix(O.MASS)[2]) <= TOL_FP64
    assert relfro(b, p.assemble_vector(2, f2)) <= TOL_FP64
    b1, vcache = GT.assemble_vector(GT.Load(dO, 2.0), float, V, reuse=True)
    GT.update_vector(b1, vcache, parameters={"f": -0.5})
    assert relfro(b1, p.assemble_vector(0, [-0.5])) <= TOL_FP64


def test_free_and_dirichlet_family_and_lifting(GT, O):
    """assemble_matrix_and_vector_with_free_and_dirichlet_columns and its
    update twin against the oracle's FREE x FREE / FREE x DIRICHLET blocks;
    the device lifting b - Ad xd against the same product in FP64 on the
    host."""
    print("case free/dirichlet family + lifting", flush=True)
    m, Om, V, p = setup(GT, O, (5, 4, 3), 2)
    dO = GT.measure(Om, 4)
    p.quadrature(4)
    nc = m.cells.shape[0]
    k1, k2 = GT.inputs.lognormal(61, nc), GT.inputs.lognormal(62, nc)
    a = GT.Laplace(dO, GT.face_constant_field(k1, Om))
    A, Ad, b, cache = GT.assemble_matrix_and_vector_with_free_and_dirichlet_columns(a, GT.Load(dO, 1.0), float, V, V,
                                                                                   reuse=True)
    GT.update_matrix_and_vector_with_free_and_dirichlet_columns(A, Ad, b, cache,
                                                                parameters={"kappa": GT.face_constant_field(k2, Om)})
    cp, rv, nz = p.assemble_matrix(O.LAPLACE, coef=k2)
    cpd, rvd, nzd = p.assemble_matrix(O.LAPLACE, coef=k2, rows=O.FREE, cols=O.DIRICHLET)
    assert np.array_equal(A.rowval, rv) and relfro(A.nzval, nz) <= TOL_FP64
    assert np.array_equal(Ad.colptr, cpd) and np.array_equal(Ad.rowval, rvd) and relfro(Ad.nzval, nzd) <= TOL_FP64
    assert relfro(b, p.assemble_vector(0, [1.0])) <= TOL_FP64
    xd = GT.inputs.normal(63, V.ndiri)
    lifted = Ad.device[0].ctx.lift_dirichlet(xd, b)
    ref = b - Ad.to_scipy() @ xd
    assert np.max(np.abs(lifted - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_solve_uses_its_own_matrix(GT):
    """Assembling a mass matrix on the same space and measure overwrites the
    device values of the (FREE, FREE) block; solve(A_laplace) must still
    solve with the Laplace values (uploaded again), and a replaced A.nzval
    is seen too."""
    print("case solve after another form on the block", flush=True)
    m = GT.cartesian_mesh((0.0, 1.0) * 2, (12, 10))
    m.coords = GT.perturb_interior_nodes(m, 0.15, 4)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 2, dirichlet_boundary=GT.boundary(m))
    dO = GT.measure(Om, 4)
    A = GT.assemble_matrix(GT.Laplace(dO), float, V, V)
    M = GT.assemble_matrix(GT.Mass(dO), float, V, V)
    b = GT.inputs.normal(71, V.nfree)
    x, st = GT.solve(A, b, rtol=1e-12)
    xl = sla.spsolve(A.to_scipy().tocsc(), b)
    assert np.linalg.norm(x - xl) / np.linalg.norm(xl) <= 1e-9
    xm, _ = GT.solve(M, b, rtol=1e-12)
    xml = sla.spsolve(M.to_scipy().tocsc(), b)
    assert np.linalg.norm(xm - xml) / np.linalg.norm(xml) <= 1e-9
    with pytest.raises(ValueError):
        A.nzval[0] = 1.0  # device-backed values change by replacement only
    A.nzval = A.nzval * 2.0
    x2, _ = GT.solve(A, b, rtol=1e-12)
    assert np.linalg.norm(x2 - xl / 2.0) / np.linalg.norm(xl / 2.0) <= 1e-9


def test_batch_refuses_stale_plans(GT):
    """A batch built before one of its problems rebuilt its pattern refuses
    to run (GTK_ESTATE) instead of scattering with stale offsets; building it
    again makes it run."""
    print("case batch generation check", flush=True)
    L = GT._lib
    m = GT.cartesian_mesh((0.0, 1.0) * 2, (10, 9))
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 2, dirichlet_boundary=GT.boundary(m))
    dO = GT.measure(Om, 4)
    A = GT.assemble_matrix(GT.Laplace(dO), float, V, V)
    ctx = A.device[0].ctx
    bt = L.Batch(0, L.GTK_LAPLACE, L.GTK_FP64)
    bt.add(ctx)
    bt.build()
    bt.run()
    ctx.build_pattern(L.GTK_FREE, L.GTK_FREE)
    with pytest.raises(L.GtkError, match="gtk_batch_build again"):
        bt.run()
    bt.build()
    bt.run()
    nz = ctx.get_values(L.GTK_FREE, L.GTK_FREE, A.nnz)
    assert relfro(nz, A.nzval) <= 1e-15
    bt.close()
