"""Nitsche boundary terms on the GPU (GalerkinToolkitExamples/src/poisson.jl:93-112)
against the oracle (oracle/nitsche.py), through the C ABI:
- the matrix of int (gamma/h) v u - v n.grad(u) - n.grad(v) u over the whole
  boundary and over one side (the pattern of the oracle's COO assembly bit
  for bit, FP64 values within 1e-13, the mixed form within 1e-6);
- Laplace plus the Nitsche terms as one matrix (the boundary block added into
  the volume block on the device);
- the load int (gamma/h) v g - n.grad(v) g of a discrete g;
- the patch test of the reference (poisson_tests.jl:31-44): the solve of the
  assembled system reproduces a linear u on a perturbed mesh."""
import numpy as np
import pytest

from nitsche import NitscheFaces

pytestmark = pytest.mark.gpu

TOL64, TOLMX = 1e-13, 1e-6


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def setup(GT, O, cells, order, perturb, group=None, seed=11):
    D = len(cells)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells)
    m.coords = GT.perturb_interior_nodes(m, perturb, seed)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, order)
    G = GT.boundary(m, [group] if group else None)
    p = O.Problem(list(cells)).space(order, dirichlet=0)
    p.set_coords(m.coords)
    p.quadrature(2 * order)
    select = None
    if group:
        # the side the group covers: the coordinate its vertices share
        cf = m.cell_faces[D - 1]
        owner = np.full(m.num_faces(D - 1), -1, np.int64)
        owner[cf.ravel()] = np.arange(cf.size)
        hit = owner[G.faces]
        lf = np.asarray(GT.refcell.GEO[m.cellkind].faces[D - 1])[hit % cf.shape[1]] - 1
        X = m.coords[np.take_along_axis(m.cells[hit // cf.shape[1]], lf, axis=1)].reshape(-1, D)
        axis = int(np.argmin(X.max(axis=0) - X.min(axis=0)))
        val = X[0, axis]
        select = lambda y: abs(y[axis] - val) < 1e-12
    return m, Om, V, G, NitscheFaces(p, order, select), p


@pytest.mark.parametrize("cells,order,perturb,group", [
    ((4, 3), 1, 0.2, None), ((3, 4), 2, 0.2, None), ((3, 3), 3, 0.15, None), ((4, 3), 2, 0.2, "1-face-2"),
    ((3, 2, 3), 1, 0.15, None), ((2, 3, 2), 2, 0.15, None), ((2, 2, 3), 2, 0.1, "2-face-5")])
def test_nitsche_matrix(GT, O, cells, order, perturb, group):
    print(f"case nitsche matrix {cells} Q{order} group={group}", flush=True)
    m, Om, V, G, nf, p = setup(GT, O, cells, order, perturb, group)
    degree = 2 * order
    gamma = 2.5
    dG = GT.measure(G, degree)
    ref = nf.matrix(degree, gamma, shape=(V.nfree, V.nfree))
    for prec, tol in ((GT.FP64, TOL64), (GT.MIXED, TOLMX)):
        A = GT.assemble_matrix(GT.Nitsche(dG, gamma), float, V, V, precision=prec)
        assert np.array_equal(A.colptr - 1, ref.indptr) and np.array_equal(A.rowval - 1, ref.indices)
        assert relfro(A.nzval, ref.data) <= tol, prec


@pytest.mark.parametrize("cells,order", [((4, 3), 2), ((2, 3, 2), 1)])
def test_laplace_plus_nitsche_and_load(GT, O, cells, order):
    import scipy.sparse as sp
    print(f"case laplace + nitsche {cells} Q{order}", flush=True)
    m, Om, V, G, nf, p = setup(GT, O, cells, order, 0.2)
    degree = 2 * order
    gamma = 4.0
    dO, dG = GT.measure(Om, degree), GT.measure(G, degree)
    n = V.nfree
    cp, rv, nz = p.assemble_matrix(O.LAPLACE)
    ref = (sp.csc_matrix((nz, rv - 1, cp - 1), shape=(n, n)) + nf.matrix(degree, gamma, shape=(n, n))).tocsc()
    ref.sort_indices()
    A = GT.assemble_matrix(GT.Laplace(dO) + GT.Nitsche(dG, gamma), float, V, V)
    assert np.array_equal(A.colptr - 1, ref.indptr) and np.array_equal(A.rowval - 1, ref.indices)
    assert relfro(A.nzval, ref.data) <= TOL64
    # the load of a discrete g: nodal values of a linear function, which the
    # (isoparametric) space reproduces, so the oracle evaluates it exactly
    D = len(cells)
    u = lambda y: 0.5 - y[0] + 2.0 * y[1] + (0.25 * y[2] if D == 3 else 0.0)
    pos = np.zeros((n, D))
    for c in range(nf.cd.shape[0]):
        for a, d in enumerate(nf.cd[c]):
            pos[d - 1] = nf.cell_pos[c, a]
    g = GT.DiscreteField(V, np.array([u(y) for y in pos]))
    b = GT.assemble_vector(GT.NitscheLoad(dG, g, gamma), float, V)
    bref = nf.load(degree, gamma, u, n=n)
    assert relfro(b, bref) <= TOL64


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_patch_solve(GT, O, precision):
    """Laplace + Nitsche with the Nitsche load of g = u, solved on the device,
    reproduces the linear u (poisson_tests.jl: error below 1e-10)."""
    print(f"case nitsche patch solve {precision}", flush=True)
    cells, order = (5, 4), 2
    m, Om, V, G, nf, p = setup(GT, O, cells, order, 0.2)
    degree = 2 * order
    gamma = 40.0
    dO, dG = GT.measure(Om, degree), GT.measure(G, degree)
    u = lambda y: 1.0 + 2.0 * y[0] - y[1]
    n = V.nfree
    pos = np.zeros((n, 2))
    for c in range(nf.cd.shape[0]):
        for a, d in enumerate(nf.cd[c]):
            pos[d - 1] = nf.cell_pos[c, a]
    ue = np.array([u(y) for y in pos])
    prec = GT.FP64 if precision == "fp64" else GT.MIXED
    A = GT.assemble_matrix(GT.Laplace(dO) + GT.Nitsche(dG, gamma), float, V, V)
    b = GT.assemble_vector(GT.NitscheLoad(dG, GT.DiscreteField(V, ue), gamma), float, V)
    x, st = GT.solve(A, b, precision=prec, rtol=1e-12)
    assert st.converged and np.abs(x - ue).max() < 1e-9


@pytest.mark.parametrize("cells,order,perturb", [((5, 4), 1, 0.2), ((4, 3), 2, 0.15), ((2, 3, 2), 1, 0.1),
                                                 ((2, 2, 2), 2, 0.0)])
def test_nitsche_matrix_simplex(GT, O, cells, order, perturb):
    """The Nitsche terms on P1/P2 triangles and tetrahedra over the whole
    boundary, FP64 and mixed, against the oracle."""
    print(f"case nitsche simplex {cells} P{order}", flush=True)
    D = len(cells)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells, simplexify=True)
    if perturb:
        m.coords = GT.perturb_interior_nodes(m, perturb, 13)
    V = GT.lagrange_space(GT.interior(m), order)
    p = O.Problem(list(cells), simplexify=True).space(order, dirichlet=0)
    p.set_coords(m.coords)
    p.quadrature(2 * order)
    nf = NitscheFaces(p, order)
    degree = 2 * order
    dG = GT.measure(GT.boundary(m), degree)
    ref = nf.matrix(degree, 3.5, shape=(V.nfree, V.nfree))
    for prec, tol in ((GT.FP64, TOL64), (GT.MIXED, TOLMX)):
        A = GT.assemble_matrix(GT.Nitsche(dG, 3.5), float, V, V, precision=prec)
        assert np.array_equal(A.colptr - 1, ref.indptr) and np.array_equal(A.rowval - 1, ref.indices)
        assert relfro(A.nzval, ref.data) <= tol, prec
