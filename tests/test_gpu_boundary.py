"""Boundary integrals on the GPU against the boundary oracle
(oracle/boundary.py, which integrates each face through its own map and its
own trace element): loads f v (constant, per face, nodal) and the mass
kappa u v over boundary faces, alone and added into a volume matrix.
FP64 within relative Frobenius error 1e-13, mixed within 1e-6; patterns
bit-exact."""
import numpy as np
import pytest
from boundary import Boundary

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-13


def relfro(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


# cells, order, simplex, ncomp, dirichlet (0 none, 1 all), group (None: whole boundary), perturb
CASES = [
    ((5, 4), 1, False, 1, 1, None, 0.2),
    ((4, 3), 2, False, 1, 0, None, 0.2),
    ((3, 4), 3, False, 1, 1, "1-face-2", 0.2),
    ((3, 2, 3), 1, False, 1, 0, None, 0.2),
    ((2, 3, 2), 2, False, 1, 1, "2-face-5", 0.2),
    ((2, 2, 2), 2, True, 3, 0, "2-face-2", 0.0),
    # scalar P1/P2 triangles and tetrahedra
    ((5, 4), 1, True, 1, 1, None, 0.2),
    ((4, 3), 2, True, 1, 0, "1-face-2", 0.1),
    ((2, 3, 2), 1, True, 1, 1, None, 0.1),
    ((2, 2, 2), 2, True, 1, 0, "2-face-5", 0.0),
]
SELECT = {None: None, "1-face-2": lambda x: abs(x[1] - 1.0) < 1e-12, "2-face-5": lambda x: abs(x[0]) < 1e-12,
          "2-face-2": lambda x: abs(x[2] - 1.0) < 1e-12}


def setup(GT, cells, order, simplex, ncomp, diri, group, perturb):
    D = len(cells)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells, simplexify=simplex)
    if perturb:
        m.coords = GT.perturb_interior_nodes(m, perturb, 9)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, order, tensor_size=(ncomp,) if ncomp > 1 else None,
                          dirichlet_boundary=GT.boundary(m) if diri else None)
    G = GT.boundary(m) if group is None else GT.boundary(m, group_names=[group])
    p = O.Problem(list(cells), simplexify=simplex).space(order, ncomp=ncomp, dirichlet=diri)
    p.set_coords(m.coords)
    B = Boundary(p, order, ncomp=ncomp, select=SELECT[group])
    assert len(B.faces) == len(G.faces)
    return m, Om, V, G, p, B


@pytest.mark.parametrize("case", CASES)
def test_boundary_loads(GT, case):
    cells, order, simplex, ncomp, diri, group, perturb = case
    print(f"case boundary load {case}", flush=True)
    m, Om, V, G, p, B = setup(GT, *case)
    degree = 2 * order
    dG = GT.measure(G, degree)
    g = [0.7, -1.2, 2.0][:ncomp]
    gf = GT.inputs.uniform(3, len(G.faces) * ncomp) - 0.5
    fn = GT.inputs.normal(4, V.nfree + V.ndiri)
    loads = [  # constant, per face, nodal (free then Dirichlet values)
        (GT.Load(dG, tuple(g) if ncomp > 1 else g[0]), 0, g),
        (GT.Load(dG, GT.face_constant_field(gf.reshape(len(G.faces), ncomp) if ncomp > 1 else gf, G)), 1, gf),
        (GT.Load(dG, GT.DiscreteField(V, fn[:V.nfree], fn[V.nfree:])), 2, fn)]
    rows_sets = [(GT.FREE, V.nfree)] + ([(GT.DIRICHLET, V.ndiri)] if V.ndiri else [])
    for rows, n in rows_sets:
        for form, kind, data in loads:
            b = GT.assemble_vector(form, float, V, rows)
            assert relfro(b, B.vector(degree, kind, data, rows=rows, n=n)) <= TOL, (rows, kind)


@pytest.mark.parametrize("case", [c for c in CASES if c[3] == 1])
@pytest.mark.parametrize("precision,tol", [("fp64", TOL), ("mixed", 1e-6)])
def test_boundary_mass(GT, case, precision, tol):
    cells, order, simplex, ncomp, diri, group, perturb = case
    print(f"case boundary mass {case} {precision}", flush=True)
    m, Om, V, G, p, B = setup(GT, *case)
    degree = 2 * order
    dG = GT.measure(G, degree)
    kappa = GT.inputs.lognormal(7, len(G.faces))
    prec = GT.FP64 if precision == "fp64" else GT.MIXED
    blocks = [(GT.FREE, GT.FREE)] + ([(GT.DIRICHLET, GT.DIRICHLET), (GT.FREE, GT.DIRICHLET)] if V.ndiri else [])
    size = {GT.FREE: V.nfree, GT.DIRICHLET: V.ndiri}
    for rows, cols in blocks:
        A = GT.assemble_matrix(GT.Mass(dG, GT.face_constant_field(kappa, G)), float, V, V, (rows, cols),
                               precision=prec)
        R = B.mass(degree, coef=kappa, rows=rows, cols=cols, shape=(size[rows], size[cols]))
        assert np.array_equal(A.colptr - 1, R.indptr) and np.array_equal(A.rowval - 1, R.indices)
        assert relfro(A.nzval, R.data) <= tol, (rows, cols)


@pytest.mark.parametrize("precision,tol", [("fp64", TOL), ("mixed", 1e-6)])
def test_boundary_mass_vector(GT, precision, tol):
    """Three-component P2 field on tetrahedron faces: the mass u . v couples
    equal components only; the pattern holds every component pair of a
    face (explicit zeros where components differ), so the values are
    compared as matrices."""
    case = CASES[5]
    print(f"case boundary mass vector {case} {precision}", flush=True)
    m, Om, V, G, p, B = setup(GT, *case)
    degree = 2 * case[1]
    dG = GT.measure(G, degree)
    kappa = GT.inputs.lognormal(8, len(G.faces))
    prec = GT.FP64 if precision == "fp64" else GT.MIXED
    A = GT.assemble_matrix(GT.Mass(dG, GT.face_constant_field(kappa, G)), float, V, V, precision=prec)
    R = B.mass(degree, coef=kappa, shape=(V.nfree, V.nfree)).toarray()
    Ad = A.to_scipy().toarray()
    assert np.linalg.norm(Ad - R) / np.linalg.norm(R) <= tol


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[4]])
def test_volume_plus_boundary(GT, case):
    """Laplace over the cells plus a Robin mass over boundary faces, added on
    the device into the volume block; and a load over cells plus one over
    the boundary."""
    cells, order, simplex, ncomp, diri, group, perturb = case
    print(f"case volume + boundary {case}", flush=True)
    m, Om, V, G, p, B = setup(GT, *case)
    degree = 2 * order
    dO, dG = GT.measure(Om, degree), GT.measure(G, degree)
    p.quadrature(degree)
    kappa = GT.inputs.lognormal(8, m.cells.shape[0])
    alpha = 3.5
    A = GT.assemble_matrix(GT.Laplace(dO, GT.face_constant_field(kappa, Om)) + GT.Mass(dG, alpha), float, V, V)
    cp, rv, nz = p.assemble_matrix(O.LAPLACE, coef=kappa)
    Rb = B.mass(degree, coef=np.full(len(G.faces), alpha), shape=(V.nfree, V.nfree)).tocoo()
    ref = nz.copy()
    for i, j, v in zip(Rb.row, Rb.col, Rb.data):  # boundary entries inside the volume pattern
        seg = slice(cp[j] - 1, cp[j + 1] - 1)
        k = np.searchsorted(rv[seg], i + 1)
        assert rv[seg][k] == i + 1
        ref[cp[j] - 1 + k] += v
    assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
    assert relfro(A.nzval, ref) <= TOL
    b = GT.assemble_vector(GT.Load(dO, 1.0) + GT.Load(dG, 2.0), float, V)
    bref = p.assemble_vector(0, [1.0]) + B.vector(degree, 0, [2.0], n=V.nfree)
    assert relfro(b, bref) <= TOL


def test_boundary_rejects(GT):
    m = GT.cartesian_mesh((0.0, 1.0) * 2, (3, 3))
    V = GT.lagrange_space(GT.interior(m), 1)
    dG = GT.measure(GT.boundary(m), 2)
    with pytest.raises(NotImplementedError):
        GT.assemble_matrix(GT.Laplace(dG), float, V, V)
    with pytest.raises(NotImplementedError):
        GT.assemble_matrix(GT.Mass(dG) + GT.Mass(dG), float, V, V)
