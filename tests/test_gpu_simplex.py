"""Scalar forms on simplex meshes (P1/P2 triangles and tetrahedra from the
reference's simplexified Cartesian meshes) on the GPU against the oracle,
through the C ABI: Laplace (with a lognormal kappa) and mass matrices in both
precisions, with and without Dirichlet rows, the FREE x DIRICHLET block, and
load vectors of the three source kinds.  Pattern bit-exact, FP64 values
within relative Frobenius error 1e-13, mixed within 1e-6; a tetrahedral
vertex has more contributors than a matrix-core batch holds, so those run on
the VALU kernel.  Also linear elasticity on the element kinds besides the
P2 tetrahedra of the benchmark (Q1/Q2 quadrilaterals, Q1 hexahedra, P1/P2
triangles, P1 tetrahedra)."""
import numpy as np
import pytest

from test_gpu_parity import TOL_FP64, TOL_MIXED, build, relfro

pytestmark = pytest.mark.gpu

CASES = [((6, 5), 1, 0.2, 1), ((5, 4), 2, 0.2, 0), ((3, 3, 2), 1, 0.15, 1), ((2, 3, 2), 2, 0.1, 2)]


@pytest.mark.parametrize("cells,order,perturb,diri", CASES)
def test_simplex_laplace_and_mass(GT, O, cells, order, perturb, diri):
    print(f"case simplex cells={cells} P{order} diri={diri}", flush=True)
    m, Om, V, p = build(GT, O, cells, order, simplex=True, diri=diri, group=1, perturb=perturb)
    degree = 2 * order
    dO = GT.measure(Om, degree)
    p.quadrature(degree)
    k = GT.inputs.lognormal(3, m.cells.shape[0])
    for form, code, coef in ((GT.Laplace(dO, GT.face_constant_field(k, Om)), O.LAPLACE, k),
                             (GT.Mass(dO), O.MASS, None)):
        cp, rv, nz = p.assemble_matrix(code, coef=coef)
        for prec, tol in ((GT.FP64, TOL_FP64), (GT.MIXED, TOL_MIXED)):
            A = GT.assemble_matrix(form, float, V, V, precision=prec)
            assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
            assert relfro(A.nzval, nz) <= tol, (code, prec)
        if V.ndiri:
            cp, rv, nz = p.assemble_matrix(code, coef=coef, rows=O.FREE, cols=O.DIRICHLET)
            A = GT.assemble_matrix(form, float, V, V, free_or_dirichlet=(GT.FREE, GT.DIRICHLET))
            assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
            assert relfro(A.nzval, nz) <= TOL_FP64


@pytest.mark.parametrize("cells,order,perturb,diri", CASES)
def test_simplex_loads(GT, O, cells, order, perturb, diri):
    print(f"case simplex loads cells={cells} P{order}", flush=True)
    m, Om, V, p = build(GT, O, cells, order, simplex=True, diri=diri, group=1, perturb=perturb)
    degree = 2 * order
    dO = GT.measure(Om, degree)
    p.quadrature(degree)
    fn = GT.inputs.normal(8, V.nfree + V.ndiri)
    fc = GT.inputs.uniform(9, m.cells.shape[0]) - 0.5
    for load, (kind, data) in ((GT.Load(dO, 2.0), (0, [2.0])),
                               (GT.Load(dO, GT.face_constant_field(fc, Om)), (1, fc)),
                               (GT.Load(dO, GT.DiscreteField(V, fn[:V.nfree], fn[V.nfree:])), (2, fn))):
        b = GT.assemble_vector(load, float, V)
        assert relfro(b, p.assemble_vector(kind, data)) <= TOL_FP64, kind


# vector elasticity on the other element kinds: cells, order, simplex, perturb
ECASES = [((5, 4), 1, False, 0.2), ((4, 3), 2, False, 0.15), ((3, 2, 3), 1, False, 0.15),
          ((5, 4), 1, True, 0.2), ((4, 3), 2, True, 0.15), ((2, 3, 2), 1, True, 0.1)]


@pytest.mark.parametrize("cells,order,simplex,perturb", ECASES)
def test_elasticity_other_elements(GT, O, cells, order, simplex, perturb):
    """sigma(eps(u)) : eps(v) with per-cell Lame parameters on Q1/Q2
    quadrilaterals, Q1 hexahedra, P1/P2 triangles and P1 tetrahedra (vector
    fields, the VALU kernel), clamped on the x = 0 side, and the body-force
    load, against the oracle."""
    print(f"case elasticity cells={cells} order={order} simplex={simplex}", flush=True)
    D = len(cells)
    group = 3 if D == 2 else 5  # the x = 0 side
    m, Om, V, p = build(GT, O, cells, order, simplex=simplex, diri=2, group=group, ncomp=D, perturb=perturb)
    degree = 2 * order
    dO = GT.measure(Om, degree)
    p.quadrature(degree)
    nc = m.cells.shape[0]
    lam = 0.5 + 1.5 * GT.inputs.uniform(41, nc)
    mu = 0.5 + 1.5 * GT.inputs.uniform(42, nc)
    form = GT.Elasticity(dO, GT.face_constant_field(lam, Om), GT.face_constant_field(mu, Om))
    cp, rv, nz = p.assemble_matrix(O.ELASTICITY, coef=np.stack([lam, mu], axis=1).ravel())
    for prec, tol in ((GT.FP64, TOL_FP64), (GT.MIXED, TOL_MIXED)):
        A = GT.assemble_matrix(form, float, V, V, precision=prec)
        assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
        assert relfro(A.nzval, nz) <= tol, prec
    f = [0.0] * (D - 1) + [-1.0]
    b = GT.assemble_vector(GT.Load(dO, tuple(f)), float, V)
    assert relfro(b, p.assemble_vector(0, f)) <= TOL_FP64
