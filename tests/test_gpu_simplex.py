"""Scalar forms on simplex meshes (P1/P2 triangles and tetrahedra from the
reference's simplexified Cartesian meshes) on the GPU against the oracle,
through the C ABI: Laplace (with a lognormal kappa) and mass matrices in both
precisions, with and without Dirichlet rows, the FREE x DIRICHLET block, and
load vectors of the three source kinds.  Pattern bit-exact, FP64 values
within relative Frobenius error 1e-13, mixed within 1e-6; a tetrahedral
vertex has more contributors than a matrix-core batch holds, so those run on
the VALU kernel."""
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
