"""GPU parity of libgtk_hip against the CPU oracle (through the C ABI).

Pattern (colptr/rowval) must be bit-exact; FP64 values within relative
Frobenius error 1e-13; mixed precision within 1e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-13
TOL_MIXED = 1e-6


def build(GT, O, cells, order, simplex=False, diri=0, group=0, ncomp=1, perturb=0.0, seed=7):
    D = len(cells)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells, simplexify=simplex)
    if perturb:
        m.coords = GT.perturb_interior_nodes(m, perturb, seed)
    Om = GT.interior(m)
    db = None if diri == 0 else (GT.boundary(m) if diri == 1 else GT.boundary(m, group_names=[f"{D-1}-face-{group}"]))
    V = GT.lagrange_space(Om, order, tensor_size=(ncomp,) if ncomp > 1 else None, dirichlet_boundary=db)
    p = O.Problem(cells, simplexify=simplex).space(order, ncomp=ncomp, dirichlet=diri, group=group)
    p.set_coords(m.coords)
    return m, Om, V, p


def relfro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


CASES = [
    # cells, order, degree, diri, perturb
    ((8, 8), 1, 2, 1, 0.0),
    ((6, 7), 2, 4, 0, 0.2),
    ((5, 6), 3, 7, 1, 0.2),
    ((4, 5, 3), 1, 2, 0, 0.2),
    ((4, 4, 4), 2, 4, 1, 0.2),
    ((3, 5, 4), 2, 5, 0, 0.0),
]


@pytest.mark.parametrize("cells,order,degree,diri,perturb", CASES)
@pytest.mark.parametrize("kappa", [False, True])
def test_laplace_matrix_fp64(GT, O, cells, order, degree, diri, perturb, kappa):
    print(f"case laplace fp64 cells={cells} order={order} degree={degree} diri={diri} kappa={kappa}", flush=True)
    m, Om, V, p = build(GT, O, cells, order, diri=diri, perturb=perturb)
    dO = GT.measure(Om, degree)
    p.quadrature(degree)
    ncell = m.cells.shape[0]
    k = GT.inputs.lognormal(11, ncell) if kappa else None
    form = GT.Laplace(dO, GT.face_constant_field(k, Om) if kappa else None)
    A = GT.assemble_matrix(form, float, V, V)
    cp, rv, nz = p.assemble_matrix(O.LAPLACE, coef=k)
    assert np.array_equal(A.colptr, cp)
    assert np.array_equal(A.rowval, rv)
    assert relfro(A.nzval, nz) <= TOL_FP64


@pytest.mark.parametrize("cells,order,degree,diri,perturb", CASES[:5])
def test_laplace_matrix_mixed(GT, O, cells, order, degree, diri, perturb):
    print(f"case laplace mixed cells={cells} order={order}", flush=True)
    m, Om, V, p = build(GT, O, cells, order, diri=diri, perturb=perturb)
    dO = GT.measure(Om, degree)
    p.quadrature(degree)
    A = GT.assemble_matrix(GT.Laplace(dO), float, V, V, precision=GT.MIXED)
    cp, rv, nz = p.assemble_matrix(O.LAPLACE)
    assert np.array_equal(A.rowval, rv)
    assert relfro(A.nzval, nz) <= TOL_MIXED


@pytest.mark.parametrize("cells,order,degree,diri,perturb", [CASES[1], CASES[3], CASES[4], CASES[5]])
@pytest.mark.parametrize("spread", [0, 6])
def test_mixed_kappa_and_mass(GT, O, cells, order, degree, diri, perturb, spread):
    """Mixed form with a per-cell coefficient, lognormal times 10^U(-s, s)
    (s = 6 spans twelve decades, so the per-cell scaling of the FP16 split
    product on Q2 hexahedra sees very different cells side by side), for the
    Laplace and the mass matrix."""
    print(f"case mixed kappa/mass cells={cells} order={order} spread={spread}", flush=True)
    m, Om, V, p = build(GT, O, cells, order, diri=diri, perturb=perturb)
    dO = GT.measure(Om, degree)
    p.quadrature(degree)
    ncell = m.cells.shape[0]
    k = GT.inputs.lognormal(13, ncell) * 10.0 ** (spread * (2 * GT.inputs.uniform(14, ncell) - 1))
    for code, form in ((O.LAPLACE, GT.Laplace(dO, GT.face_constant_field(k, Om))),
                       (O.MASS, GT.Mass(dO, GT.face_constant_field(k, Om)))):
        A = GT.assemble_matrix(form, float, V, V, precision=GT.MIXED)
        cp, rv, nz = p.assemble_matrix(code, coef=k)
        assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
        assert relfro(A.nzval, nz) <= TOL_MIXED, code


@pytest.mark.parametrize("cells,order,degree,diri,perturb", CASES[:5])
def test_mass_and_load(GT, O, cells, order, degree, diri, perturb):
    print(f"case mass/load cells={cells} order={order}", flush=True)
    m, Om, V, p = build(GT, O, cells, order, diri=diri, perturb=perturb)
    dO = GT.measure(Om, degree)
    p.quadrature(degree)
    A = GT.assemble_matrix(GT.Mass(dO), float, V, V)
    cp, rv, nz = p.assemble_matrix(O.MASS)
    assert np.array_equal(A.rowval, rv)
    assert relfro(A.nzval, nz) <= TOL_FP64
    b = GT.assemble_vector(GT.Load(dO, 1.0), float, V)
    bo = p.assemble_vector(0, [1.0])
    assert relfro(b, bo) <= TOL_FP64


def test_dirichlet_columns(GT, O):
    print("case free x dirichlet block", flush=True)
    m, Om, V, p = build(GT, O, (6, 5), 2, diri=1, perturb=0.1)
    dO = GT.measure(Om, 4)
    p.quadrature(4)
    A, Ad = GT.assemble_matrix_with_free_and_dirichlet_columns(GT.Laplace(dO), float, V, V)
    cp, rv, nz = p.assemble_matrix(O.LAPLACE, rows=O.FREE, cols=O.DIRICHLET)
    assert np.array_equal(Ad.colptr, cp) and np.array_equal(Ad.rowval, rv)
    assert relfro(Ad.nzval, nz) <= TOL_FP64


@pytest.mark.parametrize("cells,valu_waves", [((3, 2, 3), None), ((8, 6, 7), None), ((8, 6, 7), "4")])
@pytest.mark.parametrize("precision,tol", [("fp64", TOL_FP64), ("mixed", TOL_MIXED)])
def test_elasticity_p2_tets(GT, O, precision, tol, cells, valu_waves, monkeypatch):
    """P2 vector tetrahedra, per-cell Lame parameters, clamped on x = 0
    (group 2-face-5), body force (0, 0, -1); the larger mesh spans several
    tiles; GTK_VALU_WAVES=4 runs the 4-wave workgroup the library falls back
    to when long columns leave no LDS for the preferred one."""
    if valu_waves:
        monkeypatch.setenv("GTK_VALU_WAVES", valu_waves)
    print(f"case elasticity p2 tets {precision} {cells} valu_waves={valu_waves}", flush=True)
    m, Om, V, p = build(GT, O, cells, 2, simplex=True, diri=2, group=5, ncomp=3, perturb=0.1)
    dO = GT.measure(Om, 5)
    p.quadrature(5)
    nc = m.cells.shape[0]
    lam = 0.5 + 1.5 * GT.inputs.uniform(31, nc)
    mu = 0.5 + 1.5 * GT.inputs.uniform(32, nc)
    form = GT.Elasticity(dO, GT.face_constant_field(lam, Om), GT.face_constant_field(mu, Om))
    A = GT.assemble_matrix(form, float, V, V, precision=GT.FP64 if precision == "fp64" else GT.MIXED)
    cp, rv, nz = p.assemble_matrix(O.ELASTICITY, coef=np.stack([lam, mu], axis=1).ravel())
    assert np.array_equal(A.colptr, cp) and np.array_equal(A.rowval, rv)
    assert relfro(A.nzval, nz) <= tol
    b = GT.assemble_vector(GT.Load(dO, (0.0, 0.0, -1.0)), float, V)
    bo = p.assemble_vector(0, [0.0, 0.0, -1.0])
    assert relfro(b, bo) <= TOL_FP64
