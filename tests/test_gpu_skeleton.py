"""Interior-penalty skeleton terms on the GPU (GalerkinToolkitExamples/src/
poisson.jl:99-102 over GT.skeleton) against the oracle (oracle/skeleton.py),
through the C ABI:
- discontinuous spaces: the DOF ids equal the oracle's (cell by cell, the
  Dirichlet tags of the face closures);
- the skeleton matrix alone on discontinuous spaces (quadrilaterals Q1-Q3,
  hexahedra Q1/Q2, perturbed): the pattern of the oracle's COO assembly bit
  for bit, FP64 values within 1e-13 and the mixed form within 1e-6
  (relative Frobenius, and entrywise relative to the largest face-block
  entry); on continuous spaces (every jump vanishes) the pattern of the
  oracle's COO assembly and values equal to rounding, and the reference's
  continuous Poisson check with skeleton terms (assembly_tests.jl:318-345);
- Laplace + skeleton + Nitsche as one matrix on a discontinuous space (the
  volume and boundary blocks added into the skeleton block on the device);
- update_matrix with moved nodes (the face diameters follow) equals a fresh
  assembly;
- cells listing their vertices rotated (non-identity point sets) give the
  matrix of the standard vertex order, up to the renumbering of their DOFs;
- the volume load on a discontinuous space, and the device solve of a
  positive definite interior-penalty system against a direct solve;
- the reference's patch tests (poisson_tests.jl:31-44): interior penalty
  with strong and with Nitsche conditions on (0,2)^2, Neumann on y = 2,
  Q2; the systems assembled on the device reproduce u = x + y to 1e-10
  (solved on the host: the interior-penalty matrix with the Examples'
  penalty is indefinite)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from nitsche import NitscheFaces
from skeleton import DGSpace, Skeleton

pytestmark = pytest.mark.gpu

TOL64, TOLMX = 1e-13, 1e-6


def setup(GT, O, cells, order, perturb, box=None, seed=11, simplex=False):
    D = len(cells)
    box = box or [0.0, 1.0] * D
    m = GT.cartesian_mesh(tuple(box), cells, simplexify=simplex)
    if perturb:
        m.coords = GT.perturb_interior_nodes(m, perturb, seed)
    p = O.Problem(list(cells), box=list(box), simplexify=simplex).space(order, dirichlet=0)
    p.set_coords(m.coords)
    p.quadrature(2 * order)
    return m, p


@pytest.mark.parametrize("cells,order", [((3, 2), 2), ((2, 2, 2), 1)])
def test_dg_numbering(GT, O, cells, order):
    print(f"case dg numbering {cells} Q{order}", flush=True)
    m, p = setup(GT, O, cells, order, 0.0)
    V = GT.lagrange_space(GT.interior(m), order, continuous=False)
    assert np.array_equal(V.cell_dofs, DGSpace(p, order).cd)
    D = len(cells)
    G = GT.boundary(m)
    Vd = GT.lagrange_space(GT.interior(m), order, dirichlet_boundary=G, continuous=False)
    on_bnd = lambda y: bool(np.any(np.abs(y) < 1e-12) or np.any(np.abs(y - 1.0) < 1e-12))
    ref = DGSpace(p, order, dirichlet=on_bnd)
    assert np.array_equal(Vd.cell_dofs, ref.cd) and (Vd.nfree, Vd.ndiri) == (ref.nfree, ref.ndiri)


@pytest.mark.parametrize("cells,order,perturb,simplex", [
    ((4, 3), 1, 0.2, False), ((3, 3), 2, 0.2, False), ((3, 2), 3, 0.15, False), ((2, 3, 2), 1, 0.15, False),
    ((2, 2, 2), 2, 0.1, False), ((4, 3), 1, 0.2, True), ((3, 3), 2, 0.2, True), ((2, 3, 2), 1, 0.1, True),
    ((2, 2, 2), 2, 0.1, True)])
def test_skeleton_matrix(GT, O, cells, order, perturb, simplex):
    print(f"case skeleton matrix {cells} order {order} simplex={simplex}", flush=True)
    m, p = setup(GT, O, cells, order, perturb, simplex=simplex)
    V = GT.lagrange_space(GT.interior(m), order, continuous=False)
    cd = DGSpace(p, order).cd
    assert np.array_equal(V.cell_dofs, cd)
    degree = 2 * order
    gamma = 2.5
    dL = GT.measure(GT.skeleton(m), degree)
    sk = Skeleton(p, order, cd)
    n = V.nfree
    ref = sk.matrix(degree, gamma, shape=(n, n))
    scale = max(np.abs(sk.face_blocks(f, degree, gamma)[1]).max() for f in sk.geo.interior)
    for prec, tol in ((GT.FP64, TOL64), (GT.MIXED, TOLMX)):
        A = GT.assemble_matrix(GT.InteriorPenalty(dL, gamma), float, V, V, precision=prec)
        assert np.array_equal(A.colptr - 1, ref.indptr) and np.array_equal(A.rowval - 1, ref.indices)
        err = np.abs(A.nzval - ref.data).max() / scale
        assert err <= tol, (prec, err)
        assert np.linalg.norm(A.nzval - ref.data) <= tol * np.linalg.norm(ref.data), prec


@pytest.mark.parametrize("cells,order,simplex", [((3, 3), 1, False), ((3, 2), 2, False), ((2, 2, 2), 1, False),
                                                  ((3, 3), 1, True)])
def test_continuous_space_skeleton(GT, O, cells, order, simplex):
    """Skeleton terms on a continuous space (the reference's own check,
    test/assembly_tests.jl:318-345): the pattern couples both cells of every
    interior face (bit-exact against the oracle's COO assembly) and the
    values, sums of the cells' copies whose jumps cancel, agree with the
    oracle's to rounding of the face blocks."""
    print(f"case continuous skeleton {cells} order {order} simplex={simplex}", flush=True)
    m, p = setup(GT, O, cells, order, 0.15, simplex=simplex)
    V = GT.lagrange_space(GT.interior(m), order)
    assert np.array_equal(V.cell_dofs, p.cell_dofs())
    dL = GT.measure(GT.skeleton(m), 2 * order)
    sk = Skeleton(p, order, p.cell_dofs())
    ref = sk.matrix(2 * order, 2.0, shape=(V.nfree, V.nfree))
    scale = max(np.abs(sk.face_blocks(f, 2 * order, 2.0)[1]).max() for f in sk.geo.interior)
    for prec, tol in ((GT.FP64, TOL64), (GT.MIXED, TOLMX)):
        A = GT.assemble_matrix(GT.InteriorPenalty(dL, 2.0), float, V, V, precision=prec)
        assert np.array_equal(A.colptr - 1, ref.indptr) and np.array_equal(A.rowval - 1, ref.indices)
        assert np.abs(A.nzval - ref.data).max() <= tol * scale, prec


def test_reference_skeleton_poisson(GT, O):
    """test/assembly_tests.jl:318-345: Q1 on (0,1)^2 in 4x4 cells, Laplace
    plus the interior-penalty terms (gamma 1, degree 2), u = x + y on the
    whole boundary, no source: the device-assembled system reproduces u."""
    print("case reference skeleton poisson", flush=True)
    m = GT.cartesian_mesh((0.0, 1.0, 0.0, 1.0), (4, 4))
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, 1, dirichlet_boundary=GT.boundary(m))
    dO, dL = GT.measure(Om, 2), GT.measure(GT.skeleton(m), 2)
    e = np.indices([2, 2]).reshape(2, -1)[::-1].T.astype(float)
    vals = np.zeros(V.nfree + V.ndiri)
    for c in range(m.cells.shape[0]):
        X = GT.refcell.linear_shape(m.cellkind, e) @ m.coords[m.cells[c]]
        for a, d in enumerate(V.cell_dofs[c]):
            vals[d - 1 if d > 0 else V.nfree - d - 1] = X[a].sum()
    n = V.nfree
    uhd = GT.DiscreteField(V, np.zeros(n), vals[n:])
    A, b = GT.linear_problem(uhd, GT.Laplace(dO) + GT.InteriorPenalty(dL, 1.0), GT.Load(dO, 0.0))
    x = sla.spsolve(A.to_scipy().tocsc(), b)
    assert np.abs(x - vals[:n]).max() < 1e-10


@pytest.mark.parametrize("cells,order", [((3, 3), 2), ((2, 2, 2), 1)])
def test_dg_laplace_skeleton_nitsche(GT, O, cells, order):
    print(f"case dg laplace + skeleton + nitsche {cells} Q{order}", flush=True)
    m, p = setup(GT, O, cells, order, 0.15)
    V = GT.lagrange_space(GT.interior(m), order, continuous=False)
    dg = DGSpace(p, order)
    degree = 2 * order
    gamma = 3.0
    dO = GT.measure(GT.interior(m), degree)
    dL = GT.measure(GT.skeleton(m), degree)
    dG = GT.measure(GT.boundary(m), degree)
    n = V.nfree
    nf = NitscheFaces(p, order)
    nf.cd = dg.cd
    ref = (dg.laplace(degree) + Skeleton(p, order, dg.cd).matrix(degree, gamma, shape=(n, n))
           + nf.matrix(degree, gamma, shape=(n, n))).tocsc()
    ref.sort_indices()
    for prec, tol in ((GT.FP64, TOL64), (GT.MIXED, TOLMX)):
        A = GT.assemble_matrix(GT.Laplace(dO) + GT.InteriorPenalty(dL, gamma) + GT.Nitsche(dG, gamma), float, V, V,
                               precision=prec)
        assert np.array_equal(A.colptr - 1, ref.indptr) and np.array_equal(A.rowval - 1, ref.indices)
        assert np.linalg.norm(A.nzval - ref.data) <= tol * np.linalg.norm(ref.data), prec


@pytest.mark.parametrize("method,degree", [("strong", 2), ("nitsche", 4)])
def test_dg_patch(GT, O, method, degree):
    """poisson_tests.jl:31-44 with discretization_method = :interior_penalty:
    (0,2)^2 in 2x2 cells, Q2, u = x + y, Dirichlet on 1-face-1, -3, -4,
    Neumann on 1-face-2 (y = 2, g = 1); gamma = degree (degree + 1) / 10.
    degree 2 is the Examples' own; with Nitsche conditions that rule leaves
    the system singular (a quadratic jump vanishing at both Gauss points of
    every face), so that case uses 4."""
    print(f"case dg patch {method} degree {degree}", flush=True)
    m, p = setup(GT, O, (2, 2), 2, 0.0, box=[0.0, 2.0, 0.0, 2.0])
    Om = GT.interior(m)
    Gd = GT.boundary(m, ["1-face-1", "1-face-3", "1-face-4"])
    Gn = GT.boundary(m, ["1-face-2"])
    gamma = degree * (degree + 1) / 10.0
    dO, dL = GT.measure(Om, degree), GT.measure(GT.skeleton(m), degree)
    dGd, dGn = GT.measure(Gd, degree), GT.measure(Gn, degree)
    u = lambda y: float(y[0] + y[1])
    if method == "strong":
        V = GT.lagrange_space(Om, 2, dirichlet_boundary=Gd, continuous=False)
    else:
        V = GT.lagrange_space(Om, 2, continuous=False)
    dg = DGSpace(p, 2)
    pos = dg.geo.cell_pos
    vals = np.zeros(V.nfree + V.ndiri)
    for c in range(V.cell_dofs.shape[0]):
        for a in range(V.cell_dofs.shape[1]):
            d = V.cell_dofs[c, a]
            vals[d - 1 if d > 0 else V.nfree - d - 1] = u(pos[c, a])
    n = V.nfree
    a_form = GT.Laplace(dO) + GT.InteriorPenalty(dL, gamma)
    l_form = GT.Load(dGn, 1.0)
    if method == "strong":
        uhd = GT.DiscreteField(V, np.zeros(n), vals[n:])
        A, b = GT.linear_problem(uhd, a_form, l_form)
    else:
        g = GT.DiscreteField(V, vals[:n])
        A = GT.assemble_matrix(a_form + GT.Nitsche(dGd, gamma), float, V, V)
        b = GT.assemble_vector(l_form + GT.NitscheLoad(dGd, g, gamma), float, V)
    x = sla.spsolve(A.to_scipy().tocsc(), b)
    assert np.abs(x - vals[:n]).max() < 1e-10 * np.abs(vals).max()


@pytest.mark.parametrize("cells,order", [((4, 3), 2), ((2, 3, 2), 1)])
def test_dg_load_and_spd_solve(GT, O, cells, order):
    """The volume load on a discontinuous space against the oracle, and the
    device solve (Jacobi-PCG inside FP64 refinement) of a symmetric
    interior-penalty system with a large penalty (positive definite)
    against a direct solve of the same matrix."""
    print(f"case dg load + solve {cells} Q{order}", flush=True)
    m, p = setup(GT, O, cells, order, 0.15)
    V = GT.lagrange_space(GT.interior(m), order, continuous=False)
    dg = DGSpace(p, order)
    degree = 2 * order
    dO = GT.measure(GT.interior(m), degree)
    f = lambda y: 1.0 + y[0] - 2.0 * y[1]
    # a nodal f of the space: exact for this linear f
    pos = dg.geo.cell_pos
    fv = np.zeros(V.nfree)
    for c in range(V.cell_dofs.shape[0]):
        for a in range(V.cell_dofs.shape[1]):
            fv[V.cell_dofs[c, a] - 1] = f(pos[c, a])
    b = GT.assemble_vector(GT.Load(dO, GT.DiscreteField(V, fv)), float, V)
    bref = dg.load(degree, f)
    assert np.linalg.norm(b - bref) <= TOL64 * np.linalg.norm(bref)
    gamma = 20.0 * order * (order + 1)
    dL = GT.measure(GT.skeleton(m), degree)
    dG = GT.measure(GT.boundary(m), degree)
    A = GT.assemble_matrix(GT.Laplace(dO) + GT.InteriorPenalty(dL, gamma) + GT.Nitsche(dG, gamma), float, V, V)
    xd = sla.spsolve(A.to_scipy().tocsc(), b)
    for prec in (GT.FP64, GT.MIXED):
        x, st = GT.solve(A, b, precision=prec, rtol=1e-11)
        assert st.converged and np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd), prec


def test_dg_update_coordinates(GT, O):
    """update_matrix with new node coordinates refills Laplace + skeleton +
    Nitsche with the moved geometry (the face diameters too): equal to a
    fresh assembly on a mesh with those coordinates."""
    print("case dg update coordinates", flush=True)
    cells, order = (4, 3), 2
    m, _ = setup(GT, O, cells, order, 0.1, seed=3)
    m2, _ = setup(GT, O, cells, order, 0.2, seed=5)
    forms = []
    for mesh in (m, m2):
        V = GT.lagrange_space(GT.interior(mesh), order, continuous=False)
        dO = GT.measure(GT.interior(mesh), 4)
        dL = GT.measure(GT.skeleton(mesh), 4)
        dG = GT.measure(GT.boundary(mesh), 4)
        forms.append((V, GT.Laplace(dO) + GT.InteriorPenalty(dL, 3.0) + GT.Nitsche(dG, 3.0)))
    (V1, f1), (V2, f2) = forms
    A, cache = GT.assemble_matrix(f1, float, V1, V1, reuse=True)
    GT.update_matrix(A, cache, coords=m2.coords)
    B = GT.assemble_matrix(f2, float, V2, V2)
    assert np.array_equal(A.colptr, B.colptr) and np.array_equal(A.rowval, B.rowval)
    assert np.linalg.norm(A.nzval - B.nzval) <= 1e-14 * np.linalg.norm(B.nzval)


def test_rotated_cells_match_standard(GT, O):
    """Skeleton items whose two local faces list the shared vertices in
    different orders (the second cell's vertices rotated, a non-identity
    point set): the matrix equals that of the same geometry with standard
    vertex order, up to the renumbering of that cell's DOFs (matched by
    position), FP64 within 1e-13."""
    from gtk_amd.mesh import Mesh
    R = GT.refcell
    print("case skeleton rotated cells", flush=True)
    nid = lambda i, j, k: i + 3 * (j + 2 * k)
    bits2 = [(n & 1, (n >> 1) & 1) for n in range(4)]
    bits3 = [(n & 1, (n >> 1) & 1, (n >> 2) & 1) for n in range(8)]
    c2 = np.array([[i, j] for j in range(2) for i in range(3)], float)
    c3 = np.array([[i, j, k] for k in range(2) for j in range(2) for i in range(3)], float)
    cases = [
        (2, c2, [[a + 3 * b for a, b in bits2], [1 + a + 3 * b for a, b in bits2]],
         [[a + 3 * b for a, b in bits2], [(2 - a) + 3 * (1 - b) for a, b in bits2]], R.QUAD),
        (3, c3, [[nid(a, b, c) for a, b, c in bits3], [nid(1 + a, b, c) for a, b, c in bits3]],
         [[nid(a, b, c) for a, b, c in bits3], [nid(1 + a, 1 - c, b) for a, b, c in bits3]], R.HEX)]
    for D, coords, std, rot, kind in cases:
        mats, pos = [], []
        for cells in (std, rot):
            m = Mesh(D, coords, np.array(cells), kind, {}, {})
            V = GT.lagrange_space(GT.interior(m), 2, continuous=False)
            dL = GT.measure(GT.skeleton(m), 4)
            A = GT.assemble_matrix(GT.InteriorPenalty(dL, 2.0), float, V, V).to_scipy().toarray()
            e = np.indices([3] * D).reshape(D, -1)[::-1].T / 2.0  # Q2 nodes, x fastest
            P = np.zeros((V.nfree, D))
            for c in range(2):
                P[V.cell_dofs[c] - 1] = R.linear_shape(kind, e) @ m.coords[m.cells[c]]
            mats.append(A)
            pos.append(P)
        # DOF k of the standard numbering sits where DOF perm[k] of the rotated one does
        n = len(pos[0])
        perm = np.array([int(np.argmin(np.linalg.norm(pos[1] - pos[0][k], axis=1)
                                       + 10.0 * (np.arange(n) // (n // 2) != k // (n // 2)))) for k in range(n)])
        assert sorted(perm) == list(range(n))
        B = mats[1][np.ix_(perm, perm)]
        assert np.abs(B - mats[0]).max() <= 1e-13 * np.abs(mats[0]).max(), D


@pytest.mark.parametrize("cells,order", [((4, 3), 1), ((3, 3), 2), ((2, 2, 2), 1)])
def test_dg_simplex_patch_strong(GT, O, cells, order):
    """Interior penalty on triangles and tetrahedra, strong Dirichlet
    conditions on the whole boundary, assembled and lifted on the device
    (linear_problem): a linear u is reproduced to 1e-10."""
    print(f"case dg simplex patch {cells} P{order}", flush=True)
    D = len(cells)
    m, p = setup(GT, O, cells, order, 0.15, simplex=True)
    Om = GT.interior(m)
    V = GT.lagrange_space(Om, order, dirichlet_boundary=GT.boundary(m), continuous=False)
    u = lambda y: 1.0 + 2.0 * y[0] - y[1] + (0.5 * y[2] if D == 3 else 0.0)
    on_bnd = lambda y: bool(np.any(np.abs(y) < 1e-12) or np.any(np.abs(y - 1.0) < 1e-12))
    dg = DGSpace(p, order, dirichlet=on_bnd)
    assert np.array_equal(V.cell_dofs, dg.cd)
    pos = dg.geo.cell_pos
    vals = np.zeros(V.nfree + V.ndiri)
    for c in range(V.cell_dofs.shape[0]):
        for a in range(V.cell_dofs.shape[1]):
            d = V.cell_dofs[c, a]
            vals[d - 1 if d > 0 else V.nfree - d - 1] = u(pos[c, a])
    n = V.nfree
    degree = 2 * order
    dO, dL = GT.measure(Om, degree), GT.measure(GT.skeleton(m), degree)
    uhd = GT.DiscreteField(V, np.zeros(n), vals[n:])
    A, b = GT.linear_problem(uhd, GT.Laplace(dO) + GT.InteriorPenalty(dL, 10.0), GT.Load(dO, 0.0))
    x = sla.spsolve(A.to_scipy().tocsc(), b)
    assert np.abs(x - vals[:n]).max() < 1e-10 * np.abs(vals).max()
