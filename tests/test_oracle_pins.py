"""Pins of the CPU oracle (oracle/gtk_oracle.c) against what the
mathematics and the reference's own tests fix.  Each check is chosen so a
plausible slip (dropped term, wrong sign/index, transposed J) fails it."""
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from helpers import csc, dof_coordinates

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


# --- quadrature --------------------------------------------------------------
def test_gauss_legendre_closed_forms(O):
    x, w = O.gauss_legendre(1)
    assert np.allclose(x, [0.5]) and np.allclose(w, [1.0])
    x, w = O.gauss_legendre(2)
    s = 0.5 / math.sqrt(3)
    assert np.allclose(x, [0.5 - s, 0.5 + s], atol=1e-15, rtol=0) and np.allclose(w, [0.5, 0.5], atol=1e-15)
    x, w = O.gauss_legendre(3)
    s = 0.5 * math.sqrt(3 / 5)
    assert np.allclose(x, [0.5 - s, 0.5, 0.5 + s], atol=1e-15, rtol=0)
    assert np.allclose(w, [5 / 18, 8 / 18, 5 / 18], atol=1e-15, rtol=0)
    x, w = O.gauss_legendre(4)
    a = math.sqrt(3 / 7 - 2 / 7 * math.sqrt(6 / 5))
    b = math.sqrt(3 / 7 + 2 / 7 * math.sqrt(6 / 5))
    wa = (18 + math.sqrt(30)) / 36
    wb = (18 - math.sqrt(30)) / 36
    assert np.allclose(x, 0.5 + 0.5 * np.array([-b, -a, a, b]), atol=1e-15, rtol=0)
    assert np.allclose(w, 0.5 * np.array([wb, wa, wa, wb]), atol=1e-15, rtol=0)


@pytest.mark.parametrize("n", range(1, 9))
def test_gauss_legendre_exactness(O, n):
    x, w = O.gauss_legendre(n)
    for k in range(2 * n):
        assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-14


@pytest.mark.parametrize("D", [2, 3])
def test_weight_sums_cubes(O, D):
    # integration_tests.jl:27-34
    p = O.Problem([2] * D)
    p.space(1)
    x, w = p.quadrature(GOLD["gauss_weight_sums"]["degree"])
    assert abs(w.sum() - GOLD["gauss_weight_sums"]["sum"]) < 1e-14


def _simplex_moment(expo):
    # int over the unit simplex of prod x_i^a_i = prod a_i! / (sum a + D)!
    D = len(expo)
    return np.prod([math.factorial(a) for a in expo]) / math.factorial(sum(expo) + D)


@pytest.mark.parametrize("D,degree", [(3, 4), (3, 5), (2, 2), (2, 3), (2, 5), (3, 6)])
def test_simplex_rules_exact(O, D, degree):
    p = O.Problem([1] * D, simplexify=True)
    p.space(1)
    x, w = p.quadrature(degree)
    assert abs(w.sum() - 1.0 / math.factorial(D)) < 1e-13
    for e in np.ndindex(*([degree + 1] * D)):
        if sum(e) > degree:
            continue
        v = np.dot(w, np.prod(x ** np.asarray(e), axis=1))
        assert abs(v - _simplex_moment(e)) < 1e-12, (e, v)


# --- reference elements ---------------------------------------------------------
@pytest.mark.parametrize("cells,order,simplex", [((1, 1), 1, False), ((1, 1), 2, False), ((1, 1), 3, False),
                                                 ((1, 1, 1), 2, False), ((1, 1, 1), 2, True), ((1, 1), 3, True)])
def test_tabulator_identity_and_partition_of_unity(GT, O, cells, order, simplex):
    # space_tests.jl:219-220: the tabulator at the nodes is the identity
    p = O.Problem(list(cells), simplexify=simplex).space(order)
    fe = GT.refcell.Lagrange(GT.refcell.TET if simplex and len(cells) == 3 else
                             GT.refcell.TRI if simplex else
                             (GT.refcell.HEX if len(cells) == 3 else GT.refcell.QUAD), order)
    val, grad = p.tabulate(fe.nodes)
    assert np.allclose(val, np.eye(fe.nnodes), atol=1e-12)
    rng = np.random.default_rng(0)
    xi = rng.uniform(0, 1, (7, len(cells)))
    if simplex:
        xi = xi / (1.2 * xi.sum(axis=1, keepdims=True))
    val, grad = p.tabulate(xi)
    assert np.allclose(val.sum(axis=1), 1.0, atol=1e-12)
    assert np.allclose(grad.sum(axis=1), 0.0, atol=1e-10)
    # gradients against central finite differences
    h = 1e-6
    for d in range(len(cells)):
        e = np.zeros(len(cells))
        e[d] = h
        vp, _ = p.tabulate(xi + e)
        vm, _ = p.tabulate(xi - e)
        assert np.allclose((vp - vm) / (2 * h), grad[:, :, d], atol=1e-7)


def test_reference_face_nodes_q1_square(GT):
    fe = GT.refcell.Lagrange(GT.refcell.QUAD, 1)
    g = GOLD["q1_square_face_nodes"]
    for d in range(3):
        assert fe.face_nodes[d].tolist() == g[str(d)]
    gi = GOLD["q1_square_face_interior_nodes"]
    for d in range(3):
        assert fe.own_nodes[d].tolist() == gi[str(d)]


# --- numbering --------------------------------------------------------------------
def test_q2_single_cell_counts(O):
    g = GOLD["q2_unit_square_dirichlet_counts"]
    p = O.Problem([1, 1]).space(2, dirichlet=1)
    assert p.q(O.Q_NFREE) == g["num_free_dofs"] and p.q(O.Q_NDIRI) == g["num_dirichlet_dofs"]


@pytest.mark.parametrize("cells,order", [((5, 4), 1), ((4, 6), 2), ((3, 5), 3), ((3, 4, 2), 1), ((3, 2, 4), 2)])
def test_dof_counts_and_nnz_closed_forms(O, cells, order):
    D = len(cells)
    p = O.Problem(list(cells)).space(order)
    assert p.q(O.Q_NDOFS) == np.prod([order * n + 1 for n in cells])
    p.quadrature(2 * order)
    cp, rv, nz = p.assemble_matrix(O.LAPLACE)
    # nnz = prod over directions of the 1D count (k^2+2k) n + 1
    assert len(rv) == np.prod([(order * order + 2 * order) * n + 1 for n in cells])
    p = O.Problem(list(cells)).space(order, dirichlet=1)
    assert p.q(O.Q_NFREE) == np.prod([order * n - 1 for n in cells])


def test_p2_tet_nnz_closed_form(O):
    """P2 on the 6-tet split of n^3 cubes: nnz = 230 n^3 + 138 n^2 + 24 n + 1
    at five sizes (a cubic through four points, checked at a fifth)."""
    for n in (2, 3, 4, 5, 6):
        p = O.Problem([n, n, n], simplexify=True).space(2)
        p.quadrature(2)
        cp, rv, nz = p.assemble_matrix(O.MASS)
        assert len(rv) == 230 * n ** 3 + 138 * n ** 2 + 24 * n + 1, n


def test_q1_numbering_closed_form_2d(O):
    # free id of interior node (i, j), 2 <= i, j <= n: (j-2)(n-1) + (i-1)
    n = 64
    p = O.Problem([n, n]).space(1, dirichlet=1)
    cn = p.cell_nodes()
    cd = p.cell_dofs()
    node = cn - 1
    i = node % (n + 1) + 1
    j = node // (n + 1) + 1
    interior = (i >= 2) & (i <= n) & (j >= 2) & (j <= n)
    assert np.array_equal(cd[interior], ((j - 2) * (n - 1) + (i - 1))[interior])
    assert np.all(cd[~interior] < 0)
    assert p.q(O.Q_NFREE) == 3969 and p.q(O.Q_NDIRI) == 256


def test_q1_numbering_closed_form_3d(O):
    # without Dirichlet, DOF = vertex id: the 8 box corners first in
    # lexicographic order, then the other nodes in node order
    n = 12
    p = O.Problem([n, n, n]).space(1)
    N = n + 1
    corners = [a + b * N + c * N * N for c in (0, n) for b in (0, n) for a in (0, n)]
    expect = np.zeros(N ** 3, np.int64)
    expect[corners] = np.arange(1, 9)
    rest = np.setdiff1d(np.arange(N ** 3), corners)
    expect[rest] = np.arange(9, N ** 3 + 1)
    cd = p.cell_dofs()
    assert np.array_equal(cd, expect[p.cell_nodes() - 1])


@pytest.mark.parametrize("cells", [(3, 4, 2), (4, 4, 4)])
def test_topology_invariants(O, cells):
    p = O.Problem(list(cells))
    nx, ny, nz = cells
    V, E, F, C = p.q(O.Q_NVERT), p.q(O.Q_NEDGES), p.q(O.Q_NFACES2), p.q(O.Q_NCELLS)
    assert V - E + F - C == 1  # Euler characteristic of a box
    assert E == nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    cf = p.cell_faces(2)
    cnt = np.bincount(cf.ravel(), minlength=F + 1)[1:]
    nb = 2 * (nx * ny + ny * nz + nx * nz)
    assert np.sum(cnt == 1) == nb and np.sum(cnt == 2) == F - nb
    # parent (boundary) faces keep ids 1..nb
    assert np.all(cnt[:nb] == 1)
    # every edge joins two grid-adjacent vertices
    fv = p.face_vertices(1)
    vn = np.zeros(V + 1, np.int64)
    nv = p.node_vertex()
    vn[nv] = np.arange(len(nv))
    x = p.coords()[vn[fv[:, 0]]] - p.coords()[vn[fv[:, 1]]]
    h = np.array([1 / nx, 1 / ny, 1 / nz])
    assert np.allclose(np.sort(np.abs(x) / h, axis=1), [0, 0, 1])


# --- assembled values ------------------------------------------------------------------
def test_unit_square_known_answers(O):
    g = GOLD["unit_square_2x2_q1"]
    p = O.Problem([2, 2]).space(1)
    p.quadrature(2)
    b = p.assemble_vector(0, [1.0])
    assert abs(b.sum() - g["sum_load_vector_f1"]) < 1e-14
    cp, rv, nz = p.assemble_matrix(O.MASS)
    assert abs(nz.sum() - g["sum_mass_matrix"]) < 1e-14


def test_square_0_2_known_answers(O):
    g = GOLD["square_0_2_8x8"]
    p = O.Problem([8, 8], box=[0.0, 2.0, 0.0, 2.0]).space(1)
    p.quadrature(2)
    b = p.assemble_vector(0, [1.0])
    assert abs(b.sum() - g["area"]) < 1e-13
    # integral of sum(x) = sum_i (sum x)_i b_i for the P1-exact field
    x = p.coords()
    nv = p.node_vertex()
    # Q1 DOF = vertex id, so nodal values follow the vertex numbering
    u = np.zeros(len(nv))
    u[nv - 1] = x.sum(axis=1)
    assert abs(np.dot(u, b) - g["integral_sum_x"]) < 1e-12


def test_q1_stencils_closed_form(O):
    # 2D: diag 8/3, the 8 neighbours -1/3; 3D (h): diag 8h/3, face 0, edge -h/6, corner -h/12
    p = O.Problem([6, 6]).space(1)
    p.quadrature(2)
    A = csc(*p.assemble_matrix(O.LAPLACE), p.q(O.Q_NFREE)).toarray()
    nv = p.node_vertex()
    node = 3 + 3 * 7  # interior node (3, 3)
    row = A[nv[node] - 1]
    assert abs(row[nv[node] - 1] - 8 / 3) < 1e-14
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            if di or dj:
                assert abs(row[nv[node + di + 7 * dj] - 1] + 1 / 3) < 1e-14
    n = 4
    h = 1.0 / n
    p = O.Problem([n] * 3).space(1)
    p.quadrature(2)
    A = csc(*p.assemble_matrix(O.LAPLACE), p.q(O.Q_NFREE)).toarray()
    nv = p.node_vertex()
    N = n + 1
    c = 2 + 2 * N + 2 * N * N
    r = A[nv[c] - 1]
    assert abs(r[nv[c] - 1] - 8 * h / 3) < 1e-14
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                k = abs(dx) + abs(dy) + abs(dz)
                if k == 0:
                    continue
                v = r[nv[c + dx + dy * N + dz * N * N] - 1]
                assert abs(v - {1: 0.0, 2: -h / 6, 3: -h / 12}[k]) < 1e-14


@pytest.mark.parametrize("cells,order,perturb", [((4, 5), 1, 0.2), ((4, 3), 2, 0.2), ((3, 4), 3, 0.15),
                                                  ((3, 2, 3), 1, 0.2), ((2, 3, 2), 2, 0.2)])
def test_laplace_row_sums_and_symmetry(GT, O, cells, order, perturb):
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * len(cells)), cells)
    p = O.Problem(list(cells)).space(order)
    p.set_coords(GT.perturb_interior_nodes(m, perturb, 3))
    p.quadrature(2 * order)
    A = csc(*p.assemble_matrix(O.LAPLACE), p.q(O.Q_NFREE))
    assert np.max(np.abs(A.sum(axis=1))) < 1e-12
    assert abs(A - A.T).max() < 1e-14


@pytest.mark.parametrize("cells,order,perturb,simplex", [((4, 5), 1, 0.2, False), ((3, 4), 2, 0.2, False),
                                                          ((3, 3), 3, 0.15, False), ((3, 2, 3), 1, 0.2, False),
                                                          ((2, 3, 2), 2, 0.2, False), ((2, 2, 2), 2, 0.0, True),
                                                          ((3, 3), 2, 0.1, True), ((4, 3), 1, 0.2, True),
                                                          ((2, 3, 2), 1, 0.15, True), ((2, 2, 2), 2, 0.15, True)])
def test_patch_linear(GT, O, cells, order, perturb, simplex):
    """u = sum(x) is reproduced exactly (assembly_tests.jl:75-118)."""
    tol = GOLD["patch_test_tolerance"]["tol"]
    D = len(cells)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells, simplexify=simplex)
    x = GT.perturb_interior_nodes(m, perturb, 9) if perturb else m.coords
    m.coords = x
    V = GT.lagrange_space(GT.interior(m), order, dirichlet_boundary=GT.boundary(m))
    p = O.Problem(list(cells), simplexify=simplex).space(order, dirichlet=1)
    p.set_coords(x)
    p.quadrature(2 * order)
    nf, nd = p.q(O.Q_NFREE), p.q(O.Q_NDIRI)
    A = csc(*p.assemble_matrix(O.LAPLACE), nf)
    Ad = csc(*p.assemble_matrix(O.LAPLACE, rows=O.FREE, cols=O.DIRICHLET), nf)
    xf, xd = dof_coordinates(GT, m, V)
    u = spla.spsolve(A.tocsc(), -Ad @ xd.sum(axis=1))
    assert np.max(np.abs(u - xf.sum(axis=1))) < tol


@pytest.mark.parametrize("cells,order,simplex", [((3, 4), 2, False), ((3, 3), 3, False), ((2, 2, 3), 2, False),
                                                  ((3, 4), 2, True), ((2, 2, 2), 2, True)])
def test_patch_quadratic(GT, O, cells, order, simplex):
    """u = |x|^2 with f = -Lap u = -2D is exact for Q2, Q3 on affine meshes
    and for P2 on triangles and tetrahedra (affine cells)."""
    D = len(cells)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells, simplexify=simplex)
    V = GT.lagrange_space(GT.interior(m), order, dirichlet_boundary=GT.boundary(m))
    p = O.Problem(list(cells), simplexify=simplex).space(order, dirichlet=1)
    p.quadrature(2 * order)
    nf = p.q(O.Q_NFREE)
    A = csc(*p.assemble_matrix(O.LAPLACE), nf)
    Ad = csc(*p.assemble_matrix(O.LAPLACE, rows=O.FREE, cols=O.DIRICHLET), nf)
    b = p.assemble_vector(0, [-2.0 * D])
    xf, xd = dof_coordinates(GT, m, V)
    u = spla.spsolve(A.tocsc(), b - Ad @ (xd ** 2).sum(axis=1))
    assert np.max(np.abs(u - (xf ** 2).sum(axis=1))) < 1e-10


@pytest.mark.parametrize("cells,order", [((4, 3), 1), ((3, 2), 2), ((2, 3, 2), 1), ((2, 2, 2), 2)])
def test_simplex_mass_and_load(GT, O, cells, order):
    """On perturbed simplex meshes (each cell affine): sum(M) = |Omega| and
    sum(b(f=1)) = |Omega| without Dirichlet rows (problems_tests.jl:55-79),
    b(f nodal) = M f, and each cell's mass total is its volume."""
    D = len(cells)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells, simplexify=True)
    x = GT.perturb_interior_nodes(m, 0.2, 5)
    p = O.Problem(list(cells), simplexify=True).space(order)
    p.set_coords(x)
    p.quadrature(2 * order)
    nf = p.q(O.Q_NFREE)
    M = csc(*p.assemble_matrix(O.MASS), nf)
    assert abs(M.sum() - 1.0) < 1e-13
    assert abs(p.assemble_vector(0, [1.0]).sum() - 1.0) < 1e-13
    f = np.random.default_rng(3).standard_normal(nf)
    assert np.allclose(p.assemble_vector(2, f), M @ f, atol=1e-14, rtol=1e-12)
    # a per-cell coefficient that is 1 on one cell only: its volume
    nc = p.q(O.Q_NCELLS)
    xm = x[m.cells]
    e = xm[:, 1:] - xm[:, :1]
    vol = np.abs(np.linalg.det(e)) / (2.0 if D == 2 else 6.0)
    for c in (0, nc // 2, nc - 1):
        k = np.zeros(nc)
        k[c] = 1.0
        assert abs(csc(*p.assemble_matrix(O.MASS, coef=k), nf).sum() - vol[c]) < 1e-14


def test_mass_and_load_consistency(GT, O):
    """sum(M) = |Omega|; b(f nodal) = M f; b(f=1) = M 1."""
    cells = (3, 4, 2)
    m = GT.cartesian_mesh((0.0, 1.0) * 3, cells)
    x = GT.perturb_interior_nodes(m, 0.2, 4)
    p = O.Problem(list(cells)).space(2)
    p.set_coords(x)
    p.quadrature(5)
    nf = p.q(O.Q_NFREE)
    M = csc(*p.assemble_matrix(O.MASS), nf)
    assert abs(M.sum() - 1.0) < 1e-13
    f = np.random.default_rng(1).standard_normal(nf)
    b = p.assemble_vector(2, f)
    assert np.allclose(b, M @ f, atol=1e-14, rtol=1e-12)
    kap = np.random.default_rng(2).uniform(0.5, 2, p.q(O.Q_NCELLS))
    Mk = csc(*p.assemble_matrix(O.MASS, coef=kap), nf)
    b1 = p.assemble_vector(1, kap)  # per-cell f = kappa
    assert np.allclose(b1, Mk @ np.ones(nf), atol=1e-14, rtol=1e-12)


@pytest.mark.parametrize("cells,simplex", [((2, 2, 2), True), ((2, 3, 2), False)])
def test_elasticity_rigid_modes(GT, O, cells, simplex):
    """Rigid motions are in the kernel and K is symmetric (P2 tets, Q2
    hexes); the lambda/mu law itself is pinned by the uniform-strain energy
    and the linear patch solve below."""
    D = 3
    m = GT.cartesian_mesh((0.0, 1.0) * 3, cells, simplexify=simplex)
    x = GT.perturb_interior_nodes(m, 0.15, 6)
    m.coords = x
    V = GT.lagrange_space(GT.interior(m), 2, tensor_size=(3,))
    p = O.Problem(list(cells), simplexify=simplex).space(2, ncomp=3)
    p.set_coords(x)
    p.quadrature(5 if simplex else 4)
    nc = p.q(O.Q_NCELLS)
    lm = np.stack([np.full(nc, 1.3), np.full(nc, 0.7)], axis=1).ravel()
    nf = p.q(O.Q_NFREE)
    K = csc(*p.assemble_matrix(O.ELASTICITY, coef=lm), nf)
    xf, _ = dof_coordinates(GT, m, V)
    comp = np.tile(np.arange(3), nf // 3)
    # dof k is component comp[k] of its node (component-major)
    for r in [np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0])]:
        assert np.max(np.abs(K @ r[comp])) < 1e-12
    for (i, j) in [(0, 1), (1, 2), (0, 2)]:
        # infinitesimal rotation: u_i = -x_j, u_j = x_i
        u = np.zeros(nf)
        u[comp == i] = -xf[comp == i, j]
        u[comp == j] = xf[comp == j, i]
        assert np.max(np.abs(K @ u)) < 1e-11
    assert abs(K - K.T).max() < 1e-13


def _vector_dofs(GT, m, V):
    """Coordinates and component of each free and Dirichlet DOF of a
    D-component space (component-major inside each node)."""
    D = m.coords.shape[1]
    xf, xd = dof_coordinates(GT, m, V)
    return xf, np.tile(np.arange(D), V.nfree // D), xd, np.tile(np.arange(D), V.ndiri // D)


def _simplex_volumes(m):
    X = m.coords[m.cells]
    return np.abs(np.linalg.det(X[:, 1:] - X[:, :1])) / math.factorial(X.shape[2])


# (cells, order, simplex): P2 tetrahedra and Q2 hexahedra (the benchmark's
# kinds) and the other element kinds of the elasticity kernels
ELASTIC_KINDS = [((2, 2, 2), 2, True), ((2, 3, 2), 2, False), ((2, 2, 2), 1, True), ((2, 3, 2), 1, False),
                 ((3, 2), 1, True), ((3, 2), 2, True), ((3, 4), 1, False), ((2, 3), 2, False)]


def _elastic_degree(D, order, simplex):
    return 5 if (simplex and D == 3 and order == 2) else 2 * order


@pytest.mark.parametrize("cells,order,simplex", ELASTIC_KINDS)
def test_elasticity_uniform_strain_energy(GT, O, cells, order, simplex):
    """u = E x (E not symmetric) has the uniform strain eps = (E + E^T)/2 and
    the energy u^T K u = sum_c |Omega_c| (lambda_c (tr eps)^2 + 2 mu_c eps:eps)
    (sigma = lambda tr(eps) I + 2 mu eps, docs example_linear_elasticity.jl:
    86-95; in 2D the same law, plane strain).  lambda != mu per cell on
    simplices (cell volumes exact), constant lambda = 0.6, mu = 1.7 on
    quadrilaterals and hexahedra (total volume 1): a swap of lambda and mu, a
    factor on mu or a missing skew cancellation fails."""
    D = len(cells)
    m = GT.cartesian_mesh((0.0, 1.0) * D, cells, simplexify=simplex)
    m.coords = GT.perturb_interior_nodes(m, 0.15, 8)
    V = GT.lagrange_space(GT.interior(m), order, tensor_size=(D,))
    p = O.Problem(list(cells), simplexify=simplex).space(order, ncomp=D)
    p.set_coords(m.coords)
    p.quadrature(_elastic_degree(D, order, simplex))
    nc = p.q(O.Q_NCELLS)
    rng = np.random.default_rng(9)
    if simplex:
        lam, mu, vol = rng.uniform(0.5, 2.0, nc), rng.uniform(0.5, 2.0, nc), _simplex_volumes(m)
    else:
        lam, mu, vol = np.full(nc, 0.6), np.full(nc, 1.7), None
    K = csc(*p.assemble_matrix(O.ELASTICITY, coef=np.stack([lam, mu], axis=1).ravel()), p.q(O.Q_NFREE))
    xf, comp, _, _ = _vector_dofs(GT, m, V)
    E = np.array([[0.3, -1.1, 0.4], [0.7, -0.2, 0.9], [-0.5, 0.25, 1.3]])[:D, :D]
    u = np.einsum("kj,kj->k", E[comp], xf)
    eps = 0.5 * (E + E.T)
    per_volume = lambda l, k: l * np.trace(eps) ** 2 + 2.0 * k * np.sum(eps * eps)
    want = np.sum(vol * per_volume(lam, mu)) if simplex else per_volume(0.6, 1.7) * 1.0
    assert abs(u @ (K @ u) - want) <= 1e-11 * want
    # the same with lambda and mu exchanged differs (the check can tell them apart)
    swapped = np.sum(vol * per_volume(mu, lam)) if simplex else per_volume(1.7, 0.6)
    assert abs(swapped - want) > 1e-3 * want


@pytest.mark.parametrize("cells,order,simplex", ELASTIC_KINDS)
def test_elasticity_linear_patch_solve(GT, O, cells, order, simplex):
    """Constant lambda != mu, no body force, u = E x + c prescribed on the
    whole boundary: the discrete solution at the free DOFs is E x + c
    exactly (every kind reproduces linear fields; tolerance of the
    reference's patch tests, test/assembly_tests.jl:112-118)."""
    D = len(cells)
    cells = tuple(max(c, 2) for c in cells)  # at least one free node
    m = GT.cartesian_mesh((0.0, 1.0) * D, cells, simplexify=simplex)
    m.coords = GT.perturb_interior_nodes(m, 0.15, 12)
    V = GT.lagrange_space(GT.interior(m), order, tensor_size=(D,), dirichlet_boundary=GT.boundary(m))
    p = O.Problem(list(cells), simplexify=simplex).space(order, ncomp=D, dirichlet=1)
    p.set_coords(m.coords)
    p.quadrature(_elastic_degree(D, order, simplex))
    nc = p.q(O.Q_NCELLS)
    lm = np.stack([np.full(nc, 0.8), np.full(nc, 1.9)], axis=1).ravel()
    nf = p.q(O.Q_NFREE)
    assert nf > 0
    K = csc(*p.assemble_matrix(O.ELASTICITY, coef=lm), nf)
    Kd = csc(*p.assemble_matrix(O.ELASTICITY, coef=lm, rows=O.FREE, cols=O.DIRICHLET), nf)
    xf, cf, xd, cd = _vector_dofs(GT, m, V)
    E = np.array([[0.4, 0.1, -0.3], [-0.6, 0.2, 0.5], [0.9, -0.7, 0.1]])[:D, :D]
    c = np.array([0.2, -0.1, 0.05])[:D]
    g = lambda x, k: np.einsum("kj,kj->k", E[k], x) + c[k]
    u = spla.spsolve(K.tocsc(), -(Kd @ g(xd, cd)))
    assert np.max(np.abs(u - g(xf, cf))) < 1e-10
