"""The boundary-integral oracle (oracle/boundary.py) pinned by what the
mathematics fixes: the boundary measure (perimeter / surface area, also of
one group of faces), partition of unity, the 1D mass stencil of a Q1 edge,
symmetry, M 1 = b(1), b(f nodal) = M f, and exactness for a linear g."""
import numpy as np
import pytest
from boundary import Boundary

import oracle as O


def problem(cells, order, ncomp=1, dirichlet=0, simplex=False, perturb=0.0, GT=None, seed=5):
    p = O.Problem(list(cells), simplexify=simplex).space(order, ncomp=ncomp, dirichlet=dirichlet)
    if perturb:
        m = GT.cartesian_mesh(tuple([0.0, 1.0] * len(cells)), cells, simplexify=simplex)
        p.set_coords(GT.perturb_interior_nodes(m, perturb, seed))
    return p


@pytest.mark.parametrize("cells,order,simplex,area", [((4, 3), 1, False, 4.0), ((3, 2), 3, False, 4.0),
                                                      ((2, 2, 3), 2, False, 6.0), ((2, 2, 2), 2, True, 6.0),
                                                      ((3, 4), 2, True, 4.0)])
def test_measure_and_partition_of_unity(GT, cells, order, simplex, area):
    p = problem(cells, order, simplex=simplex, perturb=0.2, GT=GT)
    B = Boundary(p, order)
    n = p.q(O.Q_NFREE)
    b = B.vector(2 * order, 0, [1.0], n=n)
    assert abs(b.sum() - area) < 1e-13
    M = B.mass(2 * order, shape=(n, n))
    assert abs(M.sum() - area) < 1e-13
    assert abs(M - M.T).max() < 1e-15
    assert np.allclose(M @ np.ones(n), b, atol=1e-15)


def test_group_of_faces(GT):
    """The x = 0 face of the unit cube: area 1, and a linear g = 1 + y + 2z
    integrates exactly (Q1 interpolates it)."""
    p = problem((3, 2, 2), 1, perturb=0.2, GT=GT)
    B = Boundary(p, 1, select=lambda x: abs(x[0]) < 1e-12)
    n = p.q(O.Q_NFREE)
    assert abs(B.vector(2, 0, [1.0], n=n).sum() - 1.0) < 1e-14
    x = p.coords()
    f = 1.0 + x[:, 1] + 2.0 * x[:, 2]  # Q1: DOF = vertex, numbered as the vertices
    nv = p.node_vertex()  # 1-based vertex id of each node
    fd = np.zeros(n)
    fd[nv - 1] = f
    b = B.vector(2, 2, fd, n=n)
    assert abs(b.sum() - (1.0 + 0.5 + 1.0)) < 1e-14  # int_0^1 int_0^1 (1 + y + 2z)
    M = B.mass(2, shape=(n, n))
    assert np.allclose(b, M @ fd, atol=1e-15)


def test_q1_edge_stencil():
    """Q1 on the x = 0 edge of a 4 x 3 grid: the 1D mass matrix h/6 [2 1; 1 2]."""
    p = problem((4, 3), 1)
    B = Boundary(p, 1, select=lambda x: abs(x[0]) < 1e-12)
    n = p.q(O.Q_NFREE)
    M = B.mass(2, shape=(n, n)).toarray()
    x = p.coords()
    nv = p.node_vertex()  # 1-based vertex id = DOF id for Q1 without Dirichlet
    edge = sorted((i for i in range(len(x)) if abs(x[i, 0]) < 1e-12), key=lambda i: x[i, 1])
    ids = [int(nv[i]) - 1 for i in edge]
    h = 1.0 / 3.0
    sub = M[np.ix_(ids, ids)]
    ref = np.diag([2 * h / 6] + [4 * h / 6] * 2 + [2 * h / 6]) + np.diag([h / 6] * 3, 1) + np.diag([h / 6] * 3, -1)
    assert np.allclose(sub, ref, atol=1e-15)
    assert abs(M.sum() - 1.0) < 1e-14


def test_vector_components():
    """P2 vector tetrahedra: a constant traction integrates to area x value
    per component over the free rows of a space with no Dirichlet DOFs."""
    p = problem((2, 2, 2), 2, ncomp=3, simplex=True)
    B = Boundary(p, 2, ncomp=3, select=lambda x: abs(x[2] - 1.0) < 1e-12)
    n = p.q(O.Q_NFREE)
    b = B.vector(4, 0, [0.5, -1.0, 2.0], n=n)
    cd = p.cell_dofs()
    comp = np.zeros(n, np.int64)
    for c in range(3):
        comp[np.abs(cd[:, c::3].ravel()) - 1] = c
    for c, v in enumerate([0.5, -1.0, 2.0]):
        assert abs(b[comp == c].sum() - v) < 1e-14
