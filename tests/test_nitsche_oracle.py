"""Pins of the Nitsche oracle (oracle/nitsche.py) and of the reference
normals the kernels use, on the CPU.

- The outward normals of the reference cells' local faces equal the lists of
  GalerkinToolkit.jl src/domain.jl:226 (square), :258 (cube), :428
  (triangle), :460 (tetrahedron).
- Patch tests, the reference's own check of the method
  (GalerkinToolkitExamples/test/poisson_tests.jl:31-44: exact solutions to
  1e-10): Laplace plus the Nitsche terms on the whole boundary, with the
  volume load -Lap(u) and the Nitsche load of g = u, reproduce u at every DOF
  for linear u on perturbed meshes (Q1, Q2, quadrilaterals and hexahedra)
  and quadratic u on affine ones.  A missing or sign-flipped consistency term
  -v n.grad(u) breaks them.
- The matrix is symmetric and equals NL + NL^T - beta M_Gamma, NL the load's
  matrix and M_Gamma the boundary mass with beta = gamma / h per face (the
  boundary mass is pinned by tests/test_boundary_oracle.py): a flipped
  symmetric term, a wrong normal orientation or a wrong penalty fails it."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from boundary import Boundary
from nitsche import NitscheFaces


def test_reference_normals_match_reference_lists(GT):
    R = GT.refcell
    n1 = 1 / np.sqrt(2)
    n2 = 1 / np.sqrt(3)
    lists = {R.QUAD: [(0, -1), (0, 1), (-1, 0), (1, 0)],
             R.HEX: [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)],
             R.TRI: [(0, -1), (-1, 0), (n1, n1)],
             R.TET: [(0, 0, -1), (0, -1, 0), (-1, 0, 0), (n2, n2, n2)]}
    for kind, ref in lists.items():
        assert np.allclose(R.face_normals(kind), np.asarray(ref, np.float64), atol=1e-15), kind


def problem(GT, O, cells, order, box, perturb, seed=7):
    D = len(cells)
    m = GT.cartesian_mesh(tuple(box), cells)
    x = GT.perturb_interior_nodes(m, perturb, seed) if perturb else m.coords
    p = O.Problem(list(cells), box=list(box)).space(order, dirichlet=0)
    p.set_coords(x)
    p.quadrature(2 * order)
    return m, x, p


def dof_positions(nf):
    """Physical position of every free DOF (all DOFs: no Dirichlet set)."""
    pos = np.zeros((nf.nfree, nf.D))
    for c in range(nf.cd.shape[0]):
        for a, d in enumerate(nf.cd[c]):
            pos[d - 1] = nf.cell_pos[c, a]
    return pos


@pytest.mark.parametrize("cells,order,perturb,kind", [
    ((3, 3), 2, 0.2, "linear"), ((3, 4), 1, 0.2, "linear"), ((3, 3), 2, 0.0, "quadratic"),
    ((2, 2, 2), 1, 0.15, "linear"), ((2, 2, 2), 2, 0.0, "quadratic")])
def test_patch_exact(GT, O, cells, order, perturb, kind):
    import scipy.sparse as sp
    D = len(cells)
    box = [0.0, 2.0] * D
    m, x, p = problem(GT, O, cells, order, box, perturb)
    if kind == "linear":
        u = lambda y: 1.0 + 2.0 * y[0] - y[1] + (0.5 * y[2] if D == 3 else 0.0)
        f = 0.0
    else:
        u = lambda y: 1.0 + y[0] - y[1] + y[0] ** 2 - 0.5 * y[0] * y[1] + 0.25 * y[1] ** 2
        f = -2.5  # -Lap(u)
    degree = 2 * order
    gamma = 10.0 * order * (order + 1)
    nf = NitscheFaces(p, order)
    n = nf.nfree
    cp, rv, nz = p.assemble_matrix(O.LAPLACE)
    K = sp.csc_matrix((nz, rv - 1, cp - 1), shape=(n, n))
    N = nf.matrix(degree, gamma, shape=(n, n))
    b = p.assemble_vector(0, [f]) + nf.load(degree, gamma, u, n=n)
    xs = sla.spsolve((K + N).tocsc(), b)
    ue = np.array([u(y) for y in dof_positions(nf)])
    assert np.abs(xs - ue).max() < 1e-10 * max(1.0, np.abs(ue).max())


@pytest.mark.parametrize("cells,order,perturb", [((3, 3), 2, 0.2), ((2, 3, 2), 1, 0.15), ((2, 2, 2), 2, 0.1)])
def test_symmetry_and_penalty(GT, O, cells, order, perturb):
    D = len(cells)
    m, x, p = problem(GT, O, cells, order, [0.0, 1.0] * D, perturb)
    degree = 2 * order
    gamma = 3.0
    nf = NitscheFaces(p, order)
    n = nf.nfree
    N = nf.matrix(degree, gamma, shape=(n, n)).toarray()
    NL = nf.matrix(degree, gamma, shape=(n, n), load=True).toarray()
    scale = np.abs(N).max()
    assert np.abs(N - N.T).max() <= 1e-13 * scale
    bd = Boundary(p, order)
    h = np.array([nf.face_points(fc, degree)[2] for fc in bd.faces])
    M = bd.mass(degree, coef=gamma / h, shape=(n, n)).toarray()
    assert np.abs(N - (NL + NL.T - M)).max() <= 1e-13 * scale
    # the consistency term is present: N differs from its penalty part
    assert np.abs(N - M).max() > 1e-3 * scale
