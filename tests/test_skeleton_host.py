"""Host side of the skeleton items (assembly.skeleton_items), on the CPU:
each interior face's point sets put the face rule's points of both sides at
the same physical points, in the same order, and the two sides' outward
normals there are opposite, on Cartesian meshes and on meshes whose second
cell lists its vertices rotated (so that the two local faces list the shared
vertices in different orders: a non-identity vertex order must appear)."""
import numpy as np
import pytest


def check_items(GT, m, order):
    R = GT.refcell
    V = GT.lagrange_space(GT.interior(m), order, continuous=False)
    dL = GT.measure(GT.skeleton(m), 2 * order)
    it = GT.assembly.skeleton_items(V, dL)
    g = R.GEO[m.cellkind]
    normals = R.face_normals(m.cellkind)
    worst = 0.0
    for i, (c0, c1) in enumerate(it["cells"]):
        s0, s1 = it["item_sets"][i]
        pts, nrm = [], []
        for c, s in ((c0, s0), (c1, s1)):
            X = m.coords[m.cells[c]]
            xi = it["points"][s]
            pts.append(R.linear_shape(m.cellkind, xi) @ X)
            _, dN = R.Lagrange(m.cellkind, 1).tabulate(xi)
            n = []
            for q in range(len(xi)):
                J = X.T @ dN[q]
                v = np.linalg.solve(J.T, normals[int(it["keys"][s][0])])
                n.append(v / np.linalg.norm(v))
            nrm.append(np.array(n))
        worst = max(worst, np.abs(pts[0] - pts[1]).max())
        assert np.abs(nrm[0] + nrm[1]).max() < 1e-12, i
    assert worst < 1e-14
    return it


@pytest.mark.parametrize("cells,order", [((4, 3), 2), ((2, 3, 2), 1), ((2, 2, 2), 2)])
def test_cartesian(GT, cells, order):
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * len(cells)), cells)
    m.coords = GT.perturb_interior_nodes(m, 0.15, 9)
    check_items(GT, m, order)


def test_rotated_cells(GT):
    from gtk_amd.mesh import Mesh
    R = GT.refcell
    # 2D: the second square's vertices listed rotated by 180 degrees
    coords = np.array([[0, 0], [1, 0], [2, 0], [0, 1], [1, 1], [2, 1]], float)
    m = Mesh(2, coords, np.array([[0, 1, 3, 4], [5, 4, 2, 1]]), R.QUAD, {}, {})
    it = check_items(GT, m, 2)
    assert any(list(k[1:]) != sorted(k[1:]) for k in it["keys"])
    # 3D: the second cube's local vertex (a, b, c) at node (1 + a, 1 - c, b),
    # a quarter turn about x, the normal of the shared face
    nid = lambda i, j, k: i + 3 * (j + 2 * k)
    coords = np.array([[i, j, k] for k in range(2) for j in range(2) for i in range(3)], float)
    bits = [(n & 1, (n >> 1) & 1, (n >> 2) & 1) for n in range(8)]
    ca = [nid(a, b, c) for a, b, c in bits]
    cb = [nid(1 + a, 1 - c, b) for a, b, c in bits]
    m3 = Mesh(3, coords, np.array([ca, cb]), R.HEX, {}, {})
    it3 = check_items(GT, m3, 2)
    assert any(list(k[1:]) != sorted(k[1:]) for k in it3["keys"])
