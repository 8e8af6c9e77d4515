"""The host mirror's numbering (numpy, vectorised) against the oracle's
literal restatement of complexify + generate_dof_ids: vertex ids, d-face ids,
face vertex order, permutation ids and the signed cell -> DOF map must be
identical."""
import numpy as np
import pytest

CASES = [
    ((4, 4), 1, False, 0, 0, 1), ((4, 4), 2, False, 0, 0, 1), ((4, 5), 3, False, 0, 0, 1),
    ((5, 5), 3, False, 1, 0, 1), ((3, 3, 3), 1, False, 0, 0, 1), ((3, 4, 2), 2, False, 0, 0, 1),
    ((4, 4, 4), 2, False, 1, 0, 1), ((2, 2), 2, True, 0, 0, 1), ((3, 3), 3, True, 1, 0, 1),
    ((3, 3, 3), 2, True, 2, 5, 3), ((2, 3, 2), 1, True, 0, 0, 1), ((3, 2, 4), 2, False, 2, 5, 1),
    ((1, 1), 2, False, 1, 0, 1), ((1, 1, 1), 2, False, 1, 0, 1), ((2, 2, 2), 2, False, 0, 0, 3),
]


@pytest.mark.parametrize("cells,order,simplex,diri,group,ncomp", CASES)
def test_numbering_matches_oracle(GT, O, cells, order, simplex, diri, group, ncomp):
    D = len(cells)
    p = O.Problem(list(cells), simplexify=simplex).space(order, ncomp=ncomp, dirichlet=diri, group=group)
    m = GT.cartesian_mesh(tuple([0.0, 1.0] * D), cells, simplexify=simplex)
    db = None if diri == 0 else (GT.boundary(m) if diri == 1 else GT.boundary(m, group_names=[f"{D-1}-face-{group}"]))
    V = GT.lagrange_space(GT.interior(m), order, tensor_size=(ncomp,) if ncomp > 1 else None, dirichlet_boundary=db)
    assert np.array_equal(m.node_vertex + 1, p.node_vertex())
    assert np.array_equal(m.cell_nodes_1b(), p.cell_nodes())
    assert np.allclose(m.coords, p.coords(), rtol=0, atol=0)
    for d in range(1, D):
        assert np.array_equal(m.cell_faces[d] + 1, p.cell_faces(d))
        assert np.array_equal(m.face_vertices[d] + 1, p.face_vertices(d))
        assert np.array_equal(m.cell_perm[d] + 1, p.cell_perm(d))
    for d in range(D):
        nodes, grp = p.parent_faces(d)
        assert np.array_equal(m.parent_faces[d] + 1, nodes)
        assert np.array_equal(m.parent_groups[d], grp)
    assert V.nfree == p.q(O.Q_NFREE) and V.ndiri == p.q(O.Q_NDIRI)
    assert np.array_equal(V.cell_dofs, p.cell_dofs())


def test_full_size_q2_counts(GT):
    """The 64^3 Q2 workload: DOF count and the structure of the numbering
    (vertices, then edges, faces, cells) at full size."""
    m = GT.cartesian_mesh((0.0, 1.0) * 3, (64, 64, 64))
    V = GT.lagrange_space(GT.interior(m), 2)
    assert V.nfree == 129 ** 3 == 2146689
    assert m.num_faces(0) == 65 ** 3
    assert m.num_faces(1) == 3 * 64 * 65 * 65
    assert m.num_faces(2) == 3 * 65 * 64 * 64
    cd = V.cell_dofs
    # vertex DOFs are the vertex ids; the cell-interior DOF is the last block
    assert np.array_equal(cd[:, [0, 2, 6, 8, 18, 20, 24, 26]] - 1, m.node_vertex[m.cells])
    nv, ne, nf = m.num_faces(0), m.num_faces(1), m.num_faces(2)
    assert np.array_equal(cd[:, 13], nv + ne + nf + np.arange(m.cells.shape[0]) + 1)
