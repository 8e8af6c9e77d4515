"""Shared helpers of the test-suite (mesh-independent math only)."""
import numpy as np


def dof_coordinates(GT, mesh, space):
    """Physical coordinates of every DOF node, indexed like the reference's
    signed DOF ids: returns (free_xyz, dirichlet_xyz)."""
    fe = space.fe
    geo = GT.refcell.Lagrange(mesh.cellkind, 1)
    N, _ = geo.tabulate(fe.nodes)  # (nnodes, nv)
    X = mesh.coords[mesh.cells]  # (ncell, nv, D)
    xn = np.einsum("an,cnd->cad", N, X)  # (ncell, nnodes, D)
    nc = space.ncomp
    free = np.zeros((space.nfree, mesh.D))
    diri = np.zeros((space.ndiri, mesh.D))
    for a in range(fe.nnodes):
        for k in range(nc):
            d = space.cell_dofs[:, a * nc + k]
            free[d[d > 0] - 1] = xn[d > 0, a]
            diri[-d[d < 0] - 1] = xn[d < 0, a]
    return free, diri


def csc(colptr, rowval, nzval, m):
    import scipy.sparse as sp
    return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(m, len(colptr) - 1))
