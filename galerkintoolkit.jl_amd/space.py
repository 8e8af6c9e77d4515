"""Domains, Lagrange spaces and their DOF numbering (host side, numpy).

Mirrors GT.interior / GT.boundary (GalerkinToolkit.jl src/mesh.jl:223-271)
and GT.lagrange_space (src/space.jl:1702-1735) with the DOF numbering of
generate_dof_ids (src/space.jl:299-535): DOFs are numbered by dimension
d = 0..D, face by face in face-id order, each face receiving its own DOFs
consecutively; a cell's own DOF i on local face f gets
offset(face) + perm_pindex[i] with pindex the face permutation id; DOFs on
the closure of Dirichlet faces are moved after the free ones and encoded as
negative ids (src/space.jl:477-535).
"""
import numpy as np

from . import refcell as R

FREE, DIRICHLET = 1, 2


class Domain:
    """A set of d-faces of a mesh (a GT.MeshDomain)."""

    def __init__(self, mesh, dim, faces, name="", is_skeleton=False):
        self.mesh = mesh
        self.dim = dim
        self.faces = np.asarray(faces, np.int64)  # 0-based face ids
        self.name = name
        self.is_skeleton = is_skeleton  # interior faces integrated from both sides

    def num_dims(self):
        return self.dim

    def num_faces(self):
        return len(self.faces)


def interior(mesh):
    return Domain(mesh, mesh.D, np.arange(mesh.cells.shape[0]), "interior")


def skeleton(mesh):
    """GT.skeleton(mesh) (src/mesh.jl:236-247): the interior (D-1)-faces,
    those with two incident cells (group_interior_faces!)."""
    d = mesh.D - 1
    cnt = np.bincount(mesh.cell_faces[d].ravel(), minlength=mesh.num_faces(d))
    return Domain(mesh, d, np.nonzero(cnt == 2)[0], "skeleton", is_skeleton=True)


def boundary(mesh, group_names=None):
    """GT.boundary(mesh; group_names): without names, the (D-1)-faces with
    one incident cell; with names, the union of those parent face groups."""
    d = mesh.D - 1
    if group_names is None:
        return Domain(mesh, d, mesh.boundary_faces(), "boundary")
    groups = mesh.group_faces(d)
    faces = np.unique(np.concatenate([groups[g] for g in group_names]))
    return Domain(mesh, d, faces, ",".join(group_names))


class LagrangeSpace:
    """GT.lagrange_space(domain, order; tensor_size, dirichlet_boundary,
    continuous).  continuous=False: every DOF belongs to its cell
    (reference_face_own_dofs, src/space.jl:860-880), so step 1 numbers them
    cell by cell (cell c's local DOF a is c*nd + a + 1); the Dirichlet step is
    the same."""

    def __init__(self, domain, order, tensor_size=None, dirichlet_boundary=None, continuous=True):
        mesh = domain.mesh
        if domain.dim != mesh.D or len(domain.faces) != mesh.cells.shape[0]:
            raise NotImplementedError("spaces on the whole interior of the mesh only")
        self.domain = domain
        self.mesh = mesh
        self.order = int(order)
        self.ncomp = 1 if tensor_size is None else int(np.prod(tensor_size))
        self.fe = R.Lagrange(mesh.cellkind, self.order, self.ncomp)
        self.dirichlet_boundary = dirichlet_boundary
        self.continuous = bool(continuous)
        self._number()

    def _number(self):
        mesh, fe, nc = self.mesh, self.fe, self.ncomp
        D = mesh.D
        ncell = mesh.cells.shape[0]
        g = R.GEO[mesh.cellkind]
        # generate_dof_ids_step_1: offsets per d-face, d = 0..D
        base = 0
        offsets = {}
        for d in range(D + 1):
            cnt = fe.own_nodes[d].shape[1] * nc
            offsets[d] = (base, cnt)
            base += mesh.num_faces(d) * cnt
        ndofs = base
        cell_dofs = np.zeros((ncell, fe.ndofs), np.int64)
        if not self.continuous:
            ndofs = ncell * fe.ndofs
            cell_dofs[:] = np.arange(ndofs, dtype=np.int64).reshape(ncell, fe.ndofs) + 1
        for d in range(D + 1 if self.continuous else 0):
            b0, cnt = offsets[d]
            own = fe.own_nodes[d]  # (nlf, n_own)
            P = fe.own_perms[d]  # (np, n_own)
            nlf = own.shape[0]
            for l in range(nlf):
                face = mesh.cell_faces[d][:, l]
                pid = mesh.cell_perm[d][:, l]
                off = b0 + face * cnt
                for i in range(own.shape[1]):
                    node = own[l, i]
                    pn = P[pid, i]
                    for k in range(nc):
                        cell_dofs[:, (node - 1) * nc + k] = off + (pn - 1) * nc + k + 1
        # generate_dof_ids_step_2: Dirichlet tags on the closure of the faces
        nfree, ndiri = ndofs, 0
        db = self.dirichlet_boundary
        if db is not None:
            N = db.dim
            if N != D - 1:
                raise NotImplementedError("Dirichlet boundaries of dimension D-1 only")
            ftag = np.zeros(mesh.num_faces(N), bool)
            ftag[db.faces] = True
            dtag = np.zeros(ndofs + 1, bool)
            fn = fe.face_nodes[N]  # closure nodes of each local face
            for l in range(fn.shape[0]):
                sel = ftag[mesh.cell_faces[N][:, l]]
                if not np.any(sel):
                    continue
                cols = ((fn[l][:, None] - 1) * nc + np.arange(nc)[None, :]).ravel()
                dtag[cell_dofs[np.ix_(sel, cols)].ravel()] = True
            dtag = dtag[1:]
            perm = np.zeros(ndofs, np.int64)
            free_ids = np.nonzero(~dtag)[0]
            diri_ids = np.nonzero(dtag)[0]
            perm[free_ids] = np.arange(1, len(free_ids) + 1)
            perm[diri_ids] = -np.arange(1, len(diri_ids) + 1)
            cell_dofs = perm[cell_dofs - 1]
            nfree, ndiri = len(free_ids), len(diri_ids)
        self.cell_dofs = cell_dofs.astype(np.int32)
        self.ndofs = ndofs
        self.nfree = nfree
        self.ndiri = ndiri
        self._g = g

    # reference-like accessors
    def num_free_dofs(self):
        return self.nfree

    def num_dirichlet_dofs(self):
        return self.ndiri

    def face_dofs(self):
        """Cell -> DOF map in the reference's signed encoding (1-based)."""
        return self.cell_dofs


def lagrange_space(domain, order, tensor_size=None, dirichlet_boundary=None, continuous=True):
    return LagrangeSpace(domain, order, tensor_size=tensor_size, dirichlet_boundary=dirichlet_boundary,
                         continuous=continuous)
