"""Cartesian meshes and their face complex (host side, numpy).

Mirrors GT.cartesian_mesh (GalerkinToolkit.jl src/cartesian_mesh.jl:16-54)
with boundary faces and face groups (:97-211 hex/quad, :330-461 simplices)
and GT.complexify (src/topology.jl:867-1125).  The numbering rules are the
reference's; the algorithm is vectorised: a d-face id is assigned at the first
(n-face, local d-face) pair, in n-face order, whose vertex set is new, after the
parent (boundary) d-faces which keep ids 1..m (generate_face_boundary,
src/topology.jl:1591-1704); its vertex order is the parent's, else that of the
first n-face that has it (generate_face_vertices, :1350-1415).
"""
import itertools

import numpy as np

from . import refcell as R


class Mesh:
    """A face complex: nodes, cells (order-1 geometry) and every d-face.

    Ids are 0-based in numpy arrays; `face_nodes(d)` / `cell_nodes_1b()`
    return the reference's 1-based ids.
    """

    def __init__(self, D, coords, cells, cellkind, parent_faces, parent_groups):
        self.D = D
        self.coords = np.ascontiguousarray(coords, np.float64)
        self.cells = np.ascontiguousarray(cells, np.int64)  # 0-based node ids
        self.cellkind = cellkind
        self.parent_faces = parent_faces  # d -> (m, w) 0-based node ids
        self.parent_groups = parent_groups  # d -> (m,) local cube face (1-based)
        self._complexify()

    # -- topology -----------------------------------------------------------
    def _complexify(self):
        D, kind = self.D, self.cellkind
        nnodes = self.coords.shape[0]
        # create_vertices! (topology.jl:1034-1097): parent 0-faces first
        first = []
        seen = np.zeros(nnodes, bool)
        if 0 in self.parent_faces and len(self.parent_faces[0]):
            for n in self.parent_faces[0][:, 0]:
                if not seen[n]:
                    seen[n] = True
                    first.append(n)
        is_vertex = np.zeros(nnodes, bool)
        is_vertex[self.cells.ravel()] = True
        for d, pf in self.parent_faces.items():
            is_vertex[pf.ravel()] = True
        rest = np.nonzero(is_vertex & ~seen)[0]
        vertex_node = np.concatenate([np.asarray(first, np.int64), rest])
        node_vertex = np.full(nnodes, -1, np.int64)
        node_vertex[vertex_node] = np.arange(len(vertex_node))
        self.vertex_node = vertex_node
        self.node_vertex = node_vertex
        nv = len(vertex_node)

        self.nfaces = {0: nv, D: self.cells.shape[0]}
        self.face_vertices = {0: np.arange(nv, dtype=np.int64)[:, None], D: node_vertex[self.cells]}
        nface_dfaces = {}
        self.face_keys = {}
        for d in range(D - 1, 0, -1):
            n = d + 1
            nkind = kind if n == D else R.GEO[kind].ftype[n]
            lf = np.asarray(R.GEO[nkind].faces[d]) - 1  # (nlf, wd)
            nfv = self.face_vertices[n]
            cand = nfv[:, lf]  # (nn, nlf, wd) visiting order of the reference
            nn, nlf, wd = cand.shape
            cand = cand.reshape(nn * nlf, wd)
            pf = self.parent_faces.get(d, np.zeros((0, wd), np.int64))
            pv = node_vertex[pf] if len(pf) else np.zeros((0, wd), np.int64)
            ckey = _keys(cand, nv)
            pkey = _keys(pv, nv)
            npar = len(pv)
            # parent lookup
            ids = np.full(nn * nlf, -1, np.int64)
            if npar:
                order = np.argsort(pkey, kind="stable")
                pos = np.searchsorted(pkey[order], ckey)
                pos = np.minimum(pos, npar - 1)
                hit = pkey[order][pos] == ckey
                ids[hit] = order[pos[hit]]
            new = ids < 0
            uk, first_idx, inv = np.unique(ckey[new], return_index=True, return_inverse=True)
            # rank unique keys by first occurrence
            rank = np.empty(len(uk), np.int64)
            rank[np.argsort(first_idx, kind="stable")] = np.arange(len(uk))
            ids[new] = npar + rank[inv]
            nd = npar + len(uk)
            nface_dfaces[n] = ids.reshape(nn, nlf)
            # face vertices: parents keep theirs, others from first n-face
            fv = np.zeros((nd, wd), np.int64)
            fv[:npar] = pv
            new_pos = np.nonzero(new)[0][first_idx]
            fv[npar + rank] = cand[new_pos]
            self.face_vertices[d] = fv
            self.nfaces[d] = nd
            self.face_keys[d] = _keys(fv, nv)
        # cell -> d-face incidence (face_incidence, topology.jl:219-244)
        g = R.GEO[kind]
        self.cell_faces = {D: np.arange(self.cells.shape[0], dtype=np.int64)[:, None],
                           0: self.face_vertices[D]}
        for d in range(1, D):
            if d == D - 1:
                self.cell_faces[d] = nface_dfaces[D]
            else:
                lf = np.asarray(g.faces[d]) - 1
                cand = self.face_vertices[D][:, lf]
                nc, nlf, wd = cand.shape
                ck = _keys(cand.reshape(nc * nlf, wd), nv)
                fk = self.face_keys[d]
                order = np.argsort(fk, kind="stable")
                pos = np.searchsorted(fk[order], ck)
                if np.any(pos >= len(fk)) or np.any(fk[order][np.minimum(pos, len(fk) - 1)] != ck):
                    raise ValueError("mesh is not a face complex")
                self.cell_faces[d] = order[pos].reshape(nc, nlf)
        # permutation ids (fill_face_permutation_ids!, topology.jl:590-666)
        self.cell_perm = {}
        for d in range(0, D + 1):
            nlf = len(g.faces[d]) if d < D else 1
            if d in (0, D):
                self.cell_perm[d] = np.zeros((self.cells.shape[0], nlf), np.int64)
                continue
            perms = np.asarray(R.vertex_permutations(g.ftype[d])) - 1  # (np, wd)
            lf = np.asarray(g.faces[d]) - 1
            cellv = self.face_vertices[D][:, lf]  # (nc, nlf, wd)
            facev = self.face_vertices[d][self.cell_faces[d]]  # (nc, nlf, wd)
            pid = np.full(cellv.shape[:2], -1, np.int64)
            for p in range(len(perms) - 1, -1, -1):
                ok = np.all(facev[:, :, perms[p]] == cellv, axis=2)
                pid[ok] = p
            if np.any(pid < 0):
                raise ValueError("valid pindex not found")
            self.cell_perm[d] = pid

    # -- accessors mirroring the reference --------------------------------------
    def num_dims(self):
        return self.D

    def num_nodes(self):
        return self.coords.shape[0]

    def num_faces(self, d):
        return self.nfaces[d]

    def node_coordinates(self):
        return self.coords

    def cell_nodes_1b(self):
        return (self.cells + 1).astype(np.int32)

    def group_faces(self, d):
        """Parent face groups "d-face-i" (and "boundary" for d = D-1)."""
        groups = {}
        if d < self.D and d in self.parent_groups:
            g = self.parent_groups[d]
            for i in range(1, len(R.GEO[R.cube_of(self.cellkind, self.D)].faces[d]) + 1):
                groups[f"{d}-face-{i}"] = np.nonzero(g == i)[0]
            if d == self.D - 1 and not R.is_simplex(self.cellkind):
                groups["boundary"] = np.arange(len(g))
        if d == self.D:
            groups["interior"] = np.arange(self.cells.shape[0])
            groups[f"{self.D}-face-1"] = np.arange(self.cells.shape[0])
        return groups

    def boundary_faces(self):
        """(D-1)-faces with exactly one incident cell (group_boundary_faces!,
        src/mesh.jl:131-143), increasing id."""
        d = self.D - 1
        cnt = np.bincount(self.cell_faces[d].ravel(), minlength=self.nfaces[d])
        return np.nonzero(cnt == 1)[0]


def _keys(v, nv):
    """Order-free key of each row of vertex ids: the sorted ids packed in an
    int64 when they fit, else the sorted row viewed as one opaque record.
    Keys of one mesh compare equal exactly when the vertex sets are equal."""
    v = np.sort(np.asarray(v, np.int64), axis=1)
    w = v.shape[1]
    bits = max(1, int(nv).bit_length())
    if bits * w <= 63:
        k = np.zeros(v.shape[0], np.int64)
        for i in range(w):
            k = (k << bits) | v[:, i]
        return k
    v = np.ascontiguousarray(v)
    return v.view(np.dtype((np.void, 8 * w))).ravel()


def cartesian_mesh(domain, cells_per_dir, boundary=True, complexify=True, simplexify=False):
    """GT.cartesian_mesh(domain, cells_per_dir; simplexify).

    domain = (x0, x1, y0, y1[, z0, z1]).  Nodes are lexicographic (x fastest)
    with x = pmin + h*(i-1) (cartesian_chain, src/cartesian_mesh.jl:213-263).
    """
    if not boundary or not complexify:
        raise NotImplementedError("only boundary=True, complexify=True meshes are supported")
    cells_per_dir = tuple(int(c) for c in cells_per_dir)
    D = len(cells_per_dir)
    if D not in (2, 3):
        raise ValueError("2D or 3D meshes only")
    if any(c != 1 for c in cells_per_dir) and any(c < 2 for c in cells_per_dir):
        raise ValueError("At least 2 cells in any direction (or 1 cell in all directions)")
    npd = [c + 1 for c in cells_per_dir]
    h = [(domain[2 * d + 1] - domain[2 * d]) / cells_per_dir[d] for d in range(D)]
    idx = np.indices(npd[::-1]).reshape(D, -1)[::-1]  # idx[d] = i_d, x fastest
    coords = np.stack([domain[2 * d] + h[d] * idx[d] for d in range(D)], axis=1)
    # hex/quad cells: local nodes {0,1}^D x fastest
    cidx = np.indices(cells_per_dir[::-1]).reshape(D, -1)[::-1]
    strides = np.cumprod([1] + npd[:-1])
    base = sum(cidx[d] * strides[d] for d in range(D))
    loc = [sum(((l >> d) & 1) * strides[d] for d in range(D)) for l in range(1 << D)]
    hexn = base[:, None] + np.asarray(loc)[None, :]
    # parent boundary faces: a local d-face is kept when every node is
    # touched by at most 2^d cells (cartesian_mesh.jl:101-182)
    count = np.bincount(hexn.ravel(), minlength=coords.shape[0])
    cube = R.cube_of(R.HEX if D == 3 else R.QUAD, D)
    parent_faces, parent_groups = {}, {}
    sub = R.simplex_cube_subfaces(D) if simplexify else None
    for d in range(D):
        lfs = np.asarray(R.GEO[cube].faces[d]) - 1
        nodes = hexn[:, lfs]  # (nc, nlf, w)
        keep = np.all(count[nodes] <= (1 << d), axis=2)  # (nc, nlf)
        ci, li = np.nonzero(keep)  # cell-major, local-face minor
        if not simplexify:
            parent_faces[d] = nodes[ci, li]
            parent_groups[d] = li + 1
        else:
            rows, grp = [], []
            for c, l in zip(ci, li):
                for s in sub[d][l]:
                    rows.append(hexn[c, np.asarray(s) - 1])
                    grp.append(l + 1)
            parent_faces[d] = np.asarray(rows, np.int64).reshape(-1, d + 1)
            parent_groups[d] = np.asarray(grp, np.int64)
    if not simplexify:
        kind = R.HEX if D == 3 else R.QUAD
        cells = hexn
    else:
        kind = R.TET if D == 3 else R.TRI
        sn = np.asarray(R.SIMPLEX_NODES[D]) - 1
        cells = hexn[:, sn].reshape(-1, D + 1)
    return Mesh(D, coords, cells, kind, parent_faces, parent_groups)


def perturb_interior_nodes(mesh, amplitude, seed):
    """Displace every interior node by U(-amplitude*h, amplitude*h) per
    component (h = smallest spacing), from the counter-based generator."""
    from .inputs import uniform
    x = mesh.coords.copy()
    lo, hi = x.min(axis=0), x.max(axis=0)
    tol = 1e-12 * max(1.0, float(np.max(hi - lo)))
    interior = np.all((x > lo + tol) & (x < hi - tol), axis=1)
    h = _spacing(mesh)
    u = uniform(seed, x.size).reshape(x.shape) * 2.0 - 1.0
    x[interior] += amplitude * h * u[interior]
    return x


def _spacing(mesh):
    c = mesh.coords[mesh.cells[0]]
    e = np.abs(c[1:] - c[0])
    return float(np.min(e[e > 0]))


__all__ = ["Mesh", "cartesian_mesh", "perturb_interior_nodes"]
_ = itertools
