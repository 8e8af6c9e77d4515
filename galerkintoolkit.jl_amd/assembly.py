"""GT-style assembly entry points backed by libgtk_hip.

Mirrors GT.measure (GalerkinToolkit.jl src/problems.jl:4-25),
GT.assemble_matrix / GT.assemble_vector (src/problems.jl:244-274, :319-350),
GT.update_vector! / GT.update_matrix! (src/problems.jl:276-285, :352-361),
the *_with_free_and_dirichlet_columns and assemble_matrix_and_vector*
families with their update_*! twins (src/problems.jl:363-435) and
PartitionedSolvers_linear_problem (src/problems.jl:439-452).
The reference takes a Julia lambda and compiles it; here the weak form is
one of the integrands the library implements (Laplace with a per-cell
coefficient, mass, isotropic elasticity, and the load f . v), passed as a
form object.  The result is the reference's SparseMatrixCSC{Float64,Int32}
layout (1-based colptr/rowval) and a Float64 vector.
"""
import dataclasses
import time
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import refcell as R
from .space import DIRICHLET, FREE, LagrangeSpace

FP64, MIXED = "fp64", "mixed"


class Measure:
    """GT.measure(domain, degree): the reference rule of every cell, or of
    every face for a boundary domain (dimension D-1; the rule of the face's
    reference cell, src/quadrature.jl:36-157)."""

    def __init__(self, domain, degree):
        self.domain = domain
        self.degree = int(degree)
        mesh = domain.mesh
        if domain.dim == mesh.D:
            self.kind = mesh.cellkind
        elif domain.dim == mesh.D - 1:
            self.kind = R.GEO[mesh.cellkind].ftype[mesh.D - 1]
        else:
            raise NotImplementedError("measures on cells or on (D-1)-faces")
        self.x, self.w = R.quadrature(self.kind, self.degree)

    @property
    def on_boundary(self):
        return self.domain.dim == self.domain.mesh.D - 1 and not self.on_skeleton

    @property
    def on_skeleton(self):
        return self.domain.is_skeleton


def measure(domain, degree):
    return Measure(domain, degree)


class FaceConstantField:
    """GT.face_constant_field(values, domain): one value (or tuple) per cell."""

    def __init__(self, values, domain):
        self.values = np.ascontiguousarray(values, np.float64)
        self.domain = domain


def face_constant_field(values, domain):
    return FaceConstantField(values, domain)


class DiscreteField:
    """A field of a space given by its free then Dirichlet values."""

    def __init__(self, space, free_values, dirichlet_values=None):
        self.space = space
        fv = np.asarray(free_values, np.float64)
        dv = np.zeros(space.ndiri) if dirichlet_values is None else np.asarray(dirichlet_values, np.float64)
        self.values = np.ascontiguousarray(np.concatenate([fv, dv]))


class FormSum:
    """Sum of integrals over several measures (the reference's
    DomainContribution sum): assembled into one matrix or vector."""

    def __init__(self, terms):
        self.terms = list(terms)

    def __add__(self, other):
        return FormSum(self.terms + (other.terms if isinstance(other, FormSum) else [other]))


class _Form:
    def __add__(self, other):
        return FormSum([self] + (other.terms if isinstance(other, FormSum) else [other]))


@dataclass
class Laplace(_Form):
    """int kappa grad(u) . grad(v) dOmega."""
    dOmega: Measure
    kappa: object = None


@dataclass
class Mass(_Form):
    """int kappa u v dOmega."""
    dOmega: Measure
    kappa: object = None


@dataclass
class Elasticity(_Form):
    """int sigma(eps(u)) : eps(v) dOmega, sigma = lambda tr(eps) I + 2 mu eps."""
    dOmega: Measure
    lam: object
    mu: object


@dataclass
class Load(_Form):
    """int f . v dOmega; f constant (scalar/tuple), per cell, or discrete."""
    dOmega: Measure
    f: object = 1.0


@dataclass
class Nitsche(_Form):
    """int (gamma/h) v u - v n.grad(u) - n.grad(v) u dGamma over boundary
    faces: the Nitsche terms of GalerkinToolkitExamples/src/poisson.jl:93-107
    (scalar fields), h the face's diameter (largest distance between its
    nodes; GT.face_diameter_field, src/field.jl:488-492,
    src/accessors.jl:907-921), n the outward unit normal
    (src/accessors.jl:1009-1030)."""
    dOmega: Measure
    gamma: float = 1.0


@dataclass
class InteriorPenalty(_Form):
    """int (gamma/h) jump(v).jump(u) - jump(v).mean(grad u)
    - mean(grad v).jump(u) dLambda over the interior faces (GT.skeleton):
    the interior-penalty terms of GalerkinToolkitExamples/src/poisson.jl:
    99-102 with jump(u) = u[1] n[1] + u[2] n[2] and mean(f) = (f[1] + f[2]) / 2
    (poisson.jl:40-41), h the face's diameter, n[i] the unit normal out of
    side i's cell (scalar fields on discontinuous spaces: on a continuous
    space every jump vanishes)."""
    dLambda: Measure
    gamma: float = 1.0

    @property
    def dOmega(self):
        return self.dLambda


@dataclass
class NitscheLoad(_Form):
    """int (gamma/h) v g - n.grad(v) g dGamma (poisson.jl:108-111), g a
    DiscreteField of the space (its free then Dirichlet values)."""
    dOmega: Measure
    g: object
    gamma: float = 1.0


class SparseMatrixCSC:
    """Julia's SparseMatrixCSC{Float64,Int32} layout: 1-based colptr/rowval."""

    def __init__(self, m, n, colptr, rowval, nzval):
        self.m, self.n = int(m), int(n)
        self.colptr, self.rowval, self.nzval = colptr, rowval, nzval

    @property
    def shape(self):
        return (self.m, self.n)

    @property
    def nnz(self):
        return len(self.rowval)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.nzval, self.rowval - 1, self.colptr - 1), shape=(self.m, self.n))


def _diameters(xyz, fnodes):
    """Largest distance between the nodes of each face (fnodes: items x
    face nodes, 0-based): GT.face_diameter_field (src/field.jl:488-492,
    src/accessors.jl:907-921)."""
    X = xyz[fnodes]
    diff = X[:, :, None, :] - X[:, None, :, :]
    return np.sqrt((diff ** 2).sum(axis=3)).max(axis=(1, 2))


class _Assembler:
    """The device context of one (space, measure) pair."""

    def set_coordinates(self, xyz):
        """New node coordinates (same topology)."""
        self.ctx.update_coordinates(xyz)

    def __init__(self, space, dOmega, device=0):
        mesh = space.mesh
        self.space = space
        self.dOmega = dOmega
        self.ctx = L.Context(device)
        self.ctx.set_mesh(mesh.D, mesh.coords, mesh.cell_nodes_1b())
        self.ctx.set_space(space.ncomp, space.cell_dofs, space.nfree, space.ndiri)
        phi, gphi = space.fe.tabulate(dOmega.x)
        _, gN = R.Lagrange(mesh.cellkind, 1).tabulate(dOmega.x)
        self.ctx.set_reference(dOmega.w, phi, gphi, gN)
        self.patterns = {}
        self.values_owner = {}  # (rows, cols) -> token of the matrix whose values the device holds

    def pattern(self, rows, cols, symbolic="host"):
        """CSC pattern of the (rows, cols) block, built once: symbolic="host"
        (threaded C++) or "device" (GPU kernels; bit-identical)."""
        key = (rows, cols)
        if key not in self.patterns:
            if symbolic == "device":
                nnz, ncols, _ = self.ctx.build_pattern_device(rows, cols)
            elif symbolic == "host":
                nnz, ncols = self.ctx.build_pattern(rows, cols)
            else:
                raise ValueError("symbolic is 'host' or 'device'")
            colptr, rowval = self.ctx.get_pattern(rows, cols, nnz, ncols)
            nrows = self.space.nfree if rows == FREE else self.space.ndiri
            self.patterns[key] = (nnz, ncols, nrows, colptr, rowval)
        return self.patterns[key]


class _BoundaryAssembler(_Assembler):
    """The device context of one (space, boundary measure) pair: one item per
    face of the domain (the face's cell and its local face), the item's DOFs
    on that face's closure nodes, and the cell's shape functions and
    geometry gradients at the face rule's points mapped into the cell's
    reference coordinates, per local face."""

    def __init__(self, space, dGamma, device=0):
        mesh = space.mesh
        D = mesh.D
        self.space = space
        self.dOmega = dGamma
        faces = dGamma.domain.faces
        cf = mesh.cell_faces[D - 1]  # (ncells, nlf) face ids
        nlf = cf.shape[1]
        owner = np.full(mesh.num_faces(D - 1), -1, np.int64)
        owner[cf.ravel()] = np.arange(cf.size)  # boundary faces have one cell
        hit = owner[faces]
        if np.any(hit < 0):
            raise ValueError("a face of the boundary domain has no cell")
        cells, lface = hit // nlf, hit % nlf
        fe = space.fe
        fnodes = fe.face_nodes[D - 1]  # (nlf, nf) 1-based cell nodes
        nc = space.ncomp
        cols = ((fnodes[lface] - 1)[:, :, None] * nc + np.arange(nc)[None, None, :]).reshape(len(faces), -1)
        item_dofs = np.take_along_axis(space.cell_dofs[cells], cols, axis=1)
        g = R.GEO[mesh.cellkind]
        ft = g.ftype[D - 1]
        N = R.linear_shape(ft, dGamma.x)  # face rule -> reference cell points
        _, dNf = R.Lagrange(ft, 1).tabulate(dGamma.x[:1])
        geo = R.Lagrange(mesh.cellkind, 1)
        phi, gN, A = [], [], []
        for l, fv in enumerate(g.faces[D - 1]):
            X = np.asarray([g.verts[v - 1] for v in fv], np.float64)
            xc = N @ X
            val, _ = fe.tabulate(xc)
            phi.append(val[:, fnodes[l] - 1])
            gN.append(geo.tabulate(xc)[1])
            A.append(X.T @ dNf[0])  # (D, D-1)
        self.ctx = L.Context(device)
        self.ctx.set_mesh(D, mesh.coords, mesh.cell_nodes_1b()[cells])
        self.ctx.set_space(nc, item_dofs, space.nfree, space.ndiri)
        self.ctx.set_reference_faces(dGamma.w, np.stack(phi), np.stack(gN), np.stack(A), lface)
        self.items = len(faces)
        self.patterns = {}
        self.values_owner = {}


class _NitscheAssembler(_Assembler):
    """The device context of Nitsche terms over a boundary measure: one item
    per face of the domain with all DOFs of its cell (normal derivatives
    involve every shape function), the cell's shape functions, reference
    gradients and order-1 geometry gradients at the face rule's points mapped
    into the cell, per local face, the reference outward normals, and the
    face diameters."""

    def __init__(self, space, dGamma, device=0):
        mesh = space.mesh
        D = mesh.D
        if space.ncomp != 1:
            raise NotImplementedError("Nitsche terms: scalar fields")
        self.space = space
        self.dOmega = dGamma
        faces = dGamma.domain.faces
        cf = mesh.cell_faces[D - 1]
        nlf = cf.shape[1]
        owner = np.full(mesh.num_faces(D - 1), -1, np.int64)
        owner[cf.ravel()] = np.arange(cf.size)
        hit = owner[faces]
        if np.any(hit < 0):
            raise ValueError("a face of the boundary domain has no cell")
        cells, lface = hit // nlf, hit % nlf
        g = R.GEO[mesh.cellkind]
        ft = g.ftype[D - 1]
        N = R.linear_shape(ft, dGamma.x)
        _, dNf = R.Lagrange(ft, 1).tabulate(dGamma.x[:1])
        geo = R.Lagrange(mesh.cellkind, 1)
        fe = space.fe
        phi, gphi, gN, A = [], [], [], []
        for fv in g.faces[D - 1]:
            X = np.asarray([g.verts[v - 1] for v in fv], np.float64)
            xc = N @ X
            val, grad = fe.tabulate(xc)
            phi.append(val)
            gphi.append(grad)
            gN.append(geo.tabulate(xc)[1])
            A.append(X.T @ dNf[0])
        # face diameters: largest distance between the face's nodes
        fn = np.asarray(g.faces[D - 1])[lface] - 1  # (items, face vertices) local vertices
        self.fnodes = np.take_along_axis(mesh.cells[cells], fn, axis=1)
        self.h = _diameters(mesh.coords, self.fnodes)
        self.ctx = L.Context(device)
        self.ctx.set_mesh(D, mesh.coords, mesh.cell_nodes_1b()[cells])
        self.ctx.set_space(1, space.cell_dofs[cells], space.nfree, space.ndiri)
        self.ctx.set_reference_faces(dGamma.w, np.stack(phi), np.stack(gN), np.stack(A), lface)
        self.ctx.set_reference_face_gradients(np.stack(gphi), R.face_normals(mesh.cellkind))
        self.items = len(faces)
        self.patterns = {}
        self.values_owner = {}

    def beta(self, gamma):
        """gamma / h per item."""
        return np.ascontiguousarray(float(gamma) / self.h)

    def set_coordinates(self, xyz):
        """New node coordinates: the context's, and the face diameters."""
        self.ctx.update_coordinates(xyz)
        self.h = _diameters(xyz, self.fnodes)


def skeleton_items(space, dL):
    """Host side of the skeleton items of a space over a skeleton measure:
    one item per interior face with both its cells (the lower cell id
    first, side 0), and the point sets of the face rule: for each side, the
    rule's points in that cell's reference coordinates, side 1's listed in
    the physical order of side 0's through the vertices the two local faces
    share.  Returns a dict: cells and lfaces (items x 2), item_sets (items x
    2, the set of each side), fnodes (side 0's face vertices, mesh nodes),
    and per set: the local face and vertex order (keys) and the points in
    the cell's reference coordinates (points, sets x nq x D)."""
    mesh = space.mesh
    D = mesh.D
    faces = dL.domain.faces
    cf = mesh.cell_faces[D - 1]
    nlf = cf.shape[1]
    flat = cf.ravel()
    order = np.argsort(flat, kind="stable")  # (cell, local face) slots grouped by face, cells ascending
    sf = flat[order]
    start = np.searchsorted(sf, faces)
    if np.any(np.searchsorted(sf, faces, side="right") - start != 2):
        raise ValueError("a face of the skeleton does not have two cells")
    p0, p1 = order[start], order[start + 1]
    cells = np.stack([p0 // nlf, p1 // nlf], axis=1)
    lfaces = np.stack([p0 % nlf, p1 % nlf], axis=1)
    g = R.GEO[mesh.cellkind]
    fv = np.asarray(g.faces[D - 1]) - 1  # (nlf, face vertices)
    V0 = np.take_along_axis(mesh.cells[cells[:, 0]], fv[lfaces[:, 0]], axis=1)  # side 0's face, mesh nodes
    V1 = np.take_along_axis(mesh.cells[cells[:, 1]], fv[lfaces[:, 1]], axis=1)
    # perm[k]: position in side 1's local face of side 0's k-th vertex
    perm = np.argmax(V1[:, None, :] == V0[:, :, None], axis=2)
    if not np.all(np.take_along_axis(V1, perm, axis=1) == V0):
        raise ValueError("the two sides of a face do not share its vertices")
    nfv = fv.shape[1]
    nit = len(faces)
    # point sets: distinct (local face, vertex order) keys of the sides
    side_keys = np.concatenate([
        np.concatenate([lfaces[:, :1], np.broadcast_to(np.arange(nfv), (nit, nfv))], axis=1),
        np.concatenate([lfaces[:, 1:], perm], axis=1)])
    keys, inv = np.unique(side_keys, axis=0, return_inverse=True)
    N = R.linear_shape(g.ftype[D - 1], dL.x)
    points = np.stack([N @ np.asarray([g.verts[v] for v in fv[int(k[0])]], np.float64)[k[1:]] for k in keys])
    return dict(cells=cells, lfaces=lfaces, item_sets=inv.reshape(2, nit).T, fnodes=V0, keys=keys, points=points)


class _SkeletonAssembler(_Assembler):
    """The device context of interior-penalty terms over a skeleton measure
    (the items and point sets of skeleton_items): the cells' shape
    functions, reference gradients and order-1 geometry gradients at each
    set's points, the set's local face map and reference outward normal,
    and the face diameters."""

    def __init__(self, space, dL, device=0):
        mesh = space.mesh
        D = mesh.D
        if space.ncomp != 1:
            raise NotImplementedError("skeleton terms: scalar fields")
        if space.continuous:
            raise ValueError("a continuous space's skeleton terms go through _ContinuousSkeletonAssembler")

        self.space = space
        self.dOmega = dL
        it = skeleton_items(space, dL)
        cells = it["cells"]
        nit = len(cells)
        # every cell must lie on an interior face: the skeleton pattern then
        # holds the volume and boundary blocks of a sum
        if not np.all(np.bincount(cells.ravel(), minlength=mesh.cells.shape[0]) > 0):
            raise NotImplementedError("skeleton terms: every cell needs an interior face")
        nsets = 8 if D == 2 else 12
        if len(it["keys"]) > nsets:
            raise NotImplementedError(f"skeleton terms: {len(it['keys'])} point sets (at most {nsets})")
        g = R.GEO[mesh.cellkind]
        fv = np.asarray(g.faces[D - 1]) - 1
        _, dNf = R.Lagrange(g.ftype[D - 1], 1).tabulate(dL.x[:1])
        geo = R.Lagrange(mesh.cellkind, 1)
        fe = space.fe
        normals = R.face_normals(mesh.cellkind)
        nq, nv = len(dL.w), len(g.verts)
        phi, gphi, gN, A, nrm = [], [], [], [], []
        for key, xc in zip(it["keys"], it["points"]):
            l = int(key[0])
            X = np.asarray([g.verts[v] for v in fv[l]], np.float64)[key[1:]]
            val, grad = fe.tabulate(xc)
            phi.append(val)
            gphi.append(grad)
            # one cell's geometry gradients, then zeros: a set spans the
            # item's two cells' worth of nodes
            gN.append(np.concatenate([geo.tabulate(xc)[1].ravel(), np.zeros(nq * nv * D)]))
            A.append(X.T @ dNf[0])
            nrm.append(normals[l])
        while len(phi) < nsets:  # unused sets: copies of the first
            phi.append(phi[0]), gphi.append(gphi[0]), gN.append(gN[0]), A.append(A[0]), nrm.append(nrm[0])
        self.fnodes = it["fnodes"]
        self.h = _diameters(mesh.coords, self.fnodes)
        item_sets = it["item_sets"]
        item_nodes = mesh.cell_nodes_1b()[cells].reshape(nit, -1)
        item_dofs = space.cell_dofs[cells].reshape(nit, -1)
        item_face = (item_sets[:, 0] | (item_sets[:, 1] << 16)).astype(np.int32)
        self.ctx = L.Context(device)
        self.ctx.set_mesh(D, mesh.coords, item_nodes)
        self.ctx.set_space(1, item_dofs, space.nfree, space.ndiri)
        self.ctx.set_reference_faces(dL.w, np.stack(phi), np.stack(gN), np.stack(A), item_face)
        self.ctx.set_reference_face_gradients(np.stack(gphi), np.stack(nrm))
        self.items = nit
        self.patterns = {}
        self.values_owner = {}

    def beta(self, gamma):
        """gamma / h per item."""
        return np.ascontiguousarray(float(gamma) / self.h)

    def set_coordinates(self, xyz):
        """New node coordinates: the context's, and the face diameters."""
        self.ctx.update_coordinates(xyz)
        self.h = _diameters(xyz, self.fnodes)


class _ContinuousSkeletonAssembler(_Assembler):
    """Interior-penalty terms of a continuous space: the face items would
    list each DOF of the face twice (the column-owner plan keeps one row per
    lane), so the terms are assembled on the cells' own copies of their DOFs
    (the discontinuous space of the same order, every DOF free) and each
    nonzero of the continuous block is the sum, in order, of the copies'
    entries that land on it (gtk_gather_block), as the reference's COO
    compression sums its duplicates.  The pattern is that of the face items
    over the continuous DOFs (both cells' DOFs coupled)."""

    def __init__(self, space, dL, device=0):
        mesh = space.mesh
        if space.ncomp != 1:
            raise NotImplementedError("skeleton terms: scalar fields")
        self.space = space
        self.dOmega = dL
        self.dg_space = LagrangeSpace(space.domain, space.order, continuous=False)
        self.dg = _SkeletonAssembler(self.dg_space, dL, device)
        it = skeleton_items(space, dL)
        cells = it["cells"]
        nit = len(cells)
        self.ctx = L.Context(device)
        self.ctx.set_mesh(mesh.D, mesh.coords, mesh.cell_nodes_1b()[cells].reshape(nit, -1))
        self.ctx.set_space(1, space.cell_dofs[cells].reshape(nit, -1), space.nfree, space.ndiri)
        self.cont = space.cell_dofs.reshape(-1).astype(np.int64)  # continuous id of each copy (copy id - 1)
        self.items = nit
        self.patterns = {}
        self.values_owner = {}
        self.maps = {}

    @property
    def h(self):
        return self.dg.h

    def beta(self, gamma):
        return self.dg.beta(gamma)

    def set_coordinates(self, xyz):
        self.ctx.update_coordinates(xyz)
        self.dg.set_coordinates(xyz)

    def gather_map(self, rows, cols):
        """(ptr, idx): the copies' entries summed into each nonzero of the
        (rows, cols) block, in ascending order of their position."""
        key = (rows, cols)
        if key not in self.maps:
            nnz_d, ncols_d, _, cp_d, rv_d = self.dg.pattern(FREE, FREE)
            nnz, _, nrows, cp, rv = self.pattern(rows, cols)
            col_d = np.repeat(np.arange(ncols_d, dtype=np.int64), np.diff(cp_d))
            I = self.cont[rv_d.astype(np.int64) - 1]
            J = self.cont[col_d]
            keep = (I > 0 if rows == FREE else I < 0) & (J > 0 if cols == FREE else J < 0)
            src = np.nonzero(keep)[0]
            r, c = np.abs(I[src]), np.abs(J[src])
            tcol = np.repeat(np.arange(len(cp) - 1, dtype=np.int64), np.diff(cp))
            tkeys = tcol * (nrows + 1) + rv.astype(np.int64)
            skeys = (c - 1) * (nrows + 1) + r
            pos = np.searchsorted(tkeys, skeys)
            if np.any(pos >= nnz) or np.any(tkeys[np.minimum(pos, nnz - 1)] != skeys):
                raise RuntimeError("a copy's entry falls outside the continuous skeleton pattern")
            order = np.argsort(pos, kind="stable")
            ptr = np.zeros(nnz + 1, np.int64)
            ptr[1:] = np.cumsum(np.bincount(pos, minlength=nnz))
            self.maps[key] = (ptr, src[order].astype(np.int64))
        return self.maps[key]

    def assemble(self, gamma, prec, rows, cols, nnz, device_only):
        self.dg.pattern(FREE, FREE)
        self.dg.ctx.set_cell_coefficients(self.dg.beta(gamma), 1)
        self.dg.ctx.assemble_matrix_async(L.GTK_SKELETON, prec, FREE, FREE)
        self.dg.ctx.sync()
        self.dg.values_owner.pop((FREE, FREE), None)
        ptr, idx = self.gather_map(rows, cols)
        self.ctx.gather_block(rows, cols, self.dg.ctx, FREE, FREE, ptr, idx)
        self.values_owner.pop((rows, cols), None)
        if device_only:
            return None, None
        stats = L.GtkStats()
        stats.nnz = nnz
        return self.ctx.get_values(rows, cols, nnz), stats


def _assembler(space, dOmega, device, kind=None):
    """The context of (space, measure); kind "nitsche" (or "nitsche_load",
    a second one so the load's matrix never replaces the Nitsche matrix's
    device values) for Nitsche items over a boundary measure."""
    key = (id(dOmega), device, kind)
    cache = space.__dict__.setdefault("_assemblers", {})
    a = cache.get(key)
    if a is None or a.dOmega is not dOmega:
        if dOmega.on_skeleton:
            a = (_ContinuousSkeletonAssembler if space.continuous else _SkeletonAssembler)(space, dOmega, device)
        elif kind is not None:
            a = _NitscheAssembler(space, dOmega, device)
        else:
            a = (_BoundaryAssembler if dOmega.on_boundary else _Assembler)(space, dOmega, device)
        cache[key] = a
    return a


def _form_code(form):
    if isinstance(form, Laplace):
        return L.GTK_LAPLACE
    if isinstance(form, Mass):
        return L.GTK_MASS
    if isinstance(form, Elasticity):
        return L.GTK_ELASTICITY
    if isinstance(form, Nitsche):
        return L.GTK_NITSCHE
    if isinstance(form, InteriorPenalty):
        return L.GTK_SKELETON
    raise TypeError("unsupported bilinear form")


def _ncoef_items(form, space):
    """Cells (or boundary items) the form's coefficients run over."""
    dom = form.dOmega.domain
    return len(dom.faces) if form.dOmega.domain.dim < space.mesh.D else space.mesh.cells.shape[0]


def _coefficients(form, ncells):
    if isinstance(form, Elasticity):
        lam = form.lam.values if isinstance(form.lam, FaceConstantField) else np.full(ncells, float(form.lam))
        mu = form.mu.values if isinstance(form.mu, FaceConstantField) else np.full(ncells, float(form.mu))
        return np.ascontiguousarray(np.stack([lam, mu], axis=1).ravel()), 2
    k = form.kappa
    if k is None:
        return None, 0
    if isinstance(k, FaceConstantField):
        return k.values, 1
    return np.full(ncells, float(k)), 1


def _with_parameters(form, parameters):
    """The form with some of its coefficient fields replaced: the values the
    reference passes as `parameters` to the generated loop
    (src/problems.jl:352-361)."""
    if not parameters:
        return form
    if isinstance(form, FormSum):
        return FormSum([_with_parameters(t, parameters) for t in form.terms])
    names = {f.name for f in dataclasses.fields(form)}
    upd = {k: v for k, v in parameters.items() if k in names and k != "dOmega"}
    return dataclasses.replace(form, **upd) if upd else form


def _lock(A, a, rows, cols):
    """Marks A as the owner of the device values of block (rows, cols) of
    assembler a; A.nzval becomes read-only so that a change of the values
    means a new array (seen by solve)."""
    A.nzval.setflags(write=False)
    A._token = object()
    A._nzval_ref = A.nzval
    A.device = (a, rows, cols)
    a.values_owner[(rows, cols)] = A._token


def _refill(A, nzval):
    """Values of an existing A replaced in place (update_matrix!)."""
    A.nzval.setflags(write=True)
    A.nzval[:] = nzval
    A.nzval.setflags(write=False)


class MatrixCache:
    """What update_matrix needs to refill a matrix: the form, its block and
    precision (the reference keeps the generated loops, the allocation and
    the compress cache, src/problems.jl:319-350)."""

    def __init__(self, form, space, rows, cols, precision, device):
        self.form, self.space, self.rows, self.cols = form, space, rows, cols
        self.precision, self.device = precision, device


class VectorCache:
    """What update_vector needs to refill a vector (src/problems.jl:244-274)."""

    def __init__(self, form, space, rows, device):
        self.form, self.space, self.rows, self.device = form, space, rows, device


def _launch(a, code, prec, rows, cols, nnz, device_only):
    """The block's values assembled on the device: (nzval, stats), or
    (None, None) with device_only (no copy to host memory)."""
    if device_only:
        a.ctx.assemble_matrix_async(code, prec, rows, cols)
        a.ctx.sync()
        a.values_owner.pop((rows, cols), None)
        return None, None
    return a.ctx.assemble_matrix(code, prec, rows, cols, nnz)


def _assemble_matrix_values(form, space, rows, cols, precision, device, symbolic="host", device_only=False):
    """(assembler of the block's device values, nzval, stats, pattern);
    device_only: the values stay on the device (nzval and stats None)."""
    if isinstance(form, FormSum):
        return _assemble_matrix_sum(form, space, rows, cols, precision, device, symbolic)
    if form.dOmega.on_boundary and not isinstance(form, (Mass, Nitsche)):
        raise NotImplementedError("boundary integrals of matrices: the mass form (Robin terms) and Nitsche terms")
    if form.dOmega.on_skeleton != isinstance(form, InteriorPenalty):
        raise NotImplementedError("skeleton integrals: the interior-penalty terms (over GT.skeleton)")
    if isinstance(form, InteriorPenalty):
        a = _assembler(space, form.dOmega, device)
        pat = a.pattern(rows, cols, symbolic)
        prec = L.GTK_FP64 if precision == FP64 else L.GTK_MIXED
        if isinstance(a, _ContinuousSkeletonAssembler):
            nzval, stats = a.assemble(form.gamma, prec, rows, cols, pat[0], device_only)
            return a, nzval, stats, pat
        a.ctx.set_cell_coefficients(a.beta(form.gamma), 1)
        nzval, stats = _launch(a, L.GTK_SKELETON, prec, rows, cols, pat[0], device_only)
        return a, nzval, stats, pat
    if isinstance(form, Nitsche):
        if not form.dOmega.on_boundary:
            raise ValueError("Nitsche terms are integrals over boundary faces")
        a = _assembler(space, form.dOmega, device, "nitsche")
        pat = a.pattern(rows, cols, symbolic)
        a.ctx.set_cell_coefficients(a.beta(form.gamma), 1)
        prec = L.GTK_FP64 if precision == FP64 else L.GTK_MIXED
        nzval, stats = _launch(a, L.GTK_NITSCHE, prec, rows, cols, pat[0], device_only)
        return a, nzval, stats, pat
    a = _assembler(space, form.dOmega, device)
    pat = a.pattern(rows, cols, symbolic)
    coef, ncoef = _coefficients(form, _ncoef_items(form, space))
    a.ctx.set_cell_coefficients(coef, ncoef)
    prec = L.GTK_FP64 if precision == FP64 else L.GTK_MIXED
    nzval, stats = _launch(a, _form_code(form), prec, rows, cols, pat[0], device_only)
    return a, nzval, stats, pat


def assemble_matrix(form, T, trial_space, test_space, free_or_dirichlet=(FREE, FREE), precision=FP64, reuse=False,
                    device=0, symbolic="host"):
    """GT.assemble_matrix(a, Float64, U, V; free_or_dirichlet, reuse) on the
    GPU (src/problems.jl:319-350).  symbolic: where the pattern is built the
    first time ("host" or "device").  reuse=True returns (A, cache) for
    update_matrix."""
    if T not in (float, np.float64):
        raise TypeError("values are Float64")
    if trial_space is not test_space:
        raise NotImplementedError("trial and test spaces must be the same space")
    space = test_space
    rows, cols = free_or_dirichlet
    a, nzval, stats, pat = _assemble_matrix_values(form, space, rows, cols, precision, device, symbolic)
    nnz, ncols, nrows, colptr, rowval = pat
    A = SparseMatrixCSC(nrows, ncols, colptr, rowval, nzval)
    A.stats = stats
    _lock(A, a, rows, cols)
    if reuse:
        return A, MatrixCache(form, space, rows, cols, precision, device)
    return A


def _set_coordinates(form, space, device, coords):
    """New node coordinates (same topology) for every context the form's
    terms run on."""
    terms = form.terms if isinstance(form, FormSum) else [form]
    xyz = np.ascontiguousarray(coords, np.float64)
    if xyz.shape != space.mesh.coords.shape:
        raise ValueError("coordinates must keep the mesh's shape")
    # every context of the space on this device: the terms' and those of the
    # Nitsche matrices and loads over the same measures
    cache = space.__dict__.get("_assemblers", {})
    measures = {id(t.dOmega) for t in terms}
    for (mid, dev, _), a in cache.items():
        if dev == device and mid in measures:
            a.set_coordinates(xyz)


def update_matrix(A, cache, parameters=None, coords=None, form=None):
    """GT.update_matrix!(A, cache; parameters) (src/problems.jl:352-361): the
    values of A are recomputed in place with the same pattern.  parameters:
    coefficient fields of the form to replace (e.g. {"kappa": field});
    coords: new node coordinates of the mesh (same topology); form: a whole
    replacement form of the same kind over the same measure."""
    if form is not None:
        cache.form = form
    cache.form = _with_parameters(cache.form, parameters)
    if coords is not None:
        _set_coordinates(cache.form, cache.space, cache.device, coords)
    a, nzval, stats, pat = _assemble_matrix_values(cache.form, cache.space, cache.rows, cache.cols, cache.precision,
                                                   cache.device)
    if pat[0] != A.nnz:
        raise ValueError("the cache does not belong to this matrix")
    _refill(A, nzval)
    A.stats = stats
    _lock(A, a, cache.rows, cache.cols)
    return A


def _load_data(form, space):
    """(kind, values) of the source of a Load form."""
    if not isinstance(form, Load):
        raise TypeError("unsupported linear form")
    f = form.f
    if isinstance(f, DiscreteField):
        return L.GTK_F_NODAL, f.values
    if isinstance(f, FaceConstantField):
        return L.GTK_F_CELL, f.values.reshape(-1)
    return L.GTK_F_CONST, np.broadcast_to(np.asarray(f, np.float64), (space.ncomp,)).copy()


def _assemble_matrix_sum(fs, space, rows, cols, precision, device, symbolic):
    """One volume integral plus boundary integrals and at most one skeleton
    integral: the block whose pattern holds the others' (the skeleton's,
    which couples the cells on either side of each interior face, else the
    volume's) is assembled, then each other block is added into its values
    on the device."""
    vol = [t for t in fs.terms if t.dOmega.domain.dim == space.mesh.D]
    skel = [t for t in fs.terms if t.dOmega.on_skeleton]
    if len(vol) != 1 or len(skel) > 1:
        raise NotImplementedError("a sum of matrix integrals: one over the cells, boundary integrals and at most "
                                  "one skeleton integral")
    base = skel[0] if skel else vol[0]
    # every block stays on the device until the sum is copied out once
    a_base, _, _, pat = _assemble_matrix_values(base, space, rows, cols, precision, device, symbolic,
                                                device_only=True)
    for t in fs.terms:
        if t is base:
            continue
        a_t, _, _, _ = _assemble_matrix_values(t, space, rows, cols, precision, device, device_only=True)
        a_base.ctx.add_block(rows, cols, a_t.ctx)
    t0 = time.perf_counter()
    nzval = a_base.ctx.get_values(rows, cols, pat[0])
    stats = L.GtkStats()
    stats.nnz = pat[0]
    stats.total_ms = (time.perf_counter() - t0) * 1e3
    return a_base, nzval, stats, pat


def _nitsche_load(form, space, rows, device):
    """int beta v g - n.grad(v) g: the load's matrix (rows x FREE, and
    rows x DIRICHLET where g has Dirichlet values) applied to g's nodal values
    on the device."""
    g = form.g
    if not isinstance(g, DiscreteField) or g.space is not space:
        raise TypeError("NitscheLoad: g is a DiscreteField of the space")
    a = _assembler(space, form.dOmega, device, "nitsche_load")
    a.ctx.set_cell_coefficients(a.beta(form.gamma), 1)
    n = space.nfree if rows == FREE else space.ndiri
    b = np.zeros(n)
    for cols, x in ((FREE, g.values[:space.nfree]), (DIRICHLET, g.values[space.nfree:])):
        if len(x) == 0 or n == 0:
            continue
        a.pattern(rows, cols)
        a.ctx.assemble_matrix_async(L.GTK_NITSCHE_LOAD, L.GTK_FP64, rows, cols)  # values stay on the device
        a.ctx.sync()
        a.values_owner.pop((rows, cols), None)  # the block's device values are the load's matrix now
        b = a.ctx.block_multiply(rows, cols, x, b, 1.0)
    return b


def _assemble_vector_values(form, space, rows, device):
    if isinstance(form, FormSum):
        return sum(_assemble_vector_values(t, space, rows, device) for t in form.terms)
    if isinstance(form, NitscheLoad):
        return _nitsche_load(form, space, rows, device)
    kind, data = _load_data(form, space)
    a = _assembler(space, form.dOmega, device)
    n = space.nfree if rows == FREE else space.ndiri
    b, stats = a.ctx.assemble_vector(kind, data, rows, n)
    return b


def assemble_vector(form, T, space, free_or_dirichlet=FREE, device=0, reuse=False):
    """GT.assemble_vector(l, Float64, V; free_or_dirichlet, reuse) on the GPU
    (src/problems.jl:244-274); l may be a sum of loads over cells and
    boundary faces.  reuse=True returns (b, cache) for update_vector."""
    if T not in (float, np.float64):
        raise TypeError("values are Float64")
    b = _assemble_vector_values(form, space, free_or_dirichlet, device)
    if reuse:
        return b, VectorCache(form, space, free_or_dirichlet, device)
    return b


def update_vector(b, cache, parameters=None, coords=None):
    """GT.update_vector!(b, cache; parameters) (src/problems.jl:276-285): b
    is recomputed in place; parameters replace fields of the load (e.g.
    {"f": field}), coords the node coordinates (same topology)."""
    cache.form = _with_parameters(cache.form, parameters)
    if coords is not None:
        _set_coordinates(cache.form, cache.space, cache.device, coords)
    nb = _assemble_vector_values(cache.form, cache.space, cache.rows, cache.device)
    if nb.shape != b.shape:
        raise ValueError("the cache does not belong to this vector")
    b[:] = nb
    return b


def assemble_matrix_with_free_and_dirichlet_columns(form, T, U, V, precision=FP64, device=0, reuse=False,
                                                    free_or_dirichlet=FREE):
    """(A, Ad): the (rows, FREE) and (rows, DIRICHLET) blocks
    (src/problems.jl:363-380); with reuse=True also their cache."""
    A, ca = assemble_matrix(form, T, U, V, (free_or_dirichlet, FREE), precision=precision, device=device, reuse=True)
    Ad, cd = assemble_matrix(form, T, U, V, (free_or_dirichlet, DIRICHLET), precision=precision, device=device,
                             reuse=True)
    if reuse:
        return A, Ad, (ca, cd)
    return A, Ad


def update_matrix_with_free_and_dirichlet_columns(A, Ad, cache, parameters=None, coords=None):
    """src/problems.jl:382-389."""
    ca, cd = cache
    update_matrix(A, ca, parameters, coords)
    update_matrix(Ad, cd, parameters)
    return A, Ad


def assemble_matrix_and_vector(a, l, T, U, V, precision=FP64, device=0, reuse=False, free_or_dirichlet=(FREE, FREE)):
    """(A, b) (src/problems.jl:391-404); with reuse=True also the cache."""
    A, mc = assemble_matrix(a, T, U, V, free_or_dirichlet, precision=precision, device=device, reuse=True)
    b, vc = assemble_vector(l, T, V, free_or_dirichlet[0], device=device, reuse=True)
    if reuse:
        return A, b, {"matrix_cache": mc, "vector_cache": vc}
    return A, b


def update_matrix_and_vector(A, b, cache, parameters=None, coords=None):
    """src/problems.jl:406-411."""
    update_matrix(A, cache["matrix_cache"], parameters, coords)
    update_vector(b, cache["vector_cache"], parameters)
    return A, b


def assemble_matrix_and_vector_with_free_and_dirichlet_columns(a, l, T, U, V, precision=FP64, device=0, reuse=False,
                                                               free_or_dirichlet=FREE):
    """(A, Ad, b) (src/problems.jl:413-428); with reuse=True also the cache."""
    A, Ad, mc = assemble_matrix_with_free_and_dirichlet_columns(a, T, U, V, precision=precision, device=device,
                                                                reuse=True, free_or_dirichlet=free_or_dirichlet)
    b, vc = assemble_vector(l, T, V, free_or_dirichlet, device=device, reuse=True)
    if reuse:
        return A, Ad, b, {"matrix_cache": mc, "vector_cache": vc}
    return A, Ad, b


def update_matrix_and_vector_with_free_and_dirichlet_columns(A, Ad, b, cache, parameters=None, coords=None):
    """src/problems.jl:430-435."""
    update_matrix_with_free_and_dirichlet_columns(A, Ad, cache["matrix_cache"], parameters, coords)
    update_vector(b, cache["vector_cache"], parameters)
    return A, Ad, b


def _ensure_device_values(A):
    """Makes the device hold A's own values on its block: uploads them when
    another matrix was assembled into that block since, or A.nzval was
    replaced."""
    a, rows, cols = A.device
    if a.values_owner.get((rows, cols)) is not A._token or A.nzval is not A._nzval_ref:
        a.ctx.set_values(rows, cols, A.nzval)
        A._nzval_ref = A.nzval
        a.values_owner[(rows, cols)] = A._token


def solve(A, b, precision=FP64, rtol=1e-10, x0=None, max_outer=30, max_inner=20000, inner_rtol=1e-4):
    """Solve A x = b on the GPU for a (FREE, FREE) matrix of assemble_matrix
    (symmetric positive definite forms).  The device holds the values of the
    last assembly on A's block; if that was another matrix (another form on
    the same space and measure) or A's values were replaced, A's own values
    are uploaded first, so the solve always uses A.  precision=MIXED: FP64
    iterative refinement around an FP32 Jacobi-preconditioned CG (inner_rtol
    per step); FP64: one FP64 Jacobi-preconditioned CG.  Stands where the
    reference solves the assembled linear problem (src/problems.jl:439-462,
    PS.solve).  Returns (x, stats); raises if rtol is not reached."""
    dev = getattr(A, "device", None)
    if dev is None:
        raise ValueError("A must come from assemble_matrix (its pattern lives on the device)")
    a, rows, cols = dev
    if (rows, cols) != (FREE, FREE):
        raise ValueError("solve needs the (FREE, FREE) block")
    _ensure_device_values(A)
    prec = L.GTK_FP64 if precision == FP64 else L.GTK_MIXED
    x, st = a.ctx.solve(prec, np.asarray(b, np.float64), x0, rtol, max_outer, max_inner, inner_rtol)
    if not st.converged:
        raise L.GtkError(f"solve did not reach rtol={rtol}: relative residual {st.rel_residual:.3e} after "
                         f"{st.outer_iterations} refinement steps, {st.inner_iterations} CG iterations")
    return x, st


def linear_problem(uhd, a_form, l_form, precision=FP64, device=0):
    """GT.PartitionedSolvers_linear_problem(uhd, a, l): the (FREE, FREE)
    matrix, and the load vector lifted by the Dirichlet values of uhd,
    b = l - A_{FREE,DIRICHLET} x_D (src/problems.jl:439-452); the lifting
    runs on the device (gtk_lift_dirichlet)."""
    V = uhd.space
    A, Ad, b = assemble_matrix_and_vector_with_free_and_dirichlet_columns(a_form, l_form, float, V, V,
                                                                          precision=precision, device=device)
    xd = np.asarray(uhd.values[V.nfree:], np.float64)
    if xd.size and Ad.nnz:
        _ensure_device_values(Ad)
        b = Ad.device[0].ctx.lift_dirichlet(xd, b)
    return A, b
