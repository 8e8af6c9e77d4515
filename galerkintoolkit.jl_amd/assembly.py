"""GT-style assembly entry points backed by libgtk_hip.

Mirrors GT.measure (GalerkinToolkit.jl src/problems.jl:4-25),
GT.assemble_matrix / GT.assemble_vector (src/problems.jl:244-274, :319-350),
GT.update_matrix! (src/problems.jl:352-361) and
assemble_matrix_with_free_and_dirichlet_columns (src/problems.jl:363-380).
The reference takes a Julia lambda and compiles it; here the weak form is
one of the integrands the library implements (Laplace with a per-cell
coefficient, mass, isotropic elasticity, and the load f . v), passed as a
form object.  The result is the reference's SparseMatrixCSC{Float64,Int32}
layout (1-based colptr/rowval) and a Float64 vector.
"""
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import refcell as R
from .space import DIRICHLET, FREE

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
        return self.domain.dim == self.domain.mesh.D - 1


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


class _Assembler:
    """The device context of one (space, measure) pair."""

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


def _assembler(space, dOmega, device):
    key = (id(dOmega), device)
    cache = space.__dict__.setdefault("_assemblers", {})
    a = cache.get(key)
    if a is None or a.dOmega is not dOmega:
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
    raise TypeError("unsupported bilinear form")


def _ncoef_items(form, space):
    """Cells (or boundary items) the form's coefficients run over."""
    dom = form.dOmega.domain
    return len(dom.faces) if form.dOmega.on_boundary else space.mesh.cells.shape[0]


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


def assemble_matrix(form, T, trial_space, test_space, free_or_dirichlet=(FREE, FREE), precision=FP64, reuse=False,
                    device=0, symbolic="host"):
    """GT.assemble_matrix(a, Float64, U, V; free_or_dirichlet) on the GPU.
    symbolic: where the pattern is built the first time ("host" or
    "device")."""
    if T not in (float, np.float64):
        raise TypeError("values are Float64")
    if trial_space is not test_space:
        raise NotImplementedError("trial and test spaces must be the same space")
    space = test_space
    if isinstance(form, FormSum):
        if reuse:
            raise NotImplementedError("reuse of a sum of integrals")
        return _assemble_matrix_sum(form, T, space, free_or_dirichlet, precision, device, symbolic)
    if form.dOmega.on_boundary and not isinstance(form, Mass):
        raise NotImplementedError("boundary integrals of matrices: the mass form (Robin terms)")
    a = _assembler(space, form.dOmega, device)
    rows, cols = free_or_dirichlet
    nnz, ncols, nrows, colptr, rowval = a.pattern(rows, cols, symbolic)
    coef, ncoef = _coefficients(form, _ncoef_items(form, space))
    a.ctx.set_cell_coefficients(coef, ncoef)
    prec = L.GTK_FP64 if precision == FP64 else L.GTK_MIXED
    nzval, stats = a.ctx.assemble_matrix(_form_code(form), prec, rows, cols, nnz)
    A = SparseMatrixCSC(nrows, ncols, colptr, rowval, nzval)
    A.stats = stats
    A.device = (a, rows, cols)
    if reuse:
        return A, (a, form, rows, cols, prec)
    return A


def update_matrix(A, cache, form=None):
    """GT.update_matrix!(A, cache): refill values with the same pattern
    (new coefficients or coordinates); the pattern is not rebuilt."""
    a, form0, rows, cols, prec = cache
    form = form0 if form is None else form
    coef, ncoef = _coefficients(form, _ncoef_items(form, a.space))
    a.ctx.set_cell_coefficients(coef, ncoef)
    A.nzval, A.stats = a.ctx.assemble_matrix(_form_code(form), prec, rows, cols, A.nnz)
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


def _assemble_matrix_sum(fs, T, space, free_or_dirichlet, precision, device, symbolic):
    """One volume integral plus boundary mass integrals: the volume block is
    assembled, then each boundary block is added into its values on the
    device (the boundary pattern lies inside the volume pattern)."""
    vol = [t for t in fs.terms if not t.dOmega.on_boundary]
    bnd = [t for t in fs.terms if t.dOmega.on_boundary]
    if len(vol) != 1:
        raise NotImplementedError("a sum of matrix integrals: one over the cells plus boundary integrals")
    rows, cols = free_or_dirichlet
    A = assemble_matrix(vol[0], T, space, space, free_or_dirichlet, precision=precision, device=device,
                        symbolic=symbolic)
    a_vol = A.device[0]
    for t in bnd:
        Ab = assemble_matrix(t, T, space, space, free_or_dirichlet, precision=precision, device=device)
        a_vol.ctx.add_block(rows, cols, Ab.device[0].ctx)
    A.nzval = a_vol.ctx.get_values(rows, cols, A.nnz)
    return A


def assemble_vector(form, T, space, free_or_dirichlet=FREE, device=0):
    """GT.assemble_vector(l, Float64, V; free_or_dirichlet) on the GPU; l may
    be a sum of loads over cells and boundary faces."""
    if isinstance(form, FormSum):
        return sum(assemble_vector(t, T, space, free_or_dirichlet, device) for t in form.terms)
    kind, data = _load_data(form, space)
    a = _assembler(space, form.dOmega, device)
    n = space.nfree if free_or_dirichlet == FREE else space.ndiri
    b, stats = a.ctx.assemble_vector(kind, data, free_or_dirichlet, n)
    return b


def assemble_matrix_with_free_and_dirichlet_columns(form, T, U, V, precision=FP64, device=0):
    A = assemble_matrix(form, T, U, V, (FREE, FREE), precision=precision, device=device)
    Ad = assemble_matrix(form, T, U, V, (FREE, DIRICHLET), precision=precision, device=device)
    return A, Ad


def solve(A, b, precision=FP64, rtol=1e-10, x0=None, max_outer=30, max_inner=20000, inner_rtol=1e-4):
    """Solve A x = b on the GPU for a (FREE, FREE) matrix of assemble_matrix
    (symmetric positive definite forms), with the values of its last device
    assembly.  precision=MIXED: FP64 iterative refinement around an FP32
    Jacobi-preconditioned CG (inner_rtol per step); FP64: one FP64
    Jacobi-preconditioned CG.  Stands where the reference solves the
    assembled linear problem (src/problems.jl:439-462, PS.solve).
    Returns (x, stats); raises if rtol is not reached."""
    dev = getattr(A, "device", None)
    if dev is None:
        raise ValueError("A must come from assemble_matrix (its values live on the device)")
    a, rows, cols = dev
    if (rows, cols) != (FREE, FREE):
        raise ValueError("solve needs the (FREE, FREE) block")
    prec = L.GTK_FP64 if precision == FP64 else L.GTK_MIXED
    x, st = a.ctx.solve(prec, np.asarray(b, np.float64), x0, rtol, max_outer, max_inner, inner_rtol)
    if not st.converged:
        raise L.GtkError(f"solve did not reach rtol={rtol}: relative residual {st.rel_residual:.3e} after "
                         f"{st.outer_iterations} refinement steps, {st.inner_iterations} CG iterations")
    return x, st


def linear_problem(uhd, a_form, l_form, precision=FP64, device=0):
    """GT.PartitionedSolvers_linear_problem(uhd, a, l): the (FREE, FREE)
    matrix, and the load vector lifted by the Dirichlet values of uhd,
    b = l - A_{FREE,DIRICHLET} x_D (src/problems.jl:439-452)."""
    V = uhd.space
    A, Ad = assemble_matrix_with_free_and_dirichlet_columns(a_form, float, V, V, precision=precision, device=device)
    b = assemble_vector(l_form, float, V, device=device)
    xd = np.asarray(uhd.values[V.nfree:], np.float64)
    if xd.size and Ad.nnz:
        b = b - Ad.to_scipy() @ xd
    return A, b
