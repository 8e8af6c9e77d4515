"""Boundary integrals restated on the oracle's mesh (oracle/gtk_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product path
never imports this module.

Each boundary face is integrated through its own map: its vertices in the
mesh's face order (complexify, restated in the C oracle) and the order-1
shape functions of the face's reference cell; the weight is
sqrt(det(Jf' Jf)) w, the reference's weight for a face map
(GalerkinToolkit.jl src/accessors.jl:1000-1007) under
GT.measure(GT.boundary(mesh), degree) (src/problems.jl:4-25, boundary terms
of the generated loops src/compiler.jl:513-529).  Shape functions on the face
are the face's own Lagrange element (equispaced nodes e/k, x fastest,
src/space.jl:981-1251), evaluated at the face's rule; its nodes are matched
to the adjacent cell's DOFs by physical position, so nothing here depends on
how a face sits inside its cell.  Pure-Python loops: small meshes only.
"""
import numpy as np

from oracle import FREE, HEX, Q_CELLKIND, Q_NFREE, QUAD, TET, TRI

SEG = 1
_FACE_KIND = {QUAD: SEG, HEX: QUAD, TRI: SEG, TET: TRI}
_SIMPLEX = {SEG: True, QUAD: False, TRI: True}
_DIM = {SEG: 1, QUAD: 2, TRI: 2}


def _exponents(dim, k, simplex):
    e = np.indices([k + 1] * dim).reshape(dim, -1)[::-1].T  # x fastest
    if simplex:
        e = e[e.sum(axis=1) <= k]
    return e


def _basis(kind, k, t):
    """Lagrange basis of order k on the face kind at points t (n, dim): the
    monomials of the lattice exponents through the inverse Vandermonde."""
    dim = _DIM[kind]
    e = _exponents(dim, k, _SIMPLEX[kind] and kind != SEG)
    nodes = e / float(k)
    V = np.prod(nodes[:, None, :] ** e[None, :, :], axis=2)
    M = np.prod(np.atleast_2d(t)[:, None, :] ** e[None, :, :], axis=2)
    return M @ np.linalg.inv(V), nodes


def _linear_grad(kind):
    """Gradients of the order-1 shape functions of the face kind (constant
    for a segment and a triangle; a square face's at its corner (0,0) is
    enough for the parallelogram faces of a box, where it is constant)."""
    if kind == SEG:
        return np.array([[-1.0], [1.0]])
    if kind == TRI:
        return np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]])
    return None


def _jacobian(kind, X, t):
    """Face map gradient (D, dim) at t for vertex coordinates X (nv, D)."""
    if kind == QUAD:
        s, r = t
        g = np.array([[-(1 - r), -(1 - s)], [(1 - r), -s], [-r, (1 - s)], [r, s]])
        return X.T @ g
    return X.T @ _linear_grad(kind)


def _linear(kind, t):
    t = np.atleast_1d(t)
    if kind == SEG:
        return np.array([1 - t[0], t[0]])
    if kind == TRI:
        return np.array([1 - t[0] - t[1], t[0], t[1]])
    s, r = t
    return np.array([(1 - s) * (1 - r), s * (1 - r), (1 - s) * r, s * r])


def _rule(kind, degree):
    """Gauss-Legendre on [0,1] with (degree+2)//2 points per direction
    (tensor on the square), collapsed Gauss-Jacobi on the triangle with
    ceil((degree+1)/2) per direction (src/quadrature.jl:36-157)."""
    if kind in (SEG, QUAD):
        n = (degree + 2) // 2
        x, w = np.polynomial.legendre.leggauss(n)
        x, w = 0.5 * x + 0.5, 0.5 * w
        if kind == SEG:
            return x[:, None], w
        X = np.array([[x[i], x[j]] for j in range(n) for i in range(n)])
        W = np.array([w[i] * w[j] for j in range(n) for i in range(n)])
        return X, W
    from scipy.special import roots_jacobi
    n = int(np.ceil((degree + 1) / 2.0))
    xj, wj = roots_jacobi(n, 1, 0)
    a, wa = 0.5 * xj + 0.5, 0.25 * wj
    xg, wg = np.polynomial.legendre.leggauss(n)
    b, wb = 0.5 * xg + 0.5, 0.5 * wg
    pts, wts = [], []
    for j in range(n):
        for i in range(n):
            pts.append([a[i], (1 - a[i]) * b[j]])
            wts.append(wa[i] * wb[j])
    return np.array(pts), np.array(wts)


class Boundary:
    """Boundary faces of an oracle Problem (after .space(order, ncomp, ...)),
    optionally only those whose vertices all satisfy select(x)."""

    def __init__(self, p, order, ncomp=1, select=None):
        D = p.D
        kind = p.q(Q_CELLKIND)
        self.p, self.D, self.k, self.nc = p, D, order, ncomp
        self.fkind = _FACE_KIND[kind]
        # the oracle's ids are 1-based; 0-based below
        x = p.coords()
        nv = p.node_vertex()  # vertex id of each node (0: not a vertex)
        vnode = np.zeros(nv.max(), np.int64)
        vnode[nv[nv > 0] - 1] = np.nonzero(nv > 0)[0]
        self.vx = x[vnode]  # vertex coordinates
        cf = p.cell_faces(D - 1) - 1
        cnt = np.bincount(cf.ravel(), minlength=p.nfaces(D - 1))
        fv = p.face_vertices(D - 1) - 1
        faces = [f for f in range(p.nfaces(D - 1)) if cnt[f] == 1]
        if select is not None:
            faces = [f for f in faces if all(select(self.vx[v]) for v in fv[f])]
        self.faces = faces
        owner = {}
        for c in range(cf.shape[0]):
            for f in cf[c]:
                owner.setdefault(int(f), c)
        self.owner = owner
        self.fv = fv
        # DOF node positions of every cell (lattice e/k through the cell map)
        cd = p.cell_dofs()
        self.cd = cd
        cn = p.cell_nodes() - 1
        simplex = kind in (TRI, TET)
        e = _exponents(D, order, simplex)
        val, _ = p.tabulate(e / float(order), geometry=True)
        self.cell_pos = np.einsum("nv,cvd->cnd", val, x[cn])  # (cells, nodes, D)
        self.nfree = p.q(Q_NFREE)

    def _face_dofs(self, f):
        """Cell DOFs (one per component) of the face's Lagrange nodes, and
        the face's vertex coordinates."""
        c = self.owner[f]
        X = self.vx[self.fv[f]]
        _, tn = _basis(self.fkind, self.k, np.zeros((1, _DIM[self.fkind])))
        pos = np.array([_linear(self.fkind, t) @ X for t in tn])
        P = self.cell_pos[c]
        scale = np.abs(P).max() + 1.0
        nodes = []
        for y in pos:
            d = np.linalg.norm(P - y, axis=1)
            i = int(np.argmin(d))
            if d[i] > 1e-10 * scale:
                raise RuntimeError("face node without a cell DOF")
            nodes.append(i)
        dofs = np.array([[self.cd[c, n * self.nc + comp] for comp in range(self.nc)] for n in nodes])
        return dofs, X

    def _points(self, f, X, degree):
        t, w = _rule(self.fkind, degree)
        phi, _ = _basis(self.fkind, self.k, t)
        dG = []
        for q in range(len(w)):
            J = _jacobian(self.fkind, X, t[q])
            dG.append(np.sqrt(np.linalg.det(J.T @ J)) * w[q])
        return phi, np.array(dG)

    @staticmethod
    def _index(dof, rows):
        if rows == FREE:
            return dof - 1 if dof > 0 else -1
        return -dof - 1 if dof < 0 else -1

    def vector(self, degree, f_kind=0, fdata=(1.0,), rows=FREE, n=None):
        """b_i = sum over faces of int f . v_i; f_kind 0: one value per
        component, 1: per face (in self.faces order) and component, 2: nodal
        (free then Dirichlet values)."""
        fdata = np.asarray(fdata, np.float64)
        b = np.zeros(n)
        for fi, f in enumerate(self.faces):
            dofs, X = self._face_dofs(f)
            phi, dG = self._points(f, X, degree)
            for comp in range(self.nc):
                if f_kind == 0:
                    g = np.full(len(dG), fdata[comp])
                elif f_kind == 1:
                    g = np.full(len(dG), fdata[fi * self.nc + comp])
                else:
                    vals = np.array([fdata[d - 1] if d > 0 else fdata[self.nfree - d - 1] for d in dofs[:, comp]])
                    g = phi @ vals
                for a in range(dofs.shape[0]):
                    r = self._index(dofs[a, comp], rows)
                    if r >= 0:
                        b[r] += np.sum(g * phi[:, a] * dG)
        return b

    def mass(self, degree, coef=None, rows=FREE, cols=FREE, shape=None):
        """sum over faces of int kappa u v (kappa per face, self.faces
        order), as scipy CSC with duplicates summed."""
        import scipy.sparse as sp
        I, J, V = [], [], []
        for fi, f in enumerate(self.faces):
            dofs, X = self._face_dofs(f)
            phi, dG = self._points(f, X, degree)
            k = 1.0 if coef is None else coef[fi]
            for comp in range(self.nc):
                for bb in range(dofs.shape[0]):
                    cidx = self._index(dofs[bb, comp], cols)
                    if cidx < 0:
                        continue
                    for a in range(dofs.shape[0]):
                        ridx = self._index(dofs[a, comp], rows)
                        if ridx >= 0:
                            I.append(ridx)
                            J.append(cidx)
                            V.append(k * np.sum(phi[:, a] * phi[:, bb] * dG))
        A = sp.coo_matrix((V, (I, J)), shape=shape).tocsc()
        A.sum_duplicates()
        A.sort_indices()
        return A
