"""Discontinuous Lagrange spaces and interior-penalty skeleton terms restated
on the oracle's mesh.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product path
never imports this module.

DGSpace: GT.lagrange_space(domain, k; continuous=false).  With continuous
false every DOF is owned by its cell (reference_face_own_dofs returns all of
a cell's DOFs at d = D and none below, GalerkinToolkit.jl
src/space.jl:860-880), so generate_dof_ids_step_1 (src/space.jl:327-466)
numbers them cell by cell: local DOF a of cell c is c * nd + a + 1.  The
Dirichlet step (src/space.jl:477-535) tags, for each cell and each of its
local (D-1)-faces in the Dirichlet set, the cell's DOFs on the closure of that
local face; free DOFs keep their order as 1..nfree, tagged ones become
-1..-ndiri in order.

Skeleton: the interior-penalty terms of
GalerkinToolkitExamples/src/poisson.jl:99-102 over the interior faces
(GT.skeleton, src/mesh.jl:236-247),
    int (gamma/h) jump(v).jump(u) - jump(v).mean(grad u) - mean(grad v).jump(u) dLambda
    jump(u) = u[1] n[1] + u[2] n[2],  mean(f(u)) = (f(u[1]) + f(u[2])) / 2
(poisson.jl:40-41), h the face's diameter (largest distance between its
vertices, GT.face_diameter_field, src/field.jl:488-492) and n[i] the unit
normal of the face pointing out of side i's cell.  Each face is integrated
through its own map (its vertices in the mesh's face order, as in
boundary.py); at each point, each side's reference coordinates come from
Newton's method on that cell's geometry map and the shape functions and
gradients from the C oracle's tabulation there.  Nothing here uses local
face numbers, face permutations or reference normals.  Pure-Python loops:
small meshes only.

The volume Laplace matrix and load of a DGSpace are restated per cell from
the C oracle's quadrature and tabulation (the same sums as the C oracle's
continuous assembly, scattered through the DG DOF ids).
"""
import numpy as np

from boundary import Boundary, _basis, _DIM, _jacobian, _linear, _rule
from oracle import FREE


class Cells(Boundary):
    """Geometry helpers of an oracle Problem (after .space(order, ...)):
    node positions of every cell's Lagrange nodes, face vertex coordinates,
    and the cells around each (D-1)-face."""

    def __init__(self, p, order):
        super().__init__(p, order, 1)
        self.x = p.coords()
        self.cn = p.cell_nodes() - 1
        cf = p.cell_faces(self.D - 1) - 1
        self.cf = cf
        around = {}
        for c in range(cf.shape[0]):
            for f in cf[c]:
                around.setdefault(int(f), []).append(c)
        self.around = around
        self.interior = sorted(f for f, cs in around.items() if len(cs) == 2)

    def face_local_nodes(self, c, f):
        """Local node numbers of cell c at the Lagrange nodes of face f
        (matched by physical position)."""
        X = self.vx[self.fv[f]]
        _, tn = _basis(self.fkind, self.k, np.zeros((1, _DIM[self.fkind])))
        pos = np.array([_linear(self.fkind, t) @ X for t in tn])
        P = self.cell_pos[c]
        scale = np.abs(P).max() + 1.0
        out = []
        for y in pos:
            d = np.linalg.norm(P - y, axis=1)
            i = int(np.argmin(d))
            if d[i] > 1e-10 * scale:
                raise RuntimeError("face node without a cell node")
            out.append(i)
        return out

    def cell_ref(self, c, x):
        """Reference coordinates of the physical point x in cell c."""
        Xc = self.x[self.cn[c]]
        xi = np.full(self.D, 0.5)
        for _ in range(60):
            N, dN = self.p.tabulate(xi[None, :], geometry=True)
            r = N[0] @ Xc - x
            J = Xc.T @ dN[0]
            step = np.linalg.solve(J, r)
            xi = xi - step
            if np.abs(step).max() < 1e-15:
                break
        return xi

    def normal(self, Jf, X, c):
        """Unit normal of the face map's tangents, away from cell c."""
        if self.D == 2:
            t = Jf[:, 0]
            n = np.array([t[1], -t[0]])
        else:
            n = np.cross(Jf[:, 0], Jf[:, 1])
        n = n / np.linalg.norm(n)
        if n @ (X.mean(axis=0) - self.x[self.cn[c]].mean(axis=0)) < 0:
            n = -n
        return n

    def side(self, c, x):
        """(phi, physical gradients) of cell c's shape functions at x."""
        xi = self.cell_ref(c, x)
        phi, gref = self.p.tabulate(xi[None, :])
        _, dN = self.p.tabulate(xi[None, :], geometry=True)
        J = self.x[self.cn[c]].T @ dN[0]
        return phi[0], gref[0] @ np.linalg.inv(J)


class DGSpace:
    """DOF ids of the discontinuous space of order k on p's mesh; dirichlet:
    None, or a predicate on a face's vertex coordinates selecting the
    Dirichlet boundary faces (faces with one cell)."""

    def __init__(self, p, order, dirichlet=None):
        self.p, self.k = p, order
        self.geo = Cells(p, order)
        ncell = self.geo.cn.shape[0]
        nd = self.geo.cell_pos.shape[1]
        self.nd = nd
        ids = np.arange(ncell * nd).reshape(ncell, nd) + 1  # step_1: cell by cell
        tag = np.zeros(ncell * nd + 1, bool)
        if dirichlet is not None:
            g = self.geo
            for c in range(ncell):
                for f in g.cf[c]:
                    f = int(f)
                    if len(g.around[f]) != 1 or not all(dirichlet(g.vx[v]) for v in g.fv[f]):
                        continue
                    for a in g.face_local_nodes(c, f):
                        tag[ids[c, a]] = True
        tag = tag[1:]
        new = np.zeros(ncell * nd, np.int64)
        free, diri = np.nonzero(~tag)[0], np.nonzero(tag)[0]
        new[free] = np.arange(1, len(free) + 1)
        new[diri] = -np.arange(1, len(diri) + 1)
        self.cd = new[ids - 1].astype(np.int32)
        self.nfree, self.ndiri = len(free), len(diri)

    def laplace(self, degree, rows=FREE, cols=FREE):
        """Volume int grad(u).grad(v) as scipy CSC (duplicates summed)."""
        import scipy.sparse as sp
        g, p = self.geo, self.p
        xq, wq = p.quadrature(degree)
        phi, gref = p.tabulate(xq)
        _, dN = p.tabulate(xq, geometry=True)
        I, J, V = [], [], []
        for c in range(g.cn.shape[0]):
            Xc = g.x[g.cn[c]]
            Ae = np.zeros((self.nd, self.nd))
            for q in range(len(wq)):
                Jq = Xc.T @ dN[q]
                gr = gref[q] @ np.linalg.inv(Jq)
                Ae += abs(np.linalg.det(Jq)) * wq[q] * (gr @ gr.T)
            _scatter(I, J, V, Ae, self.cd[c], self.cd[c], rows, cols)
        return _csc(I, J, V, self._n(rows), self._n(cols))

    def load(self, degree, f, rows=FREE):
        """int f v dOmega, f a function of the physical point."""
        g, p = self.geo, self.p
        xq, wq = p.quadrature(degree)
        phi, _ = p.tabulate(xq)
        N, dN = p.tabulate(xq, geometry=True)
        b = np.zeros(self._n(rows))
        for c in range(g.cn.shape[0]):
            Xc = g.x[g.cn[c]]
            for q in range(len(wq)):
                Jq = Xc.T @ dN[q]
                dV = abs(np.linalg.det(Jq)) * wq[q]
                fv = f(N[q] @ Xc)
                for a in range(self.nd):
                    r = Boundary._index(self.cd[c, a], rows)
                    if r >= 0:
                        b[r] += dV * fv * phi[q, a]
        return b

    def _n(self, which):
        return self.nfree if which == FREE else self.ndiri


def _scatter(I, J, V, Ae, rdofs, cdofs, rows, cols):
    for b in range(len(cdofs)):
        cb = Boundary._index(cdofs[b], cols)
        if cb < 0:
            continue
        for a in range(len(rdofs)):
            ra = Boundary._index(rdofs[a], rows)
            if ra >= 0:
                I.append(ra)
                J.append(cb)
                V.append(Ae[a, b])


def _csc(I, J, V, m, n):
    import scipy.sparse as sp
    A = sp.coo_matrix((V, (I, J)), shape=(m, n)).tocsc()
    A.sum_duplicates()
    A.sort_indices()
    return A


class Skeleton:
    """Interior-penalty terms over the interior faces of p's mesh for a
    space given by its cell DOF ids cd (continuous: p.cell_dofs(); DG:
    DGSpace.cd)."""

    def __init__(self, p, order, cd):
        self.geo = Cells(p, order)
        self.cd = np.asarray(cd)

    def face_blocks(self, f, degree, gamma, terms=(1, 1, 1)):
        """(cells (c1, c2), local matrix over [c1 DOFs, c2 DOFs]); terms
        scales (penalty, jump(v).mean(grad u), mean(grad v).jump(u))."""
        g = self.geo
        c1, c2 = sorted(g.around[f])
        X = g.vx[g.fv[f]]
        h = max(np.linalg.norm(a - b) for a in X for b in X)
        beta = gamma / h
        t, w = _rule(g.fkind, degree)
        nd = self.cd.shape[1]
        A = np.zeros((2 * nd, 2 * nd))
        for q in range(len(w)):
            x = _linear(g.fkind, t[q]) @ X
            Jf = _jacobian(g.fkind, X, t[q])
            dS = np.sqrt(np.linalg.det(Jf.T @ Jf)) * w[q]
            sides = []
            for c in (c1, c2):
                phi, grad = g.side(c, x)
                sides.append((phi, grad, g.normal(Jf, X, c)))
            for i, (pi, gi, ni) in enumerate(sides):
                for j, (pj, gj, nj) in enumerate(sides):
                    # rows: test v from side i; columns: trial u from side j
                    pen = (ni @ nj) * np.outer(pi, pj)
                    jm = 0.5 * np.outer(pi, gj @ ni)   # jump(v) . mean(grad u)
                    mj = 0.5 * np.outer(gi @ nj, pj)   # mean(grad v) . jump(u)
                    A[i * nd:(i + 1) * nd, j * nd:(j + 1) * nd] += dS * (
                        terms[0] * beta * pen - terms[1] * jm - terms[2] * mj)
        return (c1, c2), A

    def matrix(self, degree, gamma, rows=FREE, cols=FREE, shape=None, terms=(1, 1, 1)):
        """Sum over the interior faces, scipy CSC, duplicates summed."""
        I, J, V = [], [], []
        for f in self.geo.interior:
            (c1, c2), A = self.face_blocks(f, degree, gamma, terms)
            dofs = np.concatenate([self.cd[c1], self.cd[c2]])
            _scatter(I, J, V, A, dofs, dofs, rows, cols)
        return _csc(I, J, V, shape[0], shape[1])
