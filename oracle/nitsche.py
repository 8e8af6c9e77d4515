"""Nitsche boundary terms restated on the oracle's mesh.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product path
never imports this module.

The terms of GalerkinToolkitExamples/src/poisson.jl:93-112 over boundary
faces (scalar fields):
    a(u, v) = int (gamma/h) v u - v n.grad(u) - n.grad(v) u dGamma
    l(v)    = int (gamma/h) v g - n.grad(v) g dGamma
with h the face's diameter, the largest distance between its nodes
(GT.face_diameter_field, src/field.jl:488-492, diameter src/accessors.jl:907-921),
and n the outward unit normal (src/accessors.jl:1009-1030).

Each face is integrated through its own map (its vertices in the mesh's face
order and the order-1 shape functions of the face's reference cell), as in
boundary.py.  The normal is the face's own: orthogonal to the face map's
tangents, pointing away from the cell's centroid.  At every point the cell's
reference coordinates are found by Newton's method on the cell's geometry
map, and the cell's shape functions, their reference gradients and the
Jacobian come from the C oracle's tabulation there; physical gradients are
J^-T grad_ref.  Nothing here uses local face numbers, face tables or
reference normals.  Pure-Python loops: small meshes only.
"""
import numpy as np

from boundary import Boundary, _jacobian, _linear, _rule


class NitscheFaces(Boundary):
    """Boundary faces of an oracle Problem with a scalar space (after
    .space(order, ...)), optionally only those whose vertices all satisfy
    select(x)."""

    def __init__(self, p, order, select=None):
        super().__init__(p, order, 1, select)
        self.x = p.coords()
        self.cn = p.cell_nodes() - 1

    def _cell_ref(self, c, x):
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

    def _normal(self, Jf, X, c):
        D = self.D
        if D == 2:
            t = Jf[:, 0]
            n = np.array([t[1], -t[0]])
        else:
            n = np.cross(Jf[:, 0], Jf[:, 1])
        n = n / np.linalg.norm(n)
        if n @ (X.mean(axis=0) - self.x[self.cn[c]].mean(axis=0)) < 0:
            n = -n
        return n

    def face_points(self, f, degree):
        """(cell, cell DOFs, h, [(phi, n.grad phi, dGamma, x) per point])."""
        c = self.owner[f]
        X = self.vx[self.fv[f]]
        h = max(np.linalg.norm(a - b) for a in X for b in X)
        t, w = _rule(self.fkind, degree)
        pts = []
        Xc = self.x[self.cn[c]]
        for q in range(len(w)):
            x = _linear(self.fkind, t[q]) @ X
            Jf = _jacobian(self.fkind, X, t[q])
            dG = np.sqrt(np.linalg.det(Jf.T @ Jf)) * w[q]
            n = self._normal(Jf, X, c)
            xi = self._cell_ref(c, x)
            phi, gref = self.p.tabulate(xi[None, :])
            _, dN = self.p.tabulate(xi[None, :], geometry=True)
            J = Xc.T @ dN[0]
            grad = gref[0] @ np.linalg.inv(J)  # rows: J^-T grad_ref
            pts.append((phi[0], grad @ n, dG, x))
        return c, self.cd[c], h, pts

    def matrix(self, degree, gamma, rows=1, cols=1, shape=None, load=False):
        """sum over faces of the Nitsche matrix (load=True: the load's
        matrix, int (gamma/h) v u - n.grad(v) u), scipy CSC, duplicates summed
        in face order."""
        import scipy.sparse as sp
        I, J, V = [], [], []
        for f in self.faces:
            c, dofs, h, pts = self.face_points(f, degree)
            beta = gamma / h
            nd = len(dofs)
            Ae = np.zeros((nd, nd))
            for phi, dn, dG, _ in pts:
                M = beta * np.outer(phi, phi) - np.outer(dn, phi)  # row a (test), column b (trial)
                if not load:
                    M -= np.outer(phi, dn)
                Ae += dG * M
            for b in range(nd):
                cb = self._index(dofs[b], cols)
                if cb < 0:
                    continue
                for a in range(nd):
                    ra = self._index(dofs[a], rows)
                    if ra >= 0:
                        I.append(ra)
                        J.append(cb)
                        V.append(Ae[a, b])
        A = sp.coo_matrix((V, (I, J)), shape=shape).tocsc()
        A.sum_duplicates()
        A.sort_indices()
        return A

    def load(self, degree, gamma, g, rows=1, n=None):
        """int (gamma/h) v g - n.grad(v) g with g a function of the physical
        point, evaluated at the face rule's points."""
        b = np.zeros(n)
        for f in self.faces:
            c, dofs, h, pts = self.face_points(f, degree)
            beta = gamma / h
            for phi, dn, dG, x in pts:
                gv = g(x)
                for a in range(len(dofs)):
                    r = self._index(dofs[a], rows)
                    if r >= 0:
                        b[r] += dG * (beta * phi[a] - dn[a]) * gv
        return b
