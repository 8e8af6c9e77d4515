"""Reference cells, Lagrange elements and quadrature rules (host side).

Reference geometry and local face numbering follow GalerkinToolkit.jl
src/domain.jl:213-268 (unit square / cube) and :425-470 (unit triangle /
tetrahedron); the simplex split of a cube is src/domain.jl:322-336.
Lagrange elements follow src/space.jl:981-1251 (lexicographic nodes e/k,
P_k keeps sum(e) <= k); their local face nodes, interior ("own") nodes and
interior-node permutations follow complexify(::LagrangeFaceSpace)
(src/space.jl:1598-1640) and node_permutations_from_mesh_face
(src/space.jl:1447-1494).  Shape functions of Q_k are evaluated as tensor
products of 1D Lagrange polynomials; P_k through the Vandermonde inverse.
"""
import itertools
from dataclasses import dataclass, field

import numpy as np

POINT, SEG, QUAD, HEX, TRI, TET = range(6)


@dataclass
class Geo:
    dim: int
    verts: list
    faces: dict  # d -> list of local vertex lists (1-based)
    ftype: dict  # d -> geometry kind of the d-faces
    simplex: bool


GEO = {
    POINT: Geo(0, [()], {0: [[1]]}, {0: POINT}, True),
    SEG: Geo(1, [(0.0,), (1.0,)], {0: [[1], [2]], 1: [[1, 2]]}, {0: POINT, 1: SEG}, True),
    QUAD: Geo(2, [(0, 0), (1, 0), (0, 1), (1, 1)],
              {0: [[1], [2], [3], [4]], 1: [[1, 2], [3, 4], [1, 3], [2, 4]], 2: [[1, 2, 3, 4]]},
              {0: POINT, 1: SEG, 2: QUAD}, False),
    HEX: Geo(3, [(i & 1, (i >> 1) & 1, (i >> 2) & 1) for i in range(8)],
             {0: [[i] for i in range(1, 9)],
              1: [[1, 2], [3, 4], [1, 3], [2, 4], [5, 6], [7, 8], [5, 7], [6, 8], [1, 5], [3, 7], [2, 6], [4, 8]],
              2: [[1, 2, 3, 4], [5, 6, 7, 8], [1, 2, 5, 6], [3, 4, 7, 8], [1, 3, 5, 7], [2, 4, 6, 8]],
              3: [list(range(1, 9))]},
             {0: POINT, 1: SEG, 2: QUAD, 3: HEX}, False),
    TRI: Geo(2, [(0, 0), (1, 0), (0, 1)],
             {0: [[1], [2], [3]], 1: [[1, 2], [1, 3], [2, 3]], 2: [[1, 2, 3]]},
             {0: POINT, 1: SEG, 2: TRI}, True),
    TET: Geo(3, [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
             {0: [[1], [2], [3], [4]], 1: [[1, 2], [1, 3], [2, 3], [1, 4], [2, 4], [3, 4]],
              2: [[1, 2, 3], [1, 2, 4], [1, 3, 4], [2, 3, 4]], 3: [[1, 2, 3, 4]]},
             {0: POINT, 1: SEG, 2: TRI, 3: TET}, True),
}

# simplex_nodes(::UnitNCube), src/domain.jl:322-336
SIMPLEX_NODES = {2: [[1, 2, 3], [4, 3, 2]],
                 3: [[1, 2, 3, 7], [1, 2, 5, 7], [2, 3, 4, 7], [2, 4, 7, 8], [2, 5, 6, 7], [2, 6, 7, 8]]}


def is_simplex(kind):
    return GEO[kind].simplex


def cube_of(kind, D):
    return HEX if D == 3 else QUAD


def linear_shape(kind, xi):
    """Order-1 shape functions (lexicographic {0,1}^D for n-cubes,
    barycentric for simplices) at the points xi (n, dim)."""
    xi = np.atleast_2d(np.asarray(xi, np.float64))
    n = xi.shape[0]
    if kind == POINT:
        return np.ones((n, 1))
    if GEO[kind].simplex:
        return np.concatenate([1.0 - xi.sum(axis=1, keepdims=True), xi], axis=1)
    D = GEO[kind].dim
    N = np.ones((n, 1 << D))
    for l in range(1 << D):
        for d in range(D):
            N[:, l] *= xi[:, d] if (l >> d) & 1 else 1.0 - xi[:, d]
    return N


def face_normals(kind):
    """Outward unit normal of each local (D-1)-face of the reference cell, in
    the local face order of GEO (the lists of src/domain.jl:226, :258 for the
    unit square and cube, :428, :460 for the simplices): orthogonal to the
    face's tangents, pointing away from the cell's centroid."""
    g = GEO[kind]
    D = g.dim
    V = np.asarray(g.verts, np.float64)
    cen = V.mean(axis=0)
    out = []
    for fv in g.faces[D - 1]:
        X = V[np.asarray(fv) - 1]
        T = (X[1:] - X[0]).T  # D x (number of face vertices - 1) tangents
        if D == 2:
            n = np.array([T[1, 0], -T[0, 0]])
        else:
            n = np.cross(T[:, 0], T[:, 1])
        n = n / np.linalg.norm(n)
        if n @ (X.mean(axis=0) - cen) < 0:
            n = -n
        out.append(n)
    return np.array(out)


def vertex_permutations(kind):
    """Admissible vertex permutations (src/domain.jl:53-100), 1-based:
    identity for D == 0 and D > 2, all permutations (lexicographic) for
    simplices and segments, the volume-preserving ones for squares."""
    g = GEO[kind]
    nv = len(g.verts)
    if g.dim == 0 or g.dim > 2:
        return [list(range(1, nv + 1))]
    perms = [list(p) for p in itertools.permutations(range(1, nv + 1))]
    if g.simplex:
        return perms
    V = np.asarray(g.verts, np.float64)
    gN = np.array([[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.5], [0.5, 0.5]])  # at the centre
    ref = abs(np.linalg.det(V.T @ gN))
    out = []
    for p in perms:
        vol = abs(np.linalg.det(V[np.asarray(p) - 1].T @ gN))
        if np.isclose(ref + vol, 2 * ref, rtol=1.4901161193847656e-8, atol=0):
            out.append(p)
    return out


def monomial_exponents(kind, k):
    D = GEO[kind].dim
    if D == 0:
        return np.zeros((1, 0), np.int64)
    e = np.indices([k + 1] * D).reshape(D, -1)[::-1].T  # x fastest
    if GEO[kind].simplex:
        e = e[e.sum(axis=1) <= k]
    return np.ascontiguousarray(e)


def _on_boundary(kind, k, e):
    g = GEO[kind]
    if g.dim == 0:
        return np.zeros(len(e), bool)
    b = np.any(e == 0, axis=1)
    if g.simplex:
        b |= e.sum(axis=1) == k
    else:
        b |= np.any(e == k, axis=1)
    return b


@dataclass
class Lagrange:
    """Reference Lagrange element of `kind`, order k, ncomp components
    (component-major local dofs, src/space.jl:1257-1286)."""
    kind: int
    order: int
    ncomp: int = 1
    expo: np.ndarray = field(init=False)
    nodes: np.ndarray = field(init=False)
    face_nodes: dict = field(init=False)   # d -> (nlf, n) 1-based cell nodes
    own_nodes: dict = field(init=False)    # d -> (nlf, n_own)
    own_perms: dict = field(init=False)    # d -> (np, n_own) 1-based node_to_pnode

    def __post_init__(self):
        g = GEO[self.kind]
        k = self.order
        self.dim = g.dim
        self.expo = monomial_exponents(self.kind, k)
        self.nodes = self.expo / float(k)
        self.nnodes = len(self.expo)
        self.ndofs = self.nnodes * self.ncomp
        keys = {tuple(int(round(k * x)) for x in row): i + 1 for i, row in enumerate(self.nodes)}
        self.face_nodes, self.own_nodes, self.own_perms = {}, {}, {}
        for d in range(g.dim):
            ft = g.ftype[d]
            fe = monomial_exponents(ft, k)
            fxi = fe / float(k) if GEO[ft].dim else np.zeros((1, 0))
            N = linear_shape(ft, fxi)
            own = np.nonzero(~_on_boundary(ft, k, fe))[0]
            fn = []
            for f in g.faces[d]:
                X = np.asarray([g.verts[v - 1] for v in f], np.float64)
                x = N @ X
                fn.append([keys[tuple(int(round(k * c)) for c in row)] for row in x])
            fn = np.asarray(fn, np.int64)
            self.face_nodes[d] = fn
            self.own_nodes[d] = fn[:, own]
            perms = vertex_permutations(ft)
            P = np.zeros((len(perms), len(own)), np.int64)
            if len(perms) == 1:
                P[0] = np.arange(1, len(own) + 1)
            else:
                Vf = np.asarray(GEO[ft].verts, np.float64)
                q = fxi[own]
                Nq = linear_shape(ft, q)
                for pi, p in enumerate(perms):
                    y = Nq @ Vf[np.asarray(p) - 1]
                    for iq in range(len(own)):
                        dist = np.linalg.norm(q - y[iq], axis=1)
                        hit = np.nonzero(dist <= 1.4901161193847656e-8 * (1 + dist))[0]
                        P[pi, iq] = hit[0] + 1 if len(hit) else 0
            self.own_perms[d] = P
        interior = np.nonzero(~_on_boundary(self.kind, k, self.expo))[0] + 1
        self.face_nodes[g.dim] = np.arange(1, self.nnodes + 1)[None, :]
        self.own_nodes[g.dim] = interior[None, :]
        self.own_perms[g.dim] = np.arange(1, len(interior) + 1)[None, :]
        if GEO[self.kind].simplex:
            A = np.prod(self.nodes[:, None, :] ** self.expo[None, :, :], axis=2)
            self._B = np.linalg.solve(A, np.eye(self.nnodes))

    def tabulate(self, x):
        """Values (n, nnodes) and reference gradients (n, nnodes, dim)."""
        x = np.atleast_2d(np.asarray(x, np.float64))
        D, k = self.dim, self.order
        if not GEO[self.kind].simplex:
            t = np.linspace(0.0, 1.0, k + 1)
            L, dL = _lagrange_1d(t, x)  # (n, D, k+1)
            val = np.ones((x.shape[0], self.nnodes))
            grad = np.ones((x.shape[0], self.nnodes, D))
            for d in range(D):
                ed = self.expo[:, d]
                val *= L[:, d, ed]
                for dd in range(D):
                    grad[:, :, dd] *= (dL if dd == d else L)[:, d, ed]
            return val, grad
        M = np.prod(x[:, None, :] ** self.expo[None, :, :], axis=2)
        G = np.zeros((x.shape[0], self.nnodes, D))
        for d in range(D):
            e = self.expo.copy()
            c = e[:, d].astype(np.float64)
            e[:, d] = np.maximum(e[:, d] - 1, 0)
            G[:, :, d] = c[None, :] * np.prod(x[:, None, :] ** e[None, :, :], axis=2)
        val = M @ self._B
        grad = np.einsum("nmd,mk->nkd", G, self._B)
        return val, grad


def _lagrange_1d(t, x):
    """1D Lagrange basis on nodes t and its derivative at every coordinate
    of x: returns (n, D, len(t)) arrays."""
    n, D = x.shape
    m = len(t)
    L = np.ones((n, D, m))
    dL = np.zeros((n, D, m))
    for i in range(m):
        for j in range(m):
            if j == i:
                continue
            f = (x - t[j]) / (t[i] - t[j])
            # derivative of the product: sum over the dropped factor
            dprod = np.ones((n, D)) / (t[i] - t[j])
            for l in range(m):
                if l == i or l == j:
                    continue
                dprod = dprod * (x - t[l]) / (t[i] - t[l])
            dL[:, :, i] += dprod
            L[:, :, i] *= f
    return L, dL


# ---------------------------------------------------------------------------
# quadrature (src/quadrature.jl)
# ---------------------------------------------------------------------------
def gauss_legendre01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * x + 0.5, 0.5 * w


def _strang3(degree):
    # strang_quadrature_4/5(::UnitSimplex{3}), src/quadrature.jl:554-634 (tabulated data)
    if degree == 4:
        a, b = 0.3994035761667992, 0.1005964238332008
        c, d = (343.0 / 7500.0) / 6.0, (56.0 / 375.0) / 6.0
        e, f, g, h = 1.0 / 4.0, 11.0 / 14.0, 1.0 / 14.0, (-148.0 / 1875.0) / 6.0
        x = [(e, e, e), (f, g, g), (g, f, g), (g, g, f), (g, g, g), (a, a, b), (a, b, a), (a, b, b),
             (b, a, a), (b, a, b), (b, b, a)]
        w = [h, c, c, c, c, d, d, d, d, d, d]
        return np.asarray(x), np.asarray(w)
    if degree == 5:
        a, b, c, d = 0.0673422422100983, 0.3108859192633005, 0.7217942490673264, 0.0927352503108912
        e, f = 0.4544962958743506, 0.0455037041256494
        p, q, r = 0.1126879257180162 / 6.0, 0.0734930431163619 / 6.0, 0.0425460207770812 / 6.0
        x = [(a, b, b), (b, a, b), (b, b, a), (b, b, b), (c, d, d), (d, c, d), (d, d, c), (d, d, d),
             (e, e, f), (e, f, e), (e, f, f), (f, e, e), (f, e, f), (f, f, e)]
        w = [p, p, p, p, q, q, q, q, r, r, r, r, r, r]
        return np.asarray(x), np.asarray(w)
    return None


def quadrature(kind, degree):
    """quadrature(geo, degree), src/quadrature.jl:36-52: tensor Gauss with
    ceil((degree+1)/2) points per direction (x fastest) on n-cubes; Strang
    rules for tetrahedra of degree 1..5 (4 and 5 tabulated here); Duffy
    (Gauss-Jacobi collapsed) otherwise."""
    g = GEO[kind]
    D = g.dim
    if not g.simplex:
        n = (degree + 2) // 2
        x1, w1 = gauss_legendre01(n)
        idx = np.indices([n] * D).reshape(D, -1)[::-1].T
        x = x1[idx]
        w = np.prod(w1[idx], axis=1)
        return np.ascontiguousarray(x), w
    if D == 3 and degree in (4, 5):
        return _strang3(degree)
    from scipy.special import roots_jacobi
    n = int(np.ceil((degree + 1.0) / 2.0))
    xs, ws = [], []
    for d in range(D - 1):
        xj, wj = roots_jacobi(n, (D - 1) - d, 0)
        xs.append(0.5 * xj + 0.5)
        ws.append(0.5 * wj)
    x1, w1 = gauss_legendre01(n)
    xs.append(x1)
    ws.append(w1)
    a = 0.5
    for d in range(D - 2, -1, -1):
        ws[d] = ws[d] * a
        a *= 0.5
    idx = np.indices([n] * D).reshape(D, -1)[::-1].T
    y = np.stack([xs[d][idx[:, d]] for d in range(D)], axis=1)
    w = np.prod(np.stack([ws[d][idx[:, d]] for d in range(D)], axis=1), axis=1)
    x = np.zeros_like(y)
    m = np.ones(len(y))
    for d in range(D):
        if d == 0:
            x[:, 0] = y[:, 0]
        else:
            m = m * (1 - y[:, d - 1])
            x[:, d] = m * y[:, d]
    return x, w


def simplex_cube_subfaces(D):
    """For each d < D and each cube d-face, the simplex d-faces of the
    simplexified unit cube lying on it (node lists, 1-based), in increasing
    simplex-face order (simplexify(::UnitNCube), src/domain.jl:270-318)."""
    from .mesh import Mesh
    cube = HEX if D == 3 else QUAD
    coords = np.asarray(GEO[cube].verts, np.float64)
    cells = np.asarray(SIMPLEX_NODES[D]) - 1
    m = Mesh(D, coords, cells, TET if D == 3 else TRI, {}, {})
    out = {}
    for d in range(D):
        fn = m.vertex_node[m.face_vertices[d]] + 1  # node ids (1-based)
        out[d] = []
        for cf in GEO[cube].faces[d]:
            s = set(cf)
            out[d].append([list(map(int, row)) for row in fn if set(row.tolist()) <= s])
    return out
