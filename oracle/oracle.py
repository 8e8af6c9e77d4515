"""ctypes wrapper of the CPU oracle (oracle/gtk_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py as the checker.  The product path never imports
this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GTK_ORACLE_LIB: another build of the same sources (the sanitizer test's)
LIB = os.environ.get("GTK_ORACLE_LIB") or os.path.join(HERE, "libgtk_oracle.so")

POINT, SEG, QUAD, HEX, TRI, TET = range(6)
LAPLACE, MASS, ELASTICITY = 0, 1, 2
FREE, DIRICHLET = 1, 2
(Q_DIM, Q_NNODES, Q_NCELLS, Q_NVPC, Q_NVERT, Q_NEDGES, Q_NFACES2, Q_NDOFS, Q_NFREE, Q_NDIRI,
 Q_DOFS_PER_CELL, Q_NNZ, Q_NQ, Q_NPARENT0, Q_NPARENT1, Q_NPARENT2, Q_CELLKIND) = range(17)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, i32p, f64p = C.c_void_p, np.ctypeslib.ndpointer(np.int32, flags="C"), np.ctypeslib.ndpointer(np.float64, flags="C")
        L.orc_cartesian.restype = vp
        L.orc_cartesian.argtypes = [C.c_int, i32p, f64p, C.c_int]
        L.orc_free.argtypes = [vp]
        L.orc_last_error.restype = C.c_char_p
        L.orc_query.restype = C.c_int64
        L.orc_query.argtypes = [vp, C.c_int]
        for name in ["orc_get_cell_nodes", "orc_get_node_vertex", "orc_get_cell_dofs"]:
            getattr(L, name).argtypes = [vp, i32p]
        for name in ["orc_set_coords", "orc_get_coords"]:
            getattr(L, name).argtypes = [vp, f64p]
        for name in ["orc_get_cell_faces", "orc_get_face_vertices", "orc_get_cell_perm"]:
            getattr(L, name).argtypes = [vp, C.c_int, i32p]
        L.orc_get_parent_faces.argtypes = [vp, C.c_int, i32p, i32p]
        L.orc_space.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_set_quadrature.argtypes = [vp, C.c_int]
        L.orc_get_quadrature.argtypes = [vp, f64p, f64p]
        L.orc_tabulate.argtypes = [vp, C.c_int, f64p, C.c_int, f64p, f64p]
        L.orc_assemble_matrix.argtypes = [vp, C.c_int, C.c_void_p, C.c_int, C.c_int]
        L.orc_get_csc.argtypes = [vp, i32p, i32p, f64p]
        L.orc_csc_ncols.restype = C.c_int64
        L.orc_csc_ncols.argtypes = [vp]
        L.orc_assemble_vector.argtypes = [vp, C.c_int, f64p, C.c_int, f64p]
        L.orc_gauss_legendre.argtypes = [C.c_int, f64p, f64p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_phase_ms.argtypes = [f64p]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def set_threads(n):
    """Threads of the matrix assembly (the result does not depend on it)."""
    lib().orc_set_threads(int(n))


def phase_ms():
    """Wall time (ms) of the last matrix assembly's phases: symbolic count,
    numeric cell loop, COO -> CSC compress."""
    out = np.zeros(3)
    lib().orc_phase_ms(out)
    return {"symbolic": out[0], "numeric": out[1], "compress": out[2]}


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    _check(lib().orc_gauss_legendre(n, x, w))
    return x, w


class Problem:
    """cartesian_mesh + complexify + lagrange_space + assembly, restated in C."""

    def __init__(self, cells, box=None, simplexify=False):
        D = len(cells)
        if box is None:
            box = [0.0, 1.0] * D
        self.D = D
        self.h = lib().orc_cartesian(D, np.asarray(cells, np.int32), np.asarray(box, np.float64), int(simplexify))
        if not self.h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def q(self, what):
        return int(lib().orc_query(self.h, what))

    # mesh / topology
    def coords(self):
        x = np.zeros((self.q(Q_NNODES), self.D))
        _check(lib().orc_get_coords(self.h, x))
        return x

    def set_coords(self, x):
        _check(lib().orc_set_coords(self.h, np.ascontiguousarray(x, np.float64)))

    def cell_nodes(self):
        out = np.zeros((self.q(Q_NCELLS), self.q(Q_NVPC)), np.int32)
        _check(lib().orc_get_cell_nodes(self.h, out))
        return out

    def node_vertex(self):
        out = np.zeros(self.q(Q_NNODES), np.int32)
        _check(lib().orc_get_node_vertex(self.h, out))
        return out

    _NLF = {QUAD: [4, 4, 1], HEX: [8, 12, 6, 1], TRI: [3, 3, 1], TET: [4, 6, 4, 1]}
    _FV = {QUAD: [1, 2, 4], HEX: [1, 2, 4, 8], TRI: [1, 2, 3], TET: [1, 2, 3, 4]}

    def nfaces(self, d):
        if d == 0:
            return self.q(Q_NVERT)
        if d == self.D:
            return self.q(Q_NCELLS)
        if d == 1:
            return self.q(Q_NEDGES)
        return self.q(Q_NFACES2)

    def cell_faces(self, d):
        k = self.q(Q_CELLKIND)
        out = np.zeros((self.q(Q_NCELLS), self._NLF[k][d]), np.int32)
        _check(lib().orc_get_cell_faces(self.h, d, out))
        return out

    def cell_perm(self, d):
        k = self.q(Q_CELLKIND)
        out = np.zeros((self.q(Q_NCELLS), self._NLF[k][d]), np.int32)
        _check(lib().orc_get_cell_perm(self.h, d, out))
        return out

    def face_vertices(self, d):
        k = self.q(Q_CELLKIND)
        out = np.zeros((self.nfaces(d), self._FV[k][d]), np.int32)
        _check(lib().orc_get_face_vertices(self.h, d, out))
        return out

    def parent_faces(self, d):
        n = self.q(Q_NPARENT0 + d)
        w = (d + 1) if self.q(Q_CELLKIND) in (TRI, TET) else (1 << d)
        nodes = np.zeros((n, w), np.int32)
        grp = np.zeros(n, np.int32)
        _check(lib().orc_get_parent_faces(self.h, d, nodes, grp))
        return nodes, grp

    # space
    def space(self, order, ncomp=1, dirichlet=0, group=0):
        _check(lib().orc_space(self.h, order, ncomp, dirichlet, group))
        self._ncomp = ncomp
        return self

    def cell_dofs(self):
        out = np.zeros((self.q(Q_NCELLS), self.q(Q_DOFS_PER_CELL)), np.int32)
        _check(lib().orc_get_cell_dofs(self.h, out))
        return out

    def quadrature(self, degree):
        _check(lib().orc_set_quadrature(self.h, degree))
        n = self.q(Q_NQ)
        x = np.zeros((n, self.D))
        w = np.zeros(n)
        lib().orc_get_quadrature(self.h, x, w)
        return x, w

    def tabulate(self, xi, geometry=False):
        xi = np.ascontiguousarray(xi, np.float64)
        n = xi.shape[0]
        ns = (1 << self.D if self.q(Q_CELLKIND) in (QUAD, HEX) else self.D + 1) if geometry else \
            self.q(Q_DOFS_PER_CELL) // max(1, self._ncomp)
        val = np.zeros((n, ns))
        grad = np.zeros((n, ns, self.D))
        _check(lib().orc_tabulate(self.h, int(geometry), xi, n, val, grad))
        return val, grad

    _ncomp = 1

    def assemble_matrix(self, form=LAPLACE, coef=None, rows=FREE, cols=FREE):
        c = None
        if coef is not None:
            coef = np.ascontiguousarray(coef, np.float64)
            c = coef.ctypes.data_as(C.c_void_p)
        _check(lib().orc_assemble_matrix(self.h, form, c, rows, cols))
        ncol = int(lib().orc_csc_ncols(self.h))
        nnz = self.q(Q_NNZ)
        colptr = np.zeros(ncol + 1, np.int32)
        rowval = np.zeros(nnz, np.int32)
        nzval = np.zeros(nnz)
        _check(lib().orc_get_csc(self.h, colptr, rowval, nzval))
        return colptr, rowval, nzval

    def assemble_vector(self, f_kind=0, fdata=(1.0,), rows=FREE):
        n = self.q(Q_NFREE) if rows == FREE else self.q(Q_NDIRI)
        b = np.zeros(n)
        _check(lib().orc_assemble_vector(self.h, f_kind, np.ascontiguousarray(fdata, np.float64), rows, b))
        return b
