"""ctypes binding of libgtk_hip.so (the C ABI declared in include/gtk_hip.h).

The library is loaded from this package directory.  There is no fallback:
if the shared library is missing or a GPU is absent, the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libgtk_hip.so")

GTK_FREE, GTK_DIRICHLET = 1, 2
GTK_LAPLACE, GTK_MASS, GTK_ELASTICITY, GTK_NITSCHE, GTK_NITSCHE_LOAD, GTK_SKELETON = 0, 1, 2, 3, 4, 5
GTK_FP64, GTK_MIXED = 0, 1
GTK_F_CONST, GTK_F_CELL, GTK_F_NODAL = 0, 1, 2
GTK_NOALLOC_MAX_DOUBLES = 8 * 1024 * 1024  # scratch of gtk_assemble_vector_noalloc

# every symbol the header declares
EXPORTS = [
    "gtk_create", "gtk_destroy", "gtk_last_error", "gtk_version", "gtk_set_mesh", "gtk_update_coordinates",
    "gtk_set_space", "gtk_set_reference", "gtk_set_reference_faces", "gtk_build_pattern", "gtk_build_pattern_device",
    "gtk_local_nnz_map", "gtk_add_block", "gtk_gather_block",
    "gtk_get_pattern", "gtk_set_cell_coefficients",
    "gtk_assemble_matrix", "gtk_get_values", "gtk_set_values", "gtk_values_device_ptr", "gtk_assemble_vector",
    "gtk_assemble_vector_noalloc",
    "gtk_plan_info", "gtk_set_launch", "gtk_lift_dirichlet", "gtk_block_multiply", "gtk_set_reference_face_gradients",
    "gtk_assemble_matrix_async", "gtk_sync", "gtk_time_matrix", "gtk_solve", "gtk_batch_create", "gtk_batch_destroy",
    "gtk_batch_last_error", "gtk_batch_add", "gtk_batch_set_tile_cells", "gtk_batch_set_waves", "gtk_batch_set_load", "gtk_batch_get_vector",
    "gtk_batch_build", "gtk_batch_run",
]


class GtkStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("cells", C.c_int64), ("nnz", C.c_int64),
                ("bytes_alg", C.c_int64), ("flops_alg", C.c_int64)]


class GtkSolveStats(C.Structure):
    _fields_ = [("outer_iterations", C.c_int), ("inner_iterations", C.c_int), ("converged", C.c_int),
                ("rel_residual", C.c_double), ("solve_ms", C.c_double)]


class GtkPlanInfo(C.Structure):
    _fields_ = [("matrix_cores", C.c_int), ("threads", C.c_int), ("grid", C.c_int), ("tiles", C.c_int),
                ("max_tile_cells", C.c_int), ("stage_cells", C.c_int), ("lds_bytes", C.c_int), ("tile_cap", C.c_int),
                ("batches", C.c_int64), ("tile_cells", C.c_int64), ("cells", C.c_int64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["halo_redundancy"] = self.tile_cells / max(1, self.cells)
        return d


class GtkError(RuntimeError):
    pass


_lib = None


def load():
    """Load libgtk_hip.so and declare the prototypes (no GPU call)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise GtkError(f"{LIBPATH} is missing: build it with `make -C {os.path.join(HERE, 'csrc')}`")
    L = C.CDLL(LIBPATH)
    vp = C.c_void_p
    f64 = np.ctypeslib.ndpointer(np.float64, flags="C")
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C")
    i = C.c_int
    i64 = C.c_int64
    L.gtk_create.argtypes = [i, C.POINTER(vp)]
    L.gtk_destroy.argtypes = [vp]
    L.gtk_last_error.argtypes = [vp]
    L.gtk_last_error.restype = C.c_char_p
    L.gtk_version.restype = C.c_char_p
    L.gtk_set_mesh.argtypes = [vp, i, i64, f64, i64, i, i32]
    L.gtk_update_coordinates.argtypes = [vp, f64]
    L.gtk_set_space.argtypes = [vp, i, i, i64, i64, i32]
    L.gtk_set_reference.argtypes = [vp, i, f64, i, f64, f64, f64]
    L.gtk_build_pattern.argtypes = [vp, i, i, C.POINTER(i64), C.POINTER(i64)]
    L.gtk_get_pattern.argtypes = [vp, i, i, i32, i32]
    L.gtk_build_pattern_device.argtypes = [vp, i, i, C.POINTER(i64), C.POINTER(i64), C.POINTER(C.c_double)]
    L.gtk_local_nnz_map.argtypes = [vp, i, i, i32]
    L.gtk_set_reference_faces.argtypes = [vp, i, i, f64, i, f64, f64, f64, i32]
    L.gtk_add_block.argtypes = [vp, i, i, vp]
    L.gtk_gather_block.argtypes = [vp, i, i, vp, i, i, np.ctypeslib.ndpointer(np.int64, flags="C"),
                                   np.ctypeslib.ndpointer(np.int64, flags="C")]
    L.gtk_set_cell_coefficients.argtypes = [vp, i, C.c_void_p]
    L.gtk_assemble_matrix.argtypes = [vp, i, i, i, i, C.c_void_p, C.POINTER(GtkStats)]
    L.gtk_get_values.argtypes = [vp, i, i, f64]
    L.gtk_set_values.argtypes = [vp, i, i, f64]
    L.gtk_lift_dirichlet.argtypes = [vp, i, f64, f64]
    L.gtk_block_multiply.argtypes = [vp, i, i, C.c_double, f64, f64]
    L.gtk_set_reference_face_gradients.argtypes = [vp, f64, f64]
    L.gtk_plan_info.argtypes = [vp, i, i, i, i, C.POINTER(GtkPlanInfo)]
    L.gtk_set_launch.argtypes = [vp, i, i, i, i, i, i]
    L.gtk_values_device_ptr.argtypes = [vp, i, i, C.POINTER(C.c_void_p)]
    L.gtk_assemble_vector.argtypes = [vp, i, f64, i, C.c_void_p, C.POINTER(GtkStats)]
    L.gtk_assemble_vector_noalloc.argtypes = [vp, i, f64, i, C.c_void_p, C.POINTER(GtkStats)]
    L.gtk_assemble_matrix_async.argtypes = [vp, i, i, i, i]
    L.gtk_sync.argtypes = [vp]
    L.gtk_time_matrix.argtypes = [vp, i, i, i, i, i, C.POINTER(C.c_double)]
    L.gtk_solve.argtypes = [vp, i, f64, f64, C.c_double, i, i, C.c_double, C.POINTER(GtkSolveStats)]
    L.gtk_batch_create.argtypes = [i, i, i, C.POINTER(vp)]
    L.gtk_batch_destroy.argtypes = [vp]
    L.gtk_batch_last_error.argtypes = [vp]
    L.gtk_batch_last_error.restype = C.c_char_p
    L.gtk_batch_add.argtypes = [vp, vp]
    L.gtk_batch_set_tile_cells.argtypes = [vp, i, i]
    L.gtk_batch_set_waves.argtypes = [vp, i]
    L.gtk_batch_set_load.argtypes = [vp, i, i, f64]
    L.gtk_batch_get_vector.argtypes = [vp, i, f64]
    L.gtk_batch_build.argtypes = [vp]
    L.gtk_batch_run.argtypes = [vp, C.POINTER(GtkStats)]
    _lib = L
    return L


class Context:
    """Owner of one gtk_ctx (device buffers, stream, plans)."""

    def __init__(self, device=0):
        self.L = load()
        self.h = C.c_void_p()
        rc = self.L.gtk_create(device, C.byref(self.h))
        if rc != 0:
            raise GtkError(f"gtk_create failed ({rc}): {self.L.gtk_last_error(None).decode()}")

    def check(self, rc, what):
        if rc != 0:
            raise GtkError(f"{what} failed ({rc}): {self.L.gtk_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.L.gtk_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # thin wrappers ------------------------------------------------------------
    def set_mesh(self, dim, xyz, cell_nodes_1b):
        xyz = np.ascontiguousarray(xyz, np.float64)
        cn = np.ascontiguousarray(cell_nodes_1b, np.int32)
        self.check(self.L.gtk_set_mesh(self.h, dim, xyz.shape[0], xyz, cn.shape[0], cn.shape[1], cn), "gtk_set_mesh")

    def update_coordinates(self, xyz):
        self.check(self.L.gtk_update_coordinates(self.h, np.ascontiguousarray(xyz, np.float64)), "gtk_update_coordinates")

    def set_space(self, ncomp, cell_dofs, nfree, ndiri):
        cd = np.ascontiguousarray(cell_dofs, np.int32)
        self.check(self.L.gtk_set_space(self.h, ncomp, cd.shape[1], nfree, ndiri, cd), "gtk_set_space")

    def set_reference(self, w, phi, grad_phi, grad_geo):
        w = np.ascontiguousarray(w, np.float64)
        phi = np.ascontiguousarray(phi, np.float64)
        self.check(self.L.gtk_set_reference(self.h, len(w), w, phi.shape[1], phi,
                                            np.ascontiguousarray(grad_phi, np.float64),
                                            np.ascontiguousarray(grad_geo, np.float64)), "gtk_set_reference")

    def set_reference_faces(self, w, phi, grad_geo, face_jac, item_face):
        phi = np.ascontiguousarray(phi, np.float64)  # (nfaces, nq, nshape)
        self.check(self.L.gtk_set_reference_faces(self.h, phi.shape[0], phi.shape[1],
                                                  np.ascontiguousarray(w, np.float64), phi.shape[2], phi,
                                                  np.ascontiguousarray(grad_geo, np.float64),
                                                  np.ascontiguousarray(face_jac, np.float64),
                                                  np.ascontiguousarray(item_face, np.int32)), "gtk_set_reference_faces")

    def set_reference_face_gradients(self, grad_phi, normals):
        self.check(self.L.gtk_set_reference_face_gradients(self.h, np.ascontiguousarray(grad_phi, np.float64),
                                                           np.ascontiguousarray(normals, np.float64)),
                   "gtk_set_reference_face_gradients")

    def block_multiply(self, rows, cols, x, y, alpha=1.0):
        """y + alpha A x for the (rows, cols) block, on the device."""
        out = np.array(y, np.float64, copy=True)
        self.check(self.L.gtk_block_multiply(self.h, rows, cols, float(alpha), np.ascontiguousarray(x, np.float64),
                                             out), "gtk_block_multiply")
        return out

    def add_block(self, rows, cols, src):
        self.check(self.L.gtk_add_block(self.h, rows, cols, src.h), "gtk_add_block")

    def gather_block(self, rows, cols, src, src_rows, src_cols, ptr, idx):
        """Values of block (rows, cols) := sums of src's values grouped by
        ptr/idx (gtk_gather_block)."""
        idx = np.ascontiguousarray(idx, np.int64)
        if idx.size == 0:
            idx = np.zeros(1, np.int64)
        self.check(self.L.gtk_gather_block(self.h, rows, cols, src.h, src_rows, src_cols,
                                           np.ascontiguousarray(ptr, np.int64), idx), "gtk_gather_block")

    def build_pattern(self, rows, cols):
        nnz, ncols = C.c_int64(), C.c_int64()
        self.check(self.L.gtk_build_pattern(self.h, rows, cols, C.byref(nnz), C.byref(ncols)), "gtk_build_pattern")
        return nnz.value, ncols.value

    def build_pattern_device(self, rows, cols):
        nnz, ncols, ms = C.c_int64(), C.c_int64(), C.c_double()
        self.check(self.L.gtk_build_pattern_device(self.h, rows, cols, C.byref(nnz), C.byref(ncols), C.byref(ms)),
                   "gtk_build_pattern_device")
        return nnz.value, ncols.value, ms.value

    def local_nnz_map(self, rows, cols, n_cells, nd):
        out = np.zeros(max(n_cells * nd * nd, 1), np.int32)
        self.check(self.L.gtk_local_nnz_map(self.h, rows, cols, out), "gtk_local_nnz_map")
        return out[:n_cells * nd * nd].reshape(n_cells, nd, nd)

    def get_pattern(self, rows, cols, nnz, ncols):
        colptr = np.zeros(ncols + 1, np.int32)
        rowval = np.zeros(max(nnz, 1), np.int32)
        self.check(self.L.gtk_get_pattern(self.h, rows, cols, colptr, rowval), "gtk_get_pattern")
        return colptr, rowval[:nnz]

    def set_cell_coefficients(self, coef, ncoef):
        if coef is None:
            self.check(self.L.gtk_set_cell_coefficients(self.h, 0, None), "gtk_set_cell_coefficients")
            return
        self._coef = np.ascontiguousarray(coef, np.float64)
        self.check(self.L.gtk_set_cell_coefficients(self.h, ncoef, self._coef.ctypes.data_as(C.c_void_p)),
                   "gtk_set_cell_coefficients")

    def assemble_matrix(self, form, precision, rows, cols, nnz):
        out = np.zeros(max(nnz, 1), np.float64)
        st = GtkStats()
        self.check(self.L.gtk_assemble_matrix(self.h, form, precision, rows, cols, out.ctypes.data_as(C.c_void_p),
                                              C.byref(st)), "gtk_assemble_matrix")
        return out[:nnz], st

    def assemble_vector(self, fkind, f, rows, n):
        out = np.zeros(max(n, 1), np.float64)
        st = GtkStats()
        self.check(self.L.gtk_assemble_vector(self.h, fkind, np.ascontiguousarray(f, np.float64), rows,
                                              out.ctypes.data_as(C.c_void_p), C.byref(st)), "gtk_assemble_vector")
        return out[:n], st

    def assemble_vector_noalloc(self, fkind, f, rows, n):
        """gtk_assemble_vector_noalloc: the same load vector with no device
        allocation (needs one gtk_assemble_vector on the row set first)."""
        out = np.zeros(max(n, 1), np.float64)
        st = GtkStats()
        self.check(self.L.gtk_assemble_vector_noalloc(self.h, fkind, np.ascontiguousarray(f, np.float64), rows,
                                                      out.ctypes.data_as(C.c_void_p), C.byref(st)),
                   "gtk_assemble_vector_noalloc")
        return out[:n], st

    def assemble_matrix_async(self, form, precision, rows, cols):
        self.check(self.L.gtk_assemble_matrix_async(self.h, form, precision, rows, cols), "gtk_assemble_matrix_async")

    def sync(self):
        self.check(self.L.gtk_sync(self.h), "gtk_sync")

    def time_matrix(self, form, precision, rows, cols, launches):
        ms = C.c_double()
        self.check(self.L.gtk_time_matrix(self.h, form, precision, rows, cols, launches, C.byref(ms)), "gtk_time_matrix")
        return ms.value

    def solve(self, precision, b, x0, rtol, max_outer, max_inner, inner_rtol):
        b = np.ascontiguousarray(b, np.float64)
        x = np.array(x0, np.float64, copy=True) if x0 is not None else np.zeros_like(b)
        st = GtkSolveStats()
        self.check(self.L.gtk_solve(self.h, precision, b, x, rtol, max_outer, max_inner, inner_rtol, C.byref(st)),
                   "gtk_solve")
        return x, st

    def get_values(self, rows, cols, nnz):
        out = np.zeros(max(nnz, 1), np.float64)
        self.check(self.L.gtk_get_values(self.h, rows, cols, out), "gtk_get_values")
        return out[:nnz]

    def set_values(self, rows, cols, nzval):
        self.check(self.L.gtk_set_values(self.h, rows, cols, np.ascontiguousarray(nzval, np.float64)),
                   "gtk_set_values")

    def lift_dirichlet(self, xd, b, rows=GTK_FREE):
        out = np.array(b, np.float64, copy=True)
        self.check(self.L.gtk_lift_dirichlet(self.h, rows, np.ascontiguousarray(xd, np.float64), out),
                   "gtk_lift_dirichlet")
        return out

    def plan_info(self, form, precision, rows, cols):
        info = GtkPlanInfo()
        self.check(self.L.gtk_plan_info(self.h, form, precision, rows, cols, C.byref(info)), "gtk_plan_info")
        return info.as_dict()

    def set_launch(self, form, precision, rows, cols, waves=0, max_cells=0):
        self.check(self.L.gtk_set_launch(self.h, form, precision, rows, cols, waves, max_cells), "gtk_set_launch")

    def values_device_ptr(self, rows, cols):
        p = C.c_void_p()
        self.check(self.L.gtk_values_device_ptr(self.h, rows, cols, C.byref(p)), "gtk_values_device_ptr")
        return p.value


class Batch:
    """Owner of one gtk_batch: problems (contexts) assembled by one
    persistent launch captured in a HIP graph."""

    def __init__(self, device, form, precision):
        self.L = load()
        self.h = C.c_void_p()
        self.ctxs = []  # keeps the contexts alive while the batch refers to them
        rc = self.L.gtk_batch_create(device, form, precision, C.byref(self.h))
        if rc != 0:
            raise GtkError(f"gtk_batch_create failed ({rc})")

    def check(self, rc, what):
        if rc != 0:
            raise GtkError(f"{what} failed ({rc}): {self.L.gtk_batch_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.L.gtk_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, ctx):
        self.check(self.L.gtk_batch_add(self.h, ctx.h), "gtk_batch_add")
        self.ctxs.append(ctx)

    def set_tile_cells(self, elem_type, max_cells):
        self.check(self.L.gtk_batch_set_tile_cells(self.h, elem_type, max_cells), "gtk_batch_set_tile_cells")

    def set_waves(self, waves):
        self.check(self.L.gtk_batch_set_waves(self.h, waves), "gtk_batch_set_waves")

    def set_load(self, problem, fkind, f):
        self.check(self.L.gtk_batch_set_load(self.h, problem, fkind, np.ascontiguousarray(f, np.float64)),
                   "gtk_batch_set_load")

    def get_vector(self, problem, n):
        out = np.zeros(max(n, 1), np.float64)
        self.check(self.L.gtk_batch_get_vector(self.h, problem, out), "gtk_batch_get_vector")
        return out[:n]

    def build(self):
        self.check(self.L.gtk_batch_build(self.h), "gtk_batch_build")
        self.built = True

    def run(self):
        st = GtkStats()
        self.check(self.L.gtk_batch_run(self.h, C.byref(st)), "gtk_batch_run")
        return st
