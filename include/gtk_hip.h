/*
 * gtk_hip.h -- C ABI of libgtk_hip.so, the MI355X (gfx950) finite-element
 * assembly library.
 *
 * The library replaces the per-cell loop that the reference generates for
 * GT.assemble_matrix / GT.assemble_vector over GT.measure on a
 * GT.lagrange_space (GalerkinToolkit.jl: src/problems.jl:244-274 and
 * :319-350 call generate_assemble_vector / generate_assemble_matrix,
 * src/compiler.jl:1097-1155, whose loop templates are
 * src/compiler.jl:1826-1997), together with the allocation protocol that
 * turns the contributions into a SparseMatrixCSC{Float64,Int32} and a
 * Vector{Float64} (src/assembly.jl:75-208, :428-588, cached CSC :592-893).
 *
 * Conventions
 *  - Every call returns 0 on success or a negative GTK_E* code; nothing
 *    throws across the ABI.  gtk_last_error() returns the message of the
 *    last failure on the context (or of the last failed gtk_create).
 *  - The caller owns every host array; set_* calls copy them to device
 *    memory owned by the context until gtk_destroy.
 *  - Indices follow the reference: 1-based Int32 node ids, DOF ids in the
 *    reference's signed encoding (>0 free, <0 Dirichlet), CSC colptr/rowval
 *    1-based Int32 exactly as SparseMatrixCSC{Float64,Int32} stores them.
 *  - A context is not thread-safe.  Calls are synchronous unless named
 *    *_async; distinct contexts may be used from distinct threads.
 */
#ifndef GTK_HIP_H
#define GTK_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gtk_ctx gtk_ctx;

/* status codes */
#define GTK_OK 0
#define GTK_EINVAL (-1)   /* invalid argument or inconsistent sizes */
#define GTK_ESTATE (-2)   /* call out of order (mesh/space/pattern missing) */
#define GTK_EDEVICE (-3)  /* HIP runtime error */
#define GTK_ENOMEM (-4)   /* host or device allocation failed */
#define GTK_EUNSUP (-5)   /* element/form/precision combination not built */

/* row / column selectors, as the reference's FREE / DIRICHLET
 * (src/field.jl:136-142, skip_dof src/assembly.jl:155-157) */
#define GTK_FREE 1
#define GTK_DIRICHLET 2

/* bilinear forms recognised by the library (the integrands of
 * src/compiler.jl:1544-1606 that the supported weak forms lower to) */
#define GTK_LAPLACE 0     /* int kappa grad(u).grad(v), kappa per cell (or 1) */
#define GTK_MASS 1        /* int kappa u v */
#define GTK_ELASTICITY 2  /* int sigma(eps(u)) : eps(v), (lambda, mu) per cell */
/* Nitsche terms on boundary items with all DOFs of their cell (scalar
 * fields; GalerkinToolkitExamples/src/poisson.jl:93-112), beta per item
 * (gamma / h of the face) as the item's coefficient, n the outward unit
 * normal: */
#define GTK_NITSCHE 3       /* int beta v u - v n.grad(u) - n.grad(v) u */
#define GTK_NITSCHE_LOAD 4  /* int beta v u - n.grad(v) u: applied to the nodal
                               values of g (gtk_block_multiply) it is the load
                               int beta v g - n.grad(v) g of the same method */
/* Interior-penalty terms on skeleton items (an interior face with both its
 * cells; scalar fields; poisson.jl:99-102 over GT.skeleton, src/mesh.jl:
 * 236-247, terms of src/compiler.jl:524-529), beta per item (gamma / h):
 * int beta jump(v).jump(u) - jump(v).mean(grad u) - mean(grad v).jump(u),
 * jump(u) = u1 n1 + u2 n2, mean(f) = (f1 + f2) / 2 (poisson.jl:40-41) */
#define GTK_SKELETON 5

/* arithmetic of the per-cell contraction */
#define GTK_FP64 0   /* everything in FP64: the plain drop-in */
#define GTK_MIXED 1  /* FP64 geometry (J, J^-1, det J) and FP64 cross-cell
                        accumulation; per-cell contraction in FP32, or on
                        FP16 matrix cores with each operand split into two
                        FP16 parts (22-bit operands, FP32 accumulation) where
                        the contraction is long (Q2 hexahedra); values within
                        1e-6 (relative Frobenius) of GTK_FP64 */

/* load-vector source kinds */
#define GTK_F_CONST 0  /* one value per component */
#define GTK_F_CELL 1   /* one value per cell and component */
#define GTK_F_NODAL 2  /* discrete field: values at free dofs then Dirichlet dofs */

typedef struct gtk_stats {
  double kernel_ms;     /* device time of the assembly launch(es), HIP events */
  double total_ms;      /* wall time of the call on the host */
  int64_t cells;        /* cells processed */
  int64_t nnz;          /* nonzeros written (matrix) or rows written (vector) */
  int64_t bytes_alg;    /* algorithmic HBM bytes of the launch (see DESIGN) */
  int64_t flops_alg;    /* algorithmic flops of the launch */
} gtk_stats;

/* Context life cycle.  device: HIP device ordinal. */
int gtk_create(int device, gtk_ctx **out);
int gtk_destroy(gtk_ctx *ctx);
const char *gtk_last_error(const gtk_ctx *ctx);
/* Library build identifier (compile date, target). */
const char *gtk_version(void);

/* Mesh: node coordinates (n_nodes*dim, node-major, as
 * GT.node_coordinates), cells of one reference type with nodes_per_cell
 * nodes of an order-1 geometry (GT.face_nodes(mesh, D), 1-based).
 * Replaces: the node_coordinates / face_nodes reads of the generated loop
 * (src/accessors.jl:638-662, :941-968). */
int gtk_set_mesh(gtk_ctx *ctx, int dim, int64_t n_nodes, const double *xyz, int64_t n_cells,
                 int nodes_per_cell, const int32_t *cell_nodes);
/* Replace node coordinates only (same topology). */
int gtk_update_coordinates(gtk_ctx *ctx, const double *xyz);

/* Space: cell -> DOF map in GT's signed encoding (GT.face_dofs of the
 * space on the cells, src/space.jl:299-535), with ncomp components per
 * node stored component-major.  dofs_per_cell = nshape*ncomp. */
int gtk_set_space(gtk_ctx *ctx, int ncomp, int dofs_per_cell, int64_t n_free, int64_t n_diri,
                  const int32_t *cell_dofs);

/* Reference tables at the quadrature points of GT.measure(Omega, degree)
 * (src/problems.jl:4-25, src/quadrature.jl:36-157): weights w[nq],
 * shape values phi[nq*nshape], reference gradients grad_phi[nq*nshape*dim],
 * order-1 geometry gradients grad_geo[nq*nodes_per_cell*dim]
 * (tabulator, src/space.jl:960-970; src/accessors.jl:486-496). */
int gtk_set_reference(gtk_ctx *ctx, int nq, const double *w, int nshape, const double *phi,
                      const double *grad_phi, const double *grad_geo);

/* Boundary integrals (face mode).  The context's "cells" are boundary items
 * (a cell and one of its local faces): gtk_set_mesh with each item's cell
 * nodes, gtk_set_space with each item's DOFs on that face's closure nodes
 * (dofs_per_cell = face nodes x ncomp), then these tables at the nq points
 * of the face rule mapped into the cell's reference coordinates, per local
 * face: w[nq] (weights of the reference face), phi[nfaces][nq][nshape] (the
 * cell's shape functions of the face's nodes), grad_geo[nfaces][nq][nv][dim]
 * (order-1 geometry gradients), face_jac[nfaces][dim][dim-1]
 * (d reference-cell point / d reference-face point), item_face[n_items]
 * (0-based local face of each item).  The measure is
 * sqrt(det(Jf' Jf)) w with Jf = J face_jac (src/accessors.jl:1000-1007 on
 * the face map).  gtk_assemble_vector then integrates f v over the items and
 * gtk_assemble_matrix(GTK_MASS) kappa u v, kappa per item; the pattern comes
 * from the items' DOFs.  Replaces the boundary-measure terms
 * (GT.measure(GT.boundary(mesh), degree), src/compiler.jl:513-529) of the
 * generated loops. */
/* Skeleton items (GTK_SKELETON): gtk_set_mesh with both cells' nodes per
 * item (side 0 then side 1, 2 x nodes_per_cell), gtk_set_space with both
 * cells' DOFs (dofs_per_cell = 2 nshape); the nfaces tables are point sets
 * (at most 16; the kernels take 8 for quadrilaterals, 12 for hexahedra):
 * each the face rule's points in one cell's reference coordinates, the
 * sets of an item's two sides listing the same physical points in the same
 * order; grad_geo keeps its stride of nq * nv * dim values per set, the
 * first half one cell's gradients [nq][nv / 2][dim], the rest zeros;
 * face_jac is the set's local face's (side 0's sets give the measure), and
 * item_face[i] = set of side 0 | set of side 1 << 16.
 * Such a context assembles GTK_SKELETON only. */
int gtk_set_reference_faces(gtk_ctx *ctx, int nfaces, int nq, const double *w, int nshape, const double *phi,
                            const double *grad_geo, const double *face_jac, const int32_t *item_face);

/* Face mode, Nitsche terms: reference gradients of the item's shape
 * functions at the face rule's points, grad_phi[nfaces][nq][nshape][dim]
 * (the points of gtk_set_reference_faces, mapped into the cell), and the
 * outward unit normal of each local face of the reference cell,
 * normals[nfaces][dim] (src/domain.jl:226, :258).  The physical normal at a
 * point is J^-T n / |J^-T n| (map_unit_normal, src/accessors.jl:1009-1030).
 * Replaces the n.grad terms of the generated boundary loops
 * (src/compiler.jl:513-529) for the Nitsche method of
 * GalerkinToolkitExamples/src/poisson.jl:93-112. */
int gtk_set_reference_face_gradients(gtk_ctx *ctx, const double *grad_phi, const double *normals);

/* Symbolic phase for a (rows, cols) block: builds the CSC pattern and the
 * column-owner work plan.  Equals the pattern compress() produces for the
 * COO triplets of allocate_matrix (src/assembly.jl:75-157, :571-575);
 * structural zeros are kept, rows sorted within columns.
 * Returns nnz and the column count. */
int gtk_build_pattern(gtk_ctx *ctx, int rows, int cols, int64_t *nnz, int64_t *ncols);
int gtk_get_pattern(gtk_ctx *ctx, int rows, int cols, int32_t *colptr, int32_t *rowval);
/* The same symbolic phase on the GPU (column incidence by counting, each
 * column's rows sorted and deduplicated by one wavefront, pointers by
 * scans); colptr/rowval are bit-identical to gtk_build_pattern's and a host
 * copy is kept for the work plan and gtk_get_pattern.  device_ms (may be
 * NULL): device time of the pattern kernels.  GTK_EUNSUP if a column
 * gathers more than 2048 candidate rows.  Replaces, on the device, the
 * symbolic pass of allocate_matrix (src/assembly.jl:75-157) and compress
 * (:571-575). */
int gtk_build_pattern_device(gtk_ctx *ctx, int rows, int cols, int64_t *nnz, int64_t *ncols, double *device_ms);
/* Cell-local-entry -> nonzero map of a built block, computed on the GPU:
 * map_out[(cell*dofs_per_cell + b)*dofs_per_cell + a] = 1-based index in
 * the block's values of entry (row of local DOF a, column of local DOF b),
 * 0 when that row or column is outside the block; n_cells*dofs_per_cell^2
 * values.  The cached map of the reference's reassembly path
 * (src/assembly.jl:694-771). */
int gtk_local_nnz_map(gtk_ctx *ctx, int rows, int cols, int32_t *map_out);
/* dst's (rows, cols) values += src's, on the device, where the two contexts
 * number the same DOFs and src's pattern lies inside dst's (a boundary block
 * inside the volume block): the sum of integrals over several measures that
 * the reference compresses into one matrix (src/assembly.jl:545-575).
 * GTK_EINVAL if a src entry is outside dst's pattern. */
int gtk_add_block(gtk_ctx *dst, int rows, int cols, gtk_ctx *src);
/* dst's (rows, cols) values set from src's (src_rows, src_cols) values
 * through a host-given grouping: value t of dst becomes the sum, in order, of
 * src values idx[ptr[t]] .. idx[ptr[t+1] - 1] (ptr has nnz(dst) + 1
 * entries, from 0).  The skeleton terms of a continuous space are assembled
 * on the cells' own copies of their DOFs (a discontinuous space) and
 * gathered onto the continuous DOFs this way: the sum over the copies that
 * the reference's COO compression performs (src/assembly.jl:571-575). */
int gtk_gather_block(gtk_ctx *dst, int rows, int cols, gtk_ctx *src, int src_rows, int src_cols, const int64_t *ptr,
                     const int64_t *idx);

/* Per-cell coefficients resident on the device: kappa (ncoef=1) or
 * (lambda, mu) (ncoef=2), n_cells*ncoef values, cell-major
 * (GT.face_constant_field, src/field.jl:60-70).  NULL resets to 1. */
int gtk_set_cell_coefficients(gtk_ctx *ctx, int ncoef, const double *coef);

/* Numeric assembly of the (rows, cols) block into the device-resident
 * value array of that pattern; nzval_out (host, nnz) may be NULL to keep
 * the values on the device.  Replaces the generated loop of
 * src/compiler.jl:1865-1917 plus contribute!/compress
 * (src/assembly.jl:189-208, :545-575). */
int gtk_assemble_matrix(gtk_ctx *ctx, int form, int precision, int rows, int cols, double *nzval_out,
                        gtk_stats *stats);
/* Copy the device values of a (rows, cols) block to the host. */
int gtk_get_values(gtk_ctx *ctx, int rows, int cols, double *nzval_out);
/* Replace the device values of a (rows, cols) block by nzval (host, nnz
 * values in the block's CSC order): a matrix whose values the caller
 * changed, or one assembled earlier and overwritten on the device by the
 * assembly of another form on the same block, before gtk_solve.  The
 * reference's solve reads the values of the SparseMatrixCSC it is given
 * (src/problems.jl:439-462); this is how a caller hands them back. */
int gtk_set_values(gtk_ctx *ctx, int rows, int cols, const double *nzval);
/* Lifting of the Dirichlet values: b -= A_(rows, DIRICHLET) xd on the
 * device, with the values last assembled into the (rows, GTK_DIRICHLET)
 * block, each row summed in ascending column order with unfused multiply
 * and add.  b: host, in/out, one value per row of the block; xd: host,
 * n_diri values.  Replaces mul!(b, Ad, xd, -1, 1) in
 * PartitionedSolvers_linear_problem (src/problems.jl:439-452). */
int gtk_lift_dirichlet(gtk_ctx *ctx, int rows, const double *xd, double *b);
/* y += A x for the (rows, cols) block with the values last assembled into
 * it, on the device: y_i += sum_k A_ik (alpha x_k), each row in ascending
 * column order with unfused multiply and add; x (n cols), y (n rows, in/out)
 * host arrays.  gtk_lift_dirichlet is alpha = -1 on (rows, GTK_DIRICHLET).
 * Replaces mul!(y, A, x, alpha, 1) on the assembled SparseMatrixCSC (the
 * reference's lifting and the Nitsche load from nodal boundary data). */
int gtk_block_multiply(gtk_ctx *ctx, int rows, int cols, double alpha, const double *x, double *y);
/* Device address of the value array of a (rows, cols) block (for callers
 * that keep the matrix on the GPU). */
int gtk_values_device_ptr(gtk_ctx *ctx, int rows, int cols, double **dptr);

/* Load vector b_i = int f . v_i over the rows set (src/problems.jl:244-274,
 * template src/compiler.jl:1927-1997, dense_vector src/assembly.jl:558-569).
 * f: GTK_F_CONST ncomp values, GTK_F_CELL n_cells*ncomp, GTK_F_NODAL
 * (n_free+n_diri) values.  b_out (host) may be NULL. */
int gtk_assemble_vector(gtk_ctx *ctx, int fkind, const double *f, int rows, double *b_out, gtk_stats *stats);

/* gtk_assemble_vector without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernel and result as gtk_assemble_vector,
 * bit for bit; the load values and the vector are staged in a per-device
 * scratch array inside the library (GTK_NOALLOC_MAX_DOUBLES doubles, 64 MiB,
 * held by every device the library runs on until the process ends) instead
 * of context buffers, and rows that no cell touches are cleared by a kernel,
 * not a memset.
 *  - GTK_EINVAL if the rows of the set (rounded up to 32) plus the load
 *    values (ncomp, n_cells*ncomp or n_free+n_diri) exceed
 *    GTK_NOALLOC_MAX_DOUBLES.
 *  - GTK_ESTATE unless gtk_assemble_vector ran once on this context and row
 *    set since its mesh, space or reference tables were last set (that call
 *    builds the tile plan, which allocates).
 * Every call on one GPU shares that GPU's scratch, whatever the context:
 * calls that overlap on one GPU (two host threads, two contexts) race.  The
 * scratch keeps the last call's contents. */
#define GTK_NOALLOC_MAX_DOUBLES (8u * 1024u * 1024u)
int gtk_assemble_vector_noalloc(gtk_ctx *ctx, int fkind, const double *f, int rows, double *b_out,
                                gtk_stats *stats);

/* Work plan of (form, precision) over a built (rows, cols) block, built
 * here if no assembly built it yet: the kernel chosen and its launch shape,
 * as the autotuner and the benchmarks report them.  No counterpart in the
 * reference (its assembly loop has no launch shape). */
typedef struct gtk_plan_info_t {
  int matrix_cores;    /* 1: matrix-core batch kernel, 0: VALU kernel */
  int threads;         /* threads per workgroup */
  int grid;            /* workgroups launched */
  int tiles;           /* tiles of the column-owner plan */
  int max_tile_cells;  /* most cells (owned and halo) of one tile */
  int stage_cells;     /* cells staged per geometry pass (matrix-core kernel) */
  int lds_bytes;       /* LDS per workgroup */
  int tile_cap;        /* cells-per-tile cap the plan was built with (0: as many as fit in LDS) */
  int64_t batches;     /* batch records (matrix-core kernel) */
  int64_t tile_cells;  /* cells summed over the tiles: tile_cells / cells is the halo redundancy */
  int64_t cells;       /* cells of the mesh */
} gtk_plan_info_t;
int gtk_plan_info(gtk_ctx *ctx, int form, int precision, int rows, int cols, gtk_plan_info_t *out);
/* Rebuild the work plan of (form, precision) on a built block with a given
 * launch shape: waves per workgroup (matrix-core kernels 8, 12 or 16; the
 * VALU kernel 4, or 0 for its default; 0 everywhere picks the default) and
 * a cap on the cells of a tile (0: as many as fit in LDS).  The knobs of the
 * single-problem autotuner; no counterpart in the reference.  The next
 * assembly of that form and precision uses the new plan. */
int gtk_set_launch(gtk_ctx *ctx, int form, int precision, int rows, int cols, int waves, int max_cells);

/* Enqueue the matrix assembly of a (rows, cols) block on the context's
 * stream without synchronising (for timing loops); gtk_sync waits. */
int gtk_assemble_matrix_async(gtk_ctx *ctx, int form, int precision, int rows, int cols);
int gtk_sync(gtk_ctx *ctx);
/* Average device time per launch of the last timed run (HIP events on the
 * context stream around n launches). */
int gtk_time_matrix(gtk_ctx *ctx, int form, int precision, int rows, int cols, int launches, double *ms_per_launch);

typedef struct gtk_solve_stats {
  int outer_iterations;  /* FP64 refinement steps */
  int inner_iterations;  /* conjugate-gradient iterations, all steps */
  int converged;         /* 1 if the FP64 relative residual reached rtol */
  double rel_residual;   /* |b - A x| / |b| in FP64 at exit */
  double solve_ms;       /* wall time of the call */
} gtk_solve_stats;

/* Solve A x = b for the (GTK_FREE, GTK_FREE) block with the values last
 * assembled (they stay on the device), A positive definite from a
 * symmetric form (Laplace, mass, elasticity): its pattern is structurally
 * symmetric, its values symmetric up to the assembly's rounding (the solve
 * uses the values as assembled, in row order).  x: initial guess on input,
 * solution on output (host arrays of n_free).
 *  - GTK_MIXED: FP64 iterative refinement; each step solves A d = r/|r| with
 *    an FP32 Jacobi-preconditioned conjugate gradient to relative residual
 *    inner_rtol (at most max_inner iterations), then x += |r| d in FP64 and
 *    r = b - A x is recomputed in FP64; stops when |r| <= rtol |b| or after
 *    max_outer steps.
 *  - GTK_FP64: one FP64 Jacobi-preconditioned conjugate gradient to rtol
 *    (max_inner iterations), refined as above.
 * Returns GTK_OK also when rtol was not reached (stats->converged says).
 * Replaces: the linear solve of the assembled system,
 * PartitionedSolvers_linear_problem + solve (src/problems.jl:439-462). */
int gtk_solve(gtk_ctx *ctx, int precision, const double *b, double *x, double rtol, int max_outer, int max_inner,
              double inner_rtol, gtk_solve_stats *stats);

/* Batched assembly: many independent problems of uneven size assembled by
 * one persistent kernel launch over a flattened (problem, tile) queue (a
 * batch of hexahedra: one launch per element type, Q1 then Q2), captured
 * once in a HIP graph.  A problem is a context on the batch's
 * device with a (GTK_FREE, GTK_FREE) pattern built and its reference tables
 * and coefficients set, on scalar Q1/Q2/Q3 quadrilaterals with (k+1)^2
 * Gauss points, or (another batch) scalar Q1/Q2 hexahedra with (k+1)^3,
 * or scalar P1/P2 triangles with degree-2k rules; form GTK_LAPLACE or
 * GTK_MASS.  Each problem's values land
 * in its own context (gtk_get_values, gtk_values_device_ptr).  Contexts
 * must outlive the batch and have no *_async work pending at a run; new
 * coefficients or coordinates set between runs are picked up, a new mesh,
 * space or pattern needs gtk_batch_build again.
 * No counterpart in the reference, which assembles one problem per
 * GT.assemble_matrix / GT.assemble_vector call (src/problems.jl:244-274,
 * :319-350); a batch replaces a loop of such calls.  elem_type of
 * gtk_batch_set_tile_cells: 0, 1, 2 quadrilateral Q1, Q2, Q3; 3, 4 hexahedral Q1, Q2; 6, 7
 * triangular P1, P2;
 * max_cells caps the cells of a tile for that type (0: as many as fit in
 * LDS), the knob the autotuner sets. */
typedef struct gtk_batch gtk_batch;
int gtk_batch_create(int device, int form, int precision, gtk_batch **out);
int gtk_batch_destroy(gtk_batch *b);
const char *gtk_batch_last_error(const gtk_batch *b);
int gtk_batch_add(gtk_batch *b, gtk_ctx *ctx);
int gtk_batch_set_tile_cells(gtk_batch *b, int elem_type, int max_cells);
/* Waves per workgroup of the launch: 8 (default) or 12, the autotuner's
 * other knob. */
int gtk_batch_set_waves(gtk_batch *b, int waves);
/* Load vector of problem i (in gtk_batch_add order) over its free rows,
 * assembled by the same launch: f as in gtk_assemble_vector.  New values of
 * the same kind take effect at the next run; a new kind needs
 * gtk_batch_build again.  gtk_batch_get_vector copies the n_free values of
 * the last run to the host. */
int gtk_batch_set_load(gtk_batch *b, int problem, int fkind, const double *f);
int gtk_batch_get_vector(gtk_batch *b, int problem, double *out);
int gtk_batch_build(gtk_batch *b);
/* One graph launch assembling every problem; stats->kernel_ms is its
 * device time. */
int gtk_batch_run(gtk_batch *b, gtk_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
