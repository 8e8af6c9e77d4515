/*
 * gtk_oracle.h -- CPU oracle for the finite-element assembly path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (GalerkinToolkit.jl v0.6.3) used as the checker for the HIP
 * library.  Only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark may load it.  The product path never links or calls it.
 *
 * Every function names the reference file:line it follows in gtk_oracle.c.
 */
#ifndef GTK_ORACLE_H
#define GTK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem orc_problem;

/* Geometry / cell kinds. */
enum { ORC_POINT = 0, ORC_SEG = 1, ORC_QUAD = 2, ORC_HEX = 3, ORC_TRI = 4, ORC_TET = 5 };
/* Forms. */
enum { ORC_LAPLACE = 0, ORC_MASS = 1, ORC_ELASTICITY = 2 };
/* Row/column selectors (reference FREE / DIRICHLET). */
enum { ORC_FREE = 1, ORC_DIRICHLET = 2 };
/* Integer queries. */
enum {
  ORC_Q_DIM = 0, ORC_Q_NNODES, ORC_Q_NCELLS, ORC_Q_NVPC, ORC_Q_NVERT, ORC_Q_NEDGES,
  ORC_Q_NFACES2, ORC_Q_NDOFS, ORC_Q_NFREE, ORC_Q_NDIRI, ORC_Q_DOFS_PER_CELL,
  ORC_Q_NNZ, ORC_Q_NQ, ORC_Q_NPARENT0, ORC_Q_NPARENT1, ORC_Q_NPARENT2, ORC_Q_CELLKIND
};

/* cartesian_mesh(domain, cells; simplexify) with boundary faces + complexify.
 * D = 2 or 3; cells[D]; box = (x0,x1,y0,y1[,z0,z1]).  Returns NULL on error. */
orc_problem *orc_cartesian(int D, const int32_t *cells, const double *box, int simplexify);
void orc_free(orc_problem *p);
const char *orc_last_error(void);
int64_t orc_query(const orc_problem *p, int what);

/* Replace node coordinates (nnodes*D, node-major).  Topology is unchanged. */
int orc_set_coords(orc_problem *p, const double *x);
int orc_get_coords(const orc_problem *p, double *x);
int orc_get_cell_nodes(const orc_problem *p, int32_t *out);      /* 1-based */
int orc_get_node_vertex(const orc_problem *p, int32_t *out);     /* per node */
int orc_get_cell_faces(const orc_problem *p, int d, int32_t *out);   /* cell -> d-face ids */
int orc_get_face_vertices(const orc_problem *p, int d, int32_t *out);
int orc_get_cell_perm(const orc_problem *p, int d, int32_t *out);    /* permutation ids */
/* parent (boundary) d-faces and their group (local face index of the cube, 1-based) */
int orc_get_parent_faces(const orc_problem *p, int d, int32_t *nodes_out, int32_t *group_out);

/* lagrange_space(interior, order; tensor_size=(ncomp,), dirichlet_boundary).
 * diri_kind: 0 = none, 1 = boundary(mesh) (faces with one cell),
 *            2 = group "(D-1)-face-<group>" of the cartesian mesh. */
int orc_space(orc_problem *p, int order, int ncomp, int diri_kind, int group);
int orc_get_cell_dofs(const orc_problem *p, int32_t *out);   /* ncells*dofs_per_cell, GT signed encoding */

/* Reference element tables (for tests): node coordinates, values/gradients of
 * the shape functions of the current space at its quadrature points. */
int orc_set_quadrature(orc_problem *p, int degree);
int orc_get_quadrature(const orc_problem *p, double *xi, double *w);
int orc_tabulate(const orc_problem *p, int geometry, const double *xi, int npts, double *val, double *grad);

/* FP64 matrix assembly in reference order, then COO -> CSC.
 * coef: per-cell coefficient(s): LAPLACE/MASS: kappa (ncells) or NULL (=1);
 * ELASTICITY: (lambda, mu) per cell (ncells*2).  rows/cols: ORC_FREE/ORC_DIRICHLET.
 * The result is kept; fetch it with orc_get_csc. */
int orc_assemble_matrix(orc_problem *p, int form, const double *coef, int rows, int cols);
/* Threads of orc_assemble_matrix (default 1; the result does not depend on
 * it) and the wall time in ms of its last call's phases: symbolic count,
 * numeric cell loop, COO -> CSC compress. */
int orc_set_threads(int n);
int orc_phase_ms(double *out3);
int orc_get_csc(const orc_problem *p, int32_t *colptr, int32_t *rowval, double *nzval);
int64_t orc_csc_ncols(const orc_problem *p);

/* FP64 load vector: b_i = int f . v_i.
 * f_kind: 0 = constant vector fconst[ncomp]; 1 = per-cell constant (fcell[ncells*ncomp]);
 *         2 = nodal discrete field fnodal[ndofs] indexed by (free dof) / (-dirichlet dof + nfree).
 * rows: ORC_FREE or ORC_DIRICHLET. */
int orc_assemble_vector(orc_problem *p, int f_kind, const double *fdata, int rows, double *b);

/* Dense helpers used by the tests. */
int orc_gauss_legendre(int n, double *x, double *w);

#ifdef __cplusplus
}
#endif
#endif
