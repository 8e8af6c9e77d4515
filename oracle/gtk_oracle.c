/*
 * gtk_oracle.c -- CPU oracle for the finite-element assembly path.
 *
 * TEST INFRASTRUCTURE ONLY.  Plain-C restatement of the reference algorithm
 * (GalerkinToolkit.jl v0.6.3) used to check the HIP library.  It is loaded by
 * tests/, by the smoke check and by the CPU-baseline leg of the benchmark,
 * never by the product path.
 *
 * Loop order, numbering rules and accumulation order follow the reference
 * functions cited at each definition.  Arithmetic that the reference delegates
 * to third-party packages is restated from the published methods:
 *   - Gauss-Legendre nodes (FastGaussQuadrature.gausslegendre): Newton on P_n;
 *   - A\I on the Vandermonde matrix (LinearAlgebra/LAPACK): LU with partial
 *     pivoting, as LAPACK getrf/getrs;
 *   - transpose(J)\g (StaticArrays): explicit 2x2 / 3x3 inverse;
 *   - COO -> CSC (PartitionedArrays.sparse_matrix): column-major sort, rows
 *     ascending, duplicates summed in COO order.
 * These agree with the packages to a few ulp, not bit for bit; values are
 * compared to a relative tolerance, integer data bit for bit.
 */
#include "gtk_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define INVALID 0
#define MAXLF 12 /* max local faces of one dimension (hex edges) */

static char g_err[512];
/* threads of orc_assemble_matrix (CPU baseline at several core counts) and
 * the wall time of its three phases: symbolic count, numeric loop, compress */
static int g_threads = 1;
static double g_phase_ms[3];
int orc_set_threads(int n) { g_threads = n > 0 ? n : 1; return 0; }
int orc_phase_ms(double *out) { memcpy(out, g_phase_ms, sizeof g_phase_ms); return 0; }
static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }
const char *orc_last_error(void) { return g_err; }

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* ------------------------------------------------------------------------ */
/* Reference geometries: domain.jl:213-268 (unit n-cubes), :425-470 (unit   */
/* simplices).  Vertex lists are 1-based as in the reference.               */
/* ------------------------------------------------------------------------ */
typedef struct {
  int dim, nv;
  double v[8][3];
  int nf[4];          /* number of local d-faces */
  int fnv[4];         /* vertices per local d-face */
  int f[4][12][8];    /* local d-face -> local vertices (1-based) */
  int ftype[4];       /* geometry kind of the d-faces */
  int simplex;
} refgeo;

static refgeo G[6];

static void init_geo(void) {
  static int done = 0;
  if (done) return;
  done = 1;
  memset(G, 0, sizeof G);
  /* point */
  G[ORC_POINT].dim = 0; G[ORC_POINT].nv = 1; G[ORC_POINT].simplex = 1;
  G[ORC_POINT].nf[0] = 1; G[ORC_POINT].fnv[0] = 1; G[ORC_POINT].f[0][0][0] = 1;
  G[ORC_POINT].ftype[0] = ORC_POINT;
  /* segment, domain.jl:191-211 */
  refgeo *s = &G[ORC_SEG];
  s->dim = 1; s->nv = 2; s->simplex = 1; s->v[1][0] = 1;
  s->nf[0] = 2; s->fnv[0] = 1; s->f[0][0][0] = 1; s->f[0][1][0] = 2; s->ftype[0] = ORC_POINT;
  s->nf[1] = 1; s->fnv[1] = 2; s->f[1][0][0] = 1; s->f[1][0][1] = 2; s->ftype[1] = ORC_SEG;
  /* quad, domain.jl:213-236 */
  refgeo *q = &G[ORC_QUAD];
  q->dim = 2; q->nv = 4;
  double qv[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
  for (int i = 0; i < 4; i++) { q->v[i][0] = qv[i][0]; q->v[i][1] = qv[i][1]; }
  q->nf[0] = 4; q->fnv[0] = 1; for (int i = 0; i < 4; i++) q->f[0][i][0] = i + 1;
  q->ftype[0] = ORC_POINT;
  int qe[4][2] = {{1, 2}, {3, 4}, {1, 3}, {2, 4}};
  q->nf[1] = 4; q->fnv[1] = 2; q->ftype[1] = ORC_SEG;
  for (int i = 0; i < 4; i++) { q->f[1][i][0] = qe[i][0]; q->f[1][i][1] = qe[i][1]; }
  q->nf[2] = 1; q->fnv[2] = 4; for (int i = 0; i < 4; i++) q->f[2][0][i] = i + 1; q->ftype[2] = ORC_QUAD;
  /* hex, domain.jl:238-268 */
  refgeo *h = &G[ORC_HEX];
  h->dim = 3; h->nv = 8;
  for (int i = 0; i < 8; i++) { h->v[i][0] = i & 1; h->v[i][1] = (i >> 1) & 1; h->v[i][2] = (i >> 2) & 1; }
  h->nf[0] = 8; h->fnv[0] = 1; for (int i = 0; i < 8; i++) h->f[0][i][0] = i + 1; h->ftype[0] = ORC_POINT;
  int he[12][2] = {{1, 2}, {3, 4}, {1, 3}, {2, 4}, {5, 6}, {7, 8}, {5, 7}, {6, 8}, {1, 5}, {3, 7}, {2, 6}, {4, 8}};
  h->nf[1] = 12; h->fnv[1] = 2; h->ftype[1] = ORC_SEG;
  for (int i = 0; i < 12; i++) { h->f[1][i][0] = he[i][0]; h->f[1][i][1] = he[i][1]; }
  int hf[6][4] = {{1, 2, 3, 4}, {5, 6, 7, 8}, {1, 2, 5, 6}, {3, 4, 7, 8}, {1, 3, 5, 7}, {2, 4, 6, 8}};
  h->nf[2] = 6; h->fnv[2] = 4; h->ftype[2] = ORC_QUAD;
  for (int i = 0; i < 6; i++) for (int k = 0; k < 4; k++) h->f[2][i][k] = hf[i][k];
  h->nf[3] = 1; h->fnv[3] = 8; for (int i = 0; i < 8; i++) h->f[3][0][i] = i + 1; h->ftype[3] = ORC_HEX;
  /* triangle, domain.jl:425-445 */
  refgeo *t = &G[ORC_TRI];
  t->dim = 2; t->nv = 3; t->simplex = 1;
  double tv[3][2] = {{0, 0}, {1, 0}, {0, 1}};
  for (int i = 0; i < 3; i++) { t->v[i][0] = tv[i][0]; t->v[i][1] = tv[i][1]; }
  t->nf[0] = 3; t->fnv[0] = 1; for (int i = 0; i < 3; i++) t->f[0][i][0] = i + 1; t->ftype[0] = ORC_POINT;
  int te[3][2] = {{1, 2}, {1, 3}, {2, 3}};
  t->nf[1] = 3; t->fnv[1] = 2; t->ftype[1] = ORC_SEG;
  for (int i = 0; i < 3; i++) { t->f[1][i][0] = te[i][0]; t->f[1][i][1] = te[i][1]; }
  t->nf[2] = 1; t->fnv[2] = 3; for (int i = 0; i < 3; i++) t->f[2][0][i] = i + 1; t->ftype[2] = ORC_TRI;
  /* tetrahedron, domain.jl:447-470 */
  refgeo *T = &G[ORC_TET];
  T->dim = 3; T->nv = 4; T->simplex = 1;
  double Tv[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 4; i++) for (int k = 0; k < 3; k++) T->v[i][k] = Tv[i][k];
  T->nf[0] = 4; T->fnv[0] = 1; for (int i = 0; i < 4; i++) T->f[0][i][0] = i + 1; T->ftype[0] = ORC_POINT;
  int Te[6][2] = {{1, 2}, {1, 3}, {2, 3}, {1, 4}, {2, 4}, {3, 4}};
  T->nf[1] = 6; T->fnv[1] = 2; T->ftype[1] = ORC_SEG;
  for (int i = 0; i < 6; i++) { T->f[1][i][0] = Te[i][0]; T->f[1][i][1] = Te[i][1]; }
  int Tf[4][3] = {{1, 2, 3}, {1, 2, 4}, {1, 3, 4}, {2, 3, 4}};
  T->nf[2] = 4; T->fnv[2] = 3; T->ftype[2] = ORC_TRI;
  for (int i = 0; i < 4; i++) for (int k = 0; k < 3; k++) T->f[2][i][k] = Tf[i][k];
  T->nf[3] = 1; T->fnv[3] = 4; for (int i = 0; i < 4; i++) T->f[3][0][i] = i + 1; T->ftype[3] = ORC_TET;
}

/* Linear (order-1) shape functions of a reference geometry at xi.
 * For n-cubes the nodes are lexicographic {0,1}^D (tensor products of 1-t, t);
 * for simplices barycentric (1 - sum, x1, x2, ...).  These are the maps used
 * by the reference to place face nodes (space.jl:1598-1640). */
static void linear_shape(int geo, const double *xi, double *N) {
  switch (geo) {
  case ORC_POINT: N[0] = 1; break;
  case ORC_SEG: N[0] = 1 - xi[0]; N[1] = xi[0]; break;
  case ORC_QUAD:
    N[0] = (1 - xi[0]) * (1 - xi[1]); N[1] = xi[0] * (1 - xi[1]);
    N[2] = (1 - xi[0]) * xi[1]; N[3] = xi[0] * xi[1]; break;
  case ORC_HEX:
    for (int i = 0; i < 8; i++) {
      double a = (i & 1) ? xi[0] : 1 - xi[0];
      double b = ((i >> 1) & 1) ? xi[1] : 1 - xi[1];
      double c = ((i >> 2) & 1) ? xi[2] : 1 - xi[2];
      N[i] = a * b * c;
    }
    break;
  case ORC_TRI: N[0] = 1 - xi[0] - xi[1]; N[1] = xi[0]; N[2] = xi[1]; break;
  case ORC_TET: N[0] = 1 - xi[0] - xi[1] - xi[2]; N[1] = xi[0]; N[2] = xi[1]; N[3] = xi[2]; break;
  }
}

/* Vertex permutations of a reference geometry, domain.jl:53-100.
 * D==0: identity; D>2: identity only; simplices (and segments): all
 * permutations in lexicographic order (Combinatorics.permutations);
 * squares: the permutations that keep sum_q |det J| (one Gauss point). */
static int next_perm(int *a, int n) {
  int i = n - 2;
  while (i >= 0 && a[i] >= a[i + 1]) i--;
  if (i < 0) return 0;
  int j = n - 1;
  while (a[j] <= a[i]) j--;
  int t = a[i]; a[i] = a[j]; a[j] = t;
  for (int l = i + 1, r = n - 1; l < r; l++, r--) { t = a[l]; a[l] = a[r]; a[r] = t; }
  return 1;
}

static int vertex_permutations(int geo, int perms[][8]) {
  init_geo();
  const refgeo *g = &G[geo];
  int nv = g->nv, np = 0;
  if (g->dim == 0 || g->dim > 2) {
    for (int i = 0; i < nv; i++) perms[0][i] = i + 1;
    return 1;
  }
  int a[8];
  for (int i = 0; i < nv; i++) a[i] = i + 1;
  if (g->simplex) {
    do { memcpy(perms[np++], a, sizeof(int) * nv); } while (next_perm(a, nv));
    return np;
  }
  /* square: compute_volume, domain.jl:102-117, with the degree-1 rule
   * (one point at the centre, weight not used) */
  double xi[2] = {0.5, 0.5};
  /* gradients of the bilinear shape functions at the centre */
  double gN[4][2] = {{-0.5, -0.5}, {0.5, -0.5}, {-0.5, 0.5}, {0.5, 0.5}};
  (void)xi;
  double refvol = 0;
  {
    double J[2][2] = {{0, 0}, {0, 0}};
    for (int n = 0; n < 4; n++)
      for (int r = 0; r < 2; r++) for (int c = 0; c < 2; c++) J[r][c] += g->v[n][r] * gN[n][c];
    refvol = fabs(J[0][0] * J[1][1] - J[0][1] * J[1][0]);
  }
  do {
    double J[2][2] = {{0, 0}, {0, 0}};
    for (int n = 0; n < 4; n++)
      for (int r = 0; r < 2; r++) for (int c = 0; c < 2; c++) J[r][c] += g->v[a[n] - 1][r] * gN[n][c];
    double vol2 = fabs(J[0][0] * J[1][1] - J[0][1] * J[1][0]);
    if (fabs((refvol + vol2) - 2 * refvol) <= 1.4901161193847656e-8 * fmax(fabs(refvol + vol2), 2 * refvol))
      memcpy(perms[np++], a, sizeof(int) * nv);
  } while (next_perm(a, nv));
  return np;
}

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre (FastGaussQuadrature.gausslegendre + gauss_legendre,       */
/* quadrature.jl:60-66): nodes ascending on [-1,1], Newton iteration on P_n, */
/* then mapped to [0,1].                                                     */
/* ------------------------------------------------------------------------ */
int orc_gauss_legendre(int n, double *x, double *w) {
  if (n < 1 || n > 64) { set_err("gauss_legendre: n out of range"); return -1; }
  for (int i = 0; i < n; i++) {
    /* root i (descending cos) -> store ascending */
    double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1, p1 = z;
      for (int k = 2; k <= n; k++) { double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
      if (n == 1) { p1 = z; p0 = 1; }
      dp = n * (z * p1 - p0) / (z * z - 1);
      double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    double p0 = 1, p1 = z;
    for (int k = 2; k <= n; k++) { double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
    dp = n * (z * p1 - p0) / (z * z - 1);
    double wi = 2.0 / ((1 - z * z) * dp * dp);
    int j = n - 1 - i;
    x[j] = 0.5 * z + 0.5;
    w[j] = 0.5 * wi;
  }
  if (n == 1) { x[0] = 0.5; w[0] = 1.0; }
  return 0;
}

/* Gauss-Jacobi (FastGaussQuadrature.gaussjacobi, quadrature.jl:68-74) for the
 * Duffy rule: Golub-Welsch restated as Newton on the Jacobi polynomial. */
static void jacobi_p(int n, double a, double b, double x, double *p, double *dp) {
  double p0 = 1, p1 = 0.5 * (a - b + (a + b + 2) * x);
  if (n == 0) { *p = 1; *dp = 0; return; }
  for (int k = 2; k <= n; k++) {
    double a1 = 2 * k * (k + a + b) * (2 * k + a + b - 2);
    double a2 = (2 * k + a + b - 1) * (a * a - b * b);
    double a3 = (2 * k + a + b - 2) * (2 * k + a + b - 1) * (2 * k + a + b);
    double a4 = 2 * (k + a - 1) * (k + b - 1) * (2 * k + a + b);
    double p2 = ((a2 + a3 * x) * p1 - a4 * p0) / a1;
    p0 = p1; p1 = p2;
  }
  *p = p1;
  /* derivative: d/dx P_n^{a,b} = 0.5 (n+a+b+1) P_{n-1}^{a+1,b+1} */
  double q0 = 1, q1 = 0.5 * ((a + 1) - (b + 1) + (a + b + 4) * x), A = a + 1, B = b + 1;
  if (n - 1 == 0) q1 = 1;
  for (int k = 2; k <= n - 1; k++) {
    double a1 = 2 * k * (k + A + B) * (2 * k + A + B - 2);
    double a2 = (2 * k + A + B - 1) * (A * A - B * B);
    double a3 = (2 * k + A + B - 2) * (2 * k + A + B - 1) * (2 * k + A + B);
    double a4 = 2 * (k + A - 1) * (k + B - 1) * (2 * k + A + B);
    double q2 = ((a2 + a3 * x) * q1 - a4 * q0) / a1;
    q0 = q1; q1 = q2;
  }
  *dp = 0.5 * (n + a + b + 1) * q1;
}

static void gauss_jacobi01(int n, double a, double b, double *x, double *w) {
  /* roots of P_n^{a,b} on [-1,1]: bracket sign changes on a fine grid, then
   * bisection; weights by the Gauss-Jacobi formula; mapped to [0,1]
   * (quadrature.jl:68-74). */
  double r[64];
  int nr = 0;
  const int M = 20000;
  double xa = -1.0, pa, dpa;
  jacobi_p(n, a, b, xa, &pa, &dpa);
  for (int s = 1; s <= M && nr < n; s++) {
    double xb = -1.0 + 2.0 * s / M, pb, dpb;
    jacobi_p(n, a, b, xb, &pb, &dpb);
    if (pa == 0.0) { r[nr++] = xa; }
    else if ((pa < 0) != (pb < 0)) {
      double lo = xa, hi = xb, plo = pa;
      for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi), pm, dpm;
        jacobi_p(n, a, b, mid, &pm, &dpm);
        if ((pm < 0) == (plo < 0)) { lo = mid; plo = pm; } else hi = mid;
        if (hi - lo < 1e-17) break;
      }
      r[nr++] = 0.5 * (lo + hi);
    }
    xa = xb; pa = pb;
  }
  double lg = lgamma(n + a + 1) + lgamma(n + b + 1) - lgamma(n + 1) - lgamma(n + a + b + 1);
  for (int k = 0; k < n; k++) {
    double p, dp;
    jacobi_p(n, a, b, r[k], &p, &dp);
    double wk = exp(lg) * pow(2.0, a + b + 1) / ((1 - r[k] * r[k]) * dp * dp);
    x[k] = 0.5 * r[k] + 0.5;
    w[k] = 0.5 * wk;
  }
}

/* ------------------------------------------------------------------------ */
/* Reference Lagrange elements, space.jl:981-1031, :1127-1177, :1204-1251.   */
/* ------------------------------------------------------------------------ */
typedef struct {
  int geo, dim, order, ncomp;
  int nnodes, ndofs;
  int expo[128][3];
  double xn[128][3];
  double *B; /* nnodes x nnodes, B[j*nnodes+k]: coefficient of monomial j in shape function k */
  int nlf[4];
  int fn_cnt[4];
  int *face_nodes[4]; /* closure nodes of local d-faces, 1-based cell nodes */
  int own_cnt[4];
  int *own_nodes[4];  /* interior (own) nodes of local d-faces */
  int nperm[4];
  int *perm[4];       /* [pindex][i] node_to_pnode of own nodes, 1-based */
} lagrange;

/* monomial_exponents (space.jl:1127-1145): CartesianIndices over 0:k, x fastest;
 * P spaces keep sum(e) <= k. */
static int monomials(int geo, int k, int e[][3]) {
  init_geo();
  int D = G[geo].dim, n = 0;
  if (D == 0) { e[0][0] = e[0][1] = e[0][2] = 0; return 1; }
  int lim[3] = {k, D > 1 ? k : 0, D > 2 ? k : 0};
  for (int c = 0; c <= lim[2]; c++)
    for (int b = 0; b <= lim[1]; b++)
      for (int a = 0; a <= lim[0]; a++) {
        if (G[geo].simplex && a + b + c > k) continue;
        e[n][0] = a; e[n][1] = b; e[n][2] = c; n++;
      }
  return n;
}

/* boundary test for interior_nodes (space.jl:1345-1364): a node is interior
 * when it touches no lower-dimensional face of its reference geometry. */
static int node_on_boundary(int geo, int k, const int *e) {
  int D = G[geo].dim;
  if (D == 0) return 0;
  int s = 0;
  for (int d = 0; d < D; d++) {
    if (e[d] == 0) return 1;
    if (!G[geo].simplex && e[d] == k) return 1;
    s += e[d];
  }
  if (G[geo].simplex && s == k) return 1;
  return 0;
}

static double ipow(double x, int e) {
  double r = 1;
  for (int i = 0; i < e; i++) r *= x;
  return r;
}

static double mono_val(const int *e, const double *x, int D) {
  double r = 1;
  for (int d = 0; d < D; d++) r *= ipow(x[d], e[d]);
  return r;
}

static void mono_grad(const int *e, const double *x, int D, double *g) {
  for (int d = 0; d < D; d++) {
    double r = 1;
    for (int dd = 0; dd < D; dd++) {
      if (dd == d) r *= e[dd] > 0 ? e[dd] * ipow(x[dd], e[dd] - 1) : 0.0;
      else r *= ipow(x[dd], e[dd]);
    }
    g[d] = r;
  }
}

/* B = A \ I with A[i][j] = monomial_j(node_i) (tabulator, space.jl:960-970);
 * LU with partial pivoting. */
static void vandermonde_inverse(lagrange *L) {
  int n = L->nnodes, D = L->dim;
  double *A = xcalloc((size_t)n * n, sizeof(double));
  int *piv = xcalloc(n, sizeof(int));
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) A[i * n + j] = mono_val(L->expo[j], L->xn[i], D);
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++) if (fabs(A[r * n + c]) > fabs(A[p * n + c])) p = r;
    piv[c] = p;
    if (p != c) for (int j = 0; j < n; j++) { double t = A[c * n + j]; A[c * n + j] = A[p * n + j]; A[p * n + j] = t; }
    for (int r = c + 1; r < n; r++) {
      A[r * n + c] /= A[c * n + c];
      for (int j = c + 1; j < n; j++) A[r * n + j] -= A[r * n + c] * A[c * n + j];
    }
  }
  L->B = xcalloc((size_t)n * n, sizeof(double));
  double *y = xcalloc(n, sizeof(double));
  for (int k = 0; k < n; k++) {
    for (int i = 0; i < n; i++) y[i] = (i == k);
    for (int c = 0; c < n; c++) if (piv[c] != c) { double t = y[c]; y[c] = y[piv[c]]; y[piv[c]] = t; }
    for (int i = 0; i < n; i++) for (int j = 0; j < i; j++) y[i] -= A[i * n + j] * y[j];
    for (int i = n - 1; i >= 0; i--) {
      for (int j = i + 1; j < n; j++) y[i] -= A[i * n + j] * y[j];
      y[i] /= A[i * n + i];
    }
    for (int j = 0; j < n; j++) L->B[j * n + k] = y[j];
  }
  free(A); free(piv); free(y);
}

static void lagrange_free(lagrange *L) {
  free(L->B);
  for (int d = 0; d < 4; d++) { free(L->face_nodes[d]); free(L->own_nodes[d]); free(L->perm[d]); }
  memset(L, 0, sizeof *L);
}

/* Node coordinates of a face element placed on a local face of the cell and
 * matched to cell nodes by rounding order*x (complexify(::LagrangeFaceSpace),
 * space.jl:1598-1640). */
static int match_node(const lagrange *L, const double *x) {
  int k = L->order;
  for (int i = 0; i < L->nnodes; i++) {
    int ok = 1;
    for (int d = 0; d < L->dim; d++)
      if (lround(k * x[d]) != lround(k * L->xn[i][d])) { ok = 0; break; }
    if (ok) return i + 1;
  }
  return 0;
}

static int lagrange_init(lagrange *L, int geo, int order, int ncomp) {
  init_geo();
  memset(L, 0, sizeof *L);
  L->geo = geo; L->dim = G[geo].dim; L->order = order; L->ncomp = ncomp;
  L->nnodes = monomials(geo, order, L->expo);
  L->ndofs = L->nnodes * ncomp;
  for (int i = 0; i < L->nnodes; i++)
    for (int d = 0; d < L->dim; d++) L->xn[i][d] = (double)L->expo[i][d] / order;
  vandermonde_inverse(L);
  int D = L->dim;
  for (int d = 0; d < D; d++) {
    const refgeo *g = &G[geo];
    int ft = g->ftype[d];
    int fexpo[128][3];
    int fn = monomials(ft, order, fexpo);
    L->nlf[d] = g->nf[d];
    L->fn_cnt[d] = fn;
    L->face_nodes[d] = xcalloc((size_t)g->nf[d] * fn, sizeof(int));
    /* interior nodes of the face element (lexicographic, findall) */
    int own[128], nown = 0;
    for (int i = 0; i < fn; i++) if (!node_on_boundary(ft, order, fexpo[i])) own[nown++] = i;
    L->own_cnt[d] = nown;
    L->own_nodes[d] = xcalloc((size_t)g->nf[d] * (nown ? nown : 1), sizeof(int));
    for (int f = 0; f < g->nf[d]; f++) {
      for (int i = 0; i < fn; i++) {
        double xi[3] = {0, 0, 0}, N[8], x[3] = {0, 0, 0};
        for (int dd = 0; dd < G[ft].dim; dd++) xi[dd] = (double)fexpo[i][dd] / order;
        linear_shape(ft, xi, N);
        for (int v = 0; v < g->fnv[d]; v++)
          for (int c = 0; c < D; c++) x[c] += N[v] * g->v[g->f[d][f][v] - 1][c];
        int node = match_node(L, x);
        if (!node) { set_err("face node matching failed"); return -1; }
        L->face_nodes[d][f * fn + i] = node;
      }
      for (int i = 0; i < nown; i++) L->own_nodes[d][f * nown + i] = L->face_nodes[d][f * fn + own[i]];
    }
    /* interior node permutations (space.jl:1447-1494) */
    int vp[24][8];
    int np = vertex_permutations(ft, vp);
    L->nperm[d] = np;
    L->perm[d] = xcalloc((size_t)np * (nown ? nown : 1), sizeof(int));
    for (int p = 0; p < np; p++) {
      if (np == 1) { for (int i = 0; i < nown; i++) L->perm[d][p * nown + i] = i + 1; continue; }
      for (int iq = 0; iq < nown; iq++) {
        double q[3] = {0, 0, 0}, N[8], y[3] = {0, 0, 0};
        for (int dd = 0; dd < G[ft].dim; dd++) q[dd] = (double)fexpo[own[iq]][dd] / order;
        linear_shape(ft, q, N);
        for (int v = 0; v < G[ft].nv; v++)
          for (int c = 0; c < G[ft].dim; c++) y[c] += N[v] * G[ft].v[vp[p][v] - 1][c];
        int pn = INVALID;
        for (int i = 0; i < nown; i++) {
          double s = 0;
          for (int c = 0; c < G[ft].dim; c++) {
            double t = (double)fexpo[own[i]][c] / order - y[c];
            s += t * t;
          }
          if (fabs(sqrt(s)) <= 1.4901161193847656e-8 * (1 + sqrt(s))) { pn = i + 1; break; }
        }
        L->perm[d][p * nown + iq] = pn;
      }
    }
  }
  /* cell interior */
  {
    int own[128], nown = 0;
    for (int i = 0; i < L->nnodes; i++) if (!node_on_boundary(geo, order, L->expo[i])) own[nown++] = i + 1;
    L->nlf[D] = 1;
    L->fn_cnt[D] = L->nnodes;
    L->face_nodes[D] = xcalloc(L->nnodes, sizeof(int));
    for (int i = 0; i < L->nnodes; i++) L->face_nodes[D][i] = i + 1;
    L->own_cnt[D] = nown;
    L->own_nodes[D] = xcalloc(nown ? nown : 1, sizeof(int));
    memcpy(L->own_nodes[D], own, sizeof(int) * nown);
    L->nperm[D] = 1;
    L->perm[D] = xcalloc(nown ? nown : 1, sizeof(int));
    for (int i = 0; i < nown; i++) L->perm[D][i] = i + 1;
  }
  return 0;
}

/* shape function values / reference gradients (tabulator, space.jl:960-970) */
static void lagrange_tabulate(const lagrange *L, const double *x, double *val, double *grad) {
  int n = L->nnodes, D = L->dim;
  double m[128], gm[128][3];
  for (int j = 0; j < n; j++) { m[j] = mono_val(L->expo[j], x, D); mono_grad(L->expo[j], x, D, gm[j]); }
  for (int k = 0; k < n; k++) {
    double v = 0, g[3] = {0, 0, 0};
    for (int j = 0; j < n; j++) {
      v += m[j] * L->B[j * n + k];
      for (int d = 0; d < D; d++) g[d] += gm[j][d] * L->B[j * n + k];
    }
    if (val) val[k] = v;
    if (grad) for (int d = 0; d < D; d++) grad[k * D + d] = g[d];
  }
}

/* ------------------------------------------------------------------------ */
/* Problem state                                                             */
/* ------------------------------------------------------------------------ */
struct orc_problem {
  int D, cellkind, simplex;
  int64_t nnodes, ncells;
  int nvc;
  double *x;
  int32_t *cell_nodes;
  /* parent faces (mesh with boundary), d < D */
  int64_t npf[3];
  int npfv[3];
  int32_t *pf_nodes[3];
  int32_t *pf_group[3];
  /* topology */
  int64_t nvert;
  int32_t *node_vertex, *vertex_node;
  int64_t nf[4];
  int fv[4];              /* vertices per d-face */
  int32_t *face_vertices[4];
  int32_t *cell_faces[4]; /* ncells * nlf(D,d) */
  int32_t *cell_perm[4];
  /* space */
  int has_space;
  lagrange L;
  int ncomp, order;
  int64_t ndofs, nfree, ndiri;
  int32_t *cell_dofs;
  /* quadrature */
  int nq;
  double qx[128][3], qw[128];
  /* CSC result */
  int64_t csc_ncols, csc_nnz;
  int32_t *colptr, *rowval;
  double *nzval;
};

static void free_csc(orc_problem *p) {
  free(p->colptr); free(p->rowval); free(p->nzval);
  p->colptr = p->rowval = NULL; p->nzval = NULL; p->csc_nnz = 0; p->csc_ncols = 0;
}

void orc_free(orc_problem *p) {
  if (!p) return;
  free(p->x); free(p->cell_nodes);
  for (int d = 0; d < 3; d++) { free(p->pf_nodes[d]); free(p->pf_group[d]); }
  free(p->node_vertex); free(p->vertex_node);
  for (int d = 0; d < 4; d++) { free(p->face_vertices[d]); free(p->cell_faces[d]); free(p->cell_perm[d]); }
  if (p->has_space) lagrange_free(&p->L);
  free(p->cell_dofs);
  free_csc(p);
  free(p);
}

/* ------------------------------------------------------------------------ */
/* Jagged helpers: generate_face_coboundary (topology.jl:443-465)            */
/* ------------------------------------------------------------------------ */
typedef struct { int64_t n; int64_t *ptr; int32_t *data; } jagged;
static void jfree(jagged *j) { free(j->ptr); free(j->data); j->ptr = NULL; j->data = NULL; }

/* face (fixed width w, 1-based ids into [1,nm]) -> coboundary lists in
 * increasing face order */
static jagged coboundary(const int32_t *f2m, int64_t nf, int w, int64_t nm) {
  jagged J;
  J.n = nm;
  J.ptr = xcalloc(nm + 1, sizeof(int64_t));
  for (int64_t f = 0; f < nf; f++) for (int i = 0; i < w; i++) J.ptr[f2m[f * w + i]]++;
  for (int64_t m = 0; m < nm; m++) J.ptr[m + 1] += J.ptr[m];
  J.data = xcalloc(J.ptr[nm] + 1, sizeof(int32_t));
  int64_t *fill = xcalloc(nm + 1, sizeof(int64_t));
  for (int64_t f = 0; f < nf; f++)
    for (int i = 0; i < w; i++) {
      int64_t m = f2m[f * w + i] - 1;
      J.data[J.ptr[m] + fill[m]++] = (int32_t)(f + 1);
    }
  free(fill);
  return J;
}

/* same_valid_ids(a,b,idsa,idsb), topology.jl:705-714: equal vertex sets */
static int same_set(const int32_t *a, int na, const int32_t *b, int nb) {
  for (int i = 0; i < na; i++) {
    int f = 0;
    for (int j = 0; j < nb; j++) if (a[i] == b[j]) { f = 1; break; }
    if (!f) return 0;
  }
  for (int j = 0; j < nb; j++) {
    int f = 0;
    for (int i = 0; i < na; i++) if (a[i] == b[j]) { f = 1; break; }
    if (!f) return 0;
  }
  return 1;
}

/* generate_face_boundary, topology.jl:1591-1704.
 * n-faces (vertices nfv, width wn, kind with local d-faces lf) produce the
 * d-faces; parent d-faces keep ids 1..npar; returns n-face -> d-face ids. */
static int32_t *generate_face_boundary(const int32_t *nface_v, int64_t nn, int wn, const jagged *v2n,
                                       const int32_t *pface_v, int64_t npar, int wd, const jagged *v2p,
                                       int nlf, const int lf[][8], int64_t *ndfaces_out) {
  int32_t *out = xcalloc((size_t)nn * nlf, sizeof(int32_t));
  int64_t newd = npar;
  int maxn = 0;
  for (int64_t v = 0; v < v2n->n; v++) { int c = (int)(v2n->ptr[v + 1] - v2n->ptr[v]); if (c > maxn) maxn = c; }
  int32_t *D1 = xcalloc(maxn + 1, sizeof(int32_t));
  for (int64_t F = 0; F < nn; F++) {
    const int32_t *Fv = nface_v + F * wn;
    for (int l = 0; l < nlf; l++) {
      if (out[F * nlf + l] != INVALID) continue;
      int32_t lv[8];
      for (int i = 0; i < wd; i++) lv[i] = Fv[lf[l][i] - 1];
      /* look for an existing parent d-face with this vertex set */
      int32_t d2 = INVALID;
      for (int i = 0; i < wd && d2 == INVALID; i++) {
        int32_t v = lv[i];
        for (int64_t k = v2p->ptr[v - 1]; k < v2p->ptr[v]; k++) {
          int32_t d1 = v2p->data[k];
          if (same_set(lv, wd, pface_v + (int64_t)(d1 - 1) * wd, wd)) { d2 = d1; break; }
        }
      }
      if (d2 == INVALID) d2 = (int32_t)(++newd);
      /* n-faces around: intersection of the vertices' coboundaries */
      int n1 = 0;
      for (int i = 0; i <= maxn; i++) D1[i] = INVALID;
      {
        int32_t v = lv[0];
        for (int64_t k = v2n->ptr[v - 1]; k < v2n->ptr[v]; k++) D1[n1++] = v2n->data[k];
      }
      for (int i = 1; i < wd; i++) {
        int32_t v = lv[i];
        for (int a = 0; a < n1; a++) {
          if (D1[a] == INVALID) continue;
          int f = 0;
          for (int64_t k = v2n->ptr[v - 1]; k < v2n->ptr[v]; k++) if (v2n->data[k] == D1[a]) { f = 1; break; }
          if (!f) D1[a] = INVALID;
        }
      }
      for (int a = 0; a < n1; a++) {
        int32_t F1 = D1[a];
        if (F1 == INVALID) continue;
        const int32_t *F1v = nface_v + (int64_t)(F1 - 1) * wn;
        int l2 = -1;
        for (int l1 = 0; l1 < nlf; l1++) {
          int32_t lv1[8];
          for (int i = 0; i < wd; i++) lv1[i] = F1v[lf[l1][i] - 1];
          if (same_set(lv, wd, lv1, wd)) { l2 = l1; break; }
        }
        if (l2 < 0) { free(D1); free(out); set_err("generate_face_boundary: local face not found"); return NULL; }
        out[(int64_t)(F1 - 1) * nlf + l2] = d2;
      }
    }
  }
  free(D1);
  *ndfaces_out = newd;
  return out;
}

/* generate_face_vertices, topology.jl:1350-1415: parent d-faces keep their
 * vertex order, other d-faces take it from the first n-face that has them. */
static int32_t *generate_face_vertices(int64_t nd, const int32_t *pface_v, int64_t npar, int wd,
                                       const int32_t *nface_v, int64_t nn, int wn, const int32_t *nface_d,
                                       int nlf, const int lf[][8]) {
  int32_t *out = xcalloc((size_t)nd * wd, sizeof(int32_t));
  char *touched = xcalloc(nd, 1);
  for (int64_t f = 0; f < npar; f++) {
    memcpy(out + f * wd, pface_v + f * wd, sizeof(int32_t) * wd);
    touched[f] = 1;
  }
  for (int64_t F = 0; F < nn; F++)
    for (int l = 0; l < nlf; l++) {
      int32_t d = nface_d[F * nlf + l];
      if (touched[d - 1]) continue;
      for (int i = 0; i < wd; i++) out[(int64_t)(d - 1) * wd + i] = nface_v[F * wn + lf[l][i] - 1];
      touched[d - 1] = 1;
    }
  free(touched);
  return out;
}

/* fill_face_boundary_mesh_topology!, topology.jl:510-588: cell -> d-face ids
 * by searching the coboundary of the cell's vertices. */
static int32_t *face_boundary_by_search(const int32_t *nface_v, int64_t nn, int wn, const int32_t *dface_v,
                                        int64_t nd, int wd, int64_t nv, int nlf, const int lf[][8]) {
  jagged v2d = coboundary(dface_v, nd, wd, nv);
  int32_t *out = xcalloc((size_t)nn * nlf, sizeof(int32_t));
  for (int64_t F = 0; F < nn; F++)
    for (int l = 0; l < nlf; l++) {
      int32_t lv[8];
      for (int i = 0; i < wd; i++) lv[i] = nface_v[F * wn + lf[l][i] - 1];
      int32_t d2 = INVALID;
      for (int i = 0; i < wd && d2 == INVALID; i++) {
        int32_t v = lv[i];
        for (int64_t k = v2d.ptr[v - 1]; k < v2d.ptr[v]; k++) {
          int32_t d1 = v2d.data[k];
          if (same_set(lv, wd, dface_v + (int64_t)(d1 - 1) * wd, wd)) { d2 = d1; break; }
        }
      }
      if (d2 == INVALID) { jfree(&v2d); free(out); set_err("mesh is not a face complex"); return NULL; }
      out[F * nlf + l] = d2;
    }
  jfree(&v2d);
  return out;
}

/* ------------------------------------------------------------------------ */
/* complexify (topology.jl:867-1125) on a mesh with cells + parent faces     */
/* ------------------------------------------------------------------------ */
static int complexify(orc_problem *p) {
  int D = p->D, ck = p->cellkind;
  const refgeo *g = &G[ck];
  /* create_vertices!, topology.jl:1034-1097: every node of every face of
   * this order-1 mesh is a vertex; parent 0-faces first, then node order. */
  p->node_vertex = xcalloc(p->nnodes, sizeof(int32_t));
  for (int64_t c = 0; c < p->ncells; c++)
    for (int i = 0; i < p->nvc; i++) p->node_vertex[p->cell_nodes[c * p->nvc + i] - 1] = -1;
  for (int d = 0; d < D; d++)
    for (int64_t f = 0; f < p->npf[d]; f++)
      for (int i = 0; i < p->npfv[d]; i++) p->node_vertex[p->pf_nodes[d][f * p->npfv[d] + i] - 1] = -1;
  int32_t v = 0;
  for (int64_t f = 0; f < p->npf[0]; f++) {
    int32_t node = p->pf_nodes[0][f];
    if (p->node_vertex[node - 1] == -1) p->node_vertex[node - 1] = ++v;
  }
  for (int64_t n = 0; n < p->nnodes; n++) if (p->node_vertex[n] == -1) p->node_vertex[n] = ++v;
  p->nvert = v;
  p->vertex_node = xcalloc(v + 1, sizeof(int32_t));
  for (int64_t n = 0; n < p->nnodes; n++) if (p->node_vertex[n] > 0) p->vertex_node[p->node_vertex[n] - 1] = (int32_t)(n + 1);
  /* cell vertices (parent_face_vertices, topology.jl:1001-1008) */
  p->fv[D] = p->nvc;
  p->nf[D] = p->ncells;
  p->face_vertices[D] = xcalloc((size_t)p->ncells * p->nvc, sizeof(int32_t));
  for (int64_t i = 0; i < p->ncells * p->nvc; i++) p->face_vertices[D][i] = p->node_vertex[p->cell_nodes[i] - 1];
  p->nf[0] = p->nvert; p->fv[0] = 1;
  p->face_vertices[0] = xcalloc(p->nvert, sizeof(int32_t));
  for (int64_t i = 0; i < p->nvert; i++) p->face_vertices[0][i] = (int32_t)(i + 1);
  /* d = D-1 .. 1: create_face_boundary! from (d+1)-faces, then
   * create_face_vertices! for the new d-faces. */
  int32_t *nface_d[4] = {0, 0, 0, 0};
  for (int d = D - 1; d >= 1; d--) {
    int n = d + 1;
    int nkind = (n == D) ? ck : g->ftype[n];
    const refgeo *gn = &G[nkind];
    int wd = gn->fnv[d], wn = gn->nv, nlf = gn->nf[d];
    jagged v2n = coboundary(p->face_vertices[n], p->nf[n], wn, p->nvert);
    int32_t *pv = xcalloc((size_t)p->npf[d] * wd + 1, sizeof(int32_t));
    for (int64_t f = 0; f < p->npf[d]; f++)
      for (int i = 0; i < wd; i++) pv[f * wd + i] = p->node_vertex[p->pf_nodes[d][f * p->npfv[d] + i] - 1];
    jagged v2p = coboundary(pv, p->npf[d], wd, p->nvert);
    int64_t nd = 0;
    nface_d[n] = generate_face_boundary(p->face_vertices[n], p->nf[n], wn, &v2n, pv, p->npf[d], wd, &v2p,
                                        nlf, gn->f[d], &nd);
    jfree(&v2n); jfree(&v2p);
    if (!nface_d[n]) { free(pv); return -1; }
    p->nf[d] = nd; p->fv[d] = wd;
    p->face_vertices[d] = generate_face_vertices(nd, pv, p->npf[d], wd, p->face_vertices[n], p->nf[n], wn,
                                                 nface_d[n], nlf, gn->f[d]);
    free(pv);
  }
  /* cell -> d-face incidences (face_incidence, topology.jl:219-244) */
  for (int d = 0; d <= D; d++) {
    int nlf = g->nf[d];
    if (d == D) {
      p->cell_faces[d] = xcalloc(p->ncells, sizeof(int32_t));
      for (int64_t c = 0; c < p->ncells; c++) p->cell_faces[d][c] = (int32_t)(c + 1);
    } else if (d == 0) {
      p->cell_faces[0] = xcalloc((size_t)p->ncells * nlf, sizeof(int32_t));
      memcpy(p->cell_faces[0], p->face_vertices[D], sizeof(int32_t) * p->ncells * nlf);
    } else if (d == D - 1) {
      p->cell_faces[d] = nface_d[D];
      nface_d[D] = NULL;
    } else {
      p->cell_faces[d] = face_boundary_by_search(p->face_vertices[D], p->ncells, p->nvc, p->face_vertices[d],
                                                 p->nf[d], p->fv[d], p->nvert, nlf, g->f[d]);
      if (!p->cell_faces[d]) return -1;
    }
  }
  for (int d = 0; d < 4; d++) free(nface_d[d]);
  /* fill_face_permutation_ids!, topology.jl:590-666 */
  for (int d = 0; d <= D; d++) {
    int nlf = g->nf[d];
    p->cell_perm[d] = xcalloc((size_t)p->ncells * nlf, sizeof(int32_t));
    if (d == 0 || d == D) {
      for (int64_t i = 0; i < p->ncells * nlf; i++) p->cell_perm[d][i] = 1;
      continue;
    }
    int vp[24][8];
    int np = vertex_permutations(g->ftype[d], vp);
    int wd = g->fnv[d];
    for (int64_t c = 0; c < p->ncells; c++)
      for (int l = 0; l < nlf; l++) {
        int32_t face = p->cell_faces[d][c * nlf + l];
        const int32_t *fvx = p->face_vertices[d] + (int64_t)(face - 1) * wd;
        int found = 0;
        for (int pi = 0; pi < np && !found; pi++) {
          int ok = 1;
          for (int cf = 0; cf < wd; cf++) {
            int32_t v1 = fvx[vp[pi][cf] - 1];
            int32_t v2 = p->face_vertices[D][c * p->nvc + g->f[d][l][cf] - 1];
            if (v1 != v2) { ok = 0; break; }
          }
          if (ok) { p->cell_perm[d][c * nlf + l] = pi + 1; found = 1; }
        }
        if (!found) { set_err("valid pindex not found"); return -1; }
      }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Cartesian meshes                                                          */
/* ------------------------------------------------------------------------ */
/* cartesian_chain / structured_simplex_chain, cartesian_mesh.jl:213-328 */
static const int simplex_nodes2[2][3] = {{1, 2, 3}, {4, 3, 2}};
static const int simplex_nodes3[6][4] = {{1, 2, 3, 7}, {1, 2, 5, 7}, {2, 3, 4, 7}, {2, 4, 7, 8}, {2, 5, 6, 7}, {2, 6, 7, 8}};

/* Sub-faces of the simplexified unit cube on each cube d-face:
 * simplexify(::UnitNCube), domain.jl:270-318.  The simplex mesh of the cube is
 * complexified (no parent faces) and each cube d-face collects the simplex
 * d-faces whose nodes it contains, in increasing simplex-face order. */
typedef struct { int nsub[12]; int sub[12][4][3]; int w; } subfaces;

static int simplex_cube_subfaces(int D, int d, subfaces *out) {
  orc_problem q;
  memset(&q, 0, sizeof q);
  q.D = D; q.cellkind = D == 2 ? ORC_TRI : ORC_TET; q.simplex = 1;
  q.nnodes = 1 << D;
  q.nvc = D + 1;
  q.ncells = D == 2 ? 2 : 6;
  q.cell_nodes = xcalloc(q.ncells * q.nvc, sizeof(int32_t));
  for (int c = 0; c < q.ncells; c++)
    for (int i = 0; i < q.nvc; i++) q.cell_nodes[c * q.nvc + i] = D == 2 ? simplex_nodes2[c][i] : simplex_nodes3[c][i];
  if (complexify(&q)) { free(q.cell_nodes); return -1; }
  const refgeo *cube = &G[D == 2 ? ORC_QUAD : ORC_HEX];
  out->w = d + 1;
  for (int cf = 0; cf < cube->nf[d]; cf++) {
    out->nsub[cf] = 0;
    for (int64_t sf = 0; sf < q.nf[d]; sf++) {
      int all = 1;
      for (int i = 0; i < d + 1; i++) {
        int32_t node = q.vertex_node[q.face_vertices[d][sf * (d + 1) + i] - 1];
        int in = 0;
        for (int k = 0; k < cube->fnv[d]; k++) if (cube->f[d][cf][k] == node) in = 1;
        if (!in) { all = 0; break; }
      }
      if (all) {
        for (int i = 0; i < d + 1; i++)
          out->sub[cf][out->nsub[cf]][i] = q.vertex_node[q.face_vertices[d][sf * (d + 1) + i] - 1];
        out->nsub[cf]++;
      }
    }
  }
  free(q.cell_nodes); free(q.node_vertex); free(q.vertex_node);
  for (int k = 0; k < 4; k++) { free(q.face_vertices[k]); free(q.cell_faces[k]); free(q.cell_perm[k]); }
  return 0;
}

/* cartesian_mesh_with_boundary (cartesian_mesh.jl:97-211) and
 * structured_simplex_mesh_with_boundary (:330-461), then complexify. */
orc_problem *orc_cartesian(int D, const int32_t *cells, const double *box, int simplexify) {
  init_geo();
  if (D < 2 || D > 3) { set_err("D must be 2 or 3"); return NULL; }
  int any_ne1 = 0, any_lt2 = 0;
  for (int d = 0; d < D; d++) { if (cells[d] != 1) any_ne1 = 1; if (cells[d] < 2) any_lt2 = 1; }
  if (any_ne1 && any_lt2) { set_err("At least 2 cells in any direction (or 1 cell in all directions)"); return NULL; }
  orc_problem *p = xcalloc(1, sizeof(orc_problem));
  p->D = D;
  p->simplex = simplexify;
  int npd[3], ncube = 1;
  p->nnodes = 1;
  for (int d = 0; d < D; d++) { npd[d] = cells[d] + 1; p->nnodes *= npd[d]; ncube *= cells[d]; }
  /* nodes: lexicographic, x fastest, x = pmin + h*(i-1) */
  p->x = xcalloc((size_t)p->nnodes * D, sizeof(double));
  double h[3];
  for (int d = 0; d < D; d++) h[d] = (box[2 * d + 1] - box[2 * d]) / cells[d];
  for (int64_t n = 0; n < p->nnodes; n++) {
    int64_t r = n;
    for (int d = 0; d < D; d++) { int i = (int)(r % npd[d]); r /= npd[d]; p->x[n * D + d] = box[2 * d] + h[d] * i; }
  }
  /* hex/quad cells of the cartesian chain */
  int nl = 1 << D;
  int32_t *hexn = xcalloc((size_t)ncube * nl, sizeof(int32_t));
  for (int64_t c = 0; c < ncube; c++) {
    int ci[3] = {0, 0, 0};
    int64_t r = c;
    for (int d = 0; d < D; d++) { ci[d] = (int)(r % cells[d]); r /= cells[d]; }
    for (int l = 0; l < nl; l++) {
      int64_t node = 0, stride = 1;
      for (int d = 0; d < D; d++) { node += (ci[d] + ((l >> d) & 1)) * stride; stride *= npd[d]; }
      hexn[c * nl + l] = (int32_t)(node + 1);
    }
  }
  /* boundary barrier: node_to_n counts hex cells per node */
  int32_t *node_n = xcalloc(p->nnodes, sizeof(int32_t));
  for (int64_t i = 0; i < (int64_t)ncube * nl; i++) node_n[hexn[i] - 1]++;
  const refgeo *cube = &G[D == 2 ? ORC_QUAD : ORC_HEX];
  for (int d = 0; d < D; d++) {
    int nmax = 1 << d;
    subfaces sf;
    if (simplexify && simplex_cube_subfaces(D, d, &sf)) { free(hexn); free(node_n); orc_free(p); return NULL; }
    int w = simplexify ? d + 1 : cube->fnv[d];
    int64_t cnt = 0;
    for (int pass = 0; pass < 2; pass++) {
      int64_t k = 0;
      for (int64_t c = 0; c < ncube; c++)
        for (int l = 0; l < cube->nf[d]; l++) {
          int isb = 1;
          for (int i = 0; i < cube->fnv[d]; i++)
            if (node_n[hexn[c * nl + cube->f[d][l][i] - 1] - 1] > nmax) { isb = 0; break; }
          if (!isb) continue;
          int nsub = simplexify ? sf.nsub[l] : 1;
          for (int s = 0; s < nsub; s++) {
            if (pass == 1) {
              for (int i = 0; i < w; i++) {
                int ln = simplexify ? sf.sub[l][s][i] : cube->f[d][l][i];
                p->pf_nodes[d][k * w + i] = hexn[c * nl + ln - 1];
              }
              p->pf_group[d][k] = l + 1;
            }
            k++;
          }
        }
      if (pass == 0) {
        cnt = k;
        p->npf[d] = cnt; p->npfv[d] = w;
        p->pf_nodes[d] = xcalloc((size_t)cnt * w + 1, sizeof(int32_t));
        p->pf_group[d] = xcalloc(cnt + 1, sizeof(int32_t));
      }
    }
  }
  free(node_n);
  if (!simplexify) {
    p->cellkind = D == 2 ? ORC_QUAD : ORC_HEX;
    p->ncells = ncube; p->nvc = nl; p->cell_nodes = hexn;
  } else {
    p->cellkind = D == 2 ? ORC_TRI : ORC_TET;
    int ns = D == 2 ? 2 : 6;
    p->nvc = D + 1; p->ncells = (int64_t)ncube * ns;
    p->cell_nodes = xcalloc((size_t)p->ncells * p->nvc, sizeof(int32_t));
    for (int64_t c = 0; c < ncube; c++)
      for (int s = 0; s < ns; s++)
        for (int i = 0; i < p->nvc; i++) {
          int ln = D == 2 ? simplex_nodes2[s][i] : simplex_nodes3[s][i];
          p->cell_nodes[(c * ns + s) * p->nvc + i] = hexn[c * nl + ln - 1];
        }
    free(hexn);
  }
  if (complexify(p)) { orc_free(p); return NULL; }
  return p;
}

/* ------------------------------------------------------------------------ */
/* DOF numbering: generate_dof_ids steps 0-2 (space.jl:299-535)              */
/* ------------------------------------------------------------------------ */
int orc_space(orc_problem *p, int order, int ncomp, int diri_kind, int group) {
  if (order < 1 || order > 4 || ncomp < 1 || ncomp > 3) { set_err("unsupported order/ncomp"); return -1; }
  if (p->has_space) { lagrange_free(&p->L); free(p->cell_dofs); p->cell_dofs = NULL; }
  free_csc(p);
  if (lagrange_init(&p->L, p->cellkind, order, ncomp)) return -1;
  p->has_space = 1; p->order = order; p->ncomp = ncomp;
  int D = p->D;
  const lagrange *L = &p->L;
  const refgeo *g = &G[p->cellkind];
  int nd = L->ndofs;
  /* step 1: offsets per d-face in id order, d = 0..D */
  int64_t off = 0;
  int64_t *doff[4];
  for (int d = 0; d <= D; d++) {
    int nlf = g->nf[d];
    doff[d] = xcalloc(p->nf[d] + 1, sizeof(int64_t));
    for (int64_t c = 0; c < p->ncells; c++)
      for (int l = 0; l < nlf; l++) doff[d][p->cell_faces[d][c * nlf + l] - 1] = (int64_t)L->own_cnt[d] * ncomp;
    for (int64_t f = 0; f < p->nf[d]; f++) { int64_t k = doff[d][f]; doff[d][f] = off; off += k; }
  }
  p->ndofs = off;
  p->cell_dofs = xcalloc((size_t)p->ncells * nd, sizeof(int32_t));
  for (int d = 0; d <= D; d++) {
    int nlf = g->nf[d], no = L->own_cnt[d];
    for (int64_t c = 0; c < p->ncells; c++) {
      int32_t *ldof = p->cell_dofs + c * nd;
      for (int l = 0; l < nlf; l++) {
        int64_t face = p->cell_faces[d][c * nlf + l];
        int pidx = p->cell_perm[d][c * nlf + l];
        const int *perm = L->perm[d] + (pidx - 1) * no;
        int64_t o = doff[d][face - 1];
        for (int i = 0; i < no; i++) {
          int node = L->own_nodes[d][l * no + i];
          int pn = perm[i];
          /* component-major dofs per node (space.jl:1257-1286, :1526-1575) */
          for (int k = 0; k < ncomp; k++) {
            int own_dof = (node - 1) * ncomp + k;
            int j = (pn - 1) * ncomp + k + 1;
            ldof[own_dof] = (int32_t)(j + o);
          }
        }
      }
    }
  }
  for (int d = 0; d <= D; d++) free(doff[d]);
  /* step 2: Dirichlet tagging on (D-1)-faces and free/Dirichlet split */
  p->nfree = p->ndofs; p->ndiri = 0;
  if (diri_kind) {
    int N = D - 1, nlf = g->nf[N];
    char *ftag = xcalloc(p->nf[N] + 1, 1);
    if (diri_kind == 1) {
      /* boundary(mesh): (D-1)-faces with exactly one cell (mesh.jl:131-143) */
      int32_t *cnt = xcalloc(p->nf[N] + 1, sizeof(int32_t));
      for (int64_t i = 0; i < p->ncells * nlf; i++) cnt[p->cell_faces[N][i] - 1]++;
      for (int64_t f = 0; f < p->nf[N]; f++) ftag[f] = cnt[f] == 1;
      free(cnt);
    } else {
      /* group "(D-1)-face-<group>": parent faces keep ids 1..npf */
      for (int64_t f = 0; f < p->npf[N]; f++) if (p->pf_group[N][f] == group) ftag[f] = 1;
    }
    char *dtag = xcalloc(p->ndofs + 1, 1);
    int fn = L->fn_cnt[N];
    for (int64_t c = 0; c < p->ncells; c++)
      for (int l = 0; l < nlf; l++) {
        if (!ftag[p->cell_faces[N][c * nlf + l] - 1]) continue;
        for (int i = 0; i < fn; i++) {
          int node = L->face_nodes[N][l * fn + i];
          for (int k = 0; k < ncomp; k++) dtag[p->cell_dofs[c * nd + (node - 1) * ncomp + k] - 1] = 1;
        }
      }
    int32_t *perm = xcalloc(p->ndofs + 1, sizeof(int32_t));
    int64_t nf = 0, ndr = 0;
    for (int64_t i = 0; i < p->ndofs; i++) if (!dtag[i]) perm[i] = (int32_t)(++nf);
    for (int64_t i = 0; i < p->ndofs; i++) if (dtag[i]) perm[i] = (int32_t)(-(++ndr));
    for (int64_t i = 0; i < p->ncells * nd; i++) p->cell_dofs[i] = perm[p->cell_dofs[i] - 1];
    p->nfree = nf; p->ndiri = ndr;
    free(perm); free(dtag); free(ftag);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Quadrature: measure(domain, degree) -> quadrature(geo, degree)            */
/* (problems.jl:4-25, quadrature.jl:36-157, :500-634)                        */
/* ------------------------------------------------------------------------ */
static int strang3(int degree, double x[][3], double *w) {
  if (degree == 4) {
    double a = 0.3994035761667992, b = 0.1005964238332008, c = (343.0 / 7500.0) / 6.0, d = (56.0 / 375.0) / 6.0;
    double e = 1.0 / 4.0, f = 11.0 / 14.0, gg = 1.0 / 14.0, h = (-148.0 / 1875.0) / 6.0;
    double X[11][3] = {{e, e, e}, {f, gg, gg}, {gg, f, gg}, {gg, gg, f}, {gg, gg, gg}, {a, a, b},
                       {a, b, a}, {a, b, b}, {b, a, a}, {b, a, b}, {b, b, a}};
    double W[11] = {h, c, c, c, c, d, d, d, d, d, d};
    for (int i = 0; i < 11; i++) { for (int k = 0; k < 3; k++) x[i][k] = X[i][k]; w[i] = W[i]; }
    return 11;
  }
  if (degree == 5) {
    double a = 0.0673422422100983, b = 0.3108859192633005, c = 0.7217942490673264, d = 0.0927352503108912;
    double e = 0.4544962958743506, f = 0.0455037041256494;
    double P = 0.1126879257180162 / 6.0, Q = 0.0734930431163619 / 6.0, R = 0.0425460207770812 / 6.0;
    double X[14][3] = {{a, b, b}, {b, a, b}, {b, b, a}, {b, b, b}, {c, d, d}, {d, c, d}, {d, d, c},
                       {d, d, d}, {e, e, f}, {e, f, e}, {e, f, f}, {f, e, e}, {f, e, f}, {f, f, e}};
    double W[14] = {P, P, P, P, Q, Q, Q, Q, R, R, R, R, R, R};
    for (int i = 0; i < 14; i++) { for (int k = 0; k < 3; k++) x[i][k] = X[i][k]; w[i] = W[i]; }
    return 14;
  }
  return 0;
}

int orc_set_quadrature(orc_problem *p, int degree) {
  int D = p->D;
  if (!p->simplex) {
    int n = (degree + 2) / 2; /* ceil((degree+1)/2) */
    double x1[64], w1[64];
    if (orc_gauss_legendre(n, x1, w1)) return -1;
    int nq = 1;
    for (int d = 0; d < D; d++) nq *= n;
    if (nq > 128) { set_err("too many quadrature points"); return -1; }
    for (int q = 0; q < nq; q++) {
      int r = q;
      double w = 1;
      /* tensor_product!(prod, ...): per-direction weights multiplied in order */
      for (int d = 0; d < D; d++) { int i = r % n; r /= n; p->qx[q][d] = x1[i]; w = (d == 0) ? w1[i] : w * w1[i]; }
      p->qw[q] = w;
    }
    p->nq = nq;
    return 0;
  }
  if (D == 3 && degree >= 4 && degree <= 5) {
    p->nq = strang3(degree, p->qx, p->qw);
    return 0;
  }
  /* duffy_quadrature, quadrature.jl:124-157 */
  int n = (int)ceil((degree + 1.0) / 2.0);
  double xs[3][64], ws[3][64];
  for (int d = 0; d < D - 1; d++) gauss_jacobi01(n, (D - 1) - d, 0, xs[d], ws[d]);
  orc_gauss_legendre(n, xs[D - 1], ws[D - 1]);
  double a = 0.5;
  for (int d = D - 2; d >= 0; d--) { for (int i = 0; i < n; i++) ws[d][i] *= a; a *= 0.5; }
  int nq = 1;
  for (int d = 0; d < D; d++) nq *= n;
  if (nq > 128) { set_err("too many quadrature points"); return -1; }
  for (int q = 0; q < nq; q++) {
    int r = q;
    double w = 1, y[3];
    for (int d = 0; d < D; d++) { int i = r % n; r /= n; y[d] = xs[d][i]; w = (d == 0) ? ws[d][i] : w * ws[d][i]; }
    /* duffy_map */
    double m = 1.0;
    for (int d = 0; d < D; d++) {
      if (d == 0) p->qx[q][0] = y[0];
      else { m *= (1 - y[d - 1]); p->qx[q][d] = m * y[d]; }
    }
    p->qw[q] = w;
  }
  p->nq = nq;
  return 0;
}

int orc_get_quadrature(const orc_problem *p, double *xi, double *w) {
  for (int q = 0; q < p->nq; q++) {
    for (int d = 0; d < p->D; d++) xi[q * p->D + d] = p->qx[q][d];
    w[q] = p->qw[q];
  }
  return p->nq;
}

/* geometry = 1: order-1 geometry element; 0: current space element */
int orc_tabulate(const orc_problem *p, int geometry, const double *xi, int npts, double *val, double *grad) {
  lagrange Lg;
  const lagrange *L = &p->L;
  if (geometry) { if (lagrange_init(&Lg, p->cellkind, 1, 1)) return -1; L = &Lg; }
  else if (!p->has_space) { set_err("no space"); return -1; }
  for (int i = 0; i < npts; i++)
    lagrange_tabulate(L, xi + i * p->D, val ? val + (int64_t)i * L->nnodes : NULL,
                      grad ? grad + (int64_t)i * L->nnodes * p->D : NULL);
  if (geometry) lagrange_free(&Lg);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Assembly: generated loop (compiler.jl:1826-1997), accessors               */
/* (accessors.jl:933-1007, :1312-1373), contribute! (assembly.jl:175-208),   */
/* compress (assembly.jl:558-575).                                           */
/* ------------------------------------------------------------------------ */
typedef struct {
  int nq, nv, ns, D;
  double *gN;   /* nq * nv * D geometry gradients */
  double *phi;  /* nq * ns values */
  double *gphi; /* nq * ns * D reference gradients */
} tables;

static int make_tables(const orc_problem *p, tables *t) {
  lagrange Lg;
  if (lagrange_init(&Lg, p->cellkind, 1, 1)) return -1;
  t->nq = p->nq; t->nv = Lg.nnodes; t->ns = p->L.nnodes; t->D = p->D;
  t->gN = xcalloc((size_t)t->nq * t->nv * t->D, sizeof(double));
  t->phi = xcalloc((size_t)t->nq * t->ns, sizeof(double));
  t->gphi = xcalloc((size_t)t->nq * t->ns * t->D, sizeof(double));
  for (int q = 0; q < p->nq; q++) {
    lagrange_tabulate(&Lg, p->qx[q], NULL, t->gN + (int64_t)q * t->nv * t->D);
    lagrange_tabulate(&p->L, p->qx[q], t->phi + (int64_t)q * t->ns, t->gphi + (int64_t)q * t->ns * t->D);
  }
  lagrange_free(&Lg);
  return 0;
}

static void free_tables(tables *t) { free(t->gN); free(t->phi); free(t->gphi); }

/* J = sum_n x_n (outer) grad N_n (sum_jacobian, accessors.jl:941-952);
 * dV = sqrt(det(J'J)) w (change_of_measure, quadrature.jl:4-6);
 * Jinv for transpose(J)\g (map_shape_function, accessors.jl:1365-1368). */
static double geometry(const orc_problem *p, const tables *t, int64_t c, int q, double J[3][3], double Ji[3][3]) {
  int D = p->D;
  for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) J[a][b] = 0;
  for (int n = 0; n < t->nv; n++) {
    const double *xn = p->x + (int64_t)(p->cell_nodes[c * p->nvc + n] - 1) * D;
    const double *g = t->gN + ((int64_t)q * t->nv + n) * D;
    for (int a = 0; a < D; a++) for (int b = 0; b < D; b++) J[a][b] += xn[a] * g[b];
  }
  double JtJ[3][3] = {{0}};
  for (int a = 0; a < D; a++)
    for (int b = 0; b < D; b++) { double s = 0; for (int k = 0; k < D; k++) s += J[k][a] * J[k][b]; JtJ[a][b] = s; }
  double det, detJ;
  if (D == 2) {
    det = JtJ[0][0] * JtJ[1][1] - JtJ[0][1] * JtJ[1][0];
    detJ = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    Ji[0][0] = J[1][1] / detJ; Ji[0][1] = -J[0][1] / detJ;
    Ji[1][0] = -J[1][0] / detJ; Ji[1][1] = J[0][0] / detJ;
  } else {
    det = JtJ[0][0] * (JtJ[1][1] * JtJ[2][2] - JtJ[1][2] * JtJ[2][1]) -
          JtJ[0][1] * (JtJ[1][0] * JtJ[2][2] - JtJ[1][2] * JtJ[2][0]) +
          JtJ[0][2] * (JtJ[1][0] * JtJ[2][1] - JtJ[1][1] * JtJ[2][0]);
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    detJ = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    Ji[0][0] = c00 / detJ;
    Ji[1][0] = c01 / detJ;
    Ji[2][0] = c02 / detJ;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / detJ;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / detJ;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / detJ;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / detJ;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / detJ;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / detJ;
  }
  return sqrt(det) * p->qw[q];
}

typedef struct { int32_t i, j; double v; } coo;

static int in_set(int32_t dof, int which) { return which == ORC_FREE ? dof > 0 : dof < 0; }
static int32_t map_dof(int32_t dof, int which) { return which == ORC_FREE ? dof : -dof; }

int orc_assemble_matrix(orc_problem *p, int form, const double *coef, int rows, int cols) {
  if (!p->has_space || !p->nq) { set_err("space or quadrature missing"); return -1; }
  if (form == ORC_ELASTICITY && p->ncomp != p->D) { set_err("elasticity needs ncomp == D"); return -1; }
  free_csc(p);
  tables t;
  if (make_tables(p, &t)) return -1;
  int D = p->D, ns = t.ns, nc = p->ncomp, nd = p->L.ndofs;
  double t0 = now_ms();
  /* symbolic pass: count (allocate_matrix_barrier_impl!, assembly.jl:119-153);
   * cell c's COO entries start at cell_off[c], so the COO order is cell-major
   * whatever the thread count */
  int64_t *cell_off = xcalloc(p->ncells + 1, sizeof(int64_t));
  for (int64_t c = 0; c < p->ncells; c++) {
    const int32_t *dofs = p->cell_dofs + c * nd;
    int64_t ni = 0, nj = 0;
    for (int a = 0; a < nd; a++) { ni += in_set(dofs[a], rows); nj += in_set(dofs[a], cols); }
    cell_off[c + 1] = cell_off[c] + ni * nj;
  }
  const int64_t ncoo = cell_off[p->ncells];
  coo *C = xcalloc(ncoo + 1, sizeof(coo));
  double t1 = now_ms();
#pragma omp parallel num_threads(g_threads)
  {
  double *be = xcalloc((size_t)nd * nd, sizeof(double));
  double *gph = xcalloc((size_t)ns * 3, sizeof(double));
#pragma omp for schedule(static)
  for (int64_t c = 0; c < p->ncells; c++) {
    const int32_t *dofs = p->cell_dofs + c * nd;
    memset(be, 0, sizeof(double) * nd * nd);
    double kap = 1.0, lam = 0, mu = 0;
    if (form == ORC_ELASTICITY) { lam = coef[2 * c]; mu = coef[2 * c + 1]; }
    else if (coef) kap = coef[c];
    for (int q = 0; q < t.nq; q++) {
      double J[3][3], Ji[3][3];
      double dV = geometry(p, &t, c, q, J, Ji);
      /* physical gradients: Ji^T applied, i.e. g_phys[a] = sum_b Ji[b][a] g_ref[b] */
      for (int s = 0; s < ns; s++) {
        const double *gr = t.gphi + ((int64_t)q * ns + s) * D;
        for (int a = 0; a < D; a++) { double v = 0; for (int b = 0; b < D; b++) v += Ji[b][a] * gr[b]; gph[s * 3 + a] = v; }
      }
      const double *phi = t.phi + (int64_t)q * ns;
      for (int j = 0; j < nd; j++) {
        for (int i = 0; i < nd; i++) {
          double v;
          if (form == ORC_LAPLACE) {
            double s = 0;
            for (int a = 0; a < D; a++) s += gph[j * 3 + a] * gph[i * 3 + a];
            v = (kap * s) * dV;
          } else if (form == ORC_MASS) {
            v = (kap * (phi[j] * phi[i])) * dV;
          } else {
            /* sigma(eps(grad u)) : eps(grad v) with u = phi_b e_cj, v = phi_a e_ci */
            int bn = j / nc, cj = j % nc, an = i / nc, ci = i % nc;
            double gu[3][3] = {{0}}, gv[3][3] = {{0}}, eu[3][3], ev[3][3];
            for (int r = 0; r < D; r++) { gu[cj][r] = gph[bn * 3 + r]; gv[ci][r] = gph[an * 3 + r]; }
            double tr = 0;
            for (int r = 0; r < D; r++)
              for (int s2 = 0; s2 < D; s2++) { eu[r][s2] = 0.5 * (gu[r][s2] + gu[s2][r]); ev[r][s2] = 0.5 * (gv[r][s2] + gv[s2][r]); }
            for (int r = 0; r < D; r++) tr += eu[r][r];
            double s = 0;
            for (int r = 0; r < D; r++)
              for (int s2 = 0; s2 < D; s2++) {
                double sig = 2 * mu * eu[r][s2] + (r == s2 ? lam * tr : 0.0);
                s += sig * ev[r][s2];
              }
            v = s * dV;
          }
          be[i + (int64_t)j * nd] += v;
        }
      }
    }
    /* contribute!: trial j outer, test i inner (assembly.jl:189-208) */
    int64_t k = cell_off[c];
    for (int j = 0; j < nd; j++) {
      if (!in_set(dofs[j], cols)) continue;
      for (int i = 0; i < nd; i++) {
        if (!in_set(dofs[i], rows)) continue;
        C[k].i = map_dof(dofs[i], rows); C[k].j = map_dof(dofs[j], cols); C[k].v = be[i + (int64_t)j * nd];
        k++;
      }
    }
  }
  free(be); free(gph);
  }
  free(cell_off); free_tables(&t);
  double t2 = now_ms();
  /* compress: stable column/row order, duplicates summed in COO order */
  int64_t ncol = cols == ORC_FREE ? p->nfree : p->ndiri;
  /* stable sort: counting sort by column, then rows sorted per column */
  int64_t *cptr = xcalloc(ncol + 2, sizeof(int64_t));
  for (int64_t e = 0; e < ncoo; e++) cptr[C[e].j]++;
  for (int64_t j = 0; j < ncol; j++) cptr[j + 1] += cptr[j];
  coo *S = xcalloc(ncoo + 1, sizeof(coo));
  {
    int64_t *fill = xcalloc(ncol + 1, sizeof(int64_t));
    for (int64_t j = 0; j <= ncol; j++) fill[j] = j ? cptr[j - 1] : 0;
    for (int64_t e = 0; e < ncoo; e++) S[fill[C[e].j]++] = C[e];
    free(fill);
  }
  free(C);
  /* insertion-sort rows inside each column keeping equal rows in COO order;
   * columns are independent */
  p->colptr = xcalloc(ncol + 1, sizeof(int32_t));
  /* cptr[j] holds the end of column j (1-based) */
  int64_t *cstart = xcalloc(ncol + 2, sizeof(int64_t));
  int64_t *ucount = xcalloc(ncol + 2, sizeof(int64_t));
  for (int64_t j = 1; j <= ncol; j++) cstart[j] = (j == 1) ? 0 : cptr[j - 1];
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1024)
  for (int64_t j = 1; j <= ncol; j++) {
    int64_t b = cstart[j], e = cptr[j];
    for (int64_t a = b + 1; a < e; a++) {
      coo x = S[a];
      int64_t m = a - 1;
      while (m >= b && S[m].i > x.i) { S[m + 1] = S[m]; m--; }
      S[m + 1] = x;
    }
    int64_t u = 0;
    for (int64_t a = b; a < e; a++) if (a == b || S[a].i != S[a - 1].i) u++;
    ucount[j] = u;
  }
  /* ucount[j] becomes the offset of column j's first nonzero */
  int64_t nnz = 0;
  for (int64_t j = 1; j <= ncol; j++) { int64_t u = ucount[j]; ucount[j] = nnz; nnz += u; }
  p->rowval = xcalloc(nnz + 1, sizeof(int32_t));
  p->nzval = xcalloc(nnz + 1, sizeof(double));
  p->colptr[0] = 1;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1024)
  for (int64_t j = 1; j <= ncol; j++) {
    int64_t b = cstart[j], e = cptr[j], w = ucount[j];
    for (int64_t a = b; a < e; a++) {
      if (a == b || S[a].i != S[a - 1].i) { p->rowval[w] = S[a].i; p->nzval[w] = S[a].v; w++; }
      else p->nzval[w - 1] += S[a].v;
    }
    p->colptr[j] = (int32_t)(w + 1);
  }
  free(S); free(cptr); free(cstart); free(ucount);
  p->csc_nnz = nnz;
  p->csc_ncols = ncol;
  double t3 = now_ms();
  g_phase_ms[0] = t1 - t0;
  g_phase_ms[1] = t2 - t1;
  g_phase_ms[2] = t3 - t2;
  return 0;
}

int64_t orc_csc_ncols(const orc_problem *p) { return p->csc_ncols; }

int orc_get_csc(const orc_problem *p, int32_t *colptr, int32_t *rowval, double *nzval) {
  if (!p->colptr) { set_err("no matrix assembled"); return -1; }
  if (colptr) memcpy(colptr, p->colptr, sizeof(int32_t) * (p->csc_ncols + 1));
  if (rowval) memcpy(rowval, p->rowval, sizeof(int32_t) * p->csc_nnz);
  if (nzval) memcpy(nzval, p->nzval, sizeof(double) * p->csc_nnz);
  return 0;
}

int orc_assemble_vector(orc_problem *p, int f_kind, const double *fdata, int rows, double *b) {
  if (!p->has_space || !p->nq) { set_err("space or quadrature missing"); return -1; }
  tables t;
  if (make_tables(p, &t)) return -1;
  int ns = t.ns, nc = p->ncomp, nd = p->L.ndofs;
  int64_t nrow = rows == ORC_FREE ? p->nfree : p->ndiri;
  memset(b, 0, sizeof(double) * nrow);
  double *be = xcalloc(nd, sizeof(double));
  for (int64_t c = 0; c < p->ncells; c++) {
    const int32_t *dofs = p->cell_dofs + c * nd;
    memset(be, 0, sizeof(double) * nd);
    for (int q = 0; q < t.nq; q++) {
      double J[3][3], Ji[3][3];
      double dV = geometry(p, &t, c, q, J, Ji);
      const double *phi = t.phi + (int64_t)q * ns;
      double f[3] = {0, 0, 0};
      for (int k = 0; k < nc; k++) {
        if (f_kind == 0) f[k] = fdata[k];
        else if (f_kind == 1) f[k] = fdata[c * nc + k];
        else {
          /* discrete field value: sum over local dofs of value * shape function (accessors.jl:1556-1563) */
          double s = 0;
          for (int a = 0; a < ns; a++) {
            int32_t dof = dofs[a * nc + k];
            double fv = dof > 0 ? fdata[dof - 1] : fdata[p->nfree + (-dof) - 1];
            s += fv * phi[a];
          }
          f[k] = s;
        }
      }
      for (int i = 0; i < nd; i++) {
        int an = i / nc, ci = i % nc;
        be[i] += (phi[an] * f[ci]) * dV;
      }
    }
    for (int i = 0; i < nd; i++) {
      if (!in_set(dofs[i], rows)) continue;
      b[map_dof(dofs[i], rows) - 1] += be[i];
    }
  }
  free(be); free_tables(&t);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Accessors                                                                 */
/* ------------------------------------------------------------------------ */
int64_t orc_query(const orc_problem *p, int what) {
  switch (what) {
  case ORC_Q_DIM: return p->D;
  case ORC_Q_NNODES: return p->nnodes;
  case ORC_Q_NCELLS: return p->ncells;
  case ORC_Q_NVPC: return p->nvc;
  case ORC_Q_NVERT: return p->nvert;
  case ORC_Q_NEDGES: return p->nf[1];
  case ORC_Q_NFACES2: return p->D == 3 ? p->nf[2] : p->ncells;
  case ORC_Q_NDOFS: return p->ndofs;
  case ORC_Q_NFREE: return p->nfree;
  case ORC_Q_NDIRI: return p->ndiri;
  case ORC_Q_DOFS_PER_CELL: return p->has_space ? p->L.ndofs : 0;
  case ORC_Q_NNZ: return p->csc_nnz;
  case ORC_Q_NQ: return p->nq;
  case ORC_Q_NPARENT0: return p->npf[0];
  case ORC_Q_NPARENT1: return p->npf[1];
  case ORC_Q_NPARENT2: return p->npf[2];
  case ORC_Q_CELLKIND: return p->cellkind;
  }
  return -1;
}

int orc_set_coords(orc_problem *p, const double *x) {
  memcpy(p->x, x, sizeof(double) * p->nnodes * p->D);
  return 0;
}
int orc_get_coords(const orc_problem *p, double *x) {
  memcpy(x, p->x, sizeof(double) * p->nnodes * p->D);
  return 0;
}
int orc_get_cell_nodes(const orc_problem *p, int32_t *out) {
  memcpy(out, p->cell_nodes, sizeof(int32_t) * p->ncells * p->nvc);
  return 0;
}
int orc_get_node_vertex(const orc_problem *p, int32_t *out) {
  memcpy(out, p->node_vertex, sizeof(int32_t) * p->nnodes);
  return 0;
}
int orc_get_cell_faces(const orc_problem *p, int d, int32_t *out) {
  if (d < 0 || d > p->D) return -1;
  memcpy(out, p->cell_faces[d], sizeof(int32_t) * p->ncells * G[p->cellkind].nf[d]);
  return 0;
}
int orc_get_face_vertices(const orc_problem *p, int d, int32_t *out) {
  if (d < 0 || d > p->D) return -1;
  memcpy(out, p->face_vertices[d], sizeof(int32_t) * p->nf[d] * p->fv[d]);
  return 0;
}
int orc_get_cell_perm(const orc_problem *p, int d, int32_t *out) {
  if (d < 0 || d > p->D) return -1;
  memcpy(out, p->cell_perm[d], sizeof(int32_t) * p->ncells * G[p->cellkind].nf[d]);
  return 0;
}
int orc_get_parent_faces(const orc_problem *p, int d, int32_t *nodes_out, int32_t *group_out) {
  if (d < 0 || d >= p->D) return -1;
  if (nodes_out) memcpy(nodes_out, p->pf_nodes[d], sizeof(int32_t) * p->npf[d] * p->npfv[d]);
  if (group_out) memcpy(group_out, p->pf_group[d], sizeof(int32_t) * p->npf[d]);
  return 0;
}
int orc_get_cell_dofs(const orc_problem *p, int32_t *out) {
  if (!p->has_space) return -1;
  memcpy(out, p->cell_dofs, sizeof(int32_t) * p->ncells * p->L.ndofs);
  return 0;
}
