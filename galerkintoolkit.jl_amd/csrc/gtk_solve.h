// gtk_solve.h -- launch interface of the solve kernels (gtk_solve.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gtk {

constexpr int kSolvePartials = 1024;  // blocks (and partial sums) of the vector kernels

// A matrix with a structurally symmetric pattern, in CSR form over the
// library's CSC pattern arrays (values gathered into row order).
struct SolveMat {
  int64_t n = 0;
  const int32_t *ptr = nullptr;  // colptr, 1-based
  const int32_t *idx = nullptr;  // rowval, 1-based
  int lanes = 16;                // lanes per row of the SpMV
  // BSR(3) copy for the FP32 iterations (null: none): block columns and the
  // FP32 values: entries e = 3i+j < 8 of a block at bval[8 * block + e]
  // (two 16-byte loads), entry 8 at bval[8 * nb + block]
  const int32_t *bcol = nullptr;
  const float *bval = nullptr;
  int64_t nb = 0;        // blocks
  int block_lanes = 32;  // lanes per block row of the BSR(3) SpMV (8, 16 or 32)
};

// Vectors of one conjugate-gradient solve (FP64 or FP32).
struct CgVectors {
  bool f64 = false;
  void *d = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  const void *dinv = nullptr;
};

// Scalars kept on the device in sc[]: [0] and [3] r.z (alternating by
// iteration), [5] r.r, [6] scratch dot.
constexpr int kSolveScalars = 8;

int solve_rows_per_lane_group(double mean_row);
void launch_spmv(const SolveMat &m, const void *val, bool val64, const void *x, bool x64, const void *b, void *y,
                 bool y64, hipStream_t st);
void launch_dot(int64_t n, const void *a, const void *b, bool f64, double *partial, double *sc, int slot,
                hipStream_t st);
void launch_cg_start(int64_t n, double s, const double *r64, const CgVectors &v, double *partial, double *sc,
                     hipStream_t st);
// One CG iteration (it counts from 0 after launch_cg_start; partial2 holds
// 2 x kSolvePartials); read_rr: also sum r.r into sc[5].
void launch_cg_iteration(const SolveMat &m, const void *val, const CgVectors &v, double *partial, double *partial2,
                         double *sc, int it, bool read_rr, hipStream_t st);
void launch_refine(int64_t n, double s, const void *d, bool f64, double *x, hipStream_t st);
void launch_diag_inv(const SolveMat &m, const double *val, void *dinv, bool f64, int *bad, hipStream_t st);
// perm[p] = CSC position of CSR entry p of the structurally symmetric
// pattern (m.ptr, m.idx); bad[0] = 1 when the pattern is not symmetric.
void launch_transpose_perm(const SolveMat &m, int32_t *perm, int *bad, hipStream_t st);
// bad[0] = 1 unless (m.ptr, m.idx) is made of 3x3 blocks (rows 3I..3I+2
// share one column list of whole aligned triples); m.n % 3 == 0.
void launch_block3_check(const SolveMat &m, int *bad, hipStream_t st);
// BSR(3) copy of a pattern that passed launch_block3_check (m.nb = nnz / 9
// blocks) from the row-ordered FP32 values; block columns when with_cols.
void launch_block3_build(const SolveMat &m, const float *val_rows, int32_t *bcol, float *bval, bool with_cols,
                         hipStream_t st);
void launch_gather_values(int64_t n, const int32_t *perm, const double *a, double *v64, float *v32, hipStream_t st);
// b_i += sum_k A_ik (alpha xd_k) for a CSC block A (values a) through its
// row view: row i's entries rptr[i] .. rptr[i+1]-1 (0-based), CSC position
// perm[k] and column col[k] (0-based), ascending columns, unfused products
// and sums (alpha = -1: the Dirichlet lifting b -= A xd)
void launch_lift(int64_t nrows, const int64_t *rptr, const int32_t *perm, const int32_t *col, const double *a,
                 double alpha, const double *xd, double *b, hipStream_t st);

}  // namespace gtk
