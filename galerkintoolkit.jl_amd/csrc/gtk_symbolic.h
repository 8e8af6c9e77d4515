// gtk_symbolic.h -- launch interface of the device symbolic phase
// (gtk_symbolic.hip): CSC pattern of a (rows, cols) block and the
// cell-local-entry -> nonzero map, built on the GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gtk {

// Longest row list of one column the device pattern builder sorts in LDS
// (contributing cells x rows per cell, before duplicates are removed).
constexpr int kSymbolicMaxList = 2048;

// Exclusive scan of n int32 counts into out[0..n] (out[0] = base,
// out[i+1] = base + sum of counts[0..i]) in int64; tmp holds
// symbolic_scan_tmp(n) int64 values.
size_t symbolic_scan_tmp(int64_t n);
void launch_scan(const int32_t *counts, int64_t n, int64_t base, int64_t *out, int64_t *tmp, hipStream_t st);

// counts[j] += number of (cell, b) whose DOF is column j of the set `cols`
// (1: free, 2: Dirichlet); counts must be zero on entry.
void launch_column_count(const int32_t *cell_dofs, int64_t n_cells, int nd, int cols, int32_t *counts,
                         hipStream_t st);
// Cells of each column (unordered within a column): inc[ptr[j] + k], fill
// must be zero on entry.
void launch_column_fill(const int32_t *cell_dofs, int64_t n_cells, int nd, int cols, const int64_t *ptr,
                        int32_t *fill, int32_t *inc, hipStream_t st);
// *out = max(*out, max of counts[0..n)).
void launch_max_count(const int32_t *counts, int64_t n, int32_t *out, hipStream_t st);
// Pass 0: len[j] = number of distinct rows (of the set `rows`) of column j.
// Pass 1: the sorted rows (1-based) at rowval[colptr[j] - 1 ...].  `cap`
// (a power of two, at most kSymbolicMaxList) bounds the candidate rows of a
// column (cells x nd); a longer column is skipped and its length reported
// in *maxlist.
void launch_column_rows(const int32_t *cell_dofs, int nd, int rows, int64_t ncols, const int64_t *ptr,
                        const int32_t *inc, int pass, int cap, int32_t *len, int32_t *maxlist,
                        const int64_t *colptr64, int32_t *rowval, hipStream_t st);
// colptr (Int32, 1-based) from the int64 scan.
void launch_colptr32(const int64_t *colptr64, int64_t n, int32_t *colptr, hipStream_t st);
// map[(cell*nd + b)*nd + a] = 1-based index of entry (row of a, column of b)
// in the CSC values, 0 when row or column is outside the block.
void launch_local_map(const int32_t *cell_dofs, int64_t n_cells, int nd, int rows, int cols, const int32_t *colptr,
                      const int32_t *rowval, int32_t *map, hipStream_t st);

// dst values += src values over the src pattern (each src column's rows are
// found in the same dst column by binary search); *missing = 1 if a src
// entry has no dst entry.
// dst[t] = the sum, in order, of src[idx[ptr[t]]] .. src[idx[ptr[t+1]-1]]
void launch_gather_block(int64_t n, const int64_t *ptr, const int64_t *idx, const double *src, double *dst,
                         hipStream_t st);
void launch_add_block(int64_t ncols, const int32_t *scolptr, const int32_t *srowval, const double *sval,
                      const int32_t *dcolptr, const int32_t *drowval, double *dval, int32_t *missing, hipStream_t st);

}  // namespace gtk
