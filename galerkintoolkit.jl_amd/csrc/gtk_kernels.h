// gtk_kernels.h -- launch interface between the C ABI (gtk_abi.cpp) and the
// gfx950 kernels (gtk_kernels.hip).
#pragma once
#include <cstddef>
#include <cstdint>

namespace gtk {

constexpr int kBlock = 256;            // threads per workgroup (4 waves of 64)
constexpr uint16_t kNoRowDev = 0xFFFF;  // contributor row outside the row set

enum { kLaplace = 0, kMass = 1, kElasticity = 2 };
enum { kFConst = 0, kFCell = 1, kFNodal = 2 };

struct MatrixArgs {
  const double *xyz;
  const int32_t *cell_nodes;
  const double *coef;
  const double *w, *phi, *gphi, *gN;
  const int32_t *tile_cell_ptr, *tile_cells, *tile_col_ptr, *col_id, *col_ctr_ptr;
  const uint16_t *ctr_slot, *ctr_b, *ctr_off;
  const int32_t *colptr;
  double *nzval;
  int max_seg;
  int acc_offset;  // byte offset of the accumulators in LDS
};

struct VectorArgs {
  const double *xyz;
  const int32_t *cell_nodes;
  const int32_t *cell_dofs;
  const double *f;
  int64_t n_free;
  const double *w, *phi, *gN;
  const int32_t *tile_cell_ptr, *tile_cells, *tile_col_ptr, *col_id, *col_ctr_ptr;
  const uint16_t *ctr_slot, *ctr_b;
  double *out;
};

struct ElemKey {
  int dim, nv, ns, nq, nc;
};

struct KernelInfo {
  const void *fn = nullptr;
  size_t table_bytes = 0;         // LDS bytes of the reference tables
  size_t geo_bytes_per_cell = 0;  // LDS bytes of staged geometry per cell
  int cols_per_wave = 1;
};

bool find_matrix_kernel(const ElemKey &key, int form, int precision, KernelInfo &out);
bool find_vector_kernel(const ElemKey &key, int fkind, KernelInfo &out);

}  // namespace gtk
