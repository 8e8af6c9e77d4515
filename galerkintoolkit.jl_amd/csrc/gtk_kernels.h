// gtk_kernels.h -- launch interface between the C ABI (gtk_abi.cpp) and the
// gfx950 kernels (gtk_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gtk {

constexpr int kBlock = 256;            // threads per workgroup (4 waves of 64)
// threads per workgroup of the VALU matrix kernel, one workgroup per CU:
// FP64 8 waves (its staged factors are twice as large, so more waves
// would leave too few cells per tile), the mixed form 12 (measured on
// P2 tets: FP64 2.59 ms at 8 waves vs 3.00 at 12; mixed 2.28 vs 1.83)
template <typename CT>
struct ValuBlock {
  static constexpr int value = sizeof(CT) == 4 ? 768 : 512;
};

// kNitsche: (beta v u - v n.grad u - n.grad v u) over boundary items;
// kNitscheLoad: (beta v u - n.grad v u), the load's matrix applied to nodal g;
// kSkeleton: the interior-penalty terms over interior faces, items of two
// cells (beta jump(v).jump(u) - jump(v).mean(grad u) - mean(grad v).jump(u))
enum { kLaplace = 0, kMass = 1, kElasticity = 2, kNitsche = 3, kNitscheLoad = 4, kSkeleton = 5 };
enum { kFConst = 0, kFCell = 1, kFNodal = 2 };

struct MatrixArgs {
  const double *xyz;
  const int32_t *cell_nodes;
  const double *coef;
  const double *w, *phi, *gphi, *gN;
  const double *rtab;  // affine simplices: reference integrals R[a][b][m][n] (NS*NS*D*D)
  const int32_t *tile_cell_ptr, *tile_cells, *tile_col_ptr, *col_id, *col_ctr_ptr;
  const int32_t *tile_batch_ptr;
  const uint16_t *ctr_slot, *ctr_b;
  const uint8_t *ctr_off;  // VALU kernel: off_stride entries per contributor, off_bytes each
  const unsigned char *rec;  // matrix-core kernels: batch records (BatchPlan)
  int rec_bytes, off_offset, endmask_words;
  const int32_t *colptr;
  double *nzval;
  int off_stride, off_bytes;
  int max_seg;
  int lds_geo;       // byte offset of the staged geometry factors in LDS
  int acc_offset;    // byte offset of the accumulators in LDS
  int acc_per_wave;  // FP64 accumulator slots per wave (matrix-core kernels)
  int stage_cells;   // cells whose coordinates fit the accumulator space (matrix-core kernels)
  int lds_claim;     // byte offset in LDS of the batch counter (matrix-core kernels)
  int ntiles;        // tiles of the plan (matrix-core kernels loop over them)
  const int32_t *item_face;  // face mode (boundary items), as in VectorArgs
  const double *face_jac;
  const double *face_normal;  // Nitsche: reference outward normal of each local face, [face][D]
};

// Many problems in one launch (matrix-core kernels of 2D quadrilaterals):
// per-problem arguments and element type, and the flattened work queue.


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
  // boundary items (face mode): local face of each item, d(reference cell
  // point)/d(reference face point) of each local face, [face][D][D-1]
  const int32_t *item_face;
  const double *face_jac;
};

// A batch of independent problems in one launch: work items (problem, t)
// are matrix tile t for t >= 0 and load-vector tile ~t for t < 0.
struct BatchedArgs {
  const MatrixArgs *problems;
  const VectorArgs *loads;  // per problem (read for vector items only)
  const int32_t *types;     // 0: Q1, 1: Q2, 2: Q3 quadrilaterals
  const int32_t *fkinds;    // load kind per problem (kFConst, kFCell, kFNodal)
  const int2 *items;  // in descending order of estimated cost within matrix and load-vector items
  int nitems;
  int *counter;      // items claimed so far in this launch (cleared by the launch's first graph node)
  int claim_offset;  // byte offset in LDS of the workgroup's two claim slots
};

// nl: local faces of the cell for boundary items (face mode), 0 for cells
struct ElemKey {
  int dim, nv, ns, nq, nc, nl;
};

struct KernelInfo {
  const void *fn = nullptr;
  size_t table_bytes = 0;         // LDS bytes of the reference tables
  size_t geo_bytes_per_cell = 0;  // LDS bytes of staged geometry per cell
  int cols_per_wave = 1;
  bool mfma = false;  // matrix-core kernel driven by batch records
  int rounds = 0;     // write-out rounds of 64 accumulator entries per batch (matrix-core kernels)
  int block = kBlock;
};

// allow_mfma: the plan's columns fit the matrix-core batches (at most
// kBatchSlots contributors per column); otherwise the VALU kernel is used.
// waves: wavefronts per workgroup of the matrix-core kernels (8, 12 or 16).
bool find_matrix_kernel(const ElemKey &key, int form, int precision, int off_bytes, bool allow_mfma, int waves,
                        KernelInfo &out);
bool find_vector_kernel(const ElemKey &key, int fkind, KernelInfo &out);
// Batched kernels: family 0 over scalar Q1/Q2/Q3 quadrilaterals with
// (k+1)^2 Gauss points (types 0, 1, 2), family 1 over scalar Q1/Q2
// hexahedra with (k+1)^3 (types 3, 4), family 2 over scalar P1/P2
// triangles with degree-2k rules (types 6, 7); batched_type gives a
// problem's type (-1: none), type / 3 its family and type % 3 its slot in
// the family's kernel; 8 or 12 waves per workgroup.  A hexahedral batch
// runs one launch per element type, on the kernels of "families"
// kHexQ1Launch and kHexQ2Launch (one tile body each, inlined).
constexpr int kBatchedTypes = 8;
constexpr int kHexQ1Launch = 3, kHexQ2Launch = 4;
int batched_type(const ElemKey &key);
bool find_batched_kernel(int form, int precision, int waves, int family, KernelInfo &out);

}  // namespace gtk
