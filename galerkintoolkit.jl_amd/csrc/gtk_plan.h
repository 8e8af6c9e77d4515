// gtk_plan.h -- host-side symbolic phase of libgtk_hip: CSC pattern of a
// (rows, cols) block and the column-owner work plan the assembly kernels run.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gtk {

constexpr uint16_t kNoRow = 0xFFFF;  // contributor row not in the row set
constexpr int kColGroup = 16;        // columns batched together by the MFMA kernels
constexpr int kBatchSlots = 16;      // contributors per matrix-core batch (16 = MFMA N)
constexpr int kBatchCols = 15;       // columns per batch (the 16th column entry is the padding column)

// CSC pattern of one (rows, cols) block, 1-based like SparseMatrixCSC.
// First slot of a batch column of cnt contributors placed after `used`
// slots: a column of more than 4 contributors starts at a multiple of 4, a
// shorter one does not straddle one.  The matrix-core kernel adds slot 4g+r
// of rank t in step (t / 4, r); with this placement the steps of one column
// come in rank (cell) order, and two slots of one column in one step would
// need ranks 4 apart in one window of 4, which cannot happen.
inline int batch_slot_start(int used, int cnt) {
  return (cnt > 4 || used % 4 + cnt > 4) ? (used + 3) & ~3 : used;
}

struct Pattern {
  int rows = 0, cols = 0;
  int64_t nrows = 0, ncols = 0, nnz = 0;
  std::vector<int32_t> colptr;  // ncols+1, 1-based
  std::vector<int32_t> rowval;  // nnz, 1-based
};

// Column-owner work plan over the DOFs of one set S (the columns of a
// matrix block, or the rows of a load vector).  A tile is a group of owned
// DOFs plus every cell that touches one of them; each DOF of S belongs to
// exactly one tile, so each output entry is produced by one workgroup in a
// fixed order (no atomics, deterministic).
struct TilePlan {
  int set = 0;                  // GTK_FREE / GTK_DIRICHLET
  int ndofs_cell = 0;           // local dofs per cell
  int64_t nset = 0;             // number of dofs in S
  int64_t uncovered = 0;        // dofs of S no cell touches (in no tile)
  int ntiles = 0;
  int max_tile_cells = 0, max_tile_cols = 0, max_ctr = 0;
  int max_tile_ctr = 0, max_tile_batches = 0;
  std::vector<int32_t> tile_cell_ptr;  // ntiles+1
  std::vector<int32_t> tile_cells;     // 0-based cell ids, ascending within a tile
  std::vector<int32_t> tile_col_ptr;   // ntiles+1, into col_*
  std::vector<int32_t> col_id;         // 0-based index in S of each planned column
  std::vector<int32_t> col_ctr_ptr;    // size ncolumns+1, into ctr_*
  std::vector<uint16_t> ctr_slot;      // tile-local cell slot of each contributor
  std::vector<uint16_t> ctr_b;         // local dof b of the contributor in its cell
  // batches: runs of consecutive planned columns of one tile whose
  // contributors total at most kBatchSlots (one column per batch when it has
  // more); batch i covers planned columns [batch_col[i], batch_col[i+1])
  std::vector<int32_t> tile_batch_ptr;  // ntiles+1
  std::vector<int32_t> batch_col;       // nbatches+1
  int max_batch_cols = 0;
};

// Row offsets of each contributor in its column segment (matrix blocks):
// `stride` entries per contributor (ndofs_cell rounded up to 16), one byte
// each when every segment is shorter than 255 rows, else two; the all-ones
// value marks a row outside the row set (or padding).
struct OffsetPlan {
  std::vector<uint8_t> data;
  int stride = 0;
  int bytes = 1;
  int max_seg = 0;
  int max_group_acc = 0;  // largest sum of segment lengths over kColGroup consecutive planned columns
  int max_batch_acc = 0;  // largest sum of segment lengths over the columns of one batch
};

// Matrix-core batch records (scalar spaces whose columns have at most
// kBatchSlots contributors).  One fixed-size record per batch, in tile
// order, holds everything a wavefront needs for the batch, so the kernel
// streams it with coalesced loads.  Slots are the batch's contributors,
// column by column, ranked within each column in ascending cell order.
//   u32  info        nslots | tmax << 8 | ncols << 16   (tmax: most contributors of one column)
//   u32  acc         sum of the segment lengths of the batch's columns
//   u32  pad[2]
//   u32  slot_sb[16] s | b << 16          (tile cell slot, local column b)
//   u32  slot_at[16] abase | t << 16      (offset of the slot's column in the batch
//                    accumulator, rank t; t = 0xFFFF for empty slots)
//   i32  delta[16]   seg0 - abase of column ci (seg0: 0-based start of the column's
//                    segment in the value array); after the ncols columns, a padding
//                    column fills the accumulator up to 64*rounds entries and maps
//                    them past the values (to the value array's tail, nnz onwards)
//   u64  endmask[rounds]  bit e set iff accumulator entry e ends a column
//   off  [dofs per cell][16]  row offset of local row a of each slot in its column
//                    segment (off_bytes each), all ones for rows outside the row set
//                    and for empty slots
struct BatchPlan {
  std::vector<uint8_t> rec;
  int64_t nbatches = 0;
  int rec_bytes = 0;
  int off_bytes = 1;
  int rounds = 0;         // write-out rounds of 64 entries: ceil(dofs per cell / 4)
  int endmask_words = 0;  // = rounds
  int off_offset = 0;  // byte offset of the row offsets in a record
  int max_batch_acc = 0;
  static constexpr int kSlotSb = 16, kSlotAt = 80, kDelta = 144, kEndmask = 208;
};

struct MeshView {
  int dim = 0;
  int64_t n_nodes = 0, n_cells = 0;
  int nodes_per_cell = 0;
  const double *xyz = nullptr;        // host, node-major
  const int32_t *cell_nodes = nullptr;  // host, 0-based
  const int32_t *cell_dofs = nullptr;   // host, GT signed encoding
  int dofs_per_cell = 0;
  int64_t n_free = 0, n_diri = 0;
};

// Builds the CSC pattern for (rows, cols); returns an error string on failure.
// nnz_limit: the most nonzeros a pattern may hold (SparseMatrixCSC{Float64,
// Int32} keeps nnz + 1 in its Int32 colptr); beyond it an error.
std::string build_pattern(const MeshView &m, int rows, int cols, Pattern &out, int nthreads,
                          int64_t nnz_limit = INT32_MAX - 1);
// Builds the tile plan over set S with at most max_cells cells, max_cols
// owned DOFs and max_ctr contributors per tile.
std::string build_tiles(const MeshView &m, int set, int max_cells, int max_cols, int max_ctr, int chunk,
                        TilePlan &out, int nthreads);
// Batch records of a tile plan over the columns of `p` (every column must
// have at most kBatchSlots contributors).
std::string build_batches(const MeshView &m, const Pattern &p, const TilePlan &t, BatchPlan &out, int nthreads);
// CSR view of a square, structurally symmetric pattern: perm[k] is the CSC
// position of the k-th entry in row-major order (row i's entries by
// ascending column), so the rows reuse colptr/rowval as pointers/indices.
// The solve builds the same permutation on the device
// (gtk_solve.hip, transpose_perm_kernel); this host form is the one the
// host tests pin against a sparse transpose.
std::string build_transpose_perm(const Pattern &p, std::vector<int32_t> &perm);
// Row offsets of every contributor of a tile plan over the columns of `p`.
std::string build_offsets(const MeshView &m, const Pattern &p, const TilePlan &t, OffsetPlan &out, int nthreads);

}  // namespace gtk
