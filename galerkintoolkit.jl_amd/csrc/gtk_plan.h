// gtk_plan.h -- host-side symbolic phase of libgtk_hip: CSC pattern of a
// (rows, cols) block and the column-owner work plan the assembly kernels run.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gtk {

constexpr uint16_t kNoRow = 0xFFFF;  // contributor row not in the row set

// CSC pattern of one (rows, cols) block, 1-based like SparseMatrixCSC.
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
  int ntiles = 0;
  int max_tile_cells = 0, max_tile_cols = 0, max_ctr = 0;
  std::vector<int32_t> tile_cell_ptr;  // ntiles+1
  std::vector<int32_t> tile_cells;     // 0-based cell ids, ascending within a tile
  std::vector<int32_t> tile_col_ptr;   // ntiles+1, into col_*
  std::vector<int32_t> col_id;         // 0-based index in S of each planned column
  std::vector<int32_t> col_ctr_ptr;    // size ncolumns+1, into ctr_*
  std::vector<uint16_t> ctr_slot;      // tile-local cell slot of each contributor
  std::vector<uint16_t> ctr_b;         // local dof b of the contributor in its cell
};

// Row offsets of each contributor in its column segment (matrix blocks).
struct OffsetPlan {
  std::vector<uint16_t> off;  // per contributor, ndofs_cell entries, kNoRow if excluded
  int max_seg = 0;
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
std::string build_pattern(const MeshView &m, int rows, int cols, Pattern &out, int nthreads);
// Builds the tile plan over set S with at most max_cells cells per tile and
// at most max_cols owned DOFs per tile.
std::string build_tiles(const MeshView &m, int set, int max_cells, int max_cols, int chunk, TilePlan &out,
                        int nthreads);
// Row offsets of every contributor of a tile plan over the columns of `p`.
std::string build_offsets(const MeshView &m, const Pattern &p, const TilePlan &t, OffsetPlan &out, int nthreads);

}  // namespace gtk
