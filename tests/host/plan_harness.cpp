// Host-only harness for the symbolic phase (galerkintoolkit.jl_amd/csrc/gtk_plan.cpp).
// It replays the scatter the assembly kernels perform, on the CPU, with the
// exactly representable local entries v(c, a, b) = 1 + c*7 + a*131 + b*977
// (doubles, integer valued), so sums are exact and comparable bit for bit
// with a COO reference assembled in the test.
#include <cstdint>
#include <cstring>
#include <string>

#include "gtk_plan.h"

extern "C" int harness_run(int dim, int64_t n_nodes, const double *xyz, int64_t n_cells, int nv,
                           const int32_t *cell_nodes0, int nd, int64_t n_free, int64_t n_diri,
                           const int32_t *cell_dofs, int rows, int cols, int max_cells, int chunk,
                           int64_t *nnz_out, int32_t *colptr, int32_t *rowval, double *nzval, int64_t cap,
                           int *ntiles, char *err, int errlen) {
  gtk::MeshView m;
  m.dim = dim; m.n_nodes = n_nodes; m.n_cells = n_cells; m.nodes_per_cell = nv;
  m.xyz = xyz; m.cell_nodes = cell_nodes0; m.cell_dofs = cell_dofs; m.dofs_per_cell = nd;
  m.n_free = n_free; m.n_diri = n_diri;
  gtk::Pattern p;
  gtk::TilePlan t;
  gtk::OffsetPlan o;
  std::string e = gtk::build_pattern(m, rows, cols, p, 4);
  if (e.empty()) e = gtk::build_tiles(m, cols, max_cells, 1 << 20, 1 << 20, chunk, t, 4);
  if (e.empty()) e = gtk::build_offsets(m, p, t, o, 4);
  if (!e.empty()) { std::strncpy(err, e.c_str(), errlen - 1); return -1; }
  *nnz_out = p.nnz;
  *ntiles = t.ntiles;
  if (p.nnz > cap) return 1;
  std::memcpy(colptr, p.colptr.data(), sizeof(int32_t) * p.colptr.size());
  std::memcpy(rowval, p.rowval.data(), sizeof(int32_t) * p.nnz);
  // batches: cover each tile's planned columns in order, at most
  // kBatchSlots contributors unless a single column has more
  if ((int)t.tile_batch_ptr.size() != t.ntiles + 1) { std::strncpy(err, "tile_batch_ptr size", errlen - 1); return -1; }
  for (int tl = 0; tl < t.ntiles; tl++) {
    int b0 = t.tile_batch_ptr[tl], b1 = t.tile_batch_ptr[tl + 1];
    if (b1 <= b0 || t.batch_col[b0] != t.tile_col_ptr[tl] || t.batch_col[b1] != t.tile_col_ptr[tl + 1]) {
      std::strncpy(err, "batches do not tile the columns", errlen - 1);
      return -1;
    }
    for (int b = b0; b < b1; b++) {
      int c0 = t.batch_col[b], c1 = t.batch_col[b + 1];
      int slots = t.col_ctr_ptr[c1] - t.col_ctr_ptr[c0];
      if (c1 <= c0 || c1 - c0 > 16 || (slots > gtk::kBatchSlots && c1 - c0 != 1)) {
        std::strncpy(err, "batch exceeds its slots", errlen - 1);
        return -1;
      }
    }
  }
  for (int64_t i = 0; i < p.nnz; i++) nzval[i] = -1.0;  // every entry must be written
  for (int tl = 0; tl < t.ntiles; tl++) {
    const int32_t *cells = t.tile_cells.data() + t.tile_cell_ptr[tl];
    int ncell = t.tile_cell_ptr[tl + 1] - t.tile_cell_ptr[tl];
    if (ncell > max_cells) { std::strncpy(err, "tile exceeds max_cells", errlen - 1); return -1; }
    for (int32_t pc = t.tile_col_ptr[tl]; pc < t.tile_col_ptr[tl + 1]; pc++) {
      int32_t j = t.col_id[pc];
      int len = p.colptr[j + 1] - p.colptr[j];
      std::vector<double> acc(len, 0.0);
      for (int32_t k = t.col_ctr_ptr[pc]; k < t.col_ctr_ptr[pc + 1]; k++) {
        int s = t.ctr_slot[k], b = t.ctr_b[k];
        if (s >= ncell) { std::strncpy(err, "slot out of range", errlen - 1); return -1; }
        int64_t c = cells[s];
        for (int a = 0; a < nd; a++) {
          const size_t pos = (size_t)k * o.stride + a;
          int off = o.bytes == 1 ? o.data[pos] : reinterpret_cast<const uint16_t *>(o.data.data())[pos];
          if (off == (o.bytes == 1 ? 0xFF : 0xFFFF)) continue;
          if (off >= len) { std::strncpy(err, "offset out of range", errlen - 1); return -1; }
          acc[off] += 1.0 + c * 7.0 + a * 131.0 + b * 977.0;
        }
      }
      for (int r = 0; r < len; r++) nzval[p.colptr[j] - 1 + r] = acc[r];
    }
  }
  return 0;
}
