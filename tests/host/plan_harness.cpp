// Host-only harness for the symbolic phase (galerkintoolkit.jl_amd/csrc/gtk_plan.cpp).
// It replays the scatter the assembly kernels perform, on the CPU, with the
// exactly representable local entries v(c, a, b) = 1 + c*7 + a*131 + b*977
// (doubles, integer valued), so sums are exact and comparable bit for bit
// with a COO reference assembled in the test.
#include <cstdint>
#include <cstring>
#include <string>

#include "gtk_plan.h"

// Replays the batch records exactly as the matrix-core kernel consumes
// them: per batch, in steps (h, r), slot 4g+r adds its rows when its rank is
// in 4h..4h+3, into the batch accumulator through the record's row offsets
// (the replay checks that no two additions of one step touch the same
// entry); then each
// accumulator entry e is written to e + delta[c], c = number of column ends
// before e in the record's end mask.  Also checks that each column's
// contributors are ranked in ascending cell order and that every nonzero is
// written exactly once.
static int replay_batches(const gtk::MeshView &m, const gtk::Pattern &p, const gtk::TilePlan &t, double *nzval,
                          char *err, int errlen) {
  auto fail = [&](const char *msg) {
    std::strncpy(err, msg, errlen - 1);
    return -1;
  };
  if (t.max_ctr > gtk::kBatchSlots) return 2;  // matrix-core batches not applicable
  gtk::BatchPlan bp;
  std::string e = gtk::build_batches(m, p, t, bp, 4);
  if (!e.empty()) return fail(e.c_str());
  const int nd = m.dofs_per_cell;
  const uint32_t none = bp.off_bytes == 1 ? 0xFF : 0xFFFF;
  std::vector<int> written(p.nnz, 0);
  for (int tl = 0; tl < t.ntiles; tl++) {
    const int32_t *cells = t.tile_cells.data() + t.tile_cell_ptr[tl];
    const int ncell = t.tile_cell_ptr[tl + 1] - t.tile_cell_ptr[tl];
    for (int32_t b = t.tile_batch_ptr[tl]; b < t.tile_batch_ptr[tl + 1]; b++) {
      const uint8_t *R = bp.rec.data() + (size_t)b * bp.rec_bytes;
      const uint32_t info = reinterpret_cast<const uint32_t *>(R)[0];
      const int nslots = info & 0xFF, tmax = (info >> 8) & 0xFF, ncols = info >> 16;
      const int acc_total = (int)reinterpret_cast<const uint32_t *>(R)[1];
      const uint32_t *sb = reinterpret_cast<const uint32_t *>(R + gtk::BatchPlan::kSlotSb);
      const uint32_t *at = reinterpret_cast<const uint32_t *>(R + gtk::BatchPlan::kSlotAt);
      const int32_t *delta = reinterpret_cast<const int32_t *>(R + gtk::BatchPlan::kDelta);
      const uint64_t *endmask = reinterpret_cast<const uint64_t *>(R + gtk::BatchPlan::kEndmask);
      const uint8_t *off = R + bp.off_offset;
      if (nslots > gtk::kBatchSlots || ncols > 16 || ncols < (acc_total > 0 ? 1 : 0) ||
          acc_total > bp.max_batch_acc)
        return fail("batch header out of range");
      std::vector<double> acc(acc_total, 0.0);
      std::vector<int> stamp(acc_total, -1);
      std::vector<int> last_cell(16, -1);
      int seen = 0, used = 0;
      for (int n = 0; n < nslots; n++) used += (at[n] >> 16) != 0xFFFF;
      // the kernel's steps: (h, r) adds slot 4g+r whose rank is in 4h..4h+3
      for (int hh = 0; hh < (tmax + 3) / 4; hh++)
        for (int rr = 0; rr < 4; rr++)
        for (int gg = 0; gg < 4; gg++) {
          const int n = 4 * gg + rr;
          const int tn = at[n] >> 16, abase = at[n] & 0xFFFF;
          const int tt = hh * 4 + rr;  // stamp of the step
          if (tn == 0xFFFF || tn / 4 != hh) continue;
          if (n >= nslots) return fail("empty slot with a rank");
          const int s = sb[n] & 0xFFFF, lb = sb[n] >> 16;
          if (s >= ncell || lb >= nd) return fail("slot word out of range");
          // the slot's column: the one whose accumulator range starts at abase
          int ci = -1;
          for (int c = 0, a0 = 0, k = 0; c < ncols; c++) {
            if (a0 == abase) ci = c;
            int len = 0;
            while (k < acc_total && !((endmask[k / 64] >> (k % 64)) & 1)) k++, len++;
            a0 += len + 1;
            k++;
          }
          if (ci < 0) return fail("slot column not found");
          if (cells[s] <= last_cell[ci]) return fail("contributors not in ascending cell order");
          last_cell[ci] = cells[s];
          seen++;
          for (int a = 0; a < nd; a++) {
            const size_t pos = (size_t)a * gtk::kBatchSlots + n;
            const uint32_t o = bp.off_bytes == 1 ? off[pos] : reinterpret_cast<const uint16_t *>(off)[pos];
            if (o == none) continue;
            const int ent = abase + (int)o;
            if (ent >= acc_total) return fail("offset out of range");
            if (stamp[ent] == tt) return fail("two additions of one rank step touch one entry");
            stamp[ent] = tt;
            acc[ent] += 1.0 + cells[s] * 7.0 + a * 131.0 + lb * 977.0;
          }
        }
      if (seen != used) return fail("slots without a rank step");
      int ends = 0;
      for (int k = 0; k < acc_total; k++) ends += (endmask[k / 64] >> (k % 64)) & 1;
      if (ends != ncols || (acc_total > 0 && !((endmask[(acc_total - 1) / 64] >> ((acc_total - 1) % 64)) & 1)))
        return fail("end mask does not close every column");
      if (ncols > gtk::kBatchCols || bp.rounds != (nd + 3) / 4) return fail("batch shape");
      // the kernel's write-out: 64*rounds entries, padding entries go past nnz
      int c = 0;
      for (int k = 0; k < 64 * bp.rounds; k++) {
        if (c > 15) return fail("column index past the delta table");
        const int64_t z = (int64_t)k + delta[c];
        if (k < acc_total) {
          if (z < 0 || z >= p.nnz) return fail("segment out of range");
          nzval[z] = acc[k];
          written[z]++;
        } else if (z < p.nnz || z >= p.nnz + 64 * bp.rounds) {
          return fail("padding entry outside the value array's tail");
        }
        c += (endmask[k / 64] >> (k % 64)) & 1;
      }
    }
  }
  for (int64_t z = 0; z < p.nnz; z++)
    if (written[z] != 1) return fail("a nonzero is not written exactly once");
  return 0;
}

extern "C" int harness_run(int dim, int64_t n_nodes, const double *xyz, int64_t n_cells, int nv,
                           const int32_t *cell_nodes0, int nd, int64_t n_free, int64_t n_diri,
                           const int32_t *cell_dofs, int rows, int cols, int max_cells, int chunk,
                           int64_t *nnz_out, int32_t *colptr, int32_t *rowval, double *nzval, int64_t cap,
                           int *ntiles, int mode, char *err, int errlen) {
  gtk::MeshView m;
  m.dim = dim; m.n_nodes = n_nodes; m.n_cells = n_cells; m.nodes_per_cell = nv;
  m.xyz = xyz; m.cell_nodes = cell_nodes0; m.cell_dofs = cell_dofs; m.dofs_per_cell = nd;
  m.n_free = n_free; m.n_diri = n_diri;
  gtk::Pattern p;
  gtk::TilePlan t;
  gtk::OffsetPlan o;
  std::string e = gtk::build_pattern(m, rows, cols, p, 4);
  if (e.empty()) e = gtk::build_tiles(m, cols, max_cells, 1 << 20, 1 << 20, chunk, t, 4);
  if (e.empty() && mode != 2) e = gtk::build_offsets(m, p, t, o, 4);  // mode 2: batch records only
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
  if (mode >= 1) return replay_batches(m, p, t, nzval, err, errlen);
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

// CSR row order of a square CSC pattern (build_transpose_perm); returns 0,
// or -1 with the error text.
extern "C" int harness_transpose(int64_t n, const int32_t *colptr, const int32_t *rowval, int32_t *perm, char *err,
                                 int errlen) {
  gtk::Pattern p;
  p.nrows = p.ncols = n;
  p.nnz = colptr[n] - 1;
  p.colptr.assign(colptr, colptr + n + 1);
  p.rowval.assign(rowval, rowval + p.nnz);
  std::vector<int32_t> out;
  std::string e = gtk::build_transpose_perm(p, out);
  if (!e.empty()) {
    std::strncpy(err, e.c_str(), errlen - 1);
    return -1;
  }
  std::memcpy(perm, out.data(), out.size() * sizeof(int32_t));
  return 0;
}

// The symbolic phase alone with a given nonzero limit: 0 and nnz, or -1 and
// the error (the Int32 overflow check of the pattern, forced with a small
// limit).
extern "C" int harness_pattern(int dim, int64_t n_nodes, const double *xyz, int64_t n_cells, int nv,
                               const int32_t *cell_nodes0, int nd, int64_t n_free, int64_t n_diri,
                               const int32_t *cell_dofs, int rows, int cols, int64_t nnz_limit, int64_t *nnz_out,
                               char *err, int errlen) {
  gtk::MeshView m;
  m.dim = dim; m.n_nodes = n_nodes; m.n_cells = n_cells; m.nodes_per_cell = nv;
  m.xyz = xyz; m.cell_nodes = cell_nodes0; m.cell_dofs = cell_dofs; m.dofs_per_cell = nd;
  m.n_free = n_free; m.n_diri = n_diri;
  gtk::Pattern p;
  std::string e = gtk::build_pattern(m, rows, cols, p, 4, nnz_limit);
  if (!e.empty()) { std::strncpy(err, e.c_str(), errlen - 1); return -1; }
  *nnz_out = p.nnz;
  return 0;
}
