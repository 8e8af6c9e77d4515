// gtk_plan.cpp -- symbolic phase of libgtk_hip (host C++).
//
// build_pattern restates what the reference's symbolic pass plus compress
// produce (GalerkinToolkit.jl src/assembly.jl:75-157 counts the COO triplets
// cell by cell with trial DOFs outer and test DOFs inner, skipping DOFs
// outside the FREE/DIRICHLET selection; :571-575 compresses them into a
// SparseMatrixCSC{Float64,Int32}): column j holds the sorted union of the row
// DOFs of every cell that contains column DOF j, structural zeros included.
// The cached-CSC back end (src/assembly.jl:694-771) builds the same pattern.
//
// build_tiles / build_offsets are this library's own work decomposition:
// every column is owned by one tile, so the kernels write each nonzero once
// from one workgroup in a fixed contributor order (cells ascending, the
// order in which the reference's COO triplets are summed).
#include "gtk_plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>

namespace gtk {

namespace {

inline int32_t map_dof(int32_t dof, int set) {
  if (set == 1) return dof > 0 ? dof : 0;  // GTK_FREE
  return dof < 0 ? -dof : 0;               // GTK_DIRICHLET
}

void parallel_for(int64_t n, int nthreads, const std::function<void(int, int64_t, int64_t)> &f) {
  if (nthreads <= 1 || n < 4096) {
    f(0, 0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; t++) {
    int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
    if (b >= e) break;
    th.emplace_back(f, t, b, e);
  }
  for (auto &x : th) x.join();
}

// (cell, local dof) incidence of every DOF of set S, cells ascending.
struct Incidence {
  std::vector<int64_t> ptr;  // nset+1
  std::vector<int32_t> cell;
  std::vector<uint16_t> b;
};

Incidence incidence(const MeshView &m, int set, int64_t nset) {
  Incidence inc;
  inc.ptr.assign(nset + 2, 0);
  const int nd = m.dofs_per_cell;
  for (int64_t c = 0; c < m.n_cells; c++)
    for (int a = 0; a < nd; a++) {
      int32_t j = map_dof(m.cell_dofs[c * nd + a], set);
      if (j) inc.ptr[j]++;
    }
  // ptr[j] currently counts DOF j (1-based) -> exclusive prefix over 0-based
  std::vector<int64_t> start(nset + 1, 0);
  for (int64_t j = 0; j < nset; j++) start[j + 1] = start[j] + inc.ptr[j + 1];
  inc.ptr.assign(start.begin(), start.end());
  inc.cell.resize(start[nset]);
  inc.b.resize(start[nset]);
  std::vector<int64_t> fill(start.begin(), start.end() - 1);
  for (int64_t c = 0; c < m.n_cells; c++)
    for (int a = 0; a < nd; a++) {
      int32_t j = map_dof(m.cell_dofs[c * nd + a], set);
      if (!j) continue;
      int64_t p = fill[j - 1]++;
      inc.cell[p] = (int32_t)c;
      inc.b[p] = (uint16_t)a;
    }
  return inc;
}

uint64_t spread3(uint64_t x) {
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}

uint64_t spread2(uint64_t x) {
  x &= 0xffffffff;
  x = (x | x << 16) & 0x0000ffff0000ffffULL;
  x = (x | x << 8) & 0x00ff00ff00ff00ffULL;
  x = (x | x << 4) & 0x0f0f0f0f0f0f0f0fULL;
  x = (x | x << 2) & 0x3333333333333333ULL;
  x = (x | x << 1) & 0x5555555555555555ULL;
  return x;
}

}  // namespace

std::string build_pattern(const MeshView &m, int rows, int cols, Pattern &out, int nthreads, int64_t nnz_limit) {
  if (rows != 1 && rows != 2) return "rows must be GTK_FREE or GTK_DIRICHLET";
  if (cols != 1 && cols != 2) return "cols must be GTK_FREE or GTK_DIRICHLET";
  out.rows = rows;
  out.cols = cols;
  out.nrows = rows == 1 ? m.n_free : m.n_diri;
  out.ncols = cols == 1 ? m.n_free : m.n_diri;
  const int64_t ncols = out.ncols;
  const int nd = m.dofs_per_cell;
  Incidence inc = incidence(m, cols, ncols);
  std::vector<int64_t> len(ncols, 0);
  std::vector<std::vector<int32_t>> part(std::max(1, nthreads));
  std::vector<int64_t> part_begin(std::max(1, nthreads), 0), part_end(std::max(1, nthreads), 0);
  parallel_for(ncols, nthreads, [&](int t, int64_t b, int64_t e) {
    std::vector<int32_t> &buf = part[t];
    part_begin[t] = b;
    part_end[t] = e;
    std::vector<int32_t> tmp;
    for (int64_t j = b; j < e; j++) {
      tmp.clear();
      for (int64_t k = inc.ptr[j]; k < inc.ptr[j + 1]; k++) {
        const int32_t *dofs = m.cell_dofs + (int64_t)inc.cell[k] * nd;
        for (int a = 0; a < nd; a++) {
          int32_t r = map_dof(dofs[a], rows);
          if (r) tmp.push_back(r);
        }
      }
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      len[j] = (int64_t)tmp.size();
      buf.insert(buf.end(), tmp.begin(), tmp.end());
    }
  });
  out.colptr.assign(ncols + 1, 0);
  int64_t acc = 0;
  out.colptr[0] = 1;
  for (int64_t j = 0; j < ncols; j++) {
    acc += len[j];
    if (acc > nnz_limit) return "nnz exceeds the Int32 range of SparseMatrixCSC{Float64,Int32}";
    out.colptr[j + 1] = (int32_t)(acc + 1);
  }
  out.nnz = acc;
  out.rowval.resize(acc);
  for (size_t t = 0; t < part.size(); t++) {
    if (part[t].empty()) continue;
    int64_t dst = out.colptr[part_begin[t]] - 1;
    std::memcpy(out.rowval.data() + dst, part[t].data(), part[t].size() * sizeof(int32_t));
  }
  return "";
}

std::string build_tiles(const MeshView &m, int set, int max_cells, int max_cols, int max_ctr, int chunk,
                        TilePlan &out, int nthreads) {
  (void)nthreads;
  if (max_cells < 1 || max_cells > 65535) return "max_cells out of range";
  out = TilePlan();
  out.set = set;
  out.ndofs_cell = m.dofs_per_cell;
  const int64_t nset = set == 1 ? m.n_free : m.n_diri;
  out.nset = nset;
  Incidence inc = incidence(m, set, nset);
  // cells in Morton order of their centroids
  const int D = m.dim, nv = m.nodes_per_cell;
  std::vector<double> cen((size_t)m.n_cells * D, 0.0);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t c = 0; c < m.n_cells; c++)
    for (int d = 0; d < D; d++) {
      double s = 0;
      for (int n = 0; n < nv; n++) s += m.xyz[(int64_t)m.cell_nodes[c * nv + n] * D + d];
      s /= nv;
      cen[c * D + d] = s;
      lo[d] = std::min(lo[d], s);
      hi[d] = std::max(hi[d], s);
    }
  std::vector<uint64_t> key(m.n_cells);
  const double grid = D == 3 ? 2097151.0 : 4294967295.0;
  for (int64_t c = 0; c < m.n_cells; c++) {
    uint64_t q[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) {
      double ext = hi[d] - lo[d];
      double t = ext > 0 ? (cen[c * D + d] - lo[d]) / ext : 0.0;
      q[d] = (uint64_t)std::llround(t * grid);
    }
    key[c] = D == 3 ? (spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2) : (spread2(q[0]) | spread2(q[1]) << 1);
  }
  std::vector<int32_t> order(m.n_cells);
  for (int64_t c = 0; c < m.n_cells; c++) order[c] = (int32_t)c;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  std::vector<int32_t> rank(m.n_cells);
  for (int64_t i = 0; i < m.n_cells; i++) rank[order[i]] = (int32_t)i;
  // owner rank of each dof = smallest Morton rank among its cells
  std::vector<int32_t> owner(nset, INT32_MAX);
  for (int64_t j = 0; j < nset; j++)
    for (int64_t k = inc.ptr[j]; k < inc.ptr[j + 1]; k++) owner[j] = std::min(owner[j], rank[inc.cell[k]]);
  // dofs ordered by (owner rank, dof id): counting sort over ranks
  std::vector<int64_t> rp(m.n_cells + 1, 0);
  for (int64_t j = 0; j < nset; j++)
    if (owner[j] != INT32_MAX) rp[owner[j] + 1]++;
  for (int64_t i = 0; i < m.n_cells; i++) rp[i + 1] += rp[i];
  std::vector<int32_t> cols_sorted(rp[m.n_cells]);
  {
    std::vector<int64_t> f(rp.begin(), rp.end() - 1);
    for (int64_t j = 0; j < nset; j++)
      if (owner[j] != INT32_MAX) cols_sorted[f[owner[j]]++] = (int32_t)j;
  }
  // greedy tiles inside each chunk of `chunk` consecutive owner ranks
  std::vector<int32_t> mark(m.n_cells, -1);
  std::vector<int32_t> cur_cells, cur_cols;
  int cur_ctr = 0;
  auto flush = [&]() {
    if (cur_cols.empty()) return;
    int tile = out.ntiles++;
    std::sort(cur_cells.begin(), cur_cells.end());
    // columns with more contributors first: the kernels batch the t-th
    // contributor of consecutive columns, so similar counts waste no slots
    std::stable_sort(cur_cols.begin(), cur_cols.end(), [&](int32_t x, int32_t y) {
      return inc.ptr[x + 1] - inc.ptr[x] > inc.ptr[y + 1] - inc.ptr[y];
    });
    out.tile_cell_ptr.push_back((int32_t)out.tile_cells.size());
    out.tile_col_ptr.push_back((int32_t)out.col_id.size());
    out.tile_batch_ptr.push_back((int32_t)out.batch_col.size());
    {
      int slots = kBatchSlots + 1, ncolb = 0;
      for (size_t i = 0; i < cur_cols.size(); i++) {
        const int cnt = (int)(inc.ptr[cur_cols[i] + 1] - inc.ptr[cur_cols[i]]);
        int start = batch_slot_start(slots, cnt);
        if (start + cnt > kBatchSlots || ncolb == kBatchCols) {
          out.batch_col.push_back((int32_t)(out.col_id.size() + i));
          start = 0;
          ncolb = 0;
        }
        slots = start + cnt;
        ncolb++;
        out.max_batch_cols = std::max(out.max_batch_cols, ncolb);
      }
    }
    for (size_t s = 0; s < cur_cells.size(); s++) mark[cur_cells[s]] = (int32_t)s;  // slot
    out.max_tile_cells = std::max<int>(out.max_tile_cells, (int)cur_cells.size());
    out.max_tile_cols = std::max<int>(out.max_tile_cols, (int)cur_cols.size());
    out.tile_cells.insert(out.tile_cells.end(), cur_cells.begin(), cur_cells.end());
    for (int32_t j : cur_cols) {
      out.col_id.push_back(j);
      out.col_ctr_ptr.push_back((int32_t)out.ctr_slot.size());
      int n = 0;
      for (int64_t k = inc.ptr[j]; k < inc.ptr[j + 1]; k++) {
        out.ctr_slot.push_back((uint16_t)mark[inc.cell[k]]);
        out.ctr_b.push_back(inc.b[k]);
        n++;
      }
      out.max_ctr = std::max(out.max_ctr, n);
    }
    for (int32_t c : cur_cells) mark[c] = -1;
    (void)tile;
    out.max_tile_ctr = std::max<int>(out.max_tile_ctr, cur_ctr);
    cur_cells.clear();
    cur_cols.clear();
    cur_ctr = 0;
  };
  std::vector<int32_t> in_tile(m.n_cells, 0);
  int32_t stamp = 0;
  int64_t chunk_id = -1;
  for (size_t i = 0; i < cols_sorted.size(); i++) {
    int32_t j = cols_sorted[i];
    int64_t cid = owner[j] / std::max(1, chunk);
    if (cid != chunk_id) {
      flush();
      chunk_id = cid;
      stamp++;
    }
    // cells this column would add
    int add = 0;
    const int nctr_j = (int)(inc.ptr[j + 1] - inc.ptr[j]);
    for (int64_t k = inc.ptr[j]; k < inc.ptr[j + 1]; k++)
      if (in_tile[inc.cell[k]] != stamp) add++;
    if (!cur_cols.empty() && ((int)cur_cells.size() + add > max_cells || (int)cur_cols.size() + 1 > max_cols ||
                              cur_ctr + nctr_j > max_ctr)) {
      flush();
      stamp++;
    }
    cur_ctr += nctr_j;
    if (add > max_cells) return "a single DOF touches more cells than a tile may hold";
    for (int64_t k = inc.ptr[j]; k < inc.ptr[j + 1]; k++) {
      int32_t c = inc.cell[k];
      if (in_tile[c] != stamp) {
        in_tile[c] = stamp;
        cur_cells.push_back(c);
      }
    }
    cur_cols.push_back(j);
  }
  flush();
  out.tile_cell_ptr.push_back((int32_t)out.tile_cells.size());
  out.tile_col_ptr.push_back((int32_t)out.col_id.size());
  out.col_ctr_ptr.push_back((int32_t)out.ctr_slot.size());
  out.tile_batch_ptr.push_back((int32_t)out.batch_col.size());
  out.batch_col.push_back((int32_t)out.col_id.size());
  // DOFs of S that no cell touches (boundary items cover only boundary DOFs)
  // are in no tile: their matrix columns are empty, their vector rows zero
  if ((int64_t)out.col_id.size() > nset) return "tile plan lists a DOF twice";
  out.uncovered = nset - (int64_t)out.col_id.size();
  for (int tl = 0; tl < out.ntiles; tl++)
    out.max_tile_batches = std::max(out.max_tile_batches, out.tile_batch_ptr[tl + 1] - out.tile_batch_ptr[tl]);
  return "";
}

std::string build_batches(const MeshView &m, const Pattern &p, const TilePlan &t, BatchPlan &out, int nthreads) {
  if (p.cols != t.set) return "tile plan set differs from the pattern columns";
  const int nd = m.dofs_per_cell;
  out = BatchPlan();
  if (t.max_ctr > kBatchSlots) return "a column has more contributors than a batch holds";
  if (t.max_tile_cells > 65535) return "tile too large for batch records";
  int max_seg = 0;
  for (int64_t j = 0; j < p.ncols; j++) max_seg = std::max(max_seg, p.colptr[j + 1] - p.colptr[j]);
  if (max_seg >= 0xFFFF) return "column segment too long for 16-bit offsets";
  out.off_bytes = max_seg < 0xFF ? 1 : 2;
  out.nbatches = (int64_t)t.batch_col.size() - 1;
  // accumulator entries of each batch (sum of its columns' segment lengths)
  std::vector<int> bacc(std::max<int64_t>(1, out.nbatches), 0);
  for (int64_t b = 0; b < out.nbatches; b++)
    for (int32_t pc = t.batch_col[b]; pc < t.batch_col[b + 1]; pc++) {
      const int32_t j = t.col_id[pc];
      bacc[b] += p.colptr[j + 1] - p.colptr[j];
    }
  if (out.nbatches > 0) out.max_batch_acc = *std::max_element(bacc.begin(), bacc.end());
  if (out.max_batch_acc >= 0xFFFF) return "batch accumulator exceeds 16-bit offsets";
  out.rounds = (nd + 3) / 4;  // 16 slots x nd rows fill at most this many rounds of 64
  if (out.max_batch_acc > 64 * out.rounds) return "batch accumulator exceeds its write-out rounds";
  out.endmask_words = out.rounds;
  out.off_offset = BatchPlan::kEndmask + 8 * out.endmask_words;
  const int off_row = kBatchSlots * out.off_bytes;  // bytes per local row
  out.rec_bytes = (out.off_offset + nd * off_row + 15) / 16 * 16;
  out.rec.assign((size_t)out.nbatches * out.rec_bytes, 0);
  std::vector<int32_t> batch_tile(std::max<int64_t>(0, out.nbatches));
  for (int tl = 0; tl < t.ntiles; tl++)
    for (int32_t b = t.tile_batch_ptr[tl]; b < t.tile_batch_ptr[tl + 1]; b++) batch_tile[b] = tl;
  for (int64_t b = 0; b < out.nbatches; b++)
    if (t.batch_col[b + 1] - t.batch_col[b] > kBatchCols) return "batch with too many columns";
  // the rows of one slot are added by the lanes of one instruction: they
  // must be distinct (no cell lists one DOF of the row set twice)
  std::vector<int64_t> dup(std::max(1, nthreads), -1);
  parallel_for(out.nbatches, nthreads, [&](int th, int64_t b0, int64_t b1) {
    std::vector<int64_t> stamp(std::max(1, max_seg), -1);
    for (int64_t b = b0; b < b1; b++) {
      uint8_t *r = out.rec.data() + (size_t)b * out.rec_bytes;
      uint32_t *sb = reinterpret_cast<uint32_t *>(r + BatchPlan::kSlotSb);
      uint32_t *at = reinterpret_cast<uint32_t *>(r + BatchPlan::kSlotAt);
      int32_t *delta = reinterpret_cast<int32_t *>(r + BatchPlan::kDelta);
      uint64_t *endmask = reinterpret_cast<uint64_t *>(r + BatchPlan::kEndmask);
      uint8_t *off = r + out.off_offset;
      std::memset(off, 0xFF, (size_t)nd * off_row);
      for (int n = 0; n < kBatchSlots; n++) at[n] = 0xFFFFu << 16;
      const int32_t *cells = t.tile_cells.data() + t.tile_cell_ptr[batch_tile[b]];
      int ns = 0, tmax = 0, abase = 0, cd = 0;
      const int ncols = t.batch_col[b + 1] - t.batch_col[b];
      for (int ci = 0; ci < ncols; ci++) {
        const int32_t pc = t.batch_col[b] + ci;
        const int32_t j = t.col_id[pc];
        const int32_t *seg = p.rowval.data() + (p.colptr[j] - 1);
        const int32_t len = p.colptr[j + 1] - p.colptr[j];
        // a column with no row of the row set (a Dirichlet vertex whose
        // cells have only Dirichlet DOFs, in a FREE x DIRICHLET block) has
        // nothing to add or write: no slots, no end bit, no delta (the
        // write-out finds column c's delta at the c-th end bit)
        if (len == 0) continue;
        delta[cd++] = p.colptr[j] - 1 - abase;
        endmask[(abase + len - 1) / 64] |= 1ull << ((abase + len - 1) % 64);
        const int cnt = t.col_ctr_ptr[pc + 1] - t.col_ctr_ptr[pc];
        tmax = std::max(tmax, cnt);
        ns = batch_slot_start(ns, cnt);  // slots skipped here stay empty
        for (int tt = 0; tt < cnt; tt++, ns++) {
          const int32_t k = t.col_ctr_ptr[pc] + tt;
          sb[ns] = (uint32_t)t.ctr_slot[k] | (uint32_t)t.ctr_b[k] << 16;
          at[ns] = (uint32_t)abase | (uint32_t)tt << 16;
          const int32_t *dofs = m.cell_dofs + (int64_t)cells[t.ctr_slot[k]] * nd;
          for (int a = 0; a < nd; a++) {
            int32_t rr = map_dof(dofs[a], p.rows);
            if (!rr) continue;
            const int o = (int)(std::lower_bound(seg, seg + len, rr) - seg);
            if (stamp[o] == k) dup[th] = cells[t.ctr_slot[k]];
            stamp[o] = k;
            const size_t pos = (size_t)a * kBatchSlots + ns;
            if (out.off_bytes == 1) off[pos] = (uint8_t)o;
            else reinterpret_cast<uint16_t *>(off)[pos] = (uint16_t)o;
          }
        }
        abase += len;
      }
      if (abase < 64 * out.rounds) {  // padding column up to the last round
        delta[cd] = (int32_t)(p.nnz - abase);
        endmask[out.rounds - 1] |= 1ull << 63;
      }
      reinterpret_cast<uint32_t *>(r)[0] = (uint32_t)ns | (uint32_t)tmax << 8 | (uint32_t)cd << 16;
      reinterpret_cast<uint32_t *>(r)[1] = (uint32_t)abase;
    }
  });
  for (int64_t c : dup)
    if (c >= 0) return "cell " + std::to_string(c) + " lists one DOF of the row set twice";
  return "";
}

std::string build_transpose_perm(const Pattern &p, std::vector<int32_t> &perm) {
  if (p.nrows != p.ncols) return "the block is not square";
  const int64_t n = p.ncols;
  // row counts of the transpose must equal the column counts
  std::vector<int64_t> cnt(n, 0);
  for (int64_t k = 0; k < p.nnz; k++) cnt[p.rowval[k] - 1]++;
  for (int64_t i = 0; i < n; i++)
    if (cnt[i] != p.colptr[i + 1] - p.colptr[i]) return "the pattern is not structurally symmetric";
  perm.assign(p.nnz, 0);
  std::vector<int64_t> next(n);
  for (int64_t i = 0; i < n; i++) next[i] = p.colptr[i] - 1;
  // columns ascending: row i receives its entries in ascending column order
  for (int64_t j = 0; j < n; j++)
    for (int64_t k = p.colptr[j] - 1; k < p.colptr[j + 1] - 1; k++) {
      const int64_t i = p.rowval[k] - 1;
      const int64_t slot = next[i]++;
      if (p.rowval[slot] - 1 != j) return "the pattern is not structurally symmetric";
      perm[slot] = (int32_t)k;
    }
  return "";
}

std::string build_offsets(const MeshView &m, const Pattern &p, const TilePlan &t, OffsetPlan &out, int nthreads) {
  if (p.cols != t.set) return "tile plan set differs from the pattern columns";
  const int nd = m.dofs_per_cell;
  const int64_t ncol_plan = (int64_t)t.col_id.size();
  int max_seg = 0;
  for (int64_t j = 0; j < p.ncols; j++) max_seg = std::max(max_seg, p.colptr[j + 1] - p.colptr[j]);
  if (max_seg >= kNoRow) return "column segment too long for 16-bit offsets";
  out.max_seg = max_seg;
  out.stride = (nd + 15) / 16 * 16;
  out.bytes = max_seg < 0xFF ? 1 : 2;
  out.data.assign(t.ctr_slot.size() * (size_t)out.stride * out.bytes, 0xFF);
  // accumulator footprint of a group of kColGroup consecutive planned columns
  int max_group = 0;
  for (int tl = 0; tl < t.ntiles; tl++)
    for (int32_t g0 = t.tile_col_ptr[tl]; g0 < t.tile_col_ptr[tl + 1]; g0 += kColGroup) {
      int sum = 0;
      for (int32_t i = g0; i < std::min(g0 + kColGroup, t.tile_col_ptr[tl + 1]); i++)
        sum += p.colptr[t.col_id[i] + 1] - p.colptr[t.col_id[i]];
      max_group = std::max(max_group, sum);
    }
  out.max_group_acc = max_group;
  int max_batch = 0;
  for (size_t b = 0; b + 1 < t.batch_col.size(); b++) {
    int sum = 0;
    for (int32_t i = t.batch_col[b]; i < t.batch_col[b + 1]; i++)
      sum += p.colptr[t.col_id[i] + 1] - p.colptr[t.col_id[i]];
    max_batch = std::max(max_batch, sum);
  }
  out.max_batch_acc = max_batch;
  // tile of each planned column (to read tile cells)
  std::vector<int32_t> col_tile(ncol_plan);
  for (int tl = 0; tl < t.ntiles; tl++)
    for (int32_t i = t.tile_col_ptr[tl]; i < t.tile_col_ptr[tl + 1]; i++) col_tile[i] = tl;
  // the VALU kernel's lanes of one contributor add their rows into the
  // column accumulator in one instruction, in program order: that is
  // deterministic only if those rows are distinct, i.e. no cell lists one
  // DOF of the row set twice
  std::vector<int64_t> dup(std::max(1, nthreads), -1);
  parallel_for(ncol_plan, nthreads, [&](int th, int64_t b, int64_t e) {
    std::vector<int64_t> stamp(std::max(1, max_seg), -1);
    for (int64_t i = b; i < e; i++) {
      int32_t j = t.col_id[i];
      const int32_t *seg = p.rowval.data() + (p.colptr[j] - 1);
      const int32_t len = p.colptr[j + 1] - p.colptr[j];
      const int32_t *cells = t.tile_cells.data() + t.tile_cell_ptr[col_tile[i]];
      for (int32_t k = t.col_ctr_ptr[i]; k < t.col_ctr_ptr[i + 1]; k++) {
        int32_t c = cells[t.ctr_slot[k]];
        const int32_t *dofs = m.cell_dofs + (int64_t)c * nd;
        for (int a = 0; a < nd; a++) {
          int32_t r = map_dof(dofs[a], p.rows);
          if (!r) continue;
          const int32_t *it = std::lower_bound(seg, seg + len, r);
          const int32_t o = (int32_t)(it - seg);
          if (stamp[o] == k) dup[th] = c;
          stamp[o] = k;
          const size_t pos = (size_t)k * out.stride + a;
          if (out.bytes == 1) out.data[pos] = (uint8_t)o;
          else reinterpret_cast<uint16_t *>(out.data.data())[pos] = (uint16_t)o;
        }
      }
    }
  });
  for (int64_t c : dup)
    if (c >= 0) return "cell " + std::to_string(c) + " lists one DOF of the row set twice";
  return std::string();
}

}  // namespace gtk
