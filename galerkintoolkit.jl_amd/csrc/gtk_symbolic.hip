// gtk_symbolic.hip -- the symbolic phase on the GPU: the CSC pattern of a
// (rows, cols) block, equal to the pattern compress() gives for the COO
// triplets of allocate_matrix (GalerkinToolkit.jl src/assembly.jl:75-157,
// :571-575: per column the distinct rows of its cells, ascending), and the
// cell-local-entry -> nonzero map the reference caches for reassembly
// (src/assembly.jl:694-771).
//
// Column incidence by counting (integer atomics: the counts are exact; the
// order of cells inside a column is not used), one wavefront per column
// gathering its cells' rows into LDS, a bitonic sort and a ballot over
// adjacent entries for the distinct rows, and exclusive scans for the
// pointers.  The output does not depend on scheduling.
#include "gtk_symbolic.h"

namespace gtk {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanBlock = kScanThreads * kScanItems;

// GT signed DOF id -> 1-based index in the set (1: free, 2: Dirichlet), 0 outside
__device__ inline int32_t in_set(int32_t dof, int set) {
  return set == 1 ? (dof > 0 ? dof : 0) : (dof < 0 ? -dof : 0);
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_block_reduce(const T *in, int64_t n, int64_t *sums) {
  __shared__ int64_t s[kScanThreads];
  const int64_t b0 = (int64_t)blockIdx.x * kScanBlock + (int64_t)threadIdx.x * kScanItems;
  int64_t acc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++)
    if (b0 + i < n) acc += (int64_t)in[b0 + i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kScanThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = s[0];
}

// out[i] = base + offsets[block] + sum of the block's elements before i;
// the thread holding element n-1 also writes out[n]
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_block_scan(const T *in, int64_t n, const int64_t *offsets,
                                                             int64_t base, int64_t *out) {
  __shared__ int64_t s[kScanThreads];
  const int tid = threadIdx.x;
  const int64_t b0 = (int64_t)blockIdx.x * kScanBlock + (int64_t)tid * kScanItems;
  int64_t v[kScanItems];
  int64_t acc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = b0 + i < n ? (int64_t)in[b0 + i] : 0;
    acc += v[i];
  }
  s[tid] = acc;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int64_t t = tid >= off ? s[tid - off] : 0;
    __syncthreads();
    s[tid] += t;
    __syncthreads();
  }
  int64_t run = base + (offsets ? offsets[blockIdx.x] : 0) + (tid ? s[tid - 1] : 0);
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    if (b0 + i < n) out[b0 + i] = run;
    run += v[i];
  }
  if (b0 <= n - 1 && n - 1 < b0 + kScanItems) out[n] = run;
}

__global__ void k_store_base(int64_t *out, int64_t base) { out[0] = base; }

int64_t blocks_of(int64_t n) { return (n + kScanBlock - 1) / kScanBlock; }

template <typename T>
void scan_rec(const T *in, int64_t n, int64_t base, int64_t *out, int64_t *tmp, hipStream_t st) {
  if (n == 0) {
    k_store_base<<<1, 1, 0, st>>>(out, base);
    return;
  }
  const int64_t nb = blocks_of(n);
  if (nb == 1) {
    k_block_scan<T><<<1, kScanThreads, 0, st>>>(in, n, nullptr, base, out);
    return;
  }
  int64_t *sums = tmp, *offs = tmp + nb, *rest = tmp + 2 * nb + 1;
  k_block_reduce<T><<<(unsigned)nb, kScanThreads, 0, st>>>(in, n, sums);
  scan_rec<int64_t>(sums, nb, 0, offs, rest, st);
  k_block_scan<T><<<(unsigned)nb, kScanThreads, 0, st>>>(in, n, offs, base, out);
}

__global__ void k_column_count(const int32_t *cell_dofs, int64_t total, int cols, int32_t *counts) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = in_set(cell_dofs[i], cols);
    if (c) atomicAdd(&counts[c - 1], 1);
  }
}

// *out = max(*out, largest count)
__global__ void k_max_count(const int32_t *counts, int64_t n, int32_t *out) {
  int32_t m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, counts[i]);
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, m);
}

__global__ void k_column_fill(const int32_t *cell_dofs, int64_t total, int nd, int cols, const int64_t *ptr,
                              int32_t *fill, int32_t *inc) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = in_set(cell_dofs[i], cols);
    if (c) {
      const int pos = atomicAdd(&fill[c - 1], 1);
      inc[ptr[c - 1] + pos] = (int32_t)(i / nd);
    }
  }
}

// candidate row e of column j (0 when the DOF is not in the row set)
__device__ inline int32_t column_candidate(const int32_t *cell_dofs, int nd, int rows, const int32_t *inc, int64_t k0,
                                           int e) {
  const int32_t cell = inc[k0 + e / nd];
  return in_set(cell_dofs[(int64_t)cell * nd + e % nd], rows);
}

// one wavefront (the whole workgroup) per column: the column's candidate
// rows (its cells' DOFs in the row set) are sorted by a bitonic network,
// in registers (one per lane, partners by lane xor) when they fit one
// wavefront, else in LDS (`cap` entries, a power of two, dynamic shared
// memory); the first of each run of equal rows is kept
__global__ __launch_bounds__(64) void k_column_rows(const int32_t *cell_dofs, int nd, int rows, int64_t ncols,
                                                    const int64_t *ptr, const int32_t *inc, int pass, int cap,
                                                    int32_t *len, int32_t *maxlist, const int64_t *colptr64,
                                                    int32_t *rowval) {
  extern __shared__ int32_t L[];
  const int tid = threadIdx.x;
  for (int64_t j = blockIdx.x; j < ncols; j += gridDim.x) {
    const int64_t k0 = ptr[j];
    const int64_t total = (ptr[j + 1] - k0) * nd;  // candidate rows, duplicates included
    if (total > (cap > 64 ? cap : 64)) {
      if (tid == 0) atomicMax(maxlist, (int32_t)(total < INT32_MAX ? total : INT32_MAX));
      continue;
    }
    int run = 0;
    if (total <= 64) {
      int32_t x = INT32_MAX;  // sorts to the end, never counted
      if (tid < total) {
        const int32_t m = column_candidate(cell_dofs, nd, rows, inc, k0, tid);
        if (m) x = m;
      }
      for (int k = 2; k <= 64; k <<= 1)
        for (int sd = k >> 1; sd > 0; sd >>= 1) {
          const int32_t y = __shfl_xor(x, sd);
          const bool up = (tid & k) == 0, lower = (tid & sd) == 0;
          x = (lower == up) ? min(x, y) : max(x, y);
        }
      const int32_t prev = __shfl(x, tid > 0 ? tid - 1 : 0);
      const bool f = x != INT32_MAX && (tid == 0 || prev != x);
      const uint64_t m = __ballot(f);
      if (pass == 1 && f) rowval[colptr64[j] - 1 + __popcll(m & ((1ull << tid) - 1))] = x;
      run = __popcll(m);
    } else {
      int P = 1, lp = 0;
      while (P < total) {
        P <<= 1;
        lp++;
      }
      for (int e = tid; e < P; e += 64) {
        int32_t r = INT32_MAX;
        if (e < total) {
          const int32_t m = column_candidate(cell_dofs, nd, rows, inc, k0, e);
          if (m) r = m;
        }
        L[e] = r;
      }
      __syncthreads();
      for (int lk = 1; lk <= lp; lk++)
        for (int ls = lk - 1; ls >= 0; ls--) {
          const int sd = 1 << ls;
          for (int i = tid; i < P / 2; i += 64) {
            // i-th pair of the step: lo has bit ls clear, hi = lo + sd
            const int lo = ((i >> ls) << (ls + 1)) | (i & (sd - 1)), hi = lo + sd;
            const bool up = ((lo >> lk) & 1) == 0;
            const int32_t x = L[lo], y = L[hi];
            if ((x > y) == up) {
              L[lo] = y;
              L[hi] = x;
            }
          }
          __syncthreads();
        }
      for (int e0 = 0; e0 < P; e0 += 64) {
        const int e = e0 + tid;
        const int32_t x = e < P ? L[e] : INT32_MAX;
        const bool f = x != INT32_MAX && (e == 0 || L[e - 1] != x);
        const uint64_t m = __ballot(f);
        if (pass == 1 && f) rowval[colptr64[j] - 1 + run + __popcll(m & ((1ull << tid) - 1))] = x;
        run += __popcll(m);
      }
      __syncthreads();  // L is reused by the next column
    }
    if (pass == 0 && tid == 0) len[j] = run;
  }
}

__global__ void k_colptr32(const int64_t *c64, int64_t n, int32_t *c32) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (int64_t)gridDim.x * blockDim.x)
    c32[i] = (int32_t)c64[i];
}

__global__ void k_local_map(const int32_t *cell_dofs, int64_t n_cells, int nd, int rows, int cols,
                            const int32_t *colptr, const int32_t *rowval, int32_t *map) {
  const int64_t total = n_cells * nd * nd;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t cell = i / ((int64_t)nd * nd);
    const int rem = (int)(i - cell * nd * nd), b = rem / nd, a = rem - b * nd;
    const int32_t c = in_set(cell_dofs[cell * nd + b], cols);
    const int32_t r = in_set(cell_dofs[cell * nd + a], rows);
    int32_t out = 0;
    if (c && r) {
      int lo = colptr[c - 1] - 1, hi = colptr[c] - 1;  // rowval[lo, hi) ascending
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rowval[mid] < r) lo = mid + 1;
        else hi = mid;
      }
      out = lo + 1;
    }
    map[i] = out;
  }
}

// one thread per column: every dst entry gets at most one src entry
// kAddLanes lanes per column: lane l takes the column's source entries
// l, l + kAddLanes, ...; the binary searches of a column's lanes read the
// same destination segment (21.2 M volume entries of a 512^2 Q2 DG problem
// into its skeleton pattern: 3.09 -> 0.47 ms against a thread per column)
constexpr int kAddLanes = 16;

__global__ void k_add_block(int64_t ncols, const int32_t *scolptr, const int32_t *srowval, const double *sval,
                            const int32_t *dcolptr, const int32_t *drowval, double *dval, int32_t *missing) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = (int)(t0 % kAddLanes);
  for (int64_t j = t0 / kAddLanes; j < ncols; j += (int64_t)gridDim.x * blockDim.x / kAddLanes) {
    const int d0 = dcolptr[j] - 1, d1 = dcolptr[j + 1] - 1;
    for (int k = scolptr[j] - 1 + lane; k < scolptr[j + 1] - 1; k += kAddLanes) {
      const int32_t r = srowval[k];
      int lo = d0, hi = d1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (drowval[mid] < r) lo = mid + 1;
        else hi = mid;
      }
      if (lo < d1 && drowval[lo] == r) dval[lo] += sval[k];
      else *missing = 1;
    }
  }
}

// dst[t] = sum over k in [ptr[t], ptr[t+1]) of src[idx[k]], in k order
__global__ void k_gather_block(int64_t n, const int64_t *ptr, const int64_t *idx, const double *src, double *dst) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int64_t k = ptr[t]; k < ptr[t + 1]; k++) s += src[idx[k]];
    dst[t] = s;
  }
}

unsigned grid_for(int64_t n, int block) {
  const int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

}  // namespace

size_t symbolic_scan_tmp(int64_t n) {
  size_t t = 0;
  int64_t m = n;
  while (blocks_of(m) > 1) {
    const int64_t nb = blocks_of(m);
    t += 2 * nb + 1;
    m = nb;
  }
  return t + 1;
}

void launch_scan(const int32_t *counts, int64_t n, int64_t base, int64_t *out, int64_t *tmp, hipStream_t st) {
  scan_rec<int32_t>(counts, n, base, out, tmp, st);
}

void launch_column_count(const int32_t *cell_dofs, int64_t n_cells, int nd, int cols, int32_t *counts,
                         hipStream_t st) {
  const int64_t total = n_cells * nd;
  k_column_count<<<grid_for(total, 256), 256, 0, st>>>(cell_dofs, total, cols, counts);
}

void launch_column_fill(const int32_t *cell_dofs, int64_t n_cells, int nd, int cols, const int64_t *ptr,
                        int32_t *fill, int32_t *inc, hipStream_t st) {
  const int64_t total = n_cells * nd;
  k_column_fill<<<grid_for(total, 256), 256, 0, st>>>(cell_dofs, total, nd, cols, ptr, fill, inc);
}

void launch_max_count(const int32_t *counts, int64_t n, int32_t *out, hipStream_t st) {
  if (n == 0) return;
  k_max_count<<<grid_for(n, 256), 256, 0, st>>>(counts, n, out);
}

void launch_column_rows(const int32_t *cell_dofs, int nd, int rows, int64_t ncols, const int64_t *ptr,
                        const int32_t *inc, int pass, int cap, int32_t *len, int32_t *maxlist,
                        const int64_t *colptr64, int32_t *rowval, hipStream_t st) {
  if (ncols == 0) return;
  const unsigned g = (unsigned)(ncols < (1 << 20) ? ncols : (1 << 20));
  const size_t shmem = cap > 64 ? (size_t)cap * sizeof(int32_t) : 0;
  k_column_rows<<<g, 64, shmem, st>>>(cell_dofs, nd, rows, ncols, ptr, inc, pass, cap, len, maxlist, colptr64,
                                      rowval);
}

void launch_colptr32(const int64_t *colptr64, int64_t n, int32_t *colptr, hipStream_t st) {
  k_colptr32<<<grid_for(n + 1, 256), 256, 0, st>>>(colptr64, n, colptr);
}

void launch_local_map(const int32_t *cell_dofs, int64_t n_cells, int nd, int rows, int cols, const int32_t *colptr,
                      const int32_t *rowval, int32_t *map, hipStream_t st) {
  const int64_t total = n_cells * nd * nd;
  if (total == 0) return;
  k_local_map<<<grid_for(total, 256), 256, 0, st>>>(cell_dofs, n_cells, nd, rows, cols, colptr, rowval, map);
}

void launch_gather_block(int64_t n, const int64_t *ptr, const int64_t *idx, const double *src, double *dst,
                         hipStream_t st) {
  if (n == 0) return;
  k_gather_block<<<grid_for(n, 256), 256, 0, st>>>(n, ptr, idx, src, dst);
}

void launch_add_block(int64_t ncols, const int32_t *scolptr, const int32_t *srowval, const double *sval,
                      const int32_t *dcolptr, const int32_t *drowval, double *dval, int32_t *missing, hipStream_t st) {
  if (ncols == 0) return;
  k_add_block<<<grid_for(ncols * kAddLanes, 256), 256, 0, st>>>(ncols, scolptr, srowval, sval, dcolptr, drowval, dval,
                                                                 missing);
}

}  // namespace gtk
