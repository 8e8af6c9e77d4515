// gtk_solve.hip -- gfx950 kernels of the mixed-precision solve of an
// assembled (FREE, FREE) block: FP64 iterative refinement around an FP32
// Jacobi-preconditioned conjugate gradient.
//
// The matrices are the library's CSC blocks of symmetric forms, whose
// patterns are structurally symmetric: (colptr, rowval) serve as CSR row
// pointers and column indices (1-based), with the values gathered into row
// order once per solve (the values themselves need not be exactly
// symmetric: a mixed-precision assembly is not).
//
// Reductions are deterministic: every block sums a fixed strided subset in a
// fixed tree order into its own partial, and one block sums the partials in
// order; dot products accumulate in FP64 for both precisions.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gtk_solve.h"

namespace gtk {

namespace {

constexpr int kRB = 256;  // threads per block of the vector kernels

__device__ inline double block_sum(double v, double *sh) {
  // wave reduction (fixed xor tree) then the waves in order
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += sh[i];
  }
  __syncthreads();
  return s;  // valid in thread 0
}

// sum of np partials in a fixed order, the same in every block, returned to
// every thread (sh: kRB / 64 + 1 doubles)
__device__ inline double block_sum_partials(int np, const double *__restrict__ partial, double *sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += kRB) v += partial[i];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) sh[kRB / 64] = v;
  __syncthreads();
  v = sh[kRB / 64];
  __syncthreads();
  return v;
}

// y = A x (or y = b - A x when b != nullptr); S lanes per row.
template <typename AT, typename XT, typename YT, int S>
__global__ __launch_bounds__(kRB) void spmv_kernel(int64_t n, const int32_t *__restrict__ ptr,
                                                   const int32_t *__restrict__ idx, const AT *__restrict__ val,
                                                   const XT *__restrict__ x, const YT *__restrict__ b,
                                                   YT *__restrict__ y) {
  using Acc = typename std::conditional<sizeof(AT) == 8 || sizeof(XT) == 8, double, float>::type;
  const int64_t row = ((int64_t)blockIdx.x * kRB + threadIdx.x) / S;
  const int sl = threadIdx.x & (S - 1);
  Acc acc = 0;
  if (row < n) {
    const int32_t k0 = ptr[row] - 1, k1 = ptr[row + 1] - 1;
    // four entries per lane in flight: index loads, then the x gathers
    int32_t k = k0 + sl;
    for (; k + 3 * S < k1; k += 4 * S) {
      int32_t c[4];
      AT v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        c[u] = idx[k + u * S];
        v[u] = val[k + u * S];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) acc += (Acc)v[u] * (Acc)x[c[u] - 1];
    }
    for (; k < k1; k += S) acc += (Acc)val[k] * (Acc)x[idx[k] - 1];
  }
#pragma unroll
  for (int o = S / 2; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, S);
  if (row < n && sl == 0) y[row] = b ? (YT)((Acc)b[row] - acc) : (YT)acc;
}

template <int S>
void launch_spmv_s(const SolveMat &m, const void *val, bool val64, const void *x, bool x64, const void *b, void *y,
                   bool y64, hipStream_t st) {
  const int64_t threads = m.n * S;
  const dim3 grid((unsigned)((threads + kRB - 1) / kRB));
  if (val64 && x64 && y64)
    spmv_kernel<double, double, double, S><<<grid, kRB, 0, st>>>(m.n, m.ptr, m.idx, (const double *)val,
                                                                  (const double *)x, (const double *)b, (double *)y);
  else
    spmv_kernel<float, float, float, S><<<grid, kRB, 0, st>>>(m.n, m.ptr, m.idx, (const float *)val,
                                                               (const float *)x, (const float *)b, (float *)y);
}

// partial[blockIdx] = sum over this block's strided share of a_i * b_i
template <typename T>
__global__ __launch_bounds__(kRB) void dot_partial_kernel(int64_t n, const T *__restrict__ a, const T *__restrict__ b,
                                                          double *__restrict__ partial) {
  __shared__ double sh[kRB / 64];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB)
    v += (double)a[i] * (double)b[i];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// sums np partials in order into sc[slot]
__global__ __launch_bounds__(kRB) void sum_partials_kernel(int np, const double *__restrict__ partial,
                                                           double *__restrict__ sc, int slot) {
  __shared__ double sh[kRB / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += kRB) v += partial[i];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) sc[slot] = v;
}

// CG scalars in sc: r.z of the current direction in [0] or [3] (the two
// slots alternate by iteration), [5] r.r when read for convergence.
//
// alpha = r.z / p.q (p.q from the partials of p . q, summed by every
// block in the same order); d += alpha p; r -= alpha q; z = dinv r;
// partials of the new r.z and of r.r
template <typename T>
__global__ __launch_bounds__(kRB) void cg_update_kernel(int64_t n, const double *__restrict__ sc, int cur,
                                                        T *__restrict__ d, T *__restrict__ r, T *__restrict__ z,
                                                        const T *__restrict__ p, const T *__restrict__ q,
                                                        const T *__restrict__ dinv, const double *__restrict__ part_pq,
                                                        double *__restrict__ part_rz, double *__restrict__ part_rr) {
  __shared__ double sh[kRB / 64 + 1];
  const double pq = block_sum_partials(kSolvePartials, part_pq, sh);
  const T alpha = (T)(pq != 0.0 ? sc[cur] / pq : 0.0);
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) {
    d[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = dinv[i] * ri;
    z[i] = zi;
    rz += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
  rz = block_sum(rz, sh);
  if (threadIdx.x == 0) part_rz[blockIdx.x] = rz;
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = rr;
}

// beta = r.z(new) / r.z; p = z + beta p; block 0 stores the new r.z in the
// other slot (read by the next iteration only)
template <typename T>
__global__ __launch_bounds__(kRB) void cg_direction_kernel(int64_t n, double *__restrict__ sc, int cur,
                                                           const double *__restrict__ part_rz, const T *__restrict__ z,
                                                           T *__restrict__ p) {
  __shared__ double sh[kRB / 64 + 1];
  const double rz_new = block_sum_partials(kSolvePartials, part_rz, sh);
  const double rz = sc[cur];
  const T beta = (T)(rz != 0.0 ? rz_new / rz : 0.0);
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) p[i] = z[i] + beta * p[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[cur == 0 ? 3 : 0] = rz_new;
}

// start of an inner solve: d = 0, r = s * r64, z = dinv r, p = z; partials of r.z
template <typename T>
__global__ __launch_bounds__(kRB) void cg_start_kernel(int64_t n, double s, const double *__restrict__ r64,
                                                       T *__restrict__ d, T *__restrict__ r, T *__restrict__ z,
                                                       T *__restrict__ p, const T *__restrict__ dinv,
                                                       double *__restrict__ part_rz) {
  __shared__ double sh[kRB / 64];
  double rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) {
    const T ri = (T)(s * r64[i]);
    d[i] = (T)0;
    r[i] = ri;
    const T zi = dinv[i] * ri;
    z[i] = zi;
    p[i] = zi;
    rz += (double)ri * (double)zi;
  }
  rz = block_sum(rz, sh);
  if (threadIdx.x == 0) part_rz[blockIdx.x] = rz;
}

// x += s * d (FP64 refinement step)
template <typename T>
__global__ __launch_bounds__(kRB) void refine_kernel(int64_t n, double s, const T *__restrict__ d,
                                                     double *__restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) x[i] += s * (double)d[i];
}

// inverse diagonal of the CSR rows (rows sorted by column index)
// 3x3 block structure of a CSR pattern (vector fields of three components
// numbered node by node): rows 3I, 3I+1, 3I+2 share one column list made of
// whole aligned triples 3J, 3J+1, 3J+2.  *bad = 1 where a row breaks it.
__global__ __launch_bounds__(kRB) void block3_check_kernel(int64_t n, const int32_t *__restrict__ ptr,
                                                           const int32_t *__restrict__ idx, int *__restrict__ bad) {
  for (int64_t r = (int64_t)blockIdx.x * kRB + threadIdx.x; r < n; r += (int64_t)gridDim.x * kRB) {
    const int64_t r0 = r - r % 3;
    const int32_t k0 = ptr[r] - 1, len = ptr[r + 1] - ptr[r];
    const int32_t h0 = ptr[r0] - 1, hlen = ptr[r0 + 1] - ptr[r0];
    bool ok = len % 3 == 0 && len == hlen;
    for (int32_t t = 0; ok && t < len; t++) {
      const int32_t c = idx[k0 + t] - 1;  // 0-based column
      ok = c % 3 == t % 3 && (t % 3 == 0 || idx[k0 + t - 1] == c) && (r == r0 || idx[h0 + t] - 1 == c);
    }
    if (!ok) *bad = 1;
  }
}

// BSR(3) copy of a CSR pattern that passed block3_check_kernel: the blocks
// of block row I are b = (ptr[3I]-1)/9 + m, m < (ptr[3I+1]-ptr[3I])/3, the
// m-th at block column (idx[ptr[3I]-1+3m]-1)/3; values by entry of the
// block, bval[(3i+j) nb + b] = A[3I+i][3J+j] (FP32, from the row-ordered
// values).  One thread per block row.
__global__ __launch_bounds__(kRB) void block3_build_kernel(int64_t nbr, int64_t nb, const int32_t *__restrict__ ptr,
                                                           const int32_t *__restrict__ idx,
                                                           const float *__restrict__ val, int32_t *__restrict__ bcol,
                                                           float *__restrict__ bval, int with_cols) {
  for (int64_t I = (int64_t)blockIdx.x * kRB + threadIdx.x; I < nbr; I += (int64_t)gridDim.x * kRB) {
    const int32_t k0 = ptr[3 * I] - 1, b0 = k0 / 9, len = (ptr[3 * I + 1] - ptr[3 * I]) / 3;
    for (int32_t m = 0; m < len; m++) {
      const int64_t b = b0 + m;
      if (with_cols) bcol[b] = (idx[k0 + 3 * m] - 1) / 3;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          const int e = 3 * i + j;
          bval[e < 8 ? 8 * b + e : 8 * nb + b] = val[ptr[3 * I + i] - 1 + 3 * m + j];
        }
    }
  }
}

// y = A x over the BSR(3) copy, FP32; S lanes per block row, two blocks per
// lane in flight, a block's first 8 values in two 16-byte loads
template <int S>
__global__ __launch_bounds__(kRB) void bsr3_spmv_kernel(int64_t nbr, int64_t nb, const int32_t *__restrict__ ptr,
                                                        const int32_t *__restrict__ bcol,
                                                        const float *__restrict__ bval, const float *__restrict__ x,
                                                        float *__restrict__ y) {
  const int64_t I = ((int64_t)blockIdx.x * kRB + threadIdx.x) / S;
  const int sl = threadIdx.x & (S - 1);
  float a0 = 0.f, a1 = 0.f, a2 = 0.f;
  if (I < nbr) {
    const int32_t b0 = (ptr[3 * I] - 1) / 9, b1 = b0 + (ptr[3 * I + 1] - ptr[3 * I]) / 3;
    for (int32_t b = b0 + sl; b < b1; b += 2 * S) {
      const int32_t bb[2] = {b, b + S};
      int32_t c[2];
      float v[2][9], xv[2][3];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const bool in = bb[u] < b1;
        c[u] = in ? 3 * bcol[bb[u]] : 0;
        const int64_t bk = in ? bb[u] : 0;  // past the row: block 0, values unused
        const float4 *q = reinterpret_cast<const float4 *>(bval + 8 * bk);
        const float4 lo = q[0], hi = q[1];
        v[u][0] = in ? lo.x : 0.f;
        v[u][1] = in ? lo.y : 0.f;
        v[u][2] = in ? lo.z : 0.f;
        v[u][3] = in ? lo.w : 0.f;
        v[u][4] = in ? hi.x : 0.f;
        v[u][5] = in ? hi.y : 0.f;
        v[u][6] = in ? hi.z : 0.f;
        v[u][7] = in ? hi.w : 0.f;
        v[u][8] = in ? bval[8 * nb + bk] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < 2; u++)
#pragma unroll
        for (int j = 0; j < 3; j++) xv[u][j] = x[c[u] + j];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        a0 = fmaf(v[u][0], xv[u][0], fmaf(v[u][1], xv[u][1], fmaf(v[u][2], xv[u][2], a0)));
        a1 = fmaf(v[u][3], xv[u][0], fmaf(v[u][4], xv[u][1], fmaf(v[u][5], xv[u][2], a1)));
        a2 = fmaf(v[u][6], xv[u][0], fmaf(v[u][7], xv[u][1], fmaf(v[u][8], xv[u][2], a2)));
      }
    }
  }
#pragma unroll
  for (int o = S / 2; o >= 1; o >>= 1) {
    a0 += __shfl_xor(a0, o, S);
    a1 += __shfl_xor(a1, o, S);
    a2 += __shfl_xor(a2, o, S);
  }
  if (I < nbr && sl == 0) {
    y[3 * I] = a0;
    y[3 * I + 1] = a1;
    y[3 * I + 2] = a2;
  }
}

// CSR view of a structurally symmetric CSC pattern: the rows of row i are
// the rows of column i, so CSR entry p (row i, column j) sits in column i at
// the position of j; perm[p] = the CSC position of entry (j, i) in column j.
// One thread per column j; every entry (i, j) finds (j, i) by a binary
// search in column i, else *bad = 1 (the pattern is not symmetric).
__global__ __launch_bounds__(kRB) void transpose_perm_kernel(int64_t n, const int32_t *__restrict__ ptr,
                                                             const int32_t *__restrict__ idx,
                                                             int32_t *__restrict__ perm, int *__restrict__ bad) {
  for (int64_t j = (int64_t)blockIdx.x * kRB + threadIdx.x; j < n; j += (int64_t)gridDim.x * kRB) {
    const int32_t want = (int32_t)(j + 1);
    for (int32_t k = ptr[j] - 1; k < ptr[j + 1] - 1; k++) {
      const int32_t i = idx[k] - 1;
      int32_t lo = ptr[i] - 1, hi = ptr[i + 1] - 1;
      const int32_t end = hi;
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (idx[mid] < want) lo = mid + 1;
        else hi = mid;
      }
      if (lo < end && idx[lo] == want) perm[lo] = k;
      else *bad = 1;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kRB) void diag_inv_kernel(int64_t n, const int32_t *__restrict__ ptr,
                                                       const int32_t *__restrict__ idx, const double *__restrict__ val,
                                                       T *__restrict__ dinv, int *__restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x;
  if (i >= n) return;
  int32_t lo = ptr[i] - 1, hi = ptr[i + 1] - 1;
  const int32_t want = (int32_t)(i + 1);
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (idx[mid] < want) lo = mid + 1;
    else hi = mid;
  }
  const double dv = (lo < ptr[i + 1] - 1 && idx[lo] == want) ? val[lo] : 0.0;
  if (!(dv > 0.0)) atomicOr(bad, 1);
  dinv[i] = dv > 0.0 ? (T)(1.0 / dv) : (T)0;
}

// CSR values from the CSC values: v64[k] = a[perm[k]], v32[k] = (float) v64[k]
__global__ __launch_bounds__(kRB) void gather_values_kernel(int64_t n, const int32_t *__restrict__ perm,
                                                            const double *__restrict__ a, double *__restrict__ v64,
                                                            float *__restrict__ v32) {
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) {
    const double v = a[perm[i]];
    v64[i] = v;
    if (v32) v32[i] = (float)v;
  }
}

// b[i] += sum_k a[perm[k]] * (-xd[col[k]]) over row i's entries in
// ascending column order, multiplications and additions unfused: the order
// and rounding of mul!(b, Ad, xd, -1, 1) on a CSC matrix
__global__ __launch_bounds__(kRB) void lift_kernel(int64_t n, const int64_t *__restrict__ rptr,
                                                   const int32_t *__restrict__ perm, const int32_t *__restrict__ col,
                                                   const double *__restrict__ a, double alpha,
                                                   const double *__restrict__ xd, double *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) {
    double acc = b[i];
    for (int64_t k = rptr[i]; k < rptr[i + 1]; k++)
      acc = __dadd_rn(acc, __dmul_rn(a[perm[k]], __dmul_rn(alpha, xd[col[k]])));
    b[i] = acc;
  }
}

}  // namespace

void launch_lift(int64_t nrows, const int64_t *rptr, const int32_t *perm, const int32_t *col, const double *a,
                 double alpha, const double *xd, double *b, hipStream_t st) {
  if (nrows == 0) return;
  const int64_t blocks = (nrows + kRB - 1) / kRB;
  lift_kernel<<<(unsigned)(blocks < 65536 ? blocks : 65536), kRB, 0, st>>>(nrows, rptr, perm, col, a, alpha, xd, b);
}

int solve_rows_per_lane_group(double mean_row) {
  return mean_row > 48 ? 16 : (mean_row > 12 ? 8 : 4);
}

void launch_spmv(const SolveMat &m, const void *val, bool val64, const void *x, bool x64, const void *b, void *y,
                 bool y64, hipStream_t st) {
  switch (m.lanes) {
    case 4: launch_spmv_s<4>(m, val, val64, x, x64, b, y, y64, st); break;
    case 8: launch_spmv_s<8>(m, val, val64, x, x64, b, y, y64, st); break;
    case 32: launch_spmv_s<32>(m, val, val64, x, x64, b, y, y64, st); break;
    case 64: launch_spmv_s<64>(m, val, val64, x, x64, b, y, y64, st); break;
    default: launch_spmv_s<16>(m, val, val64, x, x64, b, y, y64, st); break;
  }
}

void launch_dot(int64_t n, const void *a, const void *b, bool f64, double *partial, double *sc, int slot,
                hipStream_t st) {
  if (f64) dot_partial_kernel<double><<<kSolvePartials, kRB, 0, st>>>(n, (const double *)a, (const double *)b, partial);
  else dot_partial_kernel<float><<<kSolvePartials, kRB, 0, st>>>(n, (const float *)a, (const float *)b, partial);
  sum_partials_kernel<<<1, kRB, 0, st>>>(kSolvePartials, partial, sc, slot);
}

void launch_cg_start(int64_t n, double s, const double *r64, const CgVectors &v, double *partial, double *sc,
                     hipStream_t st) {
  if (v.f64)
    cg_start_kernel<double><<<kSolvePartials, kRB, 0, st>>>(n, s, r64, (double *)v.d, (double *)v.r, (double *)v.z,
                                                            (double *)v.p, (const double *)v.dinv, partial);
  else
    cg_start_kernel<float><<<kSolvePartials, kRB, 0, st>>>(n, s, r64, (float *)v.d, (float *)v.r, (float *)v.z,
                                                           (float *)v.p, (const float *)v.dinv, partial);
  sum_partials_kernel<<<1, kRB, 0, st>>>(kSolvePartials, partial, sc, 0);
}

void launch_cg_iteration(const SolveMat &m, const void *val, const CgVectors &v, double *partial, double *partial2,
                         double *sc, int it, bool read_rr, hipStream_t st) {
  const int64_t n = m.n;
  const int cur = (it & 1) ? 3 : 0;  // slot of the current r.z
  double *part_rz = partial2, *part_rr = partial2 + kSolvePartials;
  if (!v.f64 && m.bcol) {
    const int64_t nbr = m.n / 3;
    const unsigned g8 = (unsigned)((nbr * 8 + kRB - 1) / kRB), g16 = (unsigned)((nbr * 16 + kRB - 1) / kRB),
                   g32 = (unsigned)((nbr * 32 + kRB - 1) / kRB);
    const float *x = (const float *)v.p;
    float *y = (float *)v.q;
    if (m.block_lanes == 8) bsr3_spmv_kernel<8><<<g8, kRB, 0, st>>>(nbr, m.nb, m.ptr, m.bcol, m.bval, x, y);
    else if (m.block_lanes == 16) bsr3_spmv_kernel<16><<<g16, kRB, 0, st>>>(nbr, m.nb, m.ptr, m.bcol, m.bval, x, y);
    else bsr3_spmv_kernel<32><<<g32, kRB, 0, st>>>(nbr, m.nb, m.ptr, m.bcol, m.bval, x, y);
  } else {
    launch_spmv(m, val, v.f64, v.p, v.f64, nullptr, v.q, v.f64, st);
  }
  if (v.f64) dot_partial_kernel<double><<<kSolvePartials, kRB, 0, st>>>(n, (const double *)v.p, (const double *)v.q, partial);
  else dot_partial_kernel<float><<<kSolvePartials, kRB, 0, st>>>(n, (const float *)v.p, (const float *)v.q, partial);
  if (v.f64)
    cg_update_kernel<double><<<kSolvePartials, kRB, 0, st>>>(n, sc, cur, (double *)v.d, (double *)v.r, (double *)v.z,
                                                             (const double *)v.p, (const double *)v.q,
                                                             (const double *)v.dinv, partial, part_rz, part_rr);
  else
    cg_update_kernel<float><<<kSolvePartials, kRB, 0, st>>>(n, sc, cur, (float *)v.d, (float *)v.r, (float *)v.z,
                                                            (const float *)v.p, (const float *)v.q,
                                                            (const float *)v.dinv, partial, part_rz, part_rr);
  if (v.f64) cg_direction_kernel<double><<<kSolvePartials, kRB, 0, st>>>(n, sc, cur, part_rz, (const double *)v.z, (double *)v.p);
  else cg_direction_kernel<float><<<kSolvePartials, kRB, 0, st>>>(n, sc, cur, part_rz, (const float *)v.z, (float *)v.p);
  if (read_rr) sum_partials_kernel<<<1, kRB, 0, st>>>(kSolvePartials, part_rr, sc, 5);
}

void launch_refine(int64_t n, double s, const void *d, bool f64, double *x, hipStream_t st) {
  if (f64) refine_kernel<double><<<kSolvePartials, kRB, 0, st>>>(n, s, (const double *)d, x);
  else refine_kernel<float><<<kSolvePartials, kRB, 0, st>>>(n, s, (const float *)d, x);
}

void launch_diag_inv(const SolveMat &m, const double *val, void *dinv, bool f64, int *bad, hipStream_t st) {
  const dim3 grid((unsigned)((m.n + kRB - 1) / kRB));
  if (f64) diag_inv_kernel<double><<<grid, kRB, 0, st>>>(m.n, m.ptr, m.idx, val, (double *)dinv, bad);
  else diag_inv_kernel<float><<<grid, kRB, 0, st>>>(m.n, m.ptr, m.idx, val, (float *)dinv, bad);
}

void launch_block3_check(const SolveMat &m, int *bad, hipStream_t st) {
  if (m.n == 0) return;
  const int64_t blocks = (m.n + kRB - 1) / kRB;
  block3_check_kernel<<<(unsigned)(blocks < 65536 ? blocks : 65536), kRB, 0, st>>>(m.n, m.ptr, m.idx, bad);
}

void launch_block3_build(const SolveMat &m, const float *val_rows, int32_t *bcol, float *bval, bool with_cols,
                         hipStream_t st) {
  const int64_t nbr = m.n / 3;
  if (nbr == 0) return;
  const int64_t blocks = (nbr + kRB - 1) / kRB;
  block3_build_kernel<<<(unsigned)(blocks < 65536 ? blocks : 65536), kRB, 0, st>>>(nbr, m.nb, m.ptr, m.idx, val_rows,
                                                                                   bcol, bval, with_cols ? 1 : 0);
}

void launch_transpose_perm(const SolveMat &m, int32_t *perm, int *bad, hipStream_t st) {
  if (m.n == 0) return;
  const int64_t blocks = (m.n + kRB - 1) / kRB;
  transpose_perm_kernel<<<(unsigned)(blocks < 65536 ? blocks : 65536), kRB, 0, st>>>(m.n, m.ptr, m.idx, perm, bad);
}

void launch_gather_values(int64_t n, const int32_t *perm, const double *a, double *v64, float *v32, hipStream_t st) {
  gather_values_kernel<<<kSolvePartials, kRB, 0, st>>>(n, perm, a, v64, v32);
}

}  // namespace gtk
