// gtk_kernels.hip -- gfx950 kernels of libgtk_hip.
//
// One workgroup per tile of the column-owner plan (gtk_plan.h).  A tile owns
// a set of output columns (matrix) or rows (vector) and lists every cell that
// touches them.
//   phase A: per (cell, quadrature point) the geometry map of the reference
//            loop (J = sum_n x_n (x) grad N_n, det J, J^-1; accessors.jl:
//            941-968, :1000-1007) in FP64, folded with the coefficient and the
//            weight into the symmetric factor G = kappa |det J| w J^-1 J^-T
//            staged in LDS;
//   phase B: each owned column j is accumulated in LDS (FP64) from its
//            contributors (cell, local column b) in ascending cell order: the
//            local column A_c[:, b] = sum_q grad(phi_a)^T G_q grad(phi_b) is
//            scattered to the column segment through precomputed 16-bit row
//            offsets, then the segment is written once, coalesced.
// No atomics: every nonzero is produced by one workgroup in a fixed order.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gtk_kernels.h"

namespace gtk {

template <int D>
__device__ inline double inverse(const double J[D][D], double Ji[D][D]) {
  if constexpr (D == 2) {
    double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    double r = 1.0 / det;
    Ji[0][0] = J[1][1] * r;
    Ji[0][1] = -J[0][1] * r;
    Ji[1][0] = -J[1][0] * r;
    Ji[1][1] = J[0][0] * r;
    return det;
  } else {
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    double r = 1.0 / det;
    Ji[0][0] = c00 * r;
    Ji[1][0] = c01 * r;
    Ji[2][0] = c02 * r;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
    return det;
  }
}

// J at quadrature point q of `cell`; returns |det J| and J^-1.
template <int D, int NV>
__device__ inline double cell_jacobian(const double *__restrict__ xyz, const int32_t *__restrict__ cell_nodes,
                                       const double *__restrict__ gN, int64_t cell, int q, double Ji[D][D]) {
  double J[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) J[i][j] = 0.0;
#pragma unroll
  for (int n = 0; n < NV; n++) {
    const int64_t node = cell_nodes[cell * NV + n];
    double x[D], g[D];
#pragma unroll
    for (int i = 0; i < D; i++) x[i] = xyz[node * D + i];
#pragma unroll
    for (int j = 0; j < D; j++) g[j] = gN[(q * NV + n) * D + j];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) J[i][j] = fma(x[i], g[j], J[i][j]);
  }
  return fabs(inverse<D>(J, Ji));
}

// number of staged geometry factors per (cell, point)
template <int D, int FORM>
struct NGeo {
  static constexpr int value = FORM == kLaplace ? D * (D + 1) / 2 : (FORM == kMass ? 1 : D * D + 3);
};

template <int D>
__device__ inline int sym_index(int i, int j) {
  // packed symmetric storage: (0,0) (0,1) (0,2) (1,1) (1,2) (2,2) / (0,0) (0,1) (1,1)
  if (i > j) {
    int t = i;
    i = j;
    j = t;
  }
  if constexpr (D == 2) return i == 0 ? j : 2;
  else return i == 0 ? j : (i == 1 ? 2 + j : 5);
}

template <int D, int NV, int NS, int NQ, int NC, int FORM, typename CT>
__global__ __launch_bounds__(kBlock) void matrix_tile_kernel(MatrixArgs args) {
  constexpr int NG = NGeo<D, FORM>::value;
  constexpr int NDOF = NS * NC;
  constexpr int CPW = 64 / NDOF;
  constexpr int NWAVES = kBlock / 64;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  CT *gphi_s = reinterpret_cast<CT *>(lds_raw);   // [NQ][D][NS]
  CT *phi_s = gphi_s + NQ * D * NS;                // [NQ][NS]
  CT *G_s = phi_s + NQ * NS;                       // [cells][NQ][NG]
  double *acc_base = reinterpret_cast<double *>(lds_raw + args.acc_offset);

  const int tile = blockIdx.x;
  const int cell0 = args.tile_cell_ptr[tile];
  const int ncell = args.tile_cell_ptr[tile + 1] - cell0;
  const int col0 = args.tile_col_ptr[tile];
  const int ncol = args.tile_col_ptr[tile + 1] - col0;

  for (int i = threadIdx.x; i < NQ * D * NS; i += kBlock) gphi_s[i] = (CT)args.gphi[i];
  for (int i = threadIdx.x; i < NQ * NS; i += kBlock) phi_s[i] = (CT)args.phi[i];

  // phase A: FP64 geometry per (cell, point)
  for (int idx = threadIdx.x; idx < ncell * NQ; idx += kBlock) {
    const int s = idx / NQ, q = idx - s * NQ;
    const int64_t cell = args.tile_cells[cell0 + s];
    double Ji[D][D];
    const double dV = cell_jacobian<D, NV>(args.xyz, args.cell_nodes, args.gN, cell, q, Ji) * args.w[q];
    CT *G = G_s + (s * NQ + q) * NG;
    if constexpr (FORM == kLaplace) {
      const double k = args.coef ? args.coef[cell] : 1.0;
      const double kdV = k * dV;
#pragma unroll
      for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = i; j < D; j++) {
          double g = 0.0;
#pragma unroll
          for (int m = 0; m < D; m++) g = fma(Ji[i][m], Ji[j][m], g);
          G[sym_index<D>(i, j)] = (CT)(kdV * g);
        }
    } else if constexpr (FORM == kMass) {
      const double k = args.coef ? args.coef[cell] : 1.0;
      G[0] = (CT)(k * dV);
    } else {
      // elasticity: J^-1 (row-major), dV, lambda, mu
#pragma unroll
      for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) G[i * D + j] = (CT)Ji[i][j];
      G[D * D] = (CT)dV;
      G[D * D + 1] = (CT)args.coef[2 * cell];
      G[D * D + 2] = (CT)args.coef[2 * cell + 1];
    }
  }
  __syncthreads();

  // phase B: owned columns
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int cs = lane / NDOF, a = lane - cs * NDOF;
  const bool active = cs < CPW;
  double *acc = acc_base + (size_t)(wave * CPW + (active ? cs : 0)) * args.max_seg;
  const int an = a / NC, ac = a - an * NC;
  for (int base = wave * CPW; base < ncol; base += NWAVES * CPW) {
    const int ci = base + cs;
    if (!active || ci >= ncol) continue;
    const int pc = col0 + ci;
    const int64_t j = args.col_id[pc];
    const int64_t seg0 = (int64_t)args.colptr[j] - 1;
    const int len = args.colptr[j + 1] - args.colptr[j];
    for (int r = a; r < len; r += NDOF) acc[r] = 0.0;
    const int k1 = args.col_ctr_ptr[pc + 1];
    for (int k = args.col_ctr_ptr[pc]; k < k1; k++) {
      const int s = args.ctr_slot[k], b = args.ctr_b[k];
      const int bn = b / NC, bc = b - bn * NC;
      const CT *Gs = G_s + s * NQ * NG;
      CT v = 0;
      if constexpr (FORM == kLaplace) {
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          const CT *G = Gs + q * NG;
          CT gb[D], ga[D];
#pragma unroll
          for (int d = 0; d < D; d++) {
            gb[d] = gphi_s[(q * D + d) * NS + bn];
            ga[d] = gphi_s[(q * D + d) * NS + an];
          }
#pragma unroll
          for (int d = 0; d < D; d++) {
            CT wd = 0;
#pragma unroll
            for (int e = 0; e < D; e++) wd = fma(G[sym_index<D>(d, e)], gb[e], wd);
            v = fma(ga[d], wd, v);
          }
        }
      } else if constexpr (FORM == kMass) {
#pragma unroll
        for (int q = 0; q < NQ; q++) v = fma(phi_s[q * NS + an], Gs[q * NG] * phi_s[q * NS + bn], v);
      } else {
        // sigma(eps(phi_b e_bc)) : eps(phi_a e_ac)
        //   = lambda (g_b)_bc (g_a)_ac + mu (delta_{ac,bc} g_a.g_b + (g_a)_bc (g_b)_ac)
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          const CT *G = Gs + q * NG;
          CT ga[D], gb[D];
#pragma unroll
          for (int i = 0; i < D; i++) {
            CT sa = 0, sb = 0;
#pragma unroll
            for (int m = 0; m < D; m++) {
              sa = fma(G[m * D + i], gphi_s[(q * D + m) * NS + an], sa);
              sb = fma(G[m * D + i], gphi_s[(q * D + m) * NS + bn], sb);
            }
            ga[i] = sa;
            gb[i] = sb;
          }
          CT dot = 0;
#pragma unroll
          for (int i = 0; i < D; i++) dot = fma(ga[i], gb[i], dot);
          const CT lam = G[D * D + 1], mu = G[D * D + 2];
          CT t = lam * gb[bc] * ga[ac] + mu * ((ac == bc ? dot : (CT)0) + ga[bc] * gb[ac]);
          v = fma(t, G[D * D], v);
        }
      }
      const int off = args.off_bytes == 1 ? (int)args.ctr_off[(int64_t)k * args.off_stride + a]
                                          : (int)reinterpret_cast<const uint16_t *>(args.ctr_off)[(int64_t)k * args.off_stride + a];
      if (off != (args.off_bytes == 1 ? 0xFF : 0xFFFF)) acc[off] += (double)v;
    }
    for (int r = a; r < len; r += NDOF) args.nzval[seg0 + r] = acc[r];
  }
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename CT>
struct Acc4;
template <>
struct Acc4<double> {
  using type = v4d;
};
template <>
struct Acc4<float> {
  using type = v4f;
};

// D(16x16) += A(16x4) B(4x16); lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Result rows held by lane l, r = 0..3 (measured on gfx950): FP32 holds
// D[4*(l/16) + r][l%16], FP64 holds D[l/16 + 4*r][l%16].
__device__ inline v4d mfma16x16x4(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline v4f mfma16x16x4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of the A operand that lane i supplies so that, for both precisions,
// lane l ends up holding rows 4*(l/16) + r of the product: FP64 rows are
// permuted by i -> 4*(i%4) + i/4, FP32 rows are taken in order.
template <typename CT>
__device__ constexpr int a_row(int i) {
  return sizeof(CT) == 8 ? 4 * (i & 3) + (i >> 2) : i;
}

template <typename OffT>
struct Off4;
template <>
struct Off4<uint8_t> {
  using type = uint32_t;  // four 8-bit offsets
  static constexpr uint32_t none = 0xFF;
  __device__ static int get(type v, int r) { return (v >> (8 * r)) & 0xFF; }
};
template <>
struct Off4<uint16_t> {
  using type = uint2;  // four 16-bit offsets
  static constexpr uint32_t none = 0xFFFF;
  __device__ static int get(type v, int r) { return r < 2 ? (v.x >> (16 * r)) & 0xFFFF : (v.y >> (16 * (r - 2))) & 0xFFFF; }
};

// Scalar forms on the matrix cores.  A batch packs whole columns of the tile
// whose contributors total at most 16; for its 16 contributor slots the local
// columns are one product
//   A[a, n] = sum_k Bt[a, k] W[k, n],  k = (q, d),
//   Bt[a, (q,d)] = d_d phi_a(q)                     (reference table, registers)
//   W[(q,d), n]  = (G_{s_n,q} grad phi_{b_n}(q))_d   (built per batch)
// (mass form: k = q, Bt = phi_a(q), W = G phi_b(q)).  Batches own disjoint
// columns, so any wave can take any batch.  Each lane holds 4 rows per
// 16-row block of one slot and adds them into the FP64 column accumulators in
// LDS; the adds step over the contributor rank t of the slot within its
// column, so no two lanes of one step touch one column and every nonzero is
// summed in contributor order (cells ascending).  Finished segments are
// written once to HBM.  The tile's plan metadata (column segments,
// contributors, row offsets, batches) is staged in LDS by the prologue,
// overlapping the geometry phase, so the batch loop reads no global memory.
struct ColInfo {
  int seg0, len, kb, kc;  // segment start (0-based), length, first contributor (tile-relative), count
};

template <int D, int NV, int NS, int NQ, int FORM, typename CT, typename OffT>
__global__ __launch_bounds__(kMfmaBlock) void matrix_tile_mfma_kernel(MatrixArgs args) {
  constexpr int kBlock = kMfmaBlock;
  constexpr int NG = NGeo<D, FORM>::value;
  constexpr int KD = FORM == kLaplace ? D : 1;
  constexpr int K = NQ * KD;
  constexpr int KS = (K + 3) / 4;
  constexpr int MB = (NS + 15) / 16;
  constexpr int NWAVES = kBlock / 64;
  constexpr int NT = NQ * KD * NS;  // reference table entries
  using V4 = typename Acc4<CT>::type;
  using O4 = Off4<OffT>;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  CT *tab_s = reinterpret_cast<CT *>(lds_raw);  // Laplace: [NQ][D][NS] gradients, mass: [NQ][NS] values
  CT *G_s = tab_s + NT;                          // [cells][NQ][NG]
  ColInfo *col_s = reinterpret_cast<ColInfo *>(lds_raw + args.lds_col);
  uint32_t *ctr_s = reinterpret_cast<uint32_t *>(lds_raw + args.lds_ctr);  // slot | b << 16
  unsigned char *off_s = lds_raw + args.lds_off;                            // row offsets
  int *bat_s = reinterpret_cast<int *>(lds_raw + args.lds_bat);            // batch -> first column
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nl = lane & 15, g = lane >> 4;
  double *acc = reinterpret_cast<double *>(lds_raw + args.acc_offset) + (size_t)wave * args.acc_per_wave;

  const int tile = blockIdx.x;
  const int cell0 = args.tile_cell_ptr[tile];
  const int ncell = args.tile_cell_ptr[tile + 1] - cell0;
  const int col0 = args.tile_col_ptr[tile];
  const int ncol = args.tile_col_ptr[tile + 1] - col0;
  const int ctr0 = args.col_ctr_ptr[col0];
  const int nctr = args.col_ctr_ptr[col0 + ncol] - ctr0;
  const int b0 = args.tile_batch_ptr[tile];
  const int nbat = args.tile_batch_ptr[tile + 1] - b0;

  // prologue: tables, plan metadata, accumulators
  for (int i = threadIdx.x; i < NT; i += kBlock) tab_s[i] = (CT)(FORM == kLaplace ? args.gphi[i] : args.phi[i]);
  for (int i = threadIdx.x; i < ncol; i += kBlock) {
    const int pc = col0 + i;
    const int j = args.col_id[pc];
    const int c0 = args.colptr[j];
    const int k0 = args.col_ctr_ptr[pc];
    col_s[i] = ColInfo{c0 - 1, args.colptr[j + 1] - c0, k0 - ctr0, args.col_ctr_ptr[pc + 1] - k0};
  }
  for (int i = threadIdx.x; i < nctr; i += kBlock)
    ctr_s[i] = (uint32_t)args.ctr_slot[ctr0 + i] | ((uint32_t)args.ctr_b[ctr0 + i] << 16);
  {
    const int row_bytes = args.off_stride * (int)sizeof(OffT);  // multiple of 16
    const uint4 *src = reinterpret_cast<const uint4 *>(args.ctr_off + (int64_t)ctr0 * row_bytes);
    uint4 *dst = reinterpret_cast<uint4 *>(off_s);
    const int n16 = nctr * row_bytes / 16;
    for (int i = threadIdx.x; i < n16; i += kBlock) dst[i] = src[i];
  }
  for (int i = threadIdx.x; i <= nbat; i += kBlock) bat_s[i] = args.batch_col[b0 + i] - col0;
  for (int i = lane; i < args.acc_per_wave; i += 64) acc[i] = 0.0;

  // phase A: FP64 geometry per (cell, point)
  for (int idx = threadIdx.x; idx < ncell * NQ; idx += kBlock) {
    const int s = idx / NQ, q = idx - s * NQ;
    const int64_t cell = args.tile_cells[cell0 + s];
    double Ji[D][D];
    const double dV = cell_jacobian<D, NV>(args.xyz, args.cell_nodes, args.gN, cell, q, Ji) * args.w[q];
    const double kdV = (args.coef ? args.coef[cell] : 1.0) * dV;
    CT *G = G_s + (s * NQ + q) * NG;
    if constexpr (FORM == kLaplace) {
#pragma unroll
      for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = i; j < D; j++) {
          double gg = 0.0;
#pragma unroll
          for (int m = 0; m < D; m++) gg = fma(Ji[i][m], Ji[j][m], gg);
          G[sym_index<D>(i, j)] = (CT)(kdV * gg);
        }
    } else {
      G[0] = (CT)kdV;
    }
  }
  __syncthreads();

  CT afrag[MB][KS];
#pragma unroll
  for (int mb = 0; mb < MB; mb++)
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const int a = mb * 16 + a_row<CT>(nl), k = 4 * ks + g;
      afrag[mb][ks] = (a < NS && k < K) ? tab_s[k * NS + a] : (CT)0;  // [q][d][a] or [q][a]
    }
  const int ostride4 = args.off_stride >> 2;  // 4-offset words per contributor
  const typename O4::type *offw = reinterpret_cast<const typename O4::type *>(off_s);
  for (int bi = wave; bi < nbat; bi += NWAVES) {
    const int cb = bat_s[bi];
    const int m = bat_s[bi + 1] - cb;  // columns in the batch (<= 16)
    const bool colv = nl < m;
    const ColInfo ci_info = col_s[cb + (colv ? nl : 0)];
    const int len = colv ? ci_info.len : 0;
    const int kc = colv ? ci_info.kc : 0;
    // exclusive prefixes of contributor counts and segment lengths
    int slot0 = kc, abase = len;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const int t1 = __shfl_up(slot0, o, 16), t2 = __shfl_up(abase, o, 16);
      if (nl >= o) {
        slot0 += t1;
        abase += t2;
      }
    }
    const int nslots = __shfl(slot0, 15, 16);
    slot0 -= kc;
    abase -= len;
    int tmax = kc;
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) tmax = max(tmax, __shfl_xor(tmax, o, 16));
    // slot nl: its column ci and rank t_n within the column
    int ci = 0;
#pragma unroll
    for (int i = 1; i < 16; i++) {
      const int si = __shfl(slot0, i, 16);
      if (i < m && si <= nl) ci = i;
    }
    const bool valid = nl < nslots;
    const int t_n = nl - __shfl(slot0, ci, 16);
    const int k = __shfl(ci_info.kb, ci, 16) + t_n;  // tile-relative contributor
    const int my_acc = __shfl(abase, ci, 16);
    int s = 0, b = 0;
    typename O4::type o[MB];
    if (valid) {
      const uint32_t sb = ctr_s[k];
      s = sb & 0xFFFF;
      b = sb >> 16;
#pragma unroll
      for (int mb = 0; mb < MB; mb++) o[mb] = offw[k * ostride4 + mb * 4 + g];
    }
    const CT *Gs = G_s + s * NQ * NG;
    V4 c[MB];
#pragma unroll
    for (int mb = 0; mb < MB; mb++) c[mb] = V4{0, 0, 0, 0};
    // W without branches: padded k and empty slots read in-bounds entries
    // and are zeroed by a select, so every LDS read can be issued early
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const int kk = 4 * ks + g;
      const int kq = kk < K ? kk : K - 1;
      CT w;
      if constexpr (FORM == kLaplace) {
        const int q = kq / D, d = kq - q * D;
        const CT *G = Gs + q * NG;
        w = G[sym_index<D>(d, 0)] * tab_s[(q * D) * NS + b];
#pragma unroll
        for (int e = 1; e < D; e++) w = fma(G[sym_index<D>(d, e)], tab_s[(q * D + e) * NS + b], w);
      } else {
        w = Gs[kq] * tab_s[kq * NS + b];
      }
      w = (valid && kk < K) ? w : (CT)0;
#pragma unroll
      for (int mb = 0; mb < MB; mb++) c[mb] = mfma16x16x4(afrag[mb][ks], w, c[mb]);
    }
    // destinations: rows outside the row set or padding go to a private
    // per-lane slot past the accumulators
    int dst[MB][4];
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int off = valid ? O4::get(o[mb], r) : (int)O4::none;
        dst[mb][r] = off != (int)O4::none ? my_acc + off : args.acc_per_wave - 64 + lane;
      }
    for (int t = 0; t < tmax; t++) {
      if (valid && t_n == t) {
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
          for (int r = 0; r < 4; r++) atomicAdd(&acc[dst[mb][r]], (double)c[mb][r]);
      }
    }
    // write the finished segments and clear the accumulators
    for (int i = 0; i < m; i++) {
      const int li = __shfl(len, i, 16), ai = __shfl(abase, i, 16), si = __shfl(ci_info.seg0, i, 16);
      for (int r = lane; r < li; r += 64) {
        args.nzval[(int64_t)si + r] = acc[ai + r];
        acc[ai + r] = 0.0;
      }
    }
  }
}

template <int D, int NV, int NS, int NQ, int NC, int FKIND>
__global__ __launch_bounds__(kBlock) void vector_tile_kernel(VectorArgs args) {
  constexpr int NDOF = NS * NC;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double *phi_s = reinterpret_cast<double *>(lds_raw);  // [NQ][NS]
  double *F_s = phi_s + NQ * NS;                        // [cells][NQ][NC] = f . dV

  const int tile = blockIdx.x;
  const int cell0 = args.tile_cell_ptr[tile];
  const int ncell = args.tile_cell_ptr[tile + 1] - cell0;
  const int col0 = args.tile_col_ptr[tile];
  const int ncol = args.tile_col_ptr[tile + 1] - col0;
  for (int i = threadIdx.x; i < NQ * NS; i += kBlock) phi_s[i] = args.phi[i];
  __syncthreads();
  for (int idx = threadIdx.x; idx < ncell * NQ; idx += kBlock) {
    const int s = idx / NQ, q = idx - s * NQ;
    const int64_t cell = args.tile_cells[cell0 + s];
    double Ji[D][D];
    const double dV = cell_jacobian<D, NV>(args.xyz, args.cell_nodes, args.gN, cell, q, Ji) * args.w[q];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      double f;
      if constexpr (FKIND == kFConst) {
        f = args.f[c];
      } else if constexpr (FKIND == kFCell) {
        f = args.f[cell * NC + c];
      } else {
        // discrete field: sum_a f[dof(cell, a)] phi_a(q)
        f = 0.0;
        for (int n = 0; n < NS; n++) {
          const int32_t dof = args.cell_dofs[cell * NDOF + n * NC + c];
          const double fv = dof > 0 ? args.f[dof - 1] : args.f[args.n_free - dof - 1];
          f = fma(fv, phi_s[q * NS + n], f);
        }
      }
      F_s[(s * NQ + q) * NC + c] = f * dV;
    }
  }
  __syncthreads();
  for (int ci = threadIdx.x; ci < ncol; ci += kBlock) {
    const int pc = col0 + ci;
    const int64_t j = args.col_id[pc];
    double acc = 0.0;
    const int k1 = args.col_ctr_ptr[pc + 1];
    for (int k = args.col_ctr_ptr[pc]; k < k1; k++) {
      const int s = args.ctr_slot[k], b = args.ctr_b[k];
      const int bn = b / NC, bc = b - bn * NC;
      double v = 0.0;
#pragma unroll 4
      for (int q = 0; q < NQ; q++) v = fma(phi_s[q * NS + bn], F_s[(s * NQ + q) * NC + bc], v);
      acc += v;
    }
    args.out[j] = acc;
  }
}

// ---------------------------------------------------------------------------
// dispatch tables
// ---------------------------------------------------------------------------
template <int D, int NV, int NS, int NQ, int NC>
struct Elem {
  static constexpr int d = D, nv = NV, ns = NS, nq = NQ, nc = NC;
};

template <typename E, int FORM, typename CT, typename OffT>
static KernelInfo matrix_mfma_info() {
  KernelInfo k;
  k.fn = reinterpret_cast<const void *>(&matrix_tile_mfma_kernel<E::d, E::nv, E::ns, E::nq, FORM, CT, OffT>);
  k.table_bytes = (size_t)(E::nq * (FORM == kLaplace ? E::d : 1) * E::ns) * sizeof(CT);
  k.geo_bytes_per_cell = (size_t)E::nq * NGeo<E::d, FORM>::value * sizeof(CT);
  k.cols_per_wave = 16;
  k.mfma = true;
  k.block = kMfmaBlock;
  return k;
}

template <typename E, int FORM, typename CT>
static KernelInfo matrix_mfma_info(int off_bytes) {
  return off_bytes == 1 ? matrix_mfma_info<E, FORM, CT, uint8_t>() : matrix_mfma_info<E, FORM, CT, uint16_t>();
}

template <typename E, int FORM, typename CT>
static KernelInfo matrix_info() {
  KernelInfo k;
  k.fn = reinterpret_cast<const void *>(&matrix_tile_kernel<E::d, E::nv, E::ns, E::nq, E::nc, FORM, CT>);
  k.table_bytes = (size_t)(E::nq * E::d * E::ns + E::nq * E::ns) * sizeof(CT);
  k.geo_bytes_per_cell = (size_t)E::nq * NGeo<E::d, FORM>::value * sizeof(CT);
  k.cols_per_wave = 64 / (E::ns * E::nc);
  return k;
}

template <typename E, int FK>
static KernelInfo vector_info() {
  KernelInfo k;
  k.fn = reinterpret_cast<const void *>(&vector_tile_kernel<E::d, E::nv, E::ns, E::nq, E::nc, FK>);
  k.table_bytes = (size_t)E::nq * E::ns * sizeof(double);
  k.geo_bytes_per_cell = (size_t)E::nq * E::nc * sizeof(double);
  k.cols_per_wave = 64;
  return k;
}

template <typename E>
static bool match(const ElemKey &key) {
  return key.dim == E::d && key.nv == E::nv && key.ns == E::ns && key.nq == E::nq && key.nc == E::nc;
}

template <typename E>
static bool try_matrix(const ElemKey &key, int form, int precision, int off_bytes, bool allow_mfma, KernelInfo &out) {
  if (!match<E>(key)) return false;
  if constexpr (E::nc == 1) {
    if (form != kLaplace && form != kMass) return false;
    if (allow_mfma) {
      if (form == kLaplace)
        out = precision == 0 ? matrix_mfma_info<E, kLaplace, double>(off_bytes)
                             : matrix_mfma_info<E, kLaplace, float>(off_bytes);
      else
        out = precision == 0 ? matrix_mfma_info<E, kMass, double>(off_bytes) : matrix_mfma_info<E, kMass, float>(off_bytes);
    } else {
      if (form == kLaplace)
        out = precision == 0 ? matrix_info<E, kLaplace, double>() : matrix_info<E, kLaplace, float>();
      else
        out = precision == 0 ? matrix_info<E, kMass, double>() : matrix_info<E, kMass, float>();
    }
  } else {
    if (form != kElasticity || E::nc != E::d) return false;
    out = precision == 0 ? matrix_info<E, kElasticity, double>() : matrix_info<E, kElasticity, float>();
  }
  return true;
}

template <typename E>
static bool try_vector(const ElemKey &key, int fkind, KernelInfo &out) {
  if (!match<E>(key)) return false;
  if (fkind == kFConst) out = vector_info<E, kFConst>();
  else if (fkind == kFCell) out = vector_info<E, kFCell>();
  else out = vector_info<E, kFNodal>();
  return true;
}

// quad Q1..Q3 with (k+1)^2 Gauss points, hex Q1/Q2 with (k+1)^3, P2 vector tets
// with the 14-point rule, P1 triangles/tets.
using QuadQ1 = Elem<2, 4, 4, 4, 1>;
using QuadQ2 = Elem<2, 4, 9, 9, 1>;
using QuadQ3 = Elem<2, 4, 16, 16, 1>;
using HexQ1 = Elem<3, 8, 8, 8, 1>;
using HexQ2 = Elem<3, 8, 27, 27, 1>;
using TetP2v = Elem<3, 4, 10, 14, 3>;

bool find_matrix_kernel(const ElemKey &key, int form, int precision, int off_bytes, bool allow_mfma,
                        KernelInfo &out) {
  return try_matrix<QuadQ1>(key, form, precision, off_bytes, allow_mfma, out) ||
         try_matrix<QuadQ2>(key, form, precision, off_bytes, allow_mfma, out) ||
         try_matrix<QuadQ3>(key, form, precision, off_bytes, allow_mfma, out) ||
         try_matrix<HexQ1>(key, form, precision, off_bytes, allow_mfma, out) ||
         try_matrix<HexQ2>(key, form, precision, off_bytes, allow_mfma, out) ||
         try_matrix<TetP2v>(key, form, precision, off_bytes, allow_mfma, out);
}

bool find_vector_kernel(const ElemKey &key, int fkind, KernelInfo &out) {
  return try_vector<QuadQ1>(key, fkind, out) || try_vector<QuadQ2>(key, fkind, out) ||
         try_vector<QuadQ3>(key, fkind, out) || try_vector<HexQ1>(key, fkind, out) ||
         try_vector<HexQ2>(key, fkind, out) || try_vector<TetP2v>(key, fkind, out);
}

}  // namespace gtk
