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
  __device__ static type all_none() { return 0xFFFFFFFFu; }
  __device__ static int get(type v, int r) { return (v >> (8 * r)) & 0xFF; }
};
template <>
struct Off4<uint16_t> {
  using type = uint2;  // four 16-bit offsets
  static constexpr uint32_t none = 0xFFFF;
  __device__ static type all_none() { return make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu); }
  __device__ static int get(type v, int r) { return r < 2 ? (v.x >> (16 * r)) & 0xFFFF : (v.y >> (16 * (r - 2))) & 0xFFFF; }
};

// N consecutive CT values from LDS with the widest loads the alignment
// `ALIGN` (bytes) allows.
template <typename CT, int N, int ALIGN>
__device__ inline void lds_load(const CT *__restrict__ p, CT (&out)[N]) {
  constexpr int BYTES = N * (int)sizeof(CT);
  if constexpr (ALIGN % 16 == 0 && BYTES % 16 == 0) {
    using V = __attribute__((ext_vector_type(16 / sizeof(CT)))) CT;
    constexpr int L = 16 / sizeof(CT);
#pragma unroll
    for (int i = 0; i < N / L; i++) {
      V v = reinterpret_cast<const V *>(p)[i];
#pragma unroll
      for (int j = 0; j < L; j++) out[i * L + j] = v[j];
    }
  } else if constexpr (ALIGN % 8 == 0 && BYTES % 8 == 0 && sizeof(CT) <= 8) {
    using V = __attribute__((ext_vector_type(8 / sizeof(CT)))) CT;
    constexpr int L = 8 / sizeof(CT);
    if constexpr (L == 1) {
#pragma unroll
      for (int i = 0; i < N; i++) out[i] = p[i];
    } else {
#pragma unroll
      for (int i = 0; i < N / L; i++) {
        V v = reinterpret_cast<const V *>(p)[i];
#pragma unroll
        for (int j = 0; j < L; j++) out[i * L + j] = v[j];
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < N; i++) out[i] = p[i];
  }
}

// Per-batch record (layout documented with BatchPlan in gtk_plan.h).
template <typename OffT, int MB>
struct BatchRegs {
  uint32_t info;   // nslots | tmax << 8 | ncols << 16 (wave-uniform)
  uint32_t slot;   // this lane's slot (lane & 15)
  int32_t seg0;    // column (lane & 15) of the batch
  uint32_t la;
  typename Off4<OffT>::type off[MB];  // rows 16 mb + 4 g + r of the slot
};

template <typename OffT, int MB>
__device__ inline void load_batch(const unsigned char *__restrict__ rec, int rec_bytes, int off_stride, int64_t bi,
                                  int nl, int g, BatchRegs<OffT, MB> &r) {
  const unsigned char *R = rec + bi * rec_bytes;
  r.info = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(R));
  r.slot = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(R + 16) + nl);
  r.seg0 = __builtin_nontemporal_load(reinterpret_cast<const int32_t *>(R + 80) + nl);
  r.la = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(R + 144) + nl);
  const unsigned char *O = R + 208 + nl * off_stride * (int)sizeof(OffT);
#pragma unroll
  for (int mb = 0; mb < MB; mb++) {
    const int a0 = mb * 16 + 4 * g;
    r.off[mb] = a0 < off_stride ? *reinterpret_cast<const typename Off4<OffT>::type *>(O + a0 * (int)sizeof(OffT))
                                : Off4<OffT>::all_none();
  }
}

// Scalar forms on the matrix cores, one workgroup per tile, NW wavefronts.
//
// Phase A (all waves): node coordinates of the tile's cells are staged in
// LDS (in chunks, in the space the accumulators use later); each thread
// keeps one quadrature point q with its geometry gradients in registers and
// computes, per cell, J = sum_n x_n (x) grad N_n(q), det J, J^-1 in FP64
// and the factor G = kappa |det J| w_q J^-1 J^-T (mass: kappa |det J| w_q),
// stored in LDS in the contraction precision CT.
//
// Phase B (each wave independently): batches of the tile, strided over the
// waves.  A batch packs whole columns whose contributors (cell slot s,
// local column b) total at most 16; for its 16 slots one product gives the
// local columns
//   A[a, n] = sum_k Bt[a, k] W[k, n],  k = (q, d),
//   Bt[a, (q,d)] = d_d phi_a(q)                    (reference table, registers)
//   W[(q,d), n]  = (G_{s_n,q} grad phi_{b_n}(q))_d  (built per batch)
// with the k dimension ordered so that lane group g = lane/16 covers the
// quadrature points g*QG .. g*QG+QG-1 with all their directions: each lane
// reads G and grad phi_b once per point with vector LDS loads.  The 4 rows
// per 16-row block a lane holds are added into FP64 column accumulators in
// LDS through the record's row offsets, stepping over the contributor rank
// t so that every nonzero is summed in contributor order (cells ascending)
// and no two lanes of one step touch one address.  Finished segments are
// written once to HBM and the accumulators cleared.  The record of the next
// batch is loaded while the current one is computed.
template <int D, int NV, int NS, int NQ, int FORM, typename CT, typename OffT, int NW>
__global__ __launch_bounds__(NW * 64) void matrix_batch_kernel(MatrixArgs args) {
  constexpr int kB = NW * 64;
  constexpr int NG = NGeo<D, FORM>::value;
  constexpr int NGP = NG == 3 ? 4 : NG;  // staged factors per (cell, point)
  constexpr int KD = FORM == kLaplace ? D : 1;
  constexpr int TP = FORM == kLaplace ? (D == 3 ? 4 : 2) : 1;  // table entries per (q, b)
  constexpr int QG = (NQ + 3) / 4;
  constexpr int KS = QG * KD;
  constexpr int MB = (NS + 15) / 16;
  constexpr int CPR = kB / NQ;  // cells per geometry round
  using V4 = typename Acc4<CT>::type;
  using O4 = Off4<OffT>;
  static_assert(CPR >= 1, "block too small for the quadrature");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  CT *tab_s = reinterpret_cast<CT *>(lds_raw);                     // [NQ][NS][TP]
  CT *G_s = reinterpret_cast<CT *>(lds_raw + args.lds_geo);        // [cells][NQ][NGP]
  double *acc_all = reinterpret_cast<double *>(lds_raw + args.acc_offset);
  double *X_s = acc_all;  // phase A: [stage cells][NV][D]

  const int tid = threadIdx.x;
  const int tile = blockIdx.x;
  const int cell0 = args.tile_cell_ptr[tile];
  const int ncell = args.tile_cell_ptr[tile + 1] - cell0;

  for (int i = tid; i < NQ * NS * TP; i += kB) {
    const int q = i / (NS * TP), rem = i - q * NS * TP, b = rem / TP, d = rem - b * TP;
    double v;
    if constexpr (FORM == kLaplace) v = d < D ? args.gphi[(q * D + d) * NS + b] : 0.0;
    else v = args.phi[q * NS + b];
    tab_s[i] = (CT)v;
  }

  // phase A: geometry
  {
    const int tq = tid % NQ, tcs = tid / NQ;
    const bool gl = tcs < CPR;
    double gq[NV][D];
#pragma unroll
    for (int n = 0; n < NV; n++)
#pragma unroll
      for (int d = 0; d < D; d++) gq[n][d] = gl ? args.gN[(tq * NV + n) * D + d] : 0.0;
    const double wq = gl ? args.w[tq] : 0.0;
    for (int c0 = 0; c0 < ncell; c0 += args.stage_cells) {
      const int nch = min(args.stage_cells, ncell - c0);
      for (int i = tid; i < nch * NV; i += kB) {
        const int sl = i / NV, n = i - sl * NV;
        const int64_t cell = args.tile_cells[cell0 + c0 + sl];
        const int64_t node = args.cell_nodes[cell * NV + n];
#pragma unroll
        for (int d = 0; d < D; d++) X_s[i * D + d] = args.xyz[node * D + d];
      }
      __syncthreads();
      if (gl) {
        for (int sl = tcs; sl < nch; sl += CPR) {
          const double *X = X_s + sl * NV * D;
          double J[D][D];
#pragma unroll
          for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) J[i][j] = 0.0;
#pragma unroll
          for (int n = 0; n < NV; n++)
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
              for (int j = 0; j < D; j++) J[i][j] = fma(X[n * D + i], gq[n][j], J[i][j]);
          double Ji[D][D];
          const double det = inverse<D>(J, Ji);
          const int64_t cell = args.tile_cells[cell0 + c0 + sl];
          const double kdV = (args.coef ? args.coef[cell] : 1.0) * fabs(det) * wq;
          CT *G = G_s + ((c0 + sl) * NQ + tq) * NGP;
          if constexpr (FORM == kLaplace) {
            CT gv[NGP];
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
              for (int j = i; j < D; j++) {
                double gg = 0.0;
#pragma unroll
                for (int m = 0; m < D; m++) gg = fma(Ji[i][m], Ji[j][m], gg);
                gv[sym_index<D>(i, j)] = (CT)(kdV * gg);
              }
            if constexpr (NGP > NG) gv[NGP - 1] = (CT)0;
#pragma unroll
            for (int i = 0; i < NGP; i++) G[i] = gv[i];
          } else {
            G[0] = (CT)kdV;
          }
        }
      }
      __syncthreads();
    }
  }

  // phase B: batches
  const int wave = tid >> 6, lane = tid & 63;
  const int nl = lane & 15, g = lane >> 4;
  double *acc = acc_all + (size_t)wave * args.acc_per_wave;
  for (int i = lane; i < args.acc_per_wave; i += 64) acc[i] = 0.0;
  CT afrag[MB][KS];
#pragma unroll
  for (int mb = 0; mb < MB; mb++)
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const int a = mb * 16 + a_row<CT>(nl), q = g * QG + ks / KD, d = ks % KD;
      afrag[mb][ks] = (a < NS && q < NQ) ? tab_s[(q * NS + a) * TP + d] : (CT)0;
    }
  const int bt0 = args.tile_batch_ptr[tile], bt1 = args.tile_batch_ptr[tile + 1];
  int bi = bt0 + wave;
  BatchRegs<OffT, MB> nxt;
  if (bi < bt1) load_batch<OffT, MB>(args.rec, args.rec_bytes, args.off_stride, bi, nl, g, nxt);
  for (; bi < bt1; bi += NW) {
    BatchRegs<OffT, MB> cur = nxt;
    cur.info = __builtin_amdgcn_readfirstlane(cur.info);
    if (bi + NW < bt1) load_batch<OffT, MB>(args.rec, args.rec_bytes, args.off_stride, bi + NW, nl, g, nxt);
    const int nslots = cur.info & 0xFF, tmax = (cur.info >> 8) & 0xFF, ncols = cur.info >> 16;
    const bool valid = nl < nslots;
    const int s = cur.slot & 0xFFF, b = (cur.slot >> 12) & 0xFF, t_n = (cur.slot >> 20) & 0xF;
    const int ci = cur.slot >> 24;
    const int my_acc = (int)(__shfl((int)cur.la, ci, 16) >> 16);
    const CT *Gc = G_s + s * NQ * NGP;
    const CT *Tb = tab_s + b * TP;
    V4 c[MB];
#pragma unroll
    for (int mb = 0; mb < MB; mb++) c[mb] = V4{0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < QG; j++) {
      const int q = g * QG + j;
      const int qq = q < NQ ? q : NQ - 1;
      CT G[NGP], T[TP];
      lds_load<CT, NGP, (NGP * sizeof(CT)) % 16 == 0 ? 16 : 8>(Gc + qq * NGP, G);
      lds_load<CT, TP, TP * sizeof(CT) >= 16 ? 16 : (TP * sizeof(CT) >= 8 ? 8 : 4)>(Tb + qq * NS * TP, T);
      const bool on = valid && q < NQ;
      CT W[KD];
      if constexpr (FORM == kLaplace) {
        if constexpr (D == 3) {
          W[0] = G[0] * T[0] + G[1] * T[1] + G[2] * T[2];
          W[1] = G[1] * T[0] + G[3] * T[1] + G[4] * T[2];
          W[2] = G[2] * T[0] + G[4] * T[1] + G[5] * T[2];
        } else {
          W[0] = G[0] * T[0] + G[1] * T[1];
          W[1] = G[1] * T[0] + G[2] * T[1];
        }
      } else {
        W[0] = G[0] * T[0];
      }
#pragma unroll
      for (int d = 0; d < KD; d++) {
        const CT wv = on ? W[d] : (CT)0;
#pragma unroll
        for (int mb = 0; mb < MB; mb++) c[mb] = mfma16x16x4(afrag[mb][j * KD + d], wv, c[mb]);
      }
    }
    // destinations: rows outside the row set or padding go to a private
    // per-lane slot past the accumulators
    int dst[MB][4];
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int off = valid ? O4::get(cur.off[mb], r) : (int)O4::none;
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
    // write the finished segments and clear the accumulators: lane group g
    // takes columns g, g+4, ...
    for (int i = g; i < ncols; i += 4) {
      const uint32_t lai = (uint32_t)__shfl((int)cur.la, i, 16);
      const int64_t s0 = __shfl(cur.seg0, i, 16);
      const int len = lai & 0xFFFF, ab = lai >> 16;
      for (int r = nl; r < len; r += 16) {
        args.nzval[s0 + r] = acc[ab + r];
        acc[ab + r] = 0.0;
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

template <typename E, int FORM, typename CT, typename OffT, int NW>
static KernelInfo matrix_batch_info() {
  constexpr int NG = NGeo<E::d, FORM>::value;
  constexpr int NGP = NG == 3 ? 4 : NG;
  constexpr int TP = FORM == kLaplace ? (E::d == 3 ? 4 : 2) : 1;
  KernelInfo k;
  k.fn = reinterpret_cast<const void *>(&matrix_batch_kernel<E::d, E::nv, E::ns, E::nq, FORM, CT, OffT, NW>);
  k.table_bytes = (size_t)E::nq * E::ns * TP * sizeof(CT);
  k.geo_bytes_per_cell = (size_t)E::nq * NGP * sizeof(CT);
  k.cols_per_wave = 16;
  k.mfma = true;
  k.block = NW * 64;
  return k;
}

template <typename E, int FORM, typename CT>
static KernelInfo matrix_batch_info(int off_bytes, int waves) {
  if (waves == 4)
    return off_bytes == 1 ? matrix_batch_info<E, FORM, CT, uint8_t, 4>() : matrix_batch_info<E, FORM, CT, uint16_t, 4>();
  return off_bytes == 1 ? matrix_batch_info<E, FORM, CT, uint8_t, 8>() : matrix_batch_info<E, FORM, CT, uint16_t, 8>();
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
static bool try_matrix(const ElemKey &key, int form, int precision, int off_bytes, bool allow_mfma, int waves,
                       KernelInfo &out) {
  if (!match<E>(key)) return false;
  if constexpr (E::nc == 1) {
    if (form != kLaplace && form != kMass) return false;
    if (allow_mfma) {
      if (form == kLaplace)
        out = precision == 0 ? matrix_batch_info<E, kLaplace, double>(off_bytes, waves)
                             : matrix_batch_info<E, kLaplace, float>(off_bytes, waves);
      else
        out = precision == 0 ? matrix_batch_info<E, kMass, double>(off_bytes, waves)
                             : matrix_batch_info<E, kMass, float>(off_bytes, waves);
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

bool find_matrix_kernel(const ElemKey &key, int form, int precision, int off_bytes, bool allow_mfma, int waves,
                        KernelInfo &out) {
  return try_matrix<QuadQ1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadQ2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadQ3>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<HexQ1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<HexQ2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetP2v>(key, form, precision, off_bytes, allow_mfma, waves, out);
}

bool find_vector_kernel(const ElemKey &key, int fkind, KernelInfo &out) {
  return try_vector<QuadQ1>(key, fkind, out) || try_vector<QuadQ2>(key, fkind, out) ||
         try_vector<QuadQ3>(key, fkind, out) || try_vector<HexQ1>(key, fkind, out) ||
         try_vector<HexQ2>(key, fkind, out) || try_vector<TetP2v>(key, fkind, out);
}

}  // namespace gtk
