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
#include <type_traits>

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

template <int D>
__device__ inline double determinant(const double J[D][D]) {
  if constexpr (D == 2) {
    return J[0][0] * J[1][1] - J[0][1] * J[1][0];
  } else {
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    return J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  }
}

// Adjugate A = det(J) J^-1 (the transposed cofactors, the same products as
// inverse<D>); returns det J.
template <int D>
__device__ inline double adjugate(const double J[D][D], double A[D][D]) {
  if constexpr (D == 2) {
    A[0][0] = J[1][1];
    A[0][1] = -J[0][1];
    A[1][0] = -J[1][0];
    A[1][1] = J[0][0];
    return J[0][0] * J[1][1] - J[0][1] * J[1][0];
  } else {
    A[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    A[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    A[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    A[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    A[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    A[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    A[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    A[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    A[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    return J[0][0] * A[0][0] + J[0][1] * A[1][0] + J[0][2] * A[2][0];
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

// J at quadrature point q of `cell` (geometry gradients gN[q][NV][D]).
template <int D, int NV>
__device__ inline void cell_jacobian_only(const double *__restrict__ xyz, const int32_t *__restrict__ cell_nodes,
                                          const double *__restrict__ gN, int64_t cell, int q, double J[D][D]) {
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) J[i][j] = 0.0;
#pragma unroll
  for (int n = 0; n < NV; n++) {
    const int64_t node = cell_nodes[cell * NV + n];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) J[i][j] = fma(xyz[node * D + i], gN[(q * NV + n) * D + j], J[i][j]);
  }
}

// Measure factor of a boundary face at a point: sqrt(det(Jf^T Jf)) with
// Jf = J A, A = d(reference cell point)/d(reference face point), D x (D-1)
// row-major (the weight sqrt(det(J'J)) of the reference,
// src/accessors.jl:1000-1007, for the face's map).
template <int D>
__device__ inline double surface_measure(const double J[D][D], const double *__restrict__ A) {
  if constexpr (D == 2) {
    const double t0 = J[0][0] * A[0] + J[0][1] * A[1];
    const double t1 = J[1][0] * A[0] + J[1][1] * A[1];
    return sqrt(t0 * t0 + t1 * t1);
  } else {
    double u[3], v[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      u[i] = J[i][0] * A[0] + J[i][1] * A[2] + J[i][2] * A[4];
      v[i] = J[i][0] * A[1] + J[i][1] * A[3] + J[i][2] * A[5];
    }
    const double c0 = u[1] * v[2] - u[2] * v[1], c1 = u[2] * v[0] - u[0] * v[2], c2 = u[0] * v[1] - u[1] * v[0];
    return sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  }
}

// number of staged geometry factors per (cell, point)
template <int D, int FORM>
struct NGeo {
  static constexpr int value =
      FORM == kLaplace ? D * (D + 1) / 2
                       : (FORM == kMass ? 1
                                        : (FORM == kNitsche || FORM == kNitscheLoad
                                               ? D + 1
                                               : (FORM == kSkeleton ? 2 * D + 1 : D * D + 3)));
};

// lanes per column of the VALU matrix kernel: one per local row (skeleton
// items: the rows of both cells); for elasticity one per node (NC = D rows
// each)
template <int NS, int NC, int FORM>
struct ValuLanes {
  static constexpr int value = FORM == kElasticity ? NS : (FORM == kSkeleton ? 2 * NS * NC : NS * NC);
};

// elasticity on affine cells (order-1 simplices, J the same at every
// point): one staged factor set per cell instead of one per point
template <int D, int NV, int FORM, int NL>
struct CellGeo {
  static constexpr bool value = FORM == kElasticity && NV == D + 1 && NL == 0;
};

// values per staged factor set of the VALU kernel (CellGeo: J^-1, |det J|,
// lambda, mu and the symmetric M = J^-1 J^-T)
template <int D, int FORM, int NL, int NV>
struct CellGeoStride {
  static constexpr int value =
      NGeo<D, FORM>::value + (CellGeo<D, NV, FORM, NL>::value ? D * (D + 1) / 2 : 0);
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

// NL > 0: boundary items (face mode: mass, Nitsche): the tables are per
// local face, [NL][NQ][...], and the measure is the face's.  Nitsche items
// list all DOFs of their cell (normal derivatives involve every shape
// function); per (item, point) the staged factors are beta dG and dG J^-1 n,
// with dG the face measure times the weight, beta the item's coefficient
// and n = J^-T n_ref / |J^-T n_ref| the physical unit normal
// (src/accessors.jl:1009-1030), so that n . grad phi = (J^-1 n) . grad_ref phi.
// Skeleton items (kSkeleton): an interior face with its two cells; NV is
// both cells' nodes (NV / 2 each, side 0 first), the item's DOFs are both
// cells' (NS each), and the NL tables are point sets: item_face packs the
// set of side 0 (low 16 bits) and of side 1, each the face rule's points in
// that cell's reference coordinates, in the same physical order.  Per
// (item, point): beta dG and dG/2 J_i^-1 n_i per side i, n_i the unit normal
// out of side i's cell; with s = n_i . n_j = +-1 for rows of side i and
// columns of side j, the entry is
//   s (beta phi_a phi_b - phi_a (m_j . gref_b) - (m_i . gref_a) phi_b)
// (jump(v) = v_i n_i, mean(grad u) = grad u_j / 2, m_i = J_i^-1 n_i / 2).
// BS: threads per workgroup (ValuBlock<CT>, or 256 where the accumulators of
// that many waves would leave no LDS for a tile)
template <int D, int NV, int NS, int NQ, int NC, int FORM, typename CT, int NL = 0, int BS = ValuBlock<CT>::value>
__global__ __launch_bounds__(BS) void matrix_tile_kernel(MatrixArgs args) {
  constexpr int kVB = BS;
  constexpr bool kNit = FORM == kNitsche || FORM == kNitscheLoad;
  constexpr bool kSkel = FORM == kSkeleton;
  static_assert(NL == 0 || FORM == kMass || kNit || kSkel, "boundary items: mass and Nitsche forms");
  static_assert(!(kNit || kSkel) || (NL > 0 && NC == 1), "Nitsche and skeleton terms: scalar fields on face items");
  constexpr int NVC = kSkel ? NV / 2 : NV;  // nodes of one cell
  constexpr int NG = NGeo<D, FORM>::value;
  constexpr int NDOF = (kSkel ? 2 : 1) * NS * NC;
  // lanes per column: one per local row, or (elasticity) one per node, that
  // lane forming the node's NC rows from one pair of physical gradients
  constexpr int LPC = ValuLanes<NS, NC, FORM>::value;
  constexpr int RPL = NDOF / LPC;  // rows per lane
  constexpr int CPW = 64 / LPC;
  constexpr int NWAVES = kVB / 64;
  constexpr int NT = NL > 0 ? NL : 1;            // table sets
  // affine cells (simplices with order-1 geometry) have one J at every
  // point, so elasticity needs no point loop: with P = J^-T R_ab J^-1 and
  // R_ab = sum_q w_q grad phi_a(q) grad phi_b(q)^T (reference integrals per
  // node pair, in LDS), the point sums are
  //   sum_q w_q (g_a)_i (g_b)_j = P_ij,  sum_q w_q g_a . g_b = tr P
  // and one factor set per cell is staged: J^-1, |det J|, lambda, mu and
  // M = J^-1 J^-T (tr P = sum R_ab o M)
  constexpr bool kCellGeo = CellGeo<D, NV, FORM, NL>::value;
  constexpr int GQ = kCellGeo ? 1 : NQ;  // factor sets per cell
  constexpr int NGS = CellGeoStride<D, FORM, NL, NV>::value;  // values per factor set
  // gradient table ([NT][NQ][D][NS]; none for the face mass), or R
  constexpr int NGPHI = NL > 0 ? (kNit || kSkel ? NT * NQ * D * NS : 0) : (kCellGeo ? NS * NS * D * D : NQ * D * NS);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  CT *gphi_s = reinterpret_cast<CT *>(lds_raw);   // [NT][NQ][D][NS], or R [NS][NS][D][D]
  CT *phi_s = gphi_s + NGPHI;                      // [NT][NQ][NS]
  CT *G_s = phi_s + NT * NQ * NS;                  // [cells][GQ][NGS]
  double *acc_base = reinterpret_cast<double *>(lds_raw + args.acc_offset);

  const int tile = blockIdx.x;
  const int cell0 = args.tile_cell_ptr[tile];
  const int ncell = args.tile_cell_ptr[tile + 1] - cell0;
  const int col0 = args.tile_col_ptr[tile];
  const int ncol = args.tile_col_ptr[tile + 1] - col0;
  int32_t *lf_s = reinterpret_cast<int32_t *>(G_s + ncell * GQ * NGS);  // [cells] local face (NL > 0)

  if constexpr (kCellGeo) {
    // R[a][b][m][n] = sum_q w_q d_m phi_a(q) d_n phi_b(q), formed in FP64
    // once per reference table set (gtk_set_reference)
    for (int i = threadIdx.x; i < NGPHI; i += kVB) gphi_s[i] = (CT)args.rtab[i];
  } else {
    for (int i = threadIdx.x; i < NGPHI; i += kVB) gphi_s[i] = (CT)args.gphi[i];
  }
  for (int i = threadIdx.x; i < NT * NQ * NS; i += kVB) phi_s[i] = (CT)args.phi[i];
  if constexpr (NL > 0)
    for (int i = threadIdx.x; i < ncell; i += kVB) lf_s[i] = args.item_face[args.tile_cells[cell0 + i]];

  // phase A: FP64 geometry per (cell, point), or per cell (kCellGeo)
  for (int idx = threadIdx.x; idx < ncell * GQ; idx += kVB) {
    const int s = idx / GQ, q = idx - s * GQ;
    const int64_t cell = args.tile_cells[cell0 + s];
    double Ji[D][D];
    double dV;
    if constexpr (kSkel) {
      // both sides' maps at the point (the same physical point); the face
      // measure from side 0's
      const int sets = args.item_face[cell];
      CT *G = G_s + (s * GQ + q) * NGS;
      G[0] = 0;
#pragma unroll
      for (int side = 0; side < 2; side++) {
        const int l = side ? sets >> 16 : sets & 0xFFFF;
        double J[D][D];
        cell_jacobian_only<D, NVC>(args.xyz, args.cell_nodes, args.gN + (size_t)l * NQ * NV * D, 2 * cell + side, q,
                                   J);
        if (side == 0) {
          dV = surface_measure<D>(J, args.face_jac + l * D * (D - 1)) * args.w[q];
          G[0] = (CT)((args.coef ? args.coef[cell] : 1.0) * dV);
        }
        (void)inverse<D>(J, Ji);
        const double *nr = args.face_normal + l * D;
        double v[D], vv = 0.0;
#pragma unroll
        for (int i = 0; i < D; i++) {
          double t = 0.0;
#pragma unroll
          for (int j = 0; j < D; j++) t = fma(Ji[j][i], nr[j], t);
          v[i] = t;
          vv = fma(t, t, vv);
        }
        const double inv = vv > 0.0 ? 0.5 / sqrt(vv) : 0.0;
#pragma unroll
        for (int i = 0; i < D; i++) {
          double m = 0.0;
#pragma unroll
          for (int j = 0; j < D; j++) m = fma(Ji[i][j], v[j] * inv, m);
          G[1 + side * D + i] = (CT)(dV * m);
        }
      }
      continue;
    } else if constexpr (NL > 0) {
      const int l = args.item_face[cell];
      double J[D][D];
      cell_jacobian_only<D, NV>(args.xyz, args.cell_nodes, args.gN + (size_t)l * NQ * NV * D, cell, q, J);
      dV = surface_measure<D>(J, args.face_jac + l * D * (D - 1)) * args.w[q];
      if constexpr (kNit) {
        // v = J^-T n_ref, n = v / |v|, m = J^-1 n
        (void)inverse<D>(J, Ji);
        const double *nr = args.face_normal + l * D;
        double v[D], vv = 0.0;
#pragma unroll
        for (int i = 0; i < D; i++) {
          double t = 0.0;
#pragma unroll
          for (int j = 0; j < D; j++) t = fma(Ji[j][i], nr[j], t);
          v[i] = t;
          vv = fma(t, t, vv);
        }
        const double inv = vv > 0.0 ? 1.0 / sqrt(vv) : 0.0;
        CT *G = G_s + (s * GQ + q) * NGS;
        G[0] = (CT)((args.coef ? args.coef[cell] : 1.0) * dV);
#pragma unroll
        for (int i = 0; i < D; i++) {
          double m = 0.0;
#pragma unroll
          for (int j = 0; j < D; j++) m = fma(Ji[i][j], v[j] * inv, m);
          G[1 + i] = (CT)(dV * m);
        }
        continue;
      }
    } else {
      dV = cell_jacobian<D, NV>(args.xyz, args.cell_nodes, args.gN, cell, q, Ji);
      if constexpr (!kCellGeo) dV *= args.w[q];
    }
    CT *G = G_s + (s * GQ + q) * NGS;
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
      if constexpr (kCellGeo) {
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
          for (int j = i; j < D; j++) {
            double m = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) m = fma(Ji[i][k], Ji[j][k], m);
            G[D * D + 3 + sym_index<D>(i, j)] = (CT)m;
          }
      }
    }
  }
  __syncthreads();

  // phase B: owned columns
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int cs = lane / LPC, a = lane - cs * LPC;  // a: local row, or node (RPL > 1)
  const bool active = cs < CPW;
  double *acc = acc_base + (size_t)(wave * CPW + (active ? cs : 0)) * args.max_seg;
  const int an = RPL > 1 ? a : a / NC;
  const int ac = RPL > 1 ? 0 : a - an * NC;  // component of row a (one row per lane)
  for (int base = wave * CPW; base < ncol; base += NWAVES * CPW) {
    const int ci = base + cs;
    if (!active || ci >= ncol) continue;
    const int pc = col0 + ci;
    const int64_t j = args.col_id[pc];
    const int64_t seg0 = (int64_t)args.colptr[j] - 1;
    const int len = args.colptr[j + 1] - args.colptr[j];
    for (int r = a; r < len; r += LPC) acc[r] = 0.0;
    const int k1 = args.col_ctr_ptr[pc + 1];
    // contributor k+1's slot, local column and row offsets are loaded while
    // contributor k is computed
    const int none = args.off_bytes == 1 ? 0xFF : 0xFFFF;
    auto fetch = [&](int k, int &s_, int &b_, int (&o_)[RPL]) {
      s_ = args.ctr_slot[k];
      b_ = args.ctr_b[k];
#pragma unroll
      for (int r = 0; r < RPL; r++) {
        const int64_t e = (int64_t)k * args.off_stride + (RPL > 1 ? an * NC + r : a);
        o_[r] = args.off_bytes == 1 ? (int)args.ctr_off[e] : (int)reinterpret_cast<const uint16_t *>(args.ctr_off)[e];
      }
    };
    if constexpr (kCellGeo) {
      // CPS contributors per step: their cells' factors and node pairs'
      // reference integrals are all read before any is used, then the
      // contributions are formed and added in contributor order (P2 tets
      // 48^3 x 6: mixed 1.773 -> 1.668 ms, FP64 2.497 -> 2.407 ms, the values
      // bit-identical; 3 or 4 per step measured slower)
      constexpr int CPS = 2;
      const int kb = args.col_ctr_ptr[pc];
      int sP[CPS], bP[CPS], oP[CPS][RPL];
#pragma unroll
      for (int i = 0; i < CPS; i++) {
        sP[i] = bP[i] = 0;
        if (kb + i < k1) fetch(kb + i, sP[i], bP[i], oP[i]);
      }
      auto form = [&](const CT (&Gc)[NGS], const CT (&R)[D * D], const CT (&jb)[D], int b, CT (&v)[RPL]) __attribute__((always_inline)) {
        const int bc = b - (b / NC) * NC;
        CT u[D], w[D];
#pragma unroll
        for (int m = 0; m < D; m++) {
          CT su = 0, sw = 0;
#pragma unroll
          for (int n = 0; n < D; n++) {
            su = fma(R[m * D + n], jb[n], su);
            sw = fma(R[n * D + m], jb[n], sw);
          }
          u[m] = su;
          w[m] = sw;
        }
        CT tr = 0;
#pragma unroll
        for (int m = 0; m < D; m++)
#pragma unroll
          for (int n = 0; n < D; n++) tr = fma(R[m * D + n], Gc[D * D + 3 + sym_index<D>(m, n)], tr);
        const CT det = Gc[D * D], lam = Gc[D * D + 1], mu = Gc[D * D + 2];
#pragma unroll
        for (int r = 0; r < RPL; r++) {
          CT pcol = 0, prow = 0;
#pragma unroll
          for (int m = 0; m < D; m++) {
            pcol = fma(Gc[m * D + r], u[m], pcol);
            prow = fma(Gc[m * D + r], w[m], prow);
          }
          v[r] = det * (lam * pcol + mu * ((r == bc ? tr : (CT)0) + prow));
        }
      };
      for (int k = kb; k < k1; k += CPS) {
        int sc[CPS], bcol[CPS], oc[CPS][RPL];
#pragma unroll
        for (int i = 0; i < CPS; i++) {
          const bool has = k + i < k1;
          sc[i] = has ? sP[i] : sP[0];
          bcol[i] = has ? bP[i] : bP[0];
#pragma unroll
          for (int r = 0; r < RPL; r++) oc[i][r] = has ? oP[i][r] : (int)none;
        }
#pragma unroll
        for (int i = 0; i < CPS; i++)
          if (k + CPS + i < k1) fetch(k + CPS + i, sP[i], bP[i], oP[i]);
        CT Gv[CPS][NGS], Rv[CPS][D * D], jv[CPS][D];
#pragma unroll
        for (int i = 0; i < CPS; i++) {
#pragma unroll
          for (int t = 0; t < NGS; t++) Gv[i][t] = G_s[sc[i] * NGS + t];
#pragma unroll
          for (int t = 0; t < D * D; t++) Rv[i][t] = gphi_s[(an * NS + bcol[i] / NC) * D * D + t];
#pragma unroll
          for (int m = 0; m < D; m++) jv[i][m] = G_s[sc[i] * NGS + m * D + (bcol[i] - (bcol[i] / NC) * NC)];
        }
        CT vv[CPS][RPL];
#pragma unroll
        for (int i = 0; i < CPS; i++) form(Gv[i], Rv[i], jv[i], bcol[i], vv[i]);
#pragma unroll
        for (int i = 0; i < CPS; i++)
#pragma unroll
          for (int r = 0; r < RPL; r++)
            if (oc[i][r] != none) atomicAdd(&acc[oc[i][r]], (double)vv[i][r]);
      }
      for (int r = a; r < len; r += LPC) args.nzval[seg0 + r] = acc[r];
      continue;
    }
    int s_n = 0, b_n = 0, o_n[RPL];
    if (args.col_ctr_ptr[pc] < k1) fetch(args.col_ctr_ptr[pc], s_n, b_n, o_n);
    for (int k = args.col_ctr_ptr[pc]; k < k1; k++) {
      const int s = s_n, b = b_n;
      int off[RPL];
#pragma unroll
      for (int r = 0; r < RPL; r++) off[r] = o_n[r];
      if (k + 1 < k1) fetch(k + 1, s_n, b_n, o_n);
      const int bn = b / NC, bc = b - bn * NC;
      const CT *Gs = G_s + s * GQ * NGS;
      CT v[RPL];  // rows an * NC + r of the lane (RPL > 1), or row a
#pragma unroll
      for (int r = 0; r < RPL; r++) v[r] = 0;
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
            v[0] = fma(ga[d], wd, v[0]);
          }
        }
      } else if constexpr (FORM == kMass) {
        const CT *ph = phi_s + (NL > 0 ? lf_s[s] * NQ * NS : 0);
#pragma unroll
        for (int q = 0; q < NQ; q++) v[0] = fma(ph[q * NS + an], Gs[q * NG] * ph[q * NS + bn], v[0]);
      } else if constexpr (kSkel) {
        const int sets = lf_s[s];
        const int sa = an >= NS, sb = bn >= NS;  // sides of row a and column b
        const int la = sa ? sets >> 16 : sets & 0xFFFF, lb = sb ? sets >> 16 : sets & 0xFFFF;
        const int ia = an - sa * NS, ib = bn - sb * NS;
        const CT *pha = phi_s + la * NQ * NS, *phb = phi_s + lb * NQ * NS;
        const CT *gpa = gphi_s + la * NQ * D * NS, *gpb = gphi_s + lb * NQ * D * NS;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          const CT *G = Gs + q * NGS;
          CT ta = 0, tb = 0;
#pragma unroll
          for (int d = 0; d < D; d++) {
            ta = fma(G[1 + sa * D + d], gpa[(q * D + d) * NS + ia], ta);
            tb = fma(G[1 + sb * D + d], gpb[(q * D + d) * NS + ib], tb);
          }
          const CT pa = pha[q * NS + ia], pb = phb[q * NS + ib];
          v[0] += G[0] * pa * pb - pa * tb - ta * pb;
        }
        if (sa != sb) v[0] = -v[0];
      } else if constexpr (kNit) {
        // row a (test v = phi_a), column b (trial u = phi_b):
        //   beta dG phi_a phi_b - phi_a (m . gref_b) - (m . gref_a) phi_b
        // (the load's matrix drops the middle term)
        const int lf = lf_s[s];
        const CT *ph = phi_s + lf * NQ * NS;
        const CT *gp = gphi_s + lf * NQ * D * NS;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          const CT *G = Gs + q * NGS;
          CT ta = 0, tb = 0;
#pragma unroll
          for (int d = 0; d < D; d++) {
            ta = fma(G[1 + d], gp[(q * D + d) * NS + an], ta);
            tb = fma(G[1 + d], gp[(q * D + d) * NS + bn], tb);
          }
          const CT pa = ph[q * NS + an], pb = ph[q * NS + bn];
          CT t = G[0] * pa * pb - ta * pb;
          if constexpr (FORM == kNitsche) t -= pa * tb;
          v[0] += t;
        }
      } else {
        // sigma(eps(phi_b e_bc)) : eps(phi_a e_ac)
        //   = lambda (g_b)_bc (g_a)_ac + mu (delta_{ac,bc} g_a.g_b + (g_a)_bc (g_b)_ac)
        if constexpr (kCellGeo) {
          CT Gc[NGS], R[D * D];  // the cell's factors; R_ab of this lane's node and the column's node
#pragma unroll
          for (int i = 0; i < NGS; i++) Gc[i] = Gs[i];
          const CT *Rab = gphi_s + (an * NS + bn) * D * D;
#pragma unroll
          for (int i = 0; i < D * D; i++) R[i] = Rab[i];
          CT jb[D];  // column bc of J^-1
#pragma unroll
          for (int m = 0; m < D; m++) jb[m] = bc == 0 ? Gc[m * D] : (bc == 1 ? Gc[m * D + 1] : Gc[m * D + D - 1]);
          CT u[D], w[D];  // R jb and R^T jb
#pragma unroll
          for (int m = 0; m < D; m++) {
            CT su = 0, sw = 0;
#pragma unroll
            for (int n = 0; n < D; n++) {
              su = fma(R[m * D + n], jb[n], su);
              sw = fma(R[n * D + m], jb[n], sw);
            }
            u[m] = su;
            w[m] = sw;
          }
          CT tr = 0;  // sum R o M, M symmetric
#pragma unroll
          for (int m = 0; m < D; m++)
#pragma unroll
            for (int n = 0; n < D; n++) tr = fma(R[m * D + n], Gc[D * D + 3 + sym_index<D>(m, n)], tr);
          const CT det = Gc[D * D], lam = Gc[D * D + 1], mu = Gc[D * D + 2];
#pragma unroll
          for (int r = 0; r < RPL; r++) {
            CT pcol = 0, prow = 0;  // P[r][bc] and P[bc][r]
#pragma unroll
            for (int m = 0; m < D; m++) {
              pcol = fma(Gc[m * D + r], u[m], pcol);
              prow = fma(Gc[m * D + r], w[m], prow);
            }
            v[r] = det * (lam * pcol + mu * ((r == bc ? tr : (CT)0) + prow));
          }
        } else {
#pragma unroll
          for (int q = 0; q < NQ; q++) {
            const CT *G = Gs + q * NGS;
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
            const CT dV = G[D * D];
#pragma unroll
            for (int ac = 0; ac < RPL; ac++) {
              CT t = lam * gb[bc] * ga[ac] + mu * ((ac == bc ? dot : (CT)0) + ga[bc] * gb[ac]);
              v[ac] = fma(t, dV, v[ac]);
            }
          }
        }
      }
      // vector fields: the scalar forms couple equal components only
      // (u . v, grad u : grad v); the pattern holds every component pair
      if constexpr (FORM != kElasticity && NC > 1)
        if (ac != bc) v[0] = 0;
      // the lanes of a column add to distinct rows, so an LDS add with no
      // return (in program order per lane) replaces the read-modify-write
#pragma unroll
      for (int r = 0; r < RPL; r++)
        if (off[r] != none) atomicAdd(&acc[off[r]], (double)v[r]);
    }
    for (int r = a; r < len; r += LPC) args.nzval[seg0 + r] = acc[r];
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

constexpr int kBatchSlotsK = 16;  // contributor slots per matrix-core batch (MFMA N)
// Per-cell stride (in CT elements) of the staged geometry factors: the
// NQP x NGP block rounded up to an odd number of the units the batch loop
// reads (8 bytes for FP32 rows, 16 bytes for FP64 rows), so that the 16
// slots of a batch, which read 16 different cells, spread over the LDS banks.
// EXTRA: entries kept after the block (the split product's per-cell scale).
template <int NQP, int NGP, typename CT, int EXTRA = 0>
struct GeoStride {
  static constexpr int unit = (NGP * (int)sizeof(CT)) % 16 == 0 ? 16 : ((NGP * (int)sizeof(CT)) % 8 == 0 ? 8 : 4);
  static constexpr int units = ((NQP * NGP + EXTRA) * (int)sizeof(CT) + unit - 1) / unit;
  static constexpr int value = ((units | 1) * unit) / (int)sizeof(CT);
};

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

// D(16x16) += A(16x32) B(32x16), FP16 inputs, FP32 accumulation; lane l
// supplies A[l%16][8(l/16)+i] and B[8(l/16)+i][l%16], i = 0..7, and holds D
// as the FP32 16x16x4 product (pinned by tests/test_gpu_mfma_layout.py).
__device__ inline v4f mfma16x16x32(half8 a, half8 b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// x = hi + lo + e with hi, lo FP16 (round to nearest), |e| <= 2^-22 |x| for
// x in FP16's normal range; two values per conversion.  x - hi is exact in
// FP32; m1 = -1 held in a register the compiler cannot see through, so that
// x + m1 hi stays one mixed-precision FMA (FP16 operand, FP32 result)
// instead of a conversion and a subtraction
__device__ inline void split_f16(float x0, float x1, float m1, f16x2 &hi, f16x2 &lo) {
  const f32x2 x = {x0, x1};
  hi = __builtin_convertvector(x, f16x2);
  const f32x2 r = {__builtin_fmaf((float)hi[0], m1, x0), __builtin_fmaf((float)hi[1], m1, x1)};
  lo = __builtin_convertvector(r, f16x2);
}

// -1.0f that the compiler cannot constant-fold (see split_f16)
__device__ inline float opaque_minus_one() {
  float m1 = -1.0f;
  asm volatile("" : "+v"(m1));
  return m1;
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Per-batch record (layout documented with BatchPlan in gtk_plan.h), as
// lane (nl = lane & 15, g = lane >> 4) needs it.
template <typename OffT, int MB>
struct BatchRegs {
  uint32_t info, acc;  // wave-uniform header words
  uint32_t sb;         // slot whose W column this lane builds: s | b << 16
  u32x4 at;            // slots 4g .. 4g+3: abase | t << 16
  int32_t delta;       // column nl: seg0 - abase
  u32x2 em;            // end-mask word (lane & 15)
  typename Off4<OffT>::type off[MB];  // row nl + 16 mb of slots 4g .. 4g+3
};

template <typename CT, typename OffT, int NS, int MB>
__device__ inline void load_batch(const unsigned char *__restrict__ rec, const MatrixArgs &args, int64_t bi, int lane,
                                  BatchRegs<OffT, MB> &r) {
  const int nl = lane & 15, g = lane >> 4;
  const unsigned char *R = rec + bi * args.rec_bytes;
  const u32x2 h = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(R));
  r.info = h.x;
  r.acc = h.y;
  r.sb = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(R + 16) + a_row<CT>(nl));
  r.at = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(R + 80) + g);
  r.delta = __builtin_nontemporal_load(reinterpret_cast<const int32_t *>(R + 144) + nl);
  const int w = lane & 15;
  r.em = w < args.endmask_words ? __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(R + 208) + w)
                                : u32x2{0u, 0u};
#pragma unroll
  for (int mb = 0; mb < MB; mb++) {
    const int a = mb * 16 + nl;
    r.off[mb] = a < NS ? *reinterpret_cast<const typename Off4<OffT>::type *>(
                             R + args.off_offset + (a * 16 + 4 * g) * (int)sizeof(OffT))
                       : Off4<OffT>::all_none();
  }
}

// Scalar forms on the matrix cores: one tile per workgroup of NW wavefronts
// at a time.
//
// Phase A (all waves): node coordinates of the tile's cells are staged in
// LDS (in chunks, in the space the accumulators use later); each thread
// keeps one quadrature point q with its geometry gradients in registers and
// computes, per cell, J = sum_n x_n (x) grad N_n(q), det J, J^-1 in FP64
// and the factor G = kappa |det J| w_q J^-1 J^-T (mass: kappa |det J| w_q),
// stored in LDS in the contraction precision CT.
//
// Phase B (each wave independently): batches of the tile, claimed from a
// per-tile counter.  A batch packs whole columns whose contributors (cell
// slot s, local column b) total at most 16; for its 16 slots one product
// gives the local columns
//   A[n, a] = sum_k W[n, k] Bt[k, a],  k = (q, d),
//   W[n, (q,d)]  = (G_{s_n,q} grad phi_{b_n}(q))_d  (built per batch)
//   Bt[(q,d), a] = d_d phi_a(q)                    (reference table, registers)
// with the k dimension ordered so that lane group g = lane/16 covers the
// quadrature points g*QG .. g*QG+QG-1 with all their directions: each lane
// reads G and grad phi_b once per point with vector LDS loads.  The product
// runs on 16x16x4 FP64 or FP32 matrix cores, or (mixed form, long K) on
// 16x16x32 FP16 ones with W and Bt split into two FP16 parts each.  Lane
// (nl, g) ends up holding local row a = nl (+16) of slots 4g .. 4g+3.  The
// rows are added into FP64 column accumulators in LDS through the record's
// row offsets in steps (h, r): step (h, r) adds the rows of the slots 4g+r
// whose rank within their column lies in [4h, 4h+4).  Slots of one step
// belong to distinct columns, so no two lanes of one addition touch one
// address, and every nonzero is summed in contributor order (cells
// ascending); a step no lane needs is skipped.  The accumulators of a batch are then
// written flat and coalesced: entry e goes to e + delta[c], c = number of
// column ends before e (an end mask per batch), and cleared.  The record of
// the next batch is loaded while the current one is computed.
template <int D, int NV, int NS, int NQ, int FORM, typename CT, typename OffT, int NW>
struct BatchTile {
  static constexpr int kB = NW * 64;
  static constexpr int NG = NGeo<D, FORM>::value;
  static constexpr int NGP = NG == 3 ? 4 : NG;  // staged factors per (cell, point)
  static constexpr int KD = FORM == kLaplace ? D : 1;
  static constexpr int TP = FORM == kLaplace ? (D == 3 ? 4 : 2) : 1;  // table entries per (q, b)
  static constexpr int QG = (NQ + 3) / 4;
  static constexpr int NQP = 4 * QG;  // quadrature rows padded with zero rows to the four lane groups
  static constexpr int KS_ = QG * (FORM == kLaplace ? D : 1);
  static constexpr int MB_ = (NS + 15) / 16;
  // mixed form with a long K: the product on FP16 matrix cores, W and Bt
  // each split into two FP16 parts (hi*hi + hi*lo + lo*hi, FP32
  // accumulation), 32 K per instruction; the staged factors of a cell are
  // scaled by a power of two (kept after them, GS - 1) so that W's parts
  // stay in FP16's normal range
  static constexpr bool kSplit = sizeof(CT) == 4 && KS_ * MB_ >= 12;
  static constexpr int KT = (KS_ + 7) / 8;  // 32-wide K steps of the split product
  static constexpr int kScaleExp = 9;       // scaled |G| about 2^3 .. 2^7
  static constexpr int GS = GeoStride<NQP, NGP, CT, kSplit ? 1 : 0>::value;  // CT elements per cell of staged factors
  static constexpr int KS = QG * KD;
  static constexpr int MB = (NS + 15) / 16;
  static constexpr int CPR = kB / NQ;  // cells per geometry round
  static constexpr int KR = (kBatchSlotsK * NS + 63) / 64;  // write-out rounds of the largest batch
  static_assert(KR <= 16, "end mask words are read from lanes 0..15");
  static_assert(CPR >= 1, "block too small for the quadrature");
  using V4 = typename Acc4<CT>::type;
  using O4 = Off4<OffT>;

  struct Lds {
    CT *tab;     // [NQP][NS][TP]
    CT *G;       // [cells: GS][NQP][NGP]
    double *acc;  // accumulators; phase A: staged coordinates and coefficients
    double *X, *K;
    int *claim;
  };
  __device__ static Lds lds(unsigned char *raw, const MatrixArgs &a) {
    Lds l;
    l.tab = reinterpret_cast<CT *>(raw);
    l.G = reinterpret_cast<CT *>(raw + a.lds_geo);
    l.acc = reinterpret_cast<double *>(raw + a.acc_offset);
    l.X = l.acc;
    l.K = l.X + (size_t)a.stage_cells * NV * D;
    l.claim = reinterpret_cast<int *>(raw + a.lds_claim);
    return l;
  }

  // reference table entry: direction d of local column b at point q (zero
  // past the points and the directions)
  __device__ static CT table_value(const MatrixArgs &a, int q, int b, int d) {
    double v = 0.0;
    if (q < NQ) {
      if constexpr (FORM == kLaplace) v = d < D ? a.gphi[(q * D + d) * NS + b] : 0.0;
      else v = a.phi[q * NS + b];
    }
    return (CT)v;
  }

  // reference table into LDS
  __device__ static void table(const MatrixArgs &a, const Lds &l, int tid) {
    for (int i = tid; i < NQP * NS * TP; i += kB) {
      const int q = i / (NS * TP), rem = i - q * NS * TP, b = rem / TP, d = rem - b * TP;
      l.tab[i] = table_value(a, q, b, d);
    }
  }

  // staging of the (cell -> node -> coordinates, coefficient) gather of a
  // chunk, entries [from, n): SU entries per thread and pass, each level
  // issued for all of them at once
  __device__ static void stage(const MatrixArgs &a, const Lds &l, int first_cell, int from, int n, int tid) {
    constexpr int SU = 4;
    for (int base = from; base < n; base += kB * SU) {
      int64_t cellv[SU], nodev[SU];
#pragma unroll
      for (int u = 0; u < SU; u++) {
        const int i = base + u * kB + tid;
        cellv[u] = i < n ? a.tile_cells[first_cell + i / NV] : 0;
      }
#pragma unroll
      for (int u = 0; u < SU; u++) {
        const int i = base + u * kB + tid;
        nodev[u] = i < n ? a.cell_nodes[cellv[u] * NV + (i % NV)] : 0;
        if (i < n && i % NV == 0) l.K[i / NV] = a.coef ? a.coef[cellv[u]] : 1.0;
      }
#pragma unroll
      for (int u = 0; u < SU; u++) {
        const int i = base + u * kB + tid;
        if (i < n) {
#pragma unroll
          for (int d = 0; d < D; d++) l.X[i * D + d] = a.xyz[nodev[u] * D + d];
        }
      }
    }
  }

  // multilinear geometry (quadrilaterals, hexahedra: vertex n at the corner
  // whose coordinate d is bit d of n): x(xi) = sum_S c_S prod_{d in S} xi_d
  // over the bit sets S, so that dx/dxi_j = c_{j} + sum_{S > {j}} c_S
  // prod_{d in S, d != j} xi_d needs D - 1 fused multiply-adds per entry of
  // J with the point's products (the gradients' sums over the vertices
  // above S) instead of one per vertex
  static constexpr bool kMultilinear = NV == (1 << D);
  static constexpr int NSUP = (1 << (D - 1)) - 1;  // sets S > {j} for each j

  // S > {j} in a fixed order: the k-th set with bit j that is not {j}
  __device__ static constexpr int super_set(int j, int k) {
    int seen = 0;
    for (int S = 0; S < NV; S++)
      if ((S >> j & 1) && S != (1 << j)) {
        if (seen == k) return S;
        seen++;
      }
    return 0;
  }

  // geometry factors of the nch staged cells (slots c0 ..) from the staged
  // coordinates; thread tid takes quadrature point tid % NQ
  __device__ static void geometry(const MatrixArgs &a, const Lds &l, int c0, int nch, int tid) {
    if constexpr (kMultilinear) {
      // the staged vertex coordinates become the coefficients c_S (Moebius
      // transform over the bits of the vertex number; c_0 = x_0), one
      // thread per (cell, direction)
      for (int i = tid; i < nch * D; i += kB) {
        double *X = l.X + (i / D) * NV * D + i % D;
        double v[NV];
#pragma unroll
        for (int n = 0; n < NV; n++) v[n] = X[n * D];
#pragma unroll
        for (int b = 1; b < NV; b <<= 1)
#pragma unroll
          for (int n = 0; n < NV; n++)
            if (n & b) v[n] -= v[n ^ b];
#pragma unroll
        for (int n = 1; n < NV; n++) X[n * D] = v[n];
      }
      __syncthreads();
    }
    const int tq = tid % NQ, tcs = tid / NQ;
    if (tcs >= CPR) return;
    constexpr int NGQ = kMultilinear ? 1 : NV;
    double gq[NGQ][D];   // general map: the vertices' gradients at the point
    double cw[D][NSUP];  // multilinear map: the point's products for S > {j}
    if constexpr (kMultilinear) {
#pragma unroll
      for (int j = 0; j < D; j++)
#pragma unroll
        for (int k = 0; k < NSUP; k++) {
          const int S = super_set(j, k);
          double s = 0.0;
#pragma unroll
          for (int n = 0; n < NV; n++)
            if ((n & S) == S) s += a.gN[(tq * NV + n) * D + j];
          cw[j][k] = s;
        }
    } else {
#pragma unroll
      for (int n = 0; n < NV; n++)
#pragma unroll
        for (int d = 0; d < D; d++) gq[n][d] = a.gN[(tq * NV + n) * D + d];
    }
    const double wq = a.w[tq];
    for (int sl = tcs; sl < nch; sl += CPR) {
      const double *X = l.X + sl * NV * D;
      double J[D][D];
      if constexpr (kMultilinear) {
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
          for (int j = 0; j < D; j++) {
            double s = X[(1 << j) * D + i];
#pragma unroll
            for (int k = 0; k < NSUP; k++) s = fma(X[super_set(j, k) * D + i], cw[j][k], s);
            J[i][j] = s;
          }
      } else {
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
      }
      // Laplace: G = kappa |det J| w J^-1 J^-T = (kappa w / |det J|) A A^T
      // with A = adj(J); mass: kappa |det J| w
      double A[D][D];
      const double det = FORM == kLaplace ? adjugate<D>(J, A) : determinant<D>(J);
      double kdV = FORM == kLaplace ? l.K[sl] * wq / fabs(det) : l.K[sl] * fabs(det) * wq;
      if constexpr (kSplit) {
        // the cell's scale 2^e from kappa and its first edge h, the same in
        // every thread of the cell: |G| ~ kappa h^(D-2) w (Laplace),
        // kappa h^D w (mass); 2^-e is kept for the products
        double h2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; d++) {
          // x_1 - x_0 (the multilinear coefficient c_{0} holds it already)
          const double e = kMultilinear ? X[D + d] : X[D + d] - X[d];
          h2 += e * e;
        }
        const double h = sqrt(h2);
        const double mag = fabs(l.K[sl]) * (FORM == kLaplace ? (D == 3 ? h : 1.0) : (D == 3 ? h * h * h : h * h));
        int e = 0;
        if (mag > 0.0 && mag < 1e300) e = min(100, max(-100, kScaleExp - ilogb(mag)));
        kdV = ldexp(kdV, e);
        if (tq == 0) l.G[(c0 + sl) * GS + GS - 1] = (CT)ldexp(1.0, -e);
      }
      CT *G = l.G + (c0 + sl) * GS + tq * NGP;
      if constexpr (NQP > NQ) {  // padding rows of the cell: zero
        if (tq < NQP - NQ) {
#pragma unroll
          for (int i = 0; i < NGP; i++) l.G[(c0 + sl) * GS + (NQ + tq) * NGP + i] = (CT)0;
        }
      }
      if constexpr (FORM == kLaplace) {
        CT gv[NGP];
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
          for (int j = i; j < D; j++) {
            double gg = 0.0;
#pragma unroll
            for (int m = 0; m < D; m++) gg = fma(A[i][m], A[j][m], gg);
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

  // the B operand of the batches (reference gradients or values)
  __device__ static void bfrag_load(const CT *tab, int nl, int g, CT (&bfrag)[MB][KS]) {
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const int a = mb * 16 + nl, q = g * QG + ks / KD, d = ks % KD;
        bfrag[mb][ks] = (a < NS && q < NQ) ? tab[(q * NS + a) * TP + d] : (CT)0;
      }
  }

  // the B operand of the split product: for K step t, lane group g supplies
  // its values i = 8t .. 8t+7 in the order (point j = i / KD, direction
  // d = i % KD) of the FP32 product, zero past KS; hi and lo FP16 parts
  __device__ static void bfrag_split_load(const CT *tab, int nl, int g, half8 (&bh)[MB][KT], half8 (&bl)[MB][KT]) {
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
      for (int t = 0; t < KT; t++)
#pragma unroll
        for (int p = 0; p < 4; p++) {
          float v[2];
#pragma unroll
          for (int u = 0; u < 2; u++) {
            const int i = 8 * t + 2 * p + u, a = mb * 16 + nl, q = g * QG + i / KD, d = i % KD;
            v[u] = (i < KS && a < NS && q < NQ) ? (float)tab[(q * NS + a) * TP + d] : 0.0f;
          }
          f16x2 hi, lo;
          split_f16(v[0], v[1], -1.0f, hi, lo);
          bh[mb][t][2 * p] = hi[0];
          bh[mb][t][2 * p + 1] = hi[1];
          bl[mb][t][2 * p] = lo[0];
          bl[mb][t][2 * p + 1] = lo[1];
        }
  }

  // one batch: W, the matrix-core product, the rank-step additions and the
  // flat write-out
  __device__ static void batch(const MatrixArgs &a, const Lds &l, double *acc, const CT (&bfrag)[MB][kSplit ? 1 : KS],
                               const half8 (&bh)[MB][kSplit ? KT : 1], const half8 (&bl)[MB][kSplit ? KT : 1],
                               const BatchRegs<OffT, MB> &cur, int lane) {
    const int g = lane >> 4;
    const uint32_t info = __builtin_amdgcn_readfirstlane(cur.info);
    const int tmax = (info >> 8) & 0xFF;
    // W column of slot a_row(nl); empty slots read slot 0 (finite values,
    // never scattered) and padding rows are zero in both operands
    const int s = cur.sb & 0xFFFF, b = cur.sb >> 16;
    const CT *Gc = l.G + s * GS + g * QG * NGP;
    const CT *Tb = l.tab + (g * QG * NS + b) * TP;
    V4 c[MB];
#pragma unroll
    for (int mb = 0; mb < MB; mb++) c[mb] = V4{0, 0, 0, 0};
    if constexpr (kSplit) {
      const float m1 = opaque_minus_one();
      float w[KT * 8];  // this lane's W values in the split product's K order
      // the factors and table entries of point j+1 are read before point
      // j's W is formed (the scheduling barrier keeps the reads ahead):
      // Q2 hex 64^3 mixed 0.615 -> 0.609 ms
      CT Gn[NGP], Tn[TP];
      lds_load<CT, NGP, (NGP * sizeof(CT)) % 16 == 0 ? 16 : 8>(Gc, Gn);
      lds_load<CT, TP, TP * sizeof(CT) >= 16 ? 16 : (TP * sizeof(CT) >= 8 ? 8 : 4)>(Tb, Tn);
#pragma unroll
      for (int j = 0; j < QG; j++) {
        CT G[NGP], T[TP];
#pragma unroll
        for (int i = 0; i < NGP; i++) G[i] = Gn[i];
#pragma unroll
        for (int i = 0; i < TP; i++) T[i] = Tn[i];
        if (j + 1 < QG) {
          lds_load<CT, NGP, (NGP * sizeof(CT)) % 16 == 0 ? 16 : 8>(Gc + (j + 1) * NGP, Gn);
          lds_load<CT, TP, TP * sizeof(CT) >= 16 ? 16 : (TP * sizeof(CT) >= 8 ? 8 : 4)>(Tb + (j + 1) * NS * TP, Tn);
        }
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (FORM == kLaplace) {
          if constexpr (D == 3) {
            w[j * 3 + 0] = G[0] * T[0] + G[1] * T[1] + G[2] * T[2];
            w[j * 3 + 1] = G[1] * T[0] + G[3] * T[1] + G[4] * T[2];
            w[j * 3 + 2] = G[2] * T[0] + G[4] * T[1] + G[5] * T[2];
          } else {
            w[j * 2 + 0] = G[0] * T[0] + G[1] * T[1];
            w[j * 2 + 1] = G[1] * T[0] + G[2] * T[1];
          }
        } else {
          w[j] = G[0] * T[0];
        }
      }
#pragma unroll
      for (int i = KS; i < KT * 8; i++) w[i] = 0.0f;
#pragma unroll
      for (int t = 0; t < KT; t++) {
        half8 wh, wl;
#pragma unroll
        for (int p = 0; p < 4; p++) {
          f16x2 hi, lo;
          split_f16(w[8 * t + 2 * p], w[8 * t + 2 * p + 1], m1, hi, lo);
          wh[2 * p] = hi[0];
          wh[2 * p + 1] = hi[1];
          wl[2 * p] = lo[0];
          wl[2 * p + 1] = lo[1];
        }
#pragma unroll
        for (int mb = 0; mb < MB; mb++) {
          c[mb] = mfma16x16x32(wh, bh[mb][t], c[mb]);
          c[mb] = mfma16x16x32(wh, bl[mb][t], c[mb]);
          c[mb] = mfma16x16x32(wl, bh[mb][t], c[mb]);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < QG; j++) {
        CT G[NGP], T[TP];
        lds_load<CT, NGP, (NGP * sizeof(CT)) % 16 == 0 ? 16 : 8>(Gc + j * NGP, G);
        lds_load<CT, TP, TP * sizeof(CT) >= 16 ? 16 : (TP * sizeof(CT) >= 8 ? 8 : 4)>(Tb + j * NS * TP, T);
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
#pragma unroll
          for (int mb = 0; mb < MB; mb++) c[mb] = mfma16x16x4(W[d], bfrag[mb][j * KD + d], c[mb]);
        }
      }
    }
    // destinations of the 4 slots x MB rows this lane holds; -1: a row
    // outside the row set or padding (not added)
    const uint32_t at4[4] = {cur.at.x, cur.at.y, cur.at.z, cur.at.w};
    int win[4];  // window of 4 ranks of each held slot (empty slots: none)
    int dst[MB][4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      win[r] = (int)(at4[r] >> 18);
      const int ab = (int)(at4[r] & 0xFFFF);
#pragma unroll
      for (int mb = 0; mb < MB; mb++) {
        const int off = O4::get(cur.off[mb], r);
        dst[mb][r] = off != (int)O4::none ? ab + off : -1;
      }
    }
    // split product: the rows of slot 4g+r come scaled by its cell's 2^e;
    // lane 4g+r (group 0) holds that slot's cell
    double us[4] = {1.0, 1.0, 1.0, 1.0};
    if constexpr (kSplit) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int sr = __shfl((int)cur.sb, 4 * g + r) & 0xFFFF;  // 16 bits: a 24-bit multiply
        us[r] = (double)l.G[sr * GS + GS - 1];
      }
    }
    // step (h, r) adds slot 4g+r of rank 4h..4h+3: the slots of one column
    // in one step are 4 ranks apart, never in one window, so no two lanes of
    // an addition touch one entry, and the placement of the columns
    // (batch_slot_start) makes the steps of a column come in rank order
    for (int h = 0; h < (tmax + 3) >> 2; h++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        if (win[r] == h) {
#pragma unroll
          for (int mb = 0; mb < MB; mb++)
            if (dst[mb][r] >= 0) atomicAdd(&acc[dst[mb][r]], kSplit ? (double)c[mb][r] * us[r] : (double)c[mb][r]);
        }
      }
    }
    // flat write-out of the batch's KR rounds of 64 accumulator entries:
    // entry e goes to e + delta[c], c = column ends before e (the padding
    // column sends the entries past the batch's columns to the value array's
    // tail); each entry is read and cleared by one LDS exchange, all
    // exchanges before the stores, so every batch issues the same number of
    // stores and the next batch's prefetched record is waited for without
    // waiting for them (skipping the rounds wholly past the batch's entries
    // measured slower: 3.50 -> 4.09 ms on the 256-problem 2D batch)
    int ends = 0;
    double v[KR];
    int32_t z[KR];
#pragma unroll
    for (int k = 0; k < KR; k++) {
      const uint32_t lo = __builtin_amdgcn_readlane(cur.em.x, k);  // k < 16: end-mask words
      const uint32_t hi = __builtin_amdgcn_readlane(cur.em.y, k);
      const int cidx = ends + (int)__builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
      ends += __builtin_popcount(lo) + __builtin_popcount(hi);
      v[k] = __hip_atomic_exchange(&acc[k * 64 + lane], 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      z[k] = k * 64 + lane + __shfl(cur.delta, cidx);
    }
#pragma unroll
    for (int k = 0; k < KR; k++) a.nzval[(int64_t)z[k]] = v[k];
  }

  // the tile's batches, claimed from the per-tile counter (wave w starts
  // with batch w; the result does not depend on which wave runs a batch);
  // `between` runs after each batch of this wave
  template <typename F>
  __device__ static void batches(const MatrixArgs &a, const Lds &l, int tile, int tid, F &&between) {
    const int wave = tid >> 6, lane = tid & 63;
    double *acc = l.acc + (size_t)wave * a.acc_per_wave;
    for (int i = lane; i < a.acc_per_wave; i += 64) acc[i] = 0.0;
    CT bfrag[MB][kSplit ? 1 : KS];
    half8 bh[MB][kSplit ? KT : 1], bl[MB][kSplit ? KT : 1];
    if constexpr (kSplit) bfrag_split_load(l.tab, lane & 15, lane >> 4, bh, bl);
    else bfrag_load(l.tab, lane & 15, lane >> 4, bfrag);
    const int bt0 = a.tile_batch_ptr[tile], bt1 = a.tile_batch_ptr[tile + 1];
    // a claim is issued with the prefetch of a record and its result read
    // one batch later, so the atomic's round trip overlaps that batch
    auto claim_issue = [&]() -> int {
      int v = 0;
      if (lane == 0) v = atomicAdd(l.claim, 1);
      return v;
    };
    if constexpr (D == 2 && sizeof(CT) == 4) {
      // records two batches ahead: a 2D batch of the FP32 product is too
      // short to cover a record load's latency (256-problem 2D batch,
      // matrices: 3.50 -> 3.20 ms; the FP64 form unchanged, Q1 hexahedra
      // 0.745 -> 0.823 ms, so only here)
      int b1 = bt0 + wave, b2 = bt1, pend = bt1 - bt0;
      BatchRegs<OffT, MB> n1, n2;
      if (b1 < bt1) {
        load_batch<CT, OffT, NS, MB>(a.rec, a, b1, lane, n1);
        b2 = bt0 + (int)__builtin_amdgcn_readfirstlane(claim_issue());
        if (b2 < bt1) {
          load_batch<CT, OffT, NS, MB>(a.rec, a, b2, lane, n2);
          pend = claim_issue();
        }
      }
      while (b1 < bt1) {
        const BatchRegs<OffT, MB> cur = n1;
        n1 = n2;
        b1 = b2;
        if (b2 < bt1) {
          b2 = bt0 + (int)__builtin_amdgcn_readfirstlane(pend);
          if (b2 < bt1) {
            load_batch<CT, OffT, NS, MB>(a.rec, a, b2, lane, n2);
            pend = claim_issue();
          }
        }
        batch(a, l, acc, bfrag, bh, bl, cur, lane);
        between();
      }
      return;
    }
    int bi = bt0 + wave, pend = bt1 - bt0;
    BatchRegs<OffT, MB> nxt;
    if (bi < bt1) {
      load_batch<CT, OffT, NS, MB>(a.rec, a, bi, lane, nxt);
      pend = claim_issue();
    }
    while (bi < bt1) {
      const BatchRegs<OffT, MB> cur = nxt;
      bi = bt0 + (int)__builtin_amdgcn_readfirstlane(pend);
      if (bi < bt1) {
        load_batch<CT, OffT, NS, MB>(a.rec, a, bi, lane, nxt);
        pend = claim_issue();
      }
      batch(a, l, acc, bfrag, bh, bl, cur, lane);
      between();
    }
  }

  // one whole tile (table already in LDS): geometry then batches; the
  // caller provides a barrier-separated start (nothing of the previous tile
  // in use)
  __device__ static void tile_simple(const MatrixArgs &a, const Lds &l, int tile, int tid) {
    const int cell0 = a.tile_cell_ptr[tile];
    const int ncell = a.tile_cell_ptr[tile + 1] - cell0;
    if (tid == 0) *l.claim = NW;  // read after the geometry phase's barriers
    for (int c0 = 0; c0 < ncell; c0 += a.stage_cells) {
      const int nch = min(a.stage_cells, ncell - c0);
      stage(a, l, cell0 + c0, 0, nch * NV, tid);
      __syncthreads();
      geometry(a, l, c0, nch, tid);
      __syncthreads();
    }
    batches(a, l, tile, tid, [] {});
  }
};

// One problem: a persistent grid, each workgroup looping over tiles; the
// next tile's cell/node ids and coefficients are gathered into registers
// during the current tile's batches (one gather level per batch), its
// coordinates at the tile's start.
template <int D, int NV, int NS, int NQ, int FORM, typename CT, typename OffT, int NW>
__global__ __launch_bounds__(NW * 64) void matrix_batch_kernel(MatrixArgs args) {
  using BT = BatchTile<D, NV, NS, NQ, FORM, CT, OffT, NW>;
  constexpr int kB = BT::kB;
  constexpr int PF = 2;  // staging entries per thread prefetched for the next tile
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const typename BT::Lds l = BT::lds(lds_raw, args);
  BT::table(args, l, threadIdx.x);
  int32_t pc[PF], pn[PF];
  double pk[PF];
  int pre_tile = -1, npre = 0;
  for (int tile = blockIdx.x; tile < args.ntiles; tile += gridDim.x) {
    __syncthreads();  // the table is in place; the previous tile's batches are done
    // per-thread indices are rederived each tile (an opaque copy of the
    // thread id) so that their address arithmetic is not held in registers
    // across the whole tile loop
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int cell0 = args.tile_cell_ptr[tile];
    const int ncell = args.tile_cell_ptr[tile + 1] - cell0;
    if (tid == 0) *l.claim = NW;  // read after the geometry phase's barriers
    for (int c0 = 0; c0 < ncell; c0 += args.stage_cells) {
      const int nch = min(args.stage_cells, ncell - c0);
      int from = 0;
      if (c0 == 0 && pre_tile == tile) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
          const int i = u * kB + tid;
          if (i < npre) {
#pragma unroll
            for (int d = 0; d < D; d++) l.X[i * D + d] = args.xyz[(int64_t)pn[u] * D + d];
            if (i % NV == 0) l.K[i / NV] = pk[u];
          }
        }
        from = npre;
      }
      BT::stage(args, l, cell0 + c0, from, nch * NV, tid);
      __syncthreads();
      BT::geometry(args, l, c0, nch, tid);
      __syncthreads();
    }
    const int ntile = tile + gridDim.x;
    int pf_cell0 = 0, pf_n = 0;
    if (ntile < args.ntiles) {
      pf_cell0 = args.tile_cell_ptr[ntile];
      pf_n = min(min(args.tile_cell_ptr[ntile + 1] - pf_cell0, args.stage_cells) * NV, PF * kB);
#pragma unroll
      for (int u = 0; u < PF; u++) {
        const int i = u * kB + tid;
        pc[u] = i < pf_n ? args.tile_cells[pf_cell0 + i / NV] : 0;
      }
    }
    bool pf_nodes = false;  // second gather level issued
    auto pf_step = [&]() {
      if (pf_nodes) return;
#pragma unroll
      for (int u = 0; u < PF; u++) {
        const int i = u * kB + tid;
        pn[u] = i < pf_n ? args.cell_nodes[(int64_t)pc[u] * NV + (i % NV)] : 0;
        pk[u] = (i < pf_n && i % NV == 0) ? (args.coef ? args.coef[pc[u]] : 1.0) : 0.0;
      }
      pf_nodes = true;
    };
    BT::batches(args, l, tile, tid, pf_step);
    pf_step();
    pre_tile = ntile;
    npre = pf_n;
  }
}

// Many problems in one launch: a persistent grid over a flattened queue of
// (problem, tile) items; each problem's arguments (its arrays, plan and LDS
// layout) are in `problems`, its element type selects the tile body.  The
// workgroup reloads the reference table when the problem changes.
template <typename E, int FORM, typename CT, int NW>
__device__ __attribute__((always_inline)) inline void batched_tile_body(unsigned char *raw, const MatrixArgs &a,
                                                                    int tile, bool new_problem) {
  using BT = BatchTile<E::d, E::nv, E::ns, E::nq, FORM, CT, uint8_t, NW>;
  const typename BT::Lds l = BT::lds(raw, a);
  if (new_problem) {
    BT::table(a, l, threadIdx.x);
    __syncthreads();
  }
  BT::tile_simple(a, l, tile, threadIdx.x);
}

// per-cell stride (doubles) of the vector tile's f . dV array: the next odd
// count, so that lanes on different cells fall on different LDS banks
template <int N>
struct VecCellStride {
  static constexpr int value = N % 2 ? N : N + 1;
};

// One tile of the load vector.  Each cell's node coordinates (and, for a
// discrete field, its nodal values) are read from HBM once into LDS; then
// f . dV at every (cell, point) and the element vectors are formed in LDS,
// and each owned row sums its contributions in cell order.  LDS: the tables (phi, order-1 geometry
// gradients) and, per cell, (NQ*NC + NV*D + NS*NC) doubles.  BS threads; the
// caller synchronises before LDS is reused.
template <int D, int NV, int NS, int NQ, int NC, int FKIND, int BS, int NL = 0>
__device__ inline void vector_tile(const VectorArgs &args, int tile, unsigned char *lds_raw) {
  constexpr int NDOF = NS * NC;
  constexpr int NT = NL > 0 ? NL : 1;  // table sets: one per local face for boundary items
  // odd strides: lanes on different points (or cells) fall on different LDS banks
  constexpr int PS = VecCellStride<NS>::value;       // phi row
  constexpr int GR = VecCellStride<NV * D>::value;   // geometry-gradient row
  constexpr int FS = VecCellStride<NQ * NC>::value;  // f . dV of a cell
  constexpr int U = 4;                                // staging items per thread in flight
  const int cell0 = args.tile_cell_ptr[tile];
  const int ncell = args.tile_cell_ptr[tile + 1] - cell0;
  const int col0 = args.tile_col_ptr[tile];
  const int ncol = args.tile_col_ptr[tile + 1] - col0;
  double *phi_s = reinterpret_cast<double *>(lds_raw);  // [NT][NQ][PS]
  double *gN_s = phi_s + NT * NQ * PS;                  // [NT][NQ][GR]: [NV][D] used
  double *F_s = gN_s + NT * NQ * GR;                    // [cells][FS]: f . dV at the points
  double *X_s = F_s + ncell * FS;                       // [cells][NV][D]
  double *fc_s = X_s + ncell * NV * D;                  // [cells][NDOF] (discrete field)
  int32_t *lf_s = reinterpret_cast<int32_t *>(fc_s + ncell * NDOF);  // [cells] local face (NL > 0)
  int32_t *cid_s = reinterpret_cast<int32_t *>(F_s);    // cell ids while staging
  for (int i = threadIdx.x; i < NT * NQ * NS; i += BS) phi_s[(i / NS) * PS + i % NS] = args.phi[i];
  for (int i = threadIdx.x; i < NT * NQ * NV * D; i += BS) gN_s[(i / (NV * D)) * GR + i % (NV * D)] = args.gN[i];
  for (int i = threadIdx.x; i < ncell; i += BS) {
    const int32_t c = args.tile_cells[cell0 + i];
    cid_s[i] = c;
    if constexpr (NL > 0) lf_s[i] = args.item_face[c];
  }
  __syncthreads();
  // node coordinates (node id -> xyz) and nodal values (DOF id -> f, free
  // then Dirichlet), each level issued for U items before the next
  for (int i0 = threadIdx.x; i0 < ncell * NV; i0 += U * BS) {
    int64_t node[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int idx = i0 + u * BS;
      node[u] = idx < ncell * NV ? args.cell_nodes[(int64_t)cid_s[idx / NV] * NV + idx % NV] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int idx = i0 + u * BS;
      if (idx < ncell * NV) {
#pragma unroll
        for (int i = 0; i < D; i++) X_s[idx * D + i] = args.xyz[node[u] * D + i];
      }
    }
  }
  if constexpr (FKIND == kFNodal) {
    for (int i0 = threadIdx.x; i0 < ncell * NDOF; i0 += U * BS) {
      int32_t dof[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int idx = i0 + u * BS;
        dof[u] = idx < ncell * NDOF ? args.cell_dofs[(int64_t)cid_s[idx / NDOF] * NDOF + idx % NDOF] : 1;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int idx = i0 + u * BS;
        if (idx < ncell * NDOF) fc_s[idx] = dof[u] > 0 ? args.f[dof[u] - 1] : args.f[args.n_free - dof[u] - 1];
      }
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < ncell * NQ; idx += BS) {
    const int s = idx / NQ, q = idx - s * NQ;
    const int l = NL > 0 ? lf_s[s] : 0;
    const double *gq = gN_s + (l * NQ + q) * GR;
    double J[D][D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) J[i][j] = 0.0;
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
      for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) J[i][j] = fma(X_s[(s * NV + v) * D + i], gq[v * D + j], J[i][j]);
    double dV;
    if constexpr (NL > 0) dV = surface_measure<D>(J, args.face_jac + l * D * (D - 1)) * args.w[q];
    else dV = fabs(determinant<D>(J)) * args.w[q];
    const double *pq = phi_s + (l * NQ + q) * PS;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      double f;
      if constexpr (FKIND == kFConst) {
        f = args.f[c];
      } else if constexpr (FKIND == kFCell) {
        f = args.f[(int64_t)args.tile_cells[cell0 + s] * NC + c];
      } else {
        // discrete field: sum_a f[dof(cell, a)] phi_a(q)
        f = 0.0;
#pragma unroll 4
        for (int n = 0; n < NS; n++) f = fma(fc_s[s * NDOF + n * NC + c], pq[n], f);
      }
      F_s[s * FS + q * NC + c] = f * dV;
    }
  }
  __syncthreads();
  // element vectors b[s][k] = sum_q phi[q][n] F[s][q][c], k = n*NC + c, KB
  // entries per thread: a thread reads its cell's F once per point and the
  // wave shares the phi reads.  b reuses the nodal values' LDS.
  constexpr int KB = NDOF < 8 ? NDOF : 8;
  constexpr int NKB = (NDOF + KB - 1) / KB;
  double *b_s = fc_s;  // [cells][NDOF]
  for (int idx = threadIdx.x; idx < ncell * NKB; idx += BS) {
    const int s = idx / NKB, k0 = (idx - s * NKB) * KB;
    const double *ph = phi_s + (NL > 0 ? lf_s[s] * NQ * PS : 0);
    double acc[KB];
#pragma unroll
    for (int u = 0; u < KB; u++) acc[u] = 0.0;
#pragma unroll 2
    for (int q = 0; q < NQ; q++) {
#pragma unroll
      for (int u = 0; u < KB; u++) {
        const int k = k0 + u < NDOF ? k0 + u : NDOF - 1;
        const int n = k / NC, c = k - n * NC;
        acc[u] = fma(ph[q * PS + n], F_s[s * FS + q * NC + c], acc[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < KB; u++)
      if (k0 + u < NDOF) b_s[s * NDOF + k0 + u] = acc[u];
  }
  __syncthreads();
  // rows, U per thread in flight: row id and contribution range, then the
  // contributions in cell order
  for (int c0 = threadIdx.x; c0 < ncol; c0 += U * BS) {
    int j[U], k0[U], k1[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int ci = c0 + u * BS;
      const int pc = col0 + (ci < ncol ? ci : 0);
      j[u] = args.col_id[pc];
      k0[u] = args.col_ctr_ptr[pc];
      k1[u] = ci < ncol ? args.col_ctr_ptr[pc + 1] : k0[u];
    }
    double acc[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      acc[u] = 0.0;
      for (int k = k0[u]; k < k1[u]; k++) acc[u] += b_s[args.ctr_slot[k] * NDOF + args.ctr_b[k]];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (c0 + u * BS < ncol) args.out[j[u]] = acc[u];
  }
}

template <int D, int NV, int NS, int NQ, int NC, int FKIND, int NL>
__global__ __launch_bounds__(kBlock) void vector_tile_kernel(VectorArgs args) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  vector_tile<D, NV, NS, NQ, NC, FKIND, kBlock, NL>(args, blockIdx.x, lds_raw);
}

template <typename E, int BS>
__device__ __attribute__((always_inline)) inline void batched_vector_tile_body(const VectorArgs &v, int tile, int fkind,
                                                                           unsigned char *raw) {
  if (fkind == kFConst) vector_tile<E::d, E::nv, E::ns, E::nq, E::nc, kFConst, BS>(v, tile, raw);
  else if (fkind == kFCell) vector_tile<E::d, E::nv, E::ns, E::nq, E::nc, kFCell, BS>(v, tile, raw);
  else vector_tile<E::d, E::nv, E::ns, E::nq, E::nc, kFNodal, BS>(v, tile, raw);
}

// Workgroup w starts with item w and then claims the next unclaimed item
// from the launch's counter, one item ahead (the claim's round trip overlaps
// the current item), so the workgroups finish together although the items'
// costs differ by orders of magnitude; the claim is passed to all threads
// through two LDS slots used in turn.
// The tile bodies as functions the compiler may keep as calls (kernels over
// several element types) and, in a kernel of one element type, inlined:
// there a call's saved registers would go to scratch at every tile (Q1/Q2
// hexahedra: 164 bytes per lane, about as many bytes written per launch as
// the values themselves)
template <typename E, int FORM, typename CT, int NW>
__device__ inline void batched_tile(unsigned char *raw, const MatrixArgs &a, int tile, bool new_problem) {
  batched_tile_body<E, FORM, CT, NW>(raw, a, tile, new_problem);
}
template <typename E, int BS>
__device__ inline void batched_vector_tile(const VectorArgs &v, int tile, int fkind, unsigned char *raw) {
  batched_vector_tile_body<E, BS>(v, tile, fkind, raw);
}

template <int FORM, typename CT, int NW, typename E0, typename E1, typename E2>
__global__ __launch_bounds__(NW * 64) void matrix_batched_kernel(BatchedArgs ba) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int *claim = reinterpret_cast<int *>(lds_raw + ba.claim_offset);
  if (threadIdx.x == 0) claim[0] = (int)gridDim.x + atomicAdd(ba.counter, 1);
  int prev = -1;
  int item = blockIdx.x;
  for (int round = 0; item < ba.nitems; round++) {
    __syncthreads();  // the previous tile (or problem) is done with LDS; this round's claim is visible
    const int next = claim[round & 1];
    if (threadIdx.x == 0 && next < ba.nitems) claim[(round + 1) & 1] = (int)gridDim.x + atomicAdd(ba.counter, 1);
    const int2 it = ba.items[item];  // (problem, matrix tile t >= 0 or load-vector tile ~t)
    if (it.y < 0) {
      const int fk = ba.fkinds[it.x];
      if constexpr (std::is_same_v<E0, E1> && std::is_same_v<E1, E2>) {
        batched_vector_tile_body<E0, NW * 64>(ba.loads[it.x], ~it.y, fk, lds_raw);
      } else {
        switch (ba.types[it.x]) {
          case 0: batched_vector_tile<E0, NW * 64>(ba.loads[it.x], ~it.y, fk, lds_raw); break;
          case 1: batched_vector_tile<E1, NW * 64>(ba.loads[it.x], ~it.y, fk, lds_raw); break;
          default: batched_vector_tile<E2, NW * 64>(ba.loads[it.x], ~it.y, fk, lds_raw); break;
        }
      }
      prev = -1;  // the reference tables in LDS are overwritten
      item = next;
      continue;
    }
    const MatrixArgs &a = ba.problems[it.x];
    const bool np = it.x != prev;
    prev = it.x;
    if constexpr (std::is_same_v<E0, E1> && std::is_same_v<E1, E2>) {
      batched_tile_body<E0, FORM, CT, NW>(lds_raw, a, it.y, np);  // one element type: the body inlined
    } else {
      switch (ba.types[it.x]) {
        case 0: batched_tile<E0, FORM, CT, NW>(lds_raw, a, it.y, np); break;
        case 1: batched_tile<E1, FORM, CT, NW>(lds_raw, a, it.y, np); break;
        default: batched_tile<E2, FORM, CT, NW>(lds_raw, a, it.y, np); break;
      }
    }
    item = next;
  }
}

// ---------------------------------------------------------------------------
// dispatch tables
// ---------------------------------------------------------------------------
// NL > 0: boundary items of a cell with NL local faces (NS shape functions
// and NQ points of one face)
template <int D, int NV, int NS, int NQ, int NC, int NL = 0>
struct Elem {
  static constexpr int d = D, nv = NV, ns = NS, nq = NQ, nc = NC, nl = NL;
};

template <typename E, int FORM, typename CT, typename OffT, int NW>
static KernelInfo matrix_batch_info() {
  constexpr int TP = FORM == kLaplace ? (E::d == 3 ? 4 : 2) : 1;
  constexpr int NQP = 4 * ((E::nq + 3) / 4);
  KernelInfo k;
  k.fn = reinterpret_cast<const void *>(&matrix_batch_kernel<E::d, E::nv, E::ns, E::nq, FORM, CT, OffT, NW>);
  k.table_bytes = (size_t)NQP * E::ns * TP * sizeof(CT);
  k.geo_bytes_per_cell = (size_t)BatchTile<E::d, E::nv, E::ns, E::nq, FORM, CT, OffT, NW>::GS * sizeof(CT);
  k.rounds = (kBatchSlotsK * E::ns + 63) / 64;
  k.cols_per_wave = 16;
  k.mfma = true;
  k.block = NW * 64;
  return k;
}

template <typename E, int FORM, typename CT, int NW>
static KernelInfo matrix_batch_info(int off_bytes) {
  return off_bytes == 1 ? matrix_batch_info<E, FORM, CT, uint8_t, NW>() : matrix_batch_info<E, FORM, CT, uint16_t, NW>();
}

// waves per workgroup: 8 (2 per SIMD), 12 or 16 (3 or 4 per SIMD, registers
// limited to 168 or 128 by the launch bounds)
template <typename E, int FORM, typename CT>
static KernelInfo matrix_batch_info(int off_bytes, int waves) {
  if (waves == 16) return matrix_batch_info<E, FORM, CT, 16>(off_bytes);
  if (waves == 12) return matrix_batch_info<E, FORM, CT, 12>(off_bytes);
  return matrix_batch_info<E, FORM, CT, 8>(off_bytes);
}

// LDS sizes below follow the kernels' layouts (the host's tile limits use
// the same formulas)
template <typename E, int FORM, typename CT, int BS>
static KernelInfo matrix_info_bs() {
  constexpr int NT = E::nl > 0 ? E::nl : 1;
  KernelInfo k;
  k.fn = reinterpret_cast<const void *>(&matrix_tile_kernel<E::d, E::nv, E::ns, E::nq, E::nc, FORM, CT, E::nl, BS>);
  constexpr bool cell_geo = CellGeo<E::d, E::nv, FORM, E::nl>::value;
  constexpr bool nit = FORM == kNitsche || FORM == kNitscheLoad || FORM == kSkeleton;
  constexpr int ngphi = E::nl > 0 ? (nit ? NT * E::nq * E::d * E::ns : 0)
                                  : (cell_geo ? E::ns * E::ns * E::d * E::d : E::nq * E::d * E::ns);
  k.table_bytes = (size_t)(ngphi + NT * E::nq * E::ns) * sizeof(CT);
  constexpr int GQ = cell_geo ? 1 : E::nq;
  k.geo_bytes_per_cell = (size_t)GQ * CellGeoStride<E::d, FORM, E::nl, E::nv>::value * sizeof(CT) +
                         (E::nl > 0 ? sizeof(int32_t) : 0);
  k.cols_per_wave = 64 / ValuLanes<E::ns, E::nc, FORM>::value;
  k.block = BS;
  return k;
}

// waves == 4: the 256-thread workgroup (fallback for long columns)
template <typename E, int FORM, typename CT>
static KernelInfo matrix_info(int waves = 0) {
  return waves == 4 ? matrix_info_bs<E, FORM, CT, 256>() : matrix_info_bs<E, FORM, CT, ValuBlock<CT>::value>();
}

template <typename E, int FK>
static KernelInfo vector_info() {
  constexpr int NT = E::nl > 0 ? E::nl : 1;
  KernelInfo k;
  k.fn = reinterpret_cast<const void *>(&vector_tile_kernel<E::d, E::nv, E::ns, E::nq, E::nc, FK, E::nl>);
  k.table_bytes =
      (size_t)NT * E::nq * (VecCellStride<E::ns>::value + VecCellStride<E::nv * E::d>::value) * sizeof(double);
  k.geo_bytes_per_cell = (size_t)(VecCellStride<E::nq * E::nc>::value + E::nv * E::d + E::ns * E::nc +
                                  (E::nl > 0 ? 1 : 0)) * sizeof(double);
  k.cols_per_wave = 64;
  return k;
}

template <typename E>
static bool match(const ElemKey &key) {
  return key.dim == E::d && key.nv == E::nv && key.ns == E::ns && key.nq == E::nq && key.nc == E::nc &&
         key.nl == E::nl;
}

// skeleton items: two cells' nodes, a point set per (local face, vertex
// order) pair of the two sides
template <typename E>
constexpr bool is_skeleton() {
  return E::nl == (E::d == 2 ? 8 : 12);
}

template <typename E>
static bool try_matrix(const ElemKey &key, int form, int precision, int off_bytes, bool allow_mfma, int waves,
                       KernelInfo &out) {
  if (!match<E>(key)) return false;
  if constexpr (is_skeleton<E>()) {
    if (form != kSkeleton) return false;
    out = precision == 0 ? matrix_info<E, kSkeleton, double>(waves) : matrix_info<E, kSkeleton, float>(waves);
  } else if constexpr (E::nl > 0) {
    // boundary items on the VALU kernel: the mass form, and on scalar fields
    // the Nitsche terms
    if (form == kMass) {
      out = precision == 0 ? matrix_info<E, kMass, double>(waves) : matrix_info<E, kMass, float>(waves);
    } else if constexpr (E::nc == 1) {
      if (form == kNitsche)
        out = precision == 0 ? matrix_info<E, kNitsche, double>(waves) : matrix_info<E, kNitsche, float>(waves);
      else if (form == kNitscheLoad)
        out = precision == 0 ? matrix_info<E, kNitscheLoad, double>(waves)
                             : matrix_info<E, kNitscheLoad, float>(waves);
      else
        return false;
    } else {
      return false;
    }
  } else if constexpr (E::nc == 1) {
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
        out = precision == 0 ? matrix_info<E, kLaplace, double>(waves) : matrix_info<E, kLaplace, float>(waves);
      else
        out = precision == 0 ? matrix_info<E, kMass, double>(waves) : matrix_info<E, kMass, float>(waves);
    }
  } else {
    if (form != kElasticity || E::nc != E::d) return false;
    out = precision == 0 ? matrix_info<E, kElasticity, double>(waves) : matrix_info<E, kElasticity, float>(waves);
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
// Q2 quadrilaterals with integration degree 2 (2x2 Gauss points): the rule
// GalerkinToolkitExamples/src/poisson.jl keeps by default
using QuadQ2d2 = Elem<2, 4, 9, 4, 1>;
// scalar P1/P2 triangles and tetrahedra with integration degree 2k (the
// collapsed Gauss-Jacobi rules; the 11-point Strang rule for P2 tetrahedra)
using TriP1 = Elem<2, 3, 3, 4, 1>;
using TriP2 = Elem<2, 3, 6, 9, 1>;
using TetP1 = Elem<3, 4, 4, 8, 1>;
using TetP2 = Elem<3, 4, 10, 11, 1>;
// vector (elasticity) fields: Q1/Q2 quadrilaterals, Q1 hexahedra, P1/P2
// triangles, P1 tetrahedra (at most 64 DOFs per cell), degree-2k rules
using QuadQ1v = Elem<2, 4, 4, 4, 2>;
using QuadQ2v = Elem<2, 4, 9, 9, 2>;
using HexQ1v = Elem<3, 8, 8, 8, 3>;
using TriP1v = Elem<2, 3, 3, 4, 2>;
using TriP2v = Elem<2, 3, 6, 9, 2>;
using TetP1v = Elem<3, 4, 4, 8, 3>;
// boundary items: edges of Q1..Q3 quadrilaterals (k+1 Gauss points), faces
// of Q1/Q2 hexahedra ((k+1)^2), faces of P2 vector tetrahedra (9-point rule)
using QuadEdge1 = Elem<2, 4, 2, 2, 1, 4>;
using QuadEdge2 = Elem<2, 4, 3, 3, 1, 4>;
using QuadEdge3 = Elem<2, 4, 4, 4, 1, 4>;
using HexFace1 = Elem<3, 8, 4, 4, 1, 6>;
using HexFace2 = Elem<3, 8, 9, 9, 1, 6>;
using TetFace2v = Elem<3, 4, 6, 9, 3, 4>;
using QuadEdge2d2 = Elem<2, 4, 3, 2, 1, 4>;
// scalar boundary items of simplices: edges of P1/P2 triangles, faces of
// P1/P2 tetrahedra (integration degree 2k)
using TriEdge1 = Elem<2, 3, 2, 2, 1, 3>;
using TriEdge2 = Elem<2, 3, 3, 3, 1, 3>;
using TetFace1 = Elem<3, 4, 3, 4, 1, 4>;
using TetFace2 = Elem<3, 4, 6, 9, 1, 4>;
// Nitsche items: boundary faces with all DOFs of their cell, Gauss rules of
// k+1 points per direction on the face
using QuadNitsche1 = Elem<2, 4, 4, 2, 1, 4>;
using QuadNitsche2 = Elem<2, 4, 9, 3, 1, 4>;
using QuadNitsche3 = Elem<2, 4, 16, 4, 1, 4>;
using HexNitsche1 = Elem<3, 8, 8, 4, 1, 6>;
using HexNitsche2 = Elem<3, 8, 27, 9, 1, 6>;
using TriNitsche1 = Elem<2, 3, 3, 2, 1, 3>;
using TriNitsche2 = Elem<2, 3, 6, 3, 1, 3>;
using TetNitsche1 = Elem<3, 4, 4, 4, 1, 4>;
using TetNitsche2 = Elem<3, 4, 10, 9, 1, 4>;
// skeleton items: interior faces with both cells (their nodes and DOFs),
// Gauss rules of k+1 points per direction on the face, and for Q2
// quadrilaterals also 2 (integration degree 2)
using QuadSkeleton1 = Elem<2, 8, 4, 2, 1, 8>;
using QuadSkeleton2a = Elem<2, 8, 9, 2, 1, 8>;
using QuadSkeleton2 = Elem<2, 8, 9, 3, 1, 8>;
using QuadSkeleton3 = Elem<2, 8, 16, 4, 1, 8>;
using HexSkeleton1 = Elem<3, 16, 8, 4, 1, 12>;
using HexSkeleton2 = Elem<3, 16, 27, 9, 1, 12>;
using TriSkeleton1 = Elem<2, 6, 3, 2, 1, 8>;
using TriSkeleton2 = Elem<2, 6, 6, 3, 1, 8>;
using TetSkeleton1 = Elem<3, 8, 4, 4, 1, 12>;
using TetSkeleton2 = Elem<3, 8, 10, 9, 1, 12>;

bool find_matrix_kernel(const ElemKey &key, int form, int precision, int off_bytes, bool allow_mfma, int waves,
                        KernelInfo &out) {
  return try_matrix<QuadQ1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadQ2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadQ3>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<HexQ1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<HexQ2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetP2v>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadQ2d2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadEdge2d2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadEdge1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadEdge2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadEdge3>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<HexFace1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<HexFace2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetFace2v>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadNitsche1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadNitsche2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadNitsche3>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<HexNitsche1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<HexNitsche2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadSkeleton1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadSkeleton2a>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadSkeleton2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadSkeleton3>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<HexSkeleton1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<HexSkeleton2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadQ1v>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<QuadQ2v>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<HexQ1v>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TriP1v>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TriP2v>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetP1v>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TriP1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TriP2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetP1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetP2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TriSkeleton1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TriSkeleton2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetSkeleton1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetSkeleton2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TriEdge1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TriEdge2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetFace1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetFace2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TriNitsche1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TriNitsche2>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetNitsche1>(key, form, precision, off_bytes, allow_mfma, waves, out) ||
         try_matrix<TetNitsche2>(key, form, precision, off_bytes, allow_mfma, waves, out);
}

int batched_type(const ElemKey &key) {
  if (match<QuadQ1>(key)) return 0;
  if (match<QuadQ2>(key)) return 1;
  if (match<QuadQ3>(key)) return 2;
  if (match<HexQ1>(key)) return 3;
  if (match<HexQ2>(key)) return 4;
  if (match<TriP1>(key)) return 6;
  if (match<TriP2>(key)) return 7;
  return -1;
}

template <int FORM, typename CT, int NW, int FAMILY>
static KernelInfo batched_info() {
  KernelInfo k;
  if constexpr (FAMILY == 0)
    k.fn = reinterpret_cast<const void *>(&matrix_batched_kernel<FORM, CT, NW, QuadQ1, QuadQ2, QuadQ3>);
  else if constexpr (FAMILY == 1)  // the third slot is unused
    k.fn = reinterpret_cast<const void *>(&matrix_batched_kernel<FORM, CT, NW, HexQ1, HexQ2, HexQ2>);
  else if constexpr (FAMILY == kHexQ1Launch)
    k.fn = reinterpret_cast<const void *>(&matrix_batched_kernel<FORM, CT, NW, HexQ1, HexQ1, HexQ1>);
  else if constexpr (FAMILY == kHexQ2Launch)
    k.fn = reinterpret_cast<const void *>(&matrix_batched_kernel<FORM, CT, NW, HexQ2, HexQ2, HexQ2>);
  else
    k.fn = reinterpret_cast<const void *>(&matrix_batched_kernel<FORM, CT, NW, TriP1, TriP2, TriP2>);
  k.mfma = true;
  k.block = NW * 64;
  return k;
}

template <int FORM, int FAMILY>
static KernelInfo batched_info(int precision, int waves) {
  if (waves == 12)
    return precision == 0 ? batched_info<FORM, double, 12, FAMILY>() : batched_info<FORM, float, 12, FAMILY>();
  return precision == 0 ? batched_info<FORM, double, 8, FAMILY>() : batched_info<FORM, float, 8, FAMILY>();
}

template <int FORM>
static KernelInfo batched_family_info(int precision, int waves, int family) {
  if (family == 0) return batched_info<FORM, 0>(precision, waves);
  if (family == 1) return batched_info<FORM, 1>(precision, waves);
  if (family == kHexQ1Launch) return batched_info<FORM, kHexQ1Launch>(precision, waves);
  if (family == kHexQ2Launch) return batched_info<FORM, kHexQ2Launch>(precision, waves);
  return batched_info<FORM, 2>(precision, waves);
}

bool find_batched_kernel(int form, int precision, int waves, int family, KernelInfo &out) {
  if ((waves != 8 && waves != 12) || family < 0 || family > kHexQ2Launch) return false;
  if (form == kLaplace) out = batched_family_info<kLaplace>(precision, waves, family);
  else if (form == kMass) out = batched_family_info<kMass>(precision, waves, family);
  else return false;
  return true;
}

bool find_vector_kernel(const ElemKey &key, int fkind, KernelInfo &out) {
  return try_vector<QuadQ1>(key, fkind, out) || try_vector<QuadQ2>(key, fkind, out) ||
         try_vector<QuadQ3>(key, fkind, out) || try_vector<HexQ1>(key, fkind, out) ||
         try_vector<HexQ2>(key, fkind, out) || try_vector<TetP2v>(key, fkind, out) ||
         try_vector<QuadEdge1>(key, fkind, out) || try_vector<QuadEdge2>(key, fkind, out) ||
         try_vector<QuadEdge3>(key, fkind, out) || try_vector<HexFace1>(key, fkind, out) ||
         try_vector<HexFace2>(key, fkind, out) || try_vector<TetFace2v>(key, fkind, out) ||
         try_vector<QuadQ2d2>(key, fkind, out) || try_vector<QuadEdge2d2>(key, fkind, out) ||
         try_vector<TriP1>(key, fkind, out) || try_vector<TriP2>(key, fkind, out) ||
         try_vector<QuadQ1v>(key, fkind, out) || try_vector<QuadQ2v>(key, fkind, out) ||
         try_vector<HexQ1v>(key, fkind, out) || try_vector<TriP1v>(key, fkind, out) ||
         try_vector<TriP2v>(key, fkind, out) || try_vector<TetP1v>(key, fkind, out) ||
         try_vector<TetP1>(key, fkind, out) || try_vector<TetP2>(key, fkind, out) ||
         try_vector<TriEdge1>(key, fkind, out) || try_vector<TriEdge2>(key, fkind, out) ||
         try_vector<TetFace1>(key, fkind, out) || try_vector<TetFace2>(key, fkind, out);
}

}  // namespace gtk
