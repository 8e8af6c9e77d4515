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
      const uint16_t off = args.ctr_off[(int64_t)k * NDOF + a];
      if (off != kNoRowDev) acc[off] += (double)v;
    }
    for (int r = a; r < len; r += NDOF) args.nzval[seg0 + r] = acc[r];
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
static bool try_matrix(const ElemKey &key, int form, int precision, KernelInfo &out) {
  if (!match<E>(key)) return false;
  if (E::nc == 1) {
    if (form == kLaplace) out = precision == 0 ? matrix_info<E, kLaplace, double>() : matrix_info<E, kLaplace, float>();
    else if (form == kMass) out = precision == 0 ? matrix_info<E, kMass, double>() : matrix_info<E, kMass, float>();
    else return false;
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

bool find_matrix_kernel(const ElemKey &key, int form, int precision, KernelInfo &out) {
  return try_matrix<QuadQ1>(key, form, precision, out) || try_matrix<QuadQ2>(key, form, precision, out) ||
         try_matrix<QuadQ3>(key, form, precision, out) || try_matrix<HexQ1>(key, form, precision, out) ||
         try_matrix<HexQ2>(key, form, precision, out) || try_matrix<TetP2v>(key, form, precision, out);
}

bool find_vector_kernel(const ElemKey &key, int fkind, KernelInfo &out) {
  return try_vector<QuadQ1>(key, fkind, out) || try_vector<QuadQ2>(key, fkind, out) ||
         try_vector<QuadQ3>(key, fkind, out) || try_vector<HexQ1>(key, fkind, out) ||
         try_vector<HexQ2>(key, fkind, out) || try_vector<TetP2v>(key, fkind, out);
}

}  // namespace gtk
