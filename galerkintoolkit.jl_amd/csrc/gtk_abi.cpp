// gtk_abi.cpp -- C ABI of libgtk_hip (declared in include/gtk_hip.h).
//
// A context owns the device copies of the mesh, the cell -> DOF map, the
// reference tables, the per-cell coefficients, and, per (rows, cols) block,
// the CSC pattern, the column-owner plan and the value array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gtk_hip.h"
#include "gtk_kernels.h"
#include "gtk_plan.h"
#include "gtk_solve.h"
#include "gtk_symbolic.h"

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  template <typename T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

// Work plan of one (form, precision) over a block: the tiles, and either
// the batch records of the matrix-core kernel or the row offsets of the
// VALU kernel, with the kernel and its LDS layout.
struct ExecPlan {
  gtk::TilePlan tiles;
  gtk::BatchPlan bat;
  gtk::OffsetPlan offs;
  gtk::KernelInfo k;
  DevBuf tile_cell_ptr, tile_cells, tile_col_ptr, col_id, col_ctr_ptr, ctr_slot, ctr_b, ctr_off, tile_batch_ptr, rec;
  size_t lds = 0;
  int lds_geo = 0, acc_offset = 0, acc_per_wave = 0, stage_cells = 0, lds_claim = 0;
  int resident = 1;  // workgroups of the kernel resident on the device at once
  int tile_cap = 0;  // cells-per-tile cap the plan was built with (0: as many as fit in LDS)
};

// Device work space of the solve on a (FREE, FREE) block.
struct SolveWork {
  DevBuf perm, val64, val32, dinv, x, b, r, d, rr, z, p, q, partial, partial2, sc, bad;
  DevBuf bcol, bval;  // BSR(3) copy of the FP32 values (three-component fields)
  int block3 = -1;    // pattern made of aligned 3x3 blocks: -1 not checked, 0 no, 1 yes
  bool bsr_cols = false;                // block columns of the BSR(3) copy written
  bool prepared[2] = {false, false};    // solve plan built for FP64 / mixed
};

struct MatrixPlan {
  gtk::Pattern pat;
  int max_seg = 0;
  DevBuf colptr, nzval;
  DevBuf rowval;  // device copy (device symbolic phase, local map), on demand
  SolveWork sw;
  DevBuf lift_rptr, lift_perm, lift_col, lift_x, lift_b;  // row view of the block for gtk_lift_dirichlet
  std::map<int, std::unique_ptr<ExecPlan>> exec;  // key: form * 2 + precision
  bool valid = false;
};

struct VectorPlan {
  gtk::TilePlan tiles;
  DevBuf tile_cell_ptr, tile_cells, tile_col_ptr, col_id, col_ctr_ptr, ctr_slot, ctr_b, out, f;
  int64_t nrows = 0;
  bool valid = false;
};

int nthreads_host() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  return (int)std::min(16u, n);
}

}  // namespace

struct gtk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  // host copies
  int dim = 0, nv = 0;
  int64_t n_nodes = 0, n_cells = 0;
  std::vector<double> xyz;
  std::vector<int32_t> cell_nodes;  // 0-based
  int ncomp = 0, dofs_per_cell = 0;
  int64_t n_free = 0, n_diri = 0;
  std::vector<int32_t> cell_dofs;
  int nq = 0, nshape = 0;
  std::vector<double> w, phi, gphi_t, gN;  // gphi_t: [q][d][s]
  int ncoef = 0;
  // face mode: the "cells" are boundary items (cell, local face); tables per
  // local face, d(reference cell point)/d(reference face point) per face
  int nfaces = 0;
  std::vector<double> face_jac;
  std::vector<int32_t> item_face;
  // device copies
  DevBuf d_xyz, d_cell_nodes, d_cell_dofs, d_w, d_phi, d_gphi, d_gN, d_coef, d_face_jac, d_item_face, d_face_normal;
  // affine simplices: R[a][b][m][n] = sum_q w_q d_m phi_a(q) d_n phi_b(q),
  // the reference integrals of every node pair (FP64, formed once per
  // reference table set)
  DevBuf d_rtab;
  bool has_face_grad = false;  // face mode: reference gradients and normals set (Nitsche terms)
  bool faces_packed = false;   // item_face holds two point sets per item (skeleton items)
  // wavefronts per workgroup of the matrix-core kernels; 0 = per kernel:
  // 16 or 12 (four or three per SIMD) when that instantiation fits its
  // registers without scratch, else 8 (two per SIMD within 256 registers);
  // GTK_WAVES overrides
  int waves = 0;
  size_t lds_budget = 156 * 1024;  // LDS a workgroup may use (less gives several workgroups per CU)
  // largest single device allocation the context makes (GTK_MEM_LIMIT_MB;
  // 0: no cap): a larger request fails with GTK_ENOMEM before hipMalloc
  size_t alloc_cap = 0;
  std::map<int, std::unique_ptr<MatrixPlan>> mplans;
  std::map<int, std::unique_ptr<VectorPlan>> vplans;
  bool has_mesh = false, has_space = false, has_ref = false;
  // bumped by every call that replaces the mesh, the space, the reference
  // tables or a pattern: a batch built on older plans refuses to run
  uint64_t generation = 1;
};

static std::string g_create_err;

// Scratch of gtk_assemble_vector_noalloc: the staged load values and the
// load vector before its copy to the host live in this file-scope device
// array, so that call allocates nothing.  Every device the library runs on
// holds its own copy until the process ends; its address on each device is
// resolved once, in gtk_create with that device current, and host code only
// ever uses that resolved address (the array's name is a host-side stand-in).
static constexpr size_t kNoAllocDoubles = GTK_NOALLOC_MAX_DOUBLES;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
static constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

static int resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return GTK_OK;  // the no-alloc call reports it
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return GTK_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_create_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return GTK_EDEVICE;
  }
  g_scratch[device] = static_cast<double *>(p);
  return GTK_OK;
}

// Zero fill for the no-alloc path (a kernel, not a memset node, so callers
// may capture the call in a graph).
// clears the item counters of a batch's launches: the first node of its graph
__global__ void claim_reset_kernel(int *counter, int n) {
  if ((int)threadIdx.x < n) counter[threadIdx.x] = 0;
}

__global__ void zero_fill_kernel(double *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}


#define FAIL(ctx, code, msg)  \
  do {                        \
    (ctx)->err = (msg);       \
    return (code);            \
  } while (0)

#define HIPCHK(ctx, call)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                       \
      return GTK_EDEVICE;                                                                   \
    }                                                                                       \
  } while (0)

static bool over_cap(gtk_ctx *ctx, size_t bytes) {
  if (!ctx->alloc_cap || bytes <= ctx->alloc_cap) return false;
  ctx->err = "device allocation of " + std::to_string(bytes) + " bytes exceeds GTK_MEM_LIMIT_MB";
  return true;
}

template <typename T>
static int upload(gtk_ctx *ctx, DevBuf &b, const T *src, size_t count) {
  size_t bytes = count * sizeof(T);
  if (b.n < bytes || !b.p) {
    b.release();
    if (over_cap(ctx, bytes)) return GTK_ENOMEM;
    if (bytes) {
      hipError_t e = hipMalloc(&b.p, bytes);
      if (e != hipSuccess) {
        b.p = nullptr;
        ctx->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return GTK_ENOMEM;
      }
    }
    b.n = bytes;
  }
  if (bytes) HIPCHK(ctx, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return GTK_OK;
}

static int alloc(gtk_ctx *ctx, DevBuf &b, size_t bytes) {
  if (b.p && b.n >= bytes) return GTK_OK;
  b.release();
  if (!bytes) return GTK_OK;
  if (over_cap(ctx, bytes)) return GTK_ENOMEM;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    ctx->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return GTK_ENOMEM;
  }
  b.n = bytes;
  return GTK_OK;
}

static gtk::MeshView view(const gtk_ctx *ctx) {
  gtk::MeshView m;
  m.dim = ctx->dim;
  m.n_nodes = ctx->n_nodes;
  m.n_cells = ctx->n_cells;
  m.nodes_per_cell = ctx->nv;
  m.xyz = ctx->xyz.data();
  m.cell_nodes = ctx->cell_nodes.data();
  m.cell_dofs = ctx->cell_dofs.data();
  m.dofs_per_cell = ctx->dofs_per_cell;
  m.n_free = ctx->n_free;
  m.n_diri = ctx->n_diri;
  return m;
}

static gtk::ElemKey elem_key(const gtk_ctx *ctx) {
  return gtk::ElemKey{ctx->dim, ctx->nv, ctx->nshape, ctx->nq, ctx->ncomp, ctx->nfaces};
}

// LDS available to one workgroup (gfx950: 160 KiB per CU), less headroom.
static constexpr size_t kLdsBudget = 156 * 1024;
// LDS of a load-vector tile of its own launch: several workgroups share a
// CU and hide each other's gathers (Q2 hex 64^3 load vector 0.537 -> 0.417
// ms, Q1 quadrilaterals 512^2 0.031 -> 0.020 ms against tiles as large as
// LDS allows).  In a batch, where each workgroup holds a CU's LDS for the
// matrix tiles, the vector tiles stay that large (40 KB tiles there: the
// 256-problem batch 4.49 -> 5.73 ms).
static constexpr size_t kVecLdsBudget = 40 * 1024;
static constexpr int kBatchedPlan = 0x100;  // vplans key bit: the tiles of a batch

static size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// doubles past the values that receive the padding entries of a batch
// write-out (at most 16 rounds of 64)
static constexpr size_t kValueTail = 16 * 64;

// chunk: cells whose owned DOFs form a tile; about (c^(1/D)+1)^D cells are
// then needed (more around simplices)
static int tile_chunk(const gtk_ctx *ctx, int max_cells) {
  int chunk = 1;
  for (int c = 2; c <= 4096; c *= 2) {
    double side = std::pow((double)c, 1.0 / ctx->dim);
    double need = std::pow(side + 1.0, ctx->dim);
    if (ctx->nv == ctx->dim + 1) need *= 2.5;
    if (need <= max_cells) chunk = c;
  }
  return chunk;
}

// Largest cell count per tile whose staged geometry fits the LDS left by
// the tables and the accumulators of kernel k.
static int tile_cells_for(const gtk::KernelInfo &k, size_t acc_bytes, size_t budget) {
  const size_t fixed = align16(k.table_bytes) + acc_bytes + 1024;
  const size_t geo = std::max<size_t>(1, k.geo_bytes_per_cell);
  const size_t avail = budget > fixed ? budget - fixed : 8 * geo;
  return (int)std::max<size_t>(8, std::min<size_t>(4095, avail / geo));
}

// Tile limits of the vector kernel: per cell f . dV at the points, the node
// coordinates and the nodal values; tables phi and the geometry gradients.
static void vector_tile_limits(const gtk_ctx *ctx, size_t budget, int &max_cells, int &chunk) {
  const size_t fs = (size_t)ctx->nq * ctx->ncomp | 1;  // odd per-cell stride of f . dV
  const size_t geo =
      (fs + (size_t)ctx->nv * ctx->dim + (size_t)ctx->nshape * ctx->ncomp + (ctx->nfaces ? 1 : 0)) * 8;
  const size_t tables = (size_t)std::max(1, ctx->nfaces) * ctx->nq *
                        (((size_t)ctx->nshape | 1) + ((size_t)ctx->nv * ctx->dim | 1)) * 8;
  const size_t avail = budget > tables + 1024 ? budget - tables - 1024 : 8 * geo;
  max_cells = (int)std::max<size_t>(8, std::min<size_t>(4096, avail / geo));
  chunk = tile_chunk(ctx, max_cells);
}

extern "C" {

const char *gtk_version(void) { return "libgtk_hip gfx950 " __DATE__ " " __TIME__; }

int gtk_create(int device, gtk_ctx **out) {
  if (!out) {
    g_create_err = "out is NULL";
    return GTK_EINVAL;
  }
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    g_create_err = std::string("no HIP device: ") + hipGetErrorString(e);
    return GTK_EDEVICE;
  }
  if (device < 0 || device >= n) {
    g_create_err = "device ordinal out of range";
    return GTK_EINVAL;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    g_create_err = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return GTK_EDEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      g_create_err = std::string("libgtk_hip is built for gfx950, device is ") + prop.gcnArchName;
      return GTK_EUNSUP;
    }
  }
  gtk_ctx *ctx = new (std::nothrow) gtk_ctx();
  if (!ctx) {
    g_create_err = "out of host memory";
    return GTK_ENOMEM;
  }
  ctx->device = device;
  // matrix-core workgroup shape and LDS budget (tuning overrides)
  if (const char *w = std::getenv("GTK_WAVES")) {
    const long v = std::strtol(w, nullptr, 10);
    if (v == 8 || v == 12 || v == 16) ctx->waves = (int)v;
  }
  if (const char *l = std::getenv("GTK_LDS_KB")) {
    const long kb = std::strtol(l, nullptr, 10);
    if (kb >= 16 && kb <= 160) ctx->lds_budget = (size_t)kb * 1024;
  }
  if (const char *l = std::getenv("GTK_MEM_LIMIT_MB")) {
    const long mb = std::strtol(l, nullptr, 10);
    if (mb > 0) ctx->alloc_cap = (size_t)mb << 20;
  }
  if (resolve_scratch(device) != GTK_OK) {
    delete ctx;
    return GTK_EDEVICE;
  }
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
    g_create_err = "stream/event creation failed";
    delete ctx;
    return GTK_EDEVICE;
  }
  *out = ctx;
  return GTK_OK;
}

int gtk_destroy(gtk_ctx *ctx) {
  if (!ctx) return GTK_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->mplans.clear();
  ctx->vplans.clear();
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return GTK_OK;
}

const char *gtk_last_error(const gtk_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int gtk_set_mesh(gtk_ctx *ctx, int dim, int64_t n_nodes, const double *xyz, int64_t n_cells, int nodes_per_cell,
                 const int32_t *cell_nodes) {
  if (!ctx) return GTK_EINVAL;
  if (dim < 2 || dim > 3) FAIL(ctx, GTK_EINVAL, "dim must be 2 or 3");
  if (n_nodes <= 0 || n_cells <= 0 || !xyz || !cell_nodes) FAIL(ctx, GTK_EINVAL, "empty mesh or NULL array");
  if (nodes_per_cell < dim + 1 || nodes_per_cell > 16) FAIL(ctx, GTK_EINVAL, "nodes_per_cell out of range");
  if (n_nodes > INT32_MAX || n_cells > INT32_MAX) FAIL(ctx, GTK_EINVAL, "mesh too large for Int32 ids");
  std::vector<int32_t> cn((size_t)n_cells * nodes_per_cell);
  for (size_t i = 0; i < cn.size(); i++) {
    int32_t v = cell_nodes[i];
    if (v < 1 || v > n_nodes) FAIL(ctx, GTK_EINVAL, "cell node id out of range (ids are 1-based)");
    cn[i] = v - 1;
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  ctx->dim = dim;
  ctx->nv = nodes_per_cell;
  ctx->n_nodes = n_nodes;
  ctx->n_cells = n_cells;
  ctx->xyz.assign(xyz, xyz + n_nodes * dim);
  ctx->cell_nodes.swap(cn);
  int rc;
  if ((rc = upload(ctx, ctx->d_xyz, ctx->xyz.data(), ctx->xyz.size()))) return rc;
  if ((rc = upload(ctx, ctx->d_cell_nodes, ctx->cell_nodes.data(), ctx->cell_nodes.size()))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->has_mesh = true;
  ctx->has_space = false;
  ctx->generation++;
  ctx->mplans.clear();
  ctx->vplans.clear();
  return GTK_OK;
}

int gtk_update_coordinates(gtk_ctx *ctx, const double *xyz) {
  if (!ctx) return GTK_EINVAL;
  if (!ctx->has_mesh) FAIL(ctx, GTK_ESTATE, "gtk_set_mesh first");
  if (!xyz) FAIL(ctx, GTK_EINVAL, "xyz is NULL");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  ctx->xyz.assign(xyz, xyz + ctx->n_nodes * ctx->dim);
  int rc = upload(ctx, ctx->d_xyz, ctx->xyz.data(), ctx->xyz.size());
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return GTK_OK;
}

int gtk_set_space(gtk_ctx *ctx, int ncomp, int dofs_per_cell, int64_t n_free, int64_t n_diri,
                  const int32_t *cell_dofs) {
  if (!ctx) return GTK_EINVAL;
  if (!ctx->has_mesh) FAIL(ctx, GTK_ESTATE, "gtk_set_mesh first");
  if (ncomp < 1 || ncomp > 3 || dofs_per_cell < 1 || dofs_per_cell % ncomp || dofs_per_cell > 64)
    FAIL(ctx, GTK_EINVAL, "ncomp/dofs_per_cell out of range");
  if (n_free < 0 || n_diri < 0 || n_free + n_diri <= 0 || !cell_dofs) FAIL(ctx, GTK_EINVAL, "invalid dof counts");
  if (n_free > INT32_MAX || n_diri > INT32_MAX) FAIL(ctx, GTK_EINVAL, "dof count exceeds Int32");
  size_t n = (size_t)ctx->n_cells * dofs_per_cell;
  for (size_t i = 0; i < n; i++) {
    int32_t d = cell_dofs[i];
    if (d == 0 || d > n_free || -d > n_diri) FAIL(ctx, GTK_EINVAL, "cell dof out of range (GT signed encoding)");
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  ctx->ncomp = ncomp;
  ctx->dofs_per_cell = dofs_per_cell;
  ctx->n_free = n_free;
  ctx->n_diri = n_diri;
  ctx->cell_dofs.assign(cell_dofs, cell_dofs + n);
  int rc = upload(ctx, ctx->d_cell_dofs, ctx->cell_dofs.data(), n);
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->has_space = true;
  ctx->generation++;
  ctx->mplans.clear();
  ctx->vplans.clear();
  return GTK_OK;
}

int gtk_set_reference(gtk_ctx *ctx, int nq, const double *w, int nshape, const double *phi, const double *grad_phi,
                      const double *grad_geo) {
  if (!ctx) return GTK_EINVAL;
  if (!ctx->has_mesh) FAIL(ctx, GTK_ESTATE, "gtk_set_mesh first");
  if (nq < 1 || nq > 128 || nshape < 1 || nshape > 64 || !w || !phi || !grad_phi || !grad_geo)
    FAIL(ctx, GTK_EINVAL, "invalid reference tables");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int D = ctx->dim;
  ctx->nq = nq;
  ctx->nshape = nshape;
  ctx->w.assign(w, w + nq);
  ctx->phi.assign(phi, phi + (size_t)nq * nshape);
  ctx->gphi_t.assign((size_t)nq * nshape * D, 0.0);
  for (int q = 0; q < nq; q++)
    for (int s = 0; s < nshape; s++)
      for (int d = 0; d < D; d++) ctx->gphi_t[((size_t)q * D + d) * nshape + s] = grad_phi[((size_t)q * nshape + s) * D + d];
  ctx->gN.assign(grad_geo, grad_geo + (size_t)nq * ctx->nv * D);
  std::vector<double> rtab;
  if (ctx->nv == D + 1) {
    // in the order a kernel forming it per workgroup would: r = fma(w_q d_m
    // phi_a, d_n phi_b, r) over q ascending
    rtab.assign((size_t)nshape * nshape * D * D, 0.0);
    for (size_t i = 0; i < rtab.size(); i++) {
      const size_t ab = i / (D * D), mn = i % (D * D), ia = ab / nshape, ib = ab % nshape, m = mn / D, n = mn % D;
      double r = 0.0;
      for (int q = 0; q < nq; q++)
        r = std::fma(ctx->w[q] * ctx->gphi_t[((size_t)q * D + m) * nshape + ia], ctx->gphi_t[((size_t)q * D + n) * nshape + ib], r);
      rtab[i] = r;
    }
  }
  ctx->nfaces = 0;
  ctx->has_face_grad = false;
  for (auto &mp : ctx->mplans) mp.second->exec.clear();  // kernels depend on the tables
  ctx->vplans.clear();
  int rc;
  if ((rc = upload(ctx, ctx->d_w, ctx->w.data(), ctx->w.size()))) return rc;
  if ((rc = upload(ctx, ctx->d_phi, ctx->phi.data(), ctx->phi.size()))) return rc;
  if ((rc = upload(ctx, ctx->d_gphi, ctx->gphi_t.data(), ctx->gphi_t.size()))) return rc;
  if ((rc = upload(ctx, ctx->d_gN, ctx->gN.data(), ctx->gN.size()))) return rc;
  if (!rtab.empty() && (rc = upload(ctx, ctx->d_rtab, rtab.data(), rtab.size()))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->has_ref = true;
  ctx->generation++;
  return GTK_OK;
}

int gtk_set_reference_faces(gtk_ctx *ctx, int nfaces, int nq, const double *w, int nshape, const double *phi,
                            const double *grad_geo, const double *face_jac, const int32_t *item_face) {
  if (!ctx) return GTK_EINVAL;
  if (!ctx->has_mesh) FAIL(ctx, GTK_ESTATE, "gtk_set_mesh first");
  if (nfaces < 1 || nfaces > 16 || nq < 1 || nq > 128 || nshape < 1 || nshape > 64 || !w || !phi || !grad_geo ||
      !face_jac || !item_face)
    FAIL(ctx, GTK_EINVAL, "invalid face tables");
  const int D = ctx->dim;
  bool packed = false;
  for (int64_t i = 0; i < ctx->n_cells; i++) {
    const int32_t lo = item_face[i] & 0xFFFF, hi = item_face[i] >> 16;
    if (item_face[i] < 0 || lo >= nfaces || hi >= nfaces) FAIL(ctx, GTK_EINVAL, "item_face out of range");
    packed = packed || hi != 0;
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  ctx->nfaces = nfaces;
  ctx->nq = nq;
  ctx->nshape = nshape;
  ctx->w.assign(w, w + nq);
  ctx->phi.assign(phi, phi + (size_t)nfaces * nq * nshape);
  ctx->gphi_t.clear();
  ctx->gN.assign(grad_geo, grad_geo + (size_t)nfaces * nq * ctx->nv * D);
  ctx->face_jac.assign(face_jac, face_jac + (size_t)nfaces * D * (D - 1));
  ctx->item_face.assign(item_face, item_face + ctx->n_cells);
  ctx->faces_packed = packed;
  for (auto &mp : ctx->mplans) mp.second->exec.clear();
  ctx->vplans.clear();
  int rc;
  if ((rc = upload(ctx, ctx->d_w, ctx->w.data(), ctx->w.size()))) return rc;
  if ((rc = upload(ctx, ctx->d_phi, ctx->phi.data(), ctx->phi.size()))) return rc;
  if ((rc = upload(ctx, ctx->d_gN, ctx->gN.data(), ctx->gN.size()))) return rc;
  if ((rc = upload(ctx, ctx->d_face_jac, ctx->face_jac.data(), ctx->face_jac.size()))) return rc;
  if ((rc = upload(ctx, ctx->d_item_face, ctx->item_face.data(), ctx->item_face.size()))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->has_ref = true;
  ctx->has_face_grad = false;
  ctx->generation++;
  return GTK_OK;
}

int gtk_set_reference_face_gradients(gtk_ctx *ctx, const double *grad_phi, const double *normals) {
  if (!ctx) return GTK_EINVAL;
  if (!ctx->has_ref || !ctx->nfaces) FAIL(ctx, GTK_ESTATE, "gtk_set_reference_faces first");
  if (!grad_phi || !normals) FAIL(ctx, GTK_EINVAL, "grad_phi and normals are required");
  const int D = ctx->dim, L = ctx->nfaces, nq = ctx->nq, ns = ctx->nshape;
  for (int l = 0; l < L; l++) {
    double nn = 0.0;
    for (int d = 0; d < D; d++) nn += normals[l * D + d] * normals[l * D + d];
    if (!(std::fabs(nn - 1.0) < 1e-12)) FAIL(ctx, GTK_EINVAL, "reference normals must have unit length");
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // [face][q][s][d] -> [face][q][d][s], the kernels' layout
  ctx->gphi_t.assign((size_t)L * nq * D * ns, 0.0);
  for (int l = 0; l < L; l++)
    for (int q = 0; q < nq; q++)
      for (int s = 0; s < ns; s++)
        for (int d = 0; d < D; d++)
          ctx->gphi_t[(((size_t)l * nq + q) * D + d) * ns + s] = grad_phi[(((size_t)l * nq + q) * ns + s) * D + d];
  int rc;
  if ((rc = upload(ctx, ctx->d_gphi, ctx->gphi_t.data(), ctx->gphi_t.size()))) return rc;
  if ((rc = upload(ctx, ctx->d_face_normal, normals, (size_t)L * D))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (auto &mp : ctx->mplans) mp.second->exec.clear();
  ctx->has_face_grad = true;
  ctx->generation++;
  return GTK_OK;
}

int gtk_set_cell_coefficients(gtk_ctx *ctx, int ncoef, const double *coef) {
  if (!ctx) return GTK_EINVAL;
  if (!ctx->has_mesh) FAIL(ctx, GTK_ESTATE, "gtk_set_mesh first");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (!coef) {
    ctx->ncoef = 0;
    return GTK_OK;
  }
  if (ncoef < 1 || ncoef > 2) FAIL(ctx, GTK_EINVAL, "ncoef must be 1 or 2");
  int rc = upload(ctx, ctx->d_coef, coef, (size_t)ctx->n_cells * ncoef);
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->ncoef = ncoef;
  return GTK_OK;
}

static int key_of(int rows, int cols) { return rows * 4 + cols; }

static int upload_tiles(gtk_ctx *ctx, const gtk::TilePlan &t, DevBuf &tcp, DevBuf &tc, DevBuf &tcol, DevBuf &cid,
                        DevBuf &ccp, DevBuf &cs, DevBuf &cb) {
  int rc;
  if ((rc = upload(ctx, tcp, t.tile_cell_ptr.data(), t.tile_cell_ptr.size()))) return rc;
  if ((rc = upload(ctx, tc, t.tile_cells.data(), t.tile_cells.size()))) return rc;
  if ((rc = upload(ctx, tcol, t.tile_col_ptr.data(), t.tile_col_ptr.size()))) return rc;
  if ((rc = upload(ctx, cid, t.col_id.data(), t.col_id.size()))) return rc;
  if ((rc = upload(ctx, ccp, t.col_ctr_ptr.data(), t.col_ctr_ptr.size()))) return rc;
  if ((rc = upload(ctx, cs, t.ctr_slot.data(), t.ctr_slot.size()))) return rc;
  if ((rc = upload(ctx, cb, t.ctr_b.data(), t.ctr_b.size()))) return rc;
  return GTK_OK;
}

int gtk_build_pattern(gtk_ctx *ctx, int rows, int cols, int64_t *nnz, int64_t *ncols) {
  if (!ctx) return GTK_EINVAL;
  if (!ctx->has_space) FAIL(ctx, GTK_ESTATE, "gtk_set_space first");
  if ((rows != GTK_FREE && rows != GTK_DIRICHLET) || (cols != GTK_FREE && cols != GTK_DIRICHLET))
    FAIL(ctx, GTK_EINVAL, "rows/cols must be GTK_FREE or GTK_DIRICHLET");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto plan = std::make_unique<MatrixPlan>();
  gtk::MeshView m = view(ctx);
  int nt = nthreads_host();
  std::string e = gtk::build_pattern(m, rows, cols, plan->pat, nt);
  if (!e.empty()) FAIL(ctx, GTK_EINVAL, e);
  for (int64_t j = 0; j < plan->pat.ncols; j++)
    plan->max_seg = std::max(plan->max_seg, plan->pat.colptr[j + 1] - plan->pat.colptr[j]);
  int rc;
  if ((rc = upload(ctx, plan->colptr, plan->pat.colptr.data(), plan->pat.colptr.size()))) return rc;
  // values plus a tail that absorbs the padding entries of the batch write-out
  if ((rc = alloc(ctx, plan->nzval, ((size_t)plan->pat.nnz + kValueTail) * sizeof(double)))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  plan->valid = true;
  if (nnz) *nnz = plan->pat.nnz;
  if (ncols) *ncols = plan->pat.ncols;
  ctx->mplans[key_of(rows, cols)] = std::move(plan);
  ctx->generation++;
  return GTK_OK;
}

int gtk_build_pattern_device(gtk_ctx *ctx, int rows, int cols, int64_t *nnz, int64_t *ncols, double *device_ms) {
  if (!ctx) return GTK_EINVAL;
  if (!ctx->has_space) FAIL(ctx, GTK_ESTATE, "gtk_set_space first");
  if ((rows != GTK_FREE && rows != GTK_DIRICHLET) || (cols != GTK_FREE && cols != GTK_DIRICHLET))
    FAIL(ctx, GTK_EINVAL, "rows/cols must be GTK_FREE or GTK_DIRICHLET");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto plan = std::make_unique<MatrixPlan>();
  gtk::Pattern &pat = plan->pat;
  pat.rows = rows;
  pat.cols = cols;
  pat.nrows = rows == GTK_FREE ? ctx->n_free : ctx->n_diri;
  pat.ncols = cols == GTK_FREE ? ctx->n_free : ctx->n_diri;
  const int64_t nc = pat.ncols, ncell = ctx->n_cells;
  const int nd = ctx->dofs_per_cell;
  hipStream_t st = ctx->stream;
  const int32_t *cd = ctx->d_cell_dofs.as<int32_t>();
  DevBuf counts, ptr, tmp, fill, inc, len, maxlist, c64;
  int rc;
  const size_t nc1 = (size_t)std::max<int64_t>(1, nc);
  if ((rc = alloc(ctx, counts, nc1 * sizeof(int32_t))) || (rc = alloc(ctx, fill, nc1 * sizeof(int32_t))) ||
      (rc = alloc(ctx, len, nc1 * sizeof(int32_t))) || (rc = alloc(ctx, ptr, (nc1 + 1) * sizeof(int64_t))) ||
      (rc = alloc(ctx, c64, (nc1 + 1) * sizeof(int64_t))) || (rc = alloc(ctx, maxlist, 2 * sizeof(int32_t))) ||
      (rc = alloc(ctx, tmp, gtk::symbolic_scan_tmp(nc) * sizeof(int64_t))))
    return rc;
  HIPCHK(ctx, hipEventRecord(ctx->ev0, st));
  HIPCHK(ctx, hipMemsetAsync(counts.p, 0, nc1 * sizeof(int32_t), st));
  HIPCHK(ctx, hipMemsetAsync(fill.p, 0, nc1 * sizeof(int32_t), st));
  // maxlist[0]: longest candidate list of a skipped column; [1]: most cells of a column
  HIPCHK(ctx, hipMemsetAsync(maxlist.p, 0, 2 * sizeof(int32_t), st));
  gtk::launch_column_count(cd, ncell, nd, cols, counts.as<int32_t>(), st);
  gtk::launch_max_count(counts.as<int32_t>(), nc, maxlist.as<int32_t>() + 1, st);
  gtk::launch_scan(counts.as<int32_t>(), nc, 0, ptr.as<int64_t>(), tmp.as<int64_t>(), st);
  int64_t ninc = 0;
  int32_t max_cells = 0;
  HIPCHK(ctx, hipMemcpyAsync(&ninc, ptr.as<int64_t>() + nc, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipMemcpyAsync(&max_cells, maxlist.as<int32_t>() + 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
  const int64_t max_list = (int64_t)max_cells * nd;
  if (max_list > gtk::kSymbolicMaxList)
    FAIL(ctx, GTK_EUNSUP, "a column gathers " + std::to_string(max_list) + " candidate rows, more than the " +
                              std::to_string(gtk::kSymbolicMaxList) + " the device symbolic phase sorts: use gtk_build_pattern");
  int cap = 64;  // LDS entries of the sort: the longest list's power of two
  while (cap < max_list) cap <<= 1;
  if ((rc = alloc(ctx, inc, (size_t)std::max<int64_t>(1, ninc) * sizeof(int32_t)))) return rc;
  gtk::launch_column_fill(cd, ncell, nd, cols, ptr.as<int64_t>(), fill.as<int32_t>(), inc.as<int32_t>(), st);
  gtk::launch_column_rows(cd, nd, rows, nc, ptr.as<int64_t>(), inc.as<int32_t>(), 0, cap, len.as<int32_t>(),
                          maxlist.as<int32_t>(), nullptr, nullptr, st);
  gtk::launch_scan(len.as<int32_t>(), nc, 1, c64.as<int64_t>(), tmp.as<int64_t>(), st);
  int32_t too_long = 0;
  int64_t end = 1;
  HIPCHK(ctx, hipMemcpyAsync(&too_long, maxlist.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipMemcpyAsync(&end, c64.as<int64_t>() + nc, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
  if (too_long)
    FAIL(ctx, GTK_EUNSUP, "a column gathers " + std::to_string(too_long) + " candidate rows, more than the " +
                              std::to_string(gtk::kSymbolicMaxList) + " the device symbolic phase sorts: use gtk_build_pattern");
  if (end - 1 > INT32_MAX - 1) FAIL(ctx, GTK_EINVAL, "nnz exceeds the Int32 range of SparseMatrixCSC{Float64,Int32}");
  pat.nnz = end - 1;
  if ((rc = alloc(ctx, plan->rowval, (size_t)std::max<int64_t>(1, pat.nnz) * sizeof(int32_t))) ||
      (rc = alloc(ctx, plan->colptr, (nc1 + 1) * sizeof(int32_t))))
    return rc;
  gtk::launch_column_rows(cd, nd, rows, nc, ptr.as<int64_t>(), inc.as<int32_t>(), 1, cap, nullptr, maxlist.as<int32_t>(),
                          c64.as<int64_t>(), plan->rowval.as<int32_t>(), st);
  gtk::launch_colptr32(c64.as<int64_t>(), nc, plan->colptr.as<int32_t>(), st);
  HIPCHK(ctx, hipEventRecord(ctx->ev1, st));
  // host copy: the tile and batch planning and gtk_get_pattern read it
  pat.colptr.resize(nc + 1);
  pat.rowval.resize(pat.nnz);
  HIPCHK(ctx, hipMemcpyAsync(pat.colptr.data(), plan->colptr.p, (nc + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if (pat.nnz)
    HIPCHK(ctx, hipMemcpyAsync(pat.rowval.data(), plan->rowval.p, pat.nnz * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if ((rc = alloc(ctx, plan->nzval, ((size_t)pat.nnz + kValueTail) * sizeof(double)))) return rc;
  HIPCHK(ctx, hipStreamSynchronize(st));
  if (device_ms) {
    float ms = 0;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *device_ms = ms;
  }
  for (int64_t j = 0; j < nc; j++) plan->max_seg = std::max(plan->max_seg, pat.colptr[j + 1] - pat.colptr[j]);
  plan->valid = true;
  if (nnz) *nnz = pat.nnz;
  if (ncols) *ncols = pat.ncols;
  ctx->mplans[key_of(rows, cols)] = std::move(plan);
  ctx->generation++;
  return GTK_OK;
}

int gtk_add_block(gtk_ctx *dst, int rows, int cols, gtk_ctx *src) {
  if (!dst || !src) return GTK_EINVAL;
  if (dst->device != src->device) FAIL(dst, GTK_EINVAL, "the contexts are on different devices");
  auto id = dst->mplans.find(key_of(rows, cols));
  auto is = src->mplans.find(key_of(rows, cols));
  if (id == dst->mplans.end() || is == src->mplans.end()) FAIL(dst, GTK_ESTATE, "both blocks must be built");
  MatrixPlan &D = *id->second, &S = *is->second;
  if (D.pat.ncols != S.pat.ncols || D.pat.nrows != S.pat.nrows) FAIL(dst, GTK_EINVAL, "the blocks differ in shape");
  HIPCHK(dst, hipSetDevice(dst->device));
  int rc;
  if (!D.rowval.p && D.pat.nnz)
    if ((rc = upload(dst, D.rowval, D.pat.rowval.data(), D.pat.rowval.size()))) return rc;
  if (!S.rowval.p && S.pat.nnz)
    if ((rc = upload(dst, S.rowval, S.pat.rowval.data(), S.pat.rowval.size()))) return rc;
  HIPCHK(dst, hipStreamSynchronize(src->stream));  // the source values are complete
  DevBuf missing;
  if ((rc = alloc(dst, missing, sizeof(int32_t)))) return rc;
  HIPCHK(dst, hipMemsetAsync(missing.p, 0, sizeof(int32_t), dst->stream));
  gtk::launch_add_block(S.pat.ncols, S.colptr.as<int32_t>(), S.rowval.as<int32_t>(), S.nzval.as<double>(),
                        D.colptr.as<int32_t>(), D.rowval.as<int32_t>(), D.nzval.as<double>(), missing.as<int32_t>(),
                        dst->stream);
  int32_t miss = 0;
  HIPCHK(dst, hipMemcpyAsync(&miss, missing.p, sizeof(int32_t), hipMemcpyDeviceToHost, dst->stream));
  HIPCHK(dst, hipStreamSynchronize(dst->stream));
  if (miss) FAIL(dst, GTK_EINVAL, "the source block has entries outside the destination pattern");
  return GTK_OK;
}

int gtk_gather_block(gtk_ctx *dst, int rows, int cols, gtk_ctx *src, int src_rows, int src_cols, const int64_t *ptr,
                     const int64_t *idx) {
  if (!dst || !src || !ptr) return GTK_EINVAL;
  if (dst->device != src->device) FAIL(dst, GTK_EINVAL, "the contexts are on different devices");
  auto id = dst->mplans.find(key_of(rows, cols));
  auto is = src->mplans.find(key_of(src_rows, src_cols));
  if (id == dst->mplans.end() || is == src->mplans.end()) FAIL(dst, GTK_ESTATE, "both blocks must be built");
  MatrixPlan *D = id->second.get(), *S = is->second.get();
  const int64_t n = D->pat.nnz, m = ptr[n];
  if (ptr[0] != 0 || m < 0 || (m > 0 && !idx)) FAIL(dst, GTK_EINVAL, "invalid grouping");
  for (int64_t t = 0; t < n; t++)
    if (ptr[t + 1] < ptr[t]) FAIL(dst, GTK_EINVAL, "grouping pointers must not decrease");
  for (int64_t k = 0; k < m; k++)
    if (idx[k] < 0 || idx[k] >= S->pat.nnz) FAIL(dst, GTK_EINVAL, "source index out of range");
  HIPCHK(dst, hipSetDevice(dst->device));
  HIPCHK(dst, hipStreamSynchronize(src->stream));  // the source values are complete
  DevBuf dptr, didx;
  int rc;
  if ((rc = upload(dst, dptr, ptr, (size_t)n + 1))) return rc;
  if (m && (rc = upload(dst, didx, idx, (size_t)m))) return rc;
  gtk::launch_gather_block(n, dptr.as<int64_t>(), didx.as<int64_t>(), S->nzval.as<double>(), D->nzval.as<double>(),
                           dst->stream);
  HIPCHK(dst, hipStreamSynchronize(dst->stream));
  return GTK_OK;
}

int gtk_local_nnz_map(gtk_ctx *ctx, int rows, int cols, int32_t *map_out) {
  if (!ctx || !map_out) return GTK_EINVAL;
  auto it = ctx->mplans.find(key_of(rows, cols));
  if (it == ctx->mplans.end()) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern first");
  MatrixPlan &P = *it->second;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc;
  if (!P.rowval.p && P.pat.nnz)
    if ((rc = upload(ctx, P.rowval, P.pat.rowval.data(), P.pat.rowval.size()))) return rc;
  const size_t n = (size_t)ctx->n_cells * ctx->dofs_per_cell * ctx->dofs_per_cell;
  DevBuf map;
  if ((rc = alloc(ctx, map, std::max<size_t>(1, n) * sizeof(int32_t)))) return rc;
  gtk::launch_local_map(ctx->d_cell_dofs.as<int32_t>(), ctx->n_cells, ctx->dofs_per_cell, rows, cols,
                        P.colptr.as<int32_t>(), P.rowval.as<int32_t>(), map.as<int32_t>(), ctx->stream);
  if (n) HIPCHK(ctx, hipMemcpyAsync(map_out, map.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return GTK_OK;
}

int gtk_get_pattern(gtk_ctx *ctx, int rows, int cols, int32_t *colptr, int32_t *rowval) {
  if (!ctx) return GTK_EINVAL;
  auto it = ctx->mplans.find(key_of(rows, cols));
  if (it == ctx->mplans.end()) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern first");
  const gtk::Pattern &p = it->second->pat;
  if (colptr) std::memcpy(colptr, p.colptr.data(), p.colptr.size() * sizeof(int32_t));
  if (rowval) std::memcpy(rowval, p.rowval.data(), p.rowval.size() * sizeof(int32_t));
  return GTK_OK;
}

// Work plan of (form, precision) over a block, built on first use: the
// matrix-core kernel when the space is scalar and no column has more than
// kBatchSlots contributors, the VALU kernel otherwise.
// Builds the work plan of (form, precision) over a block for a kernel of
// `waves` wavefronts; max_cells > 0 caps the cells of a tile (tuning).
static int build_exec(gtk_ctx *ctx, MatrixPlan &plan, int form, int precision, int waves, int max_cells_cap,
                      std::unique_ptr<ExecPlan> &out) {
  if (!ctx->has_ref) FAIL(ctx, GTK_ESTATE, "gtk_set_reference first");
  auto ep = std::make_unique<ExecPlan>();
  const gtk::MeshView m = view(ctx);
  const int nt = nthreads_host();
  const int cols = plan.pat.cols;
  const int off_bytes = plan.max_seg < 0xFF ? 1 : 2;
  const gtk::ElemKey key_e = elem_key(ctx);
  const int nd = ctx->dofs_per_cell;
  std::string e;
  // GTK_MATRIX_KERNEL=valu forces the VALU kernel (measurement)
  const char *kenv = std::getenv("GTK_MATRIX_KERNEL");
  const bool allow_mfma = !(kenv && std::strcmp(kenv, "valu") == 0);
  bool mfma = allow_mfma && gtk::find_matrix_kernel(key_e, form, precision, off_bytes, true, waves, ep->k) &&
              ep->k.mfma;
  if (mfma) {
    // accumulator bound: a batch adds at most nd rows per slot, and the
    // write-out reads whole rounds of 64
    const size_t acc_bound = (size_t)waves * (64 * ((gtk::kBatchSlots * nd + 63) / 64)) * sizeof(double);
    int max_cells = tile_cells_for(ep->k, acc_bound, ctx->lds_budget);
    if (max_cells_cap > 0) max_cells = std::max(8, std::min(max_cells, max_cells_cap));
    e = gtk::build_tiles(m, cols, max_cells, 1 << 20, 1 << 20, tile_chunk(ctx, max_cells), ep->tiles, nt);
    if (!e.empty()) FAIL(ctx, GTK_EINVAL, e);
    if (ep->tiles.max_ctr > gtk::kBatchSlots) mfma = false;
  }
  if (mfma) {
    e = gtk::build_batches(m, plan.pat, ep->tiles, ep->bat, nt);
    if (!e.empty()) FAIL(ctx, GTK_EINVAL, e);
    if (ep->bat.endmask_words > 16) {  // the kernel reads at most 16 end-mask words
      ep->bat = gtk::BatchPlan();
      mfma = false;
    } else if (ep->bat.rounds != ep->k.rounds) {
      FAIL(ctx, GTK_EINVAL, "batch records and kernel disagree on the write-out rounds");
    }
  }
  if (mfma) {
    ep->lds_geo = (int)align16(ep->k.table_bytes);
    ep->acc_offset = (int)align16(ep->lds_geo + (size_t)ep->tiles.max_tile_cells * ep->k.geo_bytes_per_cell);
    // batch entries; at least the write-out rounds
    ep->acc_per_wave = (std::max(ep->bat.max_batch_acc, 64 * ep->bat.rounds) + 1) & ~1;
    const size_t acc_bytes = (size_t)waves * ep->acc_per_wave * sizeof(double);
    // staged coordinates and coefficient of a cell; staging reuses the
    // accumulators and the LDS the budget leaves after them, so that a tile
    // is staged in as few passes (each a gather and two barriers) as fit
    const size_t stage_per_cell = ((size_t)ctx->nv * ctx->dim + 1) * sizeof(double);
    const size_t budget = std::max<size_t>(ctx->lds_budget, (size_t)ep->acc_offset + acc_bytes + 16);
    const size_t stage_room = std::max(acc_bytes, budget - ep->acc_offset - 16);
    ep->stage_cells = (int)std::max<size_t>(
        1, std::min<size_t>(std::max(ep->tiles.max_tile_cells, 1), stage_room / stage_per_cell));
    const size_t stage_bytes = std::max(acc_bytes, (size_t)ep->stage_cells * stage_per_cell);
    ep->lds_claim = (int)(ep->acc_offset + align16(stage_bytes));
    ep->lds = ep->lds_claim + 16;
  } else {
    if (!gtk::find_matrix_kernel(key_e, form, precision, off_bytes, false, 0, ep->k))
      FAIL(ctx, GTK_EUNSUP, "no kernel for this (element, quadrature, form, precision)");
    auto acc_of = [&](const gtk::KernelInfo &k) {
      return (size_t)k.block / 64 * k.cols_per_wave * std::max(1, plan.max_seg) * sizeof(double);
    };
    // long columns: the preferred workgroup's accumulators would leave no
    // room for a tile of 8 cells; then the 4-wave workgroup
    // (GTK_VALU_WAVES=4 forces it: tests)
    const char *vw = std::getenv("GTK_VALU_WAVES");
    if (waves == 4 || (vw && std::atoi(vw) == 4) ||
        align16(ep->k.table_bytes) + acc_of(ep->k) + 1024 + 8 * ep->k.geo_bytes_per_cell > kLdsBudget) {
      gtk::KernelInfo k4;
      if (gtk::find_matrix_kernel(key_e, form, precision, off_bytes, false, 4, k4)) ep->k = k4;
    }
    const size_t acc_bytes = acc_of(ep->k);
    int max_cells = tile_cells_for(ep->k, acc_bytes, kLdsBudget);
    bool built = false;
    if (max_cells_cap > 0) {
      max_cells = std::max(8, std::min(max_cells, max_cells_cap));
    } else {
      // one workgroup per tile: a small problem is first cut into at least
      // two tiles per CU (tiles of at least 8 cells) so that it is not run
      // by a few workgroups; where a DOF touches more cells than that, the
      // LDS-bound tiles (512^2 Q2 quadrilaterals, mixed: the 2,048 Nitsche
      // items 0.233 -> 0.007 ms, the 523,264 skeleton items 1.04 -> 0.73 ms)
      int cus = 0;
      HIPCHK(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
      const int64_t spread = (ctx->n_cells + 2 * (int64_t)std::max(cus, 1) - 1) / (2 * (int64_t)std::max(cus, 1));
      const int small = (int)std::max<int64_t>(8, std::min<int64_t>(max_cells, spread));
      if (small < max_cells) {
        e = gtk::build_tiles(m, cols, small, 1 << 20, 1 << 20, tile_chunk(ctx, small), ep->tiles, nt);
        built = e.empty();
      }
    }
    if (!built) {
      e = gtk::build_tiles(m, cols, max_cells, 1 << 20, 1 << 20, tile_chunk(ctx, max_cells), ep->tiles, nt);
      if (!e.empty()) FAIL(ctx, GTK_EINVAL, e);
    }
    e = gtk::build_offsets(m, plan.pat, ep->tiles, ep->offs, nt);
    if (!e.empty()) FAIL(ctx, GTK_EINVAL, e);
    ep->acc_offset = (int)align16(ep->k.table_bytes + (size_t)ep->tiles.max_tile_cells * ep->k.geo_bytes_per_cell);
    ep->acc_per_wave = ep->k.cols_per_wave * std::max(1, plan.max_seg);
    ep->lds = ep->acc_offset + acc_bytes;
  }
  if (ep->lds > 160 * 1024) FAIL(ctx, GTK_EUNSUP, "tile does not fit in LDS");
  ep->tile_cap = max_cells_cap > 0 ? max_cells_cap : 0;
  const gtk::TilePlan &t = ep->tiles;
  int rc;
  if ((rc = upload(ctx, ep->tile_cell_ptr, t.tile_cell_ptr.data(), t.tile_cell_ptr.size()))) return rc;
  if ((rc = upload(ctx, ep->tile_cells, t.tile_cells.data(), t.tile_cells.size()))) return rc;
  if (mfma) {
    if ((rc = upload(ctx, ep->tile_batch_ptr, t.tile_batch_ptr.data(), t.tile_batch_ptr.size()))) return rc;
    if ((rc = upload(ctx, ep->rec, ep->bat.rec.data(), ep->bat.rec.size()))) return rc;
  } else {
    if ((rc = upload(ctx, ep->tile_col_ptr, t.tile_col_ptr.data(), t.tile_col_ptr.size()))) return rc;
    if ((rc = upload(ctx, ep->col_id, t.col_id.data(), t.col_id.size()))) return rc;
    if ((rc = upload(ctx, ep->col_ctr_ptr, t.col_ctr_ptr.data(), t.col_ctr_ptr.size()))) return rc;
    if ((rc = upload(ctx, ep->ctr_slot, t.ctr_slot.data(), t.ctr_slot.size()))) return rc;
    if ((rc = upload(ctx, ep->ctr_b, t.ctr_b.data(), t.ctr_b.size()))) return rc;
    if ((rc = upload(ctx, ep->ctr_off, ep->offs.data.data(), ep->offs.data.size()))) return rc;
  }
  if (ep->lds > 64 * 1024)
    HIPCHK(ctx, hipFuncSetAttribute(ep->k.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ep->lds));
  if (ep->k.mfma) {
    int per_cu = 0, cus = 0;
    HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ep->k.fn, ep->k.block, ep->lds));
    HIPCHK(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    ep->resident = std::max(1, per_cu) * std::max(1, cus);
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  out = std::move(ep);
  return GTK_OK;
}

// Default wavefronts per workgroup of the matrix-core kernel of (form,
// precision) on a block: the most (16: four per SIMD within 128 registers;
// 12: three within 168) whose instantiation needs no scratch; else 8 (two
// per SIMD within 256).  The context's GTK_WAVES setting wins.
static int default_waves(gtk_ctx *ctx, const MatrixPlan &plan, int form, int precision) {
  if (ctx->waves) return ctx->waves;
  const int off_bytes = plan.max_seg < 0xFF ? 1 : 2;
  for (const int w : {16, 12}) {
    gtk::KernelInfo kw;
    hipFuncAttributes fa;
    if (gtk::find_matrix_kernel(elem_key(ctx), form, precision, off_bytes, true, w, kw) && kw.mfma &&
        hipFuncGetAttributes(&fa, kw.fn) == hipSuccess && fa.localSizeBytes == 0)
      return w;
  }
  return 8;
}

// Work plan of (form, precision) over a block, built on first use with the
// default workgroup shape of the precision.
static int get_exec(gtk_ctx *ctx, MatrixPlan &plan, int form, int precision, ExecPlan *&out) {
  const int key = form * 2 + precision;
  auto it = plan.exec.find(key);
  if (it != plan.exec.end()) {
    out = it->second.get();
    return GTK_OK;
  }
  std::unique_ptr<ExecPlan> ep;
  int rc = build_exec(ctx, plan, form, precision, default_waves(ctx, plan, form, precision), 0, ep);
  if (rc) return rc;
  out = ep.get();
  plan.exec[key] = std::move(ep);
  return GTK_OK;
}

// Kernel arguments of a work plan.
// Tile plan and output of the load vector over a rows set, built on first use.
static int get_vector_plan(gtk_ctx *ctx, int rows, VectorPlan *&out, bool batched = false) {
  auto &slot = ctx->vplans[rows | (batched ? kBatchedPlan : 0)];
  if (!slot) {
    auto vp = std::make_unique<VectorPlan>();
    int max_cells, chunk;
    vector_tile_limits(ctx, batched ? kLdsBudget : kVecLdsBudget, max_cells, chunk);
    std::string e = gtk::build_tiles(view(ctx), rows, max_cells, 1 << 20, 1 << 20, chunk, vp->tiles, nthreads_host());
    if (!e.empty()) FAIL(ctx, GTK_EINVAL, e);
    int rc = upload_tiles(ctx, vp->tiles, vp->tile_cell_ptr, vp->tile_cells, vp->tile_col_ptr, vp->col_id,
                          vp->col_ctr_ptr, vp->ctr_slot, vp->ctr_b);
    if (rc) return rc;
    vp->nrows = rows == GTK_FREE ? ctx->n_free : ctx->n_diri;
    if ((rc = alloc(ctx, vp->out, (size_t)std::max<int64_t>(1, vp->nrows) * sizeof(double)))) return rc;
    vp->valid = true;
    slot = std::move(vp);
  }
  out = slot.get();
  return GTK_OK;
}

// Values of a load of kind fkind: one per component, per cell and component,
// or per DOF (free then Dirichlet).
static size_t load_count(const gtk_ctx *ctx, int fkind) {
  return fkind == GTK_F_CONST ? (size_t)ctx->ncomp
                              : (fkind == GTK_F_CELL ? (size_t)ctx->n_cells * ctx->ncomp
                                                     : (size_t)(ctx->n_free + ctx->n_diri));
}

static gtk::VectorArgs vector_args(const gtk_ctx *ctx, const VectorPlan &vp, const double *f) {
  gtk::VectorArgs a;
  std::memset(&a, 0, sizeof(a));
  a.xyz = ctx->d_xyz.as<double>();
  a.cell_nodes = ctx->d_cell_nodes.as<int32_t>();
  a.cell_dofs = ctx->d_cell_dofs.as<int32_t>();
  a.f = f;
  a.n_free = ctx->n_free;
  a.w = ctx->d_w.as<double>();
  a.phi = ctx->d_phi.as<double>();
  a.gN = ctx->d_gN.as<double>();
  a.tile_cell_ptr = vp.tile_cell_ptr.as<int32_t>();
  a.tile_cells = vp.tile_cells.as<int32_t>();
  a.tile_col_ptr = vp.tile_col_ptr.as<int32_t>();
  a.col_id = vp.col_id.as<int32_t>();
  a.col_ctr_ptr = vp.col_ctr_ptr.as<int32_t>();
  a.ctr_slot = vp.ctr_slot.as<uint16_t>();
  a.ctr_b = vp.ctr_b.as<uint16_t>();
  a.out = vp.out.as<double>();
  if (ctx->nfaces) {
    a.item_face = ctx->d_item_face.as<int32_t>();
    a.face_jac = ctx->d_face_jac.as<double>();
  }
  return a;
}

static gtk::MatrixArgs matrix_args(const gtk_ctx *ctx, const MatrixPlan &plan, const ExecPlan &e) {
  const ExecPlan *ep = &e;
  gtk::MatrixArgs a;
  std::memset(&a, 0, sizeof(a));  // padding too: gtk_batch_run compares records bytewise
  a.xyz = ctx->d_xyz.as<double>();
  a.cell_nodes = ctx->d_cell_nodes.as<int32_t>();
  a.coef = ctx->ncoef ? ctx->d_coef.as<double>() : nullptr;
  a.w = ctx->d_w.as<double>();
  a.phi = ctx->d_phi.as<double>();
  a.gphi = ctx->d_gphi.as<double>();
  a.gN = ctx->d_gN.as<double>();
  a.rtab = ctx->d_rtab.as<double>();
  a.tile_cell_ptr = ep->tile_cell_ptr.as<int32_t>();
  a.tile_cells = ep->tile_cells.as<int32_t>();
  a.tile_col_ptr = ep->tile_col_ptr.as<int32_t>();
  a.col_id = ep->col_id.as<int32_t>();
  a.col_ctr_ptr = ep->col_ctr_ptr.as<int32_t>();
  a.tile_batch_ptr = ep->tile_batch_ptr.as<int32_t>();
  a.ctr_slot = ep->ctr_slot.as<uint16_t>();
  a.ctr_b = ep->ctr_b.as<uint16_t>();
  a.ctr_off = ep->ctr_off.as<uint8_t>();
  a.rec = ep->rec.as<unsigned char>();
  a.rec_bytes = ep->bat.rec_bytes;
  a.off_offset = ep->bat.off_offset;
  a.endmask_words = ep->bat.endmask_words;
  a.colptr = plan.colptr.as<int32_t>();
  a.nzval = plan.nzval.as<double>();
  if (ep->k.mfma) {
    a.off_bytes = ep->bat.off_bytes;
  } else {
    a.off_stride = ep->offs.stride;
    a.off_bytes = ep->offs.bytes;
  }
  a.max_seg = std::max(1, plan.max_seg);
  a.lds_geo = ep->lds_geo;
  a.acc_offset = ep->acc_offset;
  a.acc_per_wave = ep->acc_per_wave;
  a.stage_cells = ep->stage_cells;
  a.lds_claim = ep->lds_claim;
  a.ntiles = ep->tiles.ntiles;
  if (ctx->nfaces) {
    a.item_face = ctx->d_item_face.as<int32_t>();
    a.face_jac = ctx->d_face_jac.as<double>();
    a.face_normal = ctx->d_face_normal.as<double>();
  }
  return a;
}

static int launch_matrix(gtk_ctx *ctx, MatrixPlan &plan, int form, int precision) {
  if (!ctx->has_ref) FAIL(ctx, GTK_ESTATE, "gtk_set_reference first");
  if (precision != GTK_FP64 && precision != GTK_MIXED) FAIL(ctx, GTK_EINVAL, "unknown precision");
  if (form < GTK_LAPLACE || form > GTK_SKELETON) FAIL(ctx, GTK_EINVAL, "unknown form");
  if (form == GTK_ELASTICITY && ctx->ncoef != 2) FAIL(ctx, GTK_ESTATE, "elasticity needs (lambda, mu) per cell");
  if (form != GTK_ELASTICITY && ctx->ncoef == 2) FAIL(ctx, GTK_ESTATE, "scalar forms take one coefficient per cell");
  if ((form == GTK_NITSCHE || form == GTK_NITSCHE_LOAD || form == GTK_SKELETON) && !ctx->has_face_grad)
    FAIL(ctx, GTK_ESTATE, "Nitsche terms: gtk_set_reference_faces and gtk_set_reference_face_gradients first");
  if (ctx->faces_packed && form != GTK_SKELETON)
    FAIL(ctx, GTK_ESTATE, "item_face holds point-set pairs (skeleton items): GTK_SKELETON only");
  if (form == GTK_SKELETON && (ctx->nv % 2 || ctx->dofs_per_cell != 2 * ctx->nshape * ctx->ncomp))
    FAIL(ctx, GTK_EINVAL, "skeleton items: two cells' nodes and DOFs per item");
  ExecPlan *ep = nullptr;
  int rc = get_exec(ctx, plan, form, precision, ep);
  if (rc) return rc;
  if (ep->tiles.ntiles == 0) return GTK_OK;
  gtk::MatrixArgs a = matrix_args(ctx, plan, *ep);
  void *args[] = {&a};
  // matrix-core kernels are persistent: as many workgroups as fit at once,
  // each looping over tiles
  const int grid = ep->k.mfma ? std::min(ep->tiles.ntiles, ep->resident) : ep->tiles.ntiles;
  HIPCHK(ctx, hipLaunchKernel(ep->k.fn, dim3(grid), dim3(ep->k.block), args, ep->lds, ctx->stream));
  return GTK_OK;
}

static MatrixPlan *find_plan(gtk_ctx *ctx, int rows, int cols) {
  auto it = ctx->mplans.find(key_of(rows, cols));
  return it == ctx->mplans.end() ? nullptr : it->second.get();
}

static void fill_stats(const gtk_ctx *ctx, const MatrixPlan &p, int form, int precision, gtk_stats *st) {
  st->cells = ctx->n_cells;
  st->nnz = p.pat.nnz;
  // algorithmic bytes: coordinates, cell->node map, coefficients, the plan's
  // streams read once (batch records, or offsets and contributor lists),
  // and the values written once
  int64_t plan_bytes = 0;
  auto it = p.exec.find(form * 2 + precision);
  if (it != p.exec.end()) {
    const ExecPlan &e = *it->second;
    plan_bytes = e.k.mfma ? (int64_t)e.bat.rec.size()
                          : (int64_t)e.offs.data.size() + (int64_t)e.tiles.ctr_slot.size() * 4 +
                                (int64_t)e.tiles.col_id.size() * 8;
  }
  int64_t b = (int64_t)ctx->n_nodes * ctx->dim * 8 + (int64_t)ctx->n_cells * ctx->nv * 4 +
              (int64_t)ctx->n_cells * ctx->ncoef * 8 + plan_bytes + p.pat.nnz * 8;
  st->bytes_alg = b;
  const int64_t D = ctx->dim, nd = ctx->dofs_per_cell;
  st->flops_alg = (int64_t)ctx->n_cells * ctx->nq * (2 * D * nd * nd + 2 * D * D * nd + 2 * ctx->nv * D * D);
}

int gtk_plan_info(gtk_ctx *ctx, int form, int precision, int rows, int cols, gtk_plan_info_t *out) {
  if (!ctx || !out) return GTK_EINVAL;
  MatrixPlan *p = find_plan(ctx, rows, cols);
  if (!p) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern first");
  if (precision != GTK_FP64 && precision != GTK_MIXED) FAIL(ctx, GTK_EINVAL, "unknown precision");
  if (form < GTK_LAPLACE || form > GTK_SKELETON) FAIL(ctx, GTK_EINVAL, "unknown form");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  ExecPlan *ep = nullptr;
  int rc = get_exec(ctx, *p, form, precision, ep);
  if (rc) return rc;
  std::memset(out, 0, sizeof(*out));
  out->matrix_cores = ep->k.mfma ? 1 : 0;
  out->threads = ep->k.block;
  out->grid = ep->k.mfma ? std::min(ep->tiles.ntiles, ep->resident) : ep->tiles.ntiles;
  out->tiles = ep->tiles.ntiles;
  out->max_tile_cells = ep->tiles.max_tile_cells;
  out->stage_cells = ep->stage_cells;
  out->lds_bytes = (int)ep->lds;
  out->batches = ep->k.mfma ? ep->bat.nbatches : 0;
  out->tile_cells = ep->tiles.tile_cells.size();
  out->cells = ctx->n_cells;
  out->tile_cap = ep->tile_cap;
  return GTK_OK;
}

int gtk_set_launch(gtk_ctx *ctx, int form, int precision, int rows, int cols, int waves, int max_cells) {
  if (!ctx) return GTK_EINVAL;
  MatrixPlan *p = find_plan(ctx, rows, cols);
  if (!p) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern first");
  if (precision != GTK_FP64 && precision != GTK_MIXED) FAIL(ctx, GTK_EINVAL, "unknown precision");
  if (form < GTK_LAPLACE || form > GTK_SKELETON) FAIL(ctx, GTK_EINVAL, "unknown form");
  if (waves != 0 && waves != 4 && waves != 8 && waves != 12 && waves != 16)
    FAIL(ctx, GTK_EINVAL, "waves must be 0 (default), 4, 8, 12 or 16");
  if (max_cells < 0 || max_cells > 4095) FAIL(ctx, GTK_EINVAL, "max_cells must be 0 (LDS-bound) .. 4095");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<ExecPlan> ep;
  const int w = waves == 0 || waves == 4 ? default_waves(ctx, *p, form, precision) : waves;
  int rc = build_exec(ctx, *p, form, precision, waves == 4 ? 4 : w, max_cells, ep);
  if (rc) return rc;
  if (ep->k.mfma && waves == 4) FAIL(ctx, GTK_EINVAL, "the matrix-core kernels take 8, 12 or 16 waves");
  if (!ep->k.mfma && waves != 0 && waves != 4) FAIL(ctx, GTK_EINVAL, "the VALU kernel takes 0 (its default) or 4 waves");
  p->exec[form * 2 + precision] = std::move(ep);
  ctx->generation++;
  return GTK_OK;
}

int gtk_assemble_matrix_async(gtk_ctx *ctx, int form, int precision, int rows, int cols) {
  if (!ctx) return GTK_EINVAL;
  MatrixPlan *p = find_plan(ctx, rows, cols);
  if (!p) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern first");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return launch_matrix(ctx, *p, form, precision);
}

int gtk_sync(gtk_ctx *ctx) {
  if (!ctx) return GTK_EINVAL;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return GTK_OK;
}

int gtk_assemble_matrix(gtk_ctx *ctx, int form, int precision, int rows, int cols, double *nzval_out,
                        gtk_stats *stats) {
  if (!ctx) return GTK_EINVAL;
  auto t0 = std::chrono::steady_clock::now();
  MatrixPlan *p = find_plan(ctx, rows, cols);
  if (!p) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern first");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  int rc = launch_matrix(ctx, *p, form, precision);
  if (rc) return rc;
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  if (nzval_out && p->pat.nnz)
    HIPCHK(ctx, hipMemcpyAsync(nzval_out, p->nzval.p, p->pat.nnz * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (stats) {
    float ms = 0;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    stats->kernel_ms = ms;
    fill_stats(ctx, *p, form, precision, stats);
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return GTK_OK;
}

int gtk_time_matrix(gtk_ctx *ctx, int form, int precision, int rows, int cols, int launches, double *ms_per_launch) {
  if (!ctx) return GTK_EINVAL;
  if (launches < 1 || !ms_per_launch) FAIL(ctx, GTK_EINVAL, "launches must be >= 1");
  MatrixPlan *p = find_plan(ctx, rows, cols);
  if (!p) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern first");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int i = 0; i < launches; i++) {
    int rc = launch_matrix(ctx, *p, form, precision);
    if (rc) return rc;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *ms_per_launch = ms / launches;
  return GTK_OK;
}

int gtk_get_values(gtk_ctx *ctx, int rows, int cols, double *nzval_out) {
  if (!ctx) return GTK_EINVAL;
  MatrixPlan *p = find_plan(ctx, rows, cols);
  if (!p) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern first");
  if (!nzval_out) FAIL(ctx, GTK_EINVAL, "nzval_out is NULL");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (p->pat.nnz)
    HIPCHK(ctx, hipMemcpyAsync(nzval_out, p->nzval.p, p->pat.nnz * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return GTK_OK;
}

int gtk_set_values(gtk_ctx *ctx, int rows, int cols, const double *nzval) {
  if (!ctx) return GTK_EINVAL;
  MatrixPlan *p = find_plan(ctx, rows, cols);
  if (!p) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern first");
  if (!nzval && p->pat.nnz) FAIL(ctx, GTK_EINVAL, "nzval is NULL");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (p->pat.nnz)
    HIPCHK(ctx, hipMemcpyAsync(p->nzval.p, nzval, p->pat.nnz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return GTK_OK;
}

// Row view of a block (built once per block): columns ascending, so each
// row lists its entries in ascending column order; with staging vectors for
// a product with the block.
static int row_view(gtk_ctx *ctx, MatrixPlan &P) {
  if (P.lift_rptr.p) return GTK_OK;
  const int64_t nr = P.pat.nrows, nc = P.pat.ncols, nnz = P.pat.nnz;
  std::vector<int64_t> rptr(nr + 1, 0);
  for (int64_t k = 0; k < nnz; k++) rptr[P.pat.rowval[k]]++;
  for (int64_t i = 0; i < nr; i++) rptr[i + 1] += rptr[i];
  std::vector<int64_t> next(rptr.begin(), rptr.end() - 1);
  std::vector<int32_t> perm(std::max<int64_t>(1, nnz)), col(std::max<int64_t>(1, nnz));
  for (int64_t j = 0; j < nc; j++)
    for (int64_t k = P.pat.colptr[j] - 1; k < P.pat.colptr[j + 1] - 1; k++) {
      const int64_t slot = next[P.pat.rowval[k] - 1]++;
      perm[slot] = (int32_t)k;
      col[slot] = (int32_t)j;
    }
  int rc;
  if ((rc = upload(ctx, P.lift_rptr, rptr.data(), rptr.size())) || (rc = upload(ctx, P.lift_perm, perm.data(), perm.size())) ||
      (rc = upload(ctx, P.lift_col, col.data(), col.size())) ||
      (rc = alloc(ctx, P.lift_x, std::max<int64_t>(1, nc) * sizeof(double))) ||
      (rc = alloc(ctx, P.lift_b, std::max<int64_t>(1, nr) * sizeof(double))))
    return rc;
  return GTK_OK;
}

int gtk_block_multiply(gtk_ctx *ctx, int rows, int cols, double alpha, const double *x, double *y) {
  if (!ctx) return GTK_EINVAL;
  MatrixPlan *p = find_plan(ctx, rows, cols);
  if (!p) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern(rows, cols) and gtk_assemble_matrix first");
  const int64_t nr = p->pat.nrows, nc = p->pat.ncols;
  if ((!x && nc) || (!y && nr)) FAIL(ctx, GTK_EINVAL, "x and y are required");
  if (!std::isfinite(alpha)) FAIL(ctx, GTK_EINVAL, "alpha must be finite");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc = row_view(ctx, *p);
  if (rc) return rc;
  if (nc) HIPCHK(ctx, hipMemcpyAsync(p->lift_x.p, x, nc * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  if (nr) HIPCHK(ctx, hipMemcpyAsync(p->lift_b.p, y, nr * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  gtk::launch_lift(nr, p->lift_rptr.as<int64_t>(), p->lift_perm.as<int32_t>(), p->lift_col.as<int32_t>(),
                   p->nzval.as<double>(), alpha, p->lift_x.as<double>(), p->lift_b.as<double>(), ctx->stream);
  if (nr) HIPCHK(ctx, hipMemcpyAsync(y, p->lift_b.p, nr * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return GTK_OK;
}

int gtk_lift_dirichlet(gtk_ctx *ctx, int rows, const double *xd, double *b) {
  if (!ctx) return GTK_EINVAL;
  if (!find_plan(ctx, rows, GTK_DIRICHLET))
    FAIL(ctx, GTK_ESTATE, "gtk_build_pattern(rows, GTK_DIRICHLET) and gtk_assemble_matrix first");
  return gtk_block_multiply(ctx, rows, GTK_DIRICHLET, -1.0, xd, b);
}

int gtk_values_device_ptr(gtk_ctx *ctx, int rows, int cols, double **dptr) {
  if (!ctx) return GTK_EINVAL;
  MatrixPlan *p = find_plan(ctx, rows, cols);
  if (!p) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern first");
  if (!dptr) FAIL(ctx, GTK_EINVAL, "dptr is NULL");
  *dptr = p->nzval.as<double>();
  return GTK_OK;
}

int gtk_assemble_vector(gtk_ctx *ctx, int fkind, const double *f, int rows, double *b_out, gtk_stats *stats) {
  if (!ctx) return GTK_EINVAL;
  auto t0 = std::chrono::steady_clock::now();
  if (!ctx->has_space) FAIL(ctx, GTK_ESTATE, "gtk_set_space first");
  if (!ctx->has_ref) FAIL(ctx, GTK_ESTATE, "gtk_set_reference first");
  if (rows != GTK_FREE && rows != GTK_DIRICHLET) FAIL(ctx, GTK_EINVAL, "rows must be GTK_FREE or GTK_DIRICHLET");
  if (fkind < GTK_F_CONST || fkind > GTK_F_NODAL || !f) FAIL(ctx, GTK_EINVAL, "invalid load");
  if (ctx->faces_packed) FAIL(ctx, GTK_ESTATE, "skeleton items carry no load vector");
  gtk::KernelInfo k;
  if (!gtk::find_vector_kernel(elem_key(ctx), fkind, k)) FAIL(ctx, GTK_EUNSUP, "no vector kernel for this element");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  VectorPlan *vpp = nullptr;
  int rc = get_vector_plan(ctx, rows, vpp);
  if (rc) return rc;
  VectorPlan &vp = *vpp;
  if ((rc = upload(ctx, vp.f, f, load_count(ctx, fkind)))) return rc;
  size_t lds = k.table_bytes + (size_t)vp.tiles.max_tile_cells * k.geo_bytes_per_cell;
  if (lds > 160 * 1024) FAIL(ctx, GTK_EUNSUP, "tile does not fit in LDS");
  if (lds > 64 * 1024) HIPCHK(ctx, hipFuncSetAttribute(k.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  gtk::VectorArgs a = vector_args(ctx, vp, vp.f.as<double>());
  void *args[] = {&a};
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (vp.tiles.uncovered && vp.nrows)  // rows no cell touches stay zero
    HIPCHK(ctx, hipMemsetAsync(vp.out.p, 0, vp.nrows * sizeof(double), ctx->stream));
  if (vp.tiles.ntiles)
    HIPCHK(ctx, hipLaunchKernel(k.fn, dim3(vp.tiles.ntiles), dim3(gtk::kBlock), args, lds, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  if (b_out && vp.nrows)
    HIPCHK(ctx, hipMemcpyAsync(b_out, vp.out.p, vp.nrows * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (stats) {
    float ms = 0;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    stats->kernel_ms = ms;
    stats->cells = ctx->n_cells;
    stats->nnz = vp.nrows;
    stats->bytes_alg = (int64_t)ctx->n_nodes * ctx->dim * 8 + (int64_t)ctx->n_cells * ctx->nv * 4 + vp.nrows * 8;
    stats->flops_alg = (int64_t)ctx->n_cells * ctx->nq * (2 * ctx->dofs_per_cell + 2 * ctx->nv * ctx->dim * ctx->dim);
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return GTK_OK;
}

int gtk_assemble_vector_noalloc(gtk_ctx *ctx, int fkind, const double *f, int rows, double *b_out, gtk_stats *stats) {
  if (!ctx) return GTK_EINVAL;
  auto t0 = std::chrono::steady_clock::now();
  if (!ctx->has_space) FAIL(ctx, GTK_ESTATE, "gtk_set_space first");
  if (!ctx->has_ref) FAIL(ctx, GTK_ESTATE, "gtk_set_reference first");
  if (rows != GTK_FREE && rows != GTK_DIRICHLET) FAIL(ctx, GTK_EINVAL, "rows must be GTK_FREE or GTK_DIRICHLET");
  if (fkind < GTK_F_CONST || fkind > GTK_F_NODAL || !f) FAIL(ctx, GTK_EINVAL, "invalid load");
  // scratch layout: the vector (rounded up to 256 bytes), then the load values
  const int64_t nrows = rows == GTK_FREE ? ctx->n_free : ctx->n_diri;
  const size_t nout = ((size_t)std::max<int64_t>(1, nrows) + 31) & ~(size_t)31;
  const size_t nf = load_count(ctx, fkind);
  if (nout + nf > kNoAllocDoubles)
    FAIL(ctx, GTK_EINVAL,
         "gtk_assemble_vector_noalloc: " + std::to_string(nrows) + " rows and " + std::to_string(nf) +
             " load values exceed the scratch of " + std::to_string(kNoAllocDoubles) +
             " doubles (GTK_NOALLOC_MAX_DOUBLES); use gtk_assemble_vector");
  double *s = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch[ctx->device] : nullptr;
  if (!s) FAIL(ctx, GTK_EUNSUP, "no scratch on this device ordinal");
  auto it = ctx->vplans.find(rows);
  if (it == ctx->vplans.end() || !it->second)
    FAIL(ctx, GTK_ESTATE, "gtk_assemble_vector on this row set first (it builds the tile plan)");
  const VectorPlan &vp = *it->second;
  gtk::KernelInfo k;
  if (!gtk::find_vector_kernel(elem_key(ctx), fkind, k)) FAIL(ctx, GTK_EUNSUP, "no vector kernel for this element");
  size_t lds = k.table_bytes + (size_t)vp.tiles.max_tile_cells * k.geo_bytes_per_cell;
  if (lds > 160 * 1024) FAIL(ctx, GTK_EUNSUP, "tile does not fit in LDS");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (lds > 64 * 1024) HIPCHK(ctx, hipFuncSetAttribute(k.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  double *sf = s + nout;
  gtk::VectorArgs a = vector_args(ctx, vp, sf);
  a.out = s;
  void *args[] = {&a};
  if (nf) HIPCHK(ctx, hipMemcpyAsync(sf, f, nf * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (vp.tiles.uncovered && nrows) {  // rows no cell touches stay zero; the array keeps the last call's values
    const int64_t blocks = std::min<int64_t>(1024, (nrows + 255) / 256);
    hipLaunchKernelGGL(zero_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, s, nrows);
    HIPCHK(ctx, hipGetLastError());
  }
  if (vp.tiles.ntiles)
    HIPCHK(ctx, hipLaunchKernel(k.fn, dim3(vp.tiles.ntiles), dim3(gtk::kBlock), args, lds, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  if (b_out && nrows) HIPCHK(ctx, hipMemcpyAsync(b_out, s, nrows * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (stats) {
    float ms = 0;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    stats->kernel_ms = ms;
    stats->cells = ctx->n_cells;
    stats->nnz = nrows;
    stats->bytes_alg = (int64_t)ctx->n_nodes * ctx->dim * 8 + (int64_t)ctx->n_cells * ctx->nv * 4 + nrows * 8;
    stats->flops_alg = (int64_t)ctx->n_cells * ctx->nq * (2 * ctx->dofs_per_cell + 2 * ctx->nv * ctx->dim * ctx->dim);
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return GTK_OK;
}

}  // extern "C"

// Device vectors of one solve call: the FP64 iterate, right-hand side and
// residual, the conjugate-gradient vectors (FP32 or FP64), the partial sums
// of the dot products, the scalars and the check flags.
struct SolveBufs {
  double *x = nullptr, *b = nullptr, *r = nullptr;
  void *dinv = nullptr, *d = nullptr, *rr = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  double *partial = nullptr, *partial2 = nullptr, *sc = nullptr;
  int *bad = nullptr;
};

static bool solve_try_bsr(const gtk_ctx *ctx, const MatrixPlan &P, bool f32) {
  const char *bsr_env = std::getenv("GTK_SOLVE_BSR");
  return f32 && ctx->ncomp == 3 && P.pat.ncols % 3 == 0 && P.pat.nnz > 0 && !(bsr_env && bsr_env[0] == '0');
}

static int solve_check_args(gtk_ctx *ctx, MatrixPlan *pp, int precision, double rtol, int max_outer, int max_inner,
                            double inner_rtol) {
  if (!pp) FAIL(ctx, GTK_ESTATE, "gtk_build_pattern(GTK_FREE, GTK_FREE) and gtk_assemble_matrix first");
  if (precision != GTK_FP64 && precision != GTK_MIXED) FAIL(ctx, GTK_EINVAL, "unknown precision");
  if (!(rtol > 0) || max_outer < 1 || max_inner < 1 || !(inner_rtol > 0 && inner_rtol < 1))
    FAIL(ctx, GTK_EINVAL, "rtol > 0, max_outer >= 1, max_inner >= 1, 0 < inner_rtol < 1");
  if (pp->pat.ncols != pp->pat.nrows) FAIL(ctx, GTK_EINVAL, "the block is not square");
  return GTK_OK;
}

// The solve's plan on a (FREE, FREE) block, built once per block and
// precision: the pattern's rows on the device, the CSR-position -> CSC
// position permutation, the row-ordered value arrays and, for
// three-component fields made of aligned 3x3 blocks, the BSR(3) copy.
static int solve_prepare(gtk_ctx *ctx, MatrixPlan &P, bool f32) {
  SolveWork &w = P.sw;
  const int64_t n = P.pat.ncols, nnz = P.pat.nnz;
  hipStream_t st = ctx->stream;
  int rc;
  if (!P.rowval.p && (rc = upload(ctx, P.rowval, P.pat.rowval.data(), P.pat.rowval.size()))) return rc;
  if ((rc = alloc(ctx, w.val64, std::max<int64_t>(1, nnz) * sizeof(double)))) return rc;
  if (f32 && (rc = alloc(ctx, w.val32, std::max<int64_t>(1, nnz) * sizeof(float)))) return rc;
  if ((rc = alloc(ctx, w.bad, 16))) return rc;
  gtk::SolveMat M;
  M.n = n;
  M.ptr = P.colptr.as<int32_t>();
  M.idx = P.rowval.as<int32_t>();
  if (!w.perm.p) {  // CSR position -> CSC position
    if ((rc = alloc(ctx, w.perm, std::max<int64_t>(1, nnz) * sizeof(int32_t)))) return rc;
    HIPCHK(ctx, hipMemsetAsync(w.bad.p, 0, 16, st));
    gtk::launch_transpose_perm(M, w.perm.as<int32_t>(), w.bad.as<int>(), st);
    int asym = 0;
    HIPCHK(ctx, hipMemcpyAsync(&asym, w.bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    if (asym) {
      w.perm.release();
      FAIL(ctx, GTK_EINVAL, "the pattern is not structurally symmetric");
    }
  }
  if (solve_try_bsr(ctx, P, f32)) {
    if (w.block3 < 0) {  // checked once per block
      HIPCHK(ctx, hipMemsetAsync(w.bad.p, 0, 16, st));
      gtk::launch_block3_check(M, w.bad.as<int>() + 2, st);
      int bad[4] = {0, 0, 0, 0};
      HIPCHK(ctx, hipMemcpyAsync(bad, w.bad.p, sizeof(bad), hipMemcpyDeviceToHost, st));
      HIPCHK(ctx, hipStreamSynchronize(st));
      w.block3 = bad[2] ? 0 : 1;
    }
    if (w.block3 == 1 && !w.bcol.p &&
        ((rc = alloc(ctx, w.bcol, (nnz / 9) * sizeof(int32_t))) || (rc = alloc(ctx, w.bval, nnz * sizeof(float)))))
      return rc;
  }
  w.prepared[f32 ? 1 : 0] = true;
  return GTK_OK;
}

// The solve itself on a prepared block, on device vectors only (no
// allocation): values gathered into row order, the inverse diagonal, then
// FP64 refinement around the (FP32 or FP64) Jacobi-preconditioned CG.
static int solve_core(gtk_ctx *ctx, MatrixPlan &P, bool f32, const SolveBufs &u, double rtol, int max_outer,
                      int max_inner, double inner_rtol, gtk_solve_stats *stats, std::chrono::steady_clock::time_point t0,
                      double *host_x) {
  SolveWork &w = P.sw;
  const int64_t n = P.pat.ncols, nnz = P.pat.nnz;
  gtk::SolveMat M;
  M.n = n;
  M.ptr = P.colptr.as<int32_t>();
  M.idx = P.rowval.as<int32_t>();
  M.lanes = gtk::solve_rows_per_lane_group(n ? (double)nnz / n : 1.0);
  if (const char *e = std::getenv("GTK_SPMV_LANES")) {  // measurement override
    const int v = std::atoi(e);
    if (v == 4 || v == 8 || v == 16 || v == 32 || v == 64) M.lanes = v;
  }
  hipStream_t st = ctx->stream;
  double *sc = u.sc;
  int rc;
  HIPCHK(ctx, hipMemsetAsync(u.bad, 0, 16, st));
  gtk::launch_gather_values(nnz, w.perm.as<int32_t>(), P.nzval.as<double>(), w.val64.as<double>(),
                            f32 ? w.val32.as<float>() : nullptr, st);
  gtk::launch_diag_inv(M, w.val64.as<double>(), u.dinv, !f32, u.bad, st);
  int bad = 0;
  HIPCHK(ctx, hipMemcpyAsync(&bad, u.bad, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
  if (bad) FAIL(ctx, GTK_EINVAL, "the matrix has a non-positive diagonal entry (not SPD)");
  // three-component fields numbered node by node: the FP32 iterations use
  // the BSR(3) copy (one column index per 9 values)
  if (solve_try_bsr(ctx, P, f32) && w.block3 == 1) {
    M.nb = nnz / 9;
    M.bcol = w.bcol.as<int32_t>();
    M.bval = w.bval.as<float>();
    if (const char *e = std::getenv("GTK_BSR_LANES")) {  // measurement override
      const int v = std::atoi(e);
      if (v == 8 || v == 16 || v == 32) M.block_lanes = v;
    }
    gtk::launch_block3_build(M, w.val32.as<float>(), w.bcol.as<int32_t>(), w.bval.as<float>(), !w.bsr_cols, st);
    w.bsr_cols = true;
  }
  auto read_sc = [&](int slot, double &out) -> int {
    HIPCHK(ctx, hipMemcpyAsync(&out, sc + slot, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    return GTK_OK;
  };
  // FP64 residual r = b - A x and its norm
  auto residual = [&](double &rn) -> int {
    gtk::launch_spmv(M, w.val64.p, true, u.x, true, u.b, u.r, true, st);
    gtk::launch_dot(n, u.r, u.r, true, u.partial, sc, 6, st);
    double v = 0;
    int e = read_sc(6, v);
    rn = std::sqrt(std::max(0.0, v));
    return e;
  };
  double bn2 = 0;
  gtk::launch_dot(n, u.b, u.b, true, u.partial, sc, 6, st);
  if ((rc = read_sc(6, bn2))) return rc;
  const double bn = std::sqrt(bn2);
  gtk::CgVectors v;
  v.f64 = !f32;
  v.d = u.d;
  v.r = u.rr;
  v.z = u.z;
  v.p = u.p;
  v.q = u.q;
  v.dinv = u.dinv;
  const void *val = f32 ? w.val32.p : w.val64.p;
  constexpr int kCheck = 8;  // iterations between convergence reads
  double rn = 0;
  if ((rc = residual(rn))) return rc;
  int outer = 0, inner_total = 0;
  const double target = rtol * (bn > 0 ? bn : 1.0);
  while (rn > target && outer < max_outer) {
    // inner solve of A d = r / |r| (FP32 for the mixed form, FP64 otherwise)
    const double s = rn > 0 ? 1.0 / rn : 1.0;
    gtk::launch_cg_start(n, s, u.r, v, u.partial, sc, st);
    // the FP64 form runs one CG to the outer tolerance; the mixed form to inner_rtol
    const double itol = f32 ? inner_rtol : std::max(target / rn, 1e-300);
    int it = 0;
    while (it < max_inner) {
      for (int k = 0; k < kCheck && it < max_inner; k++, it++)
        gtk::launch_cg_iteration(M, val, v, u.partial, u.partial2, sc, it, k == kCheck - 1 || it == max_inner - 1,
                                 st);
      double rr = 0;
      if ((rc = read_sc(5, rr))) return rc;
      if (!(std::sqrt(std::max(0.0, rr)) > itol)) break;
    }
    inner_total += it;
    gtk::launch_refine(n, rn, v.d, !f32, u.x, st);
    outer++;
    if ((rc = residual(rn))) return rc;
  }
  if (host_x && n) HIPCHK(ctx, hipMemcpyAsync(host_x, u.x, n * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
  if (stats) {
    stats->outer_iterations = outer;
    stats->inner_iterations = inner_total;
    stats->rel_residual = bn > 0 ? rn / bn : rn;
    stats->converged = rn <= target ? 1 : 0;
    stats->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return GTK_OK;
}

extern "C" {

int gtk_solve(gtk_ctx *ctx, int precision, const double *b, double *x, double rtol, int max_outer, int max_inner,
              double inner_rtol, gtk_solve_stats *stats) {
  if (!ctx) return GTK_EINVAL;
  auto t0 = std::chrono::steady_clock::now();
  MatrixPlan *pp = find_plan(ctx, GTK_FREE, GTK_FREE);
  int rc = solve_check_args(ctx, pp, precision, rtol, max_outer, max_inner, inner_rtol);
  if (rc) return rc;
  if (!b || !x) FAIL(ctx, GTK_EINVAL, "b and x are required");
  MatrixPlan &P = *pp;
  const int64_t n = P.pat.ncols;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  SolveWork &w = P.sw;
  const bool f32 = precision == GTK_MIXED;
  const size_t vb = f32 ? sizeof(float) : sizeof(double);
  if ((rc = solve_prepare(ctx, P, f32))) return rc;
  const size_t nb = std::max<int64_t>(1, n);
  if ((rc = alloc(ctx, w.dinv, nb * vb)) || (rc = alloc(ctx, w.x, nb * 8)) || (rc = alloc(ctx, w.b, nb * 8)) ||
      (rc = alloc(ctx, w.r, nb * 8)) || (rc = alloc(ctx, w.d, nb * vb)) || (rc = alloc(ctx, w.rr, nb * vb)) ||
      (rc = alloc(ctx, w.z, nb * vb)) || (rc = alloc(ctx, w.p, nb * vb)) || (rc = alloc(ctx, w.q, nb * vb)) ||
      (rc = alloc(ctx, w.partial, gtk::kSolvePartials * 8)) || (rc = alloc(ctx, w.partial2, 2 * gtk::kSolvePartials * 8)) ||
      (rc = alloc(ctx, w.sc, gtk::kSolveScalars * 8)))
    return rc;
  SolveBufs u;
  u.x = w.x.as<double>();
  u.b = w.b.as<double>();
  u.r = w.r.as<double>();
  u.dinv = w.dinv.p;
  u.d = w.d.p;
  u.rr = w.rr.p;
  u.z = w.z.p;
  u.p = w.p.p;
  u.q = w.q.p;
  u.partial = w.partial.as<double>();
  u.partial2 = w.partial2.as<double>();
  u.sc = w.sc.as<double>();
  u.bad = w.bad.as<int>();
  if (n) {
    HIPCHK(ctx, hipMemcpyAsync(u.b, b, n * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(u.x, x, n * 8, hipMemcpyHostToDevice, ctx->stream));
  }
  return solve_core(ctx, P, f32, u, rtol, max_outer, max_inner, inner_rtol, stats, t0, x);
}

// ---------------------------------------------------------------------------
// batches of problems in one launch
// ---------------------------------------------------------------------------

}  // extern "C"

struct gtk_batch {
  int device = 0, form = 0, precision = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  std::vector<gtk_ctx *> ctxs;
  std::vector<std::unique_ptr<ExecPlan>> plans;  // one per problem
  std::vector<uint64_t> gens;                    // each context's generation when the plans were built
  int tile_cells[gtk::kBatchedTypes] = {};     // cap on the cells of a tile per element type (0: LDS bound)
  int family = -1;                               // of the problems' element types (type / 3)
  int waves = 8;                                 // waves per workgroup
  std::vector<gtk::MatrixArgs> args;  // host copy of d_problems
  struct Load {
    int fkind = -1;  // none
    size_t count = 0;
    DevBuf f;
  };
  std::vector<std::unique_ptr<Load>> loads;  // per problem
  std::vector<gtk::VectorArgs> vargs;       // host copy of d_loads
  DevBuf d_problems, d_loads, d_types, d_fkinds, d_items, d_counter;
  // the graph's persistent launches, in order: one per element type of a
  // hexahedral batch (Q1 tiles, then Q2 tiles), else one; each claims items
  // from its own counter
  struct Launch {
    gtk::KernelInfo k;
    size_t lds = 0;
    int first = 0, nitems = 0, grid = 0;
  };
  std::vector<Launch> launches;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int64_t cells = 0, nnz = 0;
};

static void batch_release_graph(gtk_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
}

#define BFAIL(b, code, msg) \
  do {                      \
    (b)->err = (msg);       \
    return (code);          \
  } while (0)

#define BHIPCHK(b, call)                                                       \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      (b)->err = std::string(#call) + ": " + hipGetErrorString(e_);            \
      return GTK_EDEVICE;                                                      \
    }                                                                          \
  } while (0)

template <typename T>
static int batch_upload(gtk_batch *b, DevBuf &buf, const T *src, size_t count) {
  buf.release();
  const size_t bytes = std::max<size_t>(1, count) * sizeof(T);
  hipError_t e = hipMalloc(&buf.p, bytes);
  if (e != hipSuccess) {
    buf.p = nullptr;
    BFAIL(b, GTK_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  buf.n = bytes;
  if (count) BHIPCHK(b, hipMemcpy(buf.p, src, count * sizeof(T), hipMemcpyHostToDevice));
  return GTK_OK;
}

extern "C" {

const char *gtk_batch_last_error(const gtk_batch *b) { return b ? b->err.c_str() : "batch is NULL"; }

int gtk_batch_create(int device, int form, int precision, gtk_batch **out) {
  if (!out) return GTK_EINVAL;
  *out = nullptr;
  if (form != GTK_LAPLACE && form != GTK_MASS) return GTK_EUNSUP;
  if (precision != GTK_FP64 && precision != GTK_MIXED) return GTK_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return GTK_EDEVICE;
  auto b = std::make_unique<gtk_batch>();
  b->device = device;
  b->form = form;
  b->precision = precision;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return GTK_EDEVICE;
  *out = b.release();
  return GTK_OK;
}

int gtk_batch_destroy(gtk_batch *b) {
  if (!b) return GTK_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_release_graph(b);
  b->plans.clear();
  b->loads.clear();
  b->d_problems.release();
  b->d_loads.release();
  b->d_types.release();
  b->d_fkinds.release();
  b->d_items.release();
  b->d_counter.release();
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return GTK_OK;
}

int gtk_batch_add(gtk_batch *b, gtk_ctx *ctx) {
  if (!b || !ctx) return GTK_EINVAL;
  if (ctx->device != b->device) BFAIL(b, GTK_EINVAL, "the problem's context is on another device");
  if (!find_plan(ctx, GTK_FREE, GTK_FREE)) BFAIL(b, GTK_ESTATE, "gtk_build_pattern(GTK_FREE, GTK_FREE) first");
  if (!ctx->has_ref) BFAIL(b, GTK_ESTATE, "gtk_set_reference first");
  const int type = ctx->ncomp == 1 ? gtk::batched_type(elem_key(ctx)) : -1;
  if (type < 0)
    BFAIL(b, GTK_EUNSUP, "batched problems are scalar Q1/Q2/Q3 quadrilaterals with (k+1)^2 points, Q1/Q2 "
                         "hexahedra with (k+1)^3 or P1/P2 triangles with degree-2k rules");
  if (b->family >= 0 && type / 3 != b->family)
    BFAIL(b, GTK_EUNSUP, "a batch holds one cell kind");
  if (ctx->ncoef > 1) BFAIL(b, GTK_EINVAL, "scalar forms take one coefficient per cell");
  batch_release_graph(b);
  b->family = type / 3;
  b->ctxs.push_back(ctx);
  b->loads.push_back(std::make_unique<gtk_batch::Load>());
  return GTK_OK;
}

int gtk_batch_set_load(gtk_batch *b, int problem, int fkind, const double *f) {
  if (!b) return GTK_EINVAL;
  if (problem < 0 || problem >= (int)b->ctxs.size()) BFAIL(b, GTK_EINVAL, "problem index out of range");
  if (fkind < GTK_F_CONST || fkind > GTK_F_NODAL || !f) BFAIL(b, GTK_EINVAL, "invalid load");
  BHIPCHK(b, hipSetDevice(b->device));
  gtk_batch::Load &l = *b->loads[problem];
  const size_t count = load_count(b->ctxs[problem], fkind);
  if (l.fkind != fkind || l.count != count || !l.f.p) {
    // a new kind of load changes the work items: the batch is built again
    batch_release_graph(b);
    int rc = batch_upload(b, l.f, f, count);
    if (rc) return rc;
    l.fkind = fkind;
    l.count = count;
    return GTK_OK;
  }
  BHIPCHK(b, hipMemcpy(l.f.p, f, count * sizeof(double), hipMemcpyHostToDevice));
  return GTK_OK;
}

int gtk_batch_get_vector(gtk_batch *b, int problem, double *out) {
  if (!b || !out) return GTK_EINVAL;
  if (problem < 0 || problem >= (int)b->ctxs.size()) BFAIL(b, GTK_EINVAL, "problem index out of range");
  if (b->loads[problem]->fkind < 0) BFAIL(b, GTK_ESTATE, "the problem has no load (gtk_batch_set_load)");
  gtk_ctx *ctx = b->ctxs[problem];
  auto it = ctx->vplans.find(GTK_FREE | kBatchedPlan);
  if (it == ctx->vplans.end() || !it->second) BFAIL(b, GTK_ESTATE, "gtk_batch_build first");
  BHIPCHK(b, hipSetDevice(b->device));
  if (it->second->nrows)
    BHIPCHK(b, hipMemcpy(out, it->second->out.p, it->second->nrows * sizeof(double), hipMemcpyDeviceToHost));
  return GTK_OK;
}

int gtk_batch_set_tile_cells(gtk_batch *b, int elem_type, int max_cells) {
  if (!b) return GTK_EINVAL;
  if (elem_type < 0 || elem_type >= gtk::kBatchedTypes || max_cells < 0)
    BFAIL(b, GTK_EINVAL, "elem_type 0..7, max_cells >= 0");
  b->tile_cells[elem_type] = max_cells;
  batch_release_graph(b);
  return GTK_OK;
}

int gtk_batch_set_waves(gtk_batch *b, int waves) {
  if (!b) return GTK_EINVAL;
  if (waves != 8 && waves != 12) BFAIL(b, GTK_EINVAL, "waves per workgroup: 8 or 12");
  if (waves != b->waves) batch_release_graph(b);
  b->waves = waves;
  return GTK_OK;
}

int gtk_batch_build(gtk_batch *b) {
  if (!b) return GTK_EINVAL;
  if (b->ctxs.empty()) BFAIL(b, GTK_ESTATE, "gtk_batch_add problems first");
  BHIPCHK(b, hipSetDevice(b->device));
  batch_release_graph(b);
  // a hexahedral batch runs one launch per element type: a kernel holding
  // both tile bodies keeps them as calls, whose saved registers go to
  // scratch at every tile (mixed 4.22 -> 3.80 ms, FP64 6.02 -> 5.70 ms on
  // 64 uneven Q1/Q2 problems; bytes written per launch 5.0 -> 2.6 GB)
  const bool split = b->family == 1;
  const int ngroups = split ? 2 : 1;
  std::vector<gtk_batch::Launch> groups(ngroups);
  for (int g = 0; g < ngroups; g++)
    if (!gtk::find_batched_kernel(b->form, b->precision, b->waves, split ? gtk::kHexQ1Launch + g : b->family,
                                  groups[g].k))
      BFAIL(b, GTK_EUNSUP, "no batched kernel for the form");
  b->plans.clear();
  b->gens.clear();
  std::vector<gtk::MatrixArgs> args;
  std::vector<gtk::VectorArgs> vargs(b->ctxs.size());
  std::vector<int32_t> types, fkinds(b->ctxs.size(), -1), group(b->ctxs.size(), 0);
  // work items with their estimated cost (cells x shape functions per
  // cell), per launch: matrix tiles, then load-vector tiles, each by
  // descending cost so that the claims of the last workgroups are short items
  std::vector<std::vector<std::pair<int64_t, int2>>> mitems(ngroups), vitems(ngroups);
  b->cells = b->nnz = 0;
  for (size_t p = 0; p < b->ctxs.size(); p++) {
    gtk_ctx *ctx = b->ctxs[p];
    MatrixPlan *P = find_plan(ctx, GTK_FREE, GTK_FREE);
    const int type = gtk::batched_type(elem_key(ctx));
    const int g = split ? type % 3 : 0;
    group[p] = g;
    std::unique_ptr<ExecPlan> ep;
    int rc = build_exec(ctx, *P, b->form, b->precision, b->waves, b->tile_cells[type], ep);
    if (rc) BFAIL(b, rc, std::string("problem ") + std::to_string(p) + ": " + ctx->err);
    if (!ep->k.mfma || ep->bat.off_bytes != 1)
      BFAIL(b, GTK_EUNSUP, "a problem does not fit the matrix-core batches with byte offsets");
    groups[g].lds = std::max(groups[g].lds, ep->lds);
    args.push_back(matrix_args(ctx, *P, *ep));
    types.push_back(type % 3);
    const int64_t ns = elem_key(ctx).ns;
    for (int t = 0; t < ep->tiles.ntiles; t++)
      mitems[g].push_back({ns * (ep->tiles.tile_cell_ptr[t + 1] - ep->tiles.tile_cell_ptr[t]), make_int2((int)p, t)});
    b->cells += ctx->n_cells;
    b->nnz += P->pat.nnz;
    b->plans.push_back(std::move(ep));
    b->gens.push_back(ctx->generation);
  }
  for (size_t p = 0; p < b->ctxs.size(); p++) {
    const gtk_batch::Load &l = *b->loads[p];
    if (l.fkind < 0) continue;
    gtk_ctx *ctx = b->ctxs[p];
    gtk::KernelInfo vk;
    if (!gtk::find_vector_kernel(elem_key(ctx), l.fkind, vk)) BFAIL(b, GTK_EUNSUP, "no vector kernel for a problem");
    VectorPlan *vp = nullptr;
    if (get_vector_plan(ctx, GTK_FREE, vp, true))
      BFAIL(b, GTK_EINVAL, std::string("problem ") + std::to_string(p) + ": " + ctx->err);
    if (vp->tiles.uncovered) BFAIL(b, GTK_EUNSUP, "a problem's load vector has rows no cell touches");
    const int g = group[p];
    groups[g].lds = std::max(groups[g].lds, vk.table_bytes + (size_t)vp->tiles.max_tile_cells * vk.geo_bytes_per_cell);
    vargs[p] = vector_args(ctx, *vp, l.f.as<double>());
    fkinds[p] = l.fkind;
    const int64_t ns = elem_key(ctx).ns;
    for (int t = 0; t < vp->tiles.ntiles; t++)
      vitems[g].push_back({ns * (vp->tiles.tile_cell_ptr[t + 1] - vp->tiles.tile_cell_ptr[t]), make_int2((int)p, ~t)});
  }
  auto by_cost = [](const std::pair<int64_t, int2> &x, const std::pair<int64_t, int2> &y) { return x.first > y.first; };
  std::vector<int2> items;
  for (int g = 0; g < ngroups; g++) {
    std::stable_sort(mitems[g].begin(), mitems[g].end(), by_cost);
    std::stable_sort(vitems[g].begin(), vitems[g].end(), by_cost);
    groups[g].first = (int)items.size();
    for (const auto &m : mitems[g]) items.push_back(m.second);
    for (const auto &v : vitems[g]) items.push_back(v.second);
    groups[g].nitems = (int)items.size() - groups[g].first;
  }
  const std::vector<int> zeros(ngroups, 0);
  int rc;
  b->args = args;
  b->vargs = vargs;
  if ((rc = batch_upload(b, b->d_problems, args.data(), args.size()))) return rc;
  if ((rc = batch_upload(b, b->d_loads, vargs.data(), vargs.size()))) return rc;
  if ((rc = batch_upload(b, b->d_fkinds, fkinds.data(), fkinds.size()))) return rc;
  if ((rc = batch_upload(b, b->d_types, types.data(), types.size()))) return rc;
  if ((rc = batch_upload(b, b->d_items, items.data(), items.size()))) return rc;
  if ((rc = batch_upload(b, b->d_counter, zeros.data(), zeros.size()))) return rc;
  int cus = 0;
  BHIPCHK(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  b->launches.clear();
  std::vector<gtk::BatchedArgs> bas;
  for (int g = 0; g < ngroups; g++) {
    gtk_batch::Launch &L = groups[g];
    if (L.nitems == 0) continue;
    // the claim slots after the largest tile's LDS
    const int claim_offset = (int)((L.lds + 15) & ~(size_t)15);
    L.lds = (size_t)claim_offset + 16;
    if (L.lds > 160 * 1024) BFAIL(b, GTK_EUNSUP, "a problem's tile does not fit in LDS");
    if (L.lds > 64 * 1024)
      BHIPCHK(b, hipFuncSetAttribute(L.k.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds));
    int per_cu = 0;
    BHIPCHK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, L.k.fn, L.k.block, L.lds));
    L.grid = std::max(1, std::min(L.nitems, std::max(1, per_cu) * std::max(1, cus)));
    gtk::BatchedArgs ba;
    ba.problems = b->d_problems.as<gtk::MatrixArgs>();
    ba.loads = b->d_loads.as<gtk::VectorArgs>();
    ba.types = b->d_types.as<int32_t>();
    ba.fkinds = b->d_fkinds.as<int32_t>();
    ba.items = b->d_items.as<int2>() + L.first;
    ba.nitems = L.nitems;
    ba.counter = b->d_counter.as<int>() + g;
    ba.claim_offset = claim_offset;
    bas.push_back(ba);
    b->launches.push_back(L);
  }
  // the whole batch is captured once as a graph: the counters' reset, then
  // the persistent launches in order
  BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(claim_reset_kernel, dim3(1), dim3(64), 0, b->stream, b->d_counter.as<int>(), ngroups);
  hipError_t le = hipGetLastError();
  for (size_t i = 0; i < b->launches.size() && le == hipSuccess; i++) {
    void *kargs[] = {&bas[i]};
    const gtk_batch::Launch &L = b->launches[i];
    le = hipLaunchKernel(L.k.fn, dim3(L.grid), dim3(L.k.block), kargs, L.lds, b->stream);
  }
  hipError_t ce = hipStreamEndCapture(b->stream, &b->graph);
  if (le != hipSuccess) BFAIL(b, GTK_EDEVICE, std::string("hipLaunchKernel: ") + hipGetErrorString(le));
  if (ce != hipSuccess) BFAIL(b, GTK_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
  BHIPCHK(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  return GTK_OK;
}

int gtk_batch_run(gtk_batch *b, gtk_stats *stats) {
  if (!b) return GTK_EINVAL;
  if (!b->gexec) BFAIL(b, GTK_ESTATE, "gtk_batch_build first");
  auto t0 = std::chrono::steady_clock::now();
  // the batch records hold absolute offsets into each problem's values:
  // a new mesh, space, reference tables or pattern invalidates them
  for (size_t p = 0; p < b->ctxs.size(); p++)
    if (b->ctxs[p]->generation != b->gens[p])
      BFAIL(b, GTK_ESTATE, "problem " + std::to_string(p) +
                               " changed its mesh, space, reference tables or pattern: gtk_batch_build again");
  BHIPCHK(b, hipSetDevice(b->device));
  // the graph reads the per-problem records through d_problems: refresh
  // them when a context moved a buffer (new coefficients, kappa reset)
  bool changed = false;
  for (size_t p = 0; p < b->ctxs.size(); p++) {
    MatrixPlan *P = find_plan(b->ctxs[p], GTK_FREE, GTK_FREE);
    if (!P) BFAIL(b, GTK_ESTATE, "a problem's pattern is gone: gtk_batch_build again");
    const gtk::MatrixArgs a = matrix_args(b->ctxs[p], *P, *b->plans[p]);
    if (std::memcmp(&a, &b->args[p], sizeof(a)) != 0) {
      b->args[p] = a;
      changed = true;
    }
  }
  if (changed)
    BHIPCHK(b, hipMemcpy(b->d_problems.p, b->args.data(), b->args.size() * sizeof(gtk::MatrixArgs),
                         hipMemcpyHostToDevice));
  changed = false;
  for (size_t p = 0; p < b->ctxs.size(); p++) {
    const gtk_batch::Load &l = *b->loads[p];
    if (l.fkind < 0) continue;
    auto it = b->ctxs[p]->vplans.find(GTK_FREE | kBatchedPlan);
    if (it == b->ctxs[p]->vplans.end() || !it->second)
      BFAIL(b, GTK_ESTATE, "a problem's load-vector plan is gone: gtk_batch_build again");
    const gtk::VectorArgs v = vector_args(b->ctxs[p], *it->second, l.f.as<double>());
    if (std::memcmp(&v, &b->vargs[p], sizeof(v)) != 0) {
      b->vargs[p] = v;
      changed = true;
    }
  }
  if (changed)
    BHIPCHK(b, hipMemcpy(b->d_loads.p, b->vargs.data(), b->vargs.size() * sizeof(gtk::VectorArgs),
                         hipMemcpyHostToDevice));
  BHIPCHK(b, hipEventRecord(b->ev0, b->stream));
  BHIPCHK(b, hipGraphLaunch(b->gexec, b->stream));
  BHIPCHK(b, hipEventRecord(b->ev1, b->stream));
  BHIPCHK(b, hipStreamSynchronize(b->stream));
  if (stats) {
    float ms = 0;
    BHIPCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    stats->kernel_ms = ms;
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats->cells = b->cells;
    stats->nnz = b->nnz;
    stats->bytes_alg = 0;
    stats->flops_alg = 0;
  }
  return GTK_OK;
}

}  // extern "C"
