// ogs gather / scatter / gather+scatter kernels for gfx950.
//
// Same operator as libs/ogs/okl/ogsKernels.okl (a CSR with one row per
// gathered node and the local indices of its copies as columns), re-designed
// for CDNA:
//   * one workgroup (256 lanes) per row block; blocks hold whole rows and at
//     most max_block_nnz (<= 4096) column entries, staged through LDS sized at
//     launch (dynamic shared memory);
//   * phase 1: lanes stream the block's column ids (coalesced) and fetch the
//     K field values of each local node into LDS; the column ids stay in
//     registers for the write-back, so the single-CSR (Sym) kernel reads them
//     once;
//   * phase 2: one lane per row folds its entries in stored order (ascending
//     local index), so sums are bit-identical to the reference's Serial order;
//   * phase 3: lanes write results back in CSR order;
//   * no atomics: every local node belongs to exactly one row and every row to
//     exactly one block, so the in-place update is race-free;
//   * workgroup ids are remapped so that XCD x owns one contiguous range of
//     row blocks (neighbouring blocks share cache lines at their edges).
// K fields per node (AoS, stride K) are processed together for K <= 4 (one
// row fold per field, one pass over the ids); larger K runs one field per
// grid row, like the reference's @outer(1) loop.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <string>

#include "spx_internal.hpp"

#pragma clang fp contract(off)

using namespace spx;

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlockNnz = 4096;
constexpr int kPer = kMaxBlockNnz / kThreads;  // column ids held per lane

template <typename T> struct Lim;
template <> struct Lim<float> { static __device__ __forceinline__ float max() { return FLT_MAX; } };
template <> struct Lim<double> { static __device__ __forceinline__ double max() { return DBL_MAX; } };
template <> struct Lim<int> { static __device__ __forceinline__ int max() { return INT_MAX; } };
template <> struct Lim<long long> { static __device__ __forceinline__ long long max() { return LLONG_MAX; } };

// ogsOperator.cpp:37-56 (host) / ogsUtils.cpp:62-103 (device defines)
template <typename T, int OP> struct Fold;
template <typename T> struct Fold<T, SPX_ADD> {
  static __device__ __forceinline__ T init() { return T(0); }
  static __device__ __forceinline__ void apply(T &a, T b) { a += b; }
};
template <typename T> struct Fold<T, SPX_MUL> {
  static __device__ __forceinline__ T init() { return T(1); }
  static __device__ __forceinline__ void apply(T &a, T b) { a *= b; }
};
template <typename T> struct Fold<T, SPX_MAX> {
  static __device__ __forceinline__ T init() { return -Lim<T>::max(); }
  static __device__ __forceinline__ void apply(T &a, T b) { a = (b > a) ? b : a; }
};
template <typename T> struct Fold<T, SPX_MIN> {
  static __device__ __forceinline__ T init() { return Lim<T>::max(); }
  static __device__ __forceinline__ void apply(T &a, T b) { a = (b < a) ? b : a; }
};

struct Launch {
  int nblocks;
  int K;        // fields per node (stride)
  int nxcd;     // XCD remap width (1 = identity)
  int cap;      // LDS capacity in entries (>= every block's nnz)
};

__device__ __forceinline__ void flag_error(int *err, int code) {
  if (threadIdx.x == 0) err[0] = code;
}

// ---------------------------------------------------------------- Sym GS --
// gatherScatter with gStarts == sStarts, gIds == sIds (ogs::Sym).
template <typename T, int OP, int KF>
__global__ __launch_bounds__(kThreads) void k_gs_sym(Launch L, const int *__restrict__ blockStarts,
                                                     const int *__restrict__ starts,
                                                     const int *__restrict__ ids, T *__restrict__ q,
                                                     int *err) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *vals = reinterpret_cast<T *>(lds_raw);
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd);
  const int kb = blockIdx.y * KF;
  const int r0 = blockStarts[b], r1 = blockStarts[b + 1];
  const int g0 = starts[r0];
  const int nnz = starts[r1] - g0;
  if (nnz > L.cap) {
    flag_error(err, 1);
    return;
  }
  const int64_t K = L.K;
  int cid[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = threadIdx.x + j * kThreads;
    if (i < nnz) cid[j] = ids[g0 + i];
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = threadIdx.x + j * kThreads;
    if (i < nnz) {
#pragma unroll
      for (int f = 0; f < KF; ++f) vals[i * KF + f] = q[(int64_t)cid[j] * K + kb + f];
    }
  }
  __syncthreads();
  for (int r = r0 + threadIdx.x; r < r1; r += kThreads) {
    const int a = starts[r] - g0, e = starts[r + 1] - g0;
#pragma unroll
    for (int f = 0; f < KF; ++f) {
      T acc = Fold<T, OP>::init();
      for (int i = a; i < e; ++i) Fold<T, OP>::apply(acc, vals[i * KF + f]);
      for (int i = a; i < e; ++i) vals[i * KF + f] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = threadIdx.x + j * kThreads;
    if (i < nnz) {
#pragma unroll
      for (int f = 0; f < KF; ++f) q[(int64_t)cid[j] * K + kb + f] = vals[i * KF + f];
    }
  }
}

// ------------------------------------------------------------ general GS --
// gather CSR (gStarts/gIds, ngrows rows) and scatter CSR (sStarts/sIds,
// nsrows rows) differ (ogs::NoTrans, ogs::Trans).  Rows at or past ngrows
// gather nothing (their value is the op's identity); rows at or past nsrows
// scatter nothing.
template <typename T, int OP, int KF>
__global__ __launch_bounds__(kThreads) void k_gs_gen(Launch L, int ngrows, int nsrows,
                                                     const int *__restrict__ blockStarts,
                                                     const int *__restrict__ gStarts,
                                                     const int *__restrict__ gIds,
                                                     const int *__restrict__ sStarts,
                                                     const int *__restrict__ sIds, T *__restrict__ q,
                                                     int *err) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *gvals = reinterpret_cast<T *>(lds_raw);
  T *svals = gvals + (size_t)L.cap * KF;
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd);
  const int kb = blockIdx.y * KF;
  const int r0 = blockStarts[b], r1 = blockStarts[b + 1];
  const int gr0 = r0 < ngrows ? r0 : ngrows, gr1 = r1 < ngrows ? r1 : ngrows;
  const int g0 = gStarts[gr0];
  const int gnnz = gStarts[gr1] - g0;
  const int sr0 = r0 < nsrows ? r0 : nsrows, sr1 = r1 < nsrows ? r1 : nsrows;
  const int s0 = sStarts[sr0];
  const int snnz = sStarts[sr1] - s0;
  if (gnnz > L.cap || snnz > L.cap) {
    flag_error(err, 2);
    return;
  }
  const int64_t K = L.K;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = threadIdx.x + j * kThreads;
    if (i < gnnz) {
      const int64_t c = gIds[g0 + i];
#pragma unroll
      for (int f = 0; f < KF; ++f) gvals[i * KF + f] = q[c * K + kb + f];
    }
  }
  __syncthreads();
  for (int r = r0 + threadIdx.x; r < r1; r += kThreads) {
    int sa = 0, se = 0;
    if (r < nsrows) {
      sa = sStarts[r] - s0;
      se = sStarts[r + 1] - s0;
    }
    int a = 0, e = 0;
    if (r < ngrows) {
      a = gStarts[r] - g0;
      e = gStarts[r + 1] - g0;
    }
#pragma unroll
    for (int f = 0; f < KF; ++f) {
      T acc = Fold<T, OP>::init();
      for (int i = a; i < e; ++i) Fold<T, OP>::apply(acc, gvals[i * KF + f]);
      for (int i = sa; i < se; ++i) svals[i * KF + f] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = threadIdx.x + j * kThreads;
    if (i < snnz) {
      const int64_t c = sIds[s0 + i];
#pragma unroll
      for (int f = 0; f < KF; ++f) q[c * K + kb + f] = svals[i * KF + f];
    }
  }
}

// ---------------------------------------------------------------- gather --
template <typename T, int OP, int KF>
__global__ __launch_bounds__(kThreads) void k_gather(Launch L, const int *__restrict__ blockStarts,
                                                     const int *__restrict__ starts,
                                                     const int *__restrict__ ids,
                                                     const T *__restrict__ v, T *__restrict__ gv,
                                                     int *err) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *vals = reinterpret_cast<T *>(lds_raw);
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd);
  const int kb = blockIdx.y * KF;
  const int r0 = blockStarts[b], r1 = blockStarts[b + 1];
  const int g0 = starts[r0];
  const int nnz = starts[r1] - g0;
  if (nnz > L.cap) {
    flag_error(err, 3);
    return;
  }
  const int64_t K = L.K;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = threadIdx.x + j * kThreads;
    if (i < nnz) {
      const int64_t c = ids[g0 + i];
#pragma unroll
      for (int f = 0; f < KF; ++f) vals[i * KF + f] = v[c * K + kb + f];
    }
  }
  __syncthreads();
  for (int r = r0 + threadIdx.x; r < r1; r += kThreads) {
    const int a = starts[r] - g0, e = starts[r + 1] - g0;
#pragma unroll
    for (int f = 0; f < KF; ++f) {
      T acc = Fold<T, OP>::init();
      for (int i = a; i < e; ++i) Fold<T, OP>::apply(acc, vals[i * KF + f]);
      gv[(int64_t)r * K + kb + f] = acc;
    }
  }
}

// --------------------------------------------------------------- scatter --
template <typename T, int KF>
__global__ __launch_bounds__(kThreads) void k_scatter(Launch L, const int *__restrict__ blockStarts,
                                                      const int *__restrict__ starts,
                                                      const int *__restrict__ ids,
                                                      const T *__restrict__ gv, T *__restrict__ v,
                                                      int *err) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *vals = reinterpret_cast<T *>(lds_raw);
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd);
  const int kb = blockIdx.y * KF;
  const int r0 = blockStarts[b], r1 = blockStarts[b + 1];
  const int g0 = starts[r0];
  const int nnz = starts[r1] - g0;
  if (nnz > L.cap) {
    flag_error(err, 4);
    return;
  }
  const int64_t K = L.K;
  for (int r = r0 + threadIdx.x; r < r1; r += kThreads) {
    const int a = starts[r] - g0, e = starts[r + 1] - g0;
#pragma unroll
    for (int f = 0; f < KF; ++f) {
      const T x = gv[(int64_t)r * K + kb + f];
      for (int i = a; i < e; ++i) vals[i * KF + f] = x;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = threadIdx.x + j * kThreads;
    if (i < nnz) {
      const int64_t c = ids[g0 + i];
#pragma unroll
      for (int f = 0; f < KF; ++f) v[c * K + kb + f] = vals[i * KF + f];
    }
  }
}

// --------------------------------------------------------------- extract --
template <typename T>
__global__ __launch_bounds__(kThreads) void k_extract(int64_t total, int K, const int *__restrict__ ids,
                                                      const T *__restrict__ q, T *__restrict__ gq) {
  int64_t n = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * kThreads;
  for (; n < total; n += st) {
    const int64_t g = n / K;
    const int k = (int)(n - g * K);
    gq[n] = q[k + (int64_t)ids[g] * K];
  }
}

// ------------------------------------------------------------------ host --

size_t type_size(int type) {
  switch (type) {
    case SPX_FLOAT: return 4;
    case SPX_DOUBLE: return 8;
    case SPX_INT32: return 4;
    case SPX_INT64: return 8;
    default: return 0;
  }
}

spx_tuning_t ogs_defaults() {
  spx_tuning_t d{};
  d.block_threads = kThreads;
  d.xcd_remap = 1;
  d.unroll = 0;
  d.grid_blocks = 0;
  return d;
}

// Fields per pass: all K together when the LDS tile fits 64 KiB, else one.
int fields_per_pass(int K, int cap, size_t tsize, int arrays) {
  if (K >= 1 && K <= 4 && (size_t)cap * K * tsize * arrays <= 65536) return K;
  return 1;
}

int check_common(const char *name, int type, int K, int32_t nblocks, int32_t cap) {
  if (!type_size(type)) return fail(SPX_ERR_ARG, std::string(name) + ": unknown type");
  if (K < 1) return fail(SPX_ERR_ARG, std::string(name) + ": K must be >= 1");
  if (nblocks < 0) return fail(SPX_ERR_ARG, std::string(name) + ": negative nblocks");
  if (cap < 1 || cap > kMaxBlockNnz)
    return fail(SPX_ERR_ARG, std::string(name) + ": max_block_nnz must be in [1, 4096]");
  return SPX_OK;
}

Launch make_launch(const char *kname, int32_t nblocks, int K, int32_t cap) {
  spx_tuning_t t = tuning_for(kname, ogs_defaults());
  Launch L;
  L.nblocks = nblocks;
  L.K = K;
  L.nxcd = t.xcd_remap ? 8 : 1;
  L.cap = cap;
  return L;
}

template <typename T, int OP>
void launch_gs_sym(const Launch &L, int kf, const int *bs, const int *st, const int *ids, T *q) {
  dim3 grid(L.nblocks, kf == 1 ? L.K : 1);
  const size_t lds = (size_t)L.cap * kf * sizeof(T);
  int *err = kernel_error_flag();
  switch (kf) {
    case 1: hipLaunchKernelGGL((k_gs_sym<T, OP, 1>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, q, err); break;
    case 2: hipLaunchKernelGGL((k_gs_sym<T, OP, 2>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, q, err); break;
    case 3: hipLaunchKernelGGL((k_gs_sym<T, OP, 3>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, q, err); break;
    default: hipLaunchKernelGGL((k_gs_sym<T, OP, 4>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, q, err); break;
  }
}

template <typename T, int OP>
void launch_gs_gen(const Launch &L, int kf, int ngrows, int nsrows, const int *bs, const int *gs, const int *gi,
                   const int *ss, const int *si, T *q) {
  dim3 grid(L.nblocks, kf == 1 ? L.K : 1);
  const size_t lds = 2 * (size_t)L.cap * kf * sizeof(T);
  int *err = kernel_error_flag();
  switch (kf) {
    case 1: hipLaunchKernelGGL((k_gs_gen<T, OP, 1>), grid, dim3(kThreads), lds, current_stream(), L, ngrows, nsrows, bs, gs, gi, ss, si, q, err); break;
    case 2: hipLaunchKernelGGL((k_gs_gen<T, OP, 2>), grid, dim3(kThreads), lds, current_stream(), L, ngrows, nsrows, bs, gs, gi, ss, si, q, err); break;
    case 3: hipLaunchKernelGGL((k_gs_gen<T, OP, 3>), grid, dim3(kThreads), lds, current_stream(), L, ngrows, nsrows, bs, gs, gi, ss, si, q, err); break;
    default: hipLaunchKernelGGL((k_gs_gen<T, OP, 4>), grid, dim3(kThreads), lds, current_stream(), L, ngrows, nsrows, bs, gs, gi, ss, si, q, err); break;
  }
}

template <typename T, int OP>
void launch_gather(const Launch &L, int kf, const int *bs, const int *st, const int *ids, const T *v, T *gv) {
  dim3 grid(L.nblocks, kf == 1 ? L.K : 1);
  const size_t lds = (size_t)L.cap * kf * sizeof(T);
  int *err = kernel_error_flag();
  switch (kf) {
    case 1: hipLaunchKernelGGL((k_gather<T, OP, 1>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, v, gv, err); break;
    case 2: hipLaunchKernelGGL((k_gather<T, OP, 2>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, v, gv, err); break;
    case 3: hipLaunchKernelGGL((k_gather<T, OP, 3>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, v, gv, err); break;
    default: hipLaunchKernelGGL((k_gather<T, OP, 4>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, v, gv, err); break;
  }
}

template <typename T>
void launch_scatter(const Launch &L, int kf, const int *bs, const int *st, const int *ids, const T *gv, T *v) {
  dim3 grid(L.nblocks, kf == 1 ? L.K : 1);
  const size_t lds = (size_t)L.cap * kf * sizeof(T);
  int *err = kernel_error_flag();
  switch (kf) {
    case 1: hipLaunchKernelGGL((k_scatter<T, 1>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, gv, v, err); break;
    case 2: hipLaunchKernelGGL((k_scatter<T, 2>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, gv, v, err); break;
    case 3: hipLaunchKernelGGL((k_scatter<T, 3>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, gv, v, err); break;
    default: hipLaunchKernelGGL((k_scatter<T, 4>), grid, dim3(kThreads), lds, current_stream(), L, bs, st, ids, gv, v, err); break;
  }
}

#define SPX_TYPE_SWITCH(type, T, ...)                 \
  switch (type) {                                     \
    case SPX_FLOAT: { typedef float T; __VA_ARGS__; break; }       \
    case SPX_DOUBLE: { typedef double T; __VA_ARGS__; break; }     \
    case SPX_INT32: { typedef int T; __VA_ARGS__; break; }         \
    default: { typedef long long T; __VA_ARGS__; break; }          \
  }

#define SPX_OP_SWITCH(op, OPC, ...)                   \
  switch (op) {                                       \
    case SPX_ADD: { constexpr int OPC = SPX_ADD; __VA_ARGS__; break; } \
    case SPX_MUL: { constexpr int OPC = SPX_MUL; __VA_ARGS__; break; } \
    case SPX_MAX: { constexpr int OPC = SPX_MAX; __VA_ARGS__; break; } \
    default: { constexpr int OPC = SPX_MIN; __VA_ARGS__; break; }      \
  }

int check_op(const char *name, int op) {
  if (op < SPX_ADD || op > SPX_MIN) return fail(SPX_ERR_ARG, std::string(name) + ": unknown op");
  return SPX_OK;
}

}  // namespace

extern "C" {

int spx_ogs_gather(int type, int op, int K, int32_t nblocks, const int32_t *blockStarts,
                   const int32_t *rowStarts, const int32_t *colIds, int32_t cap, const void *v,
                   void *gv) {
  SPX_NEED_DEVICE();
  int rc = check_common("spx_ogs_gather", type, K, nblocks, cap);
  if (rc) return rc;
  if ((rc = check_op("spx_ogs_gather", op))) return rc;
  if (nblocks == 0) return SPX_OK;
  if (!blockStarts || !rowStarts || !colIds || !v || !gv)
    return fail(SPX_ERR_ARG, "spx_ogs_gather: null pointer");
  Launch L = make_launch("ogs_gather", nblocks, K, cap);
  const int kf = fields_per_pass(K, cap, type_size(type), 1);
  SPX_TYPE_SWITCH(type, T, SPX_OP_SWITCH(op, OPC,
      launch_gather<T, OPC>(L, kf, blockStarts, rowStarts, colIds, (const T *)v, (T *)gv)));
  SPX_LAUNCHED("spx_ogs_gather");
  return SPX_OK;
}

int spx_ogs_scatter(int type, int K, int32_t nblocks, const int32_t *blockStarts,
                    const int32_t *rowStarts, const int32_t *colIds, int32_t cap, const void *gv,
                    void *v) {
  SPX_NEED_DEVICE();
  int rc = check_common("spx_ogs_scatter", type, K, nblocks, cap);
  if (rc) return rc;
  if (nblocks == 0) return SPX_OK;
  if (!blockStarts || !rowStarts || !colIds || !v || !gv)
    return fail(SPX_ERR_ARG, "spx_ogs_scatter: null pointer");
  Launch L = make_launch("ogs_scatter", nblocks, K, cap);
  const int kf = fields_per_pass(K, cap, type_size(type), 1);
  SPX_TYPE_SWITCH(type, T, launch_scatter<T>(L, kf, blockStarts, rowStarts, colIds, (const T *)gv, (T *)v));
  SPX_LAUNCHED("spx_ogs_scatter");
  return SPX_OK;
}

int spx_ogs_gather_scatter(int type, int op, int K, int32_t nblocks, const int32_t *blockStarts,
                           int32_t ngrows, const int32_t *gStarts, const int32_t *gIds,
                           int32_t nsrows, const int32_t *sStarts, const int32_t *sIds,
                           int32_t cap, void *q) {
  SPX_NEED_DEVICE();
  int rc = check_common("spx_ogs_gather_scatter", type, K, nblocks, cap);
  if (rc) return rc;
  if ((rc = check_op("spx_ogs_gather_scatter", op))) return rc;
  if (ngrows < 0 || nsrows < 0)
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter: negative row count");
  if (nblocks == 0) return SPX_OK;
  if (!blockStarts || !gStarts || !gIds || !sStarts || !sIds || !q)
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter: null pointer");
  const bool sym = (gStarts == sStarts) && (gIds == sIds) && (ngrows == nsrows);
  if (sym) {
    Launch L = make_launch("ogs_gs", nblocks, K, cap);
    const int kf = fields_per_pass(K, cap, type_size(type), 1);
    SPX_TYPE_SWITCH(type, T, SPX_OP_SWITCH(op, OPC,
        launch_gs_sym<T, OPC>(L, kf, blockStarts, gStarts, gIds, (T *)q)));
  } else {
    Launch L = make_launch("ogs_gs", nblocks, K, cap);
    const int kf = fields_per_pass(K, cap, type_size(type), 2);
    SPX_TYPE_SWITCH(type, T, SPX_OP_SWITCH(op, OPC,
        launch_gs_gen<T, OPC>(L, kf, ngrows, nsrows, blockStarts, gStarts, gIds, sStarts, sIds, (T *)q)));
  }
  SPX_LAUNCHED("spx_ogs_gather_scatter");
  return SPX_OK;
}

int spx_ogs_extract(int type, int K, int32_t n_ids, const int32_t *ids, const void *q, void *gq) {
  SPX_NEED_DEVICE();
  if (!type_size(type)) return fail(SPX_ERR_ARG, "spx_ogs_extract: unknown type");
  if (K < 1 || n_ids < 0) return fail(SPX_ERR_ARG, "spx_ogs_extract: bad K or n_ids");
  if (n_ids == 0) return SPX_OK;
  if (!ids || !q || !gq) return fail(SPX_ERR_ARG, "spx_ogs_extract: null pointer");
  const int64_t total = (int64_t)n_ids * K;
  int64_t g = (total + kThreads - 1) / kThreads;
  if (g > 65536) g = 65536;
  SPX_TYPE_SWITCH(type, T,
      hipLaunchKernelGGL(k_extract<T>, dim3((unsigned)g), dim3(kThreads), 0, current_stream(), total, K,
                         ids, (const T *)q, (T *)gq));
  SPX_LAUNCHED("spx_ogs_extract");
  return SPX_OK;
}

}  // extern "C"
