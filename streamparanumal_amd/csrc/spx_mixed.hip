// Reduced-precision storage variants of BS8 and the BS5 CG update.
//
// Values live in HBM as FP32 or BF16 to cut the bytes moved.  Arithmetic:
//   * gather+scatter: each row's repeated-node sum is taken in FP64 (members
//     converted exactly, folded in stored order) and rounded once to the
//     storage type;
//   * CG update: x += alpha p and r -= alpha Ap in FP32, results rounded to
//     the storage type; rdotr = sum of the stored r squared, in FP64.
// FP64 storage is served by the plain kernels.
#include <hip/hip_runtime.h>

#include <string>

#include "spx_internal.hpp"
#include "spx_numeric.hpp"
#include "spx_ogs_kernels.hpp"

#pragma clang fp contract(off)

using namespace spx;
using namespace spx::ogsk;

namespace {

// ------------------------------------------------------ mixed gather+scatter --
// Sym, Add.  LDS: vals[cap*KF] of S, then 16-bit row starts.
template <typename S, int KF, int PER>
__global__ __launch_bounds__(kThreads) void k_gs_mixed(Launch L, const int *__restrict__ blockStarts,
                                                       const int *__restrict__ starts,
                                                       const int *__restrict__ ids, S *__restrict__ q) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  S *vals = reinterpret_cast<S *>(lds_raw);
  unsigned short *rs = reinterpret_cast<unsigned short *>(lds_raw + (size_t)L.cap * KF * sizeof(S));
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd);
  const int kb = blockIdx.y * KF;
  const int r0 = blockStarts[b], r1 = blockStarts[b + 1];
  const int nrows = r1 - r0;
  const int g0 = starts[r0];
  const int nnz = starts[r1] - g0;
  if (nnz > L.cap || nrows > L.cap) {
    flag_error(L.err, 5);
    return;
  }
  if (nnz == 0) return;
  const int64_t K = L.K;
  const int t = threadIdx.x;
  int cid[PER];
  load_ids<PER>(cid, ids, g0, nnz);
  stage_starts16<PER>(rs, starts, r0, nrows, g0);
  S v[PER][KF];
#pragma unroll
  for (int j = 0; j < PER; ++j)
#pragma unroll
    for (int f = 0; f < KF; ++f) v[j][f] = q[(int64_t)cid[j] * K + kb + f];
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t + j * kThreads < nnz)
#pragma unroll
      for (int f = 0; f < KF; ++f) vals[(t + j * kThreads) * KF + f] = v[j][f];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t + j * kThreads < nrows) {
      const int a = rs[t + j * kThreads], e = rs[t + j * kThreads + 1];
#pragma unroll
      for (int f = 0; f < KF; ++f) {
        double acc = 0.0;
        for (int i = a; i < e; ++i) acc += Store<S>::to_f64(vals[i * KF + f]);
        const S out = Store<S>::from_f64(acc);
        for (int i = a; i < e; ++i) vals[i * KF + f] = out;
      }
    }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t + j * kThreads < nnz)
#pragma unroll
      for (int f = 0; f < KF; ++f) q[(int64_t)cid[j] * K + kb + f] = vals[(t + j * kThreads) * KF + f];
}

template <typename S>
int launch_gs_mixed(const Launch &L, const int *bs, const int *st, const int *ids, S *q) {
  const size_t lds = (size_t)L.cap * L.kf * sizeof(S) + (size_t)(L.cap + 1) * sizeof(unsigned short);
  if (lds > 65536) return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_mixed: tile too large for LDS");
  const dim3 grid(L.nblocks, L.kf == 1 ? L.K : 1);
#define SPX_GSM(KF, PER) \
  hipLaunchKernelGGL((k_gs_mixed<S, KF, PER>), grid, dim3(kThreads), lds, L.stream, L, bs, st, ids, q)
  if (L.kf == 1 && L.per == 4) SPX_GSM(1, 4);
  else if (L.kf == 1 && L.per == 8) SPX_GSM(1, 8);
  else if (L.kf == 1 && L.per == 16) SPX_GSM(1, 16);
  else if (L.kf == 2 && L.per == 4) SPX_GSM(2, 4);
  else if (L.kf == 2 && L.per == 8) SPX_GSM(2, 8);
  else if (L.kf == 3 && L.per == 4) SPX_GSM(3, 4);
  else if (L.kf == 4 && L.per == 4) SPX_GSM(4, 4);
  else return fail(SPX_ERR_UNSUPPORTED, "spx_ogs_gather_scatter_mixed: no kernel for this tile");
#undef SPX_GSM
  return SPX_OK;
}

// ------------------------------------------------------------ mixed CG update --
// Chunks of 4 consecutive entries; lane t of block b takes chunks
// t + b*256 + j*256*nblock in j order, U chunks in flight, FP64 sum.
template <typename S> struct Vec4;
template <> struct Vec4<float> {
  typedef float4 type;
};
template <> struct Vec4<bf16_t> {
  typedef ushort4 type;
};

template <typename S>
__device__ __forceinline__ void ld4(const S *p, int64_t c, float (&o)[4]) {
  if constexpr (sizeof(S) == 4) {
    const float4 v = reinterpret_cast<const float4 *>(p)[c];
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  } else {
    const ushort4 v = reinterpret_cast<const ushort4 *>(p)[c];
    o[0] = bf16_bits_to_f32(v.x); o[1] = bf16_bits_to_f32(v.y);
    o[2] = bf16_bits_to_f32(v.z); o[3] = bf16_bits_to_f32(v.w);
  }
}

// stores the rounded values and returns them (as stored) in o
template <typename S>
__device__ __forceinline__ void st4(S *p, int64_t c, float (&o)[4]) {
  if constexpr (sizeof(S) == 4) {
    reinterpret_cast<float4 *>(p)[c] = make_float4(o[0], o[1], o[2], o[3]);
  } else {
    ushort4 v;
    v.x = f32_to_bf16_bits(o[0]); v.y = f32_to_bf16_bits(o[1]);
    v.z = f32_to_bf16_bits(o[2]); v.w = f32_to_bf16_bits(o[3]);
    reinterpret_cast<ushort4 *>(p)[c] = v;
    o[0] = bf16_bits_to_f32(v.x); o[1] = bf16_bits_to_f32(v.y);
    o[2] = bf16_bits_to_f32(v.z); o[3] = bf16_bits_to_f32(v.w);
  }
}

constexpr int kCgBlocks = 256;
constexpr int kCgU = 8;

template <typename S>
__global__ __launch_bounds__(256) void k_cg_mixed1(int64_t n, float alpha, const S *__restrict__ p,
                                                   const S *__restrict__ Ap, S *__restrict__ x, S *__restrict__ r,
                                                   double *__restrict__ partials) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  const int64_t nchunk = n / 4;
  const int64_t stride = (int64_t)256 * gridDim.x;
  int64_t c = t + (int64_t)blockIdx.x * 256;
  double acc = 0.0;
  for (; c + (int64_t)(kCgU - 1) * stride < nchunk; c += (int64_t)kCgU * stride) {
    float vp[kCgU][4], vap[kCgU][4], vx[kCgU][4], vr[kCgU][4];
#pragma unroll
    for (int u = 0; u < kCgU; ++u) {
      ld4(p, c + u * stride, vp[u]);
      ld4(Ap, c + u * stride, vap[u]);
      ld4(x, c + u * stride, vx[u]);
      ld4(r, c + u * stride, vr[u]);
    }
#pragma unroll
    for (int u = 0; u < kCgU; ++u) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float ap = alpha * vp[u][e];
        vx[u][e] = vx[u][e] + ap;
        const float aAp = alpha * vap[u][e];
        vr[u][e] = vr[u][e] - aAp;
      }
      st4(x, c + u * stride, vx[u]);
      st4(r, c + u * stride, vr[u]);
#pragma unroll
      for (int e = 0; e < 4; ++e) acc += (double)vr[u][e] * (double)vr[u][e];
    }
  }
  for (; c < nchunk; c += stride) {
    float vp[4], vap[4], vx[4], vr[4];
    ld4(p, c, vp);
    ld4(Ap, c, vap);
    ld4(x, c, vx);
    ld4(r, c, vr);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float ap = alpha * vp[e];
      vx[e] = vx[e] + ap;
      const float aAp = alpha * vap[e];
      vr[e] = vr[e] - aAp;
    }
    st4(x, c, vx);
    st4(r, c, vr);
#pragma unroll
    for (int e = 0; e < 4; ++e) acc += (double)vr[e] * (double)vr[e];
  }
  // the n % 4 tail entries: lane 0 of block 0, in order
  if (blockIdx.x == 0 && t == 0) {
    for (int64_t i = nchunk * 4; i < n; ++i) {
      const float ap = alpha * Store<S>::to_f32(p[i]);
      const float xn = Store<S>::to_f32(x[i]) + ap;
      const float aAp = alpha * Store<S>::to_f32(Ap[i]);
      const float rn = Store<S>::to_f32(r[i]) - aAp;
      x[i] = Store<S>::from_f32(xn);
      r[i] = Store<S>::from_f32(rn);
      const double rs = Store<S>::to_f64(r[i]);
      acc += rs * rs;
    }
  }
  s[t] = acc;
  __syncthreads();
#pragma unroll
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w) s[t] += s[t + w];
    __syncthreads();
  }
  if (t == 0) partials[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_sum_partials(int nb, const double *__restrict__ partials,
                                                      double *__restrict__ out) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  s[t] = t < nb ? partials[t] : 0.0;
  __syncthreads();
#pragma unroll
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w) s[t] += s[t + w];
    __syncthreads();
  }
  if (t == 0) out[0] = s[0];
}

}  // namespace

extern "C" {

int spx_ogs_gather_scatter_mixed(int storage, int K, int32_t nblocks, const int32_t *blockStarts,
                                 const int32_t *rowStarts, const int32_t *colIds, int32_t cap, void *q) {
  SPX_NEED_DEVICE();
  if (storage == SPX_STORE_F64)
    return spx_ogs_gather_scatter(SPX_DOUBLE, SPX_ADD, K, nblocks, blockStarts, 0, rowStarts, colIds, 0,
                                  rowStarts, colIds, cap, q);
  if (storage != SPX_STORE_F32 && storage != SPX_STORE_BF16)
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_mixed: unknown storage type");
  if (K < 1 || K > 65535) return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_mixed: bad K");
  if (nblocks < 0 || cap < 1 || cap > 4096) return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_mixed: bad blocks");
  if (nblocks == 0) return SPX_OK;
  if (!blockStarts || !rowStarts || !colIds || !q)
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_mixed: null pointer");
  const size_t ts = storage == SPX_STORE_F32 ? 4 : 2;
  Launch L;
  L.nblocks = nblocks;
  L.K = K;
  L.nxcd = 1;
  L.cap = cap;
  L.kf = 1;
  if (K >= 2 && K <= 4 && pick_per(K, cap) && (size_t)cap * K * ts + (size_t)(cap + 1) * 2 <= 65536) L.kf = K;
  L.per = pick_per(L.kf, cap);
  L.variant = 0;
  L.stream = current_stream();
  L.err = kernel_error_flag();
  if (!L.per) return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_mixed: tile too large");
  int rc = storage == SPX_STORE_F32 ? launch_gs_mixed(L, blockStarts, rowStarts, colIds, (float *)q)
                                    : launch_gs_mixed(L, blockStarts, rowStarts, colIds, (bf16_t *)q);
  if (rc) return rc;
  SPX_LAUNCHED("spx_ogs_gather_scatter_mixed");
  return SPX_OK;
}

int spx_bs5_full_mixed(int storage, int64_t n, double alpha, const void *p, const void *Ap, void *x, void *r,
                       double *tmp, double *rdotr) {
  SPX_NEED_DEVICE();
  if (storage != SPX_STORE_F32 && storage != SPX_STORE_BF16)
    return fail(SPX_ERR_ARG, "spx_bs5_full_mixed: storage must be SPX_STORE_F32 or SPX_STORE_BF16");
  if (n < 0) return fail(SPX_ERR_ARG, "spx_bs5_full_mixed: negative n");
  if (!tmp || !rdotr) return fail(SPX_ERR_ARG, "spx_bs5_full_mixed: null tmp/rdotr");
  if (n > 0 && (!p || !Ap || !x || !r)) return fail(SPX_ERR_ARG, "spx_bs5_full_mixed: null vector");
  const size_t al = storage == SPX_STORE_F32 ? 16 : 8;
  if (!aligned(p, al) || !aligned(Ap, al) || !aligned(x, al) || !aligned(r, al))
    return fail(SPX_ERR_ARG, "spx_bs5_full_mixed: vectors must be aligned to 4 entries");
  int64_t nb = (n / 4 + 255) / 256;
  nb = nb < 1 ? 1 : (nb > kCgBlocks ? kCgBlocks : nb);
  if (storage == SPX_STORE_F32)
    hipLaunchKernelGGL(k_cg_mixed1<float>, dim3((unsigned)nb), dim3(256), 0, current_stream(), n, (float)alpha,
                       (const float *)p, (const float *)Ap, (float *)x, (float *)r, tmp);
  else
    hipLaunchKernelGGL(k_cg_mixed1<bf16_t>, dim3((unsigned)nb), dim3(256), 0, current_stream(), n, (float)alpha,
                       (const bf16_t *)p, (const bf16_t *)Ap, (bf16_t *)x, (bf16_t *)r, tmp);
  SPX_LAUNCHED("spx_bs5_full_mixed");
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, current_stream(), (int)nb, tmp, rdotr);
  SPX_LAUNCHED("spx_bs5_full_mixed");
  return SPX_OK;
}

}  // extern "C"
