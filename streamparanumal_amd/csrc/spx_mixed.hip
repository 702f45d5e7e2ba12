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
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd, L.xcdLog2);
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
  else if (L.kf == 3 && L.per == 8) SPX_GSM(3, 8);
  else if (L.kf == 4 && L.per == 4) SPX_GSM(4, 4);
  else if (L.kf == 4 && L.per == 8) SPX_GSM(4, 8);
  else return fail(SPX_ERR_UNSUPPORTED, "spx_ogs_gather_scatter_mixed: no kernel for this tile");
#undef SPX_GSM
  return SPX_OK;
}

// ------------------------------------------ mixed gather+scatter, sorted runs --
// FP64 row sums over the stored members, one rounding (k_gs_srb's fold policy).
template <typename S> struct FoldWide {
  typedef double Acc;
  static __device__ __forceinline__ double init() { return 0.0; }
  static __device__ __forceinline__ void apply(double &a, S x) { a += Store<S>::to_f64(x); }
  static __device__ __forceinline__ S out(double a) { return Store<S>::from_f64(a); }
};

template <typename S>
int launch_gs_srb_mixed(const Launch &L, const SrbArgs &A, S *q) {
  Launch W = L;
  const bool wave = L.variant == 2 && L.cap <= 16 * 64;
  const int lanes = wave ? 64 : kThreads;
  // all fields per pass if they fit the registers and LDS, else half (K even), else one
  auto lds_of = [&](int kf) { return srb_gs_lds_bytes<S>(W.cap, kf); };
  auto fits = [&](int kf) {
    W.kf = kf;
    W.per = pick_per(kf, W.cap * (kThreads / lanes));
    return W.per > 0 && lds_of(kf) <= 65536;
  };
  const int kf0 = L.kf;
  if (!fits(kf0) && !(kf0 > 2 && kf0 % 2 == 0 && fits(kf0 / 2)) && !fits(1)) {
    if (!W.per) return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_srb_mixed: tile too large");
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_srb_mixed: tile too large for LDS");
  }
  const size_t lds = lds_of(W.kf);
  const dim3 grid(W.nblocks, W.K / W.kf);
  // non-temporal values (spx_srb_t.nontemporal) for one pass over all K
  // fields of K*sizeof(S) = 2^e bytes, value-major fields only
  const int bytes = W.K * (int)sizeof(S);
  int e = 0;
  while ((1 << e) < bytes) ++e;
  const bool ntv = !wave && W.kf == W.K && bytes <= 32 && (1 << e) == bytes && ((W.ntMask >> e) & 1) &&
                   !(W.kf > 1 && sizeof(S) == 2);
#define SPX_GSRM(KF, PER)                                                                                   \
  do {                                                                                                     \
    if constexpr (KF == 1 || sizeof(S) == 4) {                                                             \
      if (ntv) {                                                                                           \
        hipLaunchKernelGGL((k_gs_srb<S, SPX_ADD, KF, PER, kThreads, FoldWide<S>, kAll, true>), grid,         \
                           dim3(kThreads), lds, W.stream, W, SRB_ARGS(A), q, nullptr);                    \
        break;                                                                                             \
      }                                                                                                    \
    }                                                                                                      \
    if (wave)                                                                                              \
      hipLaunchKernelGGL((k_gs_srb<S, SPX_ADD, KF, PER, 64, FoldWide<S>>), grid, dim3(64), lds, W.stream, W, \
                         SRB_ARGS(A), q, nullptr);                                                         \
    else                                                                                                   \
      hipLaunchKernelGGL((k_gs_srb<S, SPX_ADD, KF, PER, kThreads, FoldWide<S>>), grid, dim3(kThreads), lds,  \
                         W.stream, W, SRB_ARGS(A), q, nullptr);                                            \
  } while (0)
  if (W.kf == 1 && W.per == 4) SPX_GSRM(1, 4);
  else if (W.kf == 1 && W.per == 8) SPX_GSRM(1, 8);
  else if (W.kf == 1 && W.per == 16) SPX_GSRM(1, 16);
  else if (W.kf == 2 && W.per == 4) SPX_GSRM(2, 4);
  else if (W.kf == 2 && W.per == 8) SPX_GSRM(2, 8);
  else if (W.kf == 3 && W.per == 4) SPX_GSRM(3, 4);
  else if (W.kf == 3 && W.per == 8) SPX_GSRM(3, 8);
  else if (W.kf == 4 && W.per == 4) SPX_GSRM(4, 4);
  else if (W.kf == 4 && W.per == 8) SPX_GSRM(4, 8);
  else return fail(SPX_ERR_UNSUPPORTED, "spx_ogs_gather_scatter_srb_mixed: no kernel for this tile");
#undef SPX_GSRM
  return SPX_OK;
}

// ------------------------------------------------------------ mixed CG update --
// Chunks of 4 consecutive entries (one float4 or 8-byte bf16 quad per vector);
// a fixed grid walks them with stride 256*gridDim, U chunks per lane in flight,
// and folds each lane's r^2 in FP64.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
template <typename S> struct Quad;
template <> struct Quad<float> {
  typedef f32x4 type;
};
template <> struct Quad<bf16_t> {
  typedef u32x2 type;
};

template <bool NT, typename V> __device__ __forceinline__ V ldq(const V *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename V> __device__ __forceinline__ void stq(V *p, V v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

__device__ __forceinline__ void unpack(const f32x4 &v, float (&o)[4]) {
  o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
__device__ __forceinline__ void unpack(const u32x2 &v, float (&o)[4]) {
  o[0] = __uint_as_float(v.x << 16);
  o[1] = __uint_as_float(v.x & 0xffff0000u);
  o[2] = __uint_as_float(v.y << 16);
  o[3] = __uint_as_float(v.y & 0xffff0000u);
}
// round to the storage type; o receives the stored values
__device__ __forceinline__ void pack(float (&o)[4], f32x4 &v) {
  v.x = o[0]; v.y = o[1]; v.z = o[2]; v.w = o[3];
}
__device__ __forceinline__ void pack(float (&o)[4], u32x2 &v) {
  const uint32_t b0 = f32_to_bf16_bits(o[0]), b1 = f32_to_bf16_bits(o[1]);
  const uint32_t b2 = f32_to_bf16_bits(o[2]), b3 = f32_to_bf16_bits(o[3]);
  v.x = b0 | (b1 << 16);
  v.y = b2 | (b3 << 16);
  o[0] = __uint_as_float(b0 << 16); o[1] = __uint_as_float(b1 << 16);
  o[2] = __uint_as_float(b2 << 16); o[3] = __uint_as_float(b3 << 16);
}

constexpr int kCgGridMax = 2048;  // partials written by stage 1

// x and p of chunk c + xs (mod the chunk count) are updated beside r and Ap
// of chunk c, so the x / p streams run at another address phase than the
// r / Ap streams (as spx_bs5_full); each x entry is still updated once from
// its own p.
template <typename S, int U, bool NT>
__global__ __launch_bounds__(256) void k_cg_mixed1(int64_t n, float alpha, const S *__restrict__ p_,
                                                   const S *__restrict__ Ap_, S *__restrict__ x_,
                                                   S *__restrict__ r_, double *__restrict__ partials,
                                                   int64_t xs) {
  typedef typename Quad<S>::type V;
  const V *__restrict__ p = reinterpret_cast<const V *>(p_);
  const V *__restrict__ Ap = reinterpret_cast<const V *>(Ap_);
  V *__restrict__ x = reinterpret_cast<V *>(x_);
  V *__restrict__ r = reinterpret_cast<V *>(r_);
  __shared__ double s[256];
  const int t = threadIdx.x;
  const int64_t nchunk = n / 4;
  const int64_t stride = (int64_t)256 * gridDim.x;
  int64_t c = t + (int64_t)blockIdx.x * 256;
  double acc = 0.0;
  auto xc = [&](int64_t cc) { const int64_t j = cc + xs; return j >= nchunk ? j - nchunk : j; };
  for (; c + (int64_t)(U - 1) * stride < nchunk; c += (int64_t)U * stride) {
    V vp[U], vap[U], vx[U], vr[U];
    int64_t cx[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cx[u] = xc(c + u * stride);
      vp[u] = ldq<NT>(p + cx[u]);
      vap[u] = ldq<NT>(Ap + c + u * stride);
      vx[u] = ldq<NT>(x + cx[u]);
      vr[u] = ldq<NT>(r + c + u * stride);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      float fp[4], fap[4], fx[4], fr[4];
      unpack(vp[u], fp);
      unpack(vap[u], fap);
      unpack(vx[u], fx);
      unpack(vr[u], fr);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float ap = alpha * fp[e];
        fx[e] = fx[e] + ap;
        const float aAp = alpha * fap[e];
        fr[e] = fr[e] - aAp;
      }
      pack(fx, vx[u]);
      pack(fr, vr[u]);
      stq<NT>(x + cx[u], vx[u]);
      stq<NT>(r + c + u * stride, vr[u]);
#pragma unroll
      for (int e = 0; e < 4; ++e) acc += (double)fr[e] * (double)fr[e];
    }
  }
  for (; c < nchunk; c += stride) {
    float fp[4], fap[4], fx[4], fr[4];
    const int64_t cx = xc(c);
    unpack(p[cx], fp);
    unpack(Ap[c], fap);
    unpack(x[cx], fx);
    unpack(r[c], fr);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float ap = alpha * fp[e];
      fx[e] = fx[e] + ap;
      const float aAp = alpha * fap[e];
      fr[e] = fr[e] - aAp;
    }
    V vx, vr;
    pack(fx, vx);
    pack(fr, vr);
    x[cx] = vx;
    r[c] = vr;
#pragma unroll
    for (int e = 0; e < 4; ++e) acc += (double)fr[e] * (double)fr[e];
  }
  // the n % 4 tail entries: lane 0 of block 0, in order
  if (blockIdx.x == 0 && t == 0) {
    for (int64_t i = nchunk * 4; i < n; ++i) {
      const float ap = alpha * Store<S>::to_f32(p_[i]);
      const float xn = Store<S>::to_f32(x_[i]) + ap;
      const float aAp = alpha * Store<S>::to_f32(Ap_[i]);
      const float rn = Store<S>::to_f32(r_[i]) - aAp;
      x_[i] = Store<S>::from_f32(xn);
      r_[i] = Store<S>::from_f32(rn);
      const double rs = Store<S>::to_f64(r_[i]);
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

// one block: lane t folds partials t, t+256, ... then a tree
__global__ __launch_bounds__(256) void k_sum_partials(int nb, const double *__restrict__ partials,
                                                      double *__restrict__ out) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  double a = 0.0;
  for (int i = t; i < nb; i += 256) a += partials[i];
  s[t] = a;
  __syncthreads();
#pragma unroll
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w) s[t] += s[t + w];
    __syncthreads();
  }
  if (t == 0) out[0] = s[0];
}

template <typename S>
int launch_cg_mixed(int64_t n, float alpha, const S *p, const S *Ap, S *x, S *r, double *tmp, double *rdotr) {
  // defaults measured on MI355X at n = 2^27: 512 blocks, streaming loads and
  // stores, 4 chunks per lane for FP32 (5.56 TB/s), 2 for BF16 (6.02 TB/s)
  spx_tuning_t d{};
  d.unroll = sizeof(S) == 4 ? 4 : 2;
  d.grid_blocks = 512;
  d.nontemporal = 1;
  d.xcd_remap = 0;
  d.variant = 0;
  const spx_tuning_t tu = tuning_for("cg_mixed", d);
  const int U = tu.unroll;
  const bool nt = tu.nontemporal > 0;
  int64_t grid = tu.grid_blocks;
  const int64_t need = (n / 4 + 255) / 256;
  if (grid > need) grid = need;
  if (grid < 1) grid = 1;
  if (grid > kCgGridMax) grid = kCgGridMax;
  const dim3 g((unsigned)grid);
  // x / p shift: half the chunks by default (variant v > 0: v chunks)
  const int64_t nchunk = n / 4;
  const int64_t xs = nchunk > 0 ? (tu.variant > 0 ? (int64_t)tu.variant : nchunk / 2) % nchunk : 0;
#define SPX_CG(UU, NTT) \
  hipLaunchKernelGGL((k_cg_mixed1<S, UU, NTT>), g, dim3(256), 0, current_stream(), n, alpha, p, Ap, x, r, tmp, xs)
  if (U == 1) { if (nt) SPX_CG(1, true); else SPX_CG(1, false); }
  else if (U == 2) { if (nt) SPX_CG(2, true); else SPX_CG(2, false); }
  else if (U == 4) { if (nt) SPX_CG(4, true); else SPX_CG(4, false); }
  else if (U == 8) { if (nt) SPX_CG(8, true); else SPX_CG(8, false); }
  else return fail(SPX_ERR_ARG, "spx_bs5_full_mixed: unroll must be 1, 2, 4 or 8");
#undef SPX_CG
  SPX_LAUNCHED("spx_bs5_full_mixed");
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, current_stream(), (int)grid, tmp, rdotr);
  return SPX_OK;
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

int spx_ogs_gather_scatter_srb_mixed(int storage, int K, const spx_srb_t *srb, void *q) {
  SPX_NEED_DEVICE();
  if (!srb) return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_srb_mixed: null layout");
  if (storage == SPX_STORE_F64) return spx_ogs_gather_scatter_srb(SPX_DOUBLE, SPX_ADD, K, srb, q);
  if (storage != SPX_STORE_F32 && storage != SPX_STORE_BF16)
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_srb_mixed: unknown storage type");
  if (K < 1 || K > 65535) return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_srb_mixed: bad K");
  if (srb->nblocks < 0 || srb->chunks_per_tile < 1 || srb->chunks_per_tile > 64)
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_srb_mixed: bad tiles");
  if (srb->nblocks == 0) return SPX_OK;
  if (!srb->chunks || !srb->runIds || !srb->links || !q)
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_srb_mixed: null pointer");
  if (srb->tile_lanes != 0 && srb->tile_lanes != 64 && srb->tile_lanes != 256)
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_srb_mixed: tile_lanes must be 0, 64 or 256");
  const spx_tuning_t tu = tuning_for("ogs_gs_srb", spx_tuning_t{0, 1, 0, 0, 0, 0});
  Launch L;
  L.nblocks = srb->nblocks;
  L.K = K;
  L.xcdLog2 = kXcdLog2Mixed;
  if (!xcd_hint(srb->xcd_remap, tu.xcd_remap, L.nxcd, L.xcdLog2))
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_srb_mixed: xcd_remap must be negative (library default), 0, 1 or a power of two up to 1024");
  L.cap = 64 * srb->chunks_per_tile;
  L.kf = (K >= 2 && K <= 4) ? K : 1;
  L.per = 0;
  L.variant = srb->tile_lanes ? (srb->tile_lanes == 64 ? 2 : 0) : tu.variant;
  if (srb->nontemporal & ~0x3e)
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_srb_mixed: nontemporal takes bits 1..5 (nodes of 2..32 bytes)");
  L.ntMask = srb->nontemporal;
  L.stream = current_stream();
  L.err = kernel_error_flag();
  SrbArgs A;
  A.chunks = reinterpret_cast<const SrbChunk *>(srb->chunks);
  A.runIds = srb->runIds;
  A.links = srb->links;
  const int rc = storage == SPX_STORE_F32 ? launch_gs_srb_mixed(L, A, (float *)q) : launch_gs_srb_mixed(L, A, (bf16_t *)q);
  if (rc) return rc;
  SPX_LAUNCHED("spx_ogs_gather_scatter_srb_mixed");
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
  int rc = storage == SPX_STORE_F32
               ? launch_cg_mixed(n, (float)alpha, (const float *)p, (const float *)Ap, (float *)x, (float *)r, tmp,
                                 rdotr)
               : launch_cg_mixed(n, (float)alpha, (const bf16_t *)p, (const bf16_t *)Ap, (bf16_t *)x, (bf16_t *)r,
                                 tmp, rdotr);
  if (rc) return rc;
  SPX_LAUNCHED("spx_bs5_full_mixed");
  return SPX_OK;
}

}  // extern "C"
