// Batched CG update over many independent vectors ("segments") of uneven
// length in one launch: x += alpha p, r -= alpha Ap on every segment, and one
// FP64 rdotr per segment.  A host-built plan cuts each segment into tiles at
// the multiples of 4096 entries; the workgroups of a fixed grid stride
// over the tile list, so at any moment the grid works on one compact front
// of each vector (contiguous per-workgroup ranges spread the concurrent
// accesses over the whole batch, which on a 35 GB batch cost 15-20 % of
// bandwidth in address translation).  Each tile leaves one FP64 partial;
// stage 2 folds each segment's partials in tile order, so every rdotr is
// deterministic and independent of the grid size and unroll.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "spx_internal.hpp"
#include "spx_numeric.hpp"

#pragma clang fp contract(off)

using namespace spx;

struct spx_cg_plan_s {
  int32_t nseg = 0;
  int32_t ntiles = 0;
  int32_t nblocks = 0;  // workgroups launched (grid-stride over the tiles)
  int64_t total = 0;
  int64_t *tileRange = nullptr;  // device, 2*ntiles: [first, end) entries
  int32_t *segTiles = nullptr;   // device, nseg+1
  double *partials = nullptr;    // device, ntiles
};

namespace {
constexpr int kTileChunks = 1024;                // chunks of 4 entries per tile: 4 per lane
constexpr int64_t kTileEntries = 4 * kTileChunks;
}  // namespace

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// four consecutive entries of a storage type, as the compute type C
template <typename S> struct Quad4;
template <> struct Quad4<float> {
  typedef f32x4 V;
  typedef float C;
  static __device__ __forceinline__ void unpack(const V &v, C (&o)[4]) {
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  }
  // rounds to storage; o receives the stored values
  static __device__ __forceinline__ V pack(C (&o)[4]) {
    V v;
    v.x = o[0]; v.y = o[1]; v.z = o[2]; v.w = o[3];
    return v;
  }
};
template <> struct Quad4<double> {
  typedef f64x4 V;
  typedef double C;
  static __device__ __forceinline__ void unpack(const V &v, C (&o)[4]) {
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  }
  static __device__ __forceinline__ V pack(C (&o)[4]) {
    V v;
    v.x = o[0]; v.y = o[1]; v.z = o[2]; v.w = o[3];
    return v;
  }
};
template <> struct Quad4<bf16_t> {
  typedef u32x2 V;
  typedef float C;
  static __device__ __forceinline__ void unpack(const V &v, C (&o)[4]) {
    o[0] = __uint_as_float(v.x << 16);
    o[1] = __uint_as_float(v.x & 0xffff0000u);
    o[2] = __uint_as_float(v.y << 16);
    o[3] = __uint_as_float(v.y & 0xffff0000u);
  }
  static __device__ __forceinline__ V pack(C (&o)[4]) {
    const uint32_t b0 = f32_to_bf16_bits(o[0]), b1 = f32_to_bf16_bits(o[1]);
    const uint32_t b2 = f32_to_bf16_bits(o[2]), b3 = f32_to_bf16_bits(o[3]);
    o[0] = __uint_as_float(b0 << 16); o[1] = __uint_as_float(b1 << 16);
    o[2] = __uint_as_float(b2 << 16); o[3] = __uint_as_float(b3 << 16);
    V v;
    v.x = b0 | (b1 << 16);
    v.y = b2 | (b3 << 16);
    return v;
  }
};

template <typename C> __device__ __forceinline__ void update4(C a, const C (&p)[4], const C (&Ap)[4], C (&x)[4], C (&r)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const C ap = a * p[e];
    x[e] = x[e] + ap;
    const C aAp = a * Ap[e];
    r[e] = r[e] - aAp;
  }
}

template <bool NT, typename V> __device__ __forceinline__ V ld(const V *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename V> __device__ __forceinline__ void st(V *p, V v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// one tile's update; lane t takes chunks c0+t, c0+t+256, ... in order
template <typename S, int U, bool NT>
__device__ __forceinline__ double cg_tile(int64_t first, int64_t end, typename Quad4<S>::C alpha,
                                          const typename Quad4<S>::V *__restrict__ p,
                                          const typename Quad4<S>::V *__restrict__ Ap,
                                          typename Quad4<S>::V *__restrict__ x, typename Quad4<S>::V *__restrict__ r,
                                          const S *__restrict__ p_, const S *__restrict__ Ap_, S *__restrict__ x_,
                                          S *__restrict__ r_) {
  typedef Quad4<S> Q;
  typedef typename Q::V V;
  typedef typename Q::C C;
  const int t = threadIdx.x;
  const int64_t c0 = first / 4, cend = end / 4;  // first is a multiple of 4
  double acc = 0.0;
  if (cend - c0 == kTileChunks) {
#pragma unroll
    for (int k = 0; k < kTileChunks / 256; k += U) {
      V vp[U], vap[U], vx[U], vr[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t c = c0 + t + (k + u) * 256;
        vp[u] = ld<NT>(p + c);
        vap[u] = ld<NT>(Ap + c);
        vx[u] = ld<NT>(x + c);
        vr[u] = ld<NT>(r + c);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t c = c0 + t + (k + u) * 256;
        C fp[4], fap[4], fx[4], fr[4];
        Q::unpack(vp[u], fp);
        Q::unpack(vap[u], fap);
        Q::unpack(vx[u], fx);
        Q::unpack(vr[u], fr);
        update4(alpha, fp, fap, fx, fr);
        st<NT>(x + c, Q::pack(fx));
        st<NT>(r + c, Q::pack(fr));
#pragma unroll
        for (int e = 0; e < 4; ++e) acc += (double)fr[e] * (double)fr[e];
      }
    }
  } else {
    for (int64_t c = c0 + t; c < cend; c += 256) {
      C fp[4], fap[4], fx[4], fr[4];
      Q::unpack(p[c], fp);
      Q::unpack(Ap[c], fap);
      Q::unpack(x[c], fx);
      Q::unpack(r[c], fr);
      update4(alpha, fp, fap, fx, fr);
      x[c] = Q::pack(fx);
      r[c] = Q::pack(fr);
#pragma unroll
      for (int e = 0; e < 4; ++e) acc += (double)fr[e] * (double)fr[e];
    }
  }
  // entries past the last whole chunk (end % 4 of them): lane 0, in order
  if (t == 0) {
    for (int64_t i = std::max(cend * 4, first); i < end; ++i) {
      C pi, Api, xi, ri;
      if constexpr (sizeof(S) == 8) {
        pi = Store<S>::to_f64(p_[i]); Api = Store<S>::to_f64(Ap_[i]);
        xi = Store<S>::to_f64(x_[i]); ri = Store<S>::to_f64(r_[i]);
      } else {
        pi = Store<S>::to_f32(p_[i]); Api = Store<S>::to_f32(Ap_[i]);
        xi = Store<S>::to_f32(x_[i]); ri = Store<S>::to_f32(r_[i]);
      }
      const C ap = alpha * pi;
      xi = xi + ap;
      const C aAp = alpha * Api;
      ri = ri - aAp;
      if constexpr (sizeof(S) == 8) {
        x_[i] = Store<S>::from_f64(xi);
        r_[i] = Store<S>::from_f64(ri);
      } else {
        x_[i] = Store<S>::from_f32(xi);
        r_[i] = Store<S>::from_f32(ri);
      }
      const double rs = Store<S>::to_f64(r_[i]);
      acc += rs * rs;
    }
  }
  return acc;
}

// the workgroups of a fixed grid stride over the tiles; one partial per tile
template <typename S, int U, bool NT>
__global__ __launch_bounds__(256) void k_cg_batch1(const int64_t *__restrict__ tileRange, int32_t ntiles,
                                                   double alpha_d, const S *__restrict__ p_,
                                                   const S *__restrict__ Ap_, S *__restrict__ x_,
                                                   S *__restrict__ r_, double *__restrict__ partials) {
  typedef Quad4<S> Q;
  typedef typename Q::V V;
  typedef typename Q::C C;
  const C alpha = (C)alpha_d;
  const V *__restrict__ p = reinterpret_cast<const V *>(p_);
  const V *__restrict__ Ap = reinterpret_cast<const V *>(Ap_);
  V *__restrict__ x = reinterpret_cast<V *>(x_);
  V *__restrict__ r = reinterpret_cast<V *>(r_);
  __shared__ double s[4];
  const int t = threadIdx.x;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t first = tileRange[2 * tile], end = tileRange[2 * tile + 1];
    double v = cg_tile<S, U, NT>(first, end, alpha, p, Ap, x, r, p_, Ap_, x_, r_);
    // fixed tree: butterfly inside each wave, then the four wave sums in order
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((t & 63) == 0) s[t >> 6] = v;
    __syncthreads();
    if (t == 0) partials[tile] = (s[0] + s[1]) + (s[2] + s[3]);
    __syncthreads();  // s is read before the next tile overwrites it
  }
}

// one workgroup per segment: its tile partials in tile order
__global__ __launch_bounds__(256) void k_cg_batch2(const int32_t *__restrict__ segTiles,
                                                   const double *__restrict__ partials, double *__restrict__ rdotr) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  const int b0 = segTiles[blockIdx.x], b1 = segTiles[blockIdx.x + 1];
  double a = 0.0;
  for (int i = b0 + t; i < b1; i += 256) a += partials[i];
  s[t] = a;
  __syncthreads();
#pragma unroll
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w) s[t] += s[t + w];
    __syncthreads();
  }
  if (t == 0) rdotr[blockIdx.x] = s[0];
}

template <typename S>
int launch_cg_batch(const spx_cg_plan_s &P, double alpha, const S *p, const S *Ap, S *x, S *r, double *rdotr) {
  spx_tuning_t d{};
  d.unroll = sizeof(S) == 2 ? 2 : (sizeof(S) == 4 ? 4 : 2);
  d.nontemporal = 1;
  d.xcd_remap = 0;
  d.variant = 0;
  const spx_tuning_t tu = tuning_for("cg_batch", d);
  const bool nt = tu.nontemporal > 0;
  const dim3 g((unsigned)P.nblocks);
  hipStream_t st = current_stream();
#define SPX_CGB(UU, NTT) \
  hipLaunchKernelGGL((k_cg_batch1<S, UU, NTT>), g, dim3(256), 0, st, P.tileRange, P.ntiles, alpha, p, Ap, x, r, \
                     P.partials)
  switch (tu.unroll) {
    case 1: if (nt) SPX_CGB(1, true); else SPX_CGB(1, false); break;
    case 2: if (nt) SPX_CGB(2, true); else SPX_CGB(2, false); break;
    case 4: if (nt) SPX_CGB(4, true); else SPX_CGB(4, false); break;
    default: return fail(SPX_ERR_ARG, "spx_cg_batch: unroll must be 1, 2 or 4");
  }
#undef SPX_CGB
  SPX_LAUNCHED("spx_cg_batch");
  hipLaunchKernelGGL(k_cg_batch2, dim3((unsigned)P.nseg), dim3(256), 0, st, P.segTiles, P.partials, rdotr);
  SPX_LAUNCHED("spx_cg_batch");
  return SPX_OK;
}

}  // namespace

extern "C" {

int spx_cg_plan_create(int32_t nseg, const int64_t *segStarts, int32_t target_blocks, spx_cg_plan_t *out) {
  SPX_NEED_DEVICE();
  if (!out) return fail(SPX_ERR_ARG, "spx_cg_plan_create: null out");
  *out = nullptr;
  if (nseg < 1 || !segStarts) return fail(SPX_ERR_ARG, "spx_cg_plan_create: need nseg >= 1 and segStarts");
  if (target_blocks < 1) return fail(SPX_ERR_ARG, "spx_cg_plan_create: target_blocks >= 1");
  for (int32_t s = 0; s < nseg; ++s) {
    if (segStarts[s + 1] < segStarts[s]) return fail(SPX_ERR_ARG, "spx_cg_plan_create: segStarts must not decrease");
    if (segStarts[s] % 4) return fail(SPX_ERR_ARG, "spx_cg_plan_create: every segment must start at a multiple of 4");
  }
  if (segStarts[0] < 0) return fail(SPX_ERR_ARG, "spx_cg_plan_create: negative start");
  const int64_t total = segStarts[nseg] - segStarts[0];
  std::vector<int64_t> range;
  std::vector<int32_t> segTiles(nseg + 1, 0);
  for (int32_t s = 0; s < nseg; ++s) {
    const int64_t a = segStarts[s], e = segStarts[s + 1];
    // tiles end on multiples of kTileEntries from the vectors' start, so full tiles cover whole
    // cache lines and neighbouring tiles share none; every segment has at least one tile (an
    // empty one for an empty segment)
    int64_t nt = 0;
    for (int64_t k0 = a; k0 < e || nt == 0; ++nt) {
      const int64_t k1 = std::min((k0 / kTileEntries + 1) * kTileEntries, e);
      range.push_back(k0);
      range.push_back(k1);
      k0 = std::max(k1, k0 + 1);
    }
    if ((int64_t)segTiles[s] + nt > INT32_MAX) return fail(SPX_ERR_ARG, "spx_cg_plan_create: too many tiles");
    segTiles[s + 1] = segTiles[s] + (int32_t)nt;
  }
  auto *P = new spx_cg_plan_s;
  P->nseg = nseg;
  P->ntiles = segTiles[nseg];
  P->nblocks = std::min(target_blocks, P->ntiles);
  P->total = total;
  if (hipMalloc((void **)&P->tileRange, range.size() * sizeof(int64_t)) != hipSuccess ||
      hipMalloc((void **)&P->segTiles, segTiles.size() * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void **)&P->partials, (size_t)P->ntiles * sizeof(double)) != hipSuccess) {
    (void)hipFree(P->tileRange);
    (void)hipFree(P->segTiles);
    (void)hipFree(P->partials);
    delete P;
    return fail(SPX_ERR_HIP, "spx_cg_plan_create: device allocation failed");
  }
  if (hipMemcpy(P->tileRange, range.data(), range.size() * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(P->segTiles, segTiles.data(), segTiles.size() * sizeof(int32_t), hipMemcpyHostToDevice) !=
          hipSuccess) {
    (void)hipFree(P->tileRange);
    (void)hipFree(P->segTiles);
    (void)hipFree(P->partials);
    delete P;
    return fail(SPX_ERR_HIP, "spx_cg_plan_create: copy failed");
  }
  *out = P;
  return SPX_OK;
}

int spx_cg_plan_destroy(spx_cg_plan_t plan) {
  if (!plan) return SPX_OK;
  (void)hipFree(plan->tileRange);
  (void)hipFree(plan->segTiles);
  (void)hipFree(plan->partials);
  delete plan;
  return SPX_OK;
}

int spx_cg_plan_blocks(spx_cg_plan_t plan) { return plan ? plan->nblocks : -1; }

int spx_cg_batch(int storage, spx_cg_plan_t plan, double alpha, const void *p, const void *Ap, void *x, void *r,
                 double *rdotr) {
  SPX_NEED_DEVICE();
  if (!plan) return fail(SPX_ERR_ARG, "spx_cg_batch: null plan");
  if (!rdotr) return fail(SPX_ERR_ARG, "spx_cg_batch: null rdotr");
  if (plan->total > 0 && (!p || !Ap || !x || !r)) return fail(SPX_ERR_ARG, "spx_cg_batch: null vector");
  const size_t al = storage == SPX_STORE_F64 ? 32 : storage == SPX_STORE_F32 ? 16 : 8;
  if (!aligned(p, al) || !aligned(Ap, al) || !aligned(x, al) || !aligned(r, al))
    return fail(SPX_ERR_ARG, "spx_cg_batch: vectors must be aligned to 4 entries");
  switch (storage) {
    case SPX_STORE_F64:
      return launch_cg_batch(*plan, alpha, (const double *)p, (const double *)Ap, (double *)x, (double *)r, rdotr);
    case SPX_STORE_F32:
      return launch_cg_batch(*plan, alpha, (const float *)p, (const float *)Ap, (float *)x, (float *)r, rdotr);
    case SPX_STORE_BF16:
      return launch_cg_batch(*plan, alpha, (const bf16_t *)p, (const bf16_t *)Ap, (bf16_t *)x, (bf16_t *)r, rdotr);
    default:
      return fail(SPX_ERR_ARG, "spx_cg_batch: unknown storage type");
  }
}

}  // extern "C"
