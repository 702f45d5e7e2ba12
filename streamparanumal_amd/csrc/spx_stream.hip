// BS0-BS5 streaming kernels for gfx950.
//
// BS1/BS2 stream 16-byte (dwordx4) loads and stores, several per thread, so a
// wave moves 1 KiB per instruction and keeps many requests in flight.
// BS3-BS5 reduce in one of two orders:
//   REFERENCE - the exact loop nest of bs3.okl/bs4.okl/bs5.okl: Nblock =
//               min(ceil(N/bs), bs) blocks, thread t of block b sums
//               a[t + b*bs + j*bs*Nblock] for j = 0,1,.. in order, a halving
//               LDS tree down to s[0]+s[1], and a second one-block pass.  The
//               serial chain per thread is kept; only its loads are issued
//               ahead (U independent loads, then U dependent adds).
//   FAST      - the same launch shape and element order with compensated
//               (Neumaier) lane sums and a compensated second pass: as fast as
//               the reference order, deterministic, and within 1e-14 of it.
#include <hip/hip_runtime.h>

#include <string>

#include "spx_internal.hpp"

#pragma clang fp contract(off)

using namespace spx;

typedef double dbl2 __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ dbl2 ld2(const dbl2 *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ void st2(dbl2 *p, dbl2 v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// ---------------------------------------------------------------------- BS0 --
// bs0.okl:27-39: one work-item; the store never happens for N != 0.
__global__ void k_bs0(int n, const double *a, double *b) {
  (void)a;
  if ((int)threadIdx.x == n && (int)blockIdx.x == n) b[0] = 0;
}

// ---------------------------------------------------------------------- BS1 --
template <int U, bool NT>
__global__ __launch_bounds__(256) void k_copy2(int64_t n2, const dbl2 *__restrict__ a,
                                               dbl2 *__restrict__ b) {
  const int64_t base = (int64_t)blockIdx.x * (U * 256) + threadIdx.x;
  dbl2 r[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i = base + u * 256;
    if (i < n2) r[u] = ld2<NT>(a + i);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i = base + u * 256;
    if (i < n2) st2<NT>(b + i, r[u]);
  }
}

// scalar path for unaligned pointers and the odd tail element
__global__ __launch_bounds__(256) void k_copy1(int64_t lo, int64_t n, const double *__restrict__ a,
                                               double *__restrict__ b) {
  int64_t i = lo + (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * 256;
  for (; i < n; i += st) b[i] = a[i];
}

// ---------------------------------------------------------------------- BS2 --
template <int U, bool NT>
__global__ __launch_bounds__(256) void k_axpy2(int64_t n2, double alpha, const dbl2 *__restrict__ x,
                                               double beta, dbl2 *__restrict__ y) {
  const int64_t base = (int64_t)blockIdx.x * (U * 256) + threadIdx.x;
  dbl2 rx[U], ry[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i = base + u * 256;
    if (i < n2) {
      rx[u] = ld2<NT>(x + i);
      ry[u] = ld2<NT>(y + i);
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i = base + u * 256;
    if (i < n2) {
      dbl2 o;
      o.x = alpha * rx[u].x + beta * ry[u].x;
      o.y = alpha * rx[u].y + beta * ry[u].y;
      st2<NT>(y + i, o);
    }
  }
}

__global__ __launch_bounds__(256) void k_axpy1(int64_t lo, int64_t n, double alpha,
                                               const double *__restrict__ x, double beta,
                                               double *__restrict__ y) {
  int64_t i = lo + (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * 256;
  for (; i < n; i += st) y[i] = alpha * x[i] + beta * y[i];
}

// --------------------------------------------------------------- reductions --
enum RedKind { RED_NORM = 0, RED_DOT = 1, RED_CG = 2, RED_CGFULL = 3 };

struct RedArgs {
  const double *a;   // NORM: a; DOT: a; CG: Ap; CGFULL: p
  const double *b;   // DOT: b; CGFULL: Ap
  double *r;         // CG, CGFULL: r (updated)
  double *x;         // CGFULL: x (updated)
  double alpha;
  // CGFULL: the x update of index i runs beside the r update of index
  // i + n - xshift (mod n): x and p then stream at another address phase
  // than r and Ap.  Each x entry is still updated once, from its own p, so
  // results do not change.
  int64_t xshift = 0;
  int64_t n = 0;
};

// CGFULL's x index for the r index i (see RedArgs::xshift)
__device__ __forceinline__ int64_t x_index(const RedArgs &A, int64_t i) {
  const int64_t j = i + A.xshift;
  return j >= A.n ? j - A.n : j;
}

// One term of the reduction at element i, applying the update when the
// kernel has one.  Loads of term i are independent of the running sum.
template <int KIND>
__device__ __forceinline__ double red_term(const RedArgs &A, int64_t i) {
  if constexpr (KIND == RED_NORM) {
    const double v = A.a[i];
    return v * v;
  } else if constexpr (KIND == RED_DOT) {
    return A.a[i] * A.b[i];
  } else if constexpr (KIND == RED_CG) {
    const double rr = A.r[i] + A.alpha * A.a[i];
    A.r[i] = rr;
    return rr * rr;
  } else {
    const int64_t ix = x_index(A, i);
    A.x[ix] = A.x[ix] + A.alpha * A.a[ix];
    const double rr = A.r[i] - A.alpha * A.b[i];
    A.r[i] = rr;
    return rr * rr;
  }
}

// halving tree of bs3.okl:48-63 on BS doubles in LDS, result s[0]+s[1]
template <int BS>
__device__ __forceinline__ double ref_tree(double *s) {
  const int t = threadIdx.x;
#pragma unroll
  for (int w = BS / 2; w >= 2; w >>= 1) {
    if (t < w) s[t] += s[t + w];
    __syncthreads();
  }
  return s[0] + s[1];
}

// REFERENCE stage 1: exactly the bsN_1 loop nest (bs3.okl:27-64).  Each
// unrolled group first issues all of its loads, then applies the updates and
// adds the U terms to the running sum in index order.
// NT: streaming (non-temporal) loads and stores; each element is touched once.
template <bool NT> __device__ __forceinline__ double ldv(const double *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT> __device__ __forceinline__ void stv(double *p, double v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// FAST mode's compensated (Neumaier) running sum: sum + c carries the
// rounding error of every add; the lane's result is sum + c.
__device__ __forceinline__ void neumaier(double &sum, double &c, double v) {
  const double t = sum + v;
  if (fabs(sum) >= fabs(v)) c += (sum - t) + v;
  else c += (v - t) + sum;
  sum = t;
}

// The running sum of one lane: plain adds in index order (the reference's
// serial chain) or compensated.
template <bool COMP> struct LaneSum;
template <> struct LaneSum<false> {
  double s = 0.0;
  __device__ __forceinline__ void add(double v) { s += v; }
  __device__ __forceinline__ double value() const { return s; }
};
template <> struct LaneSum<true> {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double v) { neumaier(s, c, v); }
  __device__ __forceinline__ double value() const { return s + c; }
};

template <int KIND, int U, bool NT, typename ACC>
__device__ __forceinline__ void ref_group(const RedArgs &A, int64_t id, int64_t stride, ACC &acc) {
  double va[U], vb[U], vr[U], vx[U];
  int64_t ix[KIND == RED_CGFULL ? U : 1];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i = id + u * stride;
    if constexpr (KIND == RED_CGFULL) {
      ix[u] = x_index(A, i);
      va[u] = ldv<NT>(A.a + ix[u]);
    } else {
      va[u] = ldv<NT>(A.a + i);
    }
    if constexpr (KIND == RED_DOT || KIND == RED_CGFULL) vb[u] = ldv<NT>(A.b + i);
    if constexpr (KIND == RED_CG || KIND == RED_CGFULL) vr[u] = ldv<NT>(A.r + i);
    if constexpr (KIND == RED_CGFULL) vx[u] = ldv<NT>(A.x + ix[u]);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t i = id + u * stride;
    double term;
    if constexpr (KIND == RED_NORM) {
      term = va[u] * va[u];
    } else if constexpr (KIND == RED_DOT) {
      term = va[u] * vb[u];
    } else if constexpr (KIND == RED_CG) {
      const double rr = vr[u] + A.alpha * va[u];
      stv<NT>(A.r + i, rr);
      term = rr * rr;
    } else {
      stv<NT>(A.x + ix[u], vx[u] + A.alpha * va[u]);
      const double rr = vr[u] - A.alpha * vb[u];
      stv<NT>(A.r + i, rr);
      term = rr * rr;
    }
    acc.add(term);
  }
}

// COMP = false: the reference order exactly.  COMP = true (FAST mode): the
// same launch shape and element order with compensated lane sums, so it runs
// at the reference order's bandwidth and is more accurate.
template <int KIND, int BS, int U, bool NT, bool COMP = false>
__global__ __launch_bounds__(BS) void k_red_ref1(int64_t n, int64_t nblock, RedArgs A,
                                                 double *__restrict__ partials) {
  __shared__ double s[BS];
  const int t = threadIdx.x;
  const int64_t stride = (int64_t)BS * nblock;
  int64_t id = t + (int64_t)blockIdx.x * BS;
  LaneSum<COMP> acc;
  for (; id + (int64_t)(U - 1) * stride < n; id += (int64_t)U * stride)
    ref_group<KIND, U, NT>(A, id, stride, acc);
  for (; id < n; id += stride) acc.add(red_term<KIND>(A, id));
  s[t] = acc.value();
  __syncthreads();
  const double blk = ref_tree<BS>(s);
  if (t == 0) partials[blockIdx.x] = blk;
}

// REFERENCE stage 2: bsN_2 (bs3.okl:66-102) with one block.
template <int BS>
__global__ __launch_bounds__(BS) void k_red_ref2(int64_t nblock, const double *__restrict__ partials,
                                                 double *__restrict__ out) {
  __shared__ double s[BS];
  const int t = threadIdx.x;
  double r = 0.0;
  for (int64_t id = t; id < nblock; id += (int64_t)BS * nblock) r += partials[id];
  s[t] = r;
  __syncthreads();
  const double v = ref_tree<BS>(s);
  if (t == 0) out[0] = v;
}


// Block sum of (s + c) over 256 threads with a fixed shape: wave-level
// shuffles then LDS across the 4 waves.
__device__ __forceinline__ double block_sum256(double v) {
  __shared__ double ws[4];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) ws[w] = v;
  __syncthreads();
  double out = 0.0;
  if (threadIdx.x == 0) out = (ws[0] + ws[1]) + (ws[2] + ws[3]);
  return out;
}

// FAST stage 2: compensated sum of the partials, then the fixed block tree.
__global__ __launch_bounds__(256) void k_red_fast2(int64_t g, const double *__restrict__ partials,
                                                   double *__restrict__ out) {
  double s = 0.0, c = 0.0;
  for (int64_t i = threadIdx.x; i < g; i += 256) neumaier(s, c, partials[i]);
  const double b = block_sum256(s + c);
  if (threadIdx.x == 0) out[0] = b;
}

// ------------------------------------------------------------------- host ----

static int vec_blocks(int64_t n2, int per_block, unsigned *out) {
  const int64_t b = (n2 + per_block - 1) / per_block;
  if (b > INT32_MAX) return fail(SPX_ERR_ARG, "vector too long for one launch");
  *out = (unsigned)(b > 0 ? b : 1);
  return SPX_OK;
}

static spx_tuning_t stream_defaults() {
  spx_tuning_t d{};
  d.block_threads = 256;
  d.xcd_remap = 0;
  d.unroll = 4;
  d.grid_blocks = 0;
  d.nontemporal = 1;  // measured on MI355X: 6.17 vs 5.65 TB/s for the 2^27 copy
  return d;
}

template <int U, bool NT>
static void launch_copy2(unsigned grid, int64_t n2, const double *a, double *b) {
  hipLaunchKernelGGL((k_copy2<U, NT>), dim3(grid), dim3(256), 0, current_stream(), n2,
                     reinterpret_cast<const dbl2 *>(a), reinterpret_cast<dbl2 *>(b));
}
template <int U, bool NT>
static void launch_axpy2(unsigned grid, int64_t n2, double alpha, const double *x, double beta,
                         double *y) {
  hipLaunchKernelGGL((k_axpy2<U, NT>), dim3(grid), dim3(256), 0, current_stream(), n2, alpha,
                     reinterpret_cast<const dbl2 *>(x), beta, reinterpret_cast<dbl2 *>(y));
}

// The FAST grid never exceeds kFastGridMax blocks, and spx_reduce_tmp_count
// reports kFastGridMax for FAST mode, so a partials buffer sized by it stays
// large enough whatever tuning is set later.
constexpr int kFastGridMax = 4096;
// Measured on MI355X at N = 2^27 (TB/s, norm / dot / update / full form):
// the reference loop shape with compensated lane sums runs like the
// reference order itself; an earlier FAST kernel (16-byte loads, 8 per lane,
// on 4096 blocks) gave 5.6 / 5.5 / 4.3 / 4.9 and on 256 blocks 7.2 / 7.1 /
// 4.4 / 5.6, against the reference order's 7.0 / 7.0 / 6.0 / 5.8.
static int fast_grid() {
  spx_tuning_t t = tuning_for("reduce_fast", stream_defaults());
  int g = t.grid_blocks > 0 ? t.grid_blocks : 256;
  return g < 1 ? 1 : (g > kFastGridMax ? kFastGridMax : g);
}

template <int KIND, int BS, bool NT, bool COMP>
static void launch_ref1_nt(int64_t n, int64_t nb, const RedArgs &A, double *tmp, int unroll) {
  if (unroll >= 32)
    hipLaunchKernelGGL((k_red_ref1<KIND, BS, 32, NT, COMP>), dim3((unsigned)nb), dim3(BS), 0, current_stream(), n,
                       nb, A, tmp);
  else if (unroll >= 16)
    hipLaunchKernelGGL((k_red_ref1<KIND, BS, 16, NT, COMP>), dim3((unsigned)nb), dim3(BS), 0, current_stream(), n,
                       nb, A, tmp);
  else
    hipLaunchKernelGGL((k_red_ref1<KIND, BS, 8, NT, COMP>), dim3((unsigned)nb), dim3(BS), 0, current_stream(), n,
                       nb, A, tmp);
}
template <int KIND, int BS, bool COMP = false>
static void launch_ref1(int64_t n, int64_t nb, const RedArgs &A, double *tmp, int unroll, bool nt) {
  if (nt) launch_ref1_nt<KIND, BS, true, COMP>(n, nb, A, tmp, unroll);
  else launch_ref1_nt<KIND, BS, false, COMP>(n, nb, A, tmp, unroll);
}

// Measured on MI355X at N = 2^27 (TB/s, unroll 16): streaming loads/stores
// give norm 6.81 / dot 6.93 / update 5.97 / full form 5.78, against 6.20 /
// 6.36 / 4.89 / 5.17 with plain ones.
// Unroll 16 for norm, dot and the full form; 8 for the update (6.07-6.10
// against 5.99 TB/s at unroll 16; norm 6.97, dot 6.95-7.00, full form 5.84
// at 16, each within 1 % or better than 8 and 32 there).
static spx_tuning_t reference_tuning(int kind) {
  spx_tuning_t d{};
  d.unroll = kind == RED_CG ? 8 : 16;
  d.nontemporal = 1;
  d.xcd_remap = 0;
  d.variant = 0;
  return tuning_for("reduce_reference", d);
}

template <int KIND>
static int run_reduction(const char *name, int64_t n, const RedArgs &A, double *tmp, double *out,
                         int mode, int bs) {
  SPX_NEED_DEVICE();
  if (n < 0) return fail(SPX_ERR_ARG, std::string(name) + ": negative n");
  if (!tmp || !out) return fail(SPX_ERR_ARG, std::string(name) + ": null tmp/out");
  if (mode == SPX_REDUCE_REFERENCE) {
    if (bs != 256 && bs != 512 && bs != 1024)
      return fail(SPX_ERR_ARG, std::string(name) + ": block_size must be 256, 512 or 1024");
    int64_t nb = (n + bs - 1) / bs;
    if (nb > bs) nb = bs;
    const spx_tuning_t t = reference_tuning(KIND);
    const int U = t.unroll > 0 ? t.unroll : 16;
    const bool nt = t.nontemporal > 0;
    if (nb > 0) {
      if (bs == 256) launch_ref1<KIND, 256>(n, nb, A, tmp, U, nt);
      else if (bs == 512) launch_ref1<KIND, 512>(n, nb, A, tmp, U, nt);
      else launch_ref1<KIND, 1024>(n, nb, A, tmp, U, nt);
      SPX_LAUNCHED(name);
    }
    // n == 0: the reference still runs stage 2 over zero partials -> 0.0
    if (bs == 256) hipLaunchKernelGGL(k_red_ref2<256>, dim3(1), dim3(256), 0, current_stream(), nb, tmp, out);
    else if (bs == 512) hipLaunchKernelGGL(k_red_ref2<512>, dim3(1), dim3(512), 0, current_stream(), nb, tmp, out);
    else hipLaunchKernelGGL(k_red_ref2<1024>, dim3(1), dim3(1024), 0, current_stream(), nb, tmp, out);
    SPX_LAUNCHED(name);
    return SPX_OK;
  }
  if (mode != SPX_REDUCE_FAST) return fail(SPX_ERR_ARG, std::string(name) + ": unknown reduction mode");
  // FAST: min(ceil(n/256), grid) blocks of 256 lanes in the reference's
  // element order, compensated lane sums, then a compensated fold of the
  // partials; deterministic for a given grid.
  int64_t nb = (n + 255) / 256;
  if (nb > fast_grid()) nb = fast_grid();
  spx_tuning_t fd = reference_tuning(KIND);  // loads per lane and streaming accesses as the reference order's
  fd.grid_blocks = 0;
  const spx_tuning_t ft = tuning_for("reduce_fast", fd);
  const int U = ft.unroll > 0 ? ft.unroll : 16;
  if (nb > 0) {
    launch_ref1<KIND, 256, true>(n, nb, A, tmp, U, ft.nontemporal > 0);
    SPX_LAUNCHED(name);
  }
  hipLaunchKernelGGL(k_red_fast2, dim3(1), dim3(256), 0, current_stream(), nb, tmp, out);
  SPX_LAUNCHED(name);
  return SPX_OK;
}

// ------------------------------------------------ allocation-free scratch ----
// Partials of the allocation-free CG update (spx_bs5_full_noalloc): one
// file-scope array per device, sized for the largest partial count any
// reduction shape here produces (kFastGridMax; REFERENCE mode needs at most
// its block size, <= 1024).  The host never passes `scratch` itself to a
// kernel (in host code the name is only a stand-in for the device symbol):
// spx_device_open takes its device address once, with the device current,
// and every launch uses that address.  Stage 1 writes every partial that
// stage 2 reads, so nothing has to start at zero; calls that overlap on one
// device (several streams or host threads) share the array and race.
constexpr int kScratchDoubles = kFastGridMax;
static_assert(kScratchDoubles == SPX_NOALLOC_SCRATCH_DOUBLES, "spx.h states the scratch size");
static_assert(kScratchDoubles >= 1024, "scratch must hold REFERENCE-mode partials");
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kScratchDevices = 64;
static double *g_scratch_ptr[kScratchDevices] = {};

namespace spx {
int scratch_setup(int device_id) {
  if (device_id < 0 || device_id >= kScratchDevices)
    return fail(SPX_ERR_UNSUPPORTED, "scratch_setup: device id " + std::to_string(device_id) + " past the table");
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(SPX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device_id] = static_cast<double *>(ptr);
  return SPX_OK;
}
}  // namespace spx

// The full form's x/p address phase shift (RedArgs::xshift), in entries.
// Default: half the vector, rounded down to 64 entries.  Measured on MI355X,
// N = 2^27, the four vectors carved 1 GiB + d apart (TB/s, no shift / half):
// d = 0 5.23 / 5.76, 512 B 5.53 / 5.74, 4 KiB 5.80 / 5.75, 64 KiB 5.33 /
// 5.70, 2 MiB 5.71 / 5.82, separate hipMalloc blocks 5.51 / 5.78 (copy 6.24).
// Tuning "bs5_full": variant v > 0 shifts by v*64 entries (mod n).
static int64_t full_xshift(int64_t n) {
  spx_tuning_t d{};
  d.variant = 0;
  const spx_tuning_t t = tuning_for("bs5_full", d);
  if (n <= 0) return 0;
  const int64_t sh = t.variant > 0 ? (int64_t)t.variant * 64 : (n / 2) / 64 * 64;
  return sh % n;
}

// One stage at a time, with the caller's block count, exactly as the OKL
// kernel pair is launched by bsNRun.cpp.
template <int KIND>
static int run_stage1(const char *name, int64_t nblocks, int64_t n, const RedArgs &A, double *partials,
                      int bs) {
  SPX_NEED_DEVICE();
  if (n < 0 || nblocks < 0) return fail(SPX_ERR_ARG, std::string(name) + ": negative size");
  if (bs != 256 && bs != 512 && bs != 1024)
    return fail(SPX_ERR_ARG, std::string(name) + ": block_size must be 256, 512 or 1024");
  if (nblocks > INT32_MAX) return fail(SPX_ERR_ARG, std::string(name) + ": too many blocks");
  if (nblocks == 0) return SPX_OK;
  if (!partials) return fail(SPX_ERR_ARG, std::string(name) + ": null partials");
  const spx_tuning_t t = reference_tuning(KIND);
  const int U = t.unroll > 0 ? t.unroll : 16;
  const bool nt = t.nontemporal > 0;
  if (bs == 256) launch_ref1<KIND, 256>(n, nblocks, A, partials, U, nt);
  else if (bs == 512) launch_ref1<KIND, 512>(n, nblocks, A, partials, U, nt);
  else launch_ref1<KIND, 1024>(n, nblocks, A, partials, U, nt);
  SPX_LAUNCHED(name);
  return SPX_OK;
}

extern "C" {

int spx_bs0(int32_t n, const double *a, double *b) {
  SPX_NEED_DEVICE();
  hipLaunchKernelGGL(k_bs0, dim3(1), dim3(1), 0, current_stream(), (int)n, a, b);
  SPX_LAUNCHED("spx_bs0");
  return SPX_OK;
}

int spx_bs1_copy(int64_t n, const double *a, double *b) {
  SPX_NEED_DEVICE();
  if (n < 0) return fail(SPX_ERR_ARG, "spx_bs1_copy: negative n");
  if (n == 0) return SPX_OK;
  if (!a || !b) return fail(SPX_ERR_ARG, "spx_bs1_copy: null pointer");
  spx_tuning_t t = tuning_for("bs1", stream_defaults());
  if (aligned(a, 16) && aligned(b, 16)) {
    const int64_t n2 = n / 2;
    const int U = t.unroll;
    const bool nt = t.nontemporal > 0;
    if (n2 > 0) {
      unsigned grid;
      int rc = vec_blocks(n2, (U >= 8 ? 8 : U >= 4 ? 4 : 2) * 256, &grid);
      if (rc) return rc;
      if (U >= 8) nt ? launch_copy2<8, true>(grid, n2, a, b) : launch_copy2<8, false>(grid, n2, a, b);
      else if (U >= 4) nt ? launch_copy2<4, true>(grid, n2, a, b) : launch_copy2<4, false>(grid, n2, a, b);
      else nt ? launch_copy2<2, true>(grid, n2, a, b) : launch_copy2<2, false>(grid, n2, a, b);
      SPX_LAUNCHED("spx_bs1_copy");
    }
    if (n & 1) {
      hipLaunchKernelGGL(k_copy1, dim3(1), dim3(256), 0, current_stream(), n - 1, n, a, b);
      SPX_LAUNCHED("spx_bs1_copy tail");
    }
  } else {
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_copy1, dim3((unsigned)g), dim3(256), 0, current_stream(), (int64_t)0, n, a, b);
    SPX_LAUNCHED("spx_bs1_copy");
  }
  return SPX_OK;
}

int spx_bs2_axpy(int64_t n, double alpha, const double *x, double beta, double *y) {
  SPX_NEED_DEVICE();
  if (n < 0) return fail(SPX_ERR_ARG, "spx_bs2_axpy: negative n");
  if (n == 0) return SPX_OK;
  if (!x || !y) return fail(SPX_ERR_ARG, "spx_bs2_axpy: null pointer");
  spx_tuning_t t = tuning_for("bs2", stream_defaults());
  if (aligned(x, 16) && aligned(y, 16)) {
    const int64_t n2 = n / 2;
    const int U = t.unroll;
    const bool nt = t.nontemporal > 0;
    if (n2 > 0) {
      unsigned grid;
      int rc = vec_blocks(n2, (U >= 8 ? 8 : U >= 4 ? 4 : 2) * 256, &grid);
      if (rc) return rc;
      if (U >= 8) nt ? launch_axpy2<8, true>(grid, n2, alpha, x, beta, y) : launch_axpy2<8, false>(grid, n2, alpha, x, beta, y);
      else if (U >= 4) nt ? launch_axpy2<4, true>(grid, n2, alpha, x, beta, y) : launch_axpy2<4, false>(grid, n2, alpha, x, beta, y);
      else nt ? launch_axpy2<2, true>(grid, n2, alpha, x, beta, y) : launch_axpy2<2, false>(grid, n2, alpha, x, beta, y);
      SPX_LAUNCHED("spx_bs2_axpy");
    }
    if (n & 1) {
      hipLaunchKernelGGL(k_axpy1, dim3(1), dim3(256), 0, current_stream(), n - 1, n, alpha, x, beta, y);
      SPX_LAUNCHED("spx_bs2_axpy tail");
    }
  } else {
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_axpy1, dim3((unsigned)g), dim3(256), 0, current_stream(), (int64_t)0, n,
                       alpha, x, beta, y);
    SPX_LAUNCHED("spx_bs2_axpy");
  }
  return SPX_OK;
}

int spx_bs3_1(int64_t nblocks, int64_t n, const double *a, double *partials, int bs) {
  if (n > 0 && nblocks > 0 && !a) return fail(SPX_ERR_ARG, "spx_bs3_1: null a");
  RedArgs A{a, nullptr, nullptr, nullptr, 0.0};
  return run_stage1<RED_NORM>("spx_bs3_1", nblocks, n, A, partials, bs);
}

int spx_bs4_1(int64_t nblocks, int64_t n, const double *a, const double *b, double *partials, int bs) {
  if (n > 0 && nblocks > 0 && (!a || !b)) return fail(SPX_ERR_ARG, "spx_bs4_1: null input");
  RedArgs A{a, b, nullptr, nullptr, 0.0};
  return run_stage1<RED_DOT>("spx_bs4_1", nblocks, n, A, partials, bs);
}

int spx_bs5_1(int64_t nblocks, int64_t n, const double *Ap, double alpha, double *r, double *partials,
              int bs) {
  if (n > 0 && nblocks > 0 && (!Ap || !r)) return fail(SPX_ERR_ARG, "spx_bs5_1: null input");
  RedArgs A{Ap, nullptr, r, nullptr, alpha};
  return run_stage1<RED_CG>("spx_bs5_1", nblocks, n, A, partials, bs);
}

int spx_reduce_2(int64_t nblocks, const double *partials, double *out, int bs) {
  SPX_NEED_DEVICE();
  if (nblocks < 0) return fail(SPX_ERR_ARG, "spx_reduce_2: negative nblocks");
  if (bs != 256 && bs != 512 && bs != 1024)
    return fail(SPX_ERR_ARG, "spx_reduce_2: block_size must be 256, 512 or 1024");
  if (!out || (nblocks > 0 && !partials)) return fail(SPX_ERR_ARG, "spx_reduce_2: null pointer");
  if (bs == 256) hipLaunchKernelGGL(k_red_ref2<256>, dim3(1), dim3(256), 0, current_stream(), nblocks, partials, out);
  else if (bs == 512) hipLaunchKernelGGL(k_red_ref2<512>, dim3(1), dim3(512), 0, current_stream(), nblocks, partials, out);
  else hipLaunchKernelGGL(k_red_ref2<1024>, dim3(1), dim3(1024), 0, current_stream(), nblocks, partials, out);
  SPX_LAUNCHED("spx_reduce_2");
  return SPX_OK;
}

int64_t spx_reduce_tmp_count(int64_t n, int mode, int block_size) {
  if (mode == SPX_REDUCE_REFERENCE) {
    if (block_size <= 0) return -1;
    int64_t nb = (n + block_size - 1) / block_size;
    if (nb > block_size) nb = block_size;
    return nb > 1 ? nb : 1;
  }
  return kFastGridMax;
}

int spx_bs3_norm(int64_t n, const double *a, double *tmp, double *out, int mode, int bs) {
  if (n > 0 && !a) return fail(SPX_ERR_ARG, "spx_bs3_norm: null a");
  RedArgs A{a, nullptr, nullptr, nullptr, 0.0};
  return run_reduction<RED_NORM>("spx_bs3_norm", n, A, tmp, out, mode, bs);
}

int spx_bs4_dot(int64_t n, const double *a, const double *b, double *tmp, double *out, int mode,
                int bs) {
  if (n > 0 && (!a || !b)) return fail(SPX_ERR_ARG, "spx_bs4_dot: null input");
  RedArgs A{a, b, nullptr, nullptr, 0.0};
  return run_reduction<RED_DOT>("spx_bs4_dot", n, A, tmp, out, mode, bs);
}

int spx_bs5_update(int64_t n, const double *Ap, double alpha, double *r, double *tmp, double *rdotr,
                   int mode, int bs) {
  if (n > 0 && (!Ap || !r)) return fail(SPX_ERR_ARG, "spx_bs5_update: null input");
  RedArgs A{Ap, nullptr, r, nullptr, alpha};
  return run_reduction<RED_CG>("spx_bs5_update", n, A, tmp, rdotr, mode, bs);
}

int spx_bs5_full(int64_t n, double alpha, const double *p, const double *Ap, double *x, double *r,
                 double *tmp, double *rdotr, int mode, int bs) {
  if (n > 0 && (!p || !Ap || !x || !r)) return fail(SPX_ERR_ARG, "spx_bs5_full: null input");
  RedArgs A{p, Ap, r, x, alpha, full_xshift(n), n};
  return run_reduction<RED_CGFULL>("spx_bs5_full", n, A, tmp, rdotr, mode, bs);
}

int spx_bs5_full_noalloc(int64_t n, double alpha, const double *p, const double *Ap, double *x, double *r,
                         double *rdotr, int mode, int bs) {
  SPX_NEED_DEVICE();
  const int dev = device_id();
  double *ptr = (dev >= 0 && dev < kScratchDevices) ? g_scratch_ptr[dev] : nullptr;
  if (!ptr) return fail(SPX_ERR_NODEVICE, "spx_bs5_full_noalloc: scratch not set up for this device");
  if (mode != SPX_REDUCE_REFERENCE && mode != SPX_REDUCE_FAST)
    return fail(SPX_ERR_ARG, "spx_bs5_full_noalloc: unknown reduction mode");
  const int64_t need = spx_reduce_tmp_count(n, mode, bs);
  if (need < 0 || need > kScratchDoubles)
    return fail(SPX_ERR_ARG, "spx_bs5_full_noalloc: " + std::to_string(need) + " partials exceed the " +
                                 std::to_string(kScratchDoubles) + "-double scratch");
  if (n > 0 && (!p || !Ap || !x || !r)) return fail(SPX_ERR_ARG, "spx_bs5_full_noalloc: null input");
  RedArgs A{p, Ap, r, x, alpha, full_xshift(n), n};
  return run_reduction<RED_CGFULL>("spx_bs5_full_noalloc", n, A, ptr, rdotr, mode, bs);
}

}  // extern "C"
