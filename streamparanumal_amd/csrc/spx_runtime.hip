// libspx runtime: device, streams, memory, events, graphs, synthetic inputs.
// Replaces the OCCA device used by libs/core (platformDeviceConfig.cpp:33-179,
// include/platform.hpp:60-176, include/memory.hpp:285-460).
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "spx_internal.hpp"
#include "spx_numeric.hpp"

namespace spx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

struct DeviceState {
  bool ready = false;
  int id = -1;
  int cus = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t current = nullptr;
  int *err_flag = nullptr;
};
static DeviceState g_dev;

bool device_ready() { return g_dev.ready; }
hipStream_t current_stream() { return g_dev.current; }
int *kernel_error_flag() { return g_dev.err_flag; }
int num_cus() { return g_dev.cus; }
int device_id() { return g_dev.ready ? g_dev.id : -1; }

static std::mutex g_tune_mu;
static std::map<std::string, spx_tuning_t> g_tune;

spx_tuning_t tuning_for(const char *kernel, spx_tuning_t d) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(kernel);
  if (it == g_tune.end()) return d;
  spx_tuning_t t = it->second;
  if (t.block_threads <= 0) t.block_threads = d.block_threads;
  if (t.xcd_remap < 0) t.xcd_remap = d.xcd_remap;
  if (t.unroll <= 0) t.unroll = d.unroll;
  if (t.grid_blocks <= 0) t.grid_blocks = d.grid_blocks;
  if (t.nontemporal < 0) t.nontemporal = d.nontemporal;
  if (t.variant < 0) t.variant = d.variant;
  return t;
}

}  // namespace spx

using namespace spx;

// ------------------------------------------------------------------ kernels --

__global__ void k_fill_bytes(uint8_t *p, size_t n, uint8_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// x_{i+skip} for the reference LCG: each thread jumps to its first element
// with the affine power (random.cpp:66-74) and then steps sequentially.
struct Affine {
  uint64_t a, c;
};
__device__ __forceinline__ Affine aff_compose(Affine f, Affine g) {
  return Affine{f.a * g.a, f.a * g.c + f.c};
}
__device__ __forceinline__ Affine aff_pow(Affine f, uint64_t n) {
  Affine r{1ULL, 0ULL};
  while (n) {
    if (n & 1ULL) r = aff_compose(r, f);
    f = aff_compose(f, f);
    n >>= 1;
  }
  return r;
}

constexpr uint64_t kLcgA = 6364136223846793005ULL;
constexpr uint64_t kLcgC = 1ULL;
constexpr int kFillPerThread = 16;

__global__ void k_fill_uniform(int64_t n, double *v, uint64_t x0, uint64_t skip) {
  const int64_t first = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kFillPerThread;
  if (first >= n) return;
  Affine j = aff_pow(Affine{kLcgA, kLcgC}, skip + (uint64_t)first);
  uint64_t x = j.a * x0 + j.c;
  // lanes of a wave write interleaved elements so stores stay coalesced
  for (int e = 0; e < kFillPerThread; ++e) {
    const int64_t i = first + e;
    if (i < n) v[i] = -1.0 + (double)(x >> 11) * 0x1p-52;
    x = kLcgA * x + kLcgC;
  }
}

// log-uniform magnitude in [1e-6, 1] with a random sign: element i takes draw
// 2(skip+i) for the magnitude 10^(-6u), u = (x >> 11) * 2^-53, and the top bit
// of draw 2(skip+i)+1 for the sign
__global__ void k_fill_loguniform(int64_t n, double *v, uint64_t x0, uint64_t skip) {
  const int64_t first = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kFillPerThread;
  if (first >= n) return;
  Affine j = aff_pow(Affine{kLcgA, kLcgC}, 2 * (skip + (uint64_t)first));
  uint64_t x = j.a * x0 + j.c;
  for (int e = 0; e < kFillPerThread; ++e) {
    const int64_t i = first + e;
    const double u = (double)(x >> 11) * 0x1p-53;
    x = kLcgA * x + kLcgC;
    const double mag = exp10(-6.0 * u);
    if (i < n) v[i] = (x >> 63) ? -mag : mag;
    x = kLcgA * x + kLcgC;
  }
}

__global__ void k_f64_to_f32(int64_t n, const double *s, float *d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) d[i] = (float)s[i];
}
__global__ void k_f32_to_f64(int64_t n, const float *s, double *d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) d[i] = (double)s[i];
}
__global__ void k_f64_to_bf16(int64_t n, const double *s, uint16_t *d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) d[i] = f64_to_bf16_bits(s[i]);
}
__global__ void k_bf16_to_f64(int64_t n, const uint16_t *s, double *d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) d[i] = (double)__uint_as_float(((uint32_t)s[i]) << 16);
}

static int grid_for(int64_t n, int threads) {
  int64_t g = (n + threads - 1) / threads;
  const int64_t cap = (int64_t)std::max(1, num_cus()) * 64;
  if (g > cap) g = cap;
  return (int)std::max<int64_t>(g, 1);
}

// --------------------------------------------------------------------- C ABI --
extern "C" {

const char *spx_last_error(void) { return g_last_error.c_str(); }
int spx_version(void) { return SPX_VERSION; }

int spx_device_count(int *count) {
  if (!count) return fail(SPX_ERR_ARG, "spx_device_count: null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return SPX_OK;
}

int spx_device_open(int device_id) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c == 0)
    return fail(SPX_ERR_NODEVICE, "spx_device_open: no HIP device visible");
  if (device_id < 0 || device_id >= c)
    return fail(SPX_ERR_ARG, "spx_device_open: device id " + std::to_string(device_id) +
                                 " out of range [0," + std::to_string(c) + ")");
  if (g_dev.ready && g_dev.id == device_id) return SPX_OK;
  if (g_dev.ready)
    return fail(SPX_ERR_UNSUPPORTED, "spx_device_open: one device per process (already opened " +
                                         std::to_string(g_dev.id) + ")");
  SPX_HIP(hipSetDevice(device_id));
  hipDeviceProp_t p;
  SPX_HIP(hipGetDeviceProperties(&p, device_id));
  g_dev.cus = p.multiProcessorCount;
  SPX_HIP(hipStreamCreateWithFlags(&g_dev.own_stream, hipStreamNonBlocking));
  SPX_HIP(hipMalloc((void **)&g_dev.err_flag, 256));
  SPX_HIP(hipMemset(g_dev.err_flag, 0, 256));
  SPX_HIP(hipDeviceSynchronize());
  if (int rc = scratch_setup(device_id)) return rc;
  g_dev.current = g_dev.own_stream;
  g_dev.id = device_id;
  g_dev.ready = true;
  return SPX_OK;
}

int spx_device_finish(void) {
  SPX_NEED_DEVICE();
  SPX_HIP(hipStreamSynchronize(g_dev.current));
  if (g_dev.current != g_dev.own_stream) SPX_HIP(hipStreamSynchronize(g_dev.own_stream));
  return SPX_OK;
}

int spx_device_props(spx_device_props_t *out) {
  SPX_NEED_DEVICE();
  if (!out) return fail(SPX_ERR_ARG, "spx_device_props: null");
  hipDeviceProp_t p;
  SPX_HIP(hipGetDeviceProperties(&p, g_dev.id));
  std::memset(out, 0, sizeof(*out));
  std::strncpy(out->name, p.name, sizeof(out->name) - 1);
  if (!out->name[0]) (void)hipDeviceGetName(out->name, (int)sizeof(out->name) - 1, g_dev.id);
  std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
  out->device_id = g_dev.id;
  out->compute_units = p.multiProcessorCount;
  out->clock_khz = p.clockRate;
  out->memory_clock_khz = p.memoryClockRate;
  out->memory_bus_width = p.memoryBusWidth;
  out->global_mem_bytes = (int64_t)p.totalGlobalMem;
  out->l2_bytes = (int64_t)p.l2CacheSize;
  out->lds_bytes_per_block = (int64_t)p.sharedMemPerBlock;
  out->warp_size = p.warpSize;
  return SPX_OK;
}

int spx_check_kernel_errors(void) {
  SPX_NEED_DEVICE();
  int h = 0;
  SPX_HIP(hipStreamSynchronize(g_dev.current));
  SPX_HIP(hipMemcpy(&h, g_dev.err_flag, sizeof(int), hipMemcpyDeviceToHost));
  if (h) {
    int z = 0;
    SPX_HIP(hipMemcpy(g_dev.err_flag, &z, sizeof(int), hipMemcpyHostToDevice));
    return fail(SPX_ERR_KERNEL, "a kernel found a row block larger than its max_block_nnz "
                                "(flag " + std::to_string(h) + "); results of that launch are incomplete");
  }
  return SPX_OK;
}

int spx_stream_create(spx_stream_t *s) {
  SPX_NEED_DEVICE();
  if (!s) return fail(SPX_ERR_ARG, "spx_stream_create: null");
  hipStream_t h;
  SPX_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
  *s = reinterpret_cast<spx_stream_t>(h);
  return SPX_OK;
}

int spx_stream_destroy(spx_stream_t s) {
  SPX_NEED_DEVICE();
  hipStream_t h = reinterpret_cast<hipStream_t>(s);
  if (!h || h == g_dev.own_stream) return fail(SPX_ERR_ARG, "spx_stream_destroy: not a user stream");
  if (g_dev.current == h) g_dev.current = g_dev.own_stream;
  SPX_HIP(hipStreamDestroy(h));
  return SPX_OK;
}

int spx_stream_finish(spx_stream_t s) {
  SPX_NEED_DEVICE();
  hipStream_t h = s ? reinterpret_cast<hipStream_t>(s) : g_dev.own_stream;
  SPX_HIP(hipStreamSynchronize(h));
  return SPX_OK;
}

int spx_set_stream(spx_stream_t s) {
  SPX_NEED_DEVICE();
  g_dev.current = s ? reinterpret_cast<hipStream_t>(s) : g_dev.own_stream;
  return SPX_OK;
}

spx_stream_t spx_get_stream(void) { return reinterpret_cast<spx_stream_t>(g_dev.current); }

int spx_malloc(void **ptr, size_t bytes) {
  SPX_NEED_DEVICE();
  if (!ptr) return fail(SPX_ERR_ARG, "spx_malloc: null ptr");
  *ptr = nullptr;
  if (bytes == 0) return SPX_OK;
  SPX_HIP(hipMalloc(ptr, bytes));
  return SPX_OK;
}

int spx_malloc_host(void **ptr, size_t bytes) {
  SPX_NEED_DEVICE();
  if (!ptr) return fail(SPX_ERR_ARG, "spx_malloc_host: null ptr");
  *ptr = nullptr;
  if (bytes == 0) return SPX_OK;
  SPX_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
  return SPX_OK;
}

int spx_free(void *ptr) {
  SPX_NEED_DEVICE();
  if (ptr) SPX_HIP(hipFree(ptr));
  return SPX_OK;
}

int spx_free_host(void *ptr) {
  SPX_NEED_DEVICE();
  if (ptr) SPX_HIP(hipHostFree(ptr));
  return SPX_OK;
}

int spx_memcpy(void *dst, const void *src, size_t bytes, int async) {
  SPX_NEED_DEVICE();
  if (bytes == 0) return SPX_OK;
  if (!dst || !src) return fail(SPX_ERR_ARG, "spx_memcpy: null pointer");
  SPX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, g_dev.current));
  if (!async) SPX_HIP(hipStreamSynchronize(g_dev.current));
  return SPX_OK;
}

int spx_fill_bytes(void *ptr, size_t bytes, uint8_t value) {
  SPX_NEED_DEVICE();
  if (bytes == 0) return SPX_OK;
  if (!ptr) return fail(SPX_ERR_ARG, "spx_fill_bytes: null pointer");
  const int threads = 256;
  hipLaunchKernelGGL(k_fill_bytes, dim3(grid_for((int64_t)bytes, threads)), dim3(threads), 0,
                     g_dev.current, (uint8_t *)ptr, bytes, value);
  SPX_LAUNCHED("spx_fill_bytes");
  return SPX_OK;
}

int spx_event_create(spx_event_t *ev) {
  SPX_NEED_DEVICE();
  if (!ev) return fail(SPX_ERR_ARG, "spx_event_create: null");
  hipEvent_t e;
  SPX_HIP(hipEventCreate(&e));
  *ev = reinterpret_cast<spx_event_t>(e);
  return SPX_OK;
}

int spx_event_destroy(spx_event_t ev) {
  SPX_NEED_DEVICE();
  if (ev) SPX_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)));
  return SPX_OK;
}

int spx_event_record(spx_event_t ev) {
  SPX_NEED_DEVICE();
  if (!ev) return fail(SPX_ERR_ARG, "spx_event_record: null");
  SPX_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), g_dev.current));
  return SPX_OK;
}

int spx_event_elapsed_ms(spx_event_t a, spx_event_t b, float *ms) {
  SPX_NEED_DEVICE();
  if (!a || !b || !ms) return fail(SPX_ERR_ARG, "spx_event_elapsed_ms: null");
  SPX_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(b)));
  SPX_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(a), reinterpret_cast<hipEvent_t>(b)));
  return SPX_OK;
}

struct spx_graph_s {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

int spx_graph_begin(void) {
  SPX_NEED_DEVICE();
  SPX_HIP(hipStreamBeginCapture(g_dev.current, hipStreamCaptureModeThreadLocal));
  return SPX_OK;
}

int spx_graph_end(spx_graph_t *out) {
  SPX_NEED_DEVICE();
  if (!out) return fail(SPX_ERR_ARG, "spx_graph_end: null");
  hipGraph_t g;
  SPX_HIP(hipStreamEndCapture(g_dev.current, &g));
  hipGraphExec_t x;
  hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(SPX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  *out = new spx_graph_s{g, x};
  return SPX_OK;
}

int spx_graph_launch(spx_graph_t g) {
  SPX_NEED_DEVICE();
  if (!g) return fail(SPX_ERR_ARG, "spx_graph_launch: null");
  SPX_HIP(hipGraphLaunch(g->exec, g_dev.current));
  return SPX_OK;
}

int spx_graph_destroy(spx_graph_t g) {
  SPX_NEED_DEVICE();
  if (!g) return SPX_OK;
  SPX_HIP(hipGraphExecDestroy(g->exec));
  SPX_HIP(hipGraphDestroy(g->graph));
  delete g;
  return SPX_OK;
}

int spx_fill_loguniform_f64(int64_t n, double *v, uint64_t seed, uint64_t skip) {
  SPX_NEED_DEVICE();
  if (n < 0) return fail(SPX_ERR_ARG, "spx_fill_loguniform_f64: negative n");
  if (n == 0) return SPX_OK;
  if (!v) return fail(SPX_ERR_ARG, "spx_fill_loguniform_f64: null v");
  const uint64_t x0 = kLcgA * (UINT64_MAX ^ seed) + kLcgC;
  const int threads = 256;
  const int64_t nthreads = (n + kFillPerThread - 1) / kFillPerThread;
  const int64_t blocks = (nthreads + threads - 1) / threads;
  if (blocks > INT32_MAX) return fail(SPX_ERR_ARG, "spx_fill_loguniform_f64: n too large");
  hipLaunchKernelGGL(k_fill_loguniform, dim3((unsigned)blocks), dim3(threads), 0, g_dev.current, n, v, x0, skip);
  SPX_LAUNCHED("spx_fill_loguniform_f64");
  return SPX_OK;
}

int spx_fill_uniform_f64(int64_t n, double *v, uint64_t seed, uint64_t skip) {
  SPX_NEED_DEVICE();
  if (n < 0) return fail(SPX_ERR_ARG, "spx_fill_uniform_f64: negative n");
  if (n == 0) return SPX_OK;
  if (!v) return fail(SPX_ERR_ARG, "spx_fill_uniform_f64: null v");
  const uint64_t x0 = kLcgA * (UINT64_MAX ^ seed) + kLcgC;
  const int threads = 256;
  const int64_t nthreads = (n + kFillPerThread - 1) / kFillPerThread;
  const int64_t blocks = (nthreads + threads - 1) / threads;
  if (blocks > INT32_MAX) return fail(SPX_ERR_ARG, "spx_fill_uniform_f64: n too large");
  hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)blocks), dim3(threads), 0, g_dev.current, n, v,
                     x0, skip);
  SPX_LAUNCHED("spx_fill_uniform_f64");
  return SPX_OK;
}

#define SPX_CONVERT(NAME, KERN, TS, TD)                                               \
  int NAME(int64_t n, const TS *s, TD *d) {                                           \
    SPX_NEED_DEVICE();                                                                \
    if (n < 0) return fail(SPX_ERR_ARG, #NAME ": negative n");                       \
    if (n == 0) return SPX_OK;                                                        \
    if (!s || !d) return fail(SPX_ERR_ARG, #NAME ": null pointer");                  \
    hipLaunchKernelGGL(KERN, dim3(grid_for(n, 256)), dim3(256), 0, g_dev.current, n, s, d); \
    SPX_LAUNCHED(#NAME);                                                              \
    return SPX_OK;                                                                    \
  }

SPX_CONVERT(spx_convert_f64_to_f32, k_f64_to_f32, double, float)
SPX_CONVERT(spx_convert_f32_to_f64, k_f32_to_f64, float, double)
SPX_CONVERT(spx_convert_f64_to_bf16, k_f64_to_bf16, double, uint16_t)
SPX_CONVERT(spx_convert_bf16_to_f64, k_bf16_to_f64, uint16_t, double)

int spx_set_tuning(const char *kernel, const spx_tuning_t *t) {
  if (!kernel || !t) return fail(SPX_ERR_ARG, "spx_set_tuning: null");
  if (t->block_threads != 0 &&
      (t->block_threads < 64 || t->block_threads > 1024 || t->block_threads % 64))
    return fail(SPX_ERR_ARG, "spx_set_tuning: block_threads must be a multiple of 64 in [64,1024]");
  if (t->unroll < 0 || t->unroll > 32) return fail(SPX_ERR_ARG, "spx_set_tuning: unroll in [0,32]");
  if (t->grid_blocks < 0) return fail(SPX_ERR_ARG, "spx_set_tuning: grid_blocks >= 0");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[kernel] = *t;
  return SPX_OK;
}

int spx_get_tuning(const char *kernel, spx_tuning_t *t) {
  if (!kernel || !t) return fail(SPX_ERR_ARG, "spx_get_tuning: null");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(kernel);
  if (it == g_tune.end()) {
    std::memset(t, 0, sizeof(*t));
    t->xcd_remap = -1;
    t->nontemporal = -1;
    t->variant = -1;
    return SPX_OK;
  }
  *t = it->second;
  return SPX_OK;
}

}  // extern "C"
