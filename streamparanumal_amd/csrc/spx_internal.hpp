// Internal helpers shared by the libspx HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/spx.h"

namespace spx {

// ---- error state -----------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define SPX_HIP(call)                                                             \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::spx::fail(SPX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Launch check: catches invalid launch configurations without waiting.
#define SPX_LAUNCHED(name)                                                        \
  do {                                                                            \
    hipError_t e_ = hipGetLastError();                                            \
    if (e_ != hipSuccess)                                                         \
      return ::spx::fail(SPX_ERR_HIP, std::string(name) + " launch: " + hipGetErrorString(e_)); \
  } while (0)

// ---- device state ------------------------------------------------------------
bool device_ready();
hipStream_t current_stream();
int *kernel_error_flag();   // device word kernels set on a violated precondition
int num_cus();
int device_id();            // the open device's HIP id, -1 when none is open
// takes the device address of the allocation-free entry point's scratch array
// (spx_stream.hip); called once by spx_device_open with the device current
int scratch_setup(int device_id);

#define SPX_NEED_DEVICE()                                                          \
  do {                                                                            \
    if (!::spx::device_ready())                                                   \
      return ::spx::fail(SPX_ERR_NODEVICE, "no device open: call spx_device_open first"); \
  } while (0)

// ---- tuning table ----------------------------------------------------------------
spx_tuning_t tuning_for(const char *kernel, spx_tuning_t defaults);

// ---- device helpers ----------------------------------------------------------------
// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with a private L2.  Remap the flat workgroup id so that XCD x takes blocks
// of B = 2^lblk consecutive tiles in turn (tiles g*nxcd*B + x*B ... + B - 1
// of group g): neighbouring tiles, which share cache lines at their edges,
// meet in one L2, while the XCDs stay within nxcd*B tiles of each other.
// (One contiguous range per XCD put the XCDs' streams a power of two apart in
// memory on power-of-two meshes, 2 % slower there.)  Blocks of 128 tiles were
// the fastest of 16..1024 for FP64 values on 64^3 p=7, 58^3 p=7, 116^3 p=3
// and 46^3 p=9; the FP32 / BF16 gather+scatter, whose tiles move half or a
// quarter of the bytes, runs fastest with 32 (64^3 p=7, us per call for
// blocks of 128 / 64 / 32 / 16: FP32 286 / 273 / 270 / 267, BF16 254 / 251 /
// 249 / 250; 232^3 p=1 BF16 230 / 232 / 219 / 236), while FP64 there holds
// at 128 (Tet p=3: 514 / 521 / 567 / 644).  Tiles past the last whole group
// keep their id.  Bijective on [0, n).
constexpr int kXcdLog2 = 7;       // blocks of 128 tiles
constexpr int kXcdLog2Mixed = 5;  // blocks of 32: the FP32 / BF16 gather+scatter
__device__ __forceinline__ int xcd_tile(int bid, int n, int nxcd, int lblk) {
  if (nxcd <= 1) return bid;
  const int B = 1 << lblk;
  const int group = nxcd * B;
  const int l = bid / nxcd, x = bid % nxcd;
  if (l >= (n / group) * B) return bid;
  return (l >> lblk) * group + x * B + (l & (B - 1));
}

// XCD remap of an spx_srb_t hint: negative the library default (defaultOn,
// with `lblk`), 0 off, 1 on with `lblk`, a power of two 2..1024: blocks of
// that many tiles.  Returns false for other values.
inline bool xcd_hint(int hint, int defaultOn, int &nxcd, int &lblk) {
  if (hint < 0) {
    nxcd = defaultOn ? 8 : 1;
  } else if (hint == 0 || hint == 1) {
    nxcd = hint ? 8 : 1;
  } else {
    if (hint < 2 || hint > 1024 || (hint & (hint - 1))) return false;
    nxcd = 8;
    lblk = 0;
    while ((1 << lblk) < hint) ++lblk;
  }
  return true;
}

inline bool aligned(const void *p, size_t a) {
  return (reinterpret_cast<uintptr_t>(p) % a) == 0;
}

}  // namespace spx
