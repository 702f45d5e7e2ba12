// Shared pieces of the BS0-BS8 drivers (BS/BSn/bsNMain.cpp, src/bsN*.cpp).
//
// Each driver keeps the reference's flags, defaults, timing loop and printed
// MATLAB-style row.  Additions are long-only flags (no short key of the
// reference is reused): --seed initialises the inputs with the reference LCG
// (the reference leaves them uninitialised), and a second printed row gives
// the kernel-only time from HIP events next to the reference's host timing.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <functional>

#include "mesh.hpp"
#include "ogs.hpp"
#include "platform.hpp"
#include "settings.hpp"
#include "timer.hpp"

namespace bs {

using namespace libp;

// Runs a driver body, turning libp::exception into a message and exit code 1.
inline int run(const std::function<void()> &body) {
  try {
    body();
    return 0;
  } catch (const libp::exception &e) {
    std::fprintf(stderr, "%s", e.what());
    return 1;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}

// The reference's ogs_t waits on the device inside every Gather, Scatter and
// GatherScatter call when MPI is not GPU-aware (libs/ogs/ogs.cpp:74,104,117,
// 243, 435), so its BS6-BS8 loops time one synchronised call after another.
// This ogs_t only enqueues (one rank has no halo to wait for), and the BSn
// rows time back-to-back calls.  timeSynced repeats `step` with a host finish
// after every call, which is what the reference's loop measures.
inline double timeSynced(platform_t &platform, int Ntests, const std::function<void()> &step) {
  step();
  platform.finish();
  timePoint_t start = GlobalPlatformTime(platform);
  for (int n = 0; n < Ntests; ++n) {
    step();
    platform.finish();
  }
  timePoint_t end = GlobalPlatformTime(platform);
  return ElapsedTime(start, end) / Ntests;
}

// --settle S (seconds, default 0 = the reference's 5 dry runs only): repeat
// the step untimed for S seconds first, so the timed loops see the GPU at its
// sustained clocks rather than ramping up after the host-side setup.
inline void addSettle(settings_t &s) {
  s.newSetting("--settle", "--settle", "SETTLE", "0", "Seconds of untimed calls before timing");
}
inline void settle(settings_t &s, platform_t &platform, const std::function<void()> &step) {
  double seconds = 0.0;
  s.getSetting("SETTLE", seconds);
  if (seconds <= 0.0) return;
  const timePoint_t t0 = PlatformTime(platform);
  do {
    for (int n = 0; n < 8; ++n) step();
  } while (ElapsedTime(t0, PlatformTime(platform)) < seconds);
}

// The BSnsync row: [N, DOFs, elapsed per synchronised call, DOFs/s, BW (GB/s)].
inline void printSynced(const char *name, const char *what, int N, hlong Ndofs, double t, size_t bytes) {
  printf("%ssync = [%d, " hlongFormat ", %5.4le, %5.4le, %6.2f]; %% %s, host finish after every call as in the "
         "reference's ogs_t [N, DOFs, elapsed, DOFs/s, BW (GB/s)] (the %s row times back-to-back calls)\n",
         name, N, Ndofs, t, Ndofs / t, bytes / (1.0e9 * t), what, name);
}

inline void addVerbose(settings_t &s) {
  s.newSetting("-v", "--verbose", "VERBOSE", "FALSE", "Enable verbose output", {"TRUE", "FALSE"});
}

inline void addSeed(settings_t &s, const char *dflt = "1") {
  s.newSetting("--seed", "--seed", "SEED", dflt, "Seed of the LCG that fills the input vectors (U[-1,1))");
}

// byte-sweep settings of BS1-BS5 (bs1Settings.cpp:35-53)
inline void addByteSettings(settings_t &s) {
  s.newSetting("-b", "--bytes", "BYTES", "0", "Array size in bytes");
  s.newSetting("-bmin", "--bmin", "BMIN", "0", "Minimum array size in bytes");
  s.newSetting("-bmax", "--bmax", "BMAX", "0", "Maximum array size in bytes");
  s.newSetting("-s", "--bstep", "BSTEP", "1048576", "Array step size in bytes");
}

struct Sweep {
  bool single = false;
  long long Nmin = 0, Nmax = 0, Nstep = 1;
};

// bs1Run.cpp:32-47: default single test of 1 GB; sizes in entries = bytes/sc
inline Sweep byteSweep(settings_t &s, long long sc) {
  size_t B = 0, Bmin = 0, Bmax = 0, Bstep = 0;
  s.getSetting("BYTES", B);
  s.getSetting("BMIN", Bmin);
  s.getSetting("BMAX", Bmax);
  s.getSetting("BSTEP", Bstep);
  if (!(B | Bmin | Bmax)) B = 1073741824;
  if (B) Bmax = B;
  Sweep w;
  w.Nmin = (long long)(Bmin / sc);
  w.Nmax = (long long)(Bmax / sc);
  w.Nstep = ((long long)(Bstep / sc) > 0) ? (long long)(Bstep / sc) : 1;
  if (B) {
    w.single = true;
    w.Nmin = w.Nmax = (long long)(B / sc);
  }
  return w;
}

inline void fillUniform(deviceMemory<dfloat> &m, unsigned long long seed, unsigned long long skip) {
  if (m.length()) SPX_CHECK(spx_fill_uniform_f64((int64_t)m.length(), m.ptr(), seed, skip));
}

// --storage: the stored value type of the reduced-precision forms (sums and
// reductions stay FP64); FP64 runs the plain kernels
inline void addStorage(settings_t &s) {
  s.newSetting("--storage", "--storage", "STORAGE", "FP64",
               "Stored value type (FP32 / BF16: FP64 row sums and reductions)", {"FP64", "FP32", "BF16"});
}
inline int storageOf(settings_t &s) {
  if (s.compareSetting("STORAGE", "FP32")) return SPX_STORE_F32;
  if (s.compareSetting("STORAGE", "BF16")) return SPX_STORE_BF16;
  return SPX_STORE_F64;
}
inline size_t storageBytes(int storage) {
  return storage == SPX_STORE_F64 ? 8 : storage == SPX_STORE_F32 ? 4 : 2;
}

// n values of U[-1,1) from the LCG (draws skip..skip+n-1), rounded to the
// storage type (nearest even) into a device array of n * storageBytes bytes
inline deviceMemory<unsigned char> stored(platform_t &platform, int storage, size_t n, unsigned long long seed,
                                          unsigned long long skip) {
  deviceMemory<unsigned char> out = platform.malloc<unsigned char>(n * storageBytes(storage));
  if (!n) return out;
  if (storage == SPX_STORE_F64) {
    SPX_CHECK(spx_fill_uniform_f64((int64_t)n, reinterpret_cast<double *>(out.ptr()), seed, skip));
    return out;
  }
  deviceMemory<dfloat> tmp = platform.malloc<dfloat>(n);
  SPX_CHECK(spx_fill_uniform_f64((int64_t)n, tmp.ptr(), seed, skip));
  if (storage == SPX_STORE_F32)
    SPX_CHECK(spx_convert_f64_to_f32((int64_t)n, tmp.ptr(), reinterpret_cast<float *>(out.ptr())));
  else
    SPX_CHECK(spx_convert_f64_to_bf16((int64_t)n, tmp.ptr(), reinterpret_cast<uint16_t *>(out.ptr())));
  platform.finish();
  return out;
}

}  // namespace bs
