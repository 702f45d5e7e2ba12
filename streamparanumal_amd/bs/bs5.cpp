// BS5: CG update, r += alpha*Ap and rdotr = sum r^2 (BS/BS5/src/bs5Run.cpp:29-103, bs5Setup.cpp:29-45).
// --reduction reference (default) runs the bs5_1/bs5_2 pair in the
// reference's order and block count; --reduction fast runs the full-machine
// deterministic reduction instead.
// --form full runs the README's update x += alpha p, r -= alpha Ap,
// rdotr = sum r^2 (README.md:49-53), which no reference kernel computes;
// its bytes are 48 per entry.  --alpha defaults to the reference's 1.0 (:55).
// --storage FP32 / BF16 runs the full form on vectors stored in that type
// (update in FP32, rdotr in FP64); -b then maps to N with 3 stored values per
// entry, and the bytes are 6 stored values per entry.
#include "bsCommon.hpp"

using namespace libp;

int main(int argc, char **argv) {
  return bs::run([&] {
    settings_t settings(comm_t::world());
    platformAddSettings(settings);
    bs::addByteSettings(settings);
    bs::addVerbose(settings);
    bs::addSeed(settings);
    settings.newSetting("--reduction", "--reduction", "REDUCTION", "reference",
                        "Summation order of the reduction", {"reference", "fast"});
    settings.newSetting("--alpha", "--alpha", "ALPHA", "1.0", "alpha of the CG update");
    settings.newSetting("--form", "--form", "FORM", "reference", "Kernel form", {"reference", "full"});
    bs::addStorage(settings);
    settings.parseSettings(argc, argv);
    if (settings.helpRequested()) return;

    platform_t platform(settings);
    if (settings.compareSetting("VERBOSE", "TRUE")) settings.report();
    // bs5Setup.cpp: p_blockSize = 1024 on CUDA, else 256
    const int blockSize = 256;
    properties_t kernelInfo = platform.props();
    kernelInfo["defines/p_blockSize"] = blockSize;
    kernel_t kernel1 = platform.buildKernel("BS/BS5/okl/bs5.okl", "bs5_1", kernelInfo);
    kernel_t kernel2 = platform.buildKernel("BS/BS5/okl/bs5.okl", "bs5_2", kernelInfo);
    const bool fast = settings.compareSetting("REDUCTION", "fast");
    const bool full = settings.compareSetting("FORM", "full");
    const int storage = bs::storageOf(settings);
    const bool mixed = storage != SPX_STORE_F64;
    LIBP_ABORT("--storage " << settings.getSetting("STORAGE") << " runs the full form: add --form full",
               mixed && !full);
    const size_t sb = bs::storageBytes(storage);

    const long long sc = 3 * (long long)sb;
    bs::Sweep w = bs::byteSweep(settings, sc);
    unsigned long long seed = 1;
    settings.getSetting("SEED", seed);
    dfloat alpha = 1.0;
    settings.getSetting("ALPHA", alpha);
    deviceMemory<dfloat> o_Ap, o_r, o_p, o_x;
    deviceMemory<unsigned char> s_p, s_Ap, s_x, s_r;  // stored FP32 / BF16 vectors
    if (mixed) {
      s_p = bs::stored(platform, storage, w.Nmax, seed, 2 * w.Nmax);
      s_Ap = bs::stored(platform, storage, w.Nmax, seed, 0);
      s_x = bs::stored(platform, storage, w.Nmax, seed, 3 * w.Nmax);
      s_r = bs::stored(platform, storage, w.Nmax, seed, w.Nmax);
    } else {
      o_Ap = platform.malloc<dfloat>(w.Nmax);
      o_r = platform.malloc<dfloat>(w.Nmax);
      bs::fillUniform(o_Ap, seed, 0);
      bs::fillUniform(o_r, seed, w.Nmax);
    }
    if (full && !mixed) {
      o_p = platform.malloc<dfloat>(w.Nmax);
      o_x = platform.malloc<dfloat>(w.Nmax);
      bs::fillUniform(o_p, seed, 2 * w.Nmax);
      bs::fillUniform(o_x, seed, 3 * w.Nmax);
    }
    const long long ntmp = std::max<long long>(std::max<long long>(blockSize, SPX_MIXED_TMP_COUNT),
                                               spx_reduce_tmp_count(w.Nmax, SPX_REDUCE_FAST, blockSize));
    deviceMemory<dfloat> o_tmp = platform.malloc<dfloat>(ntmp);
    deviceMemory<dfloat> o_out = platform.malloc<dfloat>(1);

    auto launch = [&](long long N) {
      if (mixed) {
        SPX_CHECK(spx_bs5_full_mixed(storage, N, alpha, s_p.ptr(), s_Ap.ptr(), s_x.ptr(), s_r.ptr(), o_tmp.ptr(),
                                     o_out.ptr()));
      } else if (fast) {
        if (full)
          SPX_CHECK(spx_bs5_full(N, alpha, o_p.ptr(), o_Ap.ptr(), o_x.ptr(), o_r.ptr(), o_tmp.ptr(), o_out.ptr(), SPX_REDUCE_FAST, blockSize));
        else
          SPX_CHECK(spx_bs5_update(N, o_Ap.ptr(), alpha, o_r.ptr(), o_tmp.ptr(), o_out.ptr(), SPX_REDUCE_FAST, blockSize));
      } else {
        bool continue_ref = true;
        long long Nblock = (N + blockSize - 1) / blockSize;
        Nblock = (Nblock > blockSize) ? blockSize : Nblock;  // bs5Run.cpp:80-81
        if (full) {
          SPX_CHECK(spx_bs5_full(N, alpha, o_p.ptr(), o_Ap.ptr(), o_x.ptr(), o_r.ptr(), o_tmp.ptr(), o_out.ptr(), SPX_REDUCE_REFERENCE, blockSize));
          continue_ref = false;
        } else {
          kernel1(Nblock, N, o_Ap, alpha, o_r, o_tmp);
        }
        if (continue_ref) kernel2(Nblock, o_tmp, o_out);
      }
    };
    for (int n = 0; n < 5; ++n) launch(w.Nmax);

    if (w.single) printf("BS5 = [");
    else printf("%%CG Update [DOFs, elapsed, DOFs/s, BW (GB/s)]\nBS5 = [\n");
    deviceTimer_t dt;
    double tdev = 0;
    for (long long N = w.Nmin; N <= w.Nmax; N += w.Nstep) {
      const int Ntests = 20;
      timePoint_t start = GlobalPlatformTime(platform);
      dt.start();
      for (int n = 0; n < Ntests; ++n) launch(N);
      dt.stop();
      timePoint_t end = GlobalPlatformTime(platform);
      const double elapsed = ElapsedTime(start, end) / Ntests;
      const double bytes = (full ? 6.0 : 3.0) * N * sb;
      printf("%lld %5.4e %5.4e %6.2f", N, elapsed, N / elapsed, (bytes / 1.e9) / elapsed);
      if (N < w.Nmax) printf(";\n");
      tdev = dt.seconds() / Ntests;
    }
    printf("]; %%CG Update [DOFs, elapsed, DOFs/s, BW (GB/s)]\n");
    double result = 0;
    o_out.copyTo(&result, 1);
    printf("BS5device = [%lld %5.4e %6.2f %.17g]; %%CG Update kernel-only [DOFs, elapsed, BW (GB/s), result] at the largest size (%s)\n",
           w.Nmax, tdev, (full ? 6.0 : 3.0) * w.Nmax * sb / 1e9 / tdev, result,
           mixed ? (storage == SPX_STORE_F32 ? "FP32 storage" : "BF16 storage") : (fast ? "fast order" : "reference order"));
  });
}
