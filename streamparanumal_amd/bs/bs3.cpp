// BS3: norm, sum of a^2 (BS/BS3/src/bs3Run.cpp:29-99, bs3Setup.cpp:29-45).
// --reduction reference (default) runs the bs3_1/bs3_2 pair in the
// reference's order and block count; --reduction fast runs the full-machine
// deterministic reduction instead.
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
    settings.parseSettings(argc, argv);
    if (settings.helpRequested()) return;

    platform_t platform(settings);
    if (settings.compareSetting("VERBOSE", "TRUE")) settings.report();
    // bs3Setup.cpp: p_blockSize = 256
    const int blockSize = 256;
    properties_t kernelInfo = platform.props();
    kernelInfo["defines/p_blockSize"] = blockSize;
    kernel_t kernel1 = platform.buildKernel("BS/BS3/okl/bs3.okl", "bs3_1", kernelInfo);
    kernel_t kernel2 = platform.buildKernel("BS/BS3/okl/bs3.okl", "bs3_2", kernelInfo);
    const bool fast = settings.compareSetting("REDUCTION", "fast");

    const long long sc = 1 * sizeof(dfloat);
    bs::Sweep w = bs::byteSweep(settings, sc);
    unsigned long long seed = 1;
    settings.getSetting("SEED", seed);
    deviceMemory<dfloat> o_a = platform.malloc<dfloat>(w.Nmax);
    bs::fillUniform(o_a, seed, 0);
    const long long ntmp = std::max<long long>(blockSize, spx_reduce_tmp_count(w.Nmax, SPX_REDUCE_FAST, blockSize));
    deviceMemory<dfloat> o_tmp = platform.malloc<dfloat>(ntmp);
    deviceMemory<dfloat> o_out = platform.malloc<dfloat>(1);

    auto launch = [&](long long N) {
      if (fast) {
        SPX_CHECK(spx_bs3_norm(N, o_a.ptr(), o_tmp.ptr(), o_out.ptr(), SPX_REDUCE_FAST, blockSize));
      } else {
        long long Nblock = (N + blockSize - 1) / blockSize;
        Nblock = (Nblock > blockSize) ? blockSize : Nblock;  // bs3Run.cpp:80-81
        kernel1(Nblock, N, o_a, o_tmp);
        kernel2(Nblock, o_tmp, o_out);
      }
    };
    for (int n = 0; n < 5; ++n) launch(w.Nmax);

    if (w.single) printf("BS3 = [");
    else printf("%%Norm [DOFs, elapsed, DOFs/s, BW (GB/s)]\nBS3 = [\n");
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
      const double bytes = 1.0 * N * sizeof(dfloat);
      printf("%lld %5.4e %5.4e %6.2f", N, elapsed, N / elapsed, (bytes / 1.e9) / elapsed);
      if (N < w.Nmax) printf(";\n");
      tdev = dt.seconds() / Ntests;
    }
    printf("]; %%Norm [DOFs, elapsed, DOFs/s, BW (GB/s)]\n");
    double result = 0;
    o_out.copyTo(&result, 1);
    printf("BS3device = [%lld %5.4e %6.2f %.17g]; %%Norm kernel-only [DOFs, elapsed, BW (GB/s), result] at the largest size (%s order)\n",
           w.Nmax, tdev, 1.0 * w.Nmax * sizeof(dfloat) / 1e9 / tdev, result, fast ? "fast" : "reference");
  });
}
