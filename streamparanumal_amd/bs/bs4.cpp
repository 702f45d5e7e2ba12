// BS4: dot, sum of a*b (BS/BS4/src/bs4Run.cpp:29-99, bs4Setup.cpp:29-45).
// --reduction reference (default) runs the bs4_1/bs4_2 pair in the
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
    // bs4Setup.cpp: p_blockSize = 256
    const int blockSize = 256;
    properties_t kernelInfo = platform.props();
    kernelInfo["defines/p_blockSize"] = blockSize;
    kernel_t kernel1 = platform.buildKernel("BS/BS4/okl/bs4.okl", "bs4_1", kernelInfo);
    kernel_t kernel2 = platform.buildKernel("BS/BS4/okl/bs4.okl", "bs4_2", kernelInfo);
    const bool fast = settings.compareSetting("REDUCTION", "fast");

    const long long sc = 2 * sizeof(dfloat);
    bs::Sweep w = bs::byteSweep(settings, sc);
    unsigned long long seed = 1;
    settings.getSetting("SEED", seed);
    deviceMemory<dfloat> o_a = platform.malloc<dfloat>(w.Nmax);
    deviceMemory<dfloat> o_b = platform.malloc<dfloat>(w.Nmax);
    bs::fillUniform(o_a, seed, 0);
    bs::fillUniform(o_b, seed, w.Nmax);
    const long long ntmp = std::max<long long>(blockSize, spx_reduce_tmp_count(w.Nmax, SPX_REDUCE_FAST, blockSize));
    deviceMemory<dfloat> o_tmp = platform.malloc<dfloat>(ntmp);
    deviceMemory<dfloat> o_out = platform.malloc<dfloat>(1);

    auto launch = [&](long long N) {
      if (fast) {
        SPX_CHECK(spx_bs4_dot(N, o_a.ptr(), o_b.ptr(), o_tmp.ptr(), o_out.ptr(), SPX_REDUCE_FAST, blockSize));
      } else {
        long long Nblock = (N + blockSize - 1) / blockSize;
        Nblock = (Nblock > blockSize) ? blockSize : Nblock;  // bs4Run.cpp:80-81
        kernel1(Nblock, N, o_a, o_b, o_tmp);
        kernel2(Nblock, o_tmp, o_out);
      }
    };
    for (int n = 0; n < 5; ++n) launch(w.Nmax);

    if (w.single) printf("BS4 = [");
    else printf("%%Dot [DOFs, elapsed, DOFs/s, BW (GB/s)]\nBS4 = [\n");
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
      const double bytes = 2.0 * N * sizeof(dfloat);
      printf("%lld %5.4e %5.4e %6.2f", N, elapsed, N / elapsed, (bytes / 1.e9) / elapsed);
      if (N < w.Nmax) printf(";\n");
      tdev = dt.seconds() / Ntests;
    }
    printf("]; %%Dot [DOFs, elapsed, DOFs/s, BW (GB/s)]\n");
    double result = 0;
    o_out.copyTo(&result, 1);
    printf("BS4device = [%lld %5.4e %6.2f %.17g]; %%Dot kernel-only [DOFs, elapsed, BW (GB/s), result] at the largest size (%s order)\n",
           w.Nmax, tdev, 2.0 * w.Nmax * sizeof(dfloat) / 1e9 / tdev, result, fast ? "fast" : "reference");
  });
}
