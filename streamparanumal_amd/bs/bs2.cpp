// BS2: axpy y = alpha*x + beta*y (BS/BS2/src/bs2Run.cpp:29-94).  The reference
// hard-codes alpha = beta = 1 (:52-53); --alpha/--beta default to those values.
#include "bsCommon.hpp"

using namespace libp;

int main(int argc, char **argv) {
  return bs::run([&] {
    settings_t settings(comm_t::world());
    platformAddSettings(settings);
    bs::addByteSettings(settings);
    bs::addVerbose(settings);
    bs::addSeed(settings);
    settings.newSetting("--nontemporal", "--nontemporal", "NONTEMPORAL", "TRUE",
                        "Streaming (non-temporal) loads and stores", {"TRUE", "FALSE"});
    settings.newSetting("--alpha", "--alpha", "ALPHA", "1.0", "alpha in y = alpha*x + beta*y");
    settings.newSetting("--beta", "--beta", "BETA", "1.0", "beta in y = alpha*x + beta*y");
    settings.parseSettings(argc, argv);
    if (settings.helpRequested()) return;

    platform_t platform(settings);
    if (settings.compareSetting("VERBOSE", "TRUE")) settings.report();
    {
      spx_tuning_t t{};
      t.xcd_remap = -1;
      t.variant = -1;
      t.nontemporal = settings.compareSetting("NONTEMPORAL", "TRUE") ? 1 : 0;
      SPX_CHECK(spx_set_tuning("bs2", &t));
    }
    properties_t kernelInfo = platform.props();
    kernelInfo["defines/p_blockSize"] = 256;
    kernel_t kernel = platform.buildKernel("BS/BS2/okl/bs2.okl", "bs2", kernelInfo);

    const long long sc = 3 * sizeof(dfloat);
    bs::Sweep w = bs::byteSweep(settings, sc);
    unsigned long long seed = 1;
    settings.getSetting("SEED", seed);
    dfloat alpha = 1.0, beta = 1.0;
    settings.getSetting("ALPHA", alpha);
    settings.getSetting("BETA", beta);
    deviceMemory<dfloat> o_a = platform.malloc<dfloat>(w.Nmax);
    deviceMemory<dfloat> o_b = platform.malloc<dfloat>(w.Nmax);
    bs::fillUniform(o_a, seed, 0);
    bs::fillUniform(o_b, seed, w.Nmax);
    for (int n = 0; n < 5; ++n) kernel(w.Nmax, alpha, o_a, beta, o_b);

    if (w.single) printf("BS2 = [");
    else printf("%%AXPY [DOFs, elapsed, DOFs/s, BW (GB/s)]\nBS2 = [\n");
    deviceTimer_t dt;
    double tdev = 0;
    for (long long N = w.Nmin; N <= w.Nmax; N += w.Nstep) {
      const int Ntests = 20;
      timePoint_t start = GlobalPlatformTime(platform);
      dt.start();
      for (int n = 0; n < Ntests; ++n) kernel(N, alpha, o_a, beta, o_b);
      dt.stop();
      timePoint_t end = GlobalPlatformTime(platform);
      const double elapsed = ElapsedTime(start, end) / Ntests;
      const double bytes = 3.0 * N * sizeof(dfloat);
      printf("%lld %5.4e %5.4e %6.2f", N, elapsed, N / elapsed, (bytes / 1.e9) / elapsed);
      if (N < w.Nmax) printf(";\n");
      tdev = dt.seconds() / Ntests;
    }
    printf("]; %%AXPY [DOFs, elapsed, DOFs/s, BW (GB/s)]\n");
    printf("BS2device = [%lld %5.4e %6.2f]; %%AXPY kernel-only [DOFs, elapsed, BW (GB/s)] at the largest size\n",
           w.Nmax, tdev, 3.0 * w.Nmax * sizeof(dfloat) / 1e9 / tdev);
  });
}
