// BS1: copy b = a (BS/BS1/src/bs1Run.cpp:29-91, bs1Setup.cpp:29-40).
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
    settings.parseSettings(argc, argv);
    if (settings.helpRequested()) return;

    platform_t platform(settings);
    if (settings.compareSetting("VERBOSE", "TRUE")) settings.report();
    {
      spx_tuning_t t{};
      t.xcd_remap = -1;
      t.variant = -1;
      t.nontemporal = settings.compareSetting("NONTEMPORAL", "TRUE") ? 1 : 0;
      SPX_CHECK(spx_set_tuning("bs1", &t));
    }
    properties_t kernelInfo = platform.props();
    kernelInfo["defines/p_blockSize"] = 256;
    kernel_t kernel = platform.buildKernel("BS/BS1/okl/bs1.okl", "bs1", kernelInfo);

    const long long sc = 2 * sizeof(dfloat);
    bs::Sweep w = bs::byteSweep(settings, sc);
    unsigned long long seed = 1;
    settings.getSetting("SEED", seed);
    deviceMemory<dfloat> o_a = platform.malloc<dfloat>(w.Nmax);
    deviceMemory<dfloat> o_b = platform.malloc<dfloat>(w.Nmax);
    bs::fillUniform(o_a, seed, 0);
    bs::fillUniform(o_b, seed, w.Nmax);
    for (int n = 0; n < 5; ++n) kernel(w.Nmax, o_a, o_b);

    if (w.single) printf("BS1 = [");
    else printf("%%Copy [DOFs, elapsed, DOFs/s, BW (GB/s)]\nBS1 = [\n");
    deviceTimer_t dt;
    std::vector<double> dev;
    for (long long N = w.Nmin; N <= w.Nmax; N += w.Nstep) {
      const int Ntests = 20;
      timePoint_t start = GlobalPlatformTime(platform);
      dt.start();
      for (int n = 0; n < Ntests; ++n) kernel(N, o_a, o_b);
      dt.stop();
      timePoint_t end = GlobalPlatformTime(platform);
      const double elapsed = ElapsedTime(start, end) / Ntests;
      const double bytes = 2.0 * N * sizeof(dfloat);
      printf("%lld %5.4e %5.4e %6.2f", N, elapsed, N / elapsed, (bytes / 1.e9) / elapsed);
      if (N < w.Nmax) printf(";\n");
      dev.push_back(dt.seconds() / Ntests);
    }
    printf("]; %%Copy [DOFs, elapsed, DOFs/s, BW (GB/s)]\n");
    const double t = dev.back();
    printf("BS1device = [%lld %5.4e %6.2f]; %%Copy kernel-only [DOFs, elapsed, BW (GB/s)] at the largest size\n",
           w.Nmax, t, 2.0 * w.Nmax * sizeof(dfloat) / 1e9 / t);
  });
}
