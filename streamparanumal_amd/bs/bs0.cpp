// BS0: kernel launch latency (BS/BS0/src/bs0Run.cpp:30-82, bs0Settings.cpp:35-59).
// Adds --graph TRUE: the same launches captured once as a HIP graph and
// replayed, so plain launch cost and graph replay cost can be compared.
#include "bsCommon.hpp"

using namespace libp;

int main(int argc, char **argv) {
  return bs::run([&] {
    settings_t settings(comm_t::world());
    platformAddSettings(settings);
    settings.newSetting("-n", "--Nlaunch", "NLAUNCHES", "0", "Number of small kernels to launch");
    settings.newSetting("-nmin", "--NlaunchMin", "NLAUNCHMIN", "0", "Minimum number of small kernels to launch");
    settings.newSetting("-nmax", "--NlaunchMax", "NLAUNCHMAX", "0", "Maximum number of small kernels to launch");
    settings.newSetting("-s", "--step", "NLAUNCHSTEP", "1", "Step size of small kernels to launch");
    bs::addVerbose(settings);
    settings.newSetting("--graph", "--graph", "GRAPH", "FALSE",
                        "Replay the launches as one captured HIP graph", {"TRUE", "FALSE"});
    settings.parseSettings(argc, argv);
    if (settings.helpRequested()) return;

    platform_t platform(settings);
    if (settings.compareSetting("VERBOSE", "TRUE")) settings.report();

    properties_t kernelInfo = platform.props();
    kernel_t kernel = platform.buildKernel("BS/BS0/okl/bs0.okl", "bs0", kernelInfo);

    int Nlaunches = 10, NlaunchMin = 1, NlaunchMax = 1024, NlaunchStep = 1;
    settings.getSetting("NLAUNCHES", Nlaunches);
    settings.getSetting("NLAUNCHMIN", NlaunchMin);
    settings.getSetting("NLAUNCHMAX", NlaunchMax);
    settings.getSetting("NLAUNCHSTEP", NlaunchStep);
    if (!(Nlaunches | NlaunchMin | NlaunchMax)) Nlaunches = 10;
    const bool graph = settings.compareSetting("GRAPH", "TRUE");

    const int N = 1;
    deviceMemory<dfloat> o_a = platform.malloc<dfloat>(N);
    deviceMemory<dfloat> o_b = platform.malloc<dfloat>(N);
    for (int w = 0; w < 100; ++w) kernel(N, o_a, o_b);

    if (Nlaunches) {
      NlaunchMin = NlaunchMax = Nlaunches;
      printf("BS0 = [");
    } else {
      printf("%%Launch Latency [Nlaunches, time per launch, elapsed]\nBS0 = [\n");
    }
    if (NlaunchStep < 1) NlaunchStep = 1;
    for (Nlaunches = NlaunchMin; Nlaunches <= NlaunchMax; Nlaunches += NlaunchStep) {
      double elapsed;
      if (graph) {
        spx_graph_t g;
        SPX_CHECK(spx_graph_begin());
        for (int n = 0; n < Nlaunches; ++n) kernel(N, o_a, o_b);
        SPX_CHECK(spx_graph_end(&g));
        SPX_CHECK(spx_graph_launch(g));  // first replay uploads the graph
        timePoint_t start = PlatformTime(platform);
        SPX_CHECK(spx_graph_launch(g));
        timePoint_t end = PlatformTime(platform);
        elapsed = ElapsedTime(start, end);
        SPX_CHECK(spx_graph_destroy(g));
      } else {
        timePoint_t start = PlatformTime(platform);
        for (int n = 0; n < Nlaunches; ++n) kernel(N, o_a, o_b);
        timePoint_t end = PlatformTime(platform);
        elapsed = ElapsedTime(start, end);
      }
      printf("%d %5.4e %5.4e", Nlaunches, elapsed / Nlaunches, elapsed);
      if (Nlaunches < NlaunchMax) printf(";\n");
    }
    printf("]; %%Launch Latency [Nlaunches, time per launch, elapsed]%s\n", graph ? " (graph replay)" : "");
  });
}
