// BS7: scatter, ogs.Scatter(o_q, o_gq, K, NoTrans) (BS/BS7/src/bs7Run.cpp:29-79, bs7Main.cpp:29-57).
// Adds --seed (inputs from the reference LCG), --fields K (the reference runs
// K = 1) and --block-nnz (LDS tile of the row blocks, the autotuner's knob).
#include "bsCommon.hpp"

using namespace libp;

int main(int argc, char **argv) {
  return bs::run([&] {
    settings_t settings(comm_t::world());
    platformAddSettings(settings);
    meshAddSettings(settings);
    bs::addVerbose(settings);
    bs::addSettle(settings);
    bs::addSeed(settings);
    settings.newSetting("--fields", "--fields", "FIELDS", "1", "Fields per node (K)");
    settings.newSetting("--block-nnz", "--block-nnz", "BLOCK NNZ", "0", "Max column entries per row block (0: library default)");
    settings.parseSettings(argc, argv);
    if (settings.helpRequested()) return;

    platform_t platform(settings);
    if (settings.compareSetting("VERBOSE", "TRUE")) {
      std::cout << "Settings:\n\n";
      platformReportSettings(settings);
      meshReportSettings(settings);
    }
    mesh_t mesh(platform, settings, settings.comm);
    int K = 1, blockNnz = 0;
    unsigned long long seed = 1;
    settings.getSetting("FIELDS", K);
    settings.getSetting("BLOCK NNZ", blockNnz);
    settings.getSetting("SEED", seed);
    LIBP_ABORT("--fields must be >= 1", K < 1);
    if (blockNnz > 0) {
      mesh.ogs.localOperator()->SetBlockNnz(blockNnz);
      mesh.ogs.localOperator()->SetGsBlockNnz(blockNnz);
    }

    const dlong N = mesh.Np * mesh.Nelements;
    const dlong Ngather = mesh.ogs.Ngather;
    deviceMemory<dfloat> o_q = platform.malloc<dfloat>((size_t)N * K);
    deviceMemory<dfloat> o_gq = platform.malloc<dfloat>((size_t)Ngather * K);
    bs::fillUniform(o_gq, seed, 0);
    for (int n = 0; n < 5; ++n) { mesh.ogs.Scatter(o_q, o_gq, K, ogs::NoTrans); }
    bs::settle(settings, platform, [&] { mesh.ogs.Scatter(o_q, o_gq, K, ogs::NoTrans); });

    const int Ntests = 20;
    deviceTimer_t dt;
    timePoint_t start = GlobalPlatformTime(platform);
    dt.start();
    for (int n = 0; n < Ntests; ++n) { mesh.ogs.Scatter(o_q, o_gq, K, ogs::NoTrans); }
    dt.stop();
    timePoint_t end = GlobalPlatformTime(platform);
    const double elapsedTime = ElapsedTime(start, end) / Ntests;
    const double kernelTime = dt.seconds() / Ntests;
    SPX_CHECK(spx_check_kernel_errors());
    const double syncedTime = bs::timeSynced(platform, Ntests, [&] { mesh.ogs.Scatter(o_q, o_gq, K, ogs::NoTrans); });

    hlong NtotalGlobal = (hlong)mesh.Nelements * mesh.Np;
    mesh.comm.Allreduce(NtotalGlobal);
    const hlong NgatherGlobal = mesh.ogs.NgatherGlobal;
    size_t bytesIn = 0, bytesOut = 0;
    bytesIn += (NgatherGlobal + 1) * sizeof(dlong);  // row starts
    bytesIn += NtotalGlobal * sizeof(dlong);         // local ids
    bytesIn += (size_t)K * NgatherGlobal * sizeof(dfloat);
    bytesOut += (size_t)K * NtotalGlobal * sizeof(dfloat);
    const size_t bytes = bytesIn + bytesOut;
    const hlong Ndofs = NgatherGlobal;
    const double Nflops = 0.0;
    if (mesh.rank == 0) {
      printf("BS7 = [%d, " hlongFormat ", %5.4le, %5.4le, %6.2f, %6.2f]; %% Scatter [N, DOFs, elapsed, DOFs/(ranks*s), avg BW (GB/s), avg GFLOPs] \n",
             mesh.N, Ndofs, elapsedTime, Ndofs / (mesh.size * elapsedTime), bytes / (1.0e9 * elapsedTime),
             Nflops / (1.0e9 * elapsedTime));
      printf("BS7device = [%d, " hlongFormat ", %5.4le, %5.4le, %6.2f]; %% Scatter kernel-only [N, DOFs, elapsed, DOFs/s, BW (GB/s)], K=%d, block nnz %d\n",
             mesh.N, Ndofs, kernelTime, Ndofs / kernelTime, bytes / (1.0e9 * kernelTime), K,
             (int)mesh.ogs.localOperator()->blockNnz);
      bs::printSynced("BS7", "Scatter", mesh.N, Ndofs, syncedTime, bytes);
    }
  });
}
