// BS8: gather+scatter, ogs.GatherScatter(o_q, K, Add, Sym) (BS/BS8/src/bs8Run.cpp:29-78, bs8Main.cpp:29-57).
// Adds --seed (inputs from the reference LCG), --fields K (the reference runs
// K = 1), --block-nnz (LDS tile of the row blocks, the autotuner's knob) and
// --storage FP32 / BF16 (values stored in that type, each row summed in FP64
// and rounded once; bytes count the stored values).
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
    bs::addStorage(settings);
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
    const int storage = bs::storageOf(settings);
    const size_t sb = bs::storageBytes(storage);
    deviceMemory<dfloat> o_q;
    deviceMemory<unsigned char> s_q;
    if (storage == SPX_STORE_F64) {
      o_q = platform.malloc<dfloat>((size_t)N * K);
      bs::fillUniform(o_q, seed, 0);
    } else {
      s_q = bs::stored(platform, storage, (size_t)N * K, seed, 0);
    }
    auto step = [&]() {
      if (storage == SPX_STORE_F64)
        mesh.ogs.GatherScatter(o_q, K, ogs::Add, ogs::Sym);
      else
        mesh.ogs.localOperator()->GatherScatterMixedRaw(storage, K, s_q.ptr());
    };
    for (int n = 0; n < 5; ++n) step();
    bs::settle(settings, platform, step);

    const int Ntests = 50;
    deviceTimer_t dt;
    timePoint_t start = GlobalPlatformTime(platform);
    dt.start();
    for (int n = 0; n < Ntests; ++n) step();
    dt.stop();
    timePoint_t end = GlobalPlatformTime(platform);
    const double elapsedTime = ElapsedTime(start, end) / Ntests;
    const double kernelTime = dt.seconds() / Ntests;
    SPX_CHECK(spx_check_kernel_errors());
    const double syncedTime = bs::timeSynced(platform, Ntests, step);

    hlong NtotalGlobal = (hlong)mesh.Nelements * mesh.Np;
    mesh.comm.Allreduce(NtotalGlobal);
    const hlong NgatherGlobal = mesh.ogs.NgatherGlobal;
    size_t bytesIn = 0, bytesOut = 0;
    bytesIn += (NgatherGlobal + 1) * sizeof(dlong);  // row starts
    bytesIn += NtotalGlobal * sizeof(dlong);         // local ids
    bytesIn += (size_t)K * NtotalGlobal * sb;
    bytesOut += (size_t)K * NtotalGlobal * sb;
    const size_t bytes = bytesIn + bytesOut;
    const hlong Ndofs = NgatherGlobal;
    const double Nflops = (double)K * NtotalGlobal;
    if (mesh.rank == 0) {
      printf("BS8 = [%d, " hlongFormat ", %5.4le, %5.4le, %6.2f, %6.2f]; %% GatherScatter [N, DOFs, elapsed, DOFs/(ranks*s), avg BW (GB/s), avg GFLOPs] \n",
             mesh.N, Ndofs, elapsedTime, Ndofs / (mesh.size * elapsedTime), bytes / (1.0e9 * elapsedTime),
             Nflops / (1.0e9 * elapsedTime));
      printf("BS8device = [%d, " hlongFormat ", %5.4le, %5.4le, %6.2f]; %% GatherScatter kernel-only [N, DOFs, elapsed, DOFs/s, BW (GB/s)], K=%d, block nnz %d, %s storage\n",
             mesh.N, Ndofs, kernelTime, Ndofs / kernelTime, bytes / (1.0e9 * kernelTime), K,
             (int)mesh.ogs.localOperator()->gsBlockNnz, settings.getSetting("STORAGE").c_str());
      bs::printSynced("BS8", "GatherScatter", mesh.N, Ndofs, syncedTime, bytes);
    }
  });
}
