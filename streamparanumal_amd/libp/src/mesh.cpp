// Structured hex box mesh for BS6-BS8 (libs/mesh/meshSetup.cpp:31-82,
// meshSetupBoxHex3D.cpp:31-131, meshConnectNodes.cpp:32-82,
// meshGatherScatterSetup.cpp:31-44, meshSettings.cpp:25-51).
#include "mesh.hpp"

namespace libp {

void meshAddSettings(settings_t &settings) {
  settings.newSetting("-e", "--elements", "ELEMENT TYPE", "Hex", "Type of mesh elements",
                      {"Tri", "Quad", "Tet", "Hex"});
  settings.newSetting("-nx", "--dimx", "BOX NX", "50", "Number of elements in X-dimension per rank");
  settings.newSetting("-ny", "--dimy", "BOX NY", "50", "Number of elements in Y-dimension per rank");
  settings.newSetting("-nz", "--dimz", "BOX NZ", "50", "Number of elements in Z-dimension per rank");
  settings.newSetting("-p", "--degree", "POLYNOMIAL DEGREE", "3", "Degree of polynomial finite element space",
                      {"1", "2", "3", "4", "5", "6", "7", "8", "9", "10", "11", "12", "13", "14", "15"});
}

void meshReportSettings(settings_t &settings) {
  settings.reportSetting("ELEMENT TYPE");
  settings.reportSetting("BOX NX");
  settings.reportSetting("BOX NY");
  if (settings.compareSetting("ELEMENT TYPE", "Tet") || settings.compareSetting("ELEMENT TYPE", "Hex"))
    settings.reportSetting("BOX NZ");
  settings.reportSetting("POLYNOMIAL DEGREE");
}

// Local index l = e*Np + n, e = ex + ey*nx + ez*nx*ny (meshSetupBoxHex3D.cpp:92),
// n = i + j*Nq + k*Nq^2 (meshBasisHex3D.cpp:44); the node sits at GLL grid
// point (ex*N+i, ey*N+j, ez*N+k) of a non-periodic box.  The reference labels
// copies of a grid point with the smallest 1+l among them
// (meshConnectNodes.cpp:43-81); any labelling that groups the same copies
// yields the same ogs operator, and the lexicographic one is closed-form.
void mesh_t::HexBoxGlobalIds(dlong nx, dlong ny, dlong nz, int N, hlong *ids) {
  const int Nq = N + 1;
  const hlong Np = (hlong)Nq * Nq * Nq;
  const hlong Gx = (hlong)nx * N + 1, Gy = (hlong)ny * N + 1;
#pragma omp parallel for collapse(3) schedule(static)
  for (dlong ez = 0; ez < nz; ++ez)
    for (dlong ey = 0; ey < ny; ++ey)
      for (dlong ex = 0; ex < nx; ++ex) {
        const hlong e = ex + (hlong)ey * nx + (hlong)ez * nx * ny;
        hlong *out = ids + e * Np;
        for (int k = 0; k < Nq; ++k)
          for (int j = 0; j < Nq; ++j)
            for (int i = 0; i < Nq; ++i) {
              const hlong gx = (hlong)ex * N + i, gy = (hlong)ey * N + j, gz = (hlong)ez * N + k;
              out[i + j * Nq + k * Nq * Nq] = 1 + gx + gy * Gx + gz * Gx * Gy;
            }
      }
}

void mesh_t::Setup(platform_t &_platform, settings_t &_settings, comm_t _comm) {
  platform = _platform;
  settings = _settings;
  comm = _comm;
  rank = comm.rank();
  size = comm.size();
  const std::string et = settings.getSetting("ELEMENT TYPE");
  LIBP_ABORT("Element type " << et << " is not provided: the streaming benchmarks run on hexahedra (-e Hex)",
             et != "Hex");
  settings.getSetting("POLYNOMIAL DEGREE", N);
  Nq = N + 1;
  Np = Nq * Nq * Nq;
  Nfp = Nq * Nq;
  SetupBoxHex3D();
  GatherScatterSetup();
}

void mesh_t::SetupBoxHex3D() {
  settings.getSetting("BOX NX", nx);
  settings.getSetting("BOX NY", ny);
  settings.getSetting("BOX NZ", nz);
  LIBP_ABORT("Box dimensions must be positive", nx < 1 || ny < 1 || nz < 1);
  const long long ne = (long long)nx * ny * nz;
  LIBP_ABORT("Mesh of " << ne << " elements x " << Np << " nodes exceeds the int32 local index range",
             ne * Np >= (long long)INT32_MAX);
  Nelements = (dlong)ne;
  Nnodes = (hlong)(nx + 1) * (ny + 1) * (nz + 1);
  NglobalNodes = ((hlong)nx * N + 1) * ((hlong)ny * N + 1) * ((hlong)nz * N + 1);
  globalIds.malloc((size_t)Nelements * Np);
  HexBoxGlobalIds(nx, ny, nz, N, globalIds.ptr());
}

// The box's operator is built in closed form (ogs_t::SetupHexBox): the same
// operator, and the same sign-flagged ids, as ogs.Setup over globalIds.
void mesh_t::GatherScatterSetup() {
  const bool verbose = settings.compareSetting("VERBOSE", "TRUE");
  const bool unique = true;  // flag a unique node in every gather node
  ogs.SetupHexBox(nx, ny, nz, N, comm, ogs::Signed, unique, platform, globalIds.ptr());
  if (verbose && comm.rank() == 0)
    std::cout << "ogs Setup: " << ogs.NlocalT << " gather nodes; setup " << ogs.setupSeconds() << " seconds."
              << std::endl;
  GlobalToLocal.malloc((size_t)Nelements * Np);
  ogs.SetupGlobalToLocalMapping(GlobalToLocal);
  o_GlobalToLocal = platform.malloc(GlobalToLocal);
}

}  // namespace libp
