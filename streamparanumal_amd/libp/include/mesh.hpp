// mesh_t for the BS6-BS8 drivers: a structured box of hexahedra at degree N
// with GLL-grid global node ids (libs/mesh/meshSetupBoxHex3D.cpp:31-131,
// meshConnectNodes.cpp:32-82, meshGatherScatterSetup.cpp:31-44).  Coordinates
// and geometric factors never enter the streaming benchmarks and are not built.
#pragma once

#include "ogs.hpp"
#include "platform.hpp"
#include "settings.hpp"

namespace libp {

void meshAddSettings(settings_t &settings);
void meshReportSettings(settings_t &settings);

class mesh_t {
 public:
  platform_t platform;
  settings_t settings;
  comm_t comm;

  int rank = 0, size = 1;
  int dim = 3;
  int Nverts = 8, Nfaces = 6;
  int N = 0;       // polynomial degree
  int Nq = 0;      // nodes per direction
  int Np = 0;      // nodes per element
  int Nfp = 0;     // nodes per face
  dlong nx = 0, ny = 0, nz = 0;
  dlong Nelements = 0;
  hlong Nnodes = 0;        // vertices of the box
  hlong NglobalNodes = 0;  // GLL grid nodes

  memory<hlong> globalIds;
  ogs::ogs_t ogs;
  memory<dlong> GlobalToLocal;
  deviceMemory<dlong> o_GlobalToLocal;

  mesh_t() = default;
  mesh_t(platform_t &_platform, settings_t &_settings, comm_t _comm) { Setup(_platform, _settings, _comm); }
  void Setup(platform_t &_platform, settings_t &_settings, comm_t _comm);

  // Lexicographic id of every local node: 1 + gx + gy*Gx + gz*Gx*Gy on the
  // (nx*N+1) x (ny*N+1) x (nz*N+1) GLL grid.
  static void HexBoxGlobalIds(dlong nx, dlong ny, dlong nz, int N, hlong *ids);

 private:
  void SetupBoxHex3D();
  void GatherScatterSetup();
};

}  // namespace libp
