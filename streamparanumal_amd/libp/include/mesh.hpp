// mesh_t for the BS6-BS8 drivers: a structured box of triangles,
// quadrilaterals, tetrahedra or hexahedra at degree N with lattice global node
// ids (libs/mesh/meshSetupBoxTri2D.cpp, meshSetupBoxQuad2D.cpp,
// meshSetupBoxTet3D.cpp, meshSetupBoxHex3D.cpp, meshConnectNodes.cpp:32-82,
// meshGatherScatterSetup.cpp:31-44).  Coordinates and geometric factors never
// enter the streaming benchmarks and are not built.
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

  // element types (include/mesh.hpp:116-119)
  static constexpr int TRIANGLES = 3;
  static constexpr int QUADRILATERALS = 4;
  static constexpr int TETRAHEDRA = 6;
  static constexpr int HEXAHEDRA = 12;

  int rank = 0, size = 1;
  int elementType = HEXAHEDRA;
  int dim = 3;
  int Nverts = 8, Nfaces = 6;
  int N = 0;       // polynomial degree
  int Nq = 0;      // nodes per direction
  int Np = 0;      // nodes per element
  int Nfp = 0;     // nodes per face
  dlong nx = 0, ny = 0, nz = 0;
  dlong Nelements = 0;
  hlong Nnodes = 0;        // vertices of the box
  hlong NglobalNodes = 0;  // nodes of the box's degree-N lattice

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

  // Nodes per element and elements per box of an element type (Np = 0 for an
  // unknown type); 2-D types ignore nz.
  static int NodesPerElement(int type, int N);
  static long long BoxElements(int type, dlong nx, dlong ny, dlong nz);
  // 1 + the lexicographic index of every local node's point on the box's
  // degree-N lattice (nx*N+1) x (ny*N+1) [x (nz*N+1)], for any element type.
  static void BoxGlobalIds(int type, dlong nx, dlong ny, dlong nz, int N, hlong *ids);
  // ogs_t of a box (meshGatherScatterSetup.cpp:31-44): ids hold BoxGlobalIds
  // (filled here for hexahedra) and receive the sign flags.  Hex boxes, Tri /
  // Quad boxes up to degree 3 and Tet boxes of degree 1 and 3 order their
  // gather+scatter tiles by blocks of lattice points.
  static void BoxGatherScatterSetup(ogs::ogs_t &ogs, int type, dlong nx, dlong ny, dlong nz, int N,
                                    memory<hlong> ids, comm_t comm, platform_t &platform, ogs::Kind kind,
                                    bool unique);

 private:
  void SetElementType(const std::string &eType);
  void SetupBox();
  void GatherScatterSetup();
};

}  // namespace libp
