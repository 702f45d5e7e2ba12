// Structured box meshes for BS6-BS8 (libs/mesh/meshSetup.cpp:31-82,
// meshSetElementType.cpp, meshSetupBoxTri2D.cpp, meshSetupBoxQuad2D.cpp,
// meshSetupBoxTet3D.cpp, meshSetupBoxHex3D.cpp:31-131,
// meshConnectNodes.cpp:32-82, meshGatherScatterSetup.cpp:31-44,
// meshSettings.cpp:25-51).
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

int mesh_t::NodesPerElement(int type, int N) {
  switch (type) {
    case TRIANGLES: return (N + 1) * (N + 2) / 2;
    case QUADRILATERALS: return (N + 1) * (N + 1);
    case TETRAHEDRA: return (N + 1) * (N + 2) * (N + 3) / 6;
    case HEXAHEDRA: return (N + 1) * (N + 1) * (N + 1);
    default: return 0;
  }
}

long long mesh_t::BoxElements(int type, dlong nx, dlong ny, dlong nz) {
  switch (type) {
    case TRIANGLES: return 2LL * nx * ny;
    case QUADRILATERALS: return (long long)nx * ny;
    case TETRAHEDRA: return 6LL * nx * ny * nz;
    case HEXAHEDRA: return (long long)nx * ny * nz;
    default: return 0;
  }
}

namespace {

// Lattice offsets of vertices 1..3 from vertex 0 (always the cell's low
// corner) for the two triangles of cell (i,j), e = 2*(i + j*nx) + t
// (meshSetupBoxTri2D.cpp:84-113), and the six tetrahedra of cell (i,j,k),
// e = 6*(i + j*nx + k*nx*ny) + t, whose corners are the cube vertices
// (0,3,2,7) (0,1,3,7) (0,2,6,7) (0,6,4,7) (0,5,1,7) (0,4,5,7), vertex c at
// offset (c&1, c>>1&1, c>>2) (meshSetupBoxTet3D.cpp:88-189).
constexpr int kTriOff[2][2][2] = {{{1, 0}, {1, 1}}, {{1, 1}, {0, 1}}};
constexpr int kTetCorners[6][3] = {{3, 2, 7}, {1, 3, 7}, {2, 6, 7}, {6, 4, 7}, {5, 1, 7}, {4, 5, 7}};

}  // namespace

// Node (m, n[, k]) of a simplex sits at reference coordinates r = -1 + 2m/N,
// s = -1 + 2n/N[, t = -1 + 2k/N] (meshBasisTri2D.cpp:81-92,
// meshBasisTet3D.cpp:81-94; warp and blend moves nodes within their vertex,
// edge, face or interior and keeps this order), so its physical point is
// vertex 0 + m*(v1 - v0)/N + n*(v2 - v0)/N [+ k*(v3 - v0)/N]: an integer
// point of the degree-N lattice.  Quadrilateral node i + j*Nq sits at lattice
// point (ex*N+i, ey*N+j) like the hexahedron's.  On a conforming box the
// reference's trace-matching fixpoint joins exactly the copies of one point
// (meshConnectNodes.cpp:54-81), so the lattice index labels the same groups.
void mesh_t::BoxGlobalIds(int type, dlong nx, dlong ny, dlong nz, int N, hlong *ids) {
  if (type == HEXAHEDRA) {
    HexBoxGlobalIds(nx, ny, nz, N, ids);
    return;
  }
  const int Np = NodesPerElement(type, N);
  const hlong Gx = (hlong)nx * N + 1, Gy = (hlong)ny * N + 1;
  if (type == QUADRILATERALS) {
    const int Nq = N + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (dlong ey = 0; ey < ny; ++ey)
      for (dlong ex = 0; ex < nx; ++ex) {
        hlong *out = ids + (ex + (hlong)ey * nx) * Np;
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i) out[i + j * Nq] = 1 + (hlong)ex * N + i + ((hlong)ey * N + j) * Gx;
      }
  } else if (type == TRIANGLES) {
#pragma omp parallel for collapse(2) schedule(static)
    for (dlong cy = 0; cy < ny; ++cy)
      for (dlong cx = 0; cx < nx; ++cx)
        for (int t = 0; t < 2; ++t) {
          hlong *out = ids + (2 * (cx + (hlong)cy * nx) + t) * Np;
          const int(&o)[2][2] = kTriOff[t];
          int sk = 0;
          for (int n = 0; n <= N; ++n)
            for (int m = 0; m <= N - n; ++m) {
              const hlong gx = (hlong)cx * N + m * o[0][0] + n * o[1][0];
              const hlong gy = (hlong)cy * N + m * o[0][1] + n * o[1][1];
              out[sk++] = 1 + gx + gy * Gx;
            }
        }
  } else if (type == TETRAHEDRA) {
#pragma omp parallel for collapse(3) schedule(static)
    for (dlong cz = 0; cz < nz; ++cz)
      for (dlong cy = 0; cy < ny; ++cy)
        for (dlong cx = 0; cx < nx; ++cx)
          for (int t = 0; t < 6; ++t) {
            hlong *out = ids + (6 * (cx + (hlong)cy * nx + (hlong)cz * nx * ny) + t) * Np;
            int off[3][3];
            for (int v = 0; v < 3; ++v) {
              const int c = kTetCorners[t][v];
              off[v][0] = c & 1;
              off[v][1] = (c >> 1) & 1;
              off[v][2] = c >> 2;
            }
            int sk = 0;
            for (int k = 0; k <= N; ++k)
              for (int n = 0; n <= N - k; ++n)
                for (int m = 0; m <= N - n - k; ++m) {
                  const hlong gx = (hlong)cx * N + m * off[0][0] + n * off[1][0] + k * off[2][0];
                  const hlong gy = (hlong)cy * N + m * off[0][1] + n * off[1][1] + k * off[2][1];
                  const hlong gz = (hlong)cz * N + m * off[0][2] + n * off[1][2] + k * off[2][2];
                  out[sk++] = 1 + gx + gy * Gx + gz * Gx * Gy;
                }
          }
  } else {
    LIBP_ABORT("Unknown element type " << type, true);
  }
}

void mesh_t::SetElementType(const std::string &eType) {
  if (eType == "Tri") {
    elementType = TRIANGLES;
    dim = 2;
    Nverts = 3;
    Nfaces = 3;
  } else if (eType == "Quad") {
    elementType = QUADRILATERALS;
    dim = 2;
    Nverts = 4;
    Nfaces = 4;
  } else if (eType == "Tet") {
    elementType = TETRAHEDRA;
    dim = 3;
    Nverts = 4;
    Nfaces = 4;
  } else if (eType == "Hex") {
    elementType = HEXAHEDRA;
    dim = 3;
    Nverts = 8;
    Nfaces = 6;
  } else {
    LIBP_ABORT("Unknown element type " << eType, true);
  }
}

void mesh_t::Setup(platform_t &_platform, settings_t &_settings, comm_t _comm) {
  platform = _platform;
  settings = _settings;
  comm = _comm;
  rank = comm.rank();
  size = comm.size();
  SetElementType(settings.getSetting("ELEMENT TYPE"));
  settings.getSetting("POLYNOMIAL DEGREE", N);
  Nq = N + 1;
  Np = NodesPerElement(elementType, N);
  switch (elementType) {
    case TRIANGLES:
    case QUADRILATERALS: Nfp = N + 1; break;
    case TETRAHEDRA: Nfp = (N + 1) * (N + 2) / 2; break;
    default: Nfp = Nq * Nq;
  }
  SetupBox();
  GatherScatterSetup();
}

void mesh_t::SetupBox() {
  settings.getSetting("BOX NX", nx);
  settings.getSetting("BOX NY", ny);
  if (dim == 3)
    settings.getSetting("BOX NZ", nz);
  else
    nz = 1;
  LIBP_ABORT("Box dimensions must be positive", nx < 1 || ny < 1 || nz < 1);
  const long long ne = BoxElements(elementType, nx, ny, nz);
  LIBP_ABORT("Mesh of " << ne << " elements x " << Np << " nodes exceeds the int32 local index range",
             ne * Np >= (long long)INT32_MAX);
  Nelements = (dlong)ne;
  const hlong Gx = (hlong)nx * N + 1, Gy = (hlong)ny * N + 1, Gz = (dim == 3) ? (hlong)nz * N + 1 : 1;
  Nnodes = (hlong)(nx + 1) * (ny + 1) * (dim == 3 ? nz + 1 : 1);
  NglobalNodes = Gx * Gy * Gz;
  globalIds.malloc((size_t)Nelements * Np);
  BoxGlobalIds(elementType, nx, ny, nz, N, globalIds.ptr());
}

// A hex box's operator is built in closed form (ogs_t::SetupHexBox): the same
// operator, and the same sign-flagged ids, as ogs.Setup over its ids.  The
// other element types go through ogs.Setup.
void mesh_t::BoxGatherScatterSetup(ogs::ogs_t &ogs, int type, dlong nx, dlong ny, dlong nz, int N,
                                   memory<hlong> ids, comm_t comm, platform_t &platform, ogs::Kind kind,
                                   bool unique) {
  if (type == HEXAHEDRA) {
    ogs.SetupHexBox(nx, ny, nz, N, comm, kind, unique, platform, ids.ptr());
    return;
  }
  ogs.Setup((dlong)ids.length(), ids, comm, kind, ogs::Auto, unique, false, platform);
  // triangles' simplex node order leaves short runs at every degree: their
  // scatter runs in local order (p=7, ~1e8 nodes: 323 -> 304 us), which the
  // row ratio alone does not pick at p=7 (0.68 rows per node)
  if (type == TRIANGLES) ogs.localOperator()->scatterByRows = false;
  // gather+scatter tiles over blocks of the node lattice (the ids are lattice
  // ids, 1 + gx + gy*Gx [+ gz*Gx*Gy]) where that measured faster on MI355X at
  // ~1e8 nodes, us per call: Quad p=3 379 -> 347 and Tri p=3 394 -> 368 with
  // 12 x 12 (p=1, 2 likewise short-run); Tet p=1 513 -> 414 with 4 x 2 x 2,
  // Tet p=3 626 -> 576 with 6 x 6 x 6, and later 514 -> 503 with 3 x 6 x 6
  // (FP32 storage 363 -> 359; 2 x 6 x 6 502, 3 x 9 x 6 505, 6 x 3 x 6 540,
  // 4 x 4 x 4 587).  At 2-D p=7 and Tet p=2 or 7 the CSR's order was as fast
  // or faster.
  int bx = 0, by = 0, bz = 0;
  if ((type == TRIANGLES || type == QUADRILATERALS) && N <= 3) {
    bx = 12; by = 12; bz = 1;
  } else if (type == TETRAHEDRA && N == 1) {
    bx = 4; by = 2; bz = 2;
  } else if (type == TETRAHEDRA && N == 3) {
    bx = 3; by = 6; bz = 6;
  }
  if (bx) {
    auto op = ogs.localOperator();
    const hlong Gx = (hlong)nx * N + 1, Gy = (hlong)ny * N + 1;
    const hlong Gz = type == TETRAHEDRA ? (hlong)nz * N + 1 : 1;
    memory<hlong> grid(op->NrowsT > 0 ? op->NrowsT : 1);
#pragma omp parallel for schedule(static)
    for (dlong r = 0; r < op->NrowsT; ++r) {
      const hlong id = ids[op->colIdsT[op->rowStartsT[r]]];
      grid[r] = (id < 0 ? -id : id) - 1;
    }
    op->SetGsRowOrder(ogs::LatticeBlockOrder(op->NrowsT, grid, Gx, Gy, Gz, bx, by, bz), false);
  }
}

void mesh_t::GatherScatterSetup() {
  const bool verbose = settings.compareSetting("VERBOSE", "TRUE");
  const bool unique = true;  // flag a unique node in every gather node
  BoxGatherScatterSetup(ogs, elementType, nx, ny, nz, N, globalIds, comm, platform, ogs::Signed, unique);
  if (verbose && comm.rank() == 0)
    std::cout << "ogs Setup: " << ogs.NlocalT << " gather nodes; setup " << ogs.setupSeconds() << " seconds."
              << std::endl;
  GlobalToLocal.malloc((size_t)Nelements * Np);
  ogs.SetupGlobalToLocalMapping(GlobalToLocal);
  o_GlobalToLocal = platform.malloc(GlobalToLocal);
}

}  // namespace libp
