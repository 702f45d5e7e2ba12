// ogs setup on one rank (libs/ogs/ogsSetup.cpp:87-290, 492-579).
//
// Produces the reference's operator exactly:
//   * zero ids are ignored (:121-139);
//   * members are grouped by |id|;
//   * rows are numbered in order of each group's first appearance in local
//     index order, groups holding a positive id first (:492-559);
//   * a row lists its members in increasing local index (stable SortByRow,
//     :561-579);
//   * with unique=true every group keeps one positive member, picked by the
//     reference LCG seeded with the rank (:111-113, 361-382).
// The reference reaches this with stable sorts; here the numbering is one
// linear pass over the ids with a dense table when the ids are dense (a
// structured mesh's GLL numbering is), else a sort.  The member kept positive
// under unique=true is the LCG draw of its group (groups in increasing |id|)
// taken modulo the group size over members in increasing local index; the
// reference takes the same draw over the order an unstable std::sort leaves
// (ogsSetup.cpp:347), which the C++ standard does not fix.
#include <algorithm>
#include <chrono>
#include <limits>
#include <cstdlib>
#include <numeric>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "ogs.hpp"

namespace libp {
namespace ogs {

namespace {

// The reference LCG (libs/primitives/random.cpp:45-121): x0 = A*(~0 ^ seed) + C,
// draw = abs((int)x).
void lcgInts(int64_t n, uint64_t seed, std::vector<int32_t> &out) {
  const uint64_t A = 6364136223846793005ULL, C = 1ULL;
  uint64_t x = A * (UINT64_MAX ^ seed) + C;
  out.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    uint32_t u = (uint32_t)x;
    if (u & 0x80000000u) u = 0u - u;
    out[i] = (int32_t)u;
    x = A * x + C;
  }
}

}  // namespace


LocalCSR BuildLocalCSR(const dlong N, memory<hlong> ids, const Kind kind, const bool unique, const int rank) {
  LIBP_ABORT("Invalid ogs setup requested", (kind == Unsigned && unique) || (kind == Halo && unique));
  LIBP_ABORT("ogs::Halo setups describe exchanges between ranks; this build runs one rank per GPU "
             "and provides no halo_t", kind == Halo);
  LIBP_ABORT("ogs Setup: negative N", N < 0);
  LIBP_ABORT("ogs Setup: ids hold " << ids.length() << " entries, N = " << N, (size_t)N > ids.length());
  LocalCSR out;
  // ---- nonzero positions and |id| range
  hlong maxAbs = 0;
  dlong Nids = 0;
#pragma omp parallel for reduction(max : maxAbs) reduction(+ : Nids)
  for (dlong i = 0; i < N; ++i) {
    const hlong a = std::llabs(ids[i]);
    if (a) {
      ++Nids;
      if (a > maxAbs) maxAbs = a;
    }
  }
  std::vector<dlong> pos;  // local index of the k-th nonzero id
  const bool allNonzero = (Nids == N);
  if (!allNonzero) {
    pos.reserve(Nids);
    for (dlong i = 0; i < N; ++i)
      if (ids[i]) pos.push_back(i);
  }
  auto posOf = [&](dlong k) { return allNonzero ? k : pos[k]; };

  // ---- groups in first-appearance order; rank of each group in increasing |id|
  std::vector<dlong> grp(Nids);
  std::vector<dlong> rankOf;  // only needed for unique
  dlong ngroups = 0;
  const bool dense = maxAbs < (hlong)INT32_MAX && maxAbs <= 4 * (hlong)Nids + (1 << 20);
  if (dense) {
    // first[a] = smallest k with |id| = a (atomic minimum), then groups are
    // numbered by a prefix count of the first occurrences in k order
    std::vector<dlong> first(maxAbs + 1, std::numeric_limits<dlong>::max());
#pragma omp parallel for schedule(static)
    for (dlong k = 0; k < Nids; ++k) {
      const hlong a = std::llabs(ids[posOf(k)]);
      dlong cur = __atomic_load_n(&first[a], __ATOMIC_RELAXED);
      while (k < cur && !__atomic_compare_exchange_n(&first[a], &cur, k, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
      }
    }
    std::vector<dlong> gnum(Nids);
#pragma omp parallel for schedule(static)
    for (dlong k = 0; k < Nids; ++k) gnum[k] = (first[std::llabs(ids[posOf(k)])] == k) ? 1 : 0;
    ngroups = ExclusiveScan(gnum.data(), Nids);
#pragma omp parallel for schedule(static)
    for (dlong k = 0; k < Nids; ++k) grp[k] = gnum[first[std::llabs(ids[posOf(k)])]];
    if (unique) {
      rankOf.assign(ngroups, 0);
      std::vector<dlong> present(maxAbs + 1);
#pragma omp parallel for schedule(static)
      for (hlong a = 0; a <= maxAbs; ++a) present[a] = (a > 0 && first[a] != std::numeric_limits<dlong>::max()) ? 1 : 0;
      ExclusiveScan(present.data(), (dlong)(maxAbs + 1));
#pragma omp parallel for schedule(static)
      for (hlong a = 1; a <= maxAbs; ++a)
        if (first[a] != std::numeric_limits<dlong>::max()) rankOf[gnum[first[a]]] = present[a];
    }
  } else {
    std::vector<std::pair<hlong, dlong>> kv(Nids);
#pragma omp parallel for schedule(static)
    for (dlong k = 0; k < Nids; ++k) kv[k] = {std::llabs(ids[posOf(k)]), k};
    std::sort(kv.begin(), kv.end());
    // groups in increasing |id|; first member (smallest k) opens the group
    std::vector<dlong> groupFirst;  // by |id| rank
    std::vector<dlong> sortedGroupOf(Nids);
    for (dlong i = 0; i < Nids; ++i) {
      if (i == 0 || kv[i].first != kv[i - 1].first) groupFirst.push_back(kv[i].second);
      sortedGroupOf[kv[i].second] = (dlong)groupFirst.size() - 1;
    }
    ngroups = (dlong)groupFirst.size();
    std::vector<dlong> order(ngroups);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](dlong x, dlong y) { return groupFirst[x] < groupFirst[y]; });
    std::vector<dlong> fa(ngroups);  // |id| rank -> first-appearance index
    for (dlong j = 0; j < ngroups; ++j) fa[order[j]] = j;
#pragma omp parallel for schedule(static)
    for (dlong k = 0; k < Nids; ++k) grp[k] = fa[sortedGroupOf[k]];
    if (unique) {
      rankOf.assign(ngroups, 0);
      for (dlong s = 0; s < ngroups; ++s) rankOf[fa[s]] = s;
    }
  }

  // ---- members of each group, in increasing local index: atomic placement,
  // then each group's few members sorted
  std::vector<dlong> gStart(ngroups + 1, 0);
#pragma omp parallel for schedule(static)
  for (dlong k = 0; k < Nids; ++k) __atomic_fetch_add(&gStart[grp[k]], (dlong)1, __ATOMIC_RELAXED);
  ExclusiveScan(gStart.data(), ngroups + 1);
  std::vector<dlong> members(Nids);
  {
    std::vector<dlong> cur(gStart.begin(), gStart.end() - 1);
#pragma omp parallel for schedule(static)
    for (dlong k = 0; k < Nids; ++k) members[__atomic_fetch_add(&cur[grp[k]], (dlong)1, __ATOMIC_RELAXED)] = posOf(k);
#pragma omp parallel for schedule(dynamic, 4096)
    for (dlong g = 0; g < ngroups; ++g) std::sort(members.begin() + gStart[g], members.begin() + gStart[g + 1]);
  }

  // ---- signs: unique flagging or the caller's flags
  std::vector<char> hasPos(ngroups, 0);
  bool gather_defined = true;
  if (unique) {
    std::vector<int32_t> rands;
    lcgInts(ngroups, (uint64_t)rank, rands);
#pragma omp parallel for schedule(static)
    for (dlong g = 0; g < ngroups; ++g) {
      const dlong a = gStart[g], e = gStart[g + 1];
      const dlong m = (dlong)((uint32_t)rands[rankOf[g]] % (uint32_t)(e - a));
      for (dlong i = a; i < e; ++i) {
        const hlong b = std::llabs(ids[members[i]]);
        ids[members[i]] = (i - a == m) ? b : -b;
      }
      hasPos[g] = 1;
    }
  } else {
    int undefined = 0;
#pragma omp parallel for schedule(static) reduction(| : undefined)
    for (dlong g = 0; g < ngroups; ++g) {
      int npos = 0;
      for (dlong i = gStart[g]; i < gStart[g + 1]; ++i) npos += (ids[members[i]] > 0);
      hasPos[g] = npos > 0;
      if (npos != 1) undefined = 1;
    }
    gather_defined = !undefined;
  }

  // ---- rows: positive groups first, each class in first-appearance order
  std::vector<dlong> rowOf(ngroups);
  dlong NP = 0;
  {
    std::vector<dlong> posRank(ngroups), negRank(ngroups);
#pragma omp parallel for schedule(static)
    for (dlong g = 0; g < ngroups; ++g) {
      posRank[g] = hasPos[g] ? 1 : 0;
      negRank[g] = hasPos[g] ? 0 : 1;
    }
    NP = ExclusiveScan(posRank.data(), ngroups);
    ExclusiveScan(negRank.data(), ngroups);
#pragma omp parallel for schedule(static)
    for (dlong g = 0; g < ngroups; ++g) rowOf[g] = hasPos[g] ? posRank[g] : NP + negRank[g];
  }

  // ---- T-CSR (every member) and N-CSR (positive members)
  memory<dlong> rowStartsT(ngroups + 1), colIdsT(Nids);
  {
#pragma omp parallel for schedule(static)
    for (dlong g = 0; g < ngroups; ++g) rowStartsT[rowOf[g]] = gStart[g + 1] - gStart[g];
    rowStartsT[ngroups] = 0;
    ExclusiveScan(rowStartsT.ptr(), ngroups + 1);
#pragma omp parallel for schedule(static)
    for (dlong g = 0; g < ngroups; ++g)
      std::copy(members.begin() + gStart[g], members.begin() + gStart[g + 1], colIdsT.ptr() + rowStartsT[rowOf[g]]);
  }
  memory<dlong> rowStartsN, colIdsN;
  dlong nnzN = 0;
  if (kind == Signed) {
    rowStartsN.malloc(NP + 1);
#pragma omp parallel for schedule(static)
    for (dlong r = 0; r < NP; ++r) {
      dlong c = 0;
      for (dlong i = rowStartsT[r]; i < rowStartsT[r + 1]; ++i) c += (ids[colIdsT[i]] > 0);
      rowStartsN[r] = c;
    }
    rowStartsN[NP] = 0;
    nnzN = ExclusiveScan(rowStartsN.ptr(), NP + 1);
    colIdsN.malloc(nnzN);
#pragma omp parallel for schedule(static)
    for (dlong r = 0; r < NP; ++r) {
      dlong c = rowStartsN[r];
      for (dlong i = rowStartsT[r]; i < rowStartsT[r + 1]; ++i)
        if (ids[colIdsT[i]] > 0) colIdsN[c++] = colIdsT[i];
    }
  } else {
    rowStartsN = rowStartsT;
    colIdsN = colIdsT;
    nnzN = Nids;
  }

  out.NrowsT = ngroups;
  out.nnzT = Nids;
  out.nnzN = nnzN;
  out.NrowsN = NP;
  out.rowStartsT = rowStartsT;
  out.colIdsT = colIdsT;
  out.rowStartsN = rowStartsN;
  out.colIdsN = colIdsN;
  out.gather_defined = gather_defined;
  return out;
}

void ogsBase_t::Setup(const dlong _N, memory<hlong> ids, comm_t _comm, const Kind _kind, const Method method,
                      const bool _unique, const bool verbose, platform_t &_platform) {
  (void)method;  // one rank: no exchange to choose
  Free();
  const auto t0 = std::chrono::steady_clock::now();

  platform = _platform;
  comm = _comm;
  N = _N;
  kind = _kind;
  unique = _unique;

  LocalCSR csr = BuildLocalCSR(N, ids, kind, unique, comm.rank());
  SetupFromCSR(csr);
  setupSeconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (verbose && comm.rank() == 0)
    std::cout << "ogs Setup: " << NlocalT << " gather nodes, " << gatherLocal->nnzT << " entries; setup "
              << setupSeconds_ << " seconds." << std::endl;
}

void ogsBase_t::SetupHexBox(const dlong nx, const dlong ny, const dlong nz, const int p, comm_t _comm,
                            const Kind _kind, const bool _unique, platform_t &_platform, hlong *idsOut) {
  Free();
  const auto t0 = std::chrono::steady_clock::now();
  platform = _platform;
  comm = _comm;
  kind = _kind;
  unique = _unique;
  LocalCSR csr = BuildHexBoxCSR(nx, ny, nz, p, kind, unique, comm.rank(), idsOut);
  N = csr.nnzT;
  SetupFromCSR(csr);
  // gather+scatter tiles over line tiles (HexGsLineRows) or small 3-D blocks
  // of grid points: a row's grid point from its first member (element e, node
  // a + b*Nq + c*Nq^2, elements x fastest, meshSetupBoxHex3D.cpp:87-130)
  {
    const ogsOperator_t &op = *gatherLocal;
    const hlong Nq = p + 1, Np = Nq * Nq * Nq;
    const hlong Mx = (hlong)nx * p + 1, My = (hlong)ny * p + 1, Mz = (hlong)nz * p + 1;
    memory<hlong> grid(op.NrowsT > 0 ? op.NrowsT : 1);
#pragma omp parallel for schedule(static)
    for (dlong r = 0; r < op.NrowsT; ++r) {
      const hlong l = op.colIdsT[op.rowStartsT[r]], e = l / Np, n = l % Np;
      const hlong ex = e % nx, ey = (e / nx) % ny, ez = e / ((hlong)nx * ny);
      grid[r] = (ex * p + n % Nq) + (ey * p + (n / Nq) % Nq) * Mx + (ez * p + n / (Nq * Nq)) * Mx * My;
    }
    const int rowsPerLine = HexGsLineRows(p);
    if (rowsPerLine > 0) {
      memory<dlong> order;
      memory<uint8_t> cuts;
      HexLineTiles(op.NrowsT, grid, nx, ny, nz, p, rowsPerLine, (int)op.gsBlockNnz, order, cuts);
      gatherLocal->SetGsRowOrder(order, false, cuts);
    } else {
      int bx, by, bz;
      HexGsBlockShape(p, bx, by, bz);
      gatherLocal->SetGsRowOrder(LatticeBlockOrder(op.NrowsT, grid, Mx, My, Mz, bx, by, bz), false);
    }
    // FP32 / BF16 tiles of their own where line tiles of other heights, or
    // the lattice blocks (r < 0), were faster for them than the FP64 tiles
    // (one layout where both storages take the same)
    memory<dlong> mixedOrder;
    memory<uint8_t> mixedCuts;
    int mixedRows = 0;
    for (int s : {SPX_STORE_F32, SPX_STORE_BF16}) {
      const int r = HexGsLineRowsMixed(p, s);
      if (r == 0 || r == rowsPerLine || (r < 0 && rowsPerLine <= 0)) continue;
      if (r != mixedRows) {
        if (r > 0) {
          HexLineTiles(op.NrowsT, grid, nx, ny, nz, p, r, (int)op.gsBlockNnz, mixedOrder, mixedCuts);
        } else {
          int bx, by, bz;
          HexGsBlockShape(p, bx, by, bz);
          mixedOrder = LatticeBlockOrder(op.NrowsT, grid, Mx, My, Mz, bx, by, bz);
          mixedCuts = memory<uint8_t>();
        }
        mixedRows = r;
      }
      gatherLocal->SetGsMixedRowOrder(s, mixedOrder, mixedCuts, false);
    }
  }
  setupSeconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void ogsBase_t::SetupFromCSR(LocalCSR &csr) {
  gather_defined = csr.gather_defined;

  NlocalT = csr.NrowsT;
  NlocalP = csr.NrowsN;
  NhaloT = 0;
  NhaloP = 0;
  Ngather = NlocalP + NhaloP;
  NgatherGlobal = Ngather;
  comm.Allreduce(NgatherGlobal);

  gatherLocal = std::make_shared<ogsOperator_t>(platform, kind == Unsigned ? Unsigned : Signed, csr.NrowsN,
                                                csr.NrowsT, N, csr.rowStartsT, csr.colIdsT, csr.rowStartsN,
                                                csr.colIdsN);
}

void ogsBase_t::SetupBatch(const std::vector<const ogsBase_t *> &parts, platform_t &_platform) {
  Free();
  const auto t0 = std::chrono::steady_clock::now();
  LIBP_ABORT("SetupBatch needs at least one part", parts.empty());
  platform = _platform;
  comm = comm_t::world();
  kind = parts[0]->kind;
  unique = parts[0]->unique;
  gather_defined = true;
  // Part i's T rows follow parts 0..i-1's; its N rows start at the same row
  // (empty N rows fill the gap when a part has T rows without N members), so
  // the union keeps N-CSR row r within T-CSR row r as one rank's operator
  // does (ogsOperator.cpp:621-643) and the NoTrans / Trans forms act part by
  // part.
  long long n = 0, rowsT = 0, rowsN = 0, nnzT = 0, nnzN = 0;
  for (const ogsBase_t *p : parts) {
    LIBP_ABORT("SetupBatch: a part is not set up", !p || !p->gatherLocal);
    LIBP_ABORT("SetupBatch: parts must share one Kind", p->kind != kind);
    n += p->N;
    rowsN = rowsT + p->gatherLocal->NrowsN;
    rowsT += p->gatherLocal->NrowsT;
    nnzT += p->gatherLocal->nnzT;
    nnzN += p->gatherLocal->nnzN;
    gather_defined = gather_defined && p->gather_defined;
  }
  LIBP_ABORT("SetupBatch: " << n << " local nodes in all exceed dlong", n > std::numeric_limits<dlong>::max());
  N = (dlong)n;
  memory<dlong> rsT(rowsT + 1), ciT(nnzT > 0 ? nnzT : 1), rsN(rowsN + 1), ciN(nnzN > 0 ? nnzN : 1);
  dlong nodeOff = 0, rowOffT = 0, eOffT = 0, eOffN = 0;
  rsT[0] = 0;
  rsN[0] = 0;
  for (const ogsBase_t *p : parts) {
    const ogsOperator_t &op = *p->gatherLocal;
#pragma omp parallel for schedule(static)
    for (dlong r = 0; r < op.NrowsT; ++r) rsT[rowOffT + r + 1] = eOffT + op.rowStartsT[r + 1];
#pragma omp parallel for schedule(static)
    for (dlong e = 0; e < op.nnzT; ++e) ciT[eOffT + e] = nodeOff + op.colIdsT[e];
    // N rows of this part, then empty ones up to the next part's first row
    const dlong nN = std::min<dlong>(op.NrowsN, (dlong)(rowsN - rowOffT));
#pragma omp parallel for schedule(static)
    for (dlong r = 0; r < nN; ++r) rsN[rowOffT + r + 1] = eOffN + op.rowStartsN[r + 1];
    const dlong gapEnd = std::min<dlong>(op.NrowsT, (dlong)(rowsN - rowOffT));
    for (dlong r = nN; r < gapEnd; ++r) rsN[rowOffT + r + 1] = eOffN + op.nnzN;
#pragma omp parallel for schedule(static)
    for (dlong e = 0; e < op.nnzN; ++e) ciN[eOffN + e] = nodeOff + op.colIdsN[e];
    nodeOff += p->N;
    rowOffT += op.NrowsT;
    eOffT += op.nnzT;
    eOffN += op.nnzN;
  }
  NlocalT = (dlong)rowsT;
  NlocalP = (dlong)rowsN;
  NhaloT = 0;
  NhaloP = 0;
  Ngather = NlocalP;
  NgatherGlobal = Ngather;
  gatherLocal = std::make_shared<ogsOperator_t>(platform, kind == Unsigned ? Unsigned : Signed, NlocalP, NlocalT, N,
                                                rsT, ciT, rsN, ciN);
  // the parts' gather+scatter row orders and tile cuts, side by side
  bool ordered = false, cut = false;
  for (const ogsBase_t *p : parts) {
    ordered = ordered || p->gatherLocal->gsRowOrder.length() > 0;
    cut = cut || p->gatherLocal->gsTileCut.length() > 0;
  }
  if (ordered) {
    memory<dlong> order(NlocalT);
    memory<uint8_t> cuts;
    if (cut) cuts.malloc(NlocalT > 0 ? NlocalT : 1);
    dlong off = 0;
    for (const ogsBase_t *p : parts) {
      const ogsOperator_t &op = *p->gatherLocal;
      const bool has = op.gsRowOrder.length() > 0, hasCut = op.gsTileCut.length() > 0;
#pragma omp parallel for schedule(static)
      for (dlong r = 0; r < op.NrowsT; ++r) {
        order[off + r] = off + (has ? op.gsRowOrder[r] : r);
        if (cut) cuts[off + r] = (r == 0) || (hasCut && op.gsTileCut[r]);
      }
      off += op.NrowsT;
    }
    if (cut && NlocalT == 0) cuts = memory<uint8_t>();
    gatherLocal->SetGsRowOrder(order, false, cuts);
  }
  setupSeconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void ogsBase_t::Free() {
  comm.Free();
  gatherLocal = nullptr;
  N = 0;
  NlocalT = 0;
  NhaloT = 0;
  NlocalP = 0;
  NhaloP = 0;
  Ngather = 0;
  NgatherGlobal = 0;
}

void ogsBase_t::AssertGatherDefined() { LIBP_ABORT("Gather operation not well-defined.", !gather_defined); }

void ogs_t::Setup(const dlong _N, memory<hlong> ids, comm_t _comm, const Kind _kind, const Method method,
                  const bool _unique, const bool verbose, platform_t &_platform) {
  ogsBase_t::Setup(_N, ids, _comm, _kind, method, _unique, verbose, _platform);
}

// ogsSetup.cpp:833-849: GlobalToLocal[l] = gathered row of local node l, or -1.
void ogs_t::SetupGlobalToLocalMapping(memory<dlong> GlobalToLocal) {
  LIBP_ABORT("ogs handle is not set up.", NgatherGlobal == 0);
  LIBP_ABORT("GlobalToLocal must hold N entries", GlobalToLocal.length() < (size_t)N);
  for (dlong i = 0; i < N; ++i) GlobalToLocal[i] = -1;
  const ogsOperator_t &op = *gatherLocal;
  for (dlong r = 0; r < op.NrowsT; ++r)
    for (dlong i = op.rowStartsT[r]; i < op.rowStartsT[r + 1]; ++i) GlobalToLocal[op.colIdsT[i]] = r;
}

}  // namespace ogs
}  // namespace libp
