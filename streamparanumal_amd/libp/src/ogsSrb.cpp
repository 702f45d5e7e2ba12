// Sorted-run row blocks: the Sym gather+scatter operator re-listed so that
// each row block reads and writes its local nodes in ascending order.  Built
// once from the T-CSR and its row partition; the rows, their members and the
// per-row member order are the CSR's (ogsOperator.hpp:44-70), so the device
// result is bit-identical to the CSR kernel's.
#include <algorithm>
#include <vector>

#include "ogs.hpp"

namespace libp {
namespace ogs {

SortedRunBlocks BuildSortedRunBlocks(const dlong Nrows, const memory<dlong> rowStarts, const memory<dlong> colIds,
                                     const dlong Nblocks, const memory<dlong> blockStarts, const int cap) {
  LIBP_ABORT("row block capacity must be in [1, " << maxBlockNnz << "]", cap < 1 || cap > maxBlockNnz);
  SortedRunBlocks S;
  S.nblocks = Nblocks;
  S.cap = cap;
  S.nnz = Nrows > 0 ? rowStarts[Nrows] : 0;
  S.entryStarts.malloc(Nblocks + 1);
  S.chunkStarts.malloc(Nblocks + 1);
  S.chunkStarts[0] = 0;
  for (dlong b = 0; b <= Nblocks; ++b) S.entryStarts[b] = Nrows > 0 ? rowStarts[blockStarts[b]] : 0;
  for (dlong b = 0; b < Nblocks; ++b) {
    const dlong n = S.entryStarts[b + 1] - S.entryStarts[b];
    LIBP_ABORT("row block " << b << " holds " << n << " entries, more than " << cap, n > cap);
    S.chunkStarts[b + 1] = S.chunkStarts[b] + (n + 63) / 64;
  }
  S.nchunks = S.chunkStarts[Nblocks];
  const size_t nch = S.nchunks > 0 ? S.nchunks : 1;
  S.runMask.malloc(nch);
  S.runBase.malloc(nch);
  S.linkMask.malloc(nch);
  S.linkBase.malloc(nch);
  S.chunkRow.malloc(nch);
  memory<dlong> sorted(S.nnz > 0 ? S.nnz : 1);  // local index at each position
  memory<dlong> prevPos(S.nnz > 0 ? S.nnz : 1); // previous member's position, or -1
  std::vector<dlong> nruns(Nblocks, 0), nlinks(Nblocks, 0);
  int duplicate = 0;

  // pass 1: sort each block by local index, link row members, count runs and links
#pragma omp parallel
  {
    std::vector<std::pair<dlong, dlong>> ent;
    std::vector<dlong> posOfSlot;
#pragma omp for schedule(dynamic, 64)
    for (dlong b = 0; b < Nblocks; ++b) {
      const dlong e0 = S.entryStarts[b], n = S.entryStarts[b + 1] - e0;
      ent.resize(n);
      posOfSlot.resize(n);
      for (dlong i = 0; i < n; ++i) ent[i] = {colIds[e0 + i], i};
      std::sort(ent.begin(), ent.end());
      for (dlong i = 0; i < n; ++i) {
        sorted[e0 + i] = ent[i].first;
        posOfSlot[ent[i].second] = i;
        if (i > 0 && ent[i].first == ent[i - 1].first) {
#pragma omp atomic write
          duplicate = 1;
        }
      }
      dlong links = 0;
      for (dlong r = blockStarts[b]; r < blockStarts[b + 1]; ++r)
        for (dlong g = rowStarts[r]; g < rowStarts[r + 1]; ++g) {
          const dlong pos = posOfSlot[g - e0];
          prevPos[e0 + pos] = (g == rowStarts[r]) ? -1 : posOfSlot[g - 1 - e0];
          links += (g != rowStarts[r]);
        }
      dlong runs = 0, leading = 0;
      for (dlong c = 0; c * 64 < n; ++c) {
        S.chunkRow[S.chunkStarts[b] + c] = blockStarts[b] + leading;
        uint64_t rmask = 0, lmask = 0;
        for (dlong l = 0; l < 64 && c * 64 + l < n; ++l) {
          const dlong i = e0 + c * 64 + l;
          if (l == 0 || sorted[i] != sorted[i - 1] + 1) {
            rmask |= 1ull << l;
            ++runs;
          }
          if (prevPos[i] >= 0) lmask |= 1ull << l;
          else ++leading;
        }
        S.runMask[S.chunkStarts[b] + c] = rmask;
        S.linkMask[S.chunkStarts[b] + c] = lmask;
      }
      nruns[b] = runs;
      nlinks[b] = links;
    }
  }
  LIBP_ABORT("a local node appears twice in one row block", duplicate);
  // pass 2: run and link tables
  std::vector<dlong> runStart(Nblocks + 1, 0), linkStart(Nblocks + 1, 0);
  for (dlong b = 0; b < Nblocks; ++b) {
    runStart[b + 1] = runStart[b] + nruns[b];
    linkStart[b + 1] = linkStart[b] + nlinks[b];
  }
  S.nruns = runStart[Nblocks];
  S.nlinks = linkStart[Nblocks];
  S.runIds.malloc(S.nruns > 0 ? S.nruns : 1);
  S.links.malloc(S.nlinks + 1);  // one padding word: kernels may read links[linkBase] of a chunk without links
  S.links[S.nlinks] = 0;
#pragma omp parallel for schedule(dynamic, 64)
  for (dlong b = 0; b < Nblocks; ++b) {
    const dlong e0 = S.entryStarts[b], n = S.entryStarts[b + 1] - e0;
    dlong kr = runStart[b], kl = linkStart[b];
    for (dlong c = 0; c * 64 < n; ++c) {
      const dlong cc = S.chunkStarts[b] + c;
      S.runBase[cc] = kr;
      S.linkBase[cc] = kl;
      for (dlong l = 0; l < 64 && c * 64 + l < n; ++l) {
        const dlong i = e0 + c * 64 + l;
        if (S.runMask[cc] >> l & 1ull) S.runIds[kr++] = sorted[i];
        if (S.linkMask[cc] >> l & 1ull) S.links[kl++] = (uint16_t)prevPos[i];
      }
    }
  }
  return S;
}

void SortedRunBlocks::Upload(platform_t &platform) {
  o_entryStarts = platform.malloc<dlong>(entryStarts);
  o_chunkStarts = platform.malloc<dlong>(chunkStarts);
  o_runMask = platform.malloc<uint64_t>(runMask);
  o_runBase = platform.malloc<dlong>(runBase);
  o_runIds = platform.malloc<dlong>(runIds);
  o_linkMask = platform.malloc<uint64_t>(linkMask);
  o_linkBase = platform.malloc<dlong>(linkBase);
  o_links = platform.malloc<uint16_t>(links);
  o_chunkRow = platform.malloc<dlong>(chunkRow);
}

spx_srb_t SortedRunBlocks::Device() const {
  spx_srb_t d;
  d.nblocks = nblocks;
  d.max_block_nnz = cap;
  d.xcd_remap = xcdRemap;
  d.tile_lanes = tileLanes;
  d.entryStarts = o_entryStarts.ptr();
  d.chunkStarts = o_chunkStarts.ptr();
  d.runMask = o_runMask.ptr();
  d.runBase = o_runBase.ptr();
  d.runIds = o_runIds.ptr();
  d.linkMask = o_linkMask.ptr();
  d.linkBase = o_linkBase.ptr();
  d.links = o_links.ptr();
  d.chunkRow = o_chunkRow.ptr();
  return d;
}

}  // namespace ogs
}  // namespace libp
