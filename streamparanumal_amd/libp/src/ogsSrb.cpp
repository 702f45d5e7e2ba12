// Sorted-run row blocks ("tiles"): the operator's T-CSR re-listed so that
// each tile reads and writes its local nodes in ascending order.  Built once
// from the T-CSR and its row partition; the rows, their members and the
// per-row member order are the CSR's (ogsOperator.hpp:44-70), so the device
// results are bit-identical to the CSR kernels'.
#include <algorithm>
#include <vector>

#include "ogs.hpp"

namespace libp {
namespace ogs {

int SrbChunksPerTile(int cap) {
  // a power of two of at least 4 chunks: the kernels' tiles are 64 or 256
  // lanes times 4, 8 or 16 positions
  int c = 4;
  while (c * 64 < cap) c *= 2;
  return c;
}

SortedRunBlocks BuildSortedRunBlocks(const dlong Nrows, const memory<dlong> rowStarts, const memory<dlong> colIds,
                                     const dlong Nblocks, const memory<dlong> blockStarts, const int cap,
                                     const uint8_t *memberFlag) {
  LIBP_ABORT("row block capacity must be in [1, " << maxBlockNnz << "]", cap < 1 || cap > maxBlockNnz);
  SortedRunBlocks S;
  S.nblocks = Nblocks;
  S.cpt = SrbChunksPerTile(cap);
  S.nnz = Nrows > 0 ? rowStarts[Nrows] : 0;
  S.entryStarts.malloc(Nblocks + 1);
  for (dlong b = 0; b <= Nblocks; ++b) S.entryStarts[b] = Nrows > 0 ? rowStarts[blockStarts[b]] : 0;
  for (dlong b = 0; b < Nblocks; ++b) {
    const dlong n = S.entryStarts[b + 1] - S.entryStarts[b];
    LIBP_ABORT("row block " << b << " holds " << n << " entries, more than " << cap, n > cap);
    LIBP_ABORT("row block " << b << " is empty", n == 0);
  }
  S.nchunks = Nblocks * S.cpt;
  S.chunks.malloc(S.nchunks > 0 ? S.nchunks : 1);
  if (memberFlag) S.memberN.malloc(S.nchunks > 0 ? S.nchunks : 1);
  memory<dlong> sorted(S.nnz > 0 ? S.nnz : 1);   // local index at each position
  memory<dlong> prevPos(S.nnz > 0 ? S.nnz : 1);  // previous member's position, or -1
  std::vector<dlong> nruns(Nblocks, 0), nlinks(Nblocks, 0);
  int duplicate = 0, unordered = 0;
  // per node size 2^e bytes: (tile, 128-byte line) pairs with an entry, and
  // those holding the whole line
  int64_t lineAcc[6] = {0, 0, 0, 0, 0, 0}, wholeAcc[6] = {0, 0, 0, 0, 0, 0};

  // pass 1: sort each tile by local index, link row members, masks and counts
#pragma omp parallel
  {
    std::vector<std::pair<dlong, dlong>> ent;
    std::vector<dlong> posOfSlot;
    int64_t lines[6] = {0, 0, 0, 0, 0, 0}, whole[6] = {0, 0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
    for (dlong b = 0; b < Nblocks; ++b) {
      const dlong e0 = S.entryStarts[b], n = S.entryStarts[b + 1] - e0;
      ent.resize(n);
      posOfSlot.resize(n);
      for (dlong i = 0; i < n; ++i) ent[i] = {colIds[e0 + i], i};
      std::sort(ent.begin(), ent.end());
      for (int e = 1; e <= 5; ++e) {
        const int sh = 7 - e;  // 2^sh nodes per line
        for (dlong i = 0; i < n;) {
          const int64_t line = (int64_t)ent[i].first >> sh;
          dlong j = i + 1;
          while (j < n && ((int64_t)ent[j].first >> sh) == line) ++j;
          ++lines[e];
          whole[e] += (j - i == ((dlong)1 << sh));
          i = j;
        }
      }
      for (dlong i = 0; i < n; ++i) {
        sorted[e0 + i] = ent[i].first;
        posOfSlot[ent[i].second] = i;
        if (i > 0 && ent[i].first == ent[i - 1].first) {
#pragma omp atomic write
          duplicate = 1;
        }
      }
      dlong links = 0;
      // rows numbered by their leading positions: each row needs one, in row order
      for (dlong r = blockStarts[b]; r < blockStarts[b + 1]; ++r) {
        const bool empty = rowStarts[r + 1] == rowStarts[r];
        const bool noPrev = r == blockStarts[b] || rowStarts[r] == rowStarts[r - 1];
        if (empty || (!noPrev && posOfSlot[rowStarts[r] - e0] < posOfSlot[rowStarts[r - 1] - e0])) {
#pragma omp atomic write
          unordered = 1;
        }
      }
      for (dlong r = blockStarts[b]; r < blockStarts[b + 1]; ++r)
        for (dlong g = rowStarts[r]; g < rowStarts[r + 1]; ++g) {
          const dlong pos = posOfSlot[g - e0];
          prevPos[e0 + pos] = (g == rowStarts[r]) ? -1 : posOfSlot[g - 1 - e0];
          links += (g != rowStarts[r]);
        }
      dlong runs = 0, leading = 0;
      for (dlong k = 0; k < S.cpt; ++k) {
        SrbChunkRecord &C = S.chunks[(size_t)b * S.cpt + k];
        C.runMask = 0;
        C.linkMask = 0;
        C.nvalid = (int32_t)std::max<dlong>(0, std::min<dlong>(64, n - 64 * k));
        C.chunkRow = blockStarts[b] + leading;
        uint64_t member = 0;
        for (dlong l = 0; l < C.nvalid; ++l) {
          const dlong i = e0 + 64 * k + l;
          if (l == 0 || sorted[i] != sorted[i - 1] + 1) {
            C.runMask |= 1ull << l;
            ++runs;
          }
          if (prevPos[i] >= 0) C.linkMask |= 1ull << l;
          else ++leading;
          if (memberFlag && memberFlag[sorted[i]]) member |= 1ull << l;
        }
        if (memberFlag) S.memberN[(size_t)b * S.cpt + k] = member;
        if (C.nvalid == 0) {  // padding: decodes as the tile's first node and row (never stored)
          C.runMask = 1;
          C.chunkRow = blockStarts[b];
        }
      }
      nruns[b] = runs;
      nlinks[b] = links;
    }
    for (int e = 1; e <= 5; ++e) {
#pragma omp atomic
      lineAcc[e] += lines[e];
#pragma omp atomic
      wholeAcc[e] += whole[e];
    }
  }
  LIBP_ABORT("a local node appears twice in one row block", duplicate);
  S.rowsAscending = !unordered;
  S.ntMask = 0;
  for (int e = 1; e <= 5; ++e)
    if (lineAcc[e] > 0 && (double)wholeAcc[e] >= kSrbWholeLineShare * (double)lineAcc[e]) S.ntMask |= 1 << e;
  // pass 2: run and link tables
  std::vector<dlong> runStart(Nblocks + 1, 0), linkStart(Nblocks + 1, 0);
  for (dlong b = 0; b < Nblocks; ++b) {
    runStart[b + 1] = runStart[b] + nruns[b];
    linkStart[b + 1] = linkStart[b] + nlinks[b];
  }
  S.nruns = runStart[Nblocks];
  S.nlinks = linkStart[Nblocks];
  S.runIds.malloc(S.nruns > 0 ? S.nruns : 1);
  S.links.malloc(S.nlinks + 1);  // one padding word
  S.links[S.nlinks] = 0;
#pragma omp parallel for schedule(dynamic, 64)
  for (dlong b = 0; b < Nblocks; ++b) {
    const dlong e0 = S.entryStarts[b];
    dlong kr = runStart[b], kl = linkStart[b];
    for (dlong k = 0; k < S.cpt; ++k) {
      SrbChunkRecord &C = S.chunks[(size_t)b * S.cpt + k];
      C.runBase = (int32_t)(C.nvalid ? kr : runStart[b]);
      C.linkBase = (int32_t)kl;
      for (dlong l = 0; l < C.nvalid; ++l) {
        const dlong i = e0 + 64 * k + l;
        // a run's local index less its position in the chunk: position l of
        // the run decodes as runIds[run] + l
        if (C.runMask >> l & 1ull) S.runIds[kr++] = sorted[i] - l;
        if (C.linkMask >> l & 1ull) S.links[kl++] = (uint16_t)prevPos[i];
      }
    }
  }
  return S;
}

void SortedRunBlocks::Upload(platform_t &platform) {
  o_chunks = platform.malloc<SrbChunkRecord>(chunks);
  o_runIds = platform.malloc<dlong>(runIds);
  o_links = platform.malloc<uint16_t>(links);
  o_memberN = memberN.length() ? platform.malloc<uint64_t>(memberN) : deviceMemory<uint64_t>();
}

memory<uint8_t> SignedMemberFlags(const dlong Ncols, const dlong NrowsT, const memory<dlong> rowStartsT,
                                  const memory<dlong> colIdsT, const dlong NrowsN, const memory<dlong> rowStartsN,
                                  const memory<dlong> colIdsN) {
  memory<uint8_t> flag(Ncols > 0 ? Ncols : 1);
  if (NrowsN > NrowsT) return memory<uint8_t>();
#pragma omp parallel for schedule(static)
  for (dlong i = 0; i < Ncols; ++i) flag[i] = 0;
  const dlong nnzN = NrowsN > 0 ? rowStartsN[NrowsN] : 0;
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (dlong e = 0; e < nnzN; ++e) {
    const dlong c = colIdsN[e];
    if (c < 0 || c >= Ncols) bad = 1;
    else flag[c] = 1;
  }
  if (bad) return memory<uint8_t>();
  // walk each T row's flagged members against its N row
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (dlong r = 0; r < NrowsT; ++r) {
    dlong n = r < NrowsN ? rowStartsN[r] : 0;
    const dlong nEnd = r < NrowsN ? rowStartsN[r + 1] : 0;
    for (dlong g = rowStartsT[r]; g < rowStartsT[r + 1]; ++g) {
      const dlong c = colIdsT[g];
      if (!flag[c]) continue;
      if (n >= nEnd || colIdsN[n] != c) {
        bad = 1;
        break;
      }
      ++n;
    }
    if (n != nEnd) bad = 1;
  }
  if (bad) return memory<uint8_t>();
  return flag;
}

spx_srb_t SortedRunBlocks::Device() const {
  spx_srb_t d;
  d.nblocks = nblocks;
  d.chunks_per_tile = cpt;
  d.xcd_remap = xcdRemap;
  d.tile_lanes = tileLanes;
  d.nontemporal = ntHint < 0 ? ntMask : ntHint ? 0x3e : 0;
  d.chunks = reinterpret_cast<const spx_srb_chunk_t *>(o_chunks.ptr());
  d.runIds = o_runIds.ptr();
  d.links = o_links.ptr();
  return d;
}

}  // namespace ogs
}  // namespace libp
