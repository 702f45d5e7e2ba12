// ogsOperator_t on the device (libs/ogs/ogsOperator.cpp:176-204, 295-322,
// 442-481, 597-665) over the libspx ogs kernels.
#include <algorithm>
#include <climits>
#include <vector>

#include <omp.h>

#include "ogs.hpp"

namespace libp {
namespace ogs {

void blockRowsGreedy(const dlong Nrows, const memory<dlong> rowStarts, const int cap, dlong &Nblocks,
                     memory<dlong> &blockStarts, const uint8_t *cuts) {
  Nblocks = 0;
  std::vector<dlong> starts;
  starts.push_back(0);
  long long acc = 0;
  for (dlong r = 0; r < Nrows; ++r) {
    const dlong sz = rowStarts[r + 1] - rowStarts[r];
    LIBP_ABORT("Multiplicity of a global node (" << sz << ") exceeds the row block capacity " << cap,
               sz > cap);
    if ((acc + sz > cap || (cuts && cuts[r])) && acc > 0) {
      starts.push_back(r);
      acc = 0;
    }
    acc += sz;
  }
  if (Nrows > 0) starts.push_back(Nrows);
  Nblocks = (dlong)starts.size() - 1;
  if (Nblocks < 0) Nblocks = 0;
  blockStarts.malloc(starts.size());
  std::copy(starts.begin(), starts.end(), blockStarts.ptr());
}

ogsOperator_t::ogsOperator_t(platform_t &platform_, Kind kind_, const dlong NrowsN_, const dlong NrowsT_,
                             const dlong Ncols_, memory<dlong> rowStartsT_, memory<dlong> colIdsT_,
                             memory<dlong> rowStartsN_, memory<dlong> colIdsN_, int blockNnz_,
                             int gsBlockNnz_)
    : platform(platform_), Ncols(Ncols_), NrowsN(NrowsN_), NrowsT(NrowsT_), kind(kind_) {
  rowStartsT = rowStartsT_;
  colIdsT = colIdsT_;
  nnzT = rowStartsT[NrowsT];
  o_rowStartsT = platform.malloc<dlong>(NrowsT + 1, rowStartsT);
  o_colIdsT = platform.malloc<dlong>(nnzT, colIdsT);
  if (kind == Signed) {
    rowStartsN = rowStartsN_;
    colIdsN = colIdsN_;
    nnzN = rowStartsN[NrowsN];
    o_rowStartsN = platform.malloc<dlong>(NrowsN + 1, rowStartsN);
    o_colIdsN = platform.malloc<dlong>(nnzN, colIdsN);
  } else {
    rowStartsN = rowStartsT;
    colIdsN = colIdsT;
    nnzN = nnzT;
    o_rowStartsN = o_rowStartsT;
    o_colIdsN = o_colIdsT;
  }
  if (kind == Signed) {
    memberFlagN = SignedMemberFlags(Ncols, NrowsT, rowStartsT, colIdsT, NrowsN, rowStartsN, colIdsN);
    signedSrb = memberFlagN.length() > 0;
  }
  SetBlockNnz(blockNnz_);
  // a small operator at the default capacity takes smaller tiles, so that
  // its launches still fill the GPU: halved (down to 256 entries) while it
  // would make fewer than 512 tiles.  Measured on MI355X, 8^3 p=7 (262,144
  // entries), us per gather+scatter: 2048 6.46, 1024 5.09, 512 4.81, 256 4.85
  if (gsBlockNnz_ == defaultGsBlockNnz)
    while (gsBlockNnz_ > 256 && (int64_t)nnzT < 512 * (int64_t)gsBlockNnz_) gsBlockNnz_ /= 2;
  SetGsBlockNnz(gsBlockNnz_);
  auto localRows = [&](dlong nrows, const memory<dlong> &rs, const memory<dlong> &ci) {
    memory<dlong> m(Ncols > 0 ? Ncols : 1);
#pragma omp parallel for schedule(static)
    for (dlong i = 0; i < Ncols; ++i) m[i] = -1;
#pragma omp parallel for schedule(static)
    for (dlong r = 0; r < nrows; ++r)
      for (dlong i = rs[r]; i < rs[r + 1]; ++i) m[ci[i]] = r;  // each local node in at most one row
    return platform.malloc<dlong>(Ncols, m);
  };
  o_localRowT = localRows(NrowsT, rowStartsT, colIdsT);
  o_localRowN = (kind == Signed) ? localRows(NrowsN, rowStartsN, colIdsN) : o_localRowT;
  scatterByRows = NrowsT >= kScatterByRowsMinRatio * Ncols;
}

void ogsOperator_t::SetGsBlockNnz(int nnz) {
  LIBP_ABORT("row block capacity must be in [1, " << maxBlockNnz << "]", nnz < 1 || nnz > maxBlockNnz);
  gsBlockNnz = nnz;
  blockRowsGreedy(NrowsT, rowStartsT, gsBlockNnz, NgsBlocks, gsBlockRowStarts);
  o_gsBlockRowStarts = platform.malloc<dlong>(NgsBlocks + 1, gsBlockRowStarts);
  const int xcd = gsSrb.xcdRemap, lanes = gsSrb.tileLanes;
  const int rxcd = rowSrb.xcdRemap, rlanes = rowSrb.tileLanes;
  const uint8_t *flags = signedSrb ? memberFlagN.ptr() : nullptr;
  rowSrb = BuildSortedRunBlocks(NrowsT, rowStartsT, colIdsT, NgsBlocks, gsBlockRowStarts, gsBlockNnz, flags);
  rowSrb.xcdRemap = rxcd;
  rowSrb.tileLanes = rlanes;
  rowSrb.Upload(platform);
  BuildGsSrb(xcd, lanes);
  // the FP32 / BF16 tiles at the new capacity (one layout where both share an order)
  for (int s = 0; s < 2; ++s) {
    MixedTiles &m = gsMixed[s];
    if (!m.own) continue;
    if (s == 1 && gsMixed[0].own && m.order.ptr() == gsMixed[0].order.ptr() && m.cuts.ptr() == gsMixed[0].cuts.ptr())
      m.srb = gsMixed[0].srb;
    else
      m.srb = BuildOrderedSrb(m.order, m.cuts);
  }
}

SortedRunBlocks ogsOperator_t::BuildOrderedSrb(const memory<dlong> &order, const memory<uint8_t> &cuts) {
  const uint8_t *flags = signedSrb ? memberFlagN.ptr() : nullptr;
  if (!order.length()) return rowSrb;
  // the T-CSR with its rows in this order, cut into greedy blocks
  memory<dlong> rs(NrowsT + 1), ci(nnzT > 0 ? nnzT : 1);
#pragma omp parallel for schedule(static)
  for (dlong i = 0; i < NrowsT; ++i) {
    const dlong r = order[i];
    rs[i] = rowStartsT[r + 1] - rowStartsT[r];
  }
  rs[NrowsT] = ExclusiveScan(rs.ptr(), NrowsT);
#pragma omp parallel for schedule(static)
  for (dlong i = 0; i < NrowsT; ++i) {
    const dlong r = order[i];
    std::copy(colIdsT.ptr() + rowStartsT[r], colIdsT.ptr() + rowStartsT[r + 1], ci.ptr() + rs[i]);
  }
  dlong nb = 0;
  memory<dlong> bs;
  blockRowsGreedy(NrowsT, rs, gsBlockNnz, nb, bs, cuts.length() ? cuts.ptr() : nullptr);
  SortedRunBlocks srb = BuildSortedRunBlocks(NrowsT, rs, ci, nb, bs, gsBlockNnz, flags);
  srb.Upload(platform);
  return srb;
}

void ogsOperator_t::BuildGsSrb(int xcd, int lanes) {
  const int nt = gsSrb.ntHint;
  gsSrb = BuildOrderedSrb(gsRowOrder, gsTileCut);
  gsSrb.xcdRemap = xcd;
  gsSrb.tileLanes = lanes;
  gsSrb.ntHint = nt;
}

void ogsOperator_t::SetGsMixedRowOrder(const int storage, const memory<dlong> &order, const memory<uint8_t> &cuts,
                                       const bool validate) {
  LIBP_ABORT("mixed tiles are for SPX_STORE_F32 or SPX_STORE_BF16",
             storage != SPX_STORE_F32 && storage != SPX_STORE_BF16);
  MixedTiles &m = gsMixed[storage - 1];
  if (!order.length()) {
    m = MixedTiles();
    return;
  }
  if (validate) {
    LIBP_ABORT("gather+scatter row order must list the " << NrowsT << " rows", (dlong)order.length() != NrowsT);
    std::vector<char> seen(NrowsT, 0);
    for (dlong i = 0; i < NrowsT; ++i) {
      const dlong r = order[i];
      LIBP_ABORT("gather+scatter row order is not a permutation of the rows", r < 0 || r >= NrowsT || seen[r]);
      seen[r] = 1;
    }
  }
  LIBP_ABORT("gather+scatter tile cuts need one flag per row", cuts.length() && cuts.length() != order.length());
  m.own = true;
  m.order = order;
  m.cuts = cuts;
  const MixedTiles &o = gsMixed[2 - storage];  // the other storage's tiles
  if (o.own && o.order.ptr() == order.ptr() && o.cuts.ptr() == cuts.ptr()) m.srb = o.srb;
  else m.srb = BuildOrderedSrb(order, cuts);
}

void ogsOperator_t::SetGsRowOrder(const memory<dlong> &order, const bool validate, const memory<uint8_t> &cuts) {
  if (order.length() && validate) {
    LIBP_ABORT("gather+scatter row order must list the " << NrowsT << " rows", (dlong)order.length() != NrowsT);
    std::vector<char> seen(NrowsT, 0);
    for (dlong i = 0; i < NrowsT; ++i) {
      const dlong r = order[i];
      LIBP_ABORT("gather+scatter row order is not a permutation of the rows", r < 0 || r >= NrowsT || seen[r]);
      seen[r] = 1;
    }
  }
  LIBP_ABORT("gather+scatter tile cuts need a row order with one flag per row",
             cuts.length() && (!order.length() || cuts.length() != order.length()));
  gsRowOrder = order;
  gsTileCut = cuts;
  BuildGsSrb(gsSrb.xcdRemap, gsSrb.tileLanes);
}


memory<dlong> LatticeBlockOrder(const dlong Nrows, const memory<hlong> &rowGrid, const hlong Mx, const hlong My,
                                const hlong Mz, const int bx, const int by, const int bz) {
  LIBP_ABORT("lattice blocks need positive sizes", Mx < 1 || My < 1 || Mz < 1 || bx < 1 || by < 1 || bz < 1);
  const hlong nbx = (Mx + bx - 1) / bx, nby = (My + by - 1) / by, nbz = (Mz + bz - 1) / bz;
  const hlong nkeys = nbx * nby * nbz;
  LIBP_ABORT("too many lattice blocks (" << nkeys << ")", nkeys > (hlong)INT32_MAX);
  if (Nrows <= 0) return memory<dlong>();  // no rows: the CSR's (empty) order
  std::vector<int32_t> key(Nrows);
#pragma omp parallel for schedule(static)
  for (dlong r = 0; r < Nrows; ++r) {
    const hlong g = rowGrid[r];
    const hlong i = g % Mx, j = (g / Mx) % My, k = g / (Mx * My);
    key[r] = (int32_t)(((k / bz) * nby + j / by) * nbx + i / bx);
  }
  return CountingOrder(Nrows, key, nkeys);
}

void HexLineTiles(const dlong Nrows, const memory<hlong> &rowGrid, const hlong nx, const hlong ny, const hlong nz,
                  const int p, const int rowsPerLine, const int cap, memory<dlong> &order, memory<uint8_t> &cuts) {
  LIBP_ABORT("line tiles need a box, a degree and a capacity",
             nx < 1 || ny < 1 || nz < 1 || p < 1 || rowsPerLine < 1 || cap < 1);
  order = memory<dlong>();
  cuts = memory<uint8_t>();
  if (Nrows <= 0) return;
  const hlong Mx = nx * p + 1, My = ny * p + 1, Mz = nz * p + 1;
  auto mult = [p](hlong a, hlong n) -> hlong { return (a % p == 0 && a > 0 && a < n * p) ? 2 : 1; };
  // y groups: an element's node rows j = 0..p in runs of rowsPerLine; the
  // last run of element ey and the first of ey+1 share the point p*(ey+1)
  // and merge, unless that run of ey is the shared point alone (p a multiple
  // of rowsPerLine), which then starts a group
  std::vector<int32_t> yg(My);
  int32_t ng = 0;
  for (hlong y = 0; y < My; ++y) {
    const hlong j = y % p;
    if (y == 0 || (j != 0 && j % rowsPerLine == 0) || (j == 0 && p % rowsPerLine == 0)) ++ng;
    yg[y] = ng - 1;
  }
  std::vector<hlong> ysl(ng, 0);  // y slots (members along y) per group
  for (hlong y = 0; y < My; ++y) ysl[yg[y]] += mult(y, ny);
  const hlong xsl = nx * (p + 1);  // members along a full x line
  // per z plane: consecutive groups merged while a tile holds at most cap
  // members; a group larger than cap is cut along x at element boundaries
  std::vector<int32_t> tileOf((size_t)Mz * ng), chunksOf((size_t)Mz * ng);
  std::vector<int64_t> zBase(Mz + 1, 0);
  for (hlong z = 0; z < Mz; ++z) {
    int64_t tiles = 0, acc = 0;
    for (int32_t g = 0; g < ng; ++g) {
      const int64_t P = xsl * ysl[g] * mult(z, nz);
      const size_t zg = (size_t)z * ng + g;
      if (P > cap) {
        const int64_t nch = std::min<int64_t>((P + cap - 1) / cap, nx);
        if (acc > 0) ++tiles;
        tileOf[zg] = (int32_t)tiles;
        chunksOf[zg] = (int32_t)nch;
        tiles += nch;
        acc = 0;
        continue;
      }
      if (acc > 0 && acc + P > cap) {
        ++tiles;
        acc = 0;
      }
      tileOf[zg] = (int32_t)tiles;
      chunksOf[zg] = 1;
      acc += P;
    }
    if (acc > 0) ++tiles;
    zBase[z + 1] = zBase[z] + tiles;
  }
  LIBP_ABORT("too many line tiles (" << zBase[Mz] << ")", zBase[Mz] > (int64_t)INT32_MAX);
  std::vector<int32_t> key(Nrows);
#pragma omp parallel for schedule(static)
  for (dlong r = 0; r < Nrows; ++r) {
    const hlong g = rowGrid[r];
    const hlong i = g % Mx, j = (g / Mx) % My, k = g / (Mx * My);
    const size_t zg = (size_t)k * ng + yg[j];
    const hlong ex = std::min<hlong>(i / p, nx - 1);
    key[r] = (int32_t)(zBase[k] + tileOf[zg] + (ex * chunksOf[zg]) / nx);
  }
  order = CountingOrder(Nrows, key, zBase[Mz]);
  cuts.malloc(Nrows);
#pragma omp parallel for schedule(static)
  for (dlong i = 0; i < Nrows; ++i) cuts[i] = (i == 0) || key[order[i]] != key[order[i - 1]];
}

memory<dlong> CountingOrder(const dlong Nrows, const std::vector<int32_t> &key, const hlong nkeys) {
  memory<dlong> order(Nrows);
  // stable counting sort by block: each thread counts its contiguous share of
  // the rows, the counts are scanned block-major then thread-major, and each
  // thread places its rows in row order.  The shares follow the team the
  // runtime actually gives (it may be smaller than asked for: OMP_THREAD_LIMIT,
  // OMP_DYNAMIC, nesting).
  const int Tmax = std::max(1, std::min(omp_get_max_threads(), (int)((Nrows + 65535) / 65536)));
  std::vector<int64_t> count((size_t)Tmax * nkeys, 0);
#pragma omp parallel num_threads(Tmax)
  {
    const int T = omp_get_num_threads(), t = omp_get_thread_num();
    const dlong r0 = (dlong)((int64_t)Nrows * t / T), r1 = (dlong)((int64_t)Nrows * (t + 1) / T);
    int64_t *c = count.data() + (size_t)t * nkeys;
    for (dlong r = r0; r < r1; ++r) ++c[key[r]];
#pragma omp barrier
#pragma omp single
    {
      int64_t acc = 0;
      for (hlong b = 0; b < nkeys; ++b)
        for (int u = 0; u < T; ++u) {
          int64_t &x = count[(size_t)u * nkeys + b];
          const int64_t n = x;
          x = acc;
          acc += n;
        }
    }
    for (dlong r = r0; r < r1; ++r) order[c[key[r]]++] = r;
  }
  return order;
}

int HexGsLineRows(const int p) {
  // Line tiles cover whole 128-byte lines of FP64 values where an element's
  // node rows fill them: 2 rows at p=7, 1 row at p=15.  Measured on MI355X,
  // us per call, setup's lattice blocks -> line tiles, FP64 / FP32 / BF16
  // storage: 64^3 p=7 481 / 351 / 364 -> 460 / 331 / 348 (three runs: 4-6 %
  // for FP64), 58^3 p=7 365 -> 356; with the non-temporal values whole
  // lines allow (kSrbWholeLineShare): 29^3 p=15 FP64 283 (blocks) -> 259
  // (1 row; 261 with 2), 64^3 p=7 346.  At other degrees the blocks were as
  // fast or faster for FP64 (~1e8 nodes: p=1 483 -> 653, p=8 386 -> 443,
  // p=9 370 -> 396, p=3 409 -> 404; 1 or 4 rows per tile were no better).
  return p == 7 ? 2 : p == 15 ? 1 : 0;
}

int HexGsLineRowsMixed(const int p, const int storage) {
  // Measured on MI355X at ~1e8 nodes, us per call, the FP64 tiles -> line
  // tiles: BF16 p=1 376 -> 253 (1 row), p=2 382 -> 342 (2 rows; 428 with 1),
  // p=7 267 -> 254 (1 row; 64^3: 346 -> 329); FP32 p=7 247 -> 241 (64^3: 318
  // -> 307).  FP32 lost at p=1 (363 -> 401) and p=2; at the other degrees
  // either storage was within 2.5 % or slower.  At p=15 both keep the
  // lattice blocks (-1) where FP64 takes line tiles: 29^3 FP32 / BF16 191 /
  // 186 against 197 / 191 on the FP64 line tiles.
  if (p == 15) return -1;
  if (storage == SPX_STORE_BF16) return p == 1 ? 1 : p == 2 ? 2 : p == 7 ? 1 : 0;
  if (storage == SPX_STORE_F32) return p == 7 ? 1 : 0;
  return 0;
}

void HexGsBlockShape(const int p, int &bx, int &by, int &bz) {
  // Measured on MI355X, us per call (CSR order -> blocks): ~1e8 nodes, FP64:
  // p=1 640 -> 436 (8x4x4), p=2 601 -> 447 (8x8x4), p=3 387 -> 361
  // (16x8x2), p=4 415 -> 398 (8x8x4; 410 with 16x8x2); 64^3 p=7, FP64 /
  // FP32 / BF16 storage (medians of 3 interleaved rounds): 469 / 347 / 370
  // -> 453 / 341 / 368 with 64x8x1, where 64x4x1 gives 460 / 335 / 357 and
  // 16x8x2 460 / 366 / 393.  From p=5 up the shapes tried (8x8x4 to
  // 128x4x1) differ by at most 3 %, so p >= 7 takes 64x8x1 (p=15 357 ->
  // 304; p = 7 and 15 take line tiles, HexGsLineRows).  Narrow blocks later did better at p = 1 and 2, whose tiles
  // then take the x points of several blocks (~1e8 nodes, FP64 / FP32
  // storage): p=1 8x4x4 425 / 271 -> 1x4x4 403 / 267 (2x4x4 406, 4x4x4
  // 413, 2x8x4 416, 16x4x4 440), p=2 8x8x4 443 / 283 -> 2x8x6 425 / 280
  // (2x8x4 433, 2x6x6 426, 2x10x6 427, 4x8x6 428, 2x8x8 432, 4x4x4 454),
  // and at p = 6: 64x8x1 375 / 248 ->
  // 4x8x2 367 / 233 (8x8x1 371, 16x8x1 372); at p = 3..5 they were within
  // 1 %.
  if (p <= 1) {
    bx = 1; by = 4; bz = 4;
  } else if (p == 2) {
    bx = 2; by = 8; bz = 6;
  } else if (p == 4) {
    bx = 8; by = 8; bz = 4;
  } else if (p <= 5) {
    bx = 16; by = 8; bz = 2;
  } else if (p == 6) {
    bx = 4; by = 8; bz = 2;
  } else {
    bx = 64; by = 8; bz = 1;
  }
}

void ogsOperator_t::SetBlockNnz(int nnz) {
  LIBP_ABORT("row block capacity must be in [1, " << maxBlockNnz << "]", nnz < 1 || nnz > maxBlockNnz);
  blockNnz = nnz;
  blockRowsGreedy(NrowsT, rowStartsT, blockNnz, NrowBlocksT, blockRowStartsT);
  o_blockRowStartsT = platform.malloc<dlong>(NrowBlocksT + 1, blockRowStartsT);
  if (kind == Signed) {
    blockRowsGreedy(NrowsN, rowStartsN, blockNnz, NrowBlocksN, blockRowStartsN);
    o_blockRowStartsN = platform.malloc<dlong>(NrowBlocksN + 1, blockRowStartsN);
  } else {
    NrowBlocksN = NrowBlocksT;
    blockRowStartsN = blockRowStartsT;
    o_blockRowStartsN = o_blockRowStartsT;
  }
  genBlockNnz = std::min(blockNnz, 2048);
  if (genBlockNnz == blockNnz) {
    NgenBlocks = NrowBlocksT;
    o_genBlockRowStarts = o_blockRowStartsT;
  } else {
    memory<dlong> starts;
    blockRowsGreedy(NrowsT, rowStartsT, genBlockNnz, NgenBlocks, starts);
    o_genBlockRowStarts = platform.malloc<dlong>(NgenBlocks + 1, starts);
  }
}

void ogsOperator_t::Free() { *this = ogsOperator_t(); }

void ogsOperator_t::GatherRaw(Type type, Op op, Transpose trans, int k, const void *v, void *gv) {
  const bool srbRows = NgsBlocks && rowSrb.rowsAscending;  // the layout numbers the rows
  if (trans == NoTrans && gatherSorted && srbRows && (signedSrb || kind == Unsigned)) {
    const spx_srb_t d = rowSrb.Device();
    if (kind == Unsigned) SPX_CHECK(spx_ogs_gather_srb((int)type, (int)op, k, &d, v, gv));
    else SPX_CHECK(spx_ogs_gather_srb_signed((int)type, (int)op, k, &d, rowSrb.DeviceMemberN(), NrowsN, v, gv));
  } else if (trans == NoTrans) {
    if (NrowBlocksN)
      SPX_CHECK(spx_ogs_gather((int)type, (int)op, k, NrowBlocksN, o_blockRowStartsN.ptr(), o_rowStartsN.ptr(),
                               o_colIdsN.ptr(), blockNnz, v, gv));
  } else if (trans != NoTrans && gatherSorted && srbRows) {
    const spx_srb_t d = rowSrb.Device();
    SPX_CHECK(spx_ogs_gather_srb((int)type, (int)op, k, &d, v, gv));
  } else {
    if (NrowBlocksT)
      SPX_CHECK(spx_ogs_gather((int)type, (int)op, k, NrowBlocksT, o_blockRowStartsT.ptr(), o_rowStartsT.ptr(),
                               o_colIdsT.ptr(), blockNnz, v, gv));
  }
}

void ogsOperator_t::ScatterRaw(Type type, Transpose trans, int k, const void *gv, void *v) {
  if (!scatterByRows) {
    const dlong nrows = (trans == Trans) ? NrowsN : NrowsT;
    if (nrows && Ncols)
      SPX_CHECK(spx_ogs_scatter_inverse((int)type, k, Ncols, (trans == Trans ? o_localRowN : o_localRowT).ptr(), gv, v));
    return;
  }
  const bool srbRows = NgsBlocks && rowSrb.rowsAscending;  // the layout numbers the rows
  if (trans == Trans && scatterSorted && srbRows && (signedSrb || kind == Unsigned)) {
    const spx_srb_t d = rowSrb.Device();
    if (kind == Unsigned) SPX_CHECK(spx_ogs_scatter_srb((int)type, k, &d, gv, v));
    else SPX_CHECK(spx_ogs_scatter_srb_signed((int)type, k, &d, rowSrb.DeviceMemberN(), NrowsN, gv, v));
  } else if (trans == Trans) {
    if (NrowBlocksN)
      SPX_CHECK(spx_ogs_scatter((int)type, k, NrowBlocksN, o_blockRowStartsN.ptr(), o_rowStartsN.ptr(),
                                o_colIdsN.ptr(), blockNnz, gv, v));
  } else if (trans != Trans && scatterSorted && srbRows) {
    const spx_srb_t d = rowSrb.Device();
    SPX_CHECK(spx_ogs_scatter_srb((int)type, k, &d, gv, v));
  } else {
    if (NrowBlocksT)
      SPX_CHECK(spx_ogs_scatter((int)type, k, NrowBlocksT, o_blockRowStartsT.ptr(), o_rowStartsT.ptr(),
                                o_colIdsT.ptr(), blockNnz, gv, v));
  }
}

// The reference's device GatherScatter launches the Add kernel whatever op
// is (ogsOperator.cpp:448-479) while its host version honours op
// (:411-427); here the device honours op as the API documents.
void ogsOperator_t::GatherScatterRaw(Type type, Op op, Transpose trans, int k, void *v) {
  if (trans == Sym) {
    if (NgsBlocks && gsSorted) {
      const spx_srb_t d = gsSrb.Device();
      SPX_CHECK(spx_ogs_gather_scatter_srb((int)type, (int)op, k, &d, v));
    } else if (NgsBlocks)
      SPX_CHECK(spx_ogs_gather_scatter((int)type, (int)op, k, NgsBlocks, o_gsBlockRowStarts.ptr(), NrowsT,
                                       o_rowStartsT.ptr(), o_colIdsT.ptr(), NrowsT, o_rowStartsT.ptr(),
                                       o_colIdsT.ptr(), gsBlockNnz, v));
    return;
  }
  if (NgsBlocks && gsSorted && (signedSrb || kind == Unsigned)) {
    // Unsigned: the N-CSR is the T-CSR, so both transposes are the Sym form
    const spx_srb_t d = gsSrb.Device();
    if (kind == Unsigned) SPX_CHECK(spx_ogs_gather_scatter_srb((int)type, (int)op, k, &d, v));
    else
      SPX_CHECK(spx_ogs_gather_scatter_srb_signed((int)type, (int)op, (int)trans, k, &d, gsSrb.DeviceMemberN(), v));
    return;
  }
  if (!NgenBlocks) return;
  const dlong *bs = o_genBlockRowStarts.ptr();
  if (trans == Trans) {
    SPX_CHECK(spx_ogs_gather_scatter((int)type, (int)op, k, NgenBlocks, bs, NrowsT, o_rowStartsT.ptr(),
                                     o_colIdsT.ptr(), NrowsN, o_rowStartsN.ptr(), o_colIdsN.ptr(), genBlockNnz, v));
  } else {
    SPX_CHECK(spx_ogs_gather_scatter((int)type, (int)op, k, NgenBlocks, bs, NrowsN, o_rowStartsN.ptr(),
                                     o_colIdsN.ptr(), NrowsT, o_rowStartsT.ptr(), o_colIdsT.ptr(), genBlockNnz, v));
  }
}

void ogsOperator_t::GatherScatterMixedRaw(int storage, int k, void *v) {
  if (NgsBlocks && gsSorted) {
    const bool own = (storage == SPX_STORE_F32 || storage == SPX_STORE_BF16) && gsMixed[storage - 1].own;
    spx_srb_t d = own ? gsMixed[storage - 1].srb.Device() : gsSrb.Device();
    d.xcd_remap = gsSrb.xcdRemap;
    d.tile_lanes = gsSrb.tileLanes;
    if (gsSrb.ntHint >= 0) d.nontemporal = gsSrb.ntHint ? 0x3e : 0;  // the operator's override
    SPX_CHECK(spx_ogs_gather_scatter_srb_mixed(storage, k, &d, v));
  } else if (NgsBlocks)
    SPX_CHECK(spx_ogs_gather_scatter_mixed(storage, k, NgsBlocks, o_gsBlockRowStarts.ptr(), o_rowStartsT.ptr(),
                                           o_colIdsT.ptr(), gsBlockNnz, v));
}

template <typename T>
void ogsOperator_t::Gather(deviceMemory<T> gv, const deviceMemory<T> v, const int k, const Op op,
                           const Transpose trans) {
  GatherRaw(ogsType<T>::get(), op, trans, k, v.ptr(), const_cast<deviceMemory<T> &>(gv).ptr());
}
template <typename T>
void ogsOperator_t::Scatter(deviceMemory<T> v, const deviceMemory<T> gv, const int k, const Transpose trans) {
  ScatterRaw(ogsType<T>::get(), trans, k, gv.ptr(), v.ptr());
}
template <typename T>
void ogsOperator_t::GatherScatter(deviceMemory<T> v, const int k, const Op op, const Transpose trans) {
  GatherScatterRaw(ogsType<T>::get(), op, trans, k, v.ptr());
}

#define OGS_OPERATOR_INSTANTIATE(T)                                                                          \
  template void ogsOperator_t::Gather(deviceMemory<T>, const deviceMemory<T>, const int, const Op, const Transpose); \
  template void ogsOperator_t::Scatter(deviceMemory<T>, const deviceMemory<T>, const int, const Transpose);       \
  template void ogsOperator_t::GatherScatter(deviceMemory<T>, const int, const Op, const Transpose);
OGS_OPERATOR_INSTANTIATE(float)
OGS_OPERATOR_INSTANTIATE(double)
OGS_OPERATOR_INSTANTIATE(int)
OGS_OPERATOR_INSTANTIATE(long long int)

}  // namespace ogs
}  // namespace libp
