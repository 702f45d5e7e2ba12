// The gather operator: a CSR with one row per gathered node
// (include/ogs/ogsOperator.hpp:36-138).  T-CSR holds every member of a group,
// N-CSR only the unflagged (positive) ones.  Rows are cut into row blocks of at
// most `blockNnz` column entries, the LDS tile of the gfx950 kernels.
#pragma once

#include <omp.h>

#include <vector>

#include "ogs.hpp"

namespace libp {
namespace ogs {

// Reference defaults (ogsOperator.hpp:37-38) and this build's default tiles,
// measured on MI355X at 64^3 hex p=7: CSR gather/scatter blocks of 2048
// entries, sorted-run tiles of 2048 positions.
constexpr int blockSize = 256;
constexpr int gatherNodesPerBlock = 512;
constexpr int defaultBlockNnz = 2048;
constexpr int defaultGsBlockNnz = 2048;
constexpr int maxBlockNnz = 4096;

// Sorted-run row blocks ("tiles") of the T-CSR (spx_srb_t): each tile's
// entries in ascending local index, described by cpt chunk records of 64
// positions (runs of consecutive indices, links from each row member to the
// previous one, the CSR row of the first row-leading position).
typedef spx_srb_chunk_t SrbChunkRecord;

struct SortedRunBlocks {
  dlong nblocks = 0, cpt = 0, nnz = 0, nchunks = 0, nruns = 0, nlinks = 0;
  // Row-leading positions number the rows in position order, so the gather
  // and scatter (which use row numbers) need each tile's rows to ascend by
  // their first member's local index.  True of every operator's T-CSR except
  // a Signed one without unique flagging, whose groups with no positive member
  // follow the others (ogsSetup.cpp:517-546); the gather+scatter, which needs
  // no row numbers, runs on any layout.
  bool rowsAscending = true;
  int xcdRemap = -1;  // launch hints passed in spx_srb_t (-1 / 0: library default)
  int tileLanes = 0;
  // non-temporal value access of the Sym gather+scatter (spx_srb_t.nontemporal):
  // ntMask bit e set by the builder when at least kSrbWholeLineShare of the
  // tiles' accesses to 128-byte lines of 2^e-byte nodes take whole lines;
  // ntHint -1 uses ntMask, 0 none, 1 every node size
  int ntMask = 0;
  int ntHint = -1;
  memory<dlong> entryStarts;  // host only: CSR entries of each tile
  memory<SrbChunkRecord> chunks;
  memory<dlong> runIds;
  memory<uint16_t> links;
  // N-CSR membership, one mask per chunk record (bit l: position l is an
  // N-CSR member), when built with member flags; empty otherwise
  memory<uint64_t> memberN;
  deviceMemory<SrbChunkRecord> o_chunks;
  deviceMemory<dlong> o_runIds;
  deviceMemory<uint16_t> o_links;
  deviceMemory<uint64_t> o_memberN;

  void Upload(platform_t &platform);
  spx_srb_t Device() const;
  const uint64_t *DeviceMemberN() const { return o_memberN.length() ? o_memberN.ptr() : nullptr; }
};

// chunk records per tile for a row block capacity (a power of two >= 4)
int SrbChunksPerTile(int cap);
// share of whole lines (tile, 128-byte line) among the tiles' line accesses
// from which a node size gets non-temporal values.  Measured on MI355X, FP64
// hex boxes at ~1e8 nodes, us per call with ordinary -> non-temporal values
// (share of 8-byte nodes): p=7 line tiles 310 -> 265 (1.00; 64^3: 407 ->
// 346), p=15 300 -> 282 (0.87), p=11 336 -> 374 (0.69-0.77), p=9, 10
// 349 / 343 -> 435 / 469 (0.57-0.59), p=1..6 1.2-1.8x slower (<= 0.36).
constexpr double kSrbWholeLineShare = 0.85;

// memberFlag (one byte per local node, optional): fills memberN.
SortedRunBlocks BuildSortedRunBlocks(const dlong Nrows, const memory<dlong> rowStarts, const memory<dlong> colIds,
                                     const dlong Nblocks, const memory<dlong> blockStarts, const int cap,
                                     const uint8_t *memberFlag = nullptr);

// Per local node (Ncols), 1 when it is an N-CSR member, if every N-CSR row r
// is T-CSR row r restricted to those members (same order) and T-CSR rows past
// the N-CSR's have none -- true of one rank's Setup (ogsOperator.cpp:621-643)
// and of a batch whose parts' N rows match their T rows; empty otherwise.
memory<uint8_t> SignedMemberFlags(const dlong Ncols, const dlong NrowsT, const memory<dlong> rowStartsT,
                                  const memory<dlong> colIdsT, const dlong NrowsN, const memory<dlong> rowStartsN,
                                  const memory<dlong> colIdsN);

// One launch shape of the Sym gather+scatter and its measured time.
struct GsTuning {
  int sorted;     // 1: sorted-run layout, 0: T-CSR blocks
  int blockNnz;   // row block capacity
  int tileLanes;  // sorted-run lanes per block (64 or 256)
  int xcdRemap;   // sorted-run XCD remap (0 or 1)
  double ms;
};

// One launch shape of the batched CG update (spx_cg_batch) and its time.
struct CgTuning {
  int targetBlocks;
  int unroll;
  int nontemporal;
  double ms;
};

// Times spx_cg_batch for plans of several sizes and unroll / streaming
// settings on the caller's vectors (alpha 0, so they keep their values);
// leaves the best setting in the "cg_batch" tuning and its plan in *plan.
CgTuning AutotuneCgBatch(const std::vector<int64_t> &segStarts, int storage, void *p, void *Ap, void *x, void *r,
                         void *rdotr, int reps, spx_cg_plan_t *plan, std::vector<CgTuning> *log = nullptr);



class ogsOperator_t {
 public:
  platform_t platform;

  dlong Ncols = 0;
  dlong NrowsN = 0;
  dlong NrowsT = 0;
  dlong nnzN = 0;
  dlong nnzT = 0;

  memory<dlong> rowStartsN, rowStartsT, colIdsN, colIdsT;
  deviceMemory<dlong> o_rowStartsN, o_rowStartsT, o_colIdsN, o_colIdsT;

  dlong blockNnz = defaultBlockNnz;  // LDS tile of gather / scatter (max column entries per block)
  dlong NrowBlocksN = 0, NrowBlocksT = 0;
  memory<dlong> blockRowStartsN, blockRowStartsT;
  deviceMemory<dlong> o_blockRowStartsN, o_blockRowStartsT;

  dlong gsBlockNnz = defaultGsBlockNnz;  // LDS tile of gather+scatter (row blocks of the T-CSR)
  dlong NgsBlocks = 0;
  memory<dlong> gsBlockRowStarts;
  deviceMemory<dlong> o_gsBlockRowStarts;

  // The gather+scatter runs on a sorted-run layout (gsSorted) or directly on
  // the CSR.  It needs no row numbers, so its tiles may take the rows in
  // another order (gsRowOrder, empty: the CSR's), e.g. a hex box's grid
  // points in small 3-D blocks, whose members then share elements and cache
  // lines; the gather and scatter keep the CSR's row order (rowSrb, the same
  // layout as gsSrb when gsRowOrder is empty).
  SortedRunBlocks gsSrb, rowSrb;
  memory<dlong> gsRowOrder;
  // gsTileCut[i] = 1: a gather+scatter tile starts at row gsRowOrder[i]
  // (empty: tiles are cut only by capacity).  A hex box's setup cuts tiles so
  // that each memory line of the vector is written by one tile.
  memory<uint8_t> gsTileCut;
  bool gsSorted = true;
  // The FP32 and BF16 gather+scatter (GatherScatterMixedRaw) may take a tile
  // order of their own: gsMixed[0] for SPX_STORE_F32, gsMixed[1] for
  // SPX_STORE_BF16 (own = false: the gsSrb tiles).  The launch hints are
  // gsSrb's.  A hex box's setup gives them line tiles, or lattice blocks
  // where FP64 takes line tiles, where those measured faster for the
  // narrower values (HexGsLineRowsMixed).
  struct MixedTiles {
    bool own = false;
    memory<dlong> order;
    memory<uint8_t> cuts;
    SortedRunBlocks srb;
  };
  MixedTiles gsMixed[2];
  // The T-CSR gather (Trans) and scatter (NoTrans) on the same layout or on
  // the T-CSR row blocks of blockNnz entries.  Measured on MI355X at 64^3 hex
  // p=7 (algorithmic TB/s): scatter 6.6 sorted against 4.3 on the CSR,
  // gather 6.8 against 5.6.
  bool gatherSorted = true;
  bool scatterSorted = true;
  // Signed operators whose N-CSR rows are the positive members of the same
  // T-CSR rows carry N-membership masks in gsSrb (signedSrb), and their
  // NoTrans / Trans gather+scatter, N-CSR gather and N-CSR scatter run on
  // the layout too (under gsSorted / gatherSorted / scatterSorted).
  bool signedSrb = false;
  memory<uint8_t> memberFlagN;  // per local node: N-CSR member (signedSrb only)

  // Local -> gathered row maps of the T- and N-CSR (-1: in no row), for the
  // scatter in local order (spx_ogs_scatter_inverse), whose stores are
  // coalesced whatever the rows.  It is the default when rows are short: hex
  // boxes at ~1e8 nodes, us per call (row blocks / local order): p=1 553 /
  // 236, p=3 325 / 291, p=5 328 / 288, but p=7 394 / 405 and p=15 326 / 347,
  // with rows / local nodes 0.13, 0.43, 0.58, 0.68 and 0.83.
  deviceMemory<dlong> o_localRowT, o_localRowN;
  static constexpr double kScatterByRowsMinRatio = 0.63;
  bool scatterByRows = true;  // false: scatter through the local -> row maps

  // The non-symmetric gather+scatter stages two value tiles, so its row
  // blocks hold at most min(blockNnz, 2048) entries.
  dlong genBlockNnz = defaultBlockNnz;
  dlong NgenBlocks = 0;
  deviceMemory<dlong> o_genBlockRowStarts;

  Kind kind = Signed;

  ogsOperator_t() = default;
  // CSR arrays are taken as built (rows sorted, columns ascending in a row).
  ogsOperator_t(platform_t &platform_, Kind kind_, const dlong NrowsN_, const dlong NrowsT_,
                const dlong Ncols_, memory<dlong> rowStartsT_, memory<dlong> colIdsT_,
                memory<dlong> rowStartsN_, memory<dlong> colIdsN_, int blockNnz_ = defaultBlockNnz,
                int gsBlockNnz_ = defaultGsBlockNnz);

  void Free();
  // Re-cut the row blocks for other LDS tile sizes (autotuner): gather and
  // scatter blocks, and gather+scatter blocks.
  void SetBlockNnz(int nnz);
  void SetGsBlockNnz(int nnz);
  // Rows of the gather+scatter tiles in this order (a permutation of the
  // T-CSR rows; empty: the CSR's); results do not change.  validate: check
  // that it is a permutation (callers that built one may skip it).
  // cuts: optional tile starts, one flag per entry of order (see gsTileCut).
  void SetGsRowOrder(const memory<dlong> &order, bool validate = true, const memory<uint8_t> &cuts = memory<uint8_t>());
  // gsSrb from rowSrb and gsRowOrder, with these launch hints
  void BuildGsSrb(int xcdRemap, int tileLanes);
  // Tiles of the FP32 (storage SPX_STORE_F32) or BF16 (SPX_STORE_BF16)
  // gather+scatter in this row order with these cuts, as SetGsRowOrder; an
  // empty order returns them to the gsSrb tiles.  Results do not change.
  void SetGsMixedRowOrder(int storage, const memory<dlong> &order, const memory<uint8_t> &cuts,
                          bool validate = true);
  // The T-CSR's rows in `order` (empty: the CSR's), cut into greedy blocks of
  // gsBlockNnz entries that also start where cuts[i] = 1, as sorted-run tiles
  // uploaded to the device.
  SortedRunBlocks BuildOrderedSrb(const memory<dlong> &order, const memory<uint8_t> &cuts);

  template <typename T> void Gather(deviceMemory<T> gv, const deviceMemory<T> v, const int k, const Op op, const Transpose trans);
  template <typename T> void Scatter(deviceMemory<T> v, const deviceMemory<T> gv, const int k, const Transpose trans);
  template <typename T> void GatherScatter(deviceMemory<T> v, const int k, const Op op, const Transpose trans);

  // type-erased forms used by the C API (pointers are device addresses)
  void GatherRaw(Type type, Op op, Transpose trans, int k, const void *v, void *gv);
  void ScatterRaw(Type type, Transpose trans, int k, const void *gv, void *v);
  void GatherScatterRaw(Type type, Op op, Transpose trans, int k, void *v);
  // Sym Add gatherScatter on FP32 / BF16 storage with FP64 row sums
  // (storage: SPX_STORE_*)
  void GatherScatterMixedRaw(int storage, int k, void *v);
  // Time every candidate launch shape of the Sym gather+scatter for values of
  // `type` (storage < 0) or of mixed storage `storage`, k fields; keep and
  // return the fastest if it beats the current shape by more than
  // 1 - kAutotuneMargin, else keep (and return) the current one.
  GsTuning Autotune(int type, int storage, int k, int reps, std::vector<GsTuning> *log = nullptr);
  // the candidates are timed on a scratch vector, and placement alone moved a
  // shape's time by up to 4 % between it and the caller's vector (Tet p=3:
  // 639 us on scratch, 640 on the data, against 625 for the current shape)
  static constexpr double kAutotuneMargin = 0.95;
};

// The local operator's CSR pair as built by the setup (host arrays).
struct LocalCSR {
  dlong NrowsT = 0, NrowsN = 0;
  dlong nnzT = 0, nnzN = 0;  // Unsigned: N aliases T and nnzN = nnzT (ogsOperator.cpp:644-650)
  memory<dlong> rowStartsT, colIdsT, rowStartsN, colIdsN;
  bool gather_defined = false;
};

// Host part of ogsBase_t::Setup on one rank: grouping, unique flagging
// (written back into ids), row numbering and both CSRs.
LocalCSR BuildLocalCSR(const dlong N, memory<hlong> ids, const Kind kind, const bool unique, const int rank);

// The same for the lexicographic GLL ids of an nx x ny x nz hex box at degree
// p, in closed form (no ids needed); idsOut, when given, receives the ids as
// BuildLocalCSR would leave them.
LocalCSR BuildHexBoxCSR(const dlong nx, const dlong ny, const dlong nz, const int p, const Kind kind,
                        const bool unique, const int rank, hlong *idsOut);

// Row order of a hex box's gather+scatter tiles: its grid points (i fastest)
// in blocks of bx x by x bz points, blocks in lexicographic order, rows in CSR
// order within a block.  rowGrid[r] is row r's lexicographic grid index on an
// Mx x My x Mz lattice.
memory<dlong> LatticeBlockOrder(const dlong Nrows, const memory<hlong> &rowGrid, const hlong Mx, const hlong My,
                                const hlong Mz, const int bx, const int by, const int bz);
// Stable counting sort of the rows by key[r] in [0, nkeys).
memory<dlong> CountingOrder(const dlong Nrows, const std::vector<int32_t> &key, const hlong nkeys);
// Block shape measured best on MI355X for a hex box of degree p (64^3 p=7 and
// ~1e8-node boxes at p = 1..15).
void HexGsBlockShape(const int p, int &bx, int &by, int &bz);
// Node rows per line tile a hex box of degree p uses for its gather+scatter
// tiles (HexLineTiles), or 0 for the lattice blocks of HexGsBlockShape.
int HexGsLineRows(const int p);
// Node rows per line tile of the FP32 (storage SPX_STORE_F32) or BF16
// (SPX_STORE_BF16) gather+scatter of a hex box of degree p, 0 for the tiles
// of the FP64 one, or -1 for the lattice blocks of HexGsBlockShape.
int HexGsLineRowsMixed(const int p, const int storage);
// Gather+scatter tiles of an nx x ny x nz hex box at degree p whose members
// fill whole memory lines of the vector: a line holds rowsPerLine node rows
// (fixed j, k) of one element, so a tile takes every row of a full x line of
// grid points for a set of y points that never splits such a group (the last
// rows of element ey and the first of ey+1 share a point and go together),
// in one z plane; groups are merged up to cap members per tile, and a group
// above cap is cut along x at element boundaries.  order: the rows, tile by
// tile; cuts: 1 where a tile starts (for SetGsRowOrder).
void HexLineTiles(const dlong Nrows, const memory<hlong> &rowGrid, const hlong nx, const hlong ny, const hlong nz,
                  const int p, const int rowsPerLine, const int cap, memory<dlong> &order, memory<uint8_t> &cuts);

// Parallel exclusive prefix sum of v[0..n) in place; returns the total.  The
// slices follow the team the runtime actually gives, which may be smaller than
// asked for (OMP_THREAD_LIMIT, OMP_DYNAMIC, nesting).
template <typename V>
inline V ExclusiveScan(V *v, dlong n) {
  if (n <= 0) return 0;
  int nt = 1, team = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  std::vector<V> part(nt + 1, 0);
#pragma omp parallel num_threads(nt)
  {
    int t = 0, T = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    T = omp_get_num_threads();
#endif
    if (t == 0) team = T;
    const dlong lo = (dlong)((int64_t)n * t / T), hi = (dlong)((int64_t)n * (t + 1) / T);
    V sum = 0;
    for (dlong i = lo; i < hi; ++i) sum += v[i];
    part[t + 1] = sum;
#pragma omp barrier
#pragma omp single
    for (int k = 0; k < T; ++k) part[k + 1] += part[k];
    V run = part[t];
    for (dlong i = lo; i < hi; ++i) {
      const V x = v[i];
      v[i] = run;
      run += x;
    }
  }
  return part[team];
}

// Greedy row blocking: consecutive rows, at most `cap` entries per block.
// Aborts when one row alone exceeds cap (ogsOperator.cpp:535-536).  With
// `cuts` (one flag per row), a block also starts at every flagged row.
void blockRowsGreedy(const dlong Nrows, const memory<dlong> rowStarts, const int cap, dlong &Nblocks,
                     memory<dlong> &blockStarts, const uint8_t *cuts = nullptr);

}  // namespace ogs
}  // namespace libp
