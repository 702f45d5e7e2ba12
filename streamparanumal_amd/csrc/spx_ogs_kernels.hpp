// ogs gather / scatter / gather+scatter kernels for gfx950.
//
// Same operator as libs/ogs/okl/ogsKernels.okl (a CSR with one row per
// gathered node and the local indices of its copies as columns), re-designed
// for CDNA:
//   * one workgroup (256 lanes) per row block; a block holds whole rows and at
//     most 256*PER column entries, staged through LDS sized at launch;
//   * phase 1 is straight-line code: each lane issues PER column-id loads,
//     PER+1 row-start loads and then PER value loads (indices clamped into the
//     block), so every load of the phase is in flight before the first use
//     and the wait counters stay exact;
//   * phase 2: one lane per row folds its entries in stored order (ascending
//     local index) out of LDS, so sums are bit-identical to the reference's
//     Serial order;
//   * phase 3: lanes write results back in CSR order;
//   * no atomics: every local node belongs to exactly one row and every row to
//     exactly one block, so the in-place update is race-free;
//   * workgroup ids are remapped so that each XCD takes blocks of consecutive
//     row blocks in turn (neighbouring blocks share cache lines at their
//     edges; xcd_tile).
// K fields per node (AoS, stride K) are processed together (KF = K <= 4) or
// one field per grid row (KF = 1, grid.y = K), like the reference's
// @outer(1) loop.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <type_traits>

#include "spx_internal.hpp"

#pragma clang fp contract(off)

namespace spx {
namespace ogsk {

constexpr int kThreads = 256;

struct Launch {
  int nblocks;
  int K;      // fields per node (stride)
  int nxcd;   // XCD remap width (1 = identity)
  int xcdLog2 = kXcdLog2;  // XCD remap: blocks of 2^xcdLog2 consecutive tiles
  int cap;    // LDS capacity in entries (>= every block's nnz)
  int kf;     // fields per pass
  int per;    // entries per lane (cap <= 256*per)
  int variant;  // Sym GS: 0 row ranges in registers, 1 row starts in LDS, 2 = 0 + nontemporal write-back,
                // 3 = 16-bit row starts in LDS
  hipStream_t stream;
  int *err;
  int rows = 0;  // sorted-run N-CSR gather / scatter: rows of the N-CSR
  int ntMask = 0;  // sorted-run gather+scatter: spx_srb_t.nontemporal
};

template <typename T> struct Lim;
template <> struct Lim<float> { static __device__ __forceinline__ float max() { return FLT_MAX; } };
template <> struct Lim<double> { static __device__ __forceinline__ double max() { return DBL_MAX; } };
template <> struct Lim<int> { static __device__ __forceinline__ int max() { return INT_MAX; } };
template <> struct Lim<long long> { static __device__ __forceinline__ long long max() { return LLONG_MAX; } };

// ogsOperator.cpp:37-56 (host) / ogsUtils.cpp:62-103 (device defines).
// Acc is the accumulator type, out() turns it into the stored result.
template <typename T, int OP> struct Fold;
template <typename T> struct Fold<T, SPX_ADD> {
  typedef T Acc;
  static __device__ __forceinline__ T init() { return T(0); }
  static __device__ __forceinline__ void apply(T &a, T b) { a += b; }
  static __device__ __forceinline__ T out(T a) { return a; }
};
template <typename T> struct Fold<T, SPX_MUL> {
  typedef T Acc;
  static __device__ __forceinline__ T init() { return T(1); }
  static __device__ __forceinline__ void apply(T &a, T b) { a *= b; }
  static __device__ __forceinline__ T out(T a) { return a; }
};
template <typename T> struct Fold<T, SPX_MAX> {
  typedef T Acc;
  static __device__ __forceinline__ T init() { return -Lim<T>::max(); }
  static __device__ __forceinline__ void apply(T &a, T b) { a = (b > a) ? b : a; }
  static __device__ __forceinline__ T out(T a) { return a; }
};
template <typename T> struct Fold<T, SPX_MIN> {
  typedef T Acc;
  static __device__ __forceinline__ T init() { return Lim<T>::max(); }
  static __device__ __forceinline__ void apply(T &a, T b) { a = (b < a) ? b : a; }
  static __device__ __forceinline__ T out(T a) { return a; }
};

// Non-temporal (streaming) load and store of one value; 2-byte wrapper types
// go through their bits.
template <typename T> __device__ __forceinline__ T nt_load(const T *p) {
  if constexpr (std::is_arithmetic<T>::value) {
    return __builtin_nontemporal_load(p);
  } else {
    static_assert(sizeof(T) == 2, "2-byte wrapper");
    const unsigned short b = __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p));
    T r;
    __builtin_memcpy(&r, &b, 2);
    return r;
  }
}
template <typename T> __device__ __forceinline__ void nt_store(T x, T *p) {
  if constexpr (std::is_arithmetic<T>::value) {
    __builtin_nontemporal_store(x, p);
  } else {
    static_assert(sizeof(T) == 2, "2-byte wrapper");
    unsigned short b;
    __builtin_memcpy(&b, &x, 2);
    __builtin_nontemporal_store(b, reinterpret_cast<unsigned short *>(p));
  }
}

__device__ __forceinline__ void flag_error(int *err, int code) {
  if (threadIdx.x == 0) err[0] = code;
}

__device__ __forceinline__ int clampi(int x, int hi) { return x < hi ? x : hi; }


// Column ids of entries t, t+256, ... (clamped to the block's last entry).
template <int PER>
__device__ __forceinline__ void load_ids(int (&cid)[PER], const int *__restrict__ ids, int g0, int nnz) {
  const int last = nnz > 0 ? nnz - 1 : 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) cid[j] = ids[g0 + clampi(threadIdx.x + j * kThreads, last)];
}

// Relative row starts rs[r] = starts[r0 + min(r, rlim)] - base for r in [0, nrows],
// staged through registers so the loads overlap the id loads.
template <int PER>
__device__ __forceinline__ void stage_starts(int *rs, const int *__restrict__ starts, int r0, int nrows,
                                             int rlim, int base) {
  int x[PER + 1];
#pragma unroll
  for (int j = 0; j <= PER; ++j) x[j] = starts[r0 + clampi(clampi(threadIdx.x + j * kThreads, nrows), rlim)];
#pragma unroll
  for (int j = 0; j <= PER; ++j)
    if (threadIdx.x + j * kThreads <= nrows) rs[threadIdx.x + j * kThreads] = x[j] - base;
}

// As stage_starts, into 16-bit LDS offsets (a block holds at most 4096 entries).
template <int PER>
__device__ __forceinline__ void stage_starts16(unsigned short *rs, const int *__restrict__ starts, int r0,
                                               int nrows, int base) {
  int x[PER + 1];
#pragma unroll
  for (int j = 0; j <= PER; ++j) x[j] = starts[r0 + clampi(threadIdx.x + j * kThreads, nrows)];
#pragma unroll
  for (int j = 0; j <= PER; ++j)
    if (threadIdx.x + j * kThreads <= nrows) rs[threadIdx.x + j * kThreads] = (unsigned short)(x[j] - base);
}

// Row ranges of the rows a lane folds (rows t, t+256, ...), relative to g0,
// loaded straight into registers: ra[j] = start, rb[j] = end.
template <int PER>
__device__ __forceinline__ void load_row_ranges(int (&ra)[PER], int (&rb)[PER], const int *__restrict__ starts,
                                                int r0, int nrows, int g0) {
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int r = r0 + clampi(threadIdx.x + j * kThreads, nrows);
    ra[j] = starts[r] - g0;
    rb[j] = starts[clampi(r + 1, r0 + nrows)] - g0;
  }
}

// ---------------------------------------------------------------- Sym GS --
// gatherScatter with gStarts == sStarts, gIds == sIds (ogs::Sym).
// LDS: vals[cap*KF] of T (+ rs[cap+1] ints for VAR 1).
//   VAR 0: each lane keeps the ranges of its rows in registers;
//   VAR 1: the block's row starts are staged in LDS;
//   VAR 2: as 0, with non-temporal write-back of the results;
//   VAR 3: as 1, with 16-bit row starts.
template <typename T, int OP, int KF, int PER, int VAR>
__global__ __launch_bounds__(kThreads) void k_gs_sym(Launch L, const int *__restrict__ blockStarts,
                                                     const int *__restrict__ starts,
                                                     const int *__restrict__ ids, T *__restrict__ q) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *vals = reinterpret_cast<T *>(lds_raw);
  int *rs = reinterpret_cast<int *>(vals + (size_t)L.cap * KF);
  unsigned short *rs16 = reinterpret_cast<unsigned short *>(rs);
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd, L.xcdLog2);
  const int kb = blockIdx.y * KF;
  const int r0 = blockStarts[b], r1 = blockStarts[b + 1];
  const int nrows = r1 - r0;
  const int g0 = starts[r0];
  const int nnz = starts[r1] - g0;
  if (nnz > L.cap || nrows > L.cap) {
    flag_error(L.err, 1);
    return;
  }
  if (nnz == 0) return;
  const int64_t K = L.K;
  const int t = threadIdx.x;
  int cid[PER], ra[PER], rb[PER];
  load_ids<PER>(cid, ids, g0, nnz);
  if constexpr (VAR == 1) stage_starts<PER>(rs, starts, r0, nrows, nrows, g0);
  else if constexpr (VAR == 3) stage_starts16<PER>(rs16, starts, r0, nrows, g0);
  else load_row_ranges<PER>(ra, rb, starts, r0, nrows, g0);
  T v[PER][KF];
#pragma unroll
  for (int j = 0; j < PER; ++j)
#pragma unroll
    for (int f = 0; f < KF; ++f) v[j][f] = q[(int64_t)cid[j] * K + kb + f];
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t + j * kThreads < nnz)
#pragma unroll
      for (int f = 0; f < KF; ++f) vals[(t + j * kThreads) * KF + f] = v[j][f];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t + j * kThreads < nrows) {
      int a, e;
      if constexpr (VAR == 1) {
        a = rs[t + j * kThreads];
        e = rs[t + j * kThreads + 1];
      } else if constexpr (VAR == 3) {
        a = rs16[t + j * kThreads];
        e = rs16[t + j * kThreads + 1];
      } else {
        a = ra[j];
        e = rb[j];
      }
#pragma unroll
      for (int f = 0; f < KF; ++f) {
        T acc = Fold<T, OP>::init();
        for (int i = a; i < e; ++i) Fold<T, OP>::apply(acc, vals[i * KF + f]);
        for (int i = a; i < e; ++i) vals[i * KF + f] = acc;
      }
    }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t + j * kThreads < nnz)
#pragma unroll
      for (int f = 0; f < KF; ++f) {
        T *dst = q + (int64_t)cid[j] * K + kb + f;
        const T x = vals[(t + j * kThreads) * KF + f];
        if constexpr (VAR == 2) __builtin_nontemporal_store(x, dst);
        else *dst = x;
      }
}

// ----------------------------------------------------- sorted-run tiles --
// Tile b (a row block) lists its entries in ascending local index
// ("positions") and owns the chunk records [b*cpt, (b+1)*cpt), record k for
// positions 64k..64k+63, the first nvalid of which are entries.  Position
// l of record C decodes as
//   local index = runIds[C.runBase + popc((C.runMask >> 1) & bits<l)] + l
//     (each run's entry is its first local index less its first position);
//   previous member of its row (bit l of C.linkMask set)
//              = links[C.linkBase + popc(C.linkMask & bits<l)];
//   row of a row-leading position = C.chunkRow + (l - popc(C.linkMask & bits<l)).
// Records are found from the tile index alone, so the first loads of a tile
// need no header.  Values sit in LDS at their positions; next[] links each
// member to the following one, and the first member of each row folds the
// row along them (ascending local index, the CSR order).  NT = 64 gives
// one-wave tiles: more independent tiles per CU, no cross-wave barrier.
struct SrbChunk {
  unsigned long long runMask, linkMask;
  int runBase, linkBase, chunkRow, nvalid;
};
static_assert(sizeof(SrbChunk) == 32, "chunk record layout");

struct SrbArgs {
  const SrbChunk *chunks;
  const int *runIds;
  const unsigned short *links;
};

constexpr unsigned short kEnd = 0xffff;

// The layout arrays as restrict kernel arguments: they never alias the value
// arrays the kernels store to, so the compiler may keep their loads ahead of
// those stores.
#define SRB_PARAMS \
  const SrbChunk *__restrict__ chunks, const int *__restrict__ runIds, const unsigned short *__restrict__ links
#define SRB_ARGS(A) A.chunks, A.runIds, A.links

// Positions j*NT + threadIdx.x of tile b: local index, previous member (-1:
// row-leading or not an entry), CSR row (row-leading positions), and validity
// bits.  Chunk records past the tile's cpt are replaced by its last one and
// their positions marked invalid, so every load stays inside the layout.
template <int PER, int NT>
__device__ __forceinline__ void srb_decode(const SrbChunk *__restrict__ chunks, const int *__restrict__ runIds,
                                           const unsigned short *__restrict__ links, int b, int cpt,
                                           int (&id)[PER], int (&prev)[PER], int (&row)[PER], unsigned &valid) {
  const int t = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int l = t & 63;
  // The wave's PER records are 8*PER dwords: lane l fetches dword l%8 of
  // record l/8 (+8 per extra round) in one coalesced load; the fields then
  // reach every lane by readlane.  (Scalar loads of the records would return
  // out of order, and each use would wait for all of them.)
  constexpr int ROUNDS = (PER * 8 + 63) / 64;
  const int *words = reinterpret_cast<const int *>(chunks);
  int word[ROUNDS];
#pragma unroll
  for (int h = 0; h < ROUNDS; ++h) {
    const int jr = min(h * 8 + (l >> 3), PER - 1);
    const int kc = min((NT / 64) * jr + w, cpt - 1);
    word[h] = words[((int64_t)b * cpt + kc) * 8 + (l & 7)];
  }
  valid = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int k = (NT / 64) * j + w;  // wave-uniform
    const int src = (j & 7) * 8;
    const int wv = word[j / 8];
    // the record's fields, wave-uniform: the masks act on lane l as position l
    const unsigned rmLo = __builtin_amdgcn_readlane(wv, src + 0), rmHi = __builtin_amdgcn_readlane(wv, src + 1);
    const unsigned lmLo = __builtin_amdgcn_readlane(wv, src + 2), lmHi = __builtin_amdgcn_readlane(wv, src + 3);
    const int runBase = __builtin_amdgcn_readlane(wv, src + 4);
    const int linkBase = __builtin_amdgcn_readlane(wv, src + 5);
    const int chunkRow = __builtin_amdgcn_readlane(wv, src + 6);
    const int nraw = min(__builtin_amdgcn_readlane(wv, src + 7), 64);
    const bool ok = k < cpt && l < nraw;
    // run of position l: the run starts at positions 1..l (runMask >> 1 below
    // lane l); a lane past the entries takes the last entry, so its address
    // stays inside the run table and the vector
    const unsigned long long rs = ((unsigned long long)rmHi << 32 | rmLo) >> 1;
    const int run = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(rs >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)rs, 0u));
    const int ls = min(l, max(nraw - 1, 0));
    id[j] = runIds[runBase + run] + ls;
    const unsigned long long lm = (unsigned long long)lmHi << 32 | lmLo;
    const int before = (int)__builtin_amdgcn_mbcnt_hi(lmHi, __builtin_amdgcn_mbcnt_lo(lmLo, 0u));
    const bool linked = ok && __builtin_amdgcn_inverse_ballot_w64(lm);
    prev[j] = linked ? (int)links[linkBase + before] : -1;
    row[j] = chunkRow + (ls - before);
    valid |= (ok ? 1u : 0u) << j;
  }
}

// Which members of a row take part.  An ogs::Signed operator's N-CSR row r
// holds the members of T-CSR row r whose id is positive, in the same order
// (ogsOperator.cpp:621-643), so on a layout of the T-CSR whose positions carry
// an N-membership bit: kAll uses every member (Sym; the T-CSR gather and
// scatter), kFoldN folds over the N members only (NoTrans gather+scatter, the
// N-CSR gather) and kStoreN stores to the N members only (Trans
// gather+scatter, the N-CSR scatter).  A row without N members folds to the
// op's identity.
enum { kAll = 0, kFoldN = 1, kStoreN = 2 };

// Bit j: position j*NT + threadIdx.x of tile b is an N member, read from
// memberN[record] (bit l: position l of the record).  Lane l loads dword l%2
// of the mask of record l/2 (one coalesced load for up to 32 records); the
// masks then reach every lane by readlane.  Records past the tile's cpt are
// replaced by its last one, as in srb_decode, whose validity bits mask them.
template <int PER, int NT>
__device__ __forceinline__ unsigned srb_members(const unsigned long long *__restrict__ memberN, int b, int cpt) {
  static_assert(PER <= 32, "one load covers 32 records");
  const int t = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int l = t & 63;
  const unsigned *words = reinterpret_cast<const unsigned *>(memberN);
  const int jr = min(l >> 1, PER - 1);
  const int kc = min((NT / 64) * jr + w, cpt - 1);
  const unsigned word = words[((int64_t)b * cpt + kc) * 2 + (l & 1)];
  unsigned bits = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)word, 2 * j);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)word, 2 * j + 1);
    bits |= (((l < 32 ? lo : hi) >> (l & 31)) & 1u) << j;
  }
  return bits;
}

// Values a sorted-run gather+scatter tile keeps in LDS per position: every
// field for 2-byte values (word-major, below), one field at a time otherwise.
template <typename T> constexpr int srb_lds_fields(int kf) { return (kf > 1 && sizeof(T) == 2) ? kf : 1; }

// LDS of a sorted-run gather+scatter tile of cap positions.  A 2- or 4-byte
// value folded one field at a time shares an entry with its position's next
// link (4 bytes: value, u16 link; 8 bytes: value, u32 link), so a step of a
// row walk is one LDS access at one address; 8-byte values and the
// word-major fields keep a value array and a u16 link array.
template <typename T> constexpr bool srb_gs_packed(int kf) { return srb_lds_fields<T>(kf) == 1 && sizeof(T) <= 4; }
template <typename T> constexpr size_t srb_gs_lds_bytes(int cap, int kf) {
  return srb_gs_packed<T>(kf) ? (size_t)cap * (sizeof(T) == 2 ? 4 : 8)
                              : (size_t)cap * srb_lds_fields<T>(kf) * sizeof(T) + (size_t)cap * sizeof(unsigned short);
}

template <typename T, int LF, bool PACK> struct GsLds;
template <typename T, int LF> struct GsLds<T, LF, false> {
  T *vals;
  unsigned short *next;
  __device__ GsLds(unsigned char *raw, int cap)
      : vals(reinterpret_cast<T *>(raw)), next(reinterpret_cast<unsigned short *>(vals + (size_t)cap * LF)) {}
  __device__ void set_next(int i, int n) const { next[i] = (unsigned short)n; }
  __device__ int get_next(int i) const { return next[i]; }
  __device__ void set_val(int i, int f, T x) const { vals[i * LF + f] = x; }
  __device__ T get_val(int i, int f) const { return vals[i * LF + f]; }
  __device__ T get_both(int i, int f, int &n) const {
    n = next[i];
    return vals[i * LF + f];
  }
};
template <typename T> struct GsLds<T, 1, true> {
  static_assert(sizeof(T) == 2 || sizeof(T) == 4, "packed entries hold 2- or 4-byte values");
  typedef typename std::conditional<sizeof(T) == 2, unsigned short, unsigned>::type Link;
  struct alignas(2 * sizeof(T)) Entry {
    T val;
    Link next;
  };
  Entry *e;
  __device__ GsLds(unsigned char *raw, int) : e(reinterpret_cast<Entry *>(raw)) {}
  // the next link is the low 16 bits; a 32-bit link word can also carry the
  // row's head position in its high 16 bits (set_link_head / get_head)
  __device__ void set_next(int i, int n) const { e[i].next = (Link)n; }
  __device__ int get_next(int i) const { return (int)(e[i].next & 0xffffu); }
  __device__ void set_val(int i, int, T x) const { e[i].val = x; }
  __device__ T get_val(int i, int) const { return e[i].val; }
  __device__ T get_both(int i, int, int &n) const {
    const Entry x = e[i];  // one access
    n = (int)(x.next & 0xffffu);
    return x.val;
  }
  __device__ void set_link_head(int i, int n, int h) const { e[i].next = (Link)((unsigned)n | ((unsigned)h << 16)); }
  __device__ int get_head(int i) const { return (int)((unsigned)e[i].next >> 16); }
};

template <typename T, int OP, int KF, int PER, int NT = kThreads, typename F = Fold<T, OP>, int MODE = kAll,
          bool NTV = false>
__global__ __launch_bounds__(NT) void k_gs_srb(Launch L, SRB_PARAMS, T *__restrict__ q,
                                               const unsigned long long *__restrict__ memberN) {
  // Several 4- or 8-byte fields per pass: each lane loads and stores its
  // position's KF values as one wide access (the compiler merges them) and
  // keeps them in registers, and the fields pass through LDS one at a time,
  // so a tile needs the LDS of one field whatever KF (FP64, K = 3 on
  // 2048-position tiles: 20 KB instead of 52 KB).
  //
  // Several 2-byte fields per pass go word-major instead: the wave's chunk j
  // holds 64*KF values (KF fields of 64 positions, contiguous per position);
  // access f takes words 64f .. 64f+63 of them, lane l the word w = 64f + l =
  // field w % KF of position w / KF, whose local index comes from that
  // position's lane.  One access then spans a run's consecutive words instead
  // of KF separate 2-byte accesses per lane (BF16, K = 3: 13 % faster), and
  // the LDS slot of word w is the chunk's base slot + w.
  //
  // NTV: the values are loaded and stored non-temporally.  Each value is
  // read and written once per call, so where the tiles' runs cover whole
  // 128-byte lines, streaming them keeps the caches (L2 and the Infinity
  // Cache) for the layout tables, which every call reads again; where lines
  // are shared between tiles it would cost their reuse instead.
  //
  // MODE kFoldN / kStoreN (memberN given): positions outside the N-CSR hold
  // the op's identity in the fold (an exact no-op: the sum starts from +0 and
  // so is never -0, x*1 == x, and the max / min never pass -/+ the type's
  // largest value) or keep their value.
  constexpr bool kWord = KF > 1 && sizeof(T) == 2;
  static_assert(MODE == kAll || (!kWord && std::is_same<typename F::Acc, T>::value),
                "N-CSR forms fold plain 4- and 8-byte values");
  static_assert(!NTV || (MODE == kAll && !kWord), "non-temporal values: Sym form, value-major fields");
  constexpr int LF = kWord ? KF : 1;  // fields held in LDS per position
  // several 4-byte fields: heads kept beside the links (one walk per field)
  constexpr bool kHeadLink = KF > 1 && srb_gs_packed<T>(KF) && sizeof(T) == 4;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int cap = L.cap;  // cpt * 64 positions
  const GsLds<T, LF, srb_gs_packed<T>(KF)> S(lds_raw, cap);
  T *vals = reinterpret_cast<T *>(lds_raw);  // word-major fields (kWord)
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd, L.xcdLog2);
  const int kb = blockIdx.y * KF;
  const int64_t K = L.K;
  const int t = threadIdx.x;
  if (cap >= PER * NT) {  // every slot a position (the usual tile)
#pragma unroll
    for (int j = 0; j < PER; ++j) S.set_next(t + j * NT, kEnd);
  } else {
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (t + j * NT < cap) S.set_next(t + j * NT, kEnd);
  }
  int id[PER], prev[PER], row[PER];
  unsigned valid;
  srb_decode<PER, NT>(chunks, runIds, links, b, cap / 64, id, prev, row, valid);
  (void)row;
  // positions whose value enters the fold (ld) and that store the result (st)
  unsigned ld = valid, st = valid;
  if constexpr (MODE != kAll) {
    const unsigned m = valid & srb_members<PER, NT>(memberN, b, cap / 64);
    if constexpr (MODE == kFoldN) ld = m;
    else st = m;
  }
  (void)ld;
  const int l = t & 63;
  const int cb0 = t - l;  // first position of this wave's chunk 0
  T v[PER][KF];
#pragma unroll
  for (int j = 0; j < PER; ++j)
#pragma unroll
    for (int f = 0; f < KF; ++f) {
      if constexpr (!kWord) {
        if constexpr (MODE == kFoldN) v[j][f] = ((ld >> j) & 1u) ? q[(int64_t)id[j] * K + kb + f] : F::init();
        else if constexpr (NTV) v[j][f] = nt_load(&q[(int64_t)id[j] * K + kb + f]);
        else v[j][f] = q[(int64_t)id[j] * K + kb + f];
      } else {
        const int w = f * 64 + l, src = w / KF;
        const int idw = __shfl(id[j], src, 64);
        v[j][f] = q[(int64_t)idw * K + kb + (w - src * KF)];
      }
    }
  __syncthreads();  // links initialised before they are set
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (((valid >> j) & 1u) && (unsigned)prev[j] < (unsigned)cap) S.set_next(prev[j], t + j * NT);
  // one LDS round per field (one round in all for KF == 1 and word-major)
  constexpr int ROUNDS = kWord ? 1 : KF;
#pragma unroll
  for (int f0 = 0; f0 < ROUNDS; ++f0) {
    if (f0 > 0) __syncthreads();  // the previous field has been read back
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      if constexpr (!kWord) {
        if ((valid >> j) & 1u) S.set_val(t + j * NT, 0, v[j][f0]);
      } else {
        const unsigned long long vm = __ballot((valid >> j) & 1u);
#pragma unroll
        for (int f = 0; f < KF; ++f) {
          const int w = f * 64 + l;
          if ((vm >> (w / KF)) & 1ull) vals[(cb0 + j * NT) * KF + w] = v[j][f];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (((valid >> j) & 1u) && prev[j] < 0) {
        const int h = t + j * NT;
        if constexpr (KF == 1) {
          // One walk: each member's link, once read, is replaced by the head's
          // position, and the head keeps the row value; after the barrier
          // every member takes the value from its head in one step instead
          // of a second walk writing it along the row.
          typename F::Acc acc = F::init();
          // walks end at kEnd (> cap); the step bound also ends them on a malformed layout
          for (int i = h, s = 0; i < cap && s < cap; ++s) {
            int n;
            F::apply(acc, S.get_both(i, 0, n));
            S.set_next(i, h);
            i = n;
          }
          S.set_val(h, 0, F::out(acc));
        } else if constexpr (kHeadLink) {
          // several fields, one per round: the first round's walk also
          // writes the head's position beside each member's link, and every
          // round's members read the row value from the head
          typename F::Acc acc = F::init();
          for (int i = h, s = 0; i < cap && s < cap; ++s) {
            int n;
            F::apply(acc, S.get_both(i, 0, n));
            if (f0 == 0) S.set_link_head(i, n, h);
            i = n;
          }
          S.set_val(h, 0, F::out(acc));
        } else {
#pragma unroll
          for (int f = 0; f < LF; ++f) {
            typename F::Acc acc = F::init();
            for (int i = h, s = 0; i < cap && s < cap; ++s) {
              int n;
              F::apply(acc, S.get_both(i, f, n));
              i = n;
            }
            const T r = F::out(acc);
            for (int i = h, s = 0; i < cap && s < cap; ++s) {
              S.set_val(i, f, r);
              i = S.get_next(i);
            }
          }
        }
      }
    __syncthreads();
    if constexpr (kHeadLink) {
#pragma unroll
      for (int j = 0; j < PER; ++j)
        if ((valid >> j) & 1u) {
          const int hd = S.get_head(t + j * NT);
          v[j][f0] = S.get_val((unsigned)hd < (unsigned)cap ? hd : t + j * NT, 0);
        }
    } else if constexpr (!kWord && KF > 1) {
#pragma unroll
      for (int j = 0; j < PER; ++j)
        if ((valid >> j) & 1u) v[j][f0] = S.get_val(t + j * NT, 0);
    }
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    if constexpr (KF == 1) {
      // the row value from the head (a position never walked, only on a
      // malformed layout, keeps its own)
      if ((st >> j) & 1u) {
        const int hd = S.get_next(t + j * NT);
        const T r = S.get_val((unsigned)hd < (unsigned)cap ? hd : t + j * NT, 0);
        if constexpr (NTV) nt_store(r, &q[(int64_t)id[j] * K + kb]);
        else q[(int64_t)id[j] * K + kb] = r;
      }
    } else if constexpr (!kWord) {
      if ((st >> j) & 1u)
#pragma unroll
        for (int f = 0; f < KF; ++f) {
          if constexpr (NTV) nt_store(v[j][f], &q[(int64_t)id[j] * K + kb + f]);
          else q[(int64_t)id[j] * K + kb + f] = v[j][f];
        }
    } else {
      const unsigned long long vm = __ballot((valid >> j) & 1u);
#pragma unroll
      for (int f = 0; f < KF; ++f) {
        const int w = f * 64 + l, src = w / KF;
        const int idw = __shfl(id[j], src, 64);  // every lane takes part in the shuffle
        if ((vm >> src) & 1ull) q[(int64_t)idw * K + kb + (w - src * KF)] = vals[(cb0 + j * NT) * KF + w];
      }
    }
  }
}

// gather (Trans) on the sorted-run layout: row-leading positions fold their
// row along the links and write gv at the row's CSR index.  Several fields
// pass through LDS one at a time (one field's LDS whatever KF); each lane
// loads its position's fields, and a row-leading lane stores its row's
// fields, as one wide access.  MODE kFoldN gathers the N-CSR (NoTrans): only
// N members load (the others hold the op's identity) and only rows below
// L.rows store.
template <typename T, int OP, int KF, int PER, int NT = kThreads, int MODE = kAll>
__global__ __launch_bounds__(NT) void k_gather_srb(Launch L, SRB_PARAMS, const T *__restrict__ v,
                                                   T *__restrict__ gv,
                                                   const unsigned long long *__restrict__ memberN) {
  static_assert(MODE == kAll || MODE == kFoldN, "gather modes");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *vals = reinterpret_cast<T *>(lds_raw);
  unsigned short *next = reinterpret_cast<unsigned short *>(vals + (size_t)L.cap);
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd, L.xcdLog2);
  const int kb = blockIdx.y * KF;
  const int cap = L.cap;
  const int64_t K = L.K;
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t + j * NT < cap) next[t + j * NT] = kEnd;
  int id[PER], prev[PER], row[PER];
  unsigned valid;
  srb_decode<PER, NT>(chunks, runIds, links, b, cap / 64, id, prev, row, valid);
  unsigned ld = valid;
  if constexpr (MODE == kFoldN) ld = valid & srb_members<PER, NT>(memberN, b, cap / 64);
  (void)ld;
  T x[PER][KF];
#pragma unroll
  for (int j = 0; j < PER; ++j)
#pragma unroll
    for (int f = 0; f < KF; ++f) {
      if constexpr (MODE == kFoldN) x[j][f] = ((ld >> j) & 1u) ? v[(int64_t)id[j] * K + kb + f] : Fold<T, OP>::init();
      else x[j][f] = v[(int64_t)id[j] * K + kb + f];
    }
  __syncthreads();  // next[] initialised before the links go in
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (((valid >> j) & 1u) && prev[j] >= 0 && prev[j] < cap) next[prev[j]] = (unsigned short)(t + j * NT);
#pragma unroll
  for (int f = 0; f < KF; ++f) {
    if (f > 0) __syncthreads();  // the previous field's folds are done
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if ((valid >> j) & 1u) vals[t + j * NT] = x[j][f];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (((valid >> j) & 1u) && prev[j] < 0) {
        T acc = Fold<T, OP>::init();
        for (int i = t + j * NT, s = 0; i < cap && s < cap; i = next[i], ++s) Fold<T, OP>::apply(acc, vals[i]);
        x[j][f] = acc;  // a row-leading lane's value of field f is no longer needed
      }
  }
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (((valid >> j) & 1u) && prev[j] < 0 && (MODE == kAll || row[j] < L.rows))
#pragma unroll
      for (int f = 0; f < KF; ++f) gv[(int64_t)row[j] * K + kb + f] = x[j][f];
}

// scatter (NoTrans) on the sorted-run layout: row-leading positions read
// their row's value and pass it along the links; every position stores.
// Several fields pass through LDS one at a time, as in the gather.  MODE
// kStoreN scatters the N-CSR (Trans): only rows below L.rows read and only N
// members store.
template <typename T, int KF, int PER, int NT = kThreads, int MODE = kAll>
__global__ __launch_bounds__(NT) void k_scatter_srb(Launch L, SRB_PARAMS, const T *__restrict__ gv,
                                                    T *__restrict__ v,
                                                    const unsigned long long *__restrict__ memberN) {
  static_assert(MODE == kAll || MODE == kStoreN, "scatter modes");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *vals = reinterpret_cast<T *>(lds_raw);
  unsigned short *next = reinterpret_cast<unsigned short *>(vals + (size_t)L.cap);
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd, L.xcdLog2);
  const int kb = blockIdx.y * KF;
  const int cap = L.cap;
  const int64_t K = L.K;
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t + j * NT < cap) next[t + j * NT] = kEnd;
  int id[PER], prev[PER], row[PER];
  unsigned valid;
  srb_decode<PER, NT>(chunks, runIds, links, b, cap / 64, id, prev, row, valid);
  unsigned st = valid;
  if constexpr (MODE == kStoreN) st = valid & srb_members<PER, NT>(memberN, b, cap / 64);
  // only row-leading positions read (the row index of the others may lie
  // one past the last row)
  T g[PER][KF];
#pragma unroll
  for (int j = 0; j < PER; ++j)
#pragma unroll
    for (int f = 0; f < KF; ++f)
      g[j][f] = (((valid >> j) & 1u) && prev[j] < 0 && (MODE == kAll || row[j] < L.rows))
                    ? gv[(int64_t)row[j] * K + kb + f]
                    : T(0);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (((valid >> j) & 1u) && prev[j] >= 0 && prev[j] < cap) next[prev[j]] = (unsigned short)(t + j * NT);
#pragma unroll
  for (int f = 0; f < KF; ++f) {
    __syncthreads();  // links in (f == 0) / the previous field read back
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (((valid >> j) & 1u) && prev[j] < 0)
        for (int i = t + j * NT, s = 0; i < cap && s < cap; i = next[i], ++s) vals[i] = g[j][f];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if ((valid >> j) & 1u) g[j][f] = vals[t + j * NT];  // every position's value of field f
  }
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if ((st >> j) & 1u)
#pragma unroll
      for (int f = 0; f < KF; ++f) v[(int64_t)id[j] * K + kb + f] = g[j][f];
}

// ------------------------------------------------------------ general GS --
// gather CSR (gStarts/gIds, ngrows rows) and scatter CSR (sStarts/sIds,
// nsrows rows) differ (ogs::NoTrans, ogs::Trans).  Rows at or past ngrows
// gather nothing (their value is the op's identity); rows at or past nsrows
// scatter nothing.  LDS: gvals[cap*KF], svals[cap*KF], grs[cap+1], srs[cap+1].
template <typename T, int OP, int KF, int PER>
__global__ __launch_bounds__(kThreads) void k_gs_gen(Launch L, int ngrows, int nsrows,
                                                     const int *__restrict__ blockStarts,
                                                     const int *__restrict__ gStarts,
                                                     const int *__restrict__ gIds,
                                                     const int *__restrict__ sStarts,
                                                     const int *__restrict__ sIds, T *__restrict__ q) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *gvals = reinterpret_cast<T *>(lds_raw);
  T *svals = gvals + (size_t)L.cap * KF;
  int *grs = reinterpret_cast<int *>(svals + (size_t)L.cap * KF);
  int *srs = grs + L.cap + 1;
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd, L.xcdLog2);
  const int kb = blockIdx.y * KF;
  const int r0 = blockStarts[b], r1 = blockStarts[b + 1];
  const int nrows = r1 - r0;
  const int gr0 = clampi(r0, ngrows), gr1 = clampi(r1, ngrows);
  const int sr0 = clampi(r0, nsrows), sr1 = clampi(r1, nsrows);
  const int g0 = gStarts[gr0];
  const int gnnz = gStarts[gr1] - g0;
  const int s0 = sStarts[sr0];
  const int snnz = sStarts[sr1] - s0;
  if (gnnz > L.cap || snnz > L.cap || nrows > L.cap) {
    flag_error(L.err, 2);
    return;
  }
  const int64_t K = L.K;
  const int t = threadIdx.x;
  // rows past a CSR's end read that CSR's last start: empty ranges
  if (gr0 == r0) stage_starts<PER>(grs, gStarts, r0, nrows, gr1 - r0, g0);
  else
    for (int r = t; r <= nrows; r += kThreads) grs[r] = 0;
  if (sr0 == r0) stage_starts<PER>(srs, sStarts, r0, nrows, sr1 - r0, s0);
  else
    for (int r = t; r <= nrows; r += kThreads) srs[r] = 0;
  if (gnnz > 0) {
    int cid[PER];
    load_ids<PER>(cid, gIds, g0, gnnz);
    T v[PER][KF];
#pragma unroll
    for (int j = 0; j < PER; ++j)
#pragma unroll
      for (int f = 0; f < KF; ++f) v[j][f] = q[(int64_t)cid[j] * K + kb + f];
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (t + j * kThreads < gnnz)
#pragma unroll
        for (int f = 0; f < KF; ++f) gvals[(t + j * kThreads) * KF + f] = v[j][f];
  }
  __syncthreads();
  for (int r = t; r < nrows; r += kThreads) {
    const int a = grs[r], e = grs[r + 1], sa = srs[r], se = srs[r + 1];
#pragma unroll
    for (int f = 0; f < KF; ++f) {
      T acc = Fold<T, OP>::init();
      for (int i = a; i < e; ++i) Fold<T, OP>::apply(acc, gvals[i * KF + f]);
      for (int i = sa; i < se; ++i) svals[i * KF + f] = acc;
    }
  }
  __syncthreads();
  if (snnz > 0) {
    int cid[PER];
    load_ids<PER>(cid, sIds, s0, snnz);
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (t + j * kThreads < snnz)
#pragma unroll
        for (int f = 0; f < KF; ++f) q[(int64_t)cid[j] * K + kb + f] = svals[(t + j * kThreads) * KF + f];
  }
}

// ---------------------------------------------------------------- gather --
// LDS: vals[cap*KF].  Row ranges live in registers.
template <typename T, int OP, int KF, int PER>
__global__ __launch_bounds__(kThreads) void k_gather(Launch L, const int *__restrict__ blockStarts,
                                                     const int *__restrict__ starts,
                                                     const int *__restrict__ ids,
                                                     const T *__restrict__ v, T *__restrict__ gv) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *vals = reinterpret_cast<T *>(lds_raw);
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd, L.xcdLog2);
  const int kb = blockIdx.y * KF;
  const int r0 = blockStarts[b], r1 = blockStarts[b + 1];
  const int nrows = r1 - r0;
  const int g0 = starts[r0];
  const int nnz = starts[r1] - g0;
  if (nnz > L.cap || nrows > L.cap) {
    flag_error(L.err, 3);
    return;
  }
  if (nrows == 0) return;
  const int64_t K = L.K;
  const int t = threadIdx.x;
  int ra[PER], rb[PER];
  load_row_ranges<PER>(ra, rb, starts, r0, nrows, g0);
  if (nnz > 0) {
    int cid[PER];
    load_ids<PER>(cid, ids, g0, nnz);
    T x[PER][KF];
#pragma unroll
    for (int j = 0; j < PER; ++j)
#pragma unroll
      for (int f = 0; f < KF; ++f) x[j][f] = v[(int64_t)cid[j] * K + kb + f];
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (t + j * kThreads < nnz)
#pragma unroll
        for (int f = 0; f < KF; ++f) vals[(t + j * kThreads) * KF + f] = x[j][f];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t + j * kThreads < nrows) {
      const int a = ra[j], e = rb[j];
#pragma unroll
      for (int f = 0; f < KF; ++f) {
        T acc = Fold<T, OP>::init();
        for (int i = a; i < e; ++i) Fold<T, OP>::apply(acc, vals[i * KF + f]);
        gv[(int64_t)(r0 + t + j * kThreads) * K + kb + f] = acc;
      }
    }
}

// --------------------------------------------------------------- scatter --
// LDS: vals[cap*KF].  Row ranges and row values live in registers.
template <typename T, int KF, int PER>
__global__ __launch_bounds__(kThreads) void k_scatter(Launch L, const int *__restrict__ blockStarts,
                                                      const int *__restrict__ starts,
                                                      const int *__restrict__ ids,
                                                      const T *__restrict__ gv, T *__restrict__ v) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *vals = reinterpret_cast<T *>(lds_raw);
  const int b = xcd_tile(blockIdx.x, L.nblocks, L.nxcd, L.xcdLog2);
  const int kb = blockIdx.y * KF;
  const int r0 = blockStarts[b], r1 = blockStarts[b + 1];
  const int nrows = r1 - r0;
  const int g0 = starts[r0];
  const int nnz = starts[r1] - g0;
  if (nnz > L.cap || nrows > L.cap) {
    flag_error(L.err, 4);
    return;
  }
  if (nrows == 0 || nnz == 0) return;
  const int64_t K = L.K;
  const int t = threadIdx.x;
  int cid[PER], ra[PER], rb[PER];
  load_ids<PER>(cid, ids, g0, nnz);
  load_row_ranges<PER>(ra, rb, starts, r0, nrows, g0);
  T x[PER][KF];
#pragma unroll
  for (int j = 0; j < PER; ++j)
#pragma unroll
    for (int f = 0; f < KF; ++f) x[j][f] = gv[(int64_t)(r0 + clampi(t + j * kThreads, nrows - 1)) * K + kb + f];
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t + j * kThreads < nrows) {
      const int a = ra[j], e = rb[j];
#pragma unroll
      for (int f = 0; f < KF; ++f)
        for (int i = a; i < e; ++i) vals[i * KF + f] = x[j][f];
    }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t + j * kThreads < nnz)
#pragma unroll
      for (int f = 0; f < KF; ++f) v[(int64_t)cid[j] * K + kb + f] = vals[(t + j * kThreads) * KF + f];
}

// ------------------------------------------------------- inverse scatter --
// v[k + l*K] = gv[k + g2l[l]*K] for every local node l with g2l[l] >= 0: the
// same result as the CSR scatter (each local node sits in at most one row),
// with reads of g2l and writes of v in local order.  Lane t of block b
// handles the NPL nodes b*NPL*256 + u*256 + t, so each map load and value
// store of a wave covers 64 consecutive nodes (a lane taking NPL
// consecutive nodes instead spread every store of a wave over NPL*64 nodes
// and ran at 2.6 TB/s).
template <typename T, int KF, int NPL>
__global__ __launch_bounds__(kThreads) void k_scatter_inv(int64_t n, int K, const int *__restrict__ g2l,
                                                          const T *__restrict__ gv, T *__restrict__ v) {
  const int kb = blockIdx.y * KF;
  const int64_t base = (int64_t)blockIdx.x * NPL * kThreads + threadIdx.x;
  int r[NPL];
#pragma unroll
  for (int u = 0; u < NPL; ++u) {
    const int64_t l = base + (int64_t)u * kThreads;
    r[u] = l < n ? g2l[l] : -1;
  }
  T x[NPL][KF];
#pragma unroll
  for (int u = 0; u < NPL; ++u)
#pragma unroll
    for (int f = 0; f < KF; ++f) x[u][f] = gv[(int64_t)(r[u] < 0 ? 0 : r[u]) * K + kb + f];
#pragma unroll
  for (int u = 0; u < NPL; ++u)
    if (r[u] >= 0)
#pragma unroll
      for (int f = 0; f < KF; ++f) v[(base + (int64_t)u * kThreads) * K + kb + f] = x[u][f];
}

// --------------------------------------------------------------- extract --
template <typename T>
__global__ __launch_bounds__(kThreads) void k_extract(int64_t total, int K, const int *__restrict__ ids,
                                                      const T *__restrict__ q, T *__restrict__ gq) {
  int64_t n = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * kThreads;
  for (; n < total; n += st) {
    const int64_t g = n / K;
    const int k = (int)(n - g * K);
    gq[n] = q[k + (int64_t)ids[g] * K];
  }
}

// ------------------------------------------------------------- launching --
// Supported (KF, PER) pairs; cap <= 256*PER.
//   KF = 1: PER in {4, 8, 16}   (cap up to 1024 / 2048 / 4096)
//   KF = 2: PER in {4, 8}
//   KF = 3, 4: PER = 4
inline int pick_per(int kf, int cap) {
  const int need = (cap + kThreads - 1) / kThreads;
  if (need <= 4) return 4;
  if (need <= 8 && kf <= 4) return 8;
  if (need <= 16 && kf == 1) return 16;
  return 0;
}

// LDS per workgroup: one value tile (gather, scatter, Sym GS), or two value
// tiles plus two row-start arrays (general GS).  Sym GS variant 1 adds its
// row starts (launch_gs_sym).
inline size_t lds_bytes(int cap, int kf, size_t tsize, int arrays) {
  if (arrays == 1) return (size_t)cap * kf * tsize;
  return 2 * ((size_t)cap * kf * tsize + (size_t)(cap + 1) * sizeof(int));
}

// Per-type launchers, defined by one instantiation unit per type.
int launch_gs_sym(const Launch &L, int op, const int *bs, const int *st, const int *ids, float *q);
int launch_gs_sym(const Launch &L, int op, const int *bs, const int *st, const int *ids, double *q);
int launch_gs_sym(const Launch &L, int op, const int *bs, const int *st, const int *ids, int *q);
int launch_gs_sym(const Launch &L, int op, const int *bs, const int *st, const int *ids, long long *q);
#define SPX_OGS_GEN_DECL(T)                                                                                  \
  int launch_gs_gen(const Launch &L, int op, int ngrows, int nsrows, const int *bs, const int *gs,          \
                    const int *gi, const int *ss, const int *si, T *q);                                    \
  int launch_gather(const Launch &L, int op, const int *bs, const int *st, const int *ids, const T *v, T *gv); \
  int launch_scatter(const Launch &L, const int *bs, const int *st, const int *ids, const T *gv, T *v);    \
  int launch_extract(hipStream_t s, int64_t total, int K, const int *ids, const T *q, T *gq);           \
  int launch_scatter_inv(hipStream_t s, int64_t n, int K, const int *g2l, const T *gv, T *v, int npl); \
  int launch_gs_srb(const Launch &L, int op, const SrbArgs &A, T *q, int mode,                              \
                    const unsigned long long *memberN);                                                     \
  int launch_gather_srb(const Launch &L, int op, const SrbArgs &A, const T *v, T *gv, int mode,             \
                        const unsigned long long *memberN);                                                 \
  int launch_scatter_srb(const Launch &L, const SrbArgs &A, const T *gv, T *v, int mode,                    \
                         const unsigned long long *memberN);
SPX_OGS_GEN_DECL(float)
SPX_OGS_GEN_DECL(double)
SPX_OGS_GEN_DECL(int)
SPX_OGS_GEN_DECL(long long)
#undef SPX_OGS_GEN_DECL

}  // namespace ogsk
}  // namespace spx
