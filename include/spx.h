/*
 * spx.h -- C ABI of libspx.so, the MI355X (gfx950) streaming library.
 *
 * libspx.so replaces the OCCA device that streamParanumal's libs/core drives
 * (platform_t, deviceMemory<T>, kernel_t) and the OKL kernels of the
 * streaming benchmarks BS0-BS8 and of libs/ogs.  Every declaration below cites
 * the reference interface it replaces (paths relative to the streamparanumal
 * tree).  The C++ host layer (streamparanumal_amd/libp) and the Python package
 * are written on top of these functions only.
 *
 * Conventions
 *   - Every function returns SPX_OK (0) or a negative SPX_ERR_* code; the
 *     message of the last failure on the calling thread is spx_last_error().
 *   - Pointers are plain device (or pinned host) addresses; sizes are int64
 *     element counts unless named *bytes.  The caller owns every buffer.
 *   - Kernels are enqueued on the library's current stream (spx_set_stream)
 *     and return without waiting; spx_device_finish / spx_stream_finish wait.
 *   - Index arrays (CSR row starts, column ids, block starts) are int32, the
 *     reference's dlong (include/types.h:60).
 */
#ifndef SPX_H
#define SPX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPX_VERSION 100

enum {
  SPX_OK = 0,
  SPX_ERR_ARG = -1,         /* invalid argument (size, alignment, enum)      */
  SPX_ERR_HIP = -2,         /* HIP runtime error                             */
  SPX_ERR_NODEVICE = -3,    /* no device opened / no GPU                     */
  SPX_ERR_UNSUPPORTED = -4, /* valid request this build does not implement   */
  SPX_ERR_KERNEL = -5       /* a kernel reported a violated precondition     */
};

/* ogs enums, same values as include/ogs.hpp:179-200 */
enum { SPX_FLOAT = 0, SPX_DOUBLE = 1, SPX_INT32 = 2, SPX_INT64 = 3 }; /* ogs::Type */
enum { SPX_ADD = 0, SPX_MUL = 1, SPX_MAX = 2, SPX_MIN = 3 };          /* ogs::Op   */
enum { SPX_SYM = 0, SPX_NOTRANS = 1, SPX_TRANS = 2 };                 /* ogs::Transpose */

/* reduction order for BS3-BS5 */
enum {
  SPX_REDUCE_REFERENCE = 0, /* bit-exact bs3/bs4/bs5 .okl order (Nblock<=blockSize) */
  SPX_REDUCE_FAST = 1       /* the same shape, compensated lane sums; deterministic, within 1e-14 */
};

typedef struct spx_stream_s *spx_stream_t;
typedef struct spx_event_s *spx_event_t;
typedef struct spx_graph_s *spx_graph_t;

typedef struct {
  char name[256];
  char arch[64];          /* e.g. "gfx950:sramecc+:xnack-" */
  int device_id;
  int compute_units;
  int clock_khz;
  int memory_clock_khz;
  int memory_bus_width;
  int64_t global_mem_bytes;
  int64_t l2_bytes;
  int64_t lds_bytes_per_block;
  int warp_size;
} spx_device_props_t;

/* ---------------------------------------------------------------- errors -- */
const char *spx_last_error(void);
int spx_version(void);

/* ---------------------------------------------------------------- device -- */
/* occa::getDeviceCount (include/platform.hpp:164-166) */
int spx_device_count(int *count);
/* occa::device::setup({mode:'HIP', device_id}) (libs/core/platformDeviceConfig.cpp:159) */
int spx_device_open(int device_id);
/* occa::device::finish (include/platform.hpp:152-154) */
int spx_device_finish(void);
/* properties used for reporting (platformProperties.cpp has none; new) */
int spx_device_props(spx_device_props_t *props);
/* kernels that met a violated precondition since the last call (new) */
int spx_check_kernel_errors(void);

/* ---------------------------------------------------------------- streams -- */
/* occa::device::createStream/setStream/getStream (ogsSetup.cpp:103-104, ogs.cpp:66-80) */
int spx_stream_create(spx_stream_t *stream);
int spx_stream_destroy(spx_stream_t stream);
int spx_stream_finish(spx_stream_t stream);
int spx_set_stream(spx_stream_t stream); /* NULL selects the library's own stream */
spx_stream_t spx_get_stream(void);

/* ---------------------------------------------------------------- memory -- */
/* occa::device::malloc (include/platform.hpp:95-115) */
int spx_malloc(void **ptr, size_t bytes);
/* occa::device::malloc with {"host": true} (include/platform.hpp:117-140) */
int spx_malloc_host(void **ptr, size_t bytes);
int spx_free(void *ptr);
int spx_free_host(void *ptr);
/* occa::memory::copyFrom/copyTo, sync or {"async": true} (include/memory.hpp:332-450) */
int spx_memcpy(void *dst, const void *src, size_t bytes, int async);
/* kernel-based byte fill (capturable in a graph, unlike a memset node) */
int spx_fill_bytes(void *ptr, size_t bytes, uint8_t value);

/* ---------------------------------------------------------------- timing -- */
/* device-side timing; the reference times on the host after finish (libs/core/timer.cpp:43-60) */
int spx_event_create(spx_event_t *event);
int spx_event_destroy(spx_event_t event);
int spx_event_record(spx_event_t event); /* on the current stream */
int spx_event_elapsed_ms(spx_event_t start, spx_event_t stop, float *ms);

/* ---------------------------------------------------------------- graphs -- */
/* capture of the current stream into a replayable graph (BS0 comparison; new) */
int spx_graph_begin(void);
int spx_graph_end(spx_graph_t *graph);
int spx_graph_launch(spx_graph_t graph);
int spx_graph_destroy(spx_graph_t graph);

/* ------------------------------------------------------- synthetic inputs -- */
/* the reference LCG (libs/primitives/random.cpp:45-163) mapped to U[-1,1):
   v[i] = -1 + (x_{skip+i} >> 11) * 2^-52 with x_0 = A*(UINT64_MAX ^ seed) + C */
int spx_fill_uniform_f64(int64_t n, double *v, uint64_t seed, uint64_t skip);
/* log-uniform magnitude in [1e-6, 1] with a random sign, from the same LCG:
   v[i] = +-10^(-6u), u = (x_{2(skip+i)} >> 11) * 2^-53, negative when the top
   bit of x_{2(skip+i)+1} is set (the device exp10 may differ from a host pow
   in the last bit) */
int spx_fill_loguniform_f64(int64_t n, double *v, uint64_t seed, uint64_t skip);
/* element-wise conversion for the reduced-precision storage variants */
int spx_convert_f64_to_f32(int64_t n, const double *src, float *dst);
int spx_convert_f64_to_bf16(int64_t n, const double *src, uint16_t *dst);
int spx_convert_f32_to_f64(int64_t n, const float *src, double *dst);
int spx_convert_bf16_to_f64(int64_t n, const uint16_t *src, double *dst);

/* ------------------------------------------------------------ BS0 .. BS5 -- */
/* bs0 (BS/BS0/okl/bs0.okl:27-39): one 1x1 launch that writes b[0]=0 only if N==0 */
int spx_bs0(int32_t n, const double *a, double *b);
/* bs1 (BS/BS1/okl/bs1.okl:27-34): b[i] = a[i] */
int spx_bs1_copy(int64_t n, const double *a, double *b);
/* bs2 (BS/BS2/okl/bs2.okl:27-36): y[i] = alpha*x[i] + beta*y[i], no contraction */
int spx_bs2_axpy(int64_t n, double alpha, const double *x, double beta, double *y);

/* Reductions.  `tmp` holds spx_reduce_tmp_count(n, mode, block_size) doubles of
   per-block partials; the scalar result is written to device memory `out`.
   block_size (REFERENCE mode) is the reference's p_blockSize: 256, 512 or 1024
   (bs3Setup.cpp:37-39, bs5Setup.cpp:37). */
int64_t spx_reduce_tmp_count(int64_t n, int mode, int block_size);
/* bs3_1 + bs3_2 (BS/BS3/okl/bs3.okl:27-102): out = sum a[i]^2 */
int spx_bs3_norm(int64_t n, const double *a, double *tmp, double *out, int mode, int block_size);
/* bs4_1 + bs4_2 (BS/BS4/okl/bs4.okl:27-103): out = sum a[i]*b[i] */
int spx_bs4_dot(int64_t n, const double *a, const double *b, double *tmp, double *out, int mode,
                int block_size);
/* bs5_1 + bs5_2 (BS/BS5/okl/bs5.okl:27-106): r += alpha*Ap; rdotr = sum r^2 */
int spx_bs5_update(int64_t n, const double *Ap, double alpha, double *r, double *tmp,
                   double *rdotr, int mode, int block_size);
/* The two stages separately, with the caller's block count, as bsNRun.cpp
   launches them: stage 1 writes nblocks partials, stage 2 (bs3_2 = bs4_2 =
   bs5_2, bs3.okl:66-102) folds them into out[0]. */
int spx_bs3_1(int64_t nblocks, int64_t n, const double *a, double *partials, int block_size);
int spx_bs4_1(int64_t nblocks, int64_t n, const double *a, const double *b, double *partials,
              int block_size);
int spx_bs5_1(int64_t nblocks, int64_t n, const double *Ap, double alpha, double *r,
              double *partials, int block_size);
int spx_reduce_2(int64_t nblocks, const double *partials, double *out, int block_size);
/* CG update of README.md:49-53 (no reference kernel): x += alpha*p; r -= alpha*Ap;
   rdotr = sum r^2, reduced in the bs5 order */
int spx_bs5_full(int64_t n, double alpha, const double *p, const double *Ap, double *x,
                 double *r, double *tmp, double *rdotr, int mode, int block_size);

/* spx_bs5_full with no tmp argument and no allocation, for callers that
   cannot allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
   The partials live in a library-owned array of SPX_NOALLOC_SCRATCH_DOUBLES
   doubles per device (32 KiB, held until the process ends); a request whose
   partial count (spx_reduce_tmp_count) exceeds it is refused with
   SPX_ERR_ARG, which no block size or FAST grid of this library reaches.
   Every call on a device shares that array: calls that overlap on one device
   (two streams, two host threads) race.  Results are bit-identical to
   spx_bs5_full's with the same mode and block size. */
enum { SPX_NOALLOC_SCRATCH_DOUBLES = 4096 };
int spx_bs5_full_noalloc(int64_t n, double alpha, const double *p, const double *Ap, double *x,
                         double *r, double *rdotr, int mode, int block_size);

/* ---------------------------------------------------------- ogs kernels -- */
/* The CSR operator of ogsOperator_t (include/ogs/ogsOperator.hpp:44-70): rows
   are gathered nodes, rowStarts has nrows+1 entries, colIds holds local
   indices in increasing order within a row.  blockStarts partitions the rows
   into nblocks row blocks: block b owns rows [blockStarts[b], blockStarts[b+1])
   and must hold at most max_block_nnz column entries (<= 4096) in every CSR
   it is used with.  Values of field k of local node l live at v[k + l*K]. */

/* gather (libs/ogs/okl/ogsKernels.okl:86-122; ogsOperator.cpp:176-204):
   gv[k + row*K] = op over v[k + col*K], cols of row in stored order */
int spx_ogs_gather(int type, int op, int K, int32_t nblocks, const int32_t *blockStarts,
                   const int32_t *rowStarts, const int32_t *colIds, int32_t max_block_nnz,
                   const void *v, void *gv);
/* scatter (ogsKernels.okl:127-164; ogsOperator.cpp:295-322): v[k + col*K] = gv[k + row*K] */
int spx_ogs_scatter(int type, int K, int32_t nblocks, const int32_t *blockStarts,
                    const int32_t *rowStarts, const int32_t *colIds, int32_t max_block_nnz,
                    const void *gv, void *v);
/* gatherScatter (ogsKernels.okl:32-81; ogsOperator.cpp:442-481), in place:
   per row, acc = op over q[k + g*K] for g in the gather CSR row, then
   q[k + s*K] = acc for s in the scatter CSR row.  When gStarts==sStarts and
   gIds==sIds (ogs::Sym) the single-CSR kernel runs.  ngrows / nsrows: rows of
   the gather / scatter CSR; rows past ngrows gather nothing (the op's identity
   is scattered), rows past nsrows scatter nothing.  blockStarts[nblocks] must
   not exceed max(ngrows, nsrows). */
int spx_ogs_gather_scatter(int type, int op, int K, int32_t nblocks, const int32_t *blockStarts,
                           int32_t ngrows, const int32_t *gStarts, const int32_t *gIds,
                           int32_t nsrows, const int32_t *sStarts, const int32_t *sIds,
                           int32_t max_block_nnz, void *q);
/* scatter through the local->gathered map g2l (n entries, -1 = in no row):
   v[k + l*K] = gv[k + g2l[l]*K].  Same result as spx_ogs_scatter over a CSR
   whose rows partition the local nodes (every ogs operator's do); the map is
   the reference's GlobalToLocal (ogsSetup.cpp:833-849).  gv must hold at
   least one row (K values) even if every map entry is -1. */
int spx_ogs_scatter_inverse(int type, int K, int32_t n, const int32_t *g2l, const void *gv, void *v);
/* extract (ogsKernels.okl:167-177): gq[k + n*K] = q[k + ids[n]*K], n < n_ids */
int spx_ogs_extract(int type, int K, int32_t n_ids, const int32_t *ids, const void *q, void *gq);

/* Sorted-run row blocks ("tiles"): the T-CSR of an operator re-listed for
   coalesced access (no reference counterpart; the rows, and the member order
   within each row, are those of the CSR of ogsOperator.hpp:44-70; built by
   spxh_srb_build / the ogs_t setup).  Tile b holds whole rows; its entries are
   listed in ascending local index ("positions") and it owns the chunk records
   [b*chunks_per_tile, (b+1)*chunks_per_tile), record k describing positions
   64k..64k+63, of which the first nvalid are entries.  In record C:
     runs:  bit l of C.runMask is set when position l starts a run of
            consecutive local indices (bit 0 always); run r of the record
            holds runIds[C.runBase + r] + l at its positions l, i.e. the
            entry is the run's first local index less the run's first
            position
     rows:  bit l of C.linkMask is set when position l is not the first
            (smallest) member of its row; such positions, in order, have the
            tile position of the previous member of their row in
            links[C.linkBase], links[C.linkBase+1], ...; links ends with one
            padding word.
     chunkRow: CSR row of the first row-leading position at or after
            position 0 of the record (row-leading positions hold the rows in
            CSR order).
   Row values are folded from the first member along the links, i.e. in
   ascending local index, as the CSR kernels do.  Arrays are device memory. */
typedef struct {
  uint64_t runMask, linkMask;
  int32_t runBase, linkBase, chunkRow, nvalid;
} spx_srb_chunk_t;
typedef struct {
  int32_t nblocks;
  int32_t chunks_per_tile; /* 1..64: positions per tile = 64 * chunks_per_tile <= 4096 */
  int32_t xcd_remap;       /* 1: consecutive tiles on one XCD in blocks (128 tiles; 32 for the FP32 /
                              BF16 gather+scatter), a power of two 2..1024: blocks of that many
                              tiles, 0: hardware order, negative: library default */
  int32_t tile_lanes;      /* lanes per tile: 64 (one wave, <= 1024 positions) or 256; 0: library default */
  int32_t nontemporal;     /* bit e (e = 1..5): the Sym gather+scatter (spx_ogs_gather_scatter_srb,
                              spx_ogs_gather_scatter_srb_signed with SPX_SYM,
                              spx_ogs_gather_scatter_srb_mixed) streams the values of
                              nodes of 2^e bytes (K values of the type, all K in one pass) past the
                              caches with non-temporal loads and stores (256-lane tiles; the
                              one-wave form keeps ordinary ones); the layout builders set
                              the bits for which whole 128-byte lines make up most of the tiles'
                              accesses.  0: ordinary loads and stores.  The other sorted-run
                              entry points ignore it. */
  const spx_srb_chunk_t *chunks;
  const int32_t *runIds;
  const uint16_t *links;
} spx_srb_t;
/* gatherScatter, ogs::Sym (ogsKernels.okl:32-81), on the sorted-run layout:
   same result as spx_ogs_gather_scatter on the CSR it was built from */
int spx_ogs_gather_scatter_srb(int type, int op, int K, const spx_srb_t *srb, void *q);
/* gather (ogsKernels.okl:86-122) and scatter (ogsKernels.okl:127-164) on the
   sorted-run layout of a CSR: same results as spx_ogs_gather /
   spx_ogs_scatter on that CSR; the row of a row-leading position is its
   record's chunkRow plus the row-leading positions before it in the record,
   so each tile's rows must ascend by their first member's local index
   (spxh_srb_rows_ascending) */
int spx_ogs_gather_srb(int type, int op, int K, const spx_srb_t *srb, const void *v, void *gv);
int spx_ogs_scatter_srb(int type, int K, const spx_srb_t *srb, const void *gv, void *v);
/* The N-CSR forms of an ogs::Signed operator on the sorted-run layout of its
   T-CSR.  The N-CSR's row r holds the members of T-CSR row r with a positive
   id, in the same order (ogsOperator.cpp:621-643); memberN has one mask per
   chunk record (index b*chunks_per_tile + k, like the records), bit l set when
   position l is such a member.  The layout must satisfy that: N-CSR row r is
   the set members of T-CSR row r for r < nrowsN, and later rows have none
   (spxh_srb_build_signed checks it).
     gather_scatter, SPX_NOTRANS: each row folds its N members (the op's
       identity when it has none) and stores to every member; SPX_TRANS: each
       row folds every member and stores to its N members only; SPX_SYM:
       spx_ogs_gather_scatter_srb, memberN unused (ogsOperator.cpp:340-409,
       442-481, the gatherScatter kernel on gather N / scatter T and on gather
       T / scatter N);
     gather: the NoTrans gather, gv[r] for r < nrowsN over the N members of
       row r (ogsOperator.cpp:176-204, N-CSR);
     scatter: the Trans scatter, v[m] = gv[r] for the N members m of rows
       r < nrowsN, other entries untouched (ogsOperator.cpp:295-322, N-CSR);
       gather and scatter number rows as spx_ogs_gather_srb does, so they
       need the same row order (spxh_srb_rows_ascending).
   Results are bit-identical to the CSR kernels on the N- and T-CSR. */
int spx_ogs_gather_scatter_srb_signed(int type, int op, int trans, int K, const spx_srb_t *srb,
                                      const uint64_t *memberN, void *q);
int spx_ogs_gather_srb_signed(int type, int op, int K, const spx_srb_t *srb, const uint64_t *memberN,
                              int32_t nrowsN, const void *v, void *gv);
int spx_ogs_scatter_srb_signed(int type, int K, const spx_srb_t *srb, const uint64_t *memberN, int32_t nrowsN,
                               const void *gv, void *v);

/* ------------------------------------------- reduced-precision storage -- */
/* Storage types of the mixed-precision variants (no reference counterpart:
   the reference's dfloat is one type for storage and arithmetic,
   include/types.h:36-44 and README.md:49-53). */
enum { SPX_STORE_F64 = 0, SPX_STORE_F32 = 1, SPX_STORE_BF16 = 2 };
/* partials needed by spx_bs5_full_mixed's tmp */
enum { SPX_MIXED_TMP_COUNT = 2048 };

/* ogs::Sym Add gatherScatter (ogsKernels.okl:32-81) on a vector stored as
   `storage`: each row's sum is taken in FP64 over its members in stored
   order and rounded once to the storage type.  SPX_STORE_F64 runs
   spx_ogs_gather_scatter.  Arguments as spx_ogs_gather_scatter's Sym case. */
int spx_ogs_gather_scatter_mixed(int storage, int K, int32_t nblocks, const int32_t *blockStarts,
                                 const int32_t *rowStarts, const int32_t *colIds,
                                 int32_t max_block_nnz, void *q);
/* the same on the sorted-run layout of spx_ogs_gather_scatter_srb */
int spx_ogs_gather_scatter_srb_mixed(int storage, int K, const spx_srb_t *srb, void *q);
/* Batched CG update (README.md:49-53) over nseg independent vectors laid out
   back to back ("segments": segment s is entries [segStarts[s],
   segStarts[s+1]), each start a multiple of 4), in one launch plus one
   per-segment fold: x += alpha*p, r -= alpha*Ap (FP64 storage: in FP64;
   FP32 / BF16: in FP32, alpha rounded to FP32), rdotr[s] = sum over segment s
   of the stored r squared, in FP64.  No reference counterpart (the reference
   runs one vector per launch, BS/BS5).  A plan cuts every segment into tiles
   at the multiples of 4096 entries and launches min(target_blocks,
   tiles) workgroups that stride over the tiles; each tile leaves one FP64
   partial and stage 2 folds each segment's partials in tile order, so results
   depend neither on target_blocks nor on the run.  spx_cg_plan_blocks gives
   the workgroups launched. */
typedef struct spx_cg_plan_s *spx_cg_plan_t;
int spx_cg_plan_create(int32_t nseg, const int64_t *segStarts, int32_t target_blocks, spx_cg_plan_t *plan);
int spx_cg_plan_destroy(spx_cg_plan_t plan);
int spx_cg_plan_blocks(spx_cg_plan_t plan);
int spx_cg_batch(int storage, spx_cg_plan_t plan, double alpha, const void *p, const void *Ap, void *x,
                 void *r, double *rdotr);
/* CG update (README.md:49-53) on FP32 or BF16 vectors: x += alpha*p and
   r -= alpha*Ap in FP32 (alpha rounded to FP32), results rounded to the
   storage type; *rdotr = sum of the stored r squared, accumulated in FP64.
   Vectors aligned to 4 entries; tmp holds SPX_MIXED_TMP_COUNT doubles. */
int spx_bs5_full_mixed(int storage, int64_t n, double alpha, const void *p, const void *Ap, void *x,
                       void *r, double *tmp, double *rdotr);

/* Kernel launch parameters chosen by the autotuner (0 = library default). */
typedef struct {
  int32_t block_threads; /* 64..1024, multiple of 64 */
  int32_t xcd_remap;     /* 1: map consecutive row blocks onto one XCD's L2 */
  int32_t unroll;        /* elements per thread per step in the vector kernels */
  int32_t grid_blocks;   /* fixed grid of the FAST reductions */
  int32_t nontemporal;   /* 1: streaming (non-temporal) loads/stores in BS1/BS2 */
  int32_t variant;       /* kernel variant of the family (0 = default; see the kernel source) */
} spx_tuning_t;
int spx_set_tuning(const char *kernel, const spx_tuning_t *tuning);
int spx_get_tuning(const char *kernel, spx_tuning_t *tuning);

#ifdef __cplusplus
}
#endif
#endif /* SPX_H */
