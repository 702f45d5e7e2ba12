/*
 * spx_oracle.h -- CPU oracle for the streamParanumal streaming hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * Serial-mode semantics (OCCA "Serial" runs each OKL loop nest as sequential
 * host loops).  The HIP library's results are checked against it: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the streamparanumal source tree).  Parity status per function is listed in
 * the header comment of spx_oracle.c.
 */
#ifndef SPX_ORACLE_H
#define SPX_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- input generator: the reference LCG (libs/primitives/random.cpp:45-163) ---- */
uint64_t oracle_lcg_state(uint64_t seed, uint64_t skip);
void oracle_fill_uniform(int64_t n, double *v, uint64_t seed, uint64_t skip);
void oracle_fill_loguniform(int64_t n, double *v, uint64_t seed, uint64_t skip);
void oracle_fill_rand_int(int64_t n, int32_t *v, uint64_t seed, uint64_t skip);

/* ---- BS1..BS5 vector kernels (BS/BSn/okl/bsN.okl) ---- */
void oracle_bs1_copy(int64_t n, const double *a, double *b);
void oracle_bs2_axpy(int64_t n, double alpha, const double *x, double beta, double *y);
double oracle_bs3_norm(int64_t n, const double *a, int block_size, double *partials);
double oracle_bs4_dot(int64_t n, const double *a, const double *b, int block_size, double *partials);
double oracle_bs5_update(int64_t n, const double *Ap, double alpha, double *r, int block_size,
                         double *partials);
double oracle_bs5_full(int64_t n, double alpha, const double *p, const double *Ap,
                       double *x, double *r, int block_size);
double oracle_sum_longdouble(int64_t n, const double *a, const double *b, int mode);

/* ---- hex box mesh numbering (libs/mesh/meshSetupBoxHex3D.cpp, meshConnectNodes.cpp) ---- */
void oracle_hex_ids_lex(int nx, int ny, int nz, int p, int64_t *ids);
void oracle_hex_ids_minlocal(int nx, int ny, int nz, int p, int64_t *ids);

/* ---- ogs setup (libs/ogs/ogsSetup.cpp, ogsOperator.cpp) ---- */
enum { ORACLE_OGS_UNSIGNED = 0, ORACLE_OGS_SIGNED = 1, ORACLE_OGS_HALO = 2 };
enum { ORACLE_OP_ADD = 0, ORACLE_OP_MUL = 1, ORACLE_OP_MAX = 2, ORACLE_OP_MIN = 3 };
enum { ORACLE_SYM = 0, ORACLE_NOTRANS = 1, ORACLE_TRANS = 2 };
enum { ORACLE_FLOAT = 0, ORACLE_DOUBLE = 1, ORACLE_INT32 = 2, ORACLE_INT64 = 3 };

typedef struct {
  int32_t N;            /* local entries (columns) */
  int32_t NrowsT;       /* gathered rows, all groups (NlocalT) */
  int32_t NrowsN;       /* gathered rows with a positive member (NlocalP) */
  int32_t nnzT, nnzN;
  int32_t *rowStartsT;  /* NrowsT+1 */
  int32_t *colIdsT;     /* nnzT */
  int32_t *rowStartsN;  /* NrowsN+1 */
  int32_t *colIdsN;     /* nnzN */
  int32_t NblocksT;     /* reference bisection row blocks (<=512 nnz) */
  int32_t *blockStartsT;
  int32_t NblocksN;
  int32_t *blockStartsN;
  int64_t *ids;         /* ids after setup (sign-flagged when unique) */
  int32_t gather_defined;
  int32_t error;        /* nonzero: the reference would LIBP_ABORT */
} oracle_ogs_t;

oracle_ogs_t *oracle_ogs_setup(int32_t n, const int64_t *ids, int kind, int unique);
void oracle_ogs_free(oracle_ogs_t *o);

/* ogsOperator_t host Gather/Scatter/GatherScatter (ogsOperator.cpp:61-114,223-267,340-409) */
void oracle_ogs_gather(const oracle_ogs_t *o, int type, int op, int trans, int K,
                       const void *v, void *gv);
void oracle_ogs_scatter(const oracle_ogs_t *o, int type, int trans, int K,
                        const void *gv, void *v);
void oracle_ogs_gather_scatter(const oracle_ogs_t *o, int type, int op, int trans, int K, void *v);

/* Sym Add gatherScatter over caller CSR arrays (ogsOperator.cpp:374-389) */
void oracle_csr_gs(int32_t nrows, const int32_t *rowStarts, const int32_t *colIds, int K, double *q);

/* gather (Add) and scatter over caller CSR arrays (ogsOperator.cpp:61-114, 223-267) */
void oracle_csr_gather(int32_t nrows, const int32_t *rowStarts, const int32_t *colIds, int K,
                       const double *q, double *gv);
void oracle_csr_scatter(int32_t nrows, const int32_t *rowStarts, const int32_t *colIds, int K,
                        const double *gv, double *q);

/* the OKL block-phase form of gatherScatter (ogsKernels.okl:32-81), Add, K fields */
void oracle_ogs_gs_okl(int32_t nblocks, int K, const int32_t *blockStarts,
                       const int32_t *gStarts, const int32_t *gIds,
                       const int32_t *sStarts, const int32_t *sIds, double *q);

/* ---- reduced-precision storage variants (no reference kernel; the
   semantics are this build's, restated here for bit-exact checks) ---- */
enum { ORACLE_STORE_F64 = 0, ORACLE_STORE_F32 = 1, ORACLE_STORE_BF16 = 2 };
uint16_t oracle_bf16_from_double(double x);     /* round to nearest even */
uint16_t oracle_bf16_from_float(float x);       /* round to nearest even */
double oracle_bf16_to_double(uint16_t b);
/* Sym Add gather+scatter of values stored as `storage`, each row summed in
   FP64 in stored order and rounded once to the storage type */
void oracle_gs_mixed(int storage, int32_t nrows, const int32_t *rowStarts, const int32_t *colIds, int K,
                     void *q);
/* x += alpha p, r -= alpha Ap in FP32 arithmetic on values stored as
   `storage`; returns sum over the stored r of r^2 in long double */
double oracle_cg_mixed(int storage, int64_t n, double alpha, const void *p, const void *Ap, void *x, void *r);

/* ---- CPU baseline timers (Serial = 1 thread; OpenMP = all host cores) ---- */
void oracle_time_ogs(int nx, int p, int reps, int nthr, const int *threads, double *seconds,
                     double *checksums, double *setup_seconds, int64_t *nlocal, int64_t *nglobal);
double oracle_time_bs8(int nx, int p, int reps, int threads, double *checksum,
                       double *setup_seconds, int64_t *nlocal, int64_t *nglobal);
double oracle_time_stream(int which, int64_t n, int reps, int threads);
int oracle_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
