/*
 * spx_oracle.c -- CPU oracle (TEST INFRASTRUCTURE, never linked by the product).
 *
 * Plain-C restatement of the reference's Serial-mode semantics for the
 * streaming hot path.  Build: oracle/Makefile (gcc -O3 -ffp-contract=off
 * -fopenmp).  -ffp-contract=off is required: the restatement defines
 * "a*b + c" as two roundings, never a fused multiply-add.
 *
 * Parity pins (what checks each function against something other than itself;
 * the checks live in tests/test_oracle.py):
 *   oracle_lcg_state / fills     : closed-form LCG recurrence checked against an
 *                                  independent Python big-int recurrence and
 *                                  the jump-ahead identity.
 *   oracle_bs1_copy, bs2_axpy    : closed form per element (exact products and
 *                                  sums on dyadic inputs).
 *   oracle_bs3/4/5 (ref order)   : known answers on integer inputs (exact sums),
 *                                  hand-derived order probes (2^53 absorption)
 *                                  that fail for any other summation order, and
 *                                  <=1e-14 vs an exactly rounded sum.
 *   oracle_bs5_full              : no reference kernel computes this form
 *                                  (README.md:49-53 describes it); pinned by
 *                                  the closed form per element and by its
 *                                  reduction being the BS5 order.
 *   oracle_hex_ids_*             : brute force on tiny boxes (grouping by GLL
 *                                  coordinate) and node-count closed forms.
 *   oracle_ogs_setup             : brute-force first-appearance numbering of
 *                                  tiny id lists, multiplicity KAT (1,2,4,8),
 *                                  row-order independence of the id scheme.
 *                                  The member left positive under unique=true
 *                                  is "parity unpinned": the reference picks it
 *                                  after an unstable std::sort
 *                                  (ogsSetup.cpp:347,375-381), whose tie order
 *                                  the C++ standard leaves unspecified.  The
 *                                  rule here (LCG draw over members in
 *                                  ascending local index) is documented and
 *                                  only the invariant "one positive per group"
 *                                  is checked.
 *   oracle_ogs_gather/scatter/gs : Sigma q = Sigma gv invariant, multiplicity
 *                                  KAT, order probe {1e16,1,-1e16} on a corner
 *                                  row, Scatter o Gather == GatherScatter, and
 *                                  agreement with the OKL block-phase form.
 *   oracle_csr_gather/scatter    : the same row loops over caller CSR arrays;
 *                                  multiplicity KAT, Sigma invariant, order
 *                                  probe, scatter closed form (numpy
 *                                  indexing), Scatter o Gather == csr_gs.
 */
#include "spx_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* LCG (libs/primitives/random.cpp:45-121)                                    */
/* ------------------------------------------------------------------------ */
#define LCG_A 6364136223846793005ULL
#define LCG_C 1ULL

typedef struct { uint64_t a, c; } lcg_affine;

/* composition: apply g then f  (random.cpp:56-63 RandCoeff::operator*) */
static lcg_affine lcg_compose(lcg_affine f, lcg_affine g) {
  lcg_affine r = {f.a * g.a, f.a * g.c + f.c};
  return r;
}

/* f^n by binary powering (random.cpp:66-74) */
static lcg_affine lcg_pow(lcg_affine f, uint64_t n) {
  lcg_affine r = {1ULL, 0ULL};
  while (n) {
    if (n & 1ULL) r = lcg_compose(r, f);
    f = lcg_compose(f, f);
    n >>= 1;
  }
  return r;
}

/* state of draw number `skip` of the stream seeded with `seed`
   (random.cpp:79-82: x0 = A*(UINT64_MAX ^ seed) + C) */
uint64_t oracle_lcg_state(uint64_t seed, uint64_t skip) {
  lcg_affine step = {LCG_A, LCG_C};
  uint64_t x0 = LCG_A * (UINT64_MAX ^ seed) + LCG_C;
  lcg_affine j = lcg_pow(step, skip);
  return j.a * x0 + j.c;
}

/* U[-1,1): -1 + (x>>11)*2^-52, exact in binary64 */
void oracle_fill_uniform(int64_t n, double *v, uint64_t seed, uint64_t skip) {
  uint64_t x = oracle_lcg_state(seed, skip);
  for (int64_t i = 0; i < n; ++i) {
    v[i] = -1.0 + (double)(x >> 11) * 0x1p-52;
    x = LCG_A * x + LCG_C;
  }
}

/* log-uniform magnitude in [1e-6,1], random sign: draws 2i (magnitude) and
   2i+1 (sign = top bit).  Uses libm pow, so it is generated on the host. */
void oracle_fill_loguniform(int64_t n, double *v, uint64_t seed, uint64_t skip) {
  uint64_t x = oracle_lcg_state(seed, 2 * skip);
  for (int64_t i = 0; i < n; ++i) {
    double u = (double)(x >> 11) * 0x1p-53;
    x = LCG_A * x + LCG_C;
    double mag = pow(10.0, -6.0 * u);
    v[i] = (x >> 63) ? -mag : mag;
    x = LCG_A * x + LCG_C;
  }
}

/* RandState::get<int>() = abs((int)x)  (random.cpp:96-99); abs(INT_MIN)
   stays INT_MIN as on two's-complement hardware */
void oracle_fill_rand_int(int64_t n, int32_t *v, uint64_t seed, uint64_t skip) {
  uint64_t x = oracle_lcg_state(seed, skip);
  for (int64_t i = 0; i < n; ++i) {
    uint32_t u = (uint32_t)x;
    if (u & 0x80000000u) u = 0u - u;
    v[i] = (int32_t)u;
    x = LCG_A * x + LCG_C;
  }
}

/* ------------------------------------------------------------------------ */
/* BS1..BS5                                                                  */
/* ------------------------------------------------------------------------ */

/* BS/BS1/okl/bs1.okl:27-34 */
void oracle_bs1_copy(int64_t n, const double *a, double *b) {
  for (int64_t i = 0; i < n; ++i) b[i] = a[i];
}

/* BS/BS2/okl/bs2.okl:27-36 : y = alpha*x + beta*y, two products, one add */
void oracle_bs2_axpy(int64_t n, double alpha, const double *x, double beta, double *y) {
  for (int64_t i = 0; i < n; ++i) {
    double t1 = alpha * x[i];
    double t2 = beta * y[i];
    y[i] = t1 + t2;
  }
}

/* Shared-memory tree of bs3.okl:48-63 (identical in bs4.okl, bs5.okl):
   halve while width > 2, then s[0]+s[1]. */
static double ref_block_tree(double *s, int bs) {
  for (int w = bs / 2; w >= 2; w /= 2)
    for (int t = 0; t < w; ++t) s[t] += s[t + w];
  return s[0] + s[1];
}

/* Nblock = min(ceil(N/bs), bs)   (bs3Run.cpp:80-81) */
static int64_t ref_nblock(int64_t n, int bs) {
  int64_t nb = (n + bs - 1) / bs;
  return nb > bs ? bs : nb;
}

/* bs3_2/bs4_2/bs5_2 (bs3.okl:66-102): thread t holds partial t (t<Nblock) */
static double ref_finish(int64_t nblock, const double *partials, int bs) {
  double *s = (double *)malloc(sizeof(double) * bs);
  for (int t = 0; t < bs; ++t) {
    double r = 0.0;
    for (int64_t id = t; id < nblock; id += (int64_t)bs * nblock) r += partials[id];
    s[t] = r;
  }
  double out = ref_block_tree(s, bs);
  free(s);
  return out;
}

/* bs3_1 (bs3.okl:27-64) + bs3_2 : sum of a^2 in the reference order */
double oracle_bs3_norm(int64_t n, const double *a, int bs, double *partials) {
  if (n <= 0) return 0.0;
  int64_t nb = ref_nblock(n, bs);
  double *s = (double *)malloc(sizeof(double) * bs);
  double *part = partials ? partials : (double *)malloc(sizeof(double) * nb);
  for (int64_t b = 0; b < nb; ++b) {
    for (int t = 0; t < bs; ++t) {
      double r = 0.0;
      for (int64_t id = t + b * bs; id < n; id += (int64_t)bs * nb) {
        double sq = a[id] * a[id];
        r += sq;
      }
      s[t] = r;
    }
    part[b] = ref_block_tree(s, bs);
  }
  double out = ref_finish(nb, part, bs);
  if (!partials) free(part);
  free(s);
  return out;
}

/* bs4_1 (bs4.okl:27-65) + bs4_2 */
double oracle_bs4_dot(int64_t n, const double *a, const double *b_, int bs, double *partials) {
  if (n <= 0) return 0.0;
  int64_t nb = ref_nblock(n, bs);
  double *s = (double *)malloc(sizeof(double) * bs);
  double *part = partials ? partials : (double *)malloc(sizeof(double) * nb);
  for (int64_t b = 0; b < nb; ++b) {
    for (int t = 0; t < bs; ++t) {
      double r = 0.0;
      for (int64_t id = t + b * bs; id < n; id += (int64_t)bs * nb) {
        double pr = a[id] * b_[id];
        r += pr;
      }
      s[t] = r;
    }
    part[b] = ref_block_tree(s, bs);
  }
  double out = ref_finish(nb, part, bs);
  if (!partials) free(part);
  free(s);
  return out;
}

/* bs5_1 (bs5.okl:27-67) + bs5_2 : r = r + alpha*Ap (stored), sum r^2 */
double oracle_bs5_update(int64_t n, const double *Ap, double alpha, double *r, int bs,
                         double *partials) {
  if (n <= 0) return 0.0;
  int64_t nb = ref_nblock(n, bs);
  double *s = (double *)malloc(sizeof(double) * bs);
  double *part = partials ? partials : (double *)malloc(sizeof(double) * nb);
  for (int64_t b = 0; b < nb; ++b) {
    for (int t = 0; t < bs; ++t) {
      double acc = 0.0;
      for (int64_t id = t + b * bs; id < n; id += (int64_t)bs * nb) {
        double aAp = alpha * Ap[id];
        double rr = r[id] + aAp;
        r[id] = rr;
        double sq = rr * rr;
        acc += sq;
      }
      s[t] = acc;
    }
    part[b] = ref_block_tree(s, bs);
  }
  double out = ref_finish(nb, part, bs);
  if (!partials) free(part);
  free(s);
  return out;
}

/* Full CG update of README.md:49-53 (no reference kernel): x += alpha p,
   r -= alpha Ap, rdotr = sum r^2 in the BS5 order. */
double oracle_bs5_full(int64_t n, double alpha, const double *p, const double *Ap,
                       double *x, double *r, int bs) {
  if (n <= 0) return 0.0;
  int64_t nb = ref_nblock(n, bs);
  double *s = (double *)malloc(sizeof(double) * bs);
  double *part = (double *)malloc(sizeof(double) * nb);
  for (int64_t b = 0; b < nb; ++b) {
    for (int t = 0; t < bs; ++t) {
      double acc = 0.0;
      for (int64_t id = t + b * bs; id < n; id += (int64_t)bs * nb) {
        double ap = alpha * p[id];
        x[id] = x[id] + ap;
        double aAp = alpha * Ap[id];
        double rr = r[id] - aAp;
        r[id] = rr;
        double sq = rr * rr;
        acc += sq;
      }
      s[t] = acc;
    }
    part[b] = ref_block_tree(s, bs);
  }
  double out = ref_finish(nb, part, bs);
  free(part);
  free(s);
  return out;
}

/* extended-precision sums for tolerance checks: mode 0 = sum a^2,
   1 = sum a*b, 2 = sum a */
double oracle_sum_longdouble(int64_t n, const double *a, const double *b, int mode) {
  long double acc = 0.0L;
  for (int64_t i = 0; i < n; ++i) {
    long double x = a[i];
    if (mode == 0) acc += x * x;
    else if (mode == 1) acc += x * (long double)b[i];
    else acc += x;
  }
  return (double)acc;
}

/* ------------------------------------------------------------------------ */
/* Hex box numbering                                                         */
/* ------------------------------------------------------------------------ */

/* Local index l = e*Np + n with e = ex + ey*nx + ez*nx*ny
   (meshSetupBoxHex3D.cpp:92) and n = i + j*Nq + k*Nq^2 (meshBasisHex3D.cpp:44).
   The node sits at GLL grid point (ex*p+i, ey*p+j, ez*p+k) of the
   non-periodic box (meshSetupBoxHex3D.cpp:95-99: (i+1)%NnX never wraps).
   Lexicographic id: 1 + gx + gy*Gx + gz*Gx*Gy. */
void oracle_hex_ids_lex(int nx, int ny, int nz, int p, int64_t *ids) {
  const int Nq = p + 1;
  const int64_t Np = (int64_t)Nq * Nq * Nq;
  const int64_t Gx = (int64_t)nx * p + 1, Gy = (int64_t)ny * p + 1;
  for (int ez = 0; ez < nz; ++ez)
    for (int ey = 0; ey < ny; ++ey)
      for (int ex = 0; ex < nx; ++ex) {
        const int64_t e = ex + (int64_t)ey * nx + (int64_t)ez * nx * ny;
        for (int k = 0; k < Nq; ++k)
          for (int j = 0; j < Nq; ++j)
            for (int i = 0; i < Nq; ++i) {
              const int64_t n = i + (int64_t)j * Nq + (int64_t)k * Nq * Nq;
              const int64_t gx = (int64_t)ex * p + i, gy = (int64_t)ey * p + j,
                            gz = (int64_t)ez * p + k;
              ids[e * Np + n] = 1 + gx + gy * Gx + gz * Gx * Gy;
            }
      }
}

/* Fixed point of meshConnectNodes.cpp:43-81: start from 1+l and take the
   minimum over coincident nodes until nothing changes.  Face adjacency links
   every copy of a grid node, so the fixed point is 1 + (smallest local index
   among the copies). */
void oracle_hex_ids_minlocal(int nx, int ny, int nz, int p, int64_t *ids) {
  const int64_t Nl = (int64_t)nx * ny * nz * (p + 1) * (p + 1) * (p + 1);
  const int64_t Ng = ((int64_t)nx * p + 1) * ((int64_t)ny * p + 1) * ((int64_t)nz * p + 1);
  int64_t *lex = (int64_t *)malloc(sizeof(int64_t) * Nl);
  int64_t *first = (int64_t *)malloc(sizeof(int64_t) * (Ng + 1));
  oracle_hex_ids_lex(nx, ny, nz, p, lex);
  for (int64_t g = 0; g <= Ng; ++g) first[g] = -1;
  for (int64_t l = 0; l < Nl; ++l)
    if (first[lex[l]] < 0) first[lex[l]] = l;
  for (int64_t l = 0; l < Nl; ++l) ids[l] = 1 + first[lex[l]];
  free(first);
  free(lex);
}

/* ------------------------------------------------------------------------ */
/* ogs setup (ogsSetup.cpp:87-290, 292-469, 492-579; ogsOperator.cpp:497-665) */
/* ------------------------------------------------------------------------ */

typedef struct { int64_t key; int32_t idx; } keyidx;

static int cmp_keyidx(const void *pa, const void *pb) {
  const keyidx *a = (const keyidx *)pa, *b = (const keyidx *)pb;
  if (a->key < b->key) return -1;
  if (a->key > b->key) return 1;
  return (a->idx > b->idx) - (a->idx < b->idx);
}

/* stable sort of keys; sortIds[i] = original position (primitives stableSort) */
static void stable_sort_ids(int32_t n, const int64_t *keys, int32_t *sortIds, int64_t *sorted) {
  keyidx *t = (keyidx *)malloc(sizeof(keyidx) * (n > 0 ? n : 1));
  for (int32_t i = 0; i < n; ++i) { t[i].key = keys[i]; t[i].idx = i; }
  qsort(t, (size_t)n, sizeof(keyidx), cmp_keyidx);
  for (int32_t i = 0; i < n; ++i) {
    sortIds[i] = t[i].idx;
    if (sorted) sorted[i] = t[i].key;
  }
  free(t);
}

/* runLengthEncodeConsecutive (runLengthEncode.cpp:98-123): rows sorted,
   offsets[d+1] = first index with value > d */
static void rle_consecutive(int32_t n, const int32_t *v, int32_t ngroups, int32_t *offset) {
  for (int32_t g = 0; g <= ngroups; ++g) offset[g] = 0;
  if (n <= 0) return;
  offset[0] = 0;
  for (int32_t i = 0; i <= n; ++i) {
    const int32_t a = (i == 0) ? 0 : v[i - 1];
    const int32_t b = (i == n) ? ngroups : v[i];
    for (int32_t d = a; d < b; ++d) offset[d + 1] = i;
  }
}

/* upperBound + blockRows (ogsOperator.cpp:500-594), gatherNodesPerBlock=512 */
static int32_t lower_bound_i32(int32_t first, int32_t last, const int32_t *v, int32_t val) {
  int32_t count = last - first;
  while (count > 0) {
    int32_t step = count / 2;
    int32_t mid = first + step;
    if (v[mid] < val) { first = mid + 1; count -= step + 1; }
    else count = step;
  }
  return first;
}

static int block_rows_ref(int32_t nrows, const int32_t *rowStarts, int32_t cap,
                          int32_t *nblocks_out, int32_t **blockStarts_out) {
  *nblocks_out = 0;
  *blockStarts_out = NULL;
  if (!nrows) return 0;
  int32_t maxRow = 0;
  for (int32_t r = 0; r < nrows; ++r) {
    int32_t m = rowStarts[r + 1] - rowStarts[r];
    if (m > maxRow) maxRow = m;
  }
  if (maxRow > cap) return 1; /* ogsOperator.cpp:535-536 LIBP_ABORT */
  int64_t maxNblocks = 1;
  while (maxNblocks < nrows) maxNblocks *= 2;
  int32_t *oldS = (int32_t *)malloc(sizeof(int32_t) * (maxNblocks + 1));
  int32_t *newS = (int32_t *)malloc(sizeof(int32_t) * (maxNblocks + 1));
  int32_t *sizes = (int32_t *)malloc(sizeof(int32_t) * maxNblocks);
  int64_t nb = 1;
  oldS[0] = 0; oldS[1] = nrows; newS[0] = 0;
  sizes[0] = rowStarts[nrows];
  int32_t maxSize = sizes[0];
  while (maxSize > cap) {
    newS[2 * nb] = nrows;
    for (int64_t n = 0; n < nb; ++n) {
      const int32_t start = oldS[n], end = oldS[n + 1];
      const int32_t rbs = rowStarts[end] - rowStarts[start];
      if (rbs > cap) {
        const int32_t midSize = rowStarts[start] + (rbs + 1) / 2;
        int32_t mid = lower_bound_i32(start, end, rowStarts, midSize);
        if (mid == end) --mid;
        newS[2 * n] = start; newS[2 * n + 1] = mid;
        sizes[2 * n] = rowStarts[mid] - rowStarts[start];
        sizes[2 * n + 1] = rowStarts[end] - rowStarts[mid];
      } else {
        newS[2 * n] = start; newS[2 * n + 1] = end;
        sizes[2 * n] = rbs; sizes[2 * n + 1] = 0;
      }
    }
    int32_t *tmp = oldS; oldS = newS; newS = tmp;
    nb *= 2;
    maxSize = 0;
    for (int64_t n = 0; n < nb; ++n) if (sizes[n] > maxSize) maxSize = sizes[n];
  }
  /* prim::unique over the sorted nb+1 boundaries */
  int32_t *bs = (int32_t *)malloc(sizeof(int32_t) * (nb + 1));
  int32_t nu = 0;
  for (int64_t i = 0; i <= nb; ++i)
    if (i == 0 || oldS[i] != oldS[i - 1]) bs[nu++] = oldS[i];
  *nblocks_out = nu - 1;
  *blockStarts_out = bs;
  free(oldS); free(newS); free(sizes);
  return 0;
}

oracle_ogs_t *oracle_ogs_setup(int32_t N, const int64_t *ids_in, int kind, int unique) {
  oracle_ogs_t *o = (oracle_ogs_t *)calloc(1, sizeof(oracle_ogs_t));
  o->N = N;
  o->ids = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
  if (N > 0) memcpy(o->ids, ids_in, sizeof(int64_t) * N);

  /* ogsSetup.cpp:116-118 */
  if ((kind == ORACLE_OGS_UNSIGNED && unique) || (kind == ORACLE_OGS_HALO && unique) ||
      kind == ORACLE_OGS_HALO) {
    o->error = 1;
    return o;
  }

  /* :121-139 select nonzero ids */
  int32_t Nids = 0;
  for (int32_t i = 0; i < N; ++i) Nids += (o->ids[i] != 0);
  int32_t *colIds = (int32_t *)malloc(sizeof(int32_t) * (Nids ? Nids : 1));
  int64_t *baseIds = (int64_t *)malloc(sizeof(int64_t) * (Nids ? Nids : 1));
  for (int32_t i = 0, c = 0; i < N; ++i)
    if (o->ids[i] != 0) { colIds[c] = i; baseIds[c] = o->ids[i]; ++c; }

  /* FindSharedGroups at one rank (:292-469): every group is local. */
  int64_t *absIds = (int64_t *)malloc(sizeof(int64_t) * (Nids ? Nids : 1));
  int32_t *sortIds = (int32_t *)malloc(sizeof(int32_t) * (Nids ? Nids : 1));
  int64_t *sortedAbs = (int64_t *)malloc(sizeof(int64_t) * (Nids ? Nids : 1));
  for (int32_t i = 0; i < Nids; ++i) absIds[i] = llabs(baseIds[i]);
  stable_sort_ids(Nids, absIds, sortIds, sortedAbs);
  int32_t ngroups = 0;
  for (int32_t i = 0; i < Nids; ++i) ngroups += (i == 0 || sortedAbs[i] != sortedAbs[i - 1]);
  int32_t *goff = (int32_t *)malloc(sizeof(int32_t) * (ngroups + 1));
  for (int32_t i = 0, g = 0; i < Nids; ++i)
    if (i == 0 || sortedAbs[i] != sortedAbs[i - 1]) goff[g++] = i;
  goff[ngroups] = Nids;

  int is_unique = 1;
  if (unique) {
    /* :361-382  one member per group stays positive; seedRNG(rank=0) at :113 */
    int32_t *rands = (int32_t *)malloc(sizeof(int32_t) * (ngroups ? ngroups : 1));
    oracle_fill_rand_int(ngroups, rands, 0, 0);
    for (int32_t g = 0; g < ngroups; ++g) {
      const int32_t start = goff[g], end = goff[g + 1];
      const int64_t base = sortedAbs[start];
      /* the reference computes rands[g] % size with a signed rand; a rand of
         INT_MIN would give a negative offset there, so it is taken unsigned */
      const int32_t m = (int32_t)((uint32_t)rands[g] % (uint32_t)(end - start));
      /* members in ascending local index (sortIds is a stable order) */
      for (int32_t i = start; i < end; ++i) baseIds[sortIds[i]] = -base;
      baseIds[sortIds[start + m]] = base;
    }
    free(rands);
    for (int32_t c = 0; c < Nids; ++c) o->ids[colIds[c]] = baseIds[c]; /* :149-151 */
  } else {
    for (int32_t g = 0; g < ngroups; ++g) {
      int pos = 0;
      for (int32_t i = goff[g]; i < goff[g + 1]; ++i) pos += (baseIds[sortIds[i]] > 0);
      if (pos != 1) is_unique = 0;
    }
  }
  o->gather_defined = is_unique;

  /* EnumerateGatherGroups (:492-559): first appearance, positive groups first */
  int32_t *gids = (int32_t *)calloc((size_t)(Nids ? Nids : 1), sizeof(int32_t));
  for (int32_t g = 0; g < ngroups; ++g) {
    int sign = -1;
    for (int32_t i = goff[g]; i < goff[g + 1]; ++i)
      if (baseIds[sortIds[i]] > 0) { sign = 1; break; }
    gids[sortIds[goff[g]]] = sign;
  }
  int32_t NP = 0;
  for (int32_t i = 0; i < Nids; ++i) NP += (gids[i] == 1);
  int32_t *firsts = (int32_t *)malloc(sizeof(int32_t) * (ngroups ? ngroups : 1));
  int32_t f = 0;
  for (int32_t i = 0; i < Nids; ++i) if (gids[i] == 1) firsts[f++] = i;
  for (int32_t i = 0; i < Nids; ++i) if (gids[i] == -1) firsts[f++] = i;
  for (int32_t g = 0; g < ngroups; ++g) gids[firsts[g]] = g;
  for (int32_t g = 0; g < ngroups; ++g) {
    const int32_t gid = gids[sortIds[goff[g]]];
    for (int32_t i = goff[g] + 1; i < goff[g + 1]; ++i) gids[sortIds[i]] = gid;
  }
  o->NrowsT = ngroups;
  o->NrowsN = NP;

  /* SortByRow (:561-579): stable sort by row */
  int64_t *rowKeys = (int64_t *)malloc(sizeof(int64_t) * (Nids ? Nids : 1));
  for (int32_t i = 0; i < Nids; ++i) rowKeys[i] = gids[i];
  int32_t *rs = (int32_t *)malloc(sizeof(int32_t) * (Nids ? Nids : 1));
  stable_sort_ids(Nids, rowKeys, rs, NULL);
  int32_t *rows = (int32_t *)malloc(sizeof(int32_t) * (Nids ? Nids : 1));
  int32_t *cols = (int32_t *)malloc(sizeof(int32_t) * (Nids ? Nids : 1));
  int64_t *bids = (int64_t *)malloc(sizeof(int64_t) * (Nids ? Nids : 1));
  for (int32_t i = 0; i < Nids; ++i) {
    rows[i] = gids[rs[i]];
    cols[i] = colIds[rs[i]];
    bids[i] = baseIds[rs[i]];
  }

  /* ogsOperator_t ctor (ogsOperator.cpp:597-665) */
  o->nnzT = Nids;
  o->rowStartsT = (int32_t *)malloc(sizeof(int32_t) * (o->NrowsT + 1));
  rle_consecutive(Nids, rows, o->NrowsT, o->rowStartsT);
  o->colIdsT = cols;
  if (kind == ORACLE_OGS_SIGNED) {
    int32_t nnzN = 0;
    for (int32_t i = 0; i < Nids; ++i) nnzN += (bids[i] > 0);
    int32_t *rowsN = (int32_t *)malloc(sizeof(int32_t) * (nnzN ? nnzN : 1));
    o->colIdsN = (int32_t *)malloc(sizeof(int32_t) * (nnzN ? nnzN : 1));
    for (int32_t i = 0, c = 0; i < Nids; ++i)
      if (bids[i] > 0) { rowsN[c] = rows[i]; o->colIdsN[c] = cols[i]; ++c; }
    o->nnzN = nnzN;
    o->rowStartsN = (int32_t *)malloc(sizeof(int32_t) * (o->NrowsN + 1));
    rle_consecutive(nnzN, rowsN, o->NrowsN, o->rowStartsN);
    free(rowsN);
  } else {
    o->nnzN = o->nnzT;
    o->rowStartsN = (int32_t *)malloc(sizeof(int32_t) * (o->NrowsT + 1));
    memcpy(o->rowStartsN, o->rowStartsT, sizeof(int32_t) * (o->NrowsT + 1));
    o->colIdsN = (int32_t *)malloc(sizeof(int32_t) * (Nids ? Nids : 1));
    memcpy(o->colIdsN, o->colIdsT, sizeof(int32_t) * Nids);
  }
  int err = block_rows_ref(o->NrowsT, o->rowStartsT, 512, &o->NblocksT, &o->blockStartsT);
  if (!err) {
    if (kind == ORACLE_OGS_SIGNED)
      err = block_rows_ref(o->NrowsN, o->rowStartsN, 512, &o->NblocksN, &o->blockStartsN);
    else {
      o->NblocksN = o->NblocksT;
      o->blockStartsN = (int32_t *)malloc(sizeof(int32_t) * (o->NblocksT + 1));
      if (o->blockStartsT)
        memcpy(o->blockStartsN, o->blockStartsT, sizeof(int32_t) * (o->NblocksT + 1));
    }
  }
  o->error = err ? 2 : 0;

  free(colIds); free(baseIds); free(absIds); free(sortIds); free(sortedAbs); free(goff);
  free(gids); free(firsts); free(rowKeys); free(rs); free(rows); free(bids);
  return o;
}

void oracle_ogs_free(oracle_ogs_t *o) {
  if (!o) return;
  free(o->rowStartsT); free(o->colIdsT); free(o->rowStartsN); free(o->colIdsN);
  free(o->blockStartsT); free(o->blockStartsN); free(o->ids);
  free(o);
}

/* ------------------------------------------------------------------------ */
/* ogs operator application (ogsOperator.cpp:37-56 ops, 61-114, 223-267, 340-409) */
/* ------------------------------------------------------------------------ */

#define OP_APPLY(op, acc, x)                      \
  do {                                            \
    switch (op) {                                 \
      case ORACLE_OP_ADD: acc += (x); break;      \
      case ORACLE_OP_MUL: acc *= (x); break;      \
      case ORACLE_OP_MAX: acc = ((x) > acc) ? (x) : acc; break; \
      case ORACLE_OP_MIN: acc = ((x) < acc) ? (x) : acc; break; \
    }                                             \
  } while (0)

#define DEFINE_OPS(T, NAME, TMAX)                                                         \
  static T init_##NAME(int op) {                                                          \
    switch (op) {                                                                         \
      case ORACLE_OP_MUL: return (T)1;                                                    \
      case ORACLE_OP_MAX: return -(TMAX);                                                 \
      case ORACLE_OP_MIN: return (TMAX);                                                  \
      default: return (T)0;                                                               \
    }                                                                                     \
  }                                                                                       \
  static void gather_##NAME(int32_t nrows, const int32_t *rs, const int32_t *ci, int op,  \
                            int K, const T *v, T *gv) {                                   \
    _Pragma("omp parallel for schedule(static)")                                          \
    for (int32_t n = 0; n < nrows; ++n)                                                   \
      for (int k = 0; k < K; ++k) {                                                       \
        T acc = init_##NAME(op);                                                          \
        for (int32_t g = rs[n]; g < rs[n + 1]; ++g) OP_APPLY(op, acc, v[k + (int64_t)ci[g] * K]); \
        gv[k + (int64_t)n * K] = acc;                                                     \
      }                                                                                   \
  }                                                                                       \
  static void scatter_##NAME(int32_t nrows, const int32_t *rs, const int32_t *ci, int K,  \
                             const T *gv, T *v) {                                         \
    _Pragma("omp parallel for schedule(static)")                                          \
    for (int32_t n = 0; n < nrows; ++n)                                                   \
      for (int32_t g = rs[n]; g < rs[n + 1]; ++g)                                         \
        for (int k = 0; k < K; ++k) v[k + (int64_t)ci[g] * K] = gv[k + (int64_t)n * K];   \
  }                                                                                       \
  static void gs_##NAME(int32_t nrows, const int32_t *grs, const int32_t *gci,            \
                        int32_t ngrows, const int32_t *srs, const int32_t *sci, int op,   \
                        int K, T *v) {                                                    \
    _Pragma("omp parallel for schedule(static)")                                          \
    for (int32_t n = 0; n < nrows; ++n)                                                   \
      for (int k = 0; k < K; ++k) {                                                       \
        T acc = init_##NAME(op);                                                          \
        if (n < ngrows)                                                                   \
          for (int32_t g = grs[n]; g < grs[n + 1]; ++g)                                   \
            OP_APPLY(op, acc, v[k + (int64_t)gci[g] * K]);                                \
        for (int32_t s = srs[n]; s < srs[n + 1]; ++s) v[k + (int64_t)sci[s] * K] = acc;   \
      }                                                                                   \
  }

DEFINE_OPS(float, f32, FLT_MAX)
DEFINE_OPS(double, f64, DBL_MAX)
DEFINE_OPS(int32_t, i32, INT32_MAX)
DEFINE_OPS(int64_t, i64, INT64_MAX)

void oracle_ogs_gather(const oracle_ogs_t *o, int type, int op, int trans, int K,
                       const void *v, void *gv) {
  const int useN = (trans == ORACLE_NOTRANS);
  const int32_t nrows = useN ? o->NrowsN : o->NrowsT;
  const int32_t *rs = useN ? o->rowStartsN : o->rowStartsT;
  const int32_t *ci = useN ? o->colIdsN : o->colIdsT;
  switch (type) {
    case ORACLE_FLOAT: gather_f32(nrows, rs, ci, op, K, (const float *)v, (float *)gv); break;
    case ORACLE_DOUBLE: gather_f64(nrows, rs, ci, op, K, (const double *)v, (double *)gv); break;
    case ORACLE_INT32: gather_i32(nrows, rs, ci, op, K, (const int32_t *)v, (int32_t *)gv); break;
    case ORACLE_INT64: gather_i64(nrows, rs, ci, op, K, (const int64_t *)v, (int64_t *)gv); break;
  }
}

void oracle_ogs_scatter(const oracle_ogs_t *o, int type, int trans, int K,
                        const void *gv, void *v) {
  const int useN = (trans == ORACLE_TRANS);
  const int32_t nrows = useN ? o->NrowsN : o->NrowsT;
  const int32_t *rs = useN ? o->rowStartsN : o->rowStartsT;
  const int32_t *ci = useN ? o->colIdsN : o->colIdsT;
  switch (type) {
    case ORACLE_FLOAT: scatter_f32(nrows, rs, ci, K, (const float *)gv, (float *)v); break;
    case ORACLE_DOUBLE: scatter_f64(nrows, rs, ci, K, (const double *)gv, (double *)v); break;
    case ORACLE_INT32: scatter_i32(nrows, rs, ci, K, (const int32_t *)gv, (int32_t *)v); break;
    case ORACLE_INT64: scatter_i64(nrows, rs, ci, K, (const int64_t *)gv, (int64_t *)v); break;
  }
}

/* ogsOperator.cpp:340-409.  Sym: gather T / scatter T over NrowsT rows.
   Trans: gather T / scatter N over NrowsN rows.  NoTrans: gather N / scatter T
   over NrowsT rows; rows past NrowsN have no N entries and receive the op's
   identity (the reference reads rowStartsN out of range there). */
void oracle_ogs_gather_scatter(const oracle_ogs_t *o, int type, int op, int trans, int K, void *v) {
  int32_t nrows, ngrows;
  const int32_t *grs, *gci, *srs, *sci;
  if (trans == ORACLE_TRANS) {
    nrows = o->NrowsN; ngrows = o->NrowsN;
    grs = o->rowStartsT; gci = o->colIdsT; srs = o->rowStartsN; sci = o->colIdsN;
  } else if (trans == ORACLE_SYM) {
    nrows = o->NrowsT; ngrows = o->NrowsT;
    grs = o->rowStartsT; gci = o->colIdsT; srs = o->rowStartsT; sci = o->colIdsT;
  } else {
    nrows = o->NrowsT; ngrows = o->NrowsN;
    grs = o->rowStartsN; gci = o->colIdsN; srs = o->rowStartsT; sci = o->colIdsT;
  }
  switch (type) {
    case ORACLE_FLOAT: gs_f32(nrows, grs, gci, ngrows, srs, sci, op, K, (float *)v); break;
    case ORACLE_DOUBLE: gs_f64(nrows, grs, gci, ngrows, srs, sci, op, K, (double *)v); break;
    case ORACLE_INT32: gs_i32(nrows, grs, gci, ngrows, srs, sci, op, K, (int32_t *)v); break;
    case ORACLE_INT64: gs_i64(nrows, grs, gci, ngrows, srs, sci, op, K, (int64_t *)v); break;
  }
}

/* Sym gatherScatter (Add, double) over a caller-given CSR: the host loop of
   ogsOperator.cpp:374-389 on arrays not built by oracle_ogs_setup. */
void oracle_csr_gs(int32_t nrows, const int32_t *rowStarts, const int32_t *colIds, int K, double *q) {
  gs_f64(nrows, rowStarts, colIds, nrows, rowStarts, colIds, ORACLE_OP_ADD, K, q);
}

/* Gather (Add, double) over a caller-given CSR: gv[row] = 0 + q[c0] + q[c1] ...
   over the row's columns in stored order, the host loop of
   ogsOperator.cpp:61-114 (Trans gather: the T-CSR) on arrays not built by
   oracle_ogs_setup. */
void oracle_csr_gather(int32_t nrows, const int32_t *rowStarts, const int32_t *colIds, int K,
                       const double *q, double *gv) {
  gather_f64(nrows, rowStarts, colIds, ORACLE_OP_ADD, K, q, gv);
}

/* Scatter (double) over a caller-given CSR: q[col] = gv[row] for every column
   of every row, other entries untouched (ogsOperator.cpp:223-267, NoTrans
   scatter: the T-CSR). */
void oracle_csr_scatter(int32_t nrows, const int32_t *rowStarts, const int32_t *colIds, int K,
                        const double *gv, double *q) {
  scatter_f64(nrows, rowStarts, colIds, K, gv, q);
}

/* ogsKernels.okl:32-81 block phases, p_gatherNodesPerBlock = 512, op Add */
void oracle_ogs_gs_okl(int32_t nblocks, int K, const int32_t *blockStarts,
                       const int32_t *gStarts, const int32_t *gIds,
                       const int32_t *sStarts, const int32_t *sIds, double *q) {
  double gtemp[512], stemp[512];
  for (int k = 0; k < K; ++k)
    for (int32_t b = 0; b < nblocks; ++b) {
      const int32_t bStart = blockStarts[b], bEnd = blockStarts[b + 1];
      const int32_t gStart = gStarts[bStart], sStart = sStarts[bStart];
      for (int32_t id = gStart; id < gStarts[bEnd]; ++id)
        gtemp[id - gStart] = q[k + (int64_t)gIds[id] * K];
      for (int32_t row = bStart; row < bEnd; ++row) {
        double gq = 0.0;
        for (int32_t i = gStarts[row] - gStart; i < gStarts[row + 1] - gStart; ++i) gq += gtemp[i];
        for (int32_t i = sStarts[row] - sStart; i < sStarts[row + 1] - sStart; ++i) stemp[i] = gq;
      }
      for (int32_t id = sStart; id < sStarts[bEnd]; ++id)
        q[k + (int64_t)sIds[id] * K] = stemp[id - sStart];
    }
}

/* ------------------------------------------------------------------------ */
/* Reduced-precision storage                                                 */
/* ------------------------------------------------------------------------ */

/* nearest-even bf16 of a double, decided exactly: candidates are the bf16
   values just below and above x in magnitude (bf16 values are doubles) */
uint16_t oracle_bf16_from_double(double x) {
  if (x != x) return 0x7fc0;
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  uint32_t lo = u & 0xffff0000u;
  float flo;
  memcpy(&flo, &lo, 4);
  if (fabs((double)flo) > fabs(x)) lo -= 0x10000u; /* f rounded up past x in magnitude */
  uint32_t hi = lo + 0x10000u;
  float fl, fh;
  memcpy(&fl, &lo, 4);
  memcpy(&fh, &hi, 4);
  /* above the largest bf16 the next candidate is 2^128 in magnitude, i.e.
     infinity when x rounds to it (round to nearest, IEEE 754 7.4) */
  const double fhv = ((hi & 0x7fffffffu) == 0x7f800000u) ? copysign(0x1p128, x) : (double)fh;
  const double dl = fabs(x - (double)fl), dh = fabs(fhv - x);
  uint32_t r;
  if (dl < dh) r = lo;
  else if (dh < dl) r = hi;
  else r = ((lo >> 16) & 1u) ? hi : lo;
  return (uint16_t)(r >> 16);
}

uint16_t oracle_bf16_from_float(float x) { return oracle_bf16_from_double((double)x); }

double oracle_bf16_to_double(uint16_t b) {
  uint32_t u = ((uint32_t)b) << 16;
  float f;
  memcpy(&f, &u, 4);
  return (double)f;
}

static double load_s(int storage, const void *q, int64_t i) {
  if (storage == ORACLE_STORE_F32) return (double)((const float *)q)[i];
  if (storage == ORACLE_STORE_BF16) return oracle_bf16_to_double(((const uint16_t *)q)[i]);
  return ((const double *)q)[i];
}

static void store_s(int storage, void *q, int64_t i, double v) {
  if (storage == ORACLE_STORE_F32) ((float *)q)[i] = (float)v;
  else if (storage == ORACLE_STORE_BF16) ((uint16_t *)q)[i] = oracle_bf16_from_double(v);
  else ((double *)q)[i] = v;
}

void oracle_gs_mixed(int storage, int32_t nrows, const int32_t *rs, const int32_t *ci, int K, void *q) {
  for (int32_t n = 0; n < nrows; ++n)
    for (int k = 0; k < K; ++k) {
      double acc = 0.0;
      for (int32_t g = rs[n]; g < rs[n + 1]; ++g) acc += load_s(storage, q, k + (int64_t)ci[g] * K);
      for (int32_t g = rs[n]; g < rs[n + 1]; ++g) store_s(storage, q, k + (int64_t)ci[g] * K, acc);
    }
}

static float load_f(int storage, const void *q, int64_t i) { return (float)load_s(storage, q, i); }

static void store_f(int storage, void *q, int64_t i, float v) {
  if (storage == ORACLE_STORE_BF16) ((uint16_t *)q)[i] = oracle_bf16_from_float(v);
  else if (storage == ORACLE_STORE_F32) ((float *)q)[i] = v;
  else ((double *)q)[i] = (double)v;
}

double oracle_cg_mixed(int storage, int64_t n, double alpha, const void *p, const void *Ap, void *x, void *r) {
  const float a = (float)alpha;
  long double acc = 0.0L;
  for (int64_t i = 0; i < n; ++i) {
    const float ap = a * load_f(storage, p, i);
    const float xn = load_f(storage, x, i) + ap;
    const float aAp = a * load_f(storage, Ap, i);
    const float rn = load_f(storage, r, i) - aAp;
    store_f(storage, x, i, xn);
    store_f(storage, r, i, rn);
    const long double rs = load_s(storage, r, i);
    acc += rs * rs;
  }
  return (double)acc;
}

/* ------------------------------------------------------------------------ */
/* CPU baseline timers                                                       */
/* ------------------------------------------------------------------------ */

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static void set_threads(int threads) {
#ifdef _OPENMP
  omp_set_num_threads(threads > 0 ? threads : 1);
#else
  (void)threads;
#endif
}

/* BS6 (ogs.Gather(gq, q, 1, Add, Trans)), BS7 (ogs.Scatter(q, gq, 1,
   NoTrans)) and BS8 (ogs.GatherScatter(q, 1, Add, Sym)) on one nx^3 hex box
   at degree p, q = U[-1,1) from seed 0 (bs6Run.cpp:41-50, bs7Run.cpp:41-50,
   bs8Run.cpp:41-49).  For each of the nthr thread counts threads[t], each
   kernel runs once untimed and then `reps` timed calls: seconds[3 t + k] is
   the time per call of BS6, BS7, BS8 (k = 0, 1, 2).  checksums[k] is the
   long-double sum of the first call's output: gq for BS6, and for BS7 and BS8
   q after scattering that gq (the same as one gather+scatter of q). */
void oracle_time_ogs(int nx, int p, int reps, int nthr, const int *threads, double *seconds,
                     double *checksums, double *setup_seconds, int64_t *nlocal, int64_t *nglobal) {
  const int64_t Nl = (int64_t)nx * nx * nx * (p + 1) * (p + 1) * (p + 1);
  int64_t *ids = (int64_t *)malloc(sizeof(int64_t) * Nl);
  set_threads(nthr > 0 ? threads[0] : 1);
  double t0 = now_seconds();
  oracle_hex_ids_lex(nx, nx, nx, p, ids);
  oracle_ogs_t *o = oracle_ogs_setup((int32_t)Nl, ids, ORACLE_OGS_SIGNED, 1);
  if (setup_seconds) *setup_seconds = now_seconds() - t0;
  double *q0 = (double *)malloc(sizeof(double) * Nl);
  double *q = (double *)malloc(sizeof(double) * Nl);
  double *gq = (double *)malloc(sizeof(double) * (o->NrowsT > 0 ? o->NrowsT : 1));
  oracle_fill_uniform(Nl, q0, 0, 0);
  for (int t = 0; t < nthr; ++t) {
    set_threads(threads[t]);
    for (int k = 0; k < 3; ++k) {
      memcpy(q, q0, sizeof(double) * Nl);
      if (k == 1) oracle_ogs_gather(o, ORACLE_DOUBLE, ORACLE_OP_ADD, ORACLE_TRANS, 1, q, gq);
      for (int r = -1; r < reps; ++r) {
        if (r == 0) t0 = now_seconds();
        if (k == 0) oracle_ogs_gather(o, ORACLE_DOUBLE, ORACLE_OP_ADD, ORACLE_TRANS, 1, q, gq);
        else if (k == 1) oracle_ogs_scatter(o, ORACLE_DOUBLE, ORACLE_NOTRANS, 1, gq, q);
        else oracle_ogs_gather_scatter(o, ORACLE_DOUBLE, ORACLE_OP_ADD, ORACLE_SYM, 1, q);
        if (r == -1 && t == 0 && checksums) {
          const double *out = (k == 0) ? gq : q;
          const int64_t n = (k == 0) ? o->NrowsT : Nl;
          long double sum = 0.0L;
          for (int64_t i = 0; i < n; ++i) sum += out[i];
          checksums[k] = (double)sum;
        }
      }
      seconds[3 * t + k] = (now_seconds() - t0) / (reps > 0 ? reps : 1);
    }
  }
  if (nlocal) *nlocal = Nl;
  if (nglobal) *nglobal = o->NrowsT;
  oracle_ogs_free(o);
  free(gq);
  free(q);
  free(q0);
  free(ids);
}

/* BS8 alone: seconds per call; *checksum as oracle_time_ogs's BS8 checksum */
double oracle_time_bs8(int nx, int p, int reps, int threads, double *checksum,
                       double *setup_seconds, int64_t *nlocal, int64_t *nglobal) {
  double sec[3], cs[3];
  oracle_time_ogs(nx, p, reps, 1, &threads, sec, cs, setup_seconds, nlocal, nglobal);
  if (checksum) *checksum = cs[2];
  return sec[2];
}

/* BS1..BS5 on n doubles, seconds per call.  threads==1 runs the Serial
   restatement; threads>1 runs the same element loops under OpenMP (the
   reductions then sum per-thread slices; only the timing is used). */
double oracle_time_stream(int which, int64_t n, int reps, int threads) {
  set_threads(threads);
  double *a = (double *)malloc(sizeof(double) * n);
  double *b = (double *)malloc(sizeof(double) * n);
  oracle_fill_uniform(n, a, 1, 0);
  oracle_fill_uniform(n, b, 1, (uint64_t)n);
  volatile double sink = 0.0;
  double t0 = 0.0;
  for (int r = -1; r < reps; ++r) {
    if (r == 0) t0 = now_seconds();
    if (threads <= 1) {
      switch (which) {
        case 1: oracle_bs1_copy(n, a, b); break;
        case 2: oracle_bs2_axpy(n, 0.37, a, -1.21, b); break;
        case 3: sink += oracle_bs3_norm(n, a, 256, NULL); break;
        case 4: sink += oracle_bs4_dot(n, a, b, 256, NULL); break;
        case 5: sink += oracle_bs5_update(n, a, 1e-9, b, 256, NULL); break;
      }
    } else {
      double acc = 0.0;
      switch (which) {
        case 1:
          _Pragma("omp parallel for schedule(static)")
          for (int64_t i = 0; i < n; ++i) b[i] = a[i];
          break;
        case 2:
          _Pragma("omp parallel for schedule(static)")
          for (int64_t i = 0; i < n; ++i) { double t1 = 0.37 * a[i]; double t2 = -1.21 * b[i]; b[i] = t1 + t2; }
          break;
        case 3:
          _Pragma("omp parallel for schedule(static) reduction(+:acc)")
          for (int64_t i = 0; i < n; ++i) acc += a[i] * a[i];
          break;
        case 4:
          _Pragma("omp parallel for schedule(static) reduction(+:acc)")
          for (int64_t i = 0; i < n; ++i) acc += a[i] * b[i];
          break;
        case 5:
          _Pragma("omp parallel for schedule(static) reduction(+:acc)")
          for (int64_t i = 0; i < n; ++i) { double rr = b[i] + 1e-9 * a[i]; b[i] = rr; acc += rr * rr; }
          break;
      }
      sink += acc;
    }
  }
  double dt = (now_seconds() - t0) / (reps > 0 ? reps : 1);
  (void)sink;
  free(a);
  free(b);
  return dt;
}
