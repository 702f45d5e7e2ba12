// C ABI of the ogs kernels (include/spx.h): argument checks, tile selection
// and dispatch to the per-type launchers (spx_ogs_inst.hip).
#include <string>

#include "spx_ogs_kernels.hpp"

using namespace spx;
using namespace spx::ogsk;

namespace {

constexpr int kMaxBlockNnz = 4096;

size_t type_size(int type) {
  switch (type) {
    case SPX_FLOAT: return 4;
    case SPX_DOUBLE: return 8;
    case SPX_INT32: return 4;
    case SPX_INT64: return 8;
    default: return 0;
  }
}

// Measured on MI355X (BS8, 64^3 hex at p=7).  CSR blocks: 16-bit row starts
// in LDS (variant 3) 4.9-5.0 TB/s against 4.4-5.0 for ranges in registers.
// Sorted-run tiles: 256 lanes (variant 0) with XCD remap, 5.8 TB/s at 2048
// positions, against 5.1 for one-wave tiles of 512 (variant 2); one-wave tiles
// stay ahead only below p=5.  The remap gains 2-40%.
spx_tuning_t ogs_defaults(const char *kname) {
  spx_tuning_t d{};
  const std::string k(kname);
  d.block_threads = kThreads;
  d.xcd_remap = k == "ogs_gs_srb" ? 1 : 0;
  d.variant = k == "ogs_gs" ? 3 : 0;
  return d;
}

int check_common(const char *name, int type, int K, int32_t nblocks, int32_t cap) {
  if (!type_size(type)) return fail(SPX_ERR_ARG, std::string(name) + ": unknown type");
  if (K < 1) return fail(SPX_ERR_ARG, std::string(name) + ": K must be >= 1");
  if (K > 65535) return fail(SPX_ERR_ARG, std::string(name) + ": K must be <= 65535");
  if (nblocks < 0) return fail(SPX_ERR_ARG, std::string(name) + ": negative nblocks");
  if (cap < 1 || cap > kMaxBlockNnz)
    return fail(SPX_ERR_ARG, std::string(name) + ": max_block_nnz must be in [1, 4096]");
  return SPX_OK;
}

int check_op(const char *name, int op) {
  if (op < SPX_ADD || op > SPX_MIN) return fail(SPX_ERR_ARG, std::string(name) + ": unknown op");
  return SPX_OK;
}

// Tile shape: all K fields per pass when K <= 4 and the tile fits the lanes'
// registers and 64 KiB of LDS; else one field per grid row.
int make_launch(const char *name, const char *kname, int32_t nblocks, int K, int32_t cap, int type, int arrays,
                Launch &L) {
  spx_tuning_t t = tuning_for(kname, ogs_defaults(kname));
  const size_t ts = type_size(type);
  int kf = 1;
  if (K >= 2 && K <= 4 && pick_per(K, cap) && lds_bytes(cap, K, ts, arrays) <= 65536) kf = K;
  const int per = pick_per(kf, cap);
  if (!per) return fail(SPX_ERR_ARG, std::string(name) + ": max_block_nnz " + std::to_string(cap) + " too large");
  const size_t lds = lds_bytes(cap, kf, ts, arrays);
  if (lds > 65536)
    return fail(SPX_ERR_ARG, std::string(name) + ": max_block_nnz " + std::to_string(cap) + " needs " +
                                 std::to_string(lds) + " bytes of LDS per workgroup (limit 65536)");
  L.nblocks = nblocks;
  L.K = K;
  L.nxcd = t.xcd_remap ? 8 : 1;
  L.cap = cap;
  L.kf = kf;
  L.per = per;
  L.variant = t.variant < 0 ? 0 : t.variant;
  L.stream = current_stream();
  L.err = kernel_error_flag();
  return SPX_OK;
}

template <typename F>
int by_type(int type, F &&f) {
  switch (type) {
    case SPX_FLOAT: return f((float *)nullptr);
    case SPX_DOUBLE: return f((double *)nullptr);
    case SPX_INT32: return f((int *)nullptr);
    default: return f((long long *)nullptr);
  }
}

}  // namespace

extern "C" {

int spx_ogs_gather(int type, int op, int K, int32_t nblocks, const int32_t *blockStarts,
                   const int32_t *rowStarts, const int32_t *colIds, int32_t cap, const void *v, void *gv) {
  SPX_NEED_DEVICE();
  int rc = check_common("spx_ogs_gather", type, K, nblocks, cap);
  if (rc || (rc = check_op("spx_ogs_gather", op))) return rc;
  if (nblocks == 0) return SPX_OK;
  if (!blockStarts || !rowStarts || !colIds || !v || !gv)
    return fail(SPX_ERR_ARG, "spx_ogs_gather: null pointer");
  Launch L;
  if ((rc = make_launch("spx_ogs_gather", "ogs_gather", nblocks, K, cap, type, 1, L))) return rc;
  rc = by_type(type, [&](auto *tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return launch_gather(L, op, blockStarts, rowStarts, colIds, (const T *)v, (T *)gv);
  });
  if (rc) return rc;
  SPX_LAUNCHED("spx_ogs_gather");
  return SPX_OK;
}

int spx_ogs_scatter(int type, int K, int32_t nblocks, const int32_t *blockStarts, const int32_t *rowStarts,
                    const int32_t *colIds, int32_t cap, const void *gv, void *v) {
  SPX_NEED_DEVICE();
  int rc = check_common("spx_ogs_scatter", type, K, nblocks, cap);
  if (rc) return rc;
  if (nblocks == 0) return SPX_OK;
  if (!blockStarts || !rowStarts || !colIds || !v || !gv)
    return fail(SPX_ERR_ARG, "spx_ogs_scatter: null pointer");
  Launch L;
  if ((rc = make_launch("spx_ogs_scatter", "ogs_scatter", nblocks, K, cap, type, 1, L))) return rc;
  rc = by_type(type, [&](auto *tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return launch_scatter(L, blockStarts, rowStarts, colIds, (const T *)gv, (T *)v);
  });
  if (rc) return rc;
  SPX_LAUNCHED("spx_ogs_scatter");
  return SPX_OK;
}

int spx_ogs_gather_scatter(int type, int op, int K, int32_t nblocks, const int32_t *blockStarts, int32_t ngrows,
                           const int32_t *gStarts, const int32_t *gIds, int32_t nsrows, const int32_t *sStarts,
                           const int32_t *sIds, int32_t cap, void *q) {
  SPX_NEED_DEVICE();
  int rc = check_common("spx_ogs_gather_scatter", type, K, nblocks, cap);
  if (rc || (rc = check_op("spx_ogs_gather_scatter", op))) return rc;
  if (ngrows < 0 || nsrows < 0) return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter: negative row count");
  if (nblocks == 0) return SPX_OK;
  if (!blockStarts || !gStarts || !gIds || !sStarts || !sIds || !q)
    return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter: null pointer");
  const bool sym = (gStarts == sStarts) && (gIds == sIds) && (ngrows == nsrows);
  Launch L;
  if ((rc = make_launch("spx_ogs_gather_scatter", "ogs_gs", nblocks, K, cap, type, sym ? 1 : 2, L))) return rc;
  rc = by_type(type, [&](auto *tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    if (sym) return launch_gs_sym(L, op, blockStarts, gStarts, gIds, (T *)q);
    return launch_gs_gen(L, op, ngrows, nsrows, blockStarts, gStarts, gIds, sStarts, sIds, (T *)q);
  });
  if (rc) return rc;
  SPX_LAUNCHED("spx_ogs_gather_scatter");
  return SPX_OK;
}

namespace {
// common checks and launch shape of the sorted-run entry points
int srb_launch_setup(const char *name, int type, int K, const spx_srb_t *srb, Launch &L, SrbArgs &A) {
  if (!srb) return fail(SPX_ERR_ARG, std::string(name) + ": null layout");
  if (srb->chunks_per_tile < 1 || srb->chunks_per_tile > kMaxBlockNnz / 64)
    return fail(SPX_ERR_ARG, std::string(name) + ": chunks_per_tile must be in [1, 64]");
  const int cap = 64 * srb->chunks_per_tile;
  int rc = check_common(name, type, K, srb->nblocks, cap);
  if (rc) return rc;
  if (srb->nblocks == 0) return SPX_OK;
  if (!srb->chunks || !srb->runIds || !srb->links) return fail(SPX_ERR_ARG, std::string(name) + ": null layout array");
  if (srb->tile_lanes != 0 && srb->tile_lanes != 64 && srb->tile_lanes != 256)
    return fail(SPX_ERR_ARG, std::string(name) + ": tile_lanes must be 0, 64 or 256");
  if ((rc = make_launch(name, "ogs_gs_srb", srb->nblocks, K, cap, type, 1, L))) return rc;
  if (srb->xcd_remap >= 0 && !xcd_hint(srb->xcd_remap, L.nxcd > 1, L.nxcd, L.xcdLog2))
    return fail(SPX_ERR_ARG, std::string(name) + ": xcd_remap must be negative (library default), 0, 1 or a power of two up to 1024");
  if (srb->tile_lanes) L.variant = srb->tile_lanes == 64 ? 2 : 0;
  if (srb->nontemporal & ~0x3e)
    return fail(SPX_ERR_ARG, std::string(name) + ": nontemporal takes bits 1..5 (nodes of 2..32 bytes)");
  L.ntMask = srb->nontemporal;
  A.chunks = reinterpret_cast<const SrbChunk *>(srb->chunks);
  A.runIds = srb->runIds;
  A.links = srb->links;
  return SPX_OK;
}
}  // namespace

int spx_ogs_gather_scatter_srb(int type, int op, int K, const spx_srb_t *srb, void *q) {
  SPX_NEED_DEVICE();
  Launch L;
  SrbArgs A;
  int rc = srb_launch_setup("spx_ogs_gather_scatter_srb", type, K, srb, L, A);
  if (rc || (rc = check_op("spx_ogs_gather_scatter_srb", op))) return rc;
  if (srb->nblocks == 0) return SPX_OK;
  if (!q) return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter_srb: null pointer");
  rc = by_type(type, [&](auto *tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return launch_gs_srb(L, op, A, (T *)q, kAll, nullptr);
  });
  if (rc) return rc;
  SPX_LAUNCHED("spx_ogs_gather_scatter_srb");
  return SPX_OK;
}

int spx_ogs_gather_srb(int type, int op, int K, const spx_srb_t *srb, const void *v, void *gv) {
  SPX_NEED_DEVICE();
  Launch L;
  SrbArgs A;
  int rc = srb_launch_setup("spx_ogs_gather_srb", type, K, srb, L, A);
  if (rc || (rc = check_op("spx_ogs_gather_srb", op))) return rc;
  if (srb->nblocks == 0) return SPX_OK;
  if (!v || !gv) return fail(SPX_ERR_ARG, "spx_ogs_gather_srb: null pointer");
  rc = by_type(type, [&](auto *tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return launch_gather_srb(L, op, A, (const T *)v, (T *)gv, kAll, nullptr);
  });
  if (rc) return rc;
  SPX_LAUNCHED("spx_ogs_gather_srb");
  return SPX_OK;
}

int spx_ogs_scatter_srb(int type, int K, const spx_srb_t *srb, const void *gv, void *v) {
  SPX_NEED_DEVICE();
  Launch L;
  SrbArgs A;
  int rc = srb_launch_setup("spx_ogs_scatter_srb", type, K, srb, L, A);
  if (rc) return rc;
  if (srb->nblocks == 0) return SPX_OK;
  if (!v || !gv) return fail(SPX_ERR_ARG, "spx_ogs_scatter_srb: null pointer");
  rc = by_type(type, [&](auto *tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return launch_scatter_srb(L, A, (const T *)gv, (T *)v, kAll, nullptr);
  });
  if (rc) return rc;
  SPX_LAUNCHED("spx_ogs_scatter_srb");
  return SPX_OK;
}

int spx_ogs_gather_scatter_srb_signed(int type, int op, int trans, int K, const spx_srb_t *srb,
                                      const uint64_t *memberN, void *q) {
  SPX_NEED_DEVICE();
  Launch L;
  SrbArgs A;
  const char *name = "spx_ogs_gather_scatter_srb_signed";
  int rc = srb_launch_setup(name, type, K, srb, L, A);
  if (rc || (rc = check_op(name, op))) return rc;
  if (trans != SPX_SYM && trans != SPX_NOTRANS && trans != SPX_TRANS)
    return fail(SPX_ERR_ARG, std::string(name) + ": unknown transpose");
  if (srb->nblocks == 0) return SPX_OK;
  if (!q || (trans != SPX_SYM && !memberN)) return fail(SPX_ERR_ARG, std::string(name) + ": null pointer");
  const int mode = trans == SPX_NOTRANS ? kFoldN : trans == SPX_TRANS ? kStoreN : kAll;
  rc = by_type(type, [&](auto *tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return launch_gs_srb(L, op, A, (T *)q, mode, reinterpret_cast<const unsigned long long *>(memberN));
  });
  if (rc) return rc;
  SPX_LAUNCHED(name);
  return SPX_OK;
}

int spx_ogs_gather_srb_signed(int type, int op, int K, const spx_srb_t *srb, const uint64_t *memberN,
                              int32_t nrowsN, const void *v, void *gv) {
  SPX_NEED_DEVICE();
  Launch L;
  SrbArgs A;
  const char *name = "spx_ogs_gather_srb_signed";
  int rc = srb_launch_setup(name, type, K, srb, L, A);
  if (rc || (rc = check_op(name, op))) return rc;
  if (nrowsN < 0) return fail(SPX_ERR_ARG, std::string(name) + ": negative row count");
  if (srb->nblocks == 0 || nrowsN == 0) return SPX_OK;
  if (!memberN || !v || !gv) return fail(SPX_ERR_ARG, std::string(name) + ": null pointer");
  L.rows = nrowsN;
  rc = by_type(type, [&](auto *tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return launch_gather_srb(L, op, A, (const T *)v, (T *)gv, kFoldN,
                             reinterpret_cast<const unsigned long long *>(memberN));
  });
  if (rc) return rc;
  SPX_LAUNCHED(name);
  return SPX_OK;
}

int spx_ogs_scatter_srb_signed(int type, int K, const spx_srb_t *srb, const uint64_t *memberN, int32_t nrowsN,
                               const void *gv, void *v) {
  SPX_NEED_DEVICE();
  Launch L;
  SrbArgs A;
  const char *name = "spx_ogs_scatter_srb_signed";
  int rc = srb_launch_setup(name, type, K, srb, L, A);
  if (rc) return rc;
  if (nrowsN < 0) return fail(SPX_ERR_ARG, std::string(name) + ": negative row count");
  if (srb->nblocks == 0 || nrowsN == 0) return SPX_OK;
  if (!memberN || !v || !gv) return fail(SPX_ERR_ARG, std::string(name) + ": null pointer");
  L.rows = nrowsN;
  rc = by_type(type, [&](auto *tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return launch_scatter_srb(L, A, (const T *)gv, (T *)v, kStoreN,
                              reinterpret_cast<const unsigned long long *>(memberN));
  });
  if (rc) return rc;
  SPX_LAUNCHED(name);
  return SPX_OK;
}

int spx_ogs_scatter_inverse(int type, int K, int32_t n, const int32_t *g2l, const void *gv, void *v) {
  SPX_NEED_DEVICE();
  if (!type_size(type)) return fail(SPX_ERR_ARG, "spx_ogs_scatter_inverse: unknown type");
  if (K < 1 || K > 65535 || n < 0) return fail(SPX_ERR_ARG, "spx_ogs_scatter_inverse: bad K or n");
  if (n == 0) return SPX_OK;
  if (!g2l || !gv || !v) return fail(SPX_ERR_ARG, "spx_ogs_scatter_inverse: null pointer");
  int rc = by_type(type, [&](auto *tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    spx_tuning_t t = tuning_for("ogs_scatter_inverse", spx_tuning_t{});
    return launch_scatter_inv(current_stream(), (int64_t)n, K, g2l, (const T *)gv, (T *)v, t.unroll > 0 ? t.unroll : 4);
  });
  if (rc) return rc;
  SPX_LAUNCHED("spx_ogs_scatter_inverse");
  return SPX_OK;
}

int spx_ogs_extract(int type, int K, int32_t n_ids, const int32_t *ids, const void *q, void *gq) {
  SPX_NEED_DEVICE();
  if (!type_size(type)) return fail(SPX_ERR_ARG, "spx_ogs_extract: unknown type");
  if (K < 1 || n_ids < 0) return fail(SPX_ERR_ARG, "spx_ogs_extract: bad K or n_ids");
  if (n_ids == 0) return SPX_OK;
  if (!ids || !q || !gq) return fail(SPX_ERR_ARG, "spx_ogs_extract: null pointer");
  const int64_t total = (int64_t)n_ids * K;
  int rc = by_type(type, [&](auto *tp) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return launch_extract(current_stream(), total, K, ids, (const T *)q, (T *)gq);
  });
  if (rc) return rc;
  SPX_LAUNCHED("spx_ogs_extract");
  return SPX_OK;
}

}  // extern "C"
