// C ABI of the host layer (include/spx_host.h): ogs_t and mesh ids for
// non-C++ callers.  Exceptions become SPX_ERR_* codes plus a message.
#include <string>

#include "mesh.hpp"
#include "ogs.hpp"
#include "spx_host.h"

using namespace libp;

namespace {

thread_local std::string g_err;

struct host_state {
  settings_t settings{comm_t::world()};
  platform_t platform;
  bool open = false;
};
host_state &state() {
  static host_state s;
  return s;
}

int err(int code, const std::string &m) {
  g_err = m;
  return code;
}

template <typename F>
int guarded(F &&f) {
  try {
    f();
    return SPX_OK;
  } catch (const libp::exception &e) {
    return err(SPX_ERR_ARG, e.message());
  } catch (const std::bad_alloc &) {
    return err(SPX_ERR_ARG, "host allocation failed");
  } catch (const std::exception &e) {
    return err(SPX_ERR_ARG, e.what());
  }
}

}  // namespace

struct spxh_ogs_s {
  ogs::ogs_t ogs;
};

extern "C" {

const char *spxh_last_error(void) { return g_err.c_str(); }

int spxh_platform_open(int device_id) {
  return guarded([&] {
    host_state &s = state();
    if (s.open) {
      LIBP_ABORT("platform already open on device " << s.platform.device.id(), s.platform.device.id() != device_id);
      return;
    }
    platformAddSettings(s.settings);
    s.settings.changeSetting("DEVICE NUMBER", std::to_string(device_id));
    s.platform = platform_t(s.settings);
    s.open = true;
  });
}

int spxh_ogs_setup(int32_t n, int64_t *ids, int kind, int unique, int verbose, spxh_ogs_t *out) {
  if (!out) return err(SPX_ERR_ARG, "spxh_ogs_setup: null out");
  if (n > 0 && !ids) return err(SPX_ERR_ARG, "spxh_ogs_setup: null ids");
  if (!state().open) return err(SPX_ERR_NODEVICE, "spxh_ogs_setup: call spxh_platform_open first");
  if (kind < 0 || kind > 2) return err(SPX_ERR_ARG, "spxh_ogs_setup: kind must be 0 (Unsigned), 1 (Signed) or 2 (Halo)");
  *out = nullptr;
  return guarded([&] {
    auto h = std::make_unique<spxh_ogs_s>();
    memory<hlong> m(n);
    if (n) std::copy(ids, ids + n, m.ptr());
    h->ogs.Setup(n, m, comm_t::world(), (ogs::Kind)kind, ogs::Auto, unique != 0, verbose != 0, state().platform);
    if (n) std::copy(m.ptr(), m.ptr() + n, ids);
    *out = h.release();
  });
}

int spxh_ogs_setup_hex(int32_t nx, int32_t ny, int32_t nz, int p, int kind, int unique, int64_t *ids,
                       spxh_ogs_t *out) {
  if (!out) return err(SPX_ERR_ARG, "spxh_ogs_setup_hex: null out");
  if (!state().open) return err(SPX_ERR_NODEVICE, "spxh_ogs_setup_hex: call spxh_platform_open first");
  if (kind < 0 || kind > 2) return err(SPX_ERR_ARG, "spxh_ogs_setup_hex: kind must be 0, 1 or 2");
  *out = nullptr;
  return guarded([&] {
    auto h = std::make_unique<spxh_ogs_s>();
    h->ogs.SetupHexBox(nx, ny, nz, p, comm_t::world(), (ogs::Kind)kind, unique != 0, state().platform,
                       reinterpret_cast<hlong *>(ids));
    *out = h.release();
  });
}

int spxh_ogs_setup_box(int type, int32_t nx, int32_t ny, int32_t nz, int p, int kind, int unique, int64_t *ids,
                       spxh_ogs_t *out) {
  if (!out) return err(SPX_ERR_ARG, "spxh_ogs_setup_box: null out");
  if (!state().open) return err(SPX_ERR_NODEVICE, "spxh_ogs_setup_box: call spxh_platform_open first");
  if (kind < 0 || kind > 2) return err(SPX_ERR_ARG, "spxh_ogs_setup_box: kind must be 0, 1 or 2");
  *out = nullptr;
  return guarded([&] {
    const int Np = mesh_t::NodesPerElement(type, p);
    LIBP_ABORT("spxh_ogs_setup_box: unknown element type or degree", Np <= 0 || p < 1);
    LIBP_ABORT("spxh_ogs_setup_box: box sizes must be positive", nx < 1 || ny < 1 || nz < 1);
    const long long n = mesh_t::BoxElements(type, nx, ny, nz) * Np;
    LIBP_ABORT("spxh_ogs_setup_box: " << n << " local nodes exceed the int32 range", n >= (long long)INT32_MAX);
    memory<hlong> lids((size_t)n);
    if (type != mesh_t::HEXAHEDRA) mesh_t::BoxGlobalIds(type, nx, ny, nz, p, lids.ptr());
    auto h = std::make_unique<spxh_ogs_s>();
    mesh_t::BoxGatherScatterSetup(h->ogs, type, nx, ny, nz, p, lids, comm_t::world(), state().platform,
                                  (ogs::Kind)kind, unique != 0);
    if (ids) std::copy(lids.ptr(), lids.ptr() + n, ids);
    *out = h.release();
  });
}

int spxh_ogs_setup_batch(const spxh_ogs_t *parts, int32_t nparts, spxh_ogs_t *out) {
  if (!out || !parts || nparts < 1) return err(SPX_ERR_ARG, "spxh_ogs_setup_batch: bad arguments");
  if (!state().open) return err(SPX_ERR_NODEVICE, "spxh_ogs_setup_batch: call spxh_platform_open first");
  *out = nullptr;
  return guarded([&] {
    std::vector<const ogs::ogsBase_t *> v;
    for (int32_t i = 0; i < nparts; ++i) {
      LIBP_ABORT("spxh_ogs_setup_batch: null part " << i, !parts[i]);
      v.push_back(&parts[i]->ogs);
    }
    auto h = std::make_unique<spxh_ogs_s>();
    h->ogs.SetupBatch(v, state().platform);
    *out = h.release();
  });
}

int spxh_ogs_destroy(spxh_ogs_t h) {
  delete h;
  return SPX_OK;
}

int spxh_ogs_info(spxh_ogs_t h, spxh_ogs_info_t *info) {
  if (!h || !info) return err(SPX_ERR_ARG, "spxh_ogs_info: null");
  return guarded([&] {
    const ogs::ogs_t &o = h->ogs;
    auto op = o.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    info->N = o.N;
    info->Ngather = o.Ngather;
    info->NlocalT = o.NlocalT;
    info->NlocalP = o.NlocalP;
    info->nnzT = op->nnzT;
    info->nnzN = op->nnzN;
    info->NblocksT = op->NrowBlocksT;
    info->NblocksN = op->NrowBlocksN;
    info->block_nnz = op->blockNnz;
    info->gs_block_nnz = op->gsBlockNnz;
    info->Ngs_blocks = op->NgsBlocks;
    info->gather_defined = o.gather_defined ? 1 : 0;
    info->kind = (int)o.kind;
    info->unique = o.unique ? 1 : 0;
    info->NgatherGlobal = o.NgatherGlobal;
    info->setup_seconds = o.setupSeconds();
    info->scatter_by_rows = op->scatterByRows ? 1 : 0;
    info->mixed_tiles[0] = op->gsMixed[0].own ? 1 : 0;
    info->mixed_tiles[1] = op->gsMixed[1].own ? 1 : 0;
    info->gs_nontemporal = op->gsSrb.Device().nontemporal;
  });
}

int spxh_ogs_copy_csr(spxh_ogs_t h, int which, int32_t *rowStarts, int32_t *colIds, int32_t *blockStarts) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_copy_csr: null");
  return guarded([&] {
    auto op = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    LIBP_ABORT("which must be 0 (T) or 1 (N)", which != 0 && which != 1);
    const bool N = which == 1;
    const dlong rows = N ? op->NrowsN : op->NrowsT;
    const dlong nnz = N ? op->nnzN : op->nnzT;
    const dlong nb = N ? op->NrowBlocksN : op->NrowBlocksT;
    const memory<dlong> &rs = N ? op->rowStartsN : op->rowStartsT;
    const memory<dlong> &ci = N ? op->colIdsN : op->colIdsT;
    const memory<dlong> &bs = N ? op->blockRowStartsN : op->blockRowStartsT;
    if (rowStarts) std::copy(rs.ptr(), rs.ptr() + rows + 1, rowStarts);
    if (colIds && nnz) std::copy(ci.ptr(), ci.ptr() + nnz, colIds);
    if (blockStarts && bs.isInitialized()) std::copy(bs.ptr(), bs.ptr() + nb + 1, blockStarts);
  });
}

int spxh_ogs_set_block_nnz(spxh_ogs_t h, int32_t nnz) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_set_block_nnz: null");
  return guarded([&] {
    auto op = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    op->SetBlockNnz(nnz);
  });
}

int spxh_ogs_set_gs_block_nnz(spxh_ogs_t h, int32_t nnz) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_set_gs_block_nnz: null");
  return guarded([&] {
    auto op = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    op->SetGsBlockNnz(nnz);
  });
}

int spxh_ogs_set_scatter_path(spxh_ogs_t h, int by_rows) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_set_scatter_path: null");
  return guarded([&] {
    auto op = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    op->scatterByRows = by_rows != 0;
  });
}

int spxh_ogs_set_gs_format(spxh_ogs_t h, int sorted, int xcd_remap) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_set_gs_format: null");
  if (xcd_remap > 1024 || (xcd_remap > 1 && (xcd_remap & (xcd_remap - 1))))
    return err(SPX_ERR_ARG, "spxh_ogs_set_gs_format: xcd_remap must be negative (library default), 0, 1 or a power of two up to 1024");
  return guarded([&] {
    auto op = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    op->gsSorted = sorted != 0;
    op->gsSrb.xcdRemap = xcd_remap;
    op->rowSrb.xcdRemap = xcd_remap;
  });
}

int spxh_ogs_set_gs_nontemporal(spxh_ogs_t h, int mode) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_set_gs_nontemporal: null");
  if (mode < -1 || mode > 1) return err(SPX_ERR_ARG, "spxh_ogs_set_gs_nontemporal: mode must be -1, 0 or 1");
  return guarded([&] {
    auto op = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    op->gsSrb.ntHint = mode;
  });
}

int spxh_ogs_set_gs_row_keys(spxh_ogs_t h, const int64_t *keys) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_set_gs_row_keys: null");
  return guarded([&] {
    auto op = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    memory<dlong> order;
    if (keys) {
      const dlong n = op->NrowsT;
      std::vector<std::pair<int64_t, dlong>> kr(n);
      for (dlong r = 0; r < n; ++r) kr[r] = {keys[r], r};
      std::sort(kr.begin(), kr.end());
      order.malloc(n);
      for (dlong r = 0; r < n; ++r) order[r] = kr[r].second;
    }
    op->SetGsRowOrder(order);
  });
}

namespace {
// The n rows sorted by key (stable) and a tile cut wherever the key changes.
void keys_to_tiles(const dlong n, const int64_t *keys, memory<dlong> &order, memory<uint8_t> &cuts) {
  const int64_t kmin = *std::min_element(keys, keys + n), kmax = *std::max_element(keys, keys + n);
  if (kmax - kmin < 4 * (int64_t)n + 1024) {
    // keys in a dense range (tile numbers): stable counting sort
    std::vector<int32_t> k(n);
    for (dlong r = 0; r < n; ++r) k[r] = (int32_t)(keys[r] - kmin);
    order = ogs::CountingOrder(n, k, kmax - kmin + 1);
  } else {
    std::vector<std::pair<int64_t, dlong>> kr(n);
    for (dlong r = 0; r < n; ++r) kr[r] = {keys[r], r};
    std::sort(kr.begin(), kr.end());
    order.malloc(n);
    for (dlong i = 0; i < n; ++i) order[i] = kr[i].second;
  }
  cuts.malloc(n);
  for (dlong i = 0; i < n; ++i) cuts[i] = (i == 0) || keys[order[i]] != keys[order[i - 1]];
}
}  // namespace

int spxh_ogs_set_gs_tiles(spxh_ogs_t h, const int64_t *keys) {
  if (!h || !keys) return err(SPX_ERR_ARG, "spxh_ogs_set_gs_tiles: null");
  return guarded([&] {
    auto op = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    const dlong n = op->NrowsT;
    if (n == 0) {
      op->SetGsRowOrder(memory<dlong>());
      return;
    }
    memory<dlong> order;
    memory<uint8_t> cuts;
    keys_to_tiles(n, keys, order, cuts);
    op->SetGsRowOrder(order, false, cuts);
  });
}

int spxh_ogs_set_gs_mixed_tiles(spxh_ogs_t h, int storage, const int64_t *keys) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_set_gs_mixed_tiles: null");
  if (storage != SPX_STORE_F32 && storage != SPX_STORE_BF16)
    return err(SPX_ERR_ARG, "spxh_ogs_set_gs_mixed_tiles: storage must be SPX_STORE_F32 or SPX_STORE_BF16");
  return guarded([&] {
    auto op = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    const dlong n = op->NrowsT;
    memory<dlong> order;
    memory<uint8_t> cuts;
    if (keys && n > 0) keys_to_tiles(n, keys, order, cuts);
    op->SetGsMixedRowOrder(storage, order, cuts, false);
  });
}

int spxh_ogs_autotune(spxh_ogs_t h, int type, int storage, int k, int reps, spxh_gs_tuning_t *best) {
  if (!h || !best) return err(SPX_ERR_ARG, "spxh_ogs_autotune: null");
  if (storage < -1 || storage > SPX_STORE_BF16) return err(SPX_ERR_ARG, "spxh_ogs_autotune: storage in -1..2");
  if (storage < 0 && (type < 0 || type > 3)) return err(SPX_ERR_ARG, "spxh_ogs_autotune: unknown type");
  return guarded([&] {
    auto op = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    std::vector<ogs::GsTuning> log;
    const ogs::GsTuning b = op->Autotune(type, storage, k, reps, &log);
    best->sorted = b.sorted;
    best->block_nnz = b.blockNnz;
    best->tile_lanes = b.tileLanes;
    best->xcd_remap = b.xcdRemap;
    best->ms = b.ms;
    best->candidates = (int32_t)log.size();
  });
}

int spxh_cg_batch_autotune(int32_t nseg, const int64_t *segStarts, int storage, void *p, void *Ap, void *x, void *r,
                           double *rdotr, int reps, spx_cg_plan_t *plan, spxh_cg_tuning_t *best) {
  if (nseg < 1 || !segStarts || !plan || !best) return err(SPX_ERR_ARG, "spxh_cg_batch_autotune: bad arguments");
  *plan = nullptr;
  return guarded([&] {
    std::vector<int64_t> st(segStarts, segStarts + nseg + 1);
    std::vector<ogs::CgTuning> log;
    const ogs::CgTuning b = ogs::AutotuneCgBatch(st, storage, p, Ap, x, r, rdotr, reps, plan, &log);
    best->target_blocks = b.targetBlocks;
    best->unroll = b.unroll;
    best->nontemporal = b.nontemporal;
    best->ms = b.ms;
    best->candidates = (int32_t)log.size();
  });
}

int spxh_ogs_set_gather_format(spxh_ogs_t h, int gather_sorted, int scatter_sorted) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_set_gather_format: null");
  return guarded([&] {
    auto op = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !op);
    if (gather_sorted >= 0) op->gatherSorted = gather_sorted != 0;
    if (scatter_sorted >= 0) op->scatterSorted = scatter_sorted != 0;
  });
}

int spxh_ogs_global_to_local(spxh_ogs_t h, int32_t *g2l) {
  if (!h || !g2l) return err(SPX_ERR_ARG, "spxh_ogs_global_to_local: null");
  return guarded([&] {
    memory<dlong> m(h->ogs.N);
    h->ogs.SetupGlobalToLocalMapping(m);
    std::copy(m.ptr(), m.ptr() + h->ogs.N, g2l);
  });
}

static int check_enums(int type, int op, int trans) {
  if (type < 0 || type > 3) return err(SPX_ERR_ARG, "type must be 0..3 (Float, Double, Int32, Int64)");
  if (op < 0 || op > 3) return err(SPX_ERR_ARG, "op must be 0..3 (Add, Mul, Max, Min)");
  if (trans < 0 || trans > 2) return err(SPX_ERR_ARG, "trans must be 0..2 (Sym, NoTrans, Trans)");
  return SPX_OK;
}

int spxh_ogs_gather_scatter(spxh_ogs_t h, int type, int op, int trans, int k, void *v) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_gather_scatter: null handle");
  if (int rc = check_enums(type, op, trans)) return rc;
  if (k < 1) return err(SPX_ERR_ARG, "k must be >= 1");
  return guarded([&] {
    auto o = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !o);
    o->GatherScatterRaw((ogs::Type)type, (ogs::Op)op, (ogs::Transpose)trans, k, v);
  });
}

int spxh_ogs_gather_scatter_mixed(spxh_ogs_t h, int storage, int k, void *v) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_gather_scatter_mixed: null handle");
  if (k < 1) return err(SPX_ERR_ARG, "k must be >= 1");
  return guarded([&] {
    auto o = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !o);
    o->GatherScatterMixedRaw(storage, k, v);
  });
}

int spxh_ogs_gather(spxh_ogs_t h, int type, int op, int trans, int k, const void *v, void *gv) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_gather: null handle");
  if (int rc = check_enums(type, op, trans)) return rc;
  if (k < 1) return err(SPX_ERR_ARG, "k must be >= 1");
  return guarded([&] {
    LIBP_ABORT("Gather operation not well-defined.", !h->ogs.gather_defined);
    auto o = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !o);
    o->GatherRaw((ogs::Type)type, (ogs::Op)op, (ogs::Transpose)trans, k, v, gv);
  });
}

int spxh_ogs_scatter(spxh_ogs_t h, int type, int trans, int k, const void *gv, void *v) {
  if (!h) return err(SPX_ERR_ARG, "spxh_ogs_scatter: null handle");
  if (int rc = check_enums(type, 0, trans)) return rc;
  if (k < 1) return err(SPX_ERR_ARG, "k must be >= 1");
  return guarded([&] {
    LIBP_ABORT("Gather operation not well-defined.", !h->ogs.gather_defined);
    auto o = h->ogs.localOperator();
    LIBP_ABORT("ogs not set up", !o);
    o->ScatterRaw((ogs::Type)type, (ogs::Transpose)trans, k, gv, v);
  });
}

struct spxh_csr_s {
  ogs::LocalCSR csr;
};

int spxh_csr_build(int32_t n, int64_t *ids, int kind, int unique, spxh_csr_t *out) {
  if (!out) return err(SPX_ERR_ARG, "spxh_csr_build: null out");
  if (n > 0 && !ids) return err(SPX_ERR_ARG, "spxh_csr_build: null ids");
  if (kind < 0 || kind > 2) return err(SPX_ERR_ARG, "spxh_csr_build: bad kind");
  *out = nullptr;
  return guarded([&] {
    auto h = std::make_unique<spxh_csr_s>();
    memory<hlong> m(n);
    if (n) std::copy(ids, ids + n, m.ptr());
    h->csr = ogs::BuildLocalCSR(n, m, (ogs::Kind)kind, unique != 0, 0);
    if (n) std::copy(m.ptr(), m.ptr() + n, ids);
    *out = h.release();
  });
}

int spxh_csr_build_hex(int32_t nx, int32_t ny, int32_t nz, int p, int kind, int unique, int64_t *ids,
                       spxh_csr_t *out) {
  if (!out) return err(SPX_ERR_ARG, "spxh_csr_build_hex: null out");
  if (kind < 0 || kind > 2) return err(SPX_ERR_ARG, "spxh_csr_build_hex: bad kind");
  *out = nullptr;
  return guarded([&] {
    auto h = std::make_unique<spxh_csr_s>();
    h->csr = ogs::BuildHexBoxCSR(nx, ny, nz, p, (ogs::Kind)kind, unique != 0, 0, reinterpret_cast<hlong *>(ids));
    *out = h.release();
  });
}

int spxh_csr_sizes(spxh_csr_t h, int32_t *sizes) {
  if (!h || !sizes) return err(SPX_ERR_ARG, "spxh_csr_sizes: null");
  const ogs::LocalCSR &c = h->csr;
  sizes[0] = c.NrowsT;
  sizes[1] = c.NrowsN;
  sizes[2] = c.nnzT;
  sizes[3] = c.nnzN;
  sizes[4] = c.gather_defined ? 1 : 0;
  return SPX_OK;
}

int spxh_csr_copy(spxh_csr_t h, int which, int32_t *rowStarts, int32_t *colIds) {
  if (!h) return err(SPX_ERR_ARG, "spxh_csr_copy: null");
  if (which != 0 && which != 1) return err(SPX_ERR_ARG, "which must be 0 (T) or 1 (N)");
  const ogs::LocalCSR &c = h->csr;
  const dlong rows = which ? c.NrowsN : c.NrowsT;
  const dlong nnz = which ? c.nnzN : c.nnzT;
  const memory<dlong> &rs = which ? c.rowStartsN : c.rowStartsT;
  const memory<dlong> &ci = which ? c.colIdsN : c.colIdsT;
  if (rowStarts) std::copy(rs.ptr(), rs.ptr() + rows + 1, rowStarts);
  if (colIds && nnz) std::copy(ci.ptr(), ci.ptr() + nnz, colIds);
  return SPX_OK;
}

int spxh_csr_destroy(spxh_csr_t h) {
  delete h;
  return SPX_OK;
}

struct spxh_srb_s {
  ogs::SortedRunBlocks srb;
};

int spxh_srb_build(int32_t nrows, const int32_t *rowStarts, const int32_t *colIds, int32_t nblocks,
                   const int32_t *blockStarts, int32_t cap, spxh_srb_t *out) {
  if (!out || !rowStarts || !blockStarts || nrows < 0 || nblocks < 0) return err(SPX_ERR_ARG, "spxh_srb_build: bad arguments");
  *out = nullptr;
  return guarded([&] {
    LIBP_ABORT("blockStarts must run from row 0 to row nrows",
               blockStarts[0] != 0 || blockStarts[nblocks] != nrows);
    for (int32_t b = 0; b < nblocks; ++b)
      LIBP_ABORT("blockStarts must be nondecreasing", blockStarts[b + 1] < blockStarts[b]);
    auto h = std::make_unique<spxh_srb_s>();
    memory<dlong> rs(nrows + 1), bs(nblocks + 1);
    std::copy(rowStarts, rowStarts + nrows + 1, rs.ptr());
    std::copy(blockStarts, blockStarts + nblocks + 1, bs.ptr());
    const dlong nnz = rs[nrows];
    LIBP_ABORT("colIds missing", nnz > 0 && !colIds);
    memory<dlong> ci(nnz > 0 ? nnz : 1);
    if (nnz > 0) std::copy(colIds, colIds + nnz, ci.ptr());
    h->srb = ogs::BuildSortedRunBlocks(nrows, rs, ci, nblocks, bs, cap);
    *out = h.release();
  });
}

int spxh_srb_build_signed(int32_t nrows, const int32_t *rowStarts, const int32_t *colIds, int32_t nblocks,
                          const int32_t *blockStarts, int32_t cap, int32_t ncols, int32_t nrowsN,
                          const int32_t *rowStartsN, const int32_t *colIdsN, spxh_srb_t *out) {
  if (!out || !rowStarts || !blockStarts || !rowStartsN || nrows < 0 || nblocks < 0 || ncols < 0 || nrowsN < 0)
    return err(SPX_ERR_ARG, "spxh_srb_build_signed: bad arguments");
  *out = nullptr;
  return guarded([&] {
    LIBP_ABORT("blockStarts must run from row 0 to row nrows",
               blockStarts[0] != 0 || blockStarts[nblocks] != nrows);
    for (int32_t b = 0; b < nblocks; ++b)
      LIBP_ABORT("blockStarts must be nondecreasing", blockStarts[b + 1] < blockStarts[b]);
    auto h = std::make_unique<spxh_srb_s>();
    memory<dlong> rs(nrows + 1), bs(nblocks + 1), rsN(nrowsN + 1);
    std::copy(rowStarts, rowStarts + nrows + 1, rs.ptr());
    std::copy(blockStarts, blockStarts + nblocks + 1, bs.ptr());
    std::copy(rowStartsN, rowStartsN + nrowsN + 1, rsN.ptr());
    const dlong nnz = rs[nrows], nnzN = rsN[nrowsN];
    LIBP_ABORT("colIds missing", (nnz > 0 && !colIds) || (nnzN > 0 && !colIdsN));
    for (int32_t i = 0; i < nnz; ++i) LIBP_ABORT("colIds outside [0, ncols)", colIds[i] < 0 || colIds[i] >= ncols);
    memory<dlong> ci(nnz > 0 ? nnz : 1), ciN(nnzN > 0 ? nnzN : 1);
    if (nnz > 0) std::copy(colIds, colIds + nnz, ci.ptr());
    if (nnzN > 0) std::copy(colIdsN, colIdsN + nnzN, ciN.ptr());
    memory<uint8_t> flag = ogs::SignedMemberFlags(ncols, nrows, rs, ci, nrowsN, rsN, ciN);
    LIBP_ABORT("the N-CSR's rows are not the flagged members of the same T-CSR rows", flag.length() == 0);
    h->srb = ogs::BuildSortedRunBlocks(nrows, rs, ci, nblocks, bs, cap, flag.ptr());
    *out = h.release();
  });
}

int spxh_srb_rows_ascending(spxh_srb_t h, int32_t *ascending) {
  if (!h || !ascending) return err(SPX_ERR_ARG, "spxh_srb_rows_ascending: null");
  *ascending = h->srb.rowsAscending ? 1 : 0;
  return SPX_OK;
}

int spxh_srb_copy_members(spxh_srb_t h, uint64_t *memberN) {
  if (!h || !memberN) return err(SPX_ERR_ARG, "spxh_srb_copy_members: null");
  const ogs::SortedRunBlocks &s = h->srb;
  if (!s.memberN.length()) return err(SPX_ERR_ARG, "spxh_srb_copy_members: layout built without member flags");
  if (s.nchunks) std::copy(s.memberN.ptr(), s.memberN.ptr() + s.nchunks, memberN);
  return SPX_OK;
}

int spxh_srb_nontemporal(spxh_srb_t h, int32_t *mask) {
  if (!h || !mask) return err(SPX_ERR_ARG, "spxh_srb_nontemporal: null");
  *mask = h->srb.ntMask;
  return SPX_OK;
}

int spxh_srb_sizes(spxh_srb_t h, int64_t *sizes) {
  if (!h || !sizes) return err(SPX_ERR_ARG, "spxh_srb_sizes: null");
  const ogs::SortedRunBlocks &s = h->srb;
  sizes[0] = s.nblocks;
  sizes[1] = s.nnz;
  sizes[2] = s.nchunks;
  sizes[3] = s.nruns;
  sizes[4] = s.nlinks;
  sizes[5] = s.cpt;
  return SPX_OK;
}

int spxh_srb_copy(spxh_srb_t h, spx_srb_chunk_t *chunks, int32_t *runIds, uint16_t *links, int32_t *entryStarts) {
  if (!h) return err(SPX_ERR_ARG, "spxh_srb_copy: null");
  const ogs::SortedRunBlocks &s = h->srb;
  auto cp = [](const auto &m, size_t n, auto *dst) {
    if (dst && n) std::copy(m.ptr(), m.ptr() + n, dst);
  };
  cp(s.chunks, (size_t)s.nchunks, chunks);
  cp(s.runIds, (size_t)s.nruns, runIds);
  cp(s.links, (size_t)s.nlinks + 1, links);
  cp(s.entryStarts, (size_t)s.nblocks + 1, entryStarts);
  return SPX_OK;
}

int spxh_srb_destroy(spxh_srb_t h) {
  delete h;
  return SPX_OK;
}

int spxh_block_rows(int32_t nrows, const int32_t *rowStarts, int32_t cap, int32_t *nblocks, int32_t *blockStarts) {
  if (!rowStarts || !nblocks || !blockStarts) return err(SPX_ERR_ARG, "spxh_block_rows: null");
  return guarded([&] {
    memory<dlong> rs(nrows + 1);
    std::copy(rowStarts, rowStarts + nrows + 1, rs.ptr());
    memory<dlong> bs;
    dlong nb = 0;
    ogs::blockRowsGreedy(nrows, rs, cap, nb, bs);
    *nblocks = nb;
    std::copy(bs.ptr(), bs.ptr() + nb + 1, blockStarts);
  });
}

int spxh_hex_gs_tiles(int nx, int ny, int nz, int p, int rows_per_line, int cap, int32_t nrows,
                      const int64_t *row_grid, int32_t *order, uint8_t *cuts) {
  if (nrows < 0 || (nrows > 0 && (!row_grid || !order || !cuts)))
    return err(SPX_ERR_ARG, "spxh_hex_gs_tiles: bad arguments");
  return guarded([&] {
    memory<hlong> grid(nrows > 0 ? nrows : 1);
    for (int32_t r = 0; r < nrows; ++r) grid[r] = row_grid[r];
    memory<dlong> o;
    memory<uint8_t> c;
    ogs::HexLineTiles(nrows, grid, nx, ny, nz, p, rows_per_line, cap, o, c);
    if (nrows > 0) {
      std::copy(o.ptr(), o.ptr() + nrows, order);
      std::copy(c.ptr(), c.ptr() + nrows, cuts);
    }
  });
}

int spxh_mesh_hex_ids(int nx, int ny, int nz, int p, int64_t *ids) {
  if (!ids) return err(SPX_ERR_ARG, "spxh_mesh_hex_ids: null ids");
  if (nx < 1 || ny < 1 || nz < 1 || p < 1 || p > 15) return err(SPX_ERR_ARG, "spxh_mesh_hex_ids: bad size");
  return guarded([&] { mesh_t::HexBoxGlobalIds(nx, ny, nz, p, (hlong *)ids); });
}

int spxh_mesh_box_shape(int type, int nx, int ny, int nz, int p, int64_t *nelements, int *np) {
  if (!nelements || !np) return err(SPX_ERR_ARG, "spxh_mesh_box_shape: null output");
  if (nx < 1 || ny < 1 || nz < 1 || p < 1 || p > 15) return err(SPX_ERR_ARG, "spxh_mesh_box_shape: bad size");
  if (mesh_t::NodesPerElement(type, p) == 0)
    return err(SPX_ERR_ARG, "spxh_mesh_box_shape: type must be 3 (Tri), 4 (Quad), 6 (Tet) or 12 (Hex)");
  *nelements = mesh_t::BoxElements(type, nx, ny, nz);
  *np = mesh_t::NodesPerElement(type, p);
  return SPX_OK;
}

int spxh_mesh_box_ids(int type, int nx, int ny, int nz, int p, int64_t *ids) {
  if (!ids) return err(SPX_ERR_ARG, "spxh_mesh_box_ids: null ids");
  if (nx < 1 || ny < 1 || nz < 1 || p < 1 || p > 15) return err(SPX_ERR_ARG, "spxh_mesh_box_ids: bad size");
  if (mesh_t::NodesPerElement(type, p) == 0)
    return err(SPX_ERR_ARG, "spxh_mesh_box_ids: type must be 3 (Tri), 4 (Quad), 6 (Tet) or 12 (Hex)");
  return guarded([&] { mesh_t::BoxGlobalIds(type, nx, ny, nz, p, (hlong *)ids); });
}

}  // extern "C"
