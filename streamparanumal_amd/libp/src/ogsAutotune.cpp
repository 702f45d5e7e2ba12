// Autotuner for the Sym gather+scatter and the batched CG update: every
// candidate launch shape is timed on the device with events on the operator's
// own data layout, the fastest is kept.  Every candidate computes the same
// bits (the row fold order never changes), so tuning changes only the time.
#include <algorithm>
#include <limits>
#include <vector>

#include "ogs.hpp"

namespace libp {
namespace ogs {

namespace {

// mean milliseconds of `reps` calls after two untimed ones
template <typename F>
double timeCalls(F &&call, int reps) {
  call();
  call();
  spx_event_t e0, e1;
  SPX_CHECK(spx_event_create(&e0));
  SPX_CHECK(spx_event_create(&e1));
  SPX_CHECK(spx_event_record(e0));
  for (int i = 0; i < reps; ++i) call();
  SPX_CHECK(spx_event_record(e1));
  SPX_CHECK(spx_device_finish());
  float ms = 0.f;
  SPX_CHECK(spx_event_elapsed_ms(e0, e1, &ms));
  SPX_CHECK(spx_event_destroy(e0));
  SPX_CHECK(spx_event_destroy(e1));
  return (double)ms / reps;
}

size_t elemBytes(int type, int storage) {
  if (storage >= 0) return storage == SPX_STORE_F64 ? 8 : storage == SPX_STORE_F32 ? 4 : 2;
  return (type == Double || type == Int64) ? 8 : 4;
}

}  // namespace

GsTuning ogsOperator_t::Autotune(int type, int storage, int k, int reps, std::vector<GsTuning> *log) {
  LIBP_ABORT("Autotune: k must be >= 1", k < 1);
  LIBP_ABORT("Autotune: reps must be >= 1", reps < 1);
  LIBP_ABORT("Autotune: unknown storage", storage > SPX_STORE_BF16);
  // scratch vector (values do not change the work: no data-dependent branches)
  deviceMemory<char> scratch = platform.malloc<char>((size_t)Ncols * k * elemBytes(type, storage) + 16);
  SPX_CHECK(spx_fill_bytes(scratch.ptr(), scratch.length(), 0));
  auto run = [&] {
    if (storage >= 0) GatherScatterMixedRaw(storage, k, scratch.ptr());
    else GatherScatterRaw((Type)type, Add, Sym, k, scratch.ptr());
  };
  // The current shape, timed first on its own layouts, which are kept: a
  // rebuilt layout lands elsewhere in HBM, and where the arrays sit moves
  // these kernels by a few percent (64^3 p=7: the same shape read 454 us as
  // set up and 464-470 us rebuilt), as much as the shapes differ.  So a
  // candidate replaces the current shape only when it is clearly faster.
  const dlong keepNnz = gsBlockNnz, keepBlocks = NgsBlocks;
  const memory<dlong> keepStarts = gsBlockRowStarts;
  const deviceMemory<dlong> keepO = o_gsBlockRowStarts;
  const SortedRunBlocks keepGs = gsSrb, keepRow = rowSrb;
  const MixedTiles keepMixed[2] = {gsMixed[0], gsMixed[1]};
  GsTuning current{gsSorted ? 1 : 0, (int)gsBlockNnz, gsSrb.tileLanes, gsSrb.xcdRemap, 0.0};
  current.ms = timeCalls(run, reps);
  if (log) log->push_back(current);
  std::vector<GsTuning> cand;
  for (int cap : {256, 512, 1024})
    for (int xcd : {1, 0}) cand.push_back({1, cap, 64, xcd, 0.0});
  for (int cap : {1024, 2048})
    for (int xcd : {1, 0}) cand.push_back({1, cap, 256, xcd, 0.0});
  if (storage < 0)
    for (int cap : {2048, 4096}) cand.push_back({0, cap, 256, 0, 0.0});
  std::stable_sort(cand.begin(), cand.end(), [](const GsTuning &a, const GsTuning &b) { return a.blockNnz < b.blockNnz; });
  GsTuning best{1, 0, 0, 0, std::numeric_limits<double>::infinity()};
  for (GsTuning &c : cand) {
    if (c.blockNnz != gsBlockNnz) SetGsBlockNnz(c.blockNnz);
    gsSorted = c.sorted != 0;
    gsSrb.tileLanes = c.tileLanes;
    gsSrb.xcdRemap = c.xcdRemap;
    c.ms = timeCalls(run, reps);
    if (log) log->push_back(c);
    if (c.ms < best.ms) best = c;
  }
  SPX_CHECK(spx_check_kernel_errors());
  if (!(best.ms < kAutotuneMargin * current.ms)) best = current;
  if (best.blockNnz == keepNnz) {
    // the current capacity: back to its original layouts
    gsBlockNnz = keepNnz;
    NgsBlocks = keepBlocks;
    gsBlockRowStarts = keepStarts;
    o_gsBlockRowStarts = keepO;
    gsSrb = keepGs;
    rowSrb = keepRow;
    gsMixed[0] = keepMixed[0];
    gsMixed[1] = keepMixed[1];
  } else if (best.blockNnz != gsBlockNnz) {
    SetGsBlockNnz(best.blockNnz);
  }
  gsSorted = best.sorted != 0;
  gsSrb.tileLanes = best.tileLanes;
  gsSrb.xcdRemap = best.xcdRemap;
  return best;
}

CgTuning AutotuneCgBatch(const std::vector<int64_t> &segStarts, int storage, void *p, void *Ap, void *x, void *r,
                         void *rdotr, int reps, spx_cg_plan_t *plan, std::vector<CgTuning> *log) {
  LIBP_ABORT("AutotuneCgBatch: need segments", segStarts.size() < 2);
  LIBP_ABORT("AutotuneCgBatch: reps must be >= 1", reps < 1);
  CgTuning best{0, 0, 0, std::numeric_limits<double>::infinity()};
  spx_tuning_t keep;
  SPX_CHECK(spx_get_tuning("cg_batch", &keep));
  const int nseg = (int)segStarts.size() - 1;
  for (int blocks : {2048, 8192, 32768, 131072, INT32_MAX}) {  // INT32_MAX: one workgroup per tile
    spx_cg_plan_t pl = nullptr;
    SPX_CHECK(spx_cg_plan_create(nseg, segStarts.data(), blocks, &pl));
    for (int unroll : {1, 2, 4})
      for (int nt : {1, 0}) {
        spx_tuning_t t{};
        t.xcd_remap = -1;
        t.variant = -1;
        t.unroll = unroll;
        t.nontemporal = nt;
        SPX_CHECK(spx_set_tuning("cg_batch", &t));
        // alpha 0: the vectors keep their values while timing
        const double ms = timeCalls([&] { SPX_CHECK(spx_cg_batch(storage, pl, 0.0, p, Ap, x, r, (double *)rdotr)); }, reps);
        CgTuning c{blocks, unroll, nt, ms};
        if (log) log->push_back(c);
        if (ms < best.ms) best = c;
      }
    SPX_CHECK(spx_cg_plan_destroy(pl));
  }
  spx_tuning_t t{};
  t.xcd_remap = -1;
  t.variant = -1;
  t.unroll = best.unroll;
  t.nontemporal = best.nontemporal;
  SPX_CHECK(spx_set_tuning("cg_batch", &t));
  SPX_CHECK(spx_cg_plan_create(nseg, segStarts.data(), best.targetBlocks, plan));
  return best;
}

}  // namespace ogs
}  // namespace libp
