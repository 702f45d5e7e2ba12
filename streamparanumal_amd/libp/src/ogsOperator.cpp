// ogsOperator_t on the device (libs/ogs/ogsOperator.cpp:176-204, 295-322,
// 442-481, 597-665) over the libspx ogs kernels.
#include <algorithm>

#include "ogs.hpp"

namespace libp {
namespace ogs {

void blockRowsGreedy(const dlong Nrows, const memory<dlong> rowStarts, const int cap, dlong &Nblocks,
                     memory<dlong> &blockStarts) {
  Nblocks = 0;
  std::vector<dlong> starts;
  starts.push_back(0);
  long long acc = 0;
  for (dlong r = 0; r < Nrows; ++r) {
    const dlong sz = rowStarts[r + 1] - rowStarts[r];
    LIBP_ABORT("Multiplicity of a global node (" << sz << ") exceeds the row block capacity " << cap,
               sz > cap);
    if (acc + sz > cap && acc > 0) {
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
}

void ogsOperator_t::SetGsBlockNnz(int nnz) {
  LIBP_ABORT("row block capacity must be in [1, " << maxBlockNnz << "]", nnz < 1 || nnz > maxBlockNnz);
  gsBlockNnz = nnz;
  blockRowsGreedy(NrowsT, rowStartsT, gsBlockNnz, NgsBlocks, gsBlockRowStarts);
  o_gsBlockRowStarts = platform.malloc<dlong>(NgsBlocks + 1, gsBlockRowStarts);
  const int xcd = gsSrb.xcdRemap, lanes = gsSrb.tileLanes;
  gsSrb = BuildSortedRunBlocks(NrowsT, rowStartsT, colIdsT, NgsBlocks, gsBlockRowStarts, gsBlockNnz,
                               signedSrb ? memberFlagN.ptr() : nullptr);
  gsSrb.xcdRemap = xcd;
  gsSrb.tileLanes = lanes;
  gsSrb.Upload(platform);
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
  const bool srbRows = NgsBlocks && gsSrb.rowsAscending;  // the layout numbers the rows
  if (trans == NoTrans && gatherSorted && srbRows && (signedSrb || kind == Unsigned)) {
    const spx_srb_t d = gsSrb.Device();
    if (kind == Unsigned) SPX_CHECK(spx_ogs_gather_srb((int)type, (int)op, k, &d, v, gv));
    else SPX_CHECK(spx_ogs_gather_srb_signed((int)type, (int)op, k, &d, gsSrb.DeviceMemberN(), NrowsN, v, gv));
  } else if (trans == NoTrans) {
    if (NrowBlocksN)
      SPX_CHECK(spx_ogs_gather((int)type, (int)op, k, NrowBlocksN, o_blockRowStartsN.ptr(), o_rowStartsN.ptr(),
                               o_colIdsN.ptr(), blockNnz, v, gv));
  } else if (trans != NoTrans && gatherSorted && srbRows) {
    const spx_srb_t d = gsSrb.Device();
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
  const bool srbRows = NgsBlocks && gsSrb.rowsAscending;  // the layout numbers the rows
  if (trans == Trans && scatterSorted && srbRows && (signedSrb || kind == Unsigned)) {
    const spx_srb_t d = gsSrb.Device();
    if (kind == Unsigned) SPX_CHECK(spx_ogs_scatter_srb((int)type, k, &d, gv, v));
    else SPX_CHECK(spx_ogs_scatter_srb_signed((int)type, k, &d, gsSrb.DeviceMemberN(), NrowsN, gv, v));
  } else if (trans == Trans) {
    if (NrowBlocksN)
      SPX_CHECK(spx_ogs_scatter((int)type, k, NrowBlocksN, o_blockRowStartsN.ptr(), o_rowStartsN.ptr(),
                                o_colIdsN.ptr(), blockNnz, gv, v));
  } else if (trans != Trans && scatterSorted && srbRows) {
    const spx_srb_t d = gsSrb.Device();
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
    const spx_srb_t d = gsSrb.Device();
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
