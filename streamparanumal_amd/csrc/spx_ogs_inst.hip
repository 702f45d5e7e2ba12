// Launchers of the ogs kernels for one value type, SPX_OGS_T (compiled once
// per type so the kernel instances build in parallel).
#include "spx_ogs_kernels.hpp"

#ifndef SPX_OGS_T
#error "compile with -DSPX_OGS_T=<float|double|int|long long>"
#endif

namespace spx {
namespace ogsk {

using T = SPX_OGS_T;

// Calls F.template operator()<OPC, KF, PER>() for the runtime (op, kf, per).
template <bool WITH_OP, typename F>
static int dispatch(int op, int kf, int per, F &&f) {
  auto by_shape = [&](auto opc) -> int {
    constexpr int OPC = decltype(opc)::value;
    if (kf == 1 && per == 4) return f.template operator()<OPC, 1, 4>();
    if (kf == 1 && per == 8) return f.template operator()<OPC, 1, 8>();
    if (kf == 1 && per == 16) return f.template operator()<OPC, 1, 16>();
    if (kf == 2 && per == 4) return f.template operator()<OPC, 2, 4>();
    if (kf == 2 && per == 8) return f.template operator()<OPC, 2, 8>();
    if (kf == 3 && per == 4) return f.template operator()<OPC, 3, 4>();
    if (kf == 3 && per == 8) return f.template operator()<OPC, 3, 8>();
    if (kf == 4 && per == 4) return f.template operator()<OPC, 4, 4>();
    if (kf == 4 && per == 8) return f.template operator()<OPC, 4, 8>();
    return fail(SPX_ERR_UNSUPPORTED, "ogs: no kernel for fields-per-pass " + std::to_string(kf) +
                                         " with " + std::to_string(per) + " entries per lane");
  };
  if (!WITH_OP) return by_shape(std::integral_constant<int, SPX_ADD>{});
  switch (op) {
    case SPX_ADD: return by_shape(std::integral_constant<int, SPX_ADD>{});
    case SPX_MUL: return by_shape(std::integral_constant<int, SPX_MUL>{});
    case SPX_MAX: return by_shape(std::integral_constant<int, SPX_MAX>{});
    default: return by_shape(std::integral_constant<int, SPX_MIN>{});
  }
}

// one grid row per pass of kf fields (kf divides K)
static dim3 grid_of(const Launch &L) { return dim3(L.nblocks, L.K / L.kf); }

int launch_gs_sym(const Launch &L, int op, const int *bs, const int *st, const int *ids, T *q) {
  const size_t rsb = L.variant == 1 ? (size_t)(L.cap + 1) * sizeof(int)
                    : L.variant == 3 ? (size_t)(L.cap + 1) * sizeof(unsigned short) : 0;
  const size_t lds = lds_bytes(L.cap, L.kf, sizeof(T), 1) + rsb;
  if (lds > 65536) return fail(SPX_ERR_ARG, "spx_ogs_gather_scatter: tile too large for LDS");
  return dispatch<true>(op, L.kf, L.per, [&]<int OPC, int KF, int PER>() {
    if (L.variant == 1)
      hipLaunchKernelGGL((k_gs_sym<T, OPC, KF, PER, 1>), grid_of(L), dim3(kThreads), lds, L.stream, L, bs, st, ids, q);
    else if (L.variant == 2)
      hipLaunchKernelGGL((k_gs_sym<T, OPC, KF, PER, 2>), grid_of(L), dim3(kThreads), lds, L.stream, L, bs, st, ids, q);
    else if (L.variant == 3)
      hipLaunchKernelGGL((k_gs_sym<T, OPC, KF, PER, 3>), grid_of(L), dim3(kThreads), lds, L.stream, L, bs, st, ids, q);
    else
      hipLaunchKernelGGL((k_gs_sym<T, OPC, KF, PER, 0>), grid_of(L), dim3(kThreads), lds, L.stream, L, bs, st, ids, q);
    return SPX_OK;
  });
}

// sorted-run kernels: one-wave tiles (variant 2, up to 1024 positions) or 256
// lanes.  All K fields go in one pass when the lanes' registers (pick_per)
// and the LDS hold them; otherwise half of them per pass (K even), else one
// per pass.  L.cap is the layout's positions per tile (chunks per tile * 64).
// gs: the gather+scatter kernel, whose LDS is srb_gs_lds_bytes (the gather
// and scatter kernels keep one value and one u16 link per position).
template <bool WITH_OP, typename F>
static int srb_shape(const Launch &L, int op, F &&f, bool gs = false) {
  Launch W = L;
  const bool wave = L.variant == 2 && L.cap <= 16 * 64;
  const int lanes = wave ? 64 : kThreads;
  auto lds_of = [&](int kf) {
    return gs ? srb_gs_lds_bytes<T>(W.cap, kf) : lds_bytes(W.cap, 1, sizeof(T), 1) + (size_t)W.cap * sizeof(unsigned short);
  };
  auto fits = [&](int kf) {
    W.kf = kf;
    W.per = pick_per(kf, W.cap * (kThreads / lanes));
    return W.per > 0 && lds_of(kf) <= 65536;
  };
  const int kf0 = L.kf;
  if (!fits(kf0) && !(kf0 > 2 && kf0 % 2 == 0 && fits(kf0 / 2)) && !fits(1)) {
    if (!W.per) return fail(SPX_ERR_ARG, "ogs sorted-run: tile too large");
    return fail(SPX_ERR_ARG, "ogs sorted-run: tile too large for LDS");
  }
  const size_t lds = lds_of(W.kf);
  return dispatch<WITH_OP>(op, W.kf, W.per, [&]<int OPC, int KF, int PER>() {
    return f.template operator()<OPC, KF, PER>(W, wave, lds);
  });
}

// Non-temporal values (L.ntMask, spx_srb_t.nontemporal) for a Sym pass over
// all K fields of nodes of K*sizeof(T) = 2^e bytes with bit e set.
static bool srb_nontemporal(const Launch &W) {
  const int bytes = W.K * (int)sizeof(T);
  if (W.kf != W.K || bytes > 32 || (bytes & (bytes - 1))) return false;
  int e = 0;
  while ((1 << e) < bytes) ++e;
  return (W.ntMask >> e) & 1;
}

// mode (kAll, kFoldN, kStoreN) and memberN: the N-CSR forms, on 256-lane
// tiles only (one instance set per mode).
int launch_gs_srb(const Launch &L, int op, const SrbArgs &A, T *q, int mode, const unsigned long long *memberN) {
  Launch L2 = L;
  if (mode != kAll) L2.variant = 0;
  return srb_shape<true>(
      L2, op,
      [&]<int OPC, int KF, int PER>(const Launch &W, bool wave, size_t lds) {
        if (mode == kAll && !wave && srb_nontemporal(W))
          hipLaunchKernelGGL((k_gs_srb<T, OPC, KF, PER, kThreads, Fold<T, OPC>, kAll, true>), grid_of(W),
                             dim3(kThreads), lds, W.stream, W, SRB_ARGS(A), q, nullptr);
        else if (mode == kFoldN)
          hipLaunchKernelGGL((k_gs_srb<T, OPC, KF, PER, kThreads, Fold<T, OPC>, kFoldN>), grid_of(W),
                             dim3(kThreads), lds, W.stream, W, SRB_ARGS(A), q, memberN);
        else if (mode == kStoreN)
          hipLaunchKernelGGL((k_gs_srb<T, OPC, KF, PER, kThreads, Fold<T, OPC>, kStoreN>), grid_of(W),
                             dim3(kThreads), lds, W.stream, W, SRB_ARGS(A), q, memberN);
        else if (wave)
          hipLaunchKernelGGL((k_gs_srb<T, OPC, KF, PER, 64>), grid_of(W), dim3(64), lds, W.stream, W, SRB_ARGS(A),
                             q, nullptr);
        else
          hipLaunchKernelGGL((k_gs_srb<T, OPC, KF, PER>), grid_of(W), dim3(kThreads), lds, W.stream, W,
                             SRB_ARGS(A), q, nullptr);
        return SPX_OK;
      },
      true);
}

int launch_gather_srb(const Launch &L, int op, const SrbArgs &A, const T *v, T *gv, int mode,
                      const unsigned long long *memberN) {
  Launch L2 = L;
  if (mode != kAll) L2.variant = 0;
  return srb_shape<true>(L2, op, [&]<int OPC, int KF, int PER>(const Launch &W, bool wave, size_t lds) {
    if (mode == kFoldN)
      hipLaunchKernelGGL((k_gather_srb<T, OPC, KF, PER, kThreads, kFoldN>), grid_of(W), dim3(kThreads), lds,
                         W.stream, W, SRB_ARGS(A), v, gv, memberN);
    else if (wave)
      hipLaunchKernelGGL((k_gather_srb<T, OPC, KF, PER, 64>), grid_of(W), dim3(64), lds, W.stream, W, SRB_ARGS(A), v,
                         gv, nullptr);
    else
      hipLaunchKernelGGL((k_gather_srb<T, OPC, KF, PER>), grid_of(W), dim3(kThreads), lds, W.stream, W, SRB_ARGS(A), v,
                         gv, nullptr);
    return SPX_OK;
  });
}

int launch_scatter_srb(const Launch &L, const SrbArgs &A, const T *gv, T *v, int mode,
                       const unsigned long long *memberN) {
  Launch L2 = L;
  if (mode != kAll) L2.variant = 0;
  return srb_shape<false>(L2, SPX_ADD, [&]<int OPC, int KF, int PER>(const Launch &W, bool wave, size_t lds) {
    if (mode == kStoreN)
      hipLaunchKernelGGL((k_scatter_srb<T, KF, PER, kThreads, kStoreN>), grid_of(W), dim3(kThreads), lds, W.stream,
                         W, SRB_ARGS(A), gv, v, memberN);
    else if (wave)
      hipLaunchKernelGGL((k_scatter_srb<T, KF, PER, 64>), grid_of(W), dim3(64), lds, W.stream, W, SRB_ARGS(A), gv, v,
                         nullptr);
    else
      hipLaunchKernelGGL((k_scatter_srb<T, KF, PER>), grid_of(W), dim3(kThreads), lds, W.stream, W, SRB_ARGS(A), gv,
                         v, nullptr);
    return SPX_OK;
  });
}

int launch_gs_gen(const Launch &L, int op, int ngrows, int nsrows, const int *bs, const int *gs, const int *gi,
                  const int *ss, const int *si, T *q) {
  const size_t lds = lds_bytes(L.cap, L.kf, sizeof(T), 2);
  return dispatch<true>(op, L.kf, L.per, [&]<int OPC, int KF, int PER>() {
    hipLaunchKernelGGL((k_gs_gen<T, OPC, KF, PER>), grid_of(L), dim3(kThreads), lds, L.stream, L, ngrows, nsrows,
                       bs, gs, gi, ss, si, q);
    return SPX_OK;
  });
}

int launch_gather(const Launch &L, int op, const int *bs, const int *st, const int *ids, const T *v, T *gv) {
  const size_t lds = lds_bytes(L.cap, L.kf, sizeof(T), 1);
  return dispatch<true>(op, L.kf, L.per, [&]<int OPC, int KF, int PER>() {
    hipLaunchKernelGGL((k_gather<T, OPC, KF, PER>), grid_of(L), dim3(kThreads), lds, L.stream, L, bs, st, ids, v, gv);
    return SPX_OK;
  });
}

int launch_scatter(const Launch &L, const int *bs, const int *st, const int *ids, const T *gv, T *v) {
  const size_t lds = lds_bytes(L.cap, L.kf, sizeof(T), 1);
  return dispatch<false>(SPX_ADD, L.kf, L.per, [&]<int OPC, int KF, int PER>() {
    hipLaunchKernelGGL((k_scatter<T, KF, PER>), grid_of(L), dim3(kThreads), lds, L.stream, L, bs, st, ids, gv, v);
    return SPX_OK;
  });
}

int launch_scatter_inv(hipStream_t s, int64_t n, int K, const int *g2l, const T *gv, T *v, int npl) {
  npl = npl >= 16 ? 16 : (npl >= 8 ? 8 : 4);
  const int64_t per_block = (int64_t)npl * kThreads;
  const int64_t blocks = (n + per_block - 1) / per_block;
  if (blocks > INT32_MAX) return fail(SPX_ERR_ARG, "spx_ogs_scatter_inverse: n too large");
  const int kf = (K >= 1 && K <= 4) ? K : 1;
  const dim3 grid((unsigned)blocks, kf == 1 ? K : 1);
  auto go = [&](auto npl_c) {
    constexpr int NPL = decltype(npl_c)::value;
    switch (kf) {
      case 1: hipLaunchKernelGGL((k_scatter_inv<T, 1, NPL>), grid, dim3(kThreads), 0, s, n, K, g2l, gv, v); break;
      case 2: hipLaunchKernelGGL((k_scatter_inv<T, 2, NPL>), grid, dim3(kThreads), 0, s, n, K, g2l, gv, v); break;
      case 3: hipLaunchKernelGGL((k_scatter_inv<T, 3, NPL>), grid, dim3(kThreads), 0, s, n, K, g2l, gv, v); break;
      default: hipLaunchKernelGGL((k_scatter_inv<T, 4, NPL>), grid, dim3(kThreads), 0, s, n, K, g2l, gv, v); break;
    }
  };
  if (npl == 16) go(std::integral_constant<int, 16>{});
  else if (npl == 8) go(std::integral_constant<int, 8>{});
  else go(std::integral_constant<int, 4>{});
  return SPX_OK;
}

int launch_extract(hipStream_t s, int64_t total, int K, const int *ids, const T *q, T *gq) {
  int64_t g = (total + kThreads - 1) / kThreads;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_extract<T>, dim3((unsigned)g), dim3(kThreads), 0, s, total, K, ids, q, gq);
  return SPX_OK;
}

}  // namespace ogsk
}  // namespace spx
