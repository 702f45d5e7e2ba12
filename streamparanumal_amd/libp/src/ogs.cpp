// ogs_t operations on one rank (libs/ogs/ogs.cpp:39-577).
//
// With one rank the halo operator is empty and no exchange happens, so each
// Start half only checks preconditions and each Finish half enqueues the
// local kernel on the current stream.  No host synchronisation is inserted
// (the reference's device.finish() calls at ogs.cpp:74,243,435 guard MPI
// buffers that do not exist here).  Host-memory overloads copy to the device,
// run the same kernel, and copy back.
#include "ogs.hpp"

namespace libp {
namespace ogs {

// ------------------------------------------------------------ GatherScatter --
template <typename T>
void ogs_t::GatherScatter(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans) {
  GatherScatterStart(o_v, k, op, trans);
  GatherScatterFinish(o_v, k, op, trans);
}
template <typename T>
void ogs_t::GatherScatterStart(deviceMemory<T> o_v, const int k, const Op, const Transpose) {
  LIBP_ABORT("ogs_t not set up", !gatherLocal);
  LIBP_ABORT("GatherScatter vector holds " << o_v.length() << " entries, needs " << (size_t)N * k,
             o_v.length() < (size_t)N * k);
}
template <typename T>
void ogs_t::GatherScatterFinish(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans) {
  gatherLocal->GatherScatter(o_v, k, op, trans);
}

template <typename T>
void ogs_t::GatherScatter(memory<T> v, const int k, const Op op, const Transpose trans) {
  GatherScatterStart(v, k, op, trans);
  GatherScatterFinish(v, k, op, trans);
}
template <typename T>
void ogs_t::GatherScatterStart(memory<T> v, const int k, const Op, const Transpose) {
  LIBP_ABORT("ogs_t not set up", !gatherLocal);
  LIBP_ABORT("GatherScatter vector holds " << v.length() << " entries, needs " << (size_t)N * k,
             v.length() < (size_t)N * k);
}
template <typename T>
void ogs_t::GatherScatterFinish(memory<T> v, const int k, const Op op, const Transpose trans) {
  const size_t n = (size_t)N * k;
  deviceMemory<T> o_v = platform.malloc<T>(n, v);
  gatherLocal->GatherScatter(o_v, k, op, trans);
  o_v.copyTo(v, n);
}

// ------------------------------------------------------------------ Gather --
template <typename T>
void ogs_t::Gather(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans) {
  GatherStart(o_gv, o_v, k, op, trans);
  GatherFinish(o_gv, o_v, k, op, trans);
}
template <typename T>
void ogs_t::GatherStart(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op, const Transpose trans) {
  AssertGatherDefined();
  const size_t rows = (trans == NoTrans) ? (size_t)NlocalP : (size_t)NlocalT;
  LIBP_ABORT("Gather output holds " << o_gv.length() << " entries, needs " << rows * k, o_gv.length() < rows * k);
  LIBP_ABORT("Gather input holds " << o_v.length() << " entries, needs " << (size_t)N * k,
             o_v.length() < (size_t)N * k);
}
template <typename T>
void ogs_t::GatherFinish(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans) {
  AssertGatherDefined();
  gatherLocal->Gather(o_gv, o_v, k, op, trans);
}

template <typename T>
void ogs_t::Gather(memory<T> gv, const memory<T> v, const int k, const Op op, const Transpose trans) {
  GatherStart(gv, v, k, op, trans);
  GatherFinish(gv, v, k, op, trans);
}
template <typename T>
void ogs_t::GatherStart(memory<T> gv, const memory<T> v, const int k, const Op, const Transpose trans) {
  AssertGatherDefined();
  const size_t rows = (trans == NoTrans) ? (size_t)NlocalP : (size_t)NlocalT;
  LIBP_ABORT("Gather output holds " << gv.length() << " entries, needs " << rows * k, gv.length() < rows * k);
  LIBP_ABORT("Gather input holds " << v.length() << " entries, needs " << (size_t)N * k, v.length() < (size_t)N * k);
}
template <typename T>
void ogs_t::GatherFinish(memory<T> gv, const memory<T> v, const int k, const Op op, const Transpose trans) {
  AssertGatherDefined();
  const size_t rows = (trans == NoTrans) ? (size_t)NlocalP : (size_t)NlocalT;
  deviceMemory<T> o_v = platform.malloc<T>((size_t)N * k, v);
  deviceMemory<T> o_gv = platform.malloc<T>(rows * k);
  gatherLocal->Gather(o_gv, o_v, k, op, trans);
  o_gv.copyTo(gv, rows * k);
}

// ----------------------------------------------------------------- Scatter --
template <typename T>
void ogs_t::Scatter(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans) {
  ScatterStart(o_v, o_gv, k, trans);
  ScatterFinish(o_v, o_gv, k, trans);
}
template <typename T>
void ogs_t::ScatterStart(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans) {
  AssertGatherDefined();
  const size_t rows = (trans == Trans) ? (size_t)NlocalP : (size_t)NlocalT;
  LIBP_ABORT("Scatter input holds " << o_gv.length() << " entries, needs " << rows * k, o_gv.length() < rows * k);
  LIBP_ABORT("Scatter output holds " << o_v.length() << " entries, needs " << (size_t)N * k,
             o_v.length() < (size_t)N * k);
}
template <typename T>
void ogs_t::ScatterFinish(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans) {
  AssertGatherDefined();
  gatherLocal->Scatter(o_v, o_gv, k, trans);
}

template <typename T>
void ogs_t::Scatter(memory<T> v, const memory<T> gv, const int k, const Transpose trans) {
  ScatterStart(v, gv, k, trans);
  ScatterFinish(v, gv, k, trans);
}
template <typename T>
void ogs_t::ScatterStart(memory<T> v, const memory<T> gv, const int k, const Transpose trans) {
  AssertGatherDefined();
  const size_t rows = (trans == Trans) ? (size_t)NlocalP : (size_t)NlocalT;
  LIBP_ABORT("Scatter input holds " << gv.length() << " entries, needs " << rows * k, gv.length() < rows * k);
  LIBP_ABORT("Scatter output holds " << v.length() << " entries, needs " << (size_t)N * k, v.length() < (size_t)N * k);
}
template <typename T>
void ogs_t::ScatterFinish(memory<T> v, memory<T> gv, const int k, const Transpose trans) {
  AssertGatherDefined();
  const size_t rows = (trans == Trans) ? (size_t)NlocalP : (size_t)NlocalT;
  deviceMemory<T> o_gv = platform.malloc<T>(rows * k, gv);
  deviceMemory<T> o_v = platform.malloc<T>((size_t)N * k, v);  // entries outside every row keep their value
  gatherLocal->Scatter(o_v, o_gv, k, trans);
  o_v.copyTo(v, (size_t)N * k);
}

#define OGS_INSTANTIATE(T)                                                                                     \
  template void ogs_t::GatherScatter(memory<T>, const int, const Op, const Transpose);                         \
  template void ogs_t::GatherScatterStart(memory<T>, const int, const Op, const Transpose);                    \
  template void ogs_t::GatherScatterFinish(memory<T>, const int, const Op, const Transpose);                   \
  template void ogs_t::GatherScatter(deviceMemory<T>, const int, const Op, const Transpose);                   \
  template void ogs_t::GatherScatterStart(deviceMemory<T>, const int, const Op, const Transpose);              \
  template void ogs_t::GatherScatterFinish(deviceMemory<T>, const int, const Op, const Transpose);             \
  template void ogs_t::Gather(memory<T>, const memory<T>, const int, const Op, const Transpose);               \
  template void ogs_t::GatherStart(memory<T>, const memory<T>, const int, const Op, const Transpose);          \
  template void ogs_t::GatherFinish(memory<T>, const memory<T>, const int, const Op, const Transpose);         \
  template void ogs_t::Gather(deviceMemory<T>, deviceMemory<T>, const int, const Op, const Transpose);         \
  template void ogs_t::GatherStart(deviceMemory<T>, deviceMemory<T>, const int, const Op, const Transpose);    \
  template void ogs_t::GatherFinish(deviceMemory<T>, deviceMemory<T>, const int, const Op, const Transpose);   \
  template void ogs_t::Scatter(memory<T>, const memory<T>, const int, const Transpose);                        \
  template void ogs_t::ScatterStart(memory<T>, const memory<T>, const int, const Transpose);                   \
  template void ogs_t::ScatterFinish(memory<T>, memory<T>, const int, const Transpose);                        \
  template void ogs_t::Scatter(deviceMemory<T>, deviceMemory<T>, const int, const Transpose);                  \
  template void ogs_t::ScatterStart(deviceMemory<T>, deviceMemory<T>, const int, const Transpose);             \
  template void ogs_t::ScatterFinish(deviceMemory<T>, deviceMemory<T>, const int, const Transpose);

OGS_INSTANTIATE(float)
OGS_INSTANTIATE(double)
OGS_INSTANTIATE(int)
OGS_INSTANTIATE(long long int)

}  // namespace ogs
}  // namespace libp
