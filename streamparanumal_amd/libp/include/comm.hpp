// Single-rank communicator.  The reference's comm_t wraps MPI
// (include/comm.hpp:95-579); this build runs one process per GPU with one GPU,
// so every collective is the identity on rank 0 of 1.
#pragma once

#include "core.hpp"

namespace libp {

namespace Comm {
enum op_t { Max, Min, Sum, Prod, And, Or, Xor };
inline void Init(int &, char **&) {}
inline void Finalize() {}
}  // namespace Comm

class comm_t {
 public:
  static comm_t world() { return comm_t(); }
  comm_t Dup() const { return *this; }
  void Free() {}
  int rank() const { return 0; }
  int size() const { return 1; }
  void Barrier() const {}
  template <typename T> void Allreduce(T &, const Comm::op_t = Comm::Sum) const {}
  template <typename T> void Reduce(T &, const int, const Comm::op_t = Comm::Sum) const {}
  template <typename T> void Scan(const T &snd, T &rcv, const Comm::op_t = Comm::Sum) const { rcv = snd; }
};

namespace Comm {
// the world communicator (include/comm.hpp:69-92)
inline comm_t World() { return comm_t::world(); }
}  // namespace Comm

}  // namespace libp
