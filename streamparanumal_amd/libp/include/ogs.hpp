// ogs_t: gather/scatter over a partition of local indices by global id
// (include/ogs.hpp:27-166 describes the semantics; API :212-344).
//
// Same public surface as the reference for one rank per GPU: Setup,
// SetupGlobalToLocalMapping, GatherScatter / Gather / Scatter with their
// Start/Finish halves, for host memory<T> and deviceMemory<T>,
// T in {float, double, int, long long}, Op in {Add, Mul, Max, Min},
// Transpose in {Sym, NoTrans, Trans}, K fields.  All work runs in the HIP
// kernels of libspx; host-memory overloads stage through the device.
#pragma once

#include "core.hpp"
#include "platform.hpp"

namespace libp {
namespace ogs {

/* type enum */
typedef enum { Float, Double, Int32, Int64 } Type;
/* operation enum */
typedef enum { Add, Mul, Max, Min } Op;
/* transpose switch */
typedef enum { Sym, NoTrans, Trans } Transpose;
/* method switch (exchange method; one rank exchanges nothing) */
typedef enum { Auto, Pairwise, CrystalRouter, AllToAll } Method;
/* kind enum */
typedef enum { Unsigned, Signed, Halo } Kind;

constexpr Type Dfloat = Double;
constexpr Type Dlong = Int32;
constexpr Type Hlong = Int64;

template <typename T> struct ogsType;
template <> struct ogsType<float> { static constexpr Type get() { return Float; } };
template <> struct ogsType<double> { static constexpr Type get() { return Double; } };
template <> struct ogsType<int> { static constexpr Type get() { return Int32; } };
template <> struct ogsType<long long int> { static constexpr Type get() { return Int64; } };

// Kernels are prebuilt for every (Type, Op); kept for source compatibility
// with call sites of the reference's JIT step (ogsUtils.cpp:46-123).
inline void InitializeKernels(platform_t &, const Type, const Op) {}

}  // namespace ogs
}  // namespace libp

#include "ogs/ogsOperator.hpp"

namespace libp {
namespace ogs {

class ogsBase_t {
 public:
  platform_t platform;
  comm_t comm;

  dlong N = 0;
  dlong Ngather = 0;  // total number of local positive gather nodes

  dlong NlocalT = 0;  // number of local gather nodes
  dlong NhaloT = 0;   // number of halo gather nodes (0 on one rank)
  dlong NlocalP = 0;  // number of positive local gather nodes
  dlong NhaloP = 0;   // number of positive halo gather nodes (0 on one rank)

  hlong NgatherGlobal = 0;  // global number of positive gather nodes

  Kind kind = Signed;
  bool unique = false;
  bool gather_defined = false;

  ogsBase_t() = default;
  virtual ~ogsBase_t() = default;

  virtual void Setup(const dlong _N, memory<hlong> ids, comm_t _comm, const Kind _kind,
                     const Method method, const bool _unique, const bool verbose,
                     platform_t &_platform);
  void Free();
  // The block-diagonal union of independent set-up handles (no reference
  // counterpart): local nodes of part i follow those of parts 0..i-1, its
  // rows follow theirs, so one launch gathers every part.  All parts share
  // one Kind; the total must fit dlong.
  void SetupBatch(const std::vector<const ogsBase_t *> &parts, platform_t &_platform);
  // The operator Setup builds from the lexicographic GLL ids of an
  // nx x ny x nz hex box at degree p (mesh_t::HexBoxGlobalIds), in closed
  // form without the ids; idsOut (optional, nx*ny*nz*(p+1)^3 entries)
  // receives the ids as Setup leaves them (sign-flagged with unique=true).
  void SetupHexBox(const dlong nx, const dlong ny, const dlong nz, const int p, comm_t _comm, const Kind _kind,
                   const bool _unique, platform_t &_platform, hlong *idsOut = nullptr);

  // the local gather operator (CSR on the device); null before Setup
  std::shared_ptr<ogsOperator_t> localOperator() const { return gatherLocal; }
  double setupSeconds() const { return setupSeconds_; }

 protected:
  std::shared_ptr<ogsOperator_t> gatherLocal;
  void AssertGatherDefined();
  void SetupFromCSR(LocalCSR &csr);
  double setupSeconds_ = 0.0;
};

class ogs_t : public ogsBase_t {
 public:
  ogs_t() = default;
  ~ogs_t() = default;

  void Setup(const dlong _N, memory<hlong> ids, comm_t _comm, const Kind _kind, const Method method,
             const bool _unique, const bool verbose, platform_t &_platform) override;

  void SetupGlobalToLocalMapping(memory<dlong> GlobalToLocal);

  template <typename T> void GatherScatter(memory<T> v, const int k, const Op op, const Transpose trans);
  template <typename T> void GatherScatterStart(memory<T> v, const int k, const Op op, const Transpose trans);
  template <typename T> void GatherScatterFinish(memory<T> v, const int k, const Op op, const Transpose trans);
  template <typename T> void GatherScatter(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans);
  template <typename T> void GatherScatterStart(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans);
  template <typename T> void GatherScatterFinish(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans);

  template <typename T> void Gather(memory<T> gv, const memory<T> v, const int k, const Op op, const Transpose trans);
  template <typename T> void GatherStart(memory<T> gv, const memory<T> v, const int k, const Op op, const Transpose trans);
  template <typename T> void GatherFinish(memory<T> gv, const memory<T> v, const int k, const Op op, const Transpose trans);
  template <typename T> void Gather(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans);
  template <typename T> void GatherStart(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans);
  template <typename T> void GatherFinish(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans);

  template <typename T> void Scatter(memory<T> v, const memory<T> gv, const int k, const Transpose trans);
  template <typename T> void ScatterStart(memory<T> v, const memory<T> gv, const int k, const Transpose trans);
  template <typename T> void ScatterFinish(memory<T> v, memory<T> gv, const int k, const Transpose trans);
  template <typename T> void Scatter(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans);
  template <typename T> void ScatterStart(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans);
  template <typename T> void ScatterFinish(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans);
};

}  // namespace ogs
}  // namespace libp
