// platform_t over libspx (include/platform.hpp:60-176).
//
// The reference's platform owns an occa::device set up from the -m/-d
// settings and JIT-builds OKL kernels.  Here the device is the HIP device of
// libspx and buildKernel resolves (okl file, kernel name, defines) to a
// precompiled gfx950 kernel family: the BS setup code keeps its
// platform.buildKernel(...) / kernel(args...) call sites unchanged.
#pragma once

#include <functional>

#include "comm.hpp"
#include "memory.hpp"
#include "settings.hpp"

namespace libp {

void platformAddSettings(settings_t &settings);
void platformReportSettings(settings_t &settings);

// occa::stream stand-in
class stream_t {
 public:
  stream_t() = default;
  explicit stream_t(spx_stream_t s) : s_(s), init_(true) {}
  bool isInitialized() const { return init_; }
  spx_stream_t handle() const { return s_; }

 private:
  spx_stream_t s_ = nullptr;
  bool init_ = false;
};

// occa::device stand-in: one HIP device per process
class device_t {
 public:
  std::string mode() const { return mode_; }
  void finish() { SPX_CHECK(spx_device_finish()); }
  stream_t createStream() {
    spx_stream_t s;
    SPX_CHECK(spx_stream_create(&s));
    return stream_t(s);
  }
  stream_t getStream() { return stream_t(spx_get_stream()); }
  void setStream(stream_t s) { SPX_CHECK(spx_set_stream(s.handle())); }
  int id() const { return id_; }

 private:
  friend class platform_t;
  std::string mode_ = "HIP";
  int id_ = 0;
};

// One kernel argument, type-erased.
struct kernelArg_t {
  enum Kind { Int, Float, Ptr } kind;
  long long i = 0;
  double d = 0.0;
  void *p = nullptr;
};

// occa::kernel stand-in: callable with the OKL kernel's argument list.
class kernel_t {
 public:
  using handler_t = std::function<void(const std::vector<kernelArg_t> &)>;
  kernel_t() = default;
  kernel_t(std::string name, handler_t h) : name_(std::move(name)), h_(std::move(h)) {}
  bool isInitialized() const { return static_cast<bool>(h_); }
  const std::string &name() const { return name_; }

  template <typename... Args>
  void operator()(Args &&...args) const {
    LIBP_ABORT("kernel not built", !h_);
    std::vector<kernelArg_t> v;
    v.reserve(sizeof...(Args));
    (push(v, std::forward<Args>(args)), ...);
    h_(v);
  }

 private:
  template <typename T>
  static void push(std::vector<kernelArg_t> &v, T &&a) {
    using U = std::decay_t<T>;
    kernelArg_t k{};
    if constexpr (std::is_integral_v<U>) {
      k.kind = kernelArg_t::Int;
      k.i = (long long)a;
    } else if constexpr (std::is_floating_point_v<U>) {
      k.kind = kernelArg_t::Float;
      k.d = (double)a;
    } else {
      k.kind = kernelArg_t::Ptr;
      k.p = (void *)const_cast<U &>(a).ptr();
    }
    v.push_back(k);
  }
  std::string name_;
  handler_t h_;
};

// Device-side timer (hip events on the current stream).
class deviceTimer_t {
 public:
  deviceTimer_t();
  ~deviceTimer_t();
  deviceTimer_t(const deviceTimer_t &) = delete;
  deviceTimer_t &operator=(const deviceTimer_t &) = delete;
  void start();
  void stop();
  double seconds();  // waits for the stop event

 private:
  spx_event_t a_ = nullptr, b_ = nullptr;
};

namespace internal {
class iplatform_t {
 public:
  settings_t &settings;
  properties_t props;
  explicit iplatform_t(settings_t &s) : settings(s) {}
};
}  // namespace internal

class platform_t {
 public:
  comm_t comm;
  device_t device;

  platform_t() = default;
  explicit platform_t(settings_t &settings);
  platform_t(const platform_t &other) = default;
  platform_t &operator=(const platform_t &other) = default;

  bool isInitialized() const { return iplatform_ != nullptr; }
  void assertInitialized() const { LIBP_ABORT("Platform not initialized.", !isInitialized()); }

  kernel_t buildKernel(std::string fileName, std::string kernelName, properties_t &kernelInfo);

  template <typename T>
  deviceMemory<T> malloc(const size_t count, const properties_t & = properties_t()) {
    assertInitialized();
    return deviceMemory<T>(count);
  }
  template <typename T>
  deviceMemory<T> malloc(const size_t count, const memory<T> src, const properties_t & = properties_t()) {
    assertInitialized();
    deviceMemory<T> m(count);
    if (count) m.copyFrom(src, count);
    return m;
  }
  template <typename T>
  deviceMemory<T> malloc(const memory<T> src, const properties_t & = properties_t()) {
    return malloc<T>(src.length(), src);
  }
  template <typename T>
  pinnedMemory<T> hostMalloc(const size_t count) {
    assertInitialized();
    return pinnedMemory<T>(count);
  }
  template <typename T>
  pinnedMemory<T> hostMalloc(const size_t count, const memory<T> src) {
    pinnedMemory<T> m = hostMalloc<T>(count);
    if (count) std::copy(src.ptr(), src.ptr() + count, m.ptr());
    return m;
  }
  template <typename T>
  pinnedMemory<T> hostMalloc(const memory<T> src) {
    return hostMalloc<T>(src.length(), src);
  }

  settings_t &settings() {
    assertInitialized();
    return iplatform_->settings;
  }
  properties_t &props() {
    assertInitialized();
    return iplatform_->props;
  }
  void finish() { device.finish(); }
  int rank() const { return comm.rank(); }
  int size() const { return comm.size(); }
  int getDeviceCount(const std::string mode);
  void setCacheDir(const std::string) {}  // kernels are prebuilt: no JIT cache

 private:
  void DeviceConfig();
  void DeviceProperties();
  std::shared_ptr<internal::iplatform_t> iplatform_;
};

}  // namespace libp
