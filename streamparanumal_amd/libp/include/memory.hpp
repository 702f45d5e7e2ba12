// Host, device and pinned memory handles over libspx.
//
// memory<T>       : shared host array with offset views (include/memory.hpp:35-282)
// deviceMemory<T> : refcounted device allocation from spx_malloc; views share
//                   the allocation, as occa::memory does (include/memory.hpp:285-460)
// pinnedMemory<T> : page-locked host allocation from spx_malloc_host
//                   (include/memory.hpp:465-640)
// Copies are bounds-checked like the reference and raise libp::exception.
#pragma once

#include <algorithm>
#include <map>

#include "core.hpp"
#include "spx.h"

namespace libp {

// Small key/value property bag standing in for occa::json where the
// reference passes properties ("defines/p_blockSize", {"async": true}).
class properties_t {
 public:
  properties_t() = default;
  properties_t(const std::string &key, bool v) { kv_[key] = v ? "true" : "false"; }
  properties_t(const std::string &key, int v) { kv_[key] = std::to_string(v); }
  properties_t(const std::string &key, const std::string &v) { kv_[key] = v; }

  class ref {
   public:
    ref(std::map<std::string, std::string> &m, const std::string &k) : m_(m), k_(k) {}
    ref &operator=(const std::string &v) { m_[k_] = v; return *this; }
    ref &operator=(const char *v) { m_[k_] = v; return *this; }
    ref &operator=(bool v) { m_[k_] = v ? "true" : "false"; return *this; }
    ref &operator=(int v) { m_[k_] = std::to_string(v); return *this; }
    ref &operator=(long v) { m_[k_] = std::to_string(v); return *this; }
    ref &operator=(long long v) { m_[k_] = std::to_string(v); return *this; }
    ref &operator=(double v) {
      std::ostringstream ss;
      ss.precision(17);
      ss << v;
      m_[k_] = ss.str();
      return *this;
    }
    ref &operator+=(const std::string &v) { m_[k_] += v; return *this; }
    void asObject() {}
    void asArray() {}
    operator std::string() const {
      auto it = m_.find(k_);
      return it == m_.end() ? std::string() : it->second;
    }

   private:
    std::map<std::string, std::string> &m_;
    std::string k_;
  };

  ref operator[](const std::string &key) { return ref(kv_, key); }
  bool has(const std::string &key) const { return kv_.count(key) != 0; }
  std::string get(const std::string &key, const std::string &dflt = "") const {
    auto it = kv_.find(key);
    return it == kv_.end() ? dflt : it->second;
  }
  long long getInt(const std::string &key, long long dflt) const {
    auto it = kv_.find(key);
    return it == kv_.end() ? dflt : std::stoll(it->second);
  }
  bool isAsync() const { return get("async") == "true"; }

 private:
  std::map<std::string, std::string> kv_;
};

// ---------------------------------------------------------------- host --
template <typename T>
class memory {
  template <typename U> friend class memory;

 public:
  memory() : lngth_(0), offset_(0) {}
  explicit memory(size_t n) : ptr_(new T[n ? n : 1], std::default_delete<T[]>()), lngth_(n), offset_(0) {}
  memory(size_t n, const T val) : memory(n) {
    std::fill(ptr(), ptr() + n, val);
  }

  void malloc(size_t n) { *this = memory<T>(n); }
  void calloc(size_t n) { *this = memory<T>(n, T{0}); }
  void realloc(size_t n) {
    memory<T> m(n);
    std::copy(ptr(), ptr() + std::min(n, lngth_), m.ptr());
    *this = m;
  }
  memory &swap(memory<T> &m) {
    std::swap(ptr_, m.ptr_);
    std::swap(lngth_, m.lngth_);
    std::swap(offset_, m.offset_);
    return *this;
  }
  T *ptr() { return ptr_ ? ptr_.get() + offset_ : nullptr; }
  const T *ptr() const { return ptr_ ? ptr_.get() + offset_ : nullptr; }
  T *begin() { return ptr(); }
  T *end() { return ptr() + lngth_; }
  size_t length() const { return lngth_; }
  size_t size() const { return lngth_ * sizeof(T); }
  size_t use_count() const { return ptr_.use_count(); }
  T &operator[](const ptrdiff_t i) const { return ptr_.get()[offset_ + i]; }
  bool isInitialized() const { return ptr_ != nullptr; }
  memory operator+(const ptrdiff_t off) const {
    LIBP_ABORT("libp::memory offset " << off << " outside length " << lngth_,
               off < 0 || (size_t)off > lngth_);
    memory m(*this);
    m.offset_ += off;
    m.lngth_ -= off;
    return m;
  }
  void copyFrom(const T *src, const ptrdiff_t count = -1, const ptrdiff_t off = 0) {
    const ptrdiff_t cnt = (count == -1) ? (ptrdiff_t)lngth_ - off : count;
    LIBP_ABORT("libp::memory::copyFrom Cannot have negative count (" << cnt << ")", cnt < 0);
    LIBP_ABORT("libp::memory::copyFrom Cannot have negative offset (" << off << ")", off < 0);
    LIBP_ABORT("libp::memory::copyFrom Destination memory has size [" << lngth_ << "], trying to access ["
               << off << ", " << off + cnt << "]", (size_t)(off + cnt) > lngth_);
    std::copy(src, src + cnt, ptr() + off);
  }
  void copyFrom(const memory<T> src, const ptrdiff_t count = -1, const ptrdiff_t off = 0) {
    const ptrdiff_t cnt = (count == -1) ? (ptrdiff_t)src.length() : count;
    LIBP_ABORT("libp::memory::copyFrom Source memory has size [" << src.length() << "], trying to access [0, "
               << cnt << "]", (size_t)cnt > src.length());
    copyFrom(src.ptr(), cnt, off);
  }
  void copyTo(T *dest, const ptrdiff_t count = -1, const ptrdiff_t off = 0) const {
    const ptrdiff_t cnt = (count == -1) ? (ptrdiff_t)lngth_ - off : count;
    LIBP_ABORT("libp::memory::copyTo Cannot have negative count (" << cnt << ")", cnt < 0);
    LIBP_ABORT("libp::memory::copyTo Cannot have negative offset (" << off << ")", off < 0);
    LIBP_ABORT("libp::memory::copyTo Source memory has size [" << lngth_ << "], trying to access ["
               << off << ", " << off + cnt << "]", (size_t)(off + cnt) > lngth_);
    std::copy(ptr() + off, ptr() + off + cnt, dest);
  }
  void copyTo(memory<T> dest, const ptrdiff_t count = -1, const ptrdiff_t off = 0) const {
    const ptrdiff_t cnt = (count == -1) ? (ptrdiff_t)lngth_ - off : count;
    LIBP_ABORT("libp::memory::copyTo Destination memory has size [" << dest.length() << "], trying to access [0, "
               << cnt << "]", (size_t)cnt > dest.length());
    copyTo(dest.ptr(), cnt, off);
  }
  memory<T> clone() const {
    memory<T> m(lngth_);
    m.copyFrom(*this);
    return m;
  }
  void free() { *this = memory<T>(); }

 private:
  std::shared_ptr<T> ptr_;
  size_t lngth_, offset_;
};

namespace internal {
// One device (or pinned host) allocation; freed when the last view goes.
struct spx_buffer {
  void *p = nullptr;
  bool pinned = false;
  spx_buffer(size_t bytes, bool pin) : pinned(pin) {
    int rc = pin ? spx_malloc_host(&p, bytes) : spx_malloc(&p, bytes);
    LIBP_ABORT("allocation of " << bytes << " bytes failed: " << spx_last_error(), rc != 0);
  }
  ~spx_buffer() {
    if (p) (void)(pinned ? spx_free_host(p) : spx_free(p));
  }
  spx_buffer(const spx_buffer &) = delete;
  spx_buffer &operator=(const spx_buffer &) = delete;
};

inline void copy_bytes(void *dst, const void *src, size_t bytes, const properties_t &props) {
  if (!bytes) return;
  SPX_CHECK(spx_memcpy(dst, src, bytes, props.isAsync() ? 1 : 0));
}
}  // namespace internal

// Shared implementation of deviceMemory and pinnedMemory.
template <typename T, bool PINNED>
class spxMemory {
 public:
  spxMemory() : lngth_(0), offset_(0) {}
  explicit spxMemory(size_t n)
      : buf_(std::make_shared<internal::spx_buffer>(n * sizeof(T), PINNED)), lngth_(n), offset_(0) {}

  T *ptr() { return buf_ ? static_cast<T *>(buf_->p) + offset_ : nullptr; }
  const T *ptr() const { return buf_ ? static_cast<const T *>(buf_->p) + offset_ : nullptr; }
  size_t length() const { return lngth_; }
  size_t size() const { return lngth_ * sizeof(T); }
  bool isInitialized() const { return buf_ != nullptr; }
  void free() { *this = spxMemory(); }
  size_t use_count() const { return buf_.use_count(); }

  spxMemory operator+(const ptrdiff_t off) const {
    LIBP_ABORT("memory offset " << off << " outside length " << lngth_, off < 0 || (size_t)off > lngth_);
    spxMemory m(*this);
    m.offset_ += off;
    m.lngth_ -= off;
    return m;
  }

  // host -> this
  void copyFrom(const T *src, const ptrdiff_t count = -1, const ptrdiff_t off = 0,
                const properties_t &props = properties_t()) {
    const ptrdiff_t cnt = (count == -1) ? (ptrdiff_t)lngth_ - off : count;
    check_range("copyFrom", "Destination", cnt, off, lngth_);
    internal::copy_bytes(ptr() + off, src, cnt * sizeof(T), props);
  }
  void copyFrom(const memory<T> src, const ptrdiff_t count = -1, const ptrdiff_t off = 0,
                const properties_t &props = properties_t()) {
    const ptrdiff_t cnt = (count == -1) ? (ptrdiff_t)src.length() : count;
    LIBP_ABORT("libp::memory::copyFrom Source memory has size [" << src.length() << "], trying to access [0, "
               << cnt << "]", (size_t)cnt > src.length());
    copyFrom(src.ptr(), cnt, off, props);
  }
  template <bool P2>
  void copyFrom(const spxMemory<T, P2> src, const ptrdiff_t count = -1, const ptrdiff_t off = 0,
                const properties_t &props = properties_t()) {
    const ptrdiff_t cnt = (count == -1) ? (ptrdiff_t)src.length() : count;
    LIBP_ABORT("libp::memory::copyFrom Source memory has size [" << src.length() << "], trying to access [0, "
               << cnt << "]", (size_t)cnt > src.length());
    check_range("copyFrom", "Destination", cnt, off, lngth_);
    internal::copy_bytes(ptr() + off, src.ptr(), cnt * sizeof(T), props);
  }
  // this -> host
  void copyTo(T *dest, const ptrdiff_t count = -1, const ptrdiff_t off = 0,
              const properties_t &props = properties_t()) const {
    const ptrdiff_t cnt = (count == -1) ? (ptrdiff_t)lngth_ - off : count;
    check_range("copyTo", "Source", cnt, off, lngth_);
    internal::copy_bytes(dest, ptr() + off, cnt * sizeof(T), props);
  }
  void copyTo(memory<T> dest, const ptrdiff_t count = -1, const ptrdiff_t off = 0,
              const properties_t &props = properties_t()) const {
    const ptrdiff_t cnt = (count == -1) ? (ptrdiff_t)lngth_ - off : count;
    LIBP_ABORT("libp::memory::copyTo Destination memory has size [" << dest.length() << "], trying to access [0, "
               << cnt << "]", (size_t)cnt > dest.length());
    copyTo(dest.ptr(), cnt, off, props);
  }
  template <bool P2>
  void copyTo(spxMemory<T, P2> dest, const ptrdiff_t count = -1, const ptrdiff_t off = 0,
              const properties_t &props = properties_t()) const {
    const ptrdiff_t cnt = (count == -1) ? (ptrdiff_t)lngth_ - off : count;
    check_range("copyTo", "Source", cnt, off, lngth_);
    LIBP_ABORT("libp::memory::copyTo Destination memory has size [" << dest.length() << "], trying to access [0, "
               << cnt << "]", (size_t)cnt > dest.length());
    internal::copy_bytes(dest.ptr(), ptr() + off, cnt * sizeof(T), props);
  }

  // host access, pinned memory only
  T &operator[](const ptrdiff_t i) const {
    static_assert(PINNED, "operator[] on device memory");
    return static_cast<T *>(buf_->p)[offset_ + i];
  }

 private:
  template <typename U, bool P> friend class spxMemory;
  static void check_range(const char *op, const char *which, ptrdiff_t cnt, ptrdiff_t off, size_t len) {
    LIBP_ABORT("libp::memory::" << op << " Cannot have negative count (" << cnt << ")", cnt < 0);
    LIBP_ABORT("libp::memory::" << op << " Cannot have negative offset (" << off << ")", off < 0);
    LIBP_ABORT("libp::memory::" << op << " " << which << " memory has size [" << len << "], trying to access ["
               << off << ", " << off + cnt << "]", (size_t)(off + cnt) > len);
  }

  std::shared_ptr<internal::spx_buffer> buf_;
  size_t lngth_, offset_;
};

template <typename T>
class deviceMemory : public spxMemory<T, false> {
 public:
  using spxMemory<T, false>::spxMemory;
  deviceMemory() = default;
  deviceMemory(const spxMemory<T, false> &m) : spxMemory<T, false>(m) {}
  deviceMemory operator+(const ptrdiff_t off) const { return spxMemory<T, false>::operator+(off); }
};

template <typename T>
class pinnedMemory : public spxMemory<T, true> {
 public:
  using spxMemory<T, true>::spxMemory;
  pinnedMemory() = default;
  pinnedMemory(const spxMemory<T, true> &m) : spxMemory<T, true>(m) {}
  pinnedMemory operator+(const ptrdiff_t off) const { return spxMemory<T, true>::operator+(off); }
};

}  // namespace libp
