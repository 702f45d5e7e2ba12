// Core utilities: error convention and includes shared by the host layer.
// LIBP_ABORT / LIBP_FORCE_ABORT throw libp::exception as the reference does
// (include/utils.hpp:57-112, libs/core/exception.cpp:79).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "types.h"

namespace libp {

class exception : public std::exception {
 public:
  exception(const std::string &msg, const char *file, const char *func, int line);
  const char *what() const noexcept override { return full_.c_str(); }
  const std::string &message() const { return message_; }

 private:
  std::string message_, full_;
};

}  // namespace libp

// program exit codes (include/utils.hpp:50-51)
#define LIBP_SUCCESS 0
#define LIBP_ERROR -1

#define LIBP_FORCE_ABORT(msg)                                                  \
  do {                                                                         \
    std::stringstream libp_ss_;                                                \
    libp_ss_ << msg;                                                           \
    throw ::libp::exception(libp_ss_.str(), __FILE__, __func__, __LINE__);     \
  } while (0)

#define LIBP_ABORT(msg, cond)          \
  do {                                 \
    if (cond) LIBP_FORCE_ABORT(msg);   \
  } while (0)

// Turns a libspx status into a thrown libp::exception.
#define SPX_CHECK(call)                                                        \
  do {                                                                         \
    int spx_rc_ = (call);                                                      \
    LIBP_ABORT(#call << " failed (" << spx_rc_ << "): " << spx_last_error(),   \
               spx_rc_ != 0);                                                  \
  } while (0)
