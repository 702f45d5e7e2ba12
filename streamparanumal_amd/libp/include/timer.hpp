// Host timers after device/communicator sync (include/timer.hpp:37-55,
// libs/core/timer.cpp:32-65).  Device-side event timing is in platform.hpp.
#pragma once

#include <chrono>

#include "platform.hpp"

namespace libp {

using timePoint_t = std::chrono::time_point<std::chrono::high_resolution_clock>;

timePoint_t Time();
timePoint_t GlobalTime(comm_t comm);
timePoint_t PlatformTime(platform_t &platform);
timePoint_t GlobalPlatformTime(platform_t &platform);
timePoint_t GlobalPlatformTime(platform_t &platform, comm_t comm);
double ElapsedTime(const timePoint_t start, const timePoint_t end);

}  // namespace libp
