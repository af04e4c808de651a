// hj_util.hpp -- fail-stop helpers and the host stopwatch.
#ifndef HJ_UTIL_HPP
#define HJ_UTIL_HPP

#include <cstdio>
#include <cstdlib>
#include <chrono>

// Print and exit, like DIE of share/my_utils.hpp:22-29.
#define HJ_DIE(msg)                                                              \
  do {                                                                           \
    (void)fprintf(stderr, "[ERROR] %s(%d): %s\n", __FILE__, __LINE__, (msg));    \
    exit(EXIT_FAILURE);                                                          \
  } while (0)

// Print and exit on a HIP runtime error, like CUDA_CALL of
// share/cuda_helper.hpp:10-18.
#define HIP_CALL(call)                                                           \
  do {                                                                           \
    const hipError_t hj_err_ = (call);                                           \
    if (hipSuccess != hj_err_) {                                                 \
      (void)fprintf(stderr, "HIP runtime error %d [%s] @ %s(%d)!\n",             \
                    static_cast<int>(hj_err_), hipGetErrorString(hj_err_),       \
                    __FILE__, __LINE__);                                         \
      exit(EXIT_FAILURE);                                                        \
    }                                                                            \
  } while (0)

namespace hj {
// microseconds since an arbitrary epoch (share/my_utils.cu:63-88 uses
// gettimeofday; a monotonic clock here)
inline long long timestamp_us()
{
  return std::chrono::duration_cast<std::chrono::microseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}
inline void stopwatch_reset(long long &sw) { sw = timestamp_us(); }
inline long long stopwatch_lap(long long &sw)
{
  const long long ts = timestamp_us(), lap = ts - sw;
  sw = ts;
  return lap;
}
constexpr double TS2S = 1e-6;
}  // namespace hj

#endif
