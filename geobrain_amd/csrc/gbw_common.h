// Utilities shared by the library's source files: the per-thread error text
// behind gbw_last_error(), HIP status checks, the optional per-kernel event
// timing behind gbw_set_kernel_timing() / gbw_kernel_times(), wavefront lane
// shifts, and float4 accesses at 32-bit byte offsets.
#ifndef GBW_COMMON_H
#define GBW_COMMON_H

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gbw_acoustic2d.h"

namespace gbw {

inline thread_local std::string g_last_error;

inline int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(GBW_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCH_CHECK()                                                                 \
  do {                                                                                 \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess)                                                              \
      return fail(GBW_ERR_HIP, std::string("kernel launch failed: ") + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// Optional per-kernel timing (off by default; gbw_set_kernel_timing).  Each
// step-kernel launch is bracketed by two events on the caller's stream; the
// API call that made the launches waits for its stream at its end and adds
// the elapsed times to per-kernel totals.
// ---------------------------------------------------------------------------
struct TimedLaunch {
  const char* name;
  hipEvent_t a, b;
};
inline std::mutex g_timing_mu;
inline bool g_timing = false;
inline std::map<std::string, std::pair<long long, double>> g_times;  // launches, ms
inline std::vector<hipEvent_t> g_event_pool;
inline thread_local std::vector<TimedLaunch> t_pending;
// set while an entry point that must not create runtime objects runs
inline thread_local bool t_no_timing = false;

inline hipEvent_t take_event() {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct LaunchTimer {
  const char* name;
  hipStream_t st;
  hipEvent_t a = nullptr;
  LaunchTimer(const char* n, hipStream_t s) : name(n), st(s) {
    bool on;
    {
      std::lock_guard<std::mutex> lk(g_timing_mu);
      on = g_timing && !t_no_timing;
    }
    if (on) {  // never inside a graph capture: the call could not wait for its events
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) on = false;
    }
    if (on && (a = take_event())) (void)hipEventRecord(a, st);
  }
  ~LaunchTimer() {
    if (!a) return;
    hipEvent_t b = take_event();
    if (!b) return;
    (void)hipEventRecord(b, st);
    t_pending.push_back({name, a, b});
  }
};

// called at the end of an API call: collects the timed launches
inline int flush_timing(hipStream_t st) {
  if (t_pending.empty()) return 0;
  HIP_TRY(hipStreamSynchronize(st));
  std::lock_guard<std::mutex> lk(g_timing_mu);
  for (const TimedLaunch& t : t_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      auto& v = g_times[t.name];
      v.first += 1;
      v.second += ms;
    }
    g_event_pool.push_back(t.a);
    g_event_pool.push_back(t.b);
  }
  t_pending.clear();
  return 0;
}

// lane l-1's value (lane 0: 0) and lane l+1's value (lane 63: 0) of a
// wavefront: DPP wavefront shifts by one lane (bound_ctrl writes 0 at the end)
__device__ __forceinline__ float lane_prev(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_next(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

// element i of a float4 array at a 32-bit byte offset (i < 2^28): the access
// takes a uniform base and a 32-bit lane offset instead of a 64-bit address
__device__ __forceinline__ float4 ld4u(const float4* base, unsigned i) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + (i << 4));
}
__device__ __forceinline__ void st4u(float4* base, unsigned i, const float4& v) {
  *reinterpret_cast<float4*>(reinterpret_cast<char*>(base) + (i << 4)) = v;
}

}  // namespace gbw

#endif  // GBW_COMMON_H
