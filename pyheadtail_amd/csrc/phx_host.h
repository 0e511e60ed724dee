// Host-side helpers of the C ABI: error reporting and launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdarg.h>

#include "../../include/phx.h"
#include "phx_device.h"

namespace phx {

void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

inline hipStream_t as_stream(phx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t n_chunks(int64_t n) { return phx_n_groups(n); }

// The opaque handle of phx_graph_capture / phx_batch_capture.
struct GraphHandle {
  hipGraphExec_t exec;
};

// Byte layout of the shared workspace of the reduction kernels.
struct WsLayout {
  static constexpr size_t kTicketBytes = 256;         // reserved head of the workspace

};

}  // namespace phx

// Device address of the allocation-free engine's scratch on `device`
// (current), looked up once per device (phx_engine.hip).
int phx_engine_scratch_init(int device);

#define PHX_CHECK_LAUNCH(name)                                                   \
  do {                                                                           \
    hipError_t _e = hipGetLastError();                                           \
    if (_e != hipSuccess)                                                        \
      return phx::fail((int)_e, "%s: launch failed: %s", name, hipGetErrorString(_e)); \
  } while (0)

#define PHX_CHECK_HIP(call, name)                                                \
  do {                                                                           \
    hipError_t _e = (call);                                                      \
    if (_e != hipSuccess)                                                        \
      return phx::fail((int)_e, "%s: %s", name, hipGetErrorString(_e));         \
  } while (0)
