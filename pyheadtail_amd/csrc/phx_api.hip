// Library state and error reporting of the C ABI.
#include <string.h>

#include "phx_device.h"
#include "phx_host.h"
#include "phx_slicing.h"

namespace {
thread_local char g_err[512] = "";
size_t g_max_lds = 65536;
int g_num_cus = 256;
}  // namespace

namespace phx {

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace phx

size_t phx_max_lds_bytes() { return g_max_lds; }
int phx_num_cus() { return g_num_cus; }

extern "C" const char *phx_last_error(void) { return g_err; }

extern "C" int phx_abi_version(void) { return PHX_ABI_VERSION; }

extern "C" int phx_init(int device) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count < 1)
    return phx::fail(PHX_E_NODEVICE, "phx_init: no HIP device (%s)", hipGetErrorString(e));
  if (device < 0 || device >= count)
    return phx::fail(PHX_E_INVALID, "phx_init: device %d not in 0..%d", device, count - 1);
  PHX_CHECK_HIP(hipSetDevice(device), "phx_init");
  int lds = 0;
  PHX_CHECK_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device),
                "phx_init");
  if (lds > 0) g_max_lds = (size_t)lds;
  int cus = 0;
  PHX_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device),
                "phx_init");
  if (cus > 0) g_num_cus = cus;
  return phx_engine_scratch_init(device);
}

extern "C" int64_t phx_workspace_bytes(int64_t n, int32_t n_slices) {
  if (n < 0 || n_slices < 0) return -1;
  const int64_t nc = (n + PHX_CHUNK - 1) / PHX_CHUNK;
  const int H = (int)(nc < 256 ? nc : 256);
  return (int64_t)(256 + acc_partials_bytes(n) +
                   acc_parts_bytes(H > 0 ? H : 1, n_slices, PHX_MAX_TERMS));
}

extern "C" int32_t phx_pass_group_size(void) { return PHX_CHUNK; }

extern "C" int64_t phx_sizeof(const char *name) {
  if (!name) return -1;
#define PHX_SZ(T) \
  if (strcmp(name, #T) == 0) return (int64_t)sizeof(T);
  PHX_SZ(phx_slicer)
  PHX_SZ(phx_slice_state)
  PHX_SZ(phx_term)
  PHX_SZ(phx_detuner)
  PHX_SZ(phx_tseg)
  PHX_SZ(phx_long_op)
  PHX_SZ(phx_long_seq)
  PHX_SZ(phx_linmap)
  PHX_SZ(phx_damper)
  PHX_SZ(phx_sc_slices)
  PHX_SZ(phx_spacecharge)
  PHX_SZ(phx_wake_kick)
  PHX_SZ(phx_wake_hist)
  PHX_SZ(phx_turn_program)
  PHX_SZ(phx_engine)
#undef PHX_SZ
  return -1;
}
