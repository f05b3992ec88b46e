// lx_run_static: the fused batched run with its scratch in a file-scope device array instead of
// the caller's workspace, for callers that cannot allocate device memory.
//
// The array holds the run's queue heads, survivor list, parked carries and per-world step counts
// (lx_run_scratch_layout) for at most LX_RUN_STATIC_MAX_WORLDS worlds.  Each device holds its own
// copy from the library's load until the process ends; its address on a device is looked up once,
// on the first call there, and only that address is handed to kernels.  The run clears the
// counters it needs with a kernel before each use (no memset), so the contents a previous call left
// do not matter; two calls that overlap on one device share the array and race.  The autotuner does
// not time schedules inside this call: it uses forced, recorded or default ones.
#include <mutex>

#include "lx_internal.h"

namespace {

constexpr size_t kStaticBytes = (16 + 4 * size_t(LX_RUN_STATIC_MAX_WORLDS) + 255) / 256 * 256 +
                                (sizeof(lx_carry) * size_t(LX_RUN_STATIC_MAX_WORLDS) + 255) / 256 * 256 +
                                4 * size_t(LX_RUN_STATIC_MAX_WORLDS);

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxDevices = 64;

// the device address of `scratch` on the current device (looked up once per device)
int scratch_address(void** out) {
  static void* addr[kMaxDevices] = {};
  static std::mutex mu;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return lx_fail(LX_EDEVICE, "hipGetDevice failed or device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return lx_fail(LX_EDEVICE, "lx_run_static: hipGetSymbolAddress: %s", hipGetErrorString(e));
    addr[dev] = p;
  }
  *out = addr[dev];
  return LX_OK;
}

}  // namespace

extern "C" int lx_run_static(const lx_rule* rule, const lx_run_args* a, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!a) return lx_fail(LX_EINVAL, "args is NULL");
  if (!lx_fused2d_supported(rule))
    return lx_fail(LX_ENOTSUP, "lx_run_static: only rules with a fused kernel (2-D N x N, N in {32, 64, 128}, one "
                   "channel, one kernel); use lx_run");
  if (a->B > LX_RUN_STATIC_MAX_WORLDS)
    return lx_fail(LX_EINVAL, "lx_run_static: %lld worlds, at most %d (its scratch is a fixed device array)",
                   (long long)a->B, LX_RUN_STATIC_MAX_WORLDS);
  lx_run_args b = *a;
  if ((rc = scratch_address(&b.scratch_d))) return rc;
  if (lx_run_scratch_layout(LX_RUN_STATIC_MAX_WORLDS).total > kStaticBytes)
    return lx_fail(LX_EDEVICE, "lx_run_static: scratch layout larger than its array (internal error)");
  LxTuneQuiet quiet;  // no schedule timing (events, repeated launches) inside this call
  return lx_run(rule, &b, stream);
}
