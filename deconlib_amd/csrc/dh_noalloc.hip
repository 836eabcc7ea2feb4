// dh_conv_apply_static: one convolution / correlation application whose
// workspace is a file-scope device array instead of a caller buffer, for
// callers that cannot allocate device memory (code under graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// The array lives in this translation unit.  Its host-side name is only a
// stand-in for the symbol, so the device address is looked up once per device
// with hipGetSymbolAddress (with that device current) and cached; the kernels
// receive that address, never the array's name.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/deconhip.h"

namespace dh {
int set_last_error(int code, const std::string& msg);  // dh_abi.hip: the text dh_last_error() returns
}

static __device__ __align__(256) char scratch[DH_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {nullptr};

int static_fail(int code, const std::string& msg) { return dh::set_last_error(code, msg); }

// Device address of this device's scratch array (looked up on first use).
int scratch_for_current_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return static_fail(DH_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return static_fail(DH_EUNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return static_fail(DH_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return DH_OK;
}

}  // namespace

extern "C" {

int dh_conv_apply_static(const dh_conv_plan* plan, int adjoint, const float* in, float* out, const void* otf_dev,
                         const dh_fuse_args* fuse, void* stream) {
  if (!plan) return static_fail(DH_EINVAL, "null plan");
  dh_plan_info info;
  int rc = dh_conv_plan_info(plan, &info);
  if (rc != DH_OK) return rc;
  if (info.work_bytes > (int64_t)DH_STATIC_SCRATCH_BYTES)
    return static_fail(DH_EWORKSPACE, "dh_conv_apply_static: this plan needs a " + std::to_string(info.work_bytes) +
                                          "-byte workspace, above the static scratch of " +
                                          std::to_string((long long)DH_STATIC_SCRATCH_BYTES) +
                                          " bytes; use dh_conv_apply with a caller workspace");
  void* ws = nullptr;
  rc = scratch_for_current_device(&ws);
  if (rc != DH_OK) return rc;
  // every pass writes the workspace regions it later reads (spectra, partial sums)
  // within this call, so the scratch needs no clearing between calls
  return dh_conv_apply(plan, adjoint, in, out, otf_dev, ws, info.work_bytes, fuse, stream);
}

}  // extern "C"
