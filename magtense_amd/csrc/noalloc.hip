// Allocation-free batched split-BF16 demag product.
//
// No counterpart in the reference.  mt_demag_apply_bf16x2 allocates its
// workspace (packed states and k-split partial sums) on every call; callers
// that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops, harnesses that forbid allocation) use mt_demag_apply_bf16x2_noalloc,
// which runs the same kernels with the workspace in the file-scope device
// array `scratch` below.
//
// Every device that loads this library holds its own 64 MiB copy of the
// array until the process ends.  Host code never passes the array's name to
// a kernel or copy (on the host it is only a stand-in): the address of the
// current device's copy comes from hipGetSymbolAddress on the first call on
// that device and is kept per device.  Calls that overlap on one device share
// the array and race; the array keeps the last call's contents, and the
// packed-state region is cleared with a fill kernel before each use.

#include <mutex>
#include <string>

#include "mt_common.h"
#include "mt_kernels.h"
#include "../../include/magtense_hip.h"

static constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
uint8_t* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of the current device's copy of `scratch`.
int scratch_of_current_device(uint8_t** out) {
  using namespace mt;
  int dev = -1;
  MT_HIP(hipGetDevice(&dev));
  MT_CHECK(dev >= 0 && dev < kMaxScratchDevices, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    MT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MT_CHECK(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = static_cast<uint8_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}
}  // namespace

extern "C" {

long long mt_noalloc_scratch_bytes(void) { return (long long)kNoAllocScratchBytes; }

long long mt_demag_apply_bf16x2_noalloc_bytes(int n, int nstates) {
  if (n <= 0 || nstates <= 0) return -1;
  return (long long)mt::bf16x2_work(n, n, 0, 0, 0, nstates).bytes();
}

int mt_demag_apply_bf16x2_noalloc(const void* split, int n, const float* m, float* h, int nstates, float pref) {
  using namespace mt;
  int rc = device_ok();
  if (rc) return rc;
  MT_CHECK(n > 0 && nstates > 0 && split && m && h, "invalid demag arguments");
  const Bf16x2Work w = bf16x2_work(n, n, 0, 0, 0, nstates);
  if (w.bytes() > kNoAllocScratchBytes) {
    set_error("mt_demag_apply_bf16x2_noalloc: n = " + std::to_string(n) + " needs " + std::to_string(w.bytes()) +
              " bytes of workspace, beyond the " + std::to_string(kNoAllocScratchBytes) +
              "-byte scratch array (use mt_demag_apply_bf16x2)");
    return kErrArg;
  }
  uint8_t* ws = nullptr;
  rc = scratch_of_current_device(&ws);
  if (rc) return rc;
  return apply_bf16x2_in(split, n, n, 0, 0, 0, m, h, nstates, pref, w, ws);
}

}  // extern "C"
