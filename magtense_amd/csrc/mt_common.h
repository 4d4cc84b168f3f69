// Shared definitions for the MI355X micromagnetic library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace mt {

// ---- error reporting (thread-local message, int status codes) ---------------
void set_error(const std::string& msg);
const char* last_error();

enum Status {
  kOk = 0,
  kErrArg = -1,       // invalid argument
  kErrHip = -2,       // HIP runtime failure
  kErrSolver = -3,    // RKSUITE fatal condition
  kErrAlloc = -4,     // device allocation failure
  kErrState = -5,     // API used in the wrong order
  kErrUnsupported = -6,  // input outside the supported subset (e.g. a non-prism tile)
};

// A failed allocation (hipErrorOutOfMemory) returns kErrAlloc, any other
// runtime failure kErrHip; the message names the call and its location.  The
// runtime's last-error state is cleared, so a later hipGetLastError() check
// (after a kernel launch) does not report this failure again.
#define MT_HIP(expr)                                                                              \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) {                                                                       \
      (void)hipGetLastError();                                                                    \
      ::mt::set_error(std::string("HIP error '") + hipGetErrorString(_e) + "' at " + __FILE__ + \
                      ":" + std::to_string(__LINE__) + " (" #expr ")");                          \
      return _e == hipErrorOutOfMemory ? ::mt::kErrAlloc : ::mt::kErrHip;                         \
    }                                                                                             \
  } while (0)

// Device allocation owned by scope: every return path frees it (hipFree
// waits for the device, so pending work on the buffer completes first).
template <class T>
struct DevArr {
  T* p = nullptr;
  DevArr() = default;
  DevArr(const DevArr&) = delete;
  DevArr& operator=(const DevArr&) = delete;
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t count) { return hipMalloc(&p, count * sizeof(T)); }
};

#define MT_CHECK(cond, msg)          \
  do {                               \
    if (!(cond)) {                   \
      ::mt::set_error(msg);          \
      return ::mt::kErrArg;          \
    }                                \
  } while (0)

// ---- geometry of the fragment-tiled split-BF16 demag tensor ------------------
// The MFMA consumer (v_mfma_f32_32x32x16_bf16) takes, per lane, 8 consecutive
// k-values of one row (A) or one column (B).  The tensor and the packed
// magnetization are stored directly in that order so that every fragment is
// one contiguous, fully coalesced 1 KiB read (64 lanes x 16 B).
constexpr int kTileM = 32;   // rows per A fragment
constexpr int kTileK = 16;   // k per MFMA
constexpr int kFragBytes = 1024;
constexpr int kMaxCols = 64;  // states per GEMM pass
constexpr int kTileRec = 48;  // doubles per magnetostatic tile record

__host__ __device__ inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Bytes of the split tensor for n cells (6 components, hi+lo planes).
inline int64_t split_tensor_bytes(int n) {
  int64_t np = round_up(n, kTileM);
  return np * np * 6 * 4;
}
// Bytes of the packed magnetization operand (3 components, 64 columns, hi+lo).
// Translation-invariant fragment table of a uniform grid (nx % 32 == 0): the
// A fragment of row tile I and k-chunk J depends only on the cell offset
// (x0 - x1, iy - jy, iz - jz); U = nx/8 - 2 x-offsets per (dy, dz) slab.
inline int64_t toeplitz_fragments(int nx, int ny, int nz) {
  return (int64_t)(2 * nz - 1) * (2 * ny - 1) * (nx / 8 - 2) * 12;
}
inline int64_t toeplitz_tensor_bytes(int nx, int ny, int nz) { return toeplitz_fragments(nx, ny, nz) * kFragBytes; }

inline int64_t packed_m_bytes(int n) {
  int64_t np = round_up(n, kTileM);
  return np * 3 * kMaxCols * 4;
}

}  // namespace mt
