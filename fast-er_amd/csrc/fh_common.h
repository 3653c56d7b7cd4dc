// Shared host/device helpers for libfaster_hip (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/faster_hip.h"

namespace fh {

// ---------------------------------------------------------------------------
// Errors: every C-ABI entry point returns an int status and leaves a
// thread-local message readable through fh_last_error().
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
void clear_error();

#define FH_HIP(expr)                                                          \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::fh::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,      \
                      hipGetErrorString(_e));                                 \
      return FH_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)

#define FH_TRY(expr)                 \
  do {                               \
    int _s = (expr);                 \
    if (_s != FH_OK) return _s;      \
  } while (0)

#define FH_REQUIRE(cond, ...)             \
  do {                                    \
    if (!(cond)) {                        \
      ::fh::set_error(__VA_ARGS__);       \
      return FH_ERR_ARG;                  \
    }                                     \
  } while (0)

// Stream-ordered device scratch owned by one entry-point call. Memory comes
// from HIP's stream-ordered pool and is returned on the same stream when the
// object goes out of scope.
// Keeps up to kPoolKeepBytes of freed scratch in the current device's
// stream-ordered pool instead of returning it to the driver at every
// synchronisation (the pool's default), so that the scratch of repeated
// calls is served from the pool. Once per device per process.
void keep_pool_memory();
constexpr uint64_t kPoolKeepBytes = (uint64_t)16 << 30;

class Scratch {
 public:
  explicit Scratch(hipStream_t s) : stream_(s) { keep_pool_memory(); }
  ~Scratch() { release(); }
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;

  // Allocates `bytes` (rounded up to 256) and returns FH_OK or an error.
  template <typename T>
  int alloc(T** out, size_t count) {
    size_t bytes = ((count * sizeof(T)) + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    void* p = nullptr;
    hipError_t e = hipMallocAsync(&p, bytes, stream_);
    if (e != hipSuccess) {
      set_error("hipMallocAsync(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
      return FH_ERR_HIP;
    }
    if (n_ >= kMax) {
      (void)hipFreeAsync(p, stream_);
      set_error("Scratch: too many allocations");
      return FH_ERR_INTERNAL;
    }
    ptrs_[n_++] = p;
    *out = static_cast<T*>(p);
    return FH_OK;
  }
  void release() {
    for (int i = n_ - 1; i >= 0; --i) (void)hipFreeAsync(ptrs_[i], stream_);
    n_ = 0;
  }

 private:
  static constexpr int kMax = 64;
  hipStream_t stream_;
  void* ptrs_[kMax] = {};
  int n_ = 0;
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Clears `count` elements with a kernel (no memset nodes, so callers may
// capture the work into a graph).
int fill_u32(uint32_t* p, uint32_t value, int64_t count, hipStream_t s);
int fill_u64(uint64_t* p, uint64_t value, int64_t count, hipStream_t s);

// ---------------------------------------------------------------------------
// Scan and radix sort (fh_scan_sort.hip)
// ---------------------------------------------------------------------------
// Exclusive prefix sum of `in` (n elements) into `out` (n elements); writes
// the total to `total` (device pointer, may be null). In-place is allowed.
int exclusive_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, uint64_t* total,
                       hipStream_t s);
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total,
                       hipStream_t s);

// Stable LSD radix sort of 64-bit keys (optionally carrying 64-bit values) on
// bits [begin_bit, end_bit). Results land in keys (and vals); tmp buffers of
// the same sizes are caller-provided.
int radix_sort_u64(uint64_t* keys, uint64_t* keys_tmp, uint64_t* vals, uint64_t* vals_tmp,
                   int64_t n, int begin_bit, int end_bit, hipStream_t s);
// Sorts the keys `in` into `out` (in is overwritten; distinct buffers).
int radix_sort_u64_into(uint64_t* in, uint64_t* out, int64_t n, int begin_bit, int end_bit, hipStream_t s);
// Same for 32-bit keys with 64-bit values.
int radix_sort_u32(uint32_t* keys, uint32_t* keys_tmp, uint64_t* vals, uint64_t* vals_tmp,
                   int64_t n, int begin_bit, int end_bit, hipStream_t s);
// Sorts (in, vin) into (out, vout); in and vin are overwritten. An odd
// number of 8-bit passes copies nothing.
int radix_sort_u32_into(uint32_t* in, uint32_t* out, uint64_t* vin, uint64_t* vout, int64_t n, int begin_bit,
                        int end_bit, hipStream_t s);

inline int bits_for(uint64_t max_value) {
  int b = 0;
  while (b < 64 && (max_value >> b) != 0) ++b;
  return b;
}

}  // namespace fh

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
namespace fhd {

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

// Order-preserving unsigned key of a value: a signed int64 with the sign bit
// flipped; for the bit pattern of a float64, every NaN becomes one positive
// quiet NaN and -0.0 becomes +0.0 (np.unique treats each as one value), then
// the IEEE sign trick, so NaN sorts last as in np.unique.
__device__ __forceinline__ uint64_t value_key(int64_t v, bool is_f64) {
  uint64_t u = (uint64_t)v;
  if (is_f64) {
    const uint64_t mag = u & 0x7FFFFFFFFFFFFFFFull;
    if (mag > 0x7FF0000000000000ull) u = 0x7FF8000000000000ull;
    else if (mag == 0) u = 0;
    return (u >> 63) ? ~u : (u ^ 0x8000000000000000ull);
  }
  return u ^ 0x8000000000000000ull;
}

// Key of the canonical NaN (the largest float64 key).
constexpr uint64_t kNanKey = 0xFFF8000000000000ull;

// v reduced over the workgroup by op (blockDim.x a multiple of 64, at most
// 1024); the result is valid in thread 0. Every thread must call it. Kernels
// that fold a grid-wide extreme or count into one global word reduce here
// first and add one atomic per workgroup: one per wave on a single address
// serialises at the memory side.
template <typename T, typename Op>
__device__ __forceinline__ T block_reduce(T v, Op op) {
  __shared__ T part[16];
  for (int d = 32; d >= 1; d >>= 1) v = op(v, (T)__shfl_xor(v, d));
  if (__lane_id() == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) v = op(v, part[i]);
  __syncthreads();  // part is reused by the next call
  return v;
}
struct OpMax {
  __device__ unsigned long long operator()(unsigned long long a, unsigned long long b) const { return a > b ? a : b; }
};
struct OpMin {
  __device__ unsigned long long operator()(unsigned long long a, unsigned long long b) const { return a < b ? a : b; }
};
struct OpSum {
  __device__ unsigned long long operator()(unsigned long long a, unsigned long long b) const { return a + b; }
};

}  // namespace fhd
