// Error reporting, version and fill kernels shared by every entry point.
#include <mutex>

#include "fh_common.h"

namespace fh {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = 0; }

void keep_pool_memory() {
  static std::mutex mu;
  static bool done[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
  std::lock_guard<std::mutex> lock(mu);
  if (done[dev]) return;
  done[dev] = true;
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess || pool == nullptr) return;
  uint64_t keep = kPoolKeepBytes;
  (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, T value, int64_t count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = value;
}

template <typename T>
static int fill_impl(T* p, T value, int64_t count, hipStream_t s) {
  if (count <= 0) return FH_OK;
  int64_t blocks = ceil_div(count, 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, value, count);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fill_u32(uint32_t* p, uint32_t value, int64_t count, hipStream_t s) {
  return fill_impl<uint32_t>(p, value, count, s);
}
int fill_u64(uint64_t* p, uint64_t value, int64_t count, hipStream_t s) {
  return fill_impl<uint64_t>(p, value, count, s);
}

// Throughput probe of the comparison kernel's match recurrence: per thread,
// 8 independent dwords, each step one v_bitop3 (or3), one add and one
// v_bitop3 (select) on register operands only, so the VALU issue rate is the
// only limit. One dword per thread is written so nothing is dead.
constexpr int kProbeChains = 8;
constexpr int kProbeSteps = 8;  // unrolled steps per iteration

__global__ void __launch_bounds__(256) k_probe_valu(uint32_t* __restrict__ out, int iters, uint32_t seed) {
  uint32_t m[kProbeChains], na[kProbeChains];
#pragma unroll
  for (int k = 0; k < kProbeChains; ++k) {
    m[k] = (seed * (uint32_t)(k + 1)) ^ threadIdx.x;
    na[k] = 0;
  }
  const uint32_t w = seed >> 7;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int s = 0; s < kProbeSteps; ++s) {
#pragma unroll
      for (int u = 0; u < kProbeChains; ++u) {
        const uint32_t nc = __builtin_amdgcn_bitop3_b32(m[(s + u) % kProbeChains], w, na[u], 0xFE);
        const uint32_t t = nc + 0x01010101u;
        na[u] = __builtin_amdgcn_bitop3_b32(nc, na[u], t, 0xCA);
      }
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < kProbeChains; ++k) r ^= na[k];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = r;
}

}  // namespace fh

extern "C" int fh_probe_valu(uint32_t* out, int32_t blocks, int32_t iters, double* lane_ops, fh_stream_t stream) {
  fh::clear_error();
  FH_REQUIRE(out != nullptr && blocks > 0 && iters > 0, "fh_probe_valu: bad arguments");
  hipLaunchKernelGGL(fh::k_probe_valu, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, out, iters,
                     0x9E3779B9u);
  FH_HIP(hipGetLastError());
  if (lane_ops)
    *lane_ops = 3.0 * fh::kProbeChains * fh::kProbeSteps * (double)iters * (double)blocks * 256.0;
  return FH_OK;
}

extern "C" const char* fh_last_error(void) { return fh::g_err; }

extern "C" int fh_version(void) { return 100; }
