// Error reporting, version and fill kernels shared by every entry point.
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

}  // namespace fh

extern "C" const char* fh_last_error(void) { return fh::g_err; }

extern "C" int fh_version(void) { return 100; }
