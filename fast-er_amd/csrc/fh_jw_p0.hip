// Prefix-form comparison instances for string-1 lengths 1..16.
#include "fh_jw_prefix.h"

namespace fh {

int launch_jw_prefix_r0(int kind, const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0,
                        int64_t nt, hipStream_t s) {
  return launch_prefix_kind<1, 16>(kind, args, lds_bytes, w0, nw, t0, nt, s);
}

int launch_eval_prefix(const JwArgs* args, int64_t n, hipStream_t s) {
  if (n <= 0) return FH_OK;
  const int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 1 << 16);
  hipLaunchKernelGGL(k_eval_prefix, dim3((unsigned)blocks), dim3(256), 0, s, args, n);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

}  // namespace fh
