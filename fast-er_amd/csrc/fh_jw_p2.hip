// Prefix-form comparison instance for string-1 lengths 33..64 (one rolled
// pass per tile kind).
#include "fh_jw_prefix.h"

namespace fh {

int launch_jw_prefix_r2(int kind, const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0,
                        int64_t nt, hipStream_t s) {
  return launch_prefix_kind<33, 64>(kind, args, lds_bytes, w0, nw, t0, nt, s);
}

}  // namespace fh
