// Comparison-kernel instances for string-1 lengths 1..16 (every mask kind).
#include "fh_jw_core.h"

namespace fh {

int launch_jw_tiles_r0(int kind, const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0, int64_t nt,
                       hipStream_t s) {
  return launch_kind<1, 16>(kind, args, lds_bytes, w0, nw, t0, nt, s);
}

}  // namespace fh
