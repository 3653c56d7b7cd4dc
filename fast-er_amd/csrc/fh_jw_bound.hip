// All-pairs Jaro-Winkler, bound-filtered form: an int8 MFMA pass computes an
// upper bound of every pair's match count and only the pairs whose bound
// reaches the pruning threshold are scored exactly.
//
// The greedy match count m of a pair (faster/comparison.py:48-64) never
// exceeds a maximum matching of its characters, which never exceeds
// sum_c min(n_1(c), n_2(c)) over symbols c (counts n_1, n_2 of c in the two
// strings); grouping symbols into 32 buckets only raises that sum. Per bucket
// b, min(x, y) <= [x >= 1][y >= 1] + max(x - 1, 0) [y >= 2], so the bound is
// the dot product of two 64-byte vectors per string:
//   string 1: bytes 0..31 = [x_b >= 1], bytes 32..63 = max(x_b - 1, 0)
//   string 2: bytes 0..31 = [y_b >= 1], bytes 32..63 = [y_b >= 2]
// and the bound of all pairs is an int8 matrix product, one
// v_mfma_i32_32x32x32_i8 pair per 32 x 32 block of pairs. A pair can reach
// level 1 only if m >= mmin(len1, len2) (the host table of fh_jw_core.h), so
// pairs with bound < mmin are level 0; the rest go to the exact restatement
// of the reference function (candidate_level). Results equal the bit-parallel
// form's exactly: both apply the same exact scorer to supersets of the pairs
// that reach level 1.
//
// Layout: string 1 ("A") vectors [sorted index][64 B] in HBM, loaded once
// per wave into VGPRs (each wave holds 128 strings: two 64-string units of
// one length each). String 2 ("B") vectors in tiles of kBTile strings of one
// length, stored [tile][plane * 2 + half][string][16 B] so a wave reads one
// 32-string block of one plane with a single conflict-free ds_read_b128 per
// lane. MFMA operand layout (gfx950, checked on the device): lane l supplies
// row/column l % 32 and bytes 16 * (l / 32) .. + 15 of the 32-byte K slice;
// accumulator r of lane l is C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
#include "fh_jw_core.h"

namespace fh {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kBWaves = 8;                 // waves per workgroup
constexpr int kBThreads = kBWaves * 64;
constexpr int kUnitsPerWave = 2;           // 64-string string-1 units per wave
constexpr int kBQCap = 256;                // candidate queue entries per wave

// Per-string count vectors. One wave per string: lane j counts bucket j % 32
// and writes byte j of the string's vector (string 1) or its tile slot
// (string 2). Strings past a tile's count get zero vectors.
__global__ void __launch_bounds__(256) k_planes_a(const uint8_t* __restrict__ chars, const int64_t* __restrict__ off,
                                                  const uint64_t* __restrict__ sorted, int64_t n,
                                                  const uint8_t* __restrict__ amap, int K,
                                                  uint8_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int b = lane & 31;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t s = (((int64_t)blockIdx.x * blockDim.x) >> 6) + (threadIdx.x >> 6); s < n; s += nw) {
    const int64_t e = (int64_t)sorted[s];
    const int64_t o = off[e];
    const int len = (int)min<int64_t>(off[e + 1] - o, kLongLen);
    int cnt = 0;
    for (int i = 0; i < len; ++i) {
      const int code = amap[chars[o + i]];
      cnt += (code < K && (code & 31) == b) ? 1 : 0;
    }
    out[(size_t)s * 64 + lane] = (uint8_t)(lane < 32 ? (cnt >= 1) : max(cnt - 1, 0));
  }
}

__global__ void __launch_bounds__(256) k_planes_b(const uint8_t* __restrict__ chars, const int64_t* __restrict__ off,
                                                  const uint64_t* __restrict__ sorted,
                                                  const TileMeta* __restrict__ tiles, int64_t ntiles,
                                                  const uint8_t* __restrict__ amap, int K,
                                                  uint8_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int b = lane & 31, p = lane >> 5;
  const int64_t total = ntiles * kBTile;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t g = (((int64_t)blockIdx.x * blockDim.x) >> 6) + (threadIdx.x >> 6); g < total; g += nw) {
    const int64_t t = g / kBTile;
    const int sit = (int)(g - t * kBTile);
    const TileMeta tm = tiles[t];
    int cnt = 0;
    if (sit < tm.count) {
      const int64_t e = (int64_t)sorted[tm.s0 + sit];
      const int64_t o = off[e];
      const int len = (int)min<int64_t>(off[e + 1] - o, kLongLen);
      for (int i = 0; i < len; ++i) {
        const int code = amap[chars[o + i]];
        cnt += (code < K && (code & 31) == b) ? 1 : 0;
      }
    }
    const size_t dst = (size_t)t * kBTile * 64 + ((size_t)(p * 2 + (b >> 4)) * kBTile + sit) * 16 + (b & 15);
    out[dst] = (uint8_t)(p == 0 ? (cnt >= 1) : (cnt >= 2));
  }
}

// Candidate entries: (string-1 sorted index << 32) | string-2 sorted index.
// Each wave collects them in an LDS queue and appends full queues to the
// global candidate buffer with one atomic per batch; entries past the
// buffer's capacity are counted but not stored (the host then redoes the
// round in smaller parts).
__device__ __noinline__ void spill_queue(const JwArgs& a, const uint64_t* q, int qn) {
  unsigned long long base = 0;
  if (__lane_id() == 0) base = atomicAdd(a.cand_count, (unsigned long long)qn);
  base = __shfl(base, 0, 64);
  for (int k = (int)__lane_id(); k < qn; k += 64)
    if (base + (unsigned long long)k < (unsigned long long)a.cand_cap) a.cand[base + k] = q[k];
}

__device__ __forceinline__ void push_candidate(const JwArgs& a, uint64_t* q, int& qn, bool pred, uint64_t entry) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  if (qn + 64 > kBQCap) {
    spill_queue(a, q, qn);
    qn = 0;
  }
  if (pred) q[qn + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u))] =
      entry;
  qn = __builtin_amdgcn_readfirstlane(qn + (int)__popcll(mask));
}

__global__ void __launch_bounds__(256) k_eval_candidates(const JwArgs* __restrict__ ap, int64_t n) {
  const JwArgs& a = *ap;
  // Wave-uniform trip count, so every lane reaches the per-wave emission.
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e0 = (int64_t)blockIdx.x * blockDim.x; e0 < n; e0 += stride) {
    const int64_t e = e0 + threadIdx.x;
    int lvl = 0;
    uint64_t x = 0, y = 0;
    if (e < n) {
      const uint64_t v = a.cand[e];
      lvl = candidate_level<false>(a, (int64_t)(v >> 32), (int64_t)(v & 0xFFFFFFFFull), x, y);
    }
    emit_levels_wave(a, x, y, lvl);
  }
}

template <int MODE>
__global__ void __launch_bounds__(kBThreads, 2) k_jw_bound(const JwArgs* __restrict__ ap, int64_t unit0, int64_t nunits,
                                                       int64_t tile0, int64_t ntiles, int tiles_per_chunk) {
  const JwArgs& a = *ap;
  extern __shared__ uint32_t lds_raw[];
  v4i* planes = reinterpret_cast<v4i*>(lds_raw);  // [4][kBTile] x 16 B
  uint64_t* queue = reinterpret_cast<uint64_t*>(lds_raw + 4 * kBTile * 4);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int half = lane >> 5;
  uint64_t* q = queue + wave * kBQCap;
  int qn = 0;

  // This wave's two string-1 units (wave-uniform metadata).
  const int64_t u_first = unit0 + ((int64_t)blockIdx.x * kBWaves + wave) * kUnitsPerWave;
  const int64_t u_end = unit0 + nunits;
  int lenA[kUnitsPerWave], cntA[kUnitsPerWave];
  int64_t s0a[kUnitsPerWave];
#pragma unroll
  for (int k = 0; k < kUnitsPerWave; ++k) {
    WaveMeta wm;
    wm.len = 0;
    wm.count = 0;
    wm.s0 = 0;
    if (u_first + k < u_end) wm = a.waves[u_first + k];
    lenA[k] = __builtin_amdgcn_readfirstlane(wm.len);
    cntA[k] = __builtin_amdgcn_readfirstlane(wm.count);
    s0a[k] = wm.s0;
  }
  // String-1 fragments: block j = rows 32 j .. 32 j + 31 of the wave (unit j / 2).
  v4i af[2 * kUnitsPerWave][2];
#pragma unroll
  for (int j = 0; j < 2 * kUnitsPerWave; ++j) {
    const int k = j >> 1;
    const int r = 32 * (j & 1) + (lane & 31);
    v4i z = {0, 0, 0, 0};
    af[j][0] = z;
    af[j][1] = z;
    if (r < cntA[k]) {
      const v4i* src = reinterpret_cast<const v4i*>(a.planes_a + (size_t)(s0a[k] + r) * 64);
      af[j][0] = src[half];
      af[j][1] = src[2 + half];
    }
  }

  const int64_t t_begin = tile0 + (int64_t)blockIdx.y * tiles_per_chunk;
  const int64_t t_end = min(t_begin + tiles_per_chunk, tile0 + ntiles);
  for (int64_t t = t_begin; t < t_end; ++t) {
    const TileMeta tm = a.tiles[t];
    const int nblk = __builtin_amdgcn_readfirstlane((tm.count + 31) >> 5);
    __syncthreads();
    {
      const v4i* src = reinterpret_cast<const v4i*>(a.planes_b + (size_t)t * kBTile * 64);
      const int used = nblk * 32;
#pragma unroll
      for (int seg = 0; seg < 4; ++seg)
        for (int i = threadIdx.x; i < used; i += kBThreads) planes[seg * kBTile + i] = src[seg * kBTile + i];
    }
    __syncthreads();
    const v4i* cur = planes;
    const int lenB = __builtin_amdgcn_readfirstlane(tm.len);
    const int cntB = __builtin_amdgcn_readfirstlane(tm.count);
    const int64_t s0b = tm.s0;
    int mm[kUnitsPerWave];
    bool act[kUnitsPerWave], diag[kUnitsPerWave];
#pragma unroll
    for (int k = 0; k < kUnitsPerWave; ++k) {
      mm[k] = cntA[k] > 0 ? a.mmin[lenA[k] * (kLongLen + 1) + lenB] : 255;
      act[k] = cntA[k] > 0 && mm[k] <= min(lenA[k], lenB);
      diag[k] = false;
      if (MODE == FH_MODE_DEDUP && act[k]) {
        if (lenA[k] < lenB) act[k] = false;
        if (lenA[k] == lenB) {
          if (s0a[k] + cntA[k] - 1 <= s0b) act[k] = false;
          diag[k] = true;
        }
      }
    }
    bool any = false;
#pragma unroll
    for (int k = 0; k < kUnitsPerWave; ++k) any = any || act[k];
    if (any) {
      for (int cb = 0; cb < nblk; ++cb) {
        const int col = cb * 32 + (lane & 31);
        const v4i b0 = cur[(0 * 2 + half) * kBTile + col];
        const v4i b1 = cur[(1 * 2 + half) * kBTile + col];
        // All eight products first (the MFMA pipeline overlaps them), then
        // the per-block maxima. Block j = rows 32 j .. 32 j + 31 of the wave.
        v16i acc[2 * kUnitsPerWave];
#pragma unroll
        for (int j = 0; j < 2 * kUnitsPerWave; ++j) {
#pragma unroll
          for (int r = 0; r < 16; ++r) acc[j][r] = 0;
          acc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[j][0], b0, acc[j], 0, 0, 0);
          acc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[j][1], b1, acc[j], 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 2 * kUnitsPerWave; ++j) {
          const int k = j >> 1;
          if (!act[k]) continue;
          int mx = acc[j][0];
#pragma unroll
          for (int r = 1; r < 16; ++r) mx = max(mx, acc[j][r]);
          if (__ballot(mx >= mm[k]) == 0) continue;
          // A candidate in this 32 x 32 block: the pairs of every register.
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const bool c = acc[j][r] >= mm[k];
            if (__ballot(c) == 0) continue;
            const int row = 32 * (j & 1) + 8 * (r >> 2) + 4 * half + (r & 3);
            bool pred = c && row < cntA[k] && col < cntB;
            if (MODE == FH_MODE_DEDUP && diag[k]) pred = pred && (s0a[k] + row > s0b + col);
            push_candidate(a, q, qn, pred, ((uint64_t)(s0a[k] + row) << 32) | (uint64_t)(s0b + col));
          }
        }
      }
    }
  }
  if (qn > 0) spill_queue(a, q, qn);
}

}  // namespace

size_t jw_bound_lds_bytes() { return (size_t)4 * kBTile * 16 + (size_t)kBWaves * kBQCap * sizeof(uint64_t); }

int jw_bound_build_planes(const uint8_t* chars_a, const int64_t* off_a, const uint64_t* sorted_a, int64_t n_a,
                          const uint8_t* chars_b, const int64_t* off_b, const uint64_t* sorted_b,
                          const TileMeta* tiles, int64_t ntiles, const uint8_t* amap, int K, uint8_t* planes_a,
                          uint8_t* planes_b, hipStream_t s) {
  if (n_a > 0) {
    const int64_t blocks = std::min<int64_t>(ceil_div(n_a, 4), 1 << 16);
    hipLaunchKernelGGL(k_planes_a, dim3((unsigned)blocks), dim3(256), 0, s, chars_a, off_a, sorted_a, n_a, amap, K,
                       planes_a);
    FH_HIP(hipGetLastError());
  }
  if (ntiles > 0) {
    const int64_t blocks = std::min<int64_t>(ceil_div(ntiles * kBTile, 4), 1 << 16);
    hipLaunchKernelGGL(k_planes_b, dim3((unsigned)blocks), dim3(256), 0, s, chars_b, off_b, sorted_b, tiles, ntiles,
                       amap, K, planes_b);
    FH_HIP(hipGetLastError());
  }
  return FH_OK;
}

int launch_eval_candidates(const JwArgs* args, int64_t n, hipStream_t s) {
  if (n <= 0) return FH_OK;
  const int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 1 << 16);
  hipLaunchKernelGGL(k_eval_candidates, dim3((unsigned)blocks), dim3(256), 0, s, args, n);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int launch_jw_bound(int mode, const JwArgs* args, int64_t u0, int64_t nu, int64_t t0, int64_t nt, hipStream_t s) {
  if (nu <= 0 || nt <= 0) return FH_OK;
  const int64_t gx = ceil_div(nu, (int64_t)kBWaves * kUnitsPerWave);
  // Fewer tiles per workgroup when there are few string-1 units, so that
  // the launch still fills the GPU.
  int tpc = 16;
  while (tpc > 1 && gx * ceil_div(nt, tpc) < 2048) tpc /= 2;
  int64_t gy = ceil_div(nt, tpc);
  while (gy > 65535) {
    tpc *= 2;
    gy = ceil_div(nt, tpc);
  }
  FH_REQUIRE(gx < (1ll << 31), "too many string-1 units");
  const size_t lds = jw_bound_lds_bytes();
  if (mode == FH_MODE_DEDUP)
    hipLaunchKernelGGL(k_jw_bound<FH_MODE_DEDUP>, dim3((unsigned)gx, (unsigned)gy), dim3(kBThreads), lds, s, args,
                       u0, nu, t0, nt, tpc);
  else
    hipLaunchKernelGGL(k_jw_bound<FH_MODE_LINK>, dim3((unsigned)gx, (unsigned)gy), dim3(kBThreads), lds, s, args,
                       u0, nu, t0, nt, tpc);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

}  // namespace fh
