// All-pairs Jaro-Winkler comparison with the string-1 loop inside the wave
// and shared prefixes computed once (the default bit-parallel form).
//
// Replaces, like k_jw_tiles, the reference's comparison kernel
// (faster/comparison.py:133-186, deduplication copy
// faster/deduplication.py:132-196) and its thresholding
// (faster/comparison.py:321-338); the match recurrence, mask table and
// pruning bound are the ones of fh_jw_core.h, the survivors are scored here
// (k_eval_prefix) from both values' row codes.
//
// Transposed roles. A wave still owns 64 string-1 values of one length LA,
// but instead of one per lane it walks them one after another, and its lanes
// cover the staged string-2 tile: lane l holds dwords [4l, 4l + 4) of every
// mask row (the tile is 256 dwords wide), read with one conflict-free
// ds_read_b128 per character, as all lanes read the same row. The string-1
// character, its row and its window are wave-uniform (scalar registers).
//
// Prefix reuse. The recurrence state after the first k characters of string
// 1 depends only on those characters and the tile. The values arrive sorted
// (byte-lexicographically within a length), so consecutive values share
// prefixes: the wave keeps the states after 1..D characters of the previous
// value and restarts each value at min(lcp, D), the length of the prefix it
// shares with the previous one. On the synthetic 1M-value columns this skips
// 36% of the recurrence steps at D = 3, on the reference's example names
// 34-38%.
//
// Every pair's match count is still computed exactly; only recomputation of
// an identical prefix state is skipped, so the levels are those of the
// per-lane form.
#pragma once

#include "fh_jw_core.h"

namespace fh {

// Launch over waves [w0, w0 + nw) of string-1 lengths 1..16 (r0) or 17..32
// (r1) against tiles [t0, t0 + nt) of one packed kind or kW32 (tile width
// 256 dwords).
int launch_jw_prefix_r0(int kind, const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0,
                        int64_t nt, hipStream_t s);
int launch_jw_prefix_r1(int kind, const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0,
                        int64_t nt, hipStream_t s);
// The same for string-1 lengths 33..64 (r2): one rolled pass for every
// length (these values are rare; their steps are not unrolled).
int launch_jw_prefix_r2(int kind, const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0,
                        int64_t nt, hipStream_t s);
// Scores the first n entries of args->cand (prefix-form survivors: both
// values of at most 32 bytes) and appends their level keys.
int launch_eval_prefix(const JwArgs* args, int64_t n, hipStream_t s);

constexpr int kPrefixTileWidth = 256;  // dwords: 64 lanes x 4
constexpr int kPrefixMaxLen = 64;      // longest string-1 value of the prefix form
// Row codes of the string-1 values, four per dword: words 0-7 (characters
// 0-31) of every value at codes_a[sorted index * 8], then words 8-15 of the
// values over 32 bytes at codes_a[n_a * 8 + (sorted index - codes_hi0) * 8],
// so that the common values keep a 32-byte stride.
constexpr int kPrefixCodeWords = 8;
// LDS bytes of one workgroup of the prefix form: the tile (rows of 256
// dwords, no padding: every lane reads the same row) and the survivor queues.
inline size_t jw_prefix_lds_bytes(int krows) {
  return (size_t)krows * kPrefixTileWidth * 4 + (size_t)kWaves * kQCap * sizeof(uint64_t);
}

namespace {

#ifndef FH_JW_PREFIX_DEPTH
#define FH_JW_PREFIX_DEPTH 3
#endif
#ifndef FH_JW_PREFIX_WAVES_PER_EU
#define FH_JW_PREFIX_WAVES_PER_EU 8
#endif
constexpr int kPrefixDepth = FH_JW_PREFIX_DEPTH;
static_assert(kPrefixDepth >= 1 && kPrefixDepth <= 4, "prefix states are found from the first code dword");

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const u32x4 lds_u32x4;

// A wave-uniform 64-bit value in scalar registers.
__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Row `code` of the staged tile at this lane's four dwords.
__device__ __forceinline__ u32x4 prefix_row(uint32_t lane_addr, uint32_t code) {
  return *(lds_u32x4*)(uintptr_t)(lane_addr + code * (uint32_t)(kPrefixTileWidth * 4));
}

// One character step of the recurrence on the lane's four dwords:
//   nc = row | nwin | navail; navail = nc ? navail : nc + ONES (bitwise).
template <uint32_t ONES>
__device__ __forceinline__ void prefix_step(const u32x4 r, uint32_t nwin, const uint32_t (&in)[4],
                                            uint32_t (&out)[4]) {
  uint32_t nc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) nc[u] = or3(r[u], nwin, in[u]);
#pragma unroll
  for (int u = 0; u < 4; ++u) out[u] = select_bits(nc[u], in[u], nc[u] + ONES);
}

// Code of character i of the current string 1 (four codes per dword).
template <int NW>
__device__ __forceinline__ uint32_t prefix_code(const uint32_t (&cw)[NW], int i) {
  return (cw[i >> 2] >> (8 * (i & 3))) & 0xFFu;
}

// Match count of string k of dword `na` (kW32: one string per dword).
template <int KIND>
__device__ __forceinline__ uint32_t prefix_count(uint32_t na, int k) {
  if constexpr (packed_kind(KIND))
    return (uint32_t)lane_count<KIND>(na, k);
  else
    return (uint32_t)__popc(na);
}

// Moves the wave's queue to the candidate buffer a.cand as
// (string-1 sorted index << 32) | string-2 sorted index, with one atomic per
// flush; entries past a.cand_cap are counted only (the caller redoes the
// round with a larger buffer). The candidates are scored afterwards by
// k_eval_prefix: scoring them here, inside the string loop (inlined, or as a
// call), holds the scorer's registers across the whole loop and measured
// slower (611 ms against 541 ms per 1M x 1M step with the call).
__device__ __forceinline__ void prefix_flush(const JwArgs& a, const uint64_t* q, int qn, uint64_t s0a) {
  unsigned long long base = 0;
  if (__lane_id() == 0) base = atomicAdd(a.cand_count, (unsigned long long)qn);
  base = __shfl(base, 0, 64);
  for (int k = (int)__lane_id(); k < qn; k += 64) {
    const uint64_t v = q[k];
    if (base + (unsigned long long)k < (unsigned long long)a.cand_cap)
      a.cand[base + k] = ((s0a + (v & 63ull)) << 32) | (v >> 6);
  }
}

// Queues string-2 sorted index sb against string 1 number s of the wave
// where `pred` holds (all lanes call it); flushes first when full.
__device__ __forceinline__ void prefix_push(const JwArgs& a, PassCtx& c, bool pred, uint64_t sb, int s) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  if (c.qn + 64 > kQCap) {
    prefix_flush(a, c.q, c.qn, (uint64_t)c.s0a);
    c.qn = 0;
  }
  if (pred) c.q[c.qn + __popcll(mask & fhd::lanemask_lt())] = (sb << 6) | (uint64_t)s;
  c.qn += __popcll(mask);
}

// Bit kind_lane_shift(k, l) set for every lane l of a packed kind.
__host__ __device__ constexpr uint32_t kind_rep(int k) {
  uint32_t v = 0;
  for (int l = 0; l < kind_lanes(k); ++l) v |= 1u << kind_lane_shift(k, l);
  return v;
}

// Complemented window of string-1 position I against string-2 values of
// length lenB (<= 32) in every string of a dword of kind KIND, computed with
// scalar ops (the window, a bit-field mask, replicated into the lanes by one
// multiply: the copies do not overlap).
template <int I, int KIND>
__device__ __forceinline__ uint32_t prefix_nwin(int w, int lenB) {
  const int lo = max(0, I - w);
  const int hi = min(lenB - 1, I + w);
  const int width = hi - lo + 1;
  uint32_t win = width <= 0 ? 0u : (width >= 32 ? ~0u : ((1u << width) - 1u)) << lo;
  if constexpr (packed_kind(KIND)) win *= kind_rep(KIND);
  return ~win;
}

template <int I, int KIND>
__device__ __forceinline__ uint32_t nwin_of(int w, int lenB) {
  return __builtin_amdgcn_readfirstlane(prefix_nwin<I, KIND>(w, lenB));
}

// Steps I..LA-1 on na, without read-ahead (the conditional-block form).
template <int I, int LA, int KIND, int NW>
struct PrefixTail {
  __device__ static __forceinline__ void go(const uint32_t (&cw)[NW], int w, int lenB, uint32_t lane_addr,
                                            uint32_t (&na)[4]) {
    if constexpr (I < LA) {
      constexpr uint32_t ONES = packed_kind(KIND) ? kind_ones(KIND) : 1u;
      prefix_step<ONES>(prefix_row(lane_addr, prefix_code(cw, I)), nwin_of<I, KIND>(w, lenB), na, na);
      PrefixTail<I + 1, LA, KIND, NW>::go(cw, w, lenB, lane_addr, na);
    }
  }
};

// The 64 (or cntA) string-1 values of one length LA against the staged tile.
// Lane l holds the row codes of value l in `codes`; each value's are taken
// with v_readlane (scalar loads would share the LDS reads' wait counter).
template <int LA, int KIND, int CW>
__device__ __forceinline__ void prefix_pass(const JwArgs& a, PassCtx& c, int cntA, const uint32_t (&codes)[CW]) {
  constexpr bool PK = packed_kind(KIND);
  constexpr int NL = PK ? kind_lanes(KIND) : 1;
  constexpr uint32_t ONES = PK ? kind_ones(KIND) : 1u;
  constexpr int DE = kPrefixDepth < LA - 1 ? kPrefixDepth : LA - 1;  // prefix states kept
  constexpr int NW = (LA + 3) / 4;
  const int lenB = c.lenB;
  const int w = c.w;
  const int lane = (int)__lane_id();
  const uint32_t lane_addr = lds_addr(c.lds) + (uint32_t)lane * 16u;
  // Bit b = u * NL + k of a lane's masks stands for string-2 tile index
  // NL * (4 * lane + u) + k = 4 * NL * lane + b; bits past the tile's
  // strings are never queued.
  const int jlane = 4 * NL * lane;
  const uint32_t valid = low_bits(c.cntB - jlane);
  const int mm = c.mm;
  uint32_t st[DE > 0 ? DE : 1][4];
  const uint32_t zero[4] = {0u, 0u, 0u, 0u};
  static_assert(NW <= CW, "codes of every character");
  uint32_t prev0 = 0;
#pragma unroll 1
  for (int s = 0; s < cntA; ++s) {
    uint32_t cw[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) cw[k] = __builtin_amdgcn_readlane(codes[k], s);
    // Restart at the shared prefix with the previous value, at most DE.
    int j0 = 0;
    if (s > 0) {
      const uint32_t x = cw[0] ^ prev0;
      j0 = x ? (int)(__builtin_ctz(x) >> 3) : 4;
      j0 = j0 < DE ? j0 : DE;
    }
    prev0 = cw[0];
    // Steps j0..DE-1 rebuild the prefix states the new value does not share
    // (each step reads its row and waits: read-ahead across these
    // conditional steps and the tail measured 2% slower, as did one copy of
    // the chain per restart depth), then steps DE..LA-1 run on na.
    uint32_t na[4];
    if constexpr (DE == 0) {
      prefix_step<ONES>(prefix_row(lane_addr, prefix_code(cw, 0)), nwin_of<0, KIND>(w, lenB), zero, na);
    } else {
      if (j0 == 0)
        prefix_step<ONES>(prefix_row(lane_addr, prefix_code(cw, 0)), nwin_of<0, KIND>(w, lenB), zero, st[0]);
      if constexpr (DE >= 2) {
        if (j0 <= 1)
          prefix_step<ONES>(prefix_row(lane_addr, prefix_code(cw, 1)), nwin_of<1, KIND>(w, lenB), st[0], st[1]);
      }
      if constexpr (DE >= 3) {
        if (j0 <= 2)
          prefix_step<ONES>(prefix_row(lane_addr, prefix_code(cw, 2)), nwin_of<2, KIND>(w, lenB), st[1], st[2]);
      }
      if constexpr (DE >= 4) {
        if (j0 <= 3)
          prefix_step<ONES>(prefix_row(lane_addr, prefix_code(cw, 3)), nwin_of<3, KIND>(w, lenB), st[2], st[3]);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) na[u] = st[DE - 1][u];
    }
    PrefixTail<(DE > 0 ? DE : 1), LA, KIND, NW>::go(cw, w, lenB, lane_addr, na);
    // Survivor test: the largest match count over the lane's 4 * NL pairs.
    uint32_t mx = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < NL; ++k) mx = max(mx, prefix_count<KIND>(na[u], k));
    if (__ballot((int)mx >= mm) == 0) continue;
    // Rare path: the pairs reaching mm, then the index checks.
    uint32_t pass = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < NL; ++k) pass |= (uint32_t)((int)prefix_count<KIND>(na[u], k) >= mm) << (u * NL + k);
    pass &= valid;
    // Deduplication triangle: string-2 sorted index below string 1's.
    if (c.diag) pass &= low_bits(c.s0a + s - c.s0b - jlane);
#pragma unroll 1
    while (__ballot(pass != 0u) != 0) {
      const bool pred = pass != 0u;
      const int b = pred ? __builtin_ctz(pass) : 0;
      prefix_push(a, c, pred, (uint64_t)(c.s0b + jlane + b), s);
      pass &= pass - 1u;
    }
  }
}

// Complemented window of string-1 position i (run-time form of prefix_nwin).
template <int KIND>
__device__ __forceinline__ uint32_t nwin_rt(int i, int w, int lenB) {
  const int lo = max(0, i - w);
  const int hi = min(lenB - 1, i + w);
  const int width = hi - lo + 1;
  uint32_t win = width <= 0 ? 0u : (width >= 32 ? ~0u : ((1u << width) - 1u)) << lo;
  if constexpr (packed_kind(KIND)) win *= kind_rep(KIND);
  return __builtin_amdgcn_readfirstlane(~win);
}

// prefix_pass for string-1 values of 33..64 bytes (rare), rolled: each
// step takes its row code from the value's code words with a scalar load
// and computes its window on the scalar unit.
template <int KIND>
__device__ __forceinline__ void prefix_pass_long(const JwArgs& a, PassCtx& c, int cntA, int LA) {
  constexpr bool PK = packed_kind(KIND);
  constexpr int NL = PK ? kind_lanes(KIND) : 1;
  constexpr uint32_t ONES = PK ? kind_ones(KIND) : 1u;
  constexpr int DE = kPrefixDepth;  // LA > 32 > DE
  const int lenB = c.lenB;
  const int w = c.w;
  const int lane = (int)__lane_id();
  const uint32_t lane_addr = lds_addr(c.lds) + (uint32_t)lane * 16u;
  const int jlane = 4 * NL * lane;
  const uint32_t valid = low_bits(c.cntB - jlane);
  const int mm = c.mm;
  uint32_t st[DE][4];
  const uint32_t zero[4] = {0u, 0u, 0u, 0u};
  uint32_t prev0 = 0;
#pragma unroll 1
  for (int s = 0; s < cntA; ++s) {
    const uint32_t* cp = a.codes_a + (size_t)(c.s0a + s) * kPrefixCodeWords;
    const uint32_t* cph = a.codes_a + a.codes_hi_base + (size_t)(c.s0a + s - a.codes_hi0) * kPrefixCodeWords - 8;
    const uint32_t w0 = __builtin_amdgcn_readfirstlane(cp[0]);
    int j0 = 0;
    if (s > 0) {
      const uint32_t x = w0 ^ prev0;
      j0 = x ? (int)(__builtin_ctz(x) >> 3) : 4;
      j0 = j0 < DE ? j0 : DE;
    }
    prev0 = w0;
    if (j0 == 0) prefix_step<ONES>(prefix_row(lane_addr, w0 & 0xFFu), nwin_rt<KIND>(0, w, lenB), zero, st[0]);
    if (j0 <= 1)
      prefix_step<ONES>(prefix_row(lane_addr, (w0 >> 8) & 0xFFu), nwin_rt<KIND>(1, w, lenB), st[0], st[1]);
    if (j0 <= 2)
      prefix_step<ONES>(prefix_row(lane_addr, (w0 >> 16) & 0xFFu), nwin_rt<KIND>(2, w, lenB), st[1], st[2]);
    uint32_t na[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) na[u] = st[DE - 1][u];
#pragma unroll 1
    for (int i = DE; i < LA; ++i) {
      const uint32_t wd = __builtin_amdgcn_readfirstlane(i < 32 ? cp[i >> 2] : cph[i >> 2]);
      prefix_step<ONES>(prefix_row(lane_addr, (wd >> (8 * (i & 3))) & 0xFFu), nwin_rt<KIND>(i, w, lenB), na, na);
    }
    uint32_t mx = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < NL; ++k) mx = max(mx, prefix_count<KIND>(na[u], k));
    if (__ballot((int)mx >= mm) == 0) continue;
    uint32_t pass = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < NL; ++k) pass |= (uint32_t)((int)prefix_count<KIND>(na[u], k) >= mm) << (u * NL + k);
    pass &= valid;
    if (c.diag) pass &= low_bits(c.s0a + s - c.s0b - jlane);
#pragma unroll 1
    while (__ballot(pass != 0u) != 0) {
      const bool pred = pass != 0u;
      const int b = pred ? __builtin_ctz(pass) : 0;
      prefix_push(a, c, pred, (uint64_t)(c.s0b + jlane + b), s);
      pass &= pass - 1u;
    }
  }
}

template <int LO, int HI, int KIND, int CW>
struct PrefixDispatch {
  __device__ static __forceinline__ void run(int lenA, const JwArgs& a, PassCtx& c, int cntA,
                                             const uint32_t (&codes)[CW]) {
    if (lenA == LO)
      prefix_pass<LO, KIND, CW>(a, c, cntA, codes);
    else
      PrefixDispatch<LO + 1, HI, KIND, CW>::run(lenA, a, c, cntA, codes);
  }
};
template <int HI, int KIND, int CW>
struct PrefixDispatch<HI, HI, KIND, CW> {
  __device__ static __forceinline__ void run(int lenA, const JwArgs& a, PassCtx& c, int cntA,
                                             const uint32_t (&codes)[CW]) {
    if (lenA == HI) prefix_pass<HI, KIND, CW>(a, c, cntA, codes);
  }
};

// This workgroup's waves [wave_first, wave_first + kWaves) ∩ [.., wave_end)
// against tiles [t_begin, t_end) of kind KIND.
template <int LMIN, int LMAX, int KIND, bool LONG>
__device__ __forceinline__ void run_item_prefix(const JwArgs& a, uint32_t* lds, int64_t wave_first, int64_t wave_end,
                                                int64_t t_begin, int64_t t_end) {
  const int krows = a.krows;
  uint64_t* queue = reinterpret_cast<uint64_t*>(lds + (size_t)a.krows_max * kPrefixTileWidth);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = __lane_id();
  const int64_t wid = wave_first + wave;
  const bool wave_ok = wid < wave_end;
  WaveMeta wm;
  wm.len = 0;
  wm.count = 0;
  wm.s0 = 0;
  if (wave_ok) wm = a.waves[wid];
  const int lenA = __builtin_amdgcn_readfirstlane(wm.len);
  const int cntA = __builtin_amdgcn_readfirstlane(wm.count);

  PassCtx c;
  c.lds = lds;
  c.q = queue + wave * kQCap;
  c.qn = 0;
  c.lane_ok = lane < cntA;
  c.s0a = uniform_i64(wm.s0);

  // Row codes of string 1 number `lane` (four per dword).
  constexpr int CW = LONG ? 1 : (LMAX + 3) / 4;
  uint32_t codes[CW];
#pragma unroll
  for (int k = 0; k < CW; ++k)
    codes[k] = (LONG || !c.lane_ok) ? 0u : a.codes_a[(size_t)(c.s0a + lane) * kPrefixCodeWords + k];

  for (int64_t t = t_begin; t < t_end; ++t) {
    const TileMeta tm = a.tiles[t];
    __syncthreads();
    {
      const u32x4* src = reinterpret_cast<const u32x4*>(a.table + (size_t)t * krows * kPrefixTileWidth);
      u32x4* dst = reinterpret_cast<u32x4*>(lds);
      const int total = krows * (kPrefixTileWidth / 4);
      for (int idx = threadIdx.x; idx < total; idx += kThreads) dst[idx] = src[idx];
    }
    __syncthreads();
    if (!wave_ok) continue;
    c.lenB = __builtin_amdgcn_readfirstlane(tm.len);
    c.cntB = __builtin_amdgcn_readfirstlane(tm.count);
    c.s0b = uniform_i64(tm.s0);
    c.diag = false;
    if (a.mode == FH_MODE_DEDUP) {
      if (lenA < c.lenB) continue;
      if (lenA == c.lenB) {
        if (c.s0a + cntA - 1 <= c.s0b) continue;  // no string-1 index above any string-2 index
        c.diag = true;
      }
    }
    c.mm = a.mmin[lenA * (kLongLen + 1) + c.lenB];
    if (c.mm > min(lenA, c.lenB)) continue;
    c.w = max(lenA, c.lenB) / 2 - 1;
    if constexpr (LONG)
      prefix_pass_long<KIND>(a, c, cntA, lenA);
    else
      PrefixDispatch<LMIN, LMAX, KIND, CW>::run(lenA, a, c, cntA, codes);
  }
  if (wave_ok && c.qn > 0) prefix_flush(a, c.q, c.qn, (uint64_t)wm.s0);
}

// Row codes of the string-1 values [s_begin, s_end) in sorted order (every
// value of at most kPrefixMaxLen bytes), four per dword; positions past the value's
// length hold the all-blocked row krows - 1.
__global__ void __launch_bounds__(256) k_prefix_codes(const uint8_t* __restrict__ chars,
                                                      const int64_t* __restrict__ off,
                                                      const uint64_t* __restrict__ sorted, int64_t s_begin,
                                                      int64_t s_end, const uint8_t* __restrict__ amap, int krows,
                                                      int64_t hi0, int64_t hi_base, uint32_t* __restrict__ codes) {
  const int64_t n = (s_end - s_begin) * 16;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t si = s_begin + g / 16;
    const int q = (int)(g % 16);
    if (q >= kPrefixCodeWords && si < hi0) continue;  // values of up to 32 bytes: words 0-7 only
    const int64_t e = (int64_t)sorted[si];
    const int64_t o = off[e];
    const int len = (int)(off[e + 1] - o);
    uint32_t w = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = 4 * q + k;
      const uint32_t code = i < len ? (uint32_t)amap[chars[o + i]] : (uint32_t)(krows - 1);
      w |= code << (8 * k);
    }
    if (q < kPrefixCodeWords)
      codes[si * kPrefixCodeWords + q] = w;
    else
      codes[hi_base + (si - hi0) * kPrefixCodeWords + (q - kPrefixCodeWords)] = w;
  }
}

// Largest class c with start[c] <= x (the class of sorted index x; empty
// classes share their start with the next one).
__device__ __forceinline__ int class_of(const int64_t* start, int64_t x) {
  int lo = 0, hi = kNumClasses - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Work accounting: per string-1 length, the recurrence steps one tile pass
// skips, summed over the values [s_begin, s_end): for every value but the
// first of its wave (waves start every 64 values from its length's first
// sorted index), min(shared prefix of row codes with the previous value,
// prefix states kept), as the pass computes it.
__global__ void __launch_bounds__(256) k_prefix_restarts(const uint32_t* __restrict__ codes,
                                                         const int64_t* __restrict__ cls_start_a, int64_t s_begin,
                                                         int64_t s_end, unsigned long long* __restrict__ skipped) {
  __shared__ int64_t starts[kNumClasses];
  __shared__ unsigned long long h[kPrefixMaxLen + 1];
  for (int i = threadIdx.x; i < kNumClasses; i += blockDim.x) starts[i] = cls_start_a[i];
  for (int i = threadIdx.x; i <= kPrefixMaxLen; i += blockDim.x) h[i] = 0;
  __syncthreads();
  for (int64_t si = s_begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; si < s_end;
       si += (int64_t)gridDim.x * blockDim.x) {
    const int la = class_of(starts, si);
    if (la < 1 || la > kPrefixMaxLen || (si - starts[la]) % 64 == 0) continue;
    const uint32_t x = codes[si * kPrefixCodeWords] ^ codes[(si - 1) * kPrefixCodeWords];
    const int depth = min(kPrefixDepth, la - 1);
    const int j0 = min(x ? (int)(__builtin_ctz(x) >> 3) : 4, depth);
    if (j0 > 0) atomicAdd(&h[la], (unsigned long long)j0);
  }
  __syncthreads();
  for (int i = threadIdx.x; i <= kPrefixMaxLen; i += blockDim.x)
    if (h[i]) atomicAdd(&skipped[i], h[i]);
}

// Bit k (k = 0..3) set where byte k of `word` equals the byte that `rep`
// repeats in all four bytes: bit 7 of each byte of y is set where x's byte is
// nonzero (the low seven bits carry into it, or it was set). The zero flags
// of bytes 0-2 are gathered into bits 21-23 by one 24-bit multiply (full
// rate; its partial products do not overlap), byte 3's flag is bit 31.
__device__ __forceinline__ uint32_t equal_bytes(uint32_t word, uint32_t rep) {
  const uint32_t x = word ^ rep;
  const uint32_t y = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;
  const uint32_t z = ~y & 0x80808080u;
  const uint32_t lo3 = (__umul24(z >> 7, 0x00204081u) >> 21) & 0x7u;
  return lo3 | ((z >> 28) & 0x8u);
}

// Level of string 1 (row codes c1, length la) against string 2 (row codes
// c2, length lb), both of at most 32 bytes: the reference's greedy match,
// transpositions and prefix (faster/comparison.py:48-123) on per-character
// position masks built in registers: row i = the positions of string 2 whose
// code equals string 1's code i. Codes are the alphabet map of bytes present
// on both sides; every other byte and every position past a value's end
// holds the last code, which matches nothing.
// `lmax` is wave-uniform and at least la and lb of every lane that calls:
// the unrolled steps past it are skipped by scalar branches.
__device__ __forceinline__ int codes_level(const JwArgs& a, const uint32_t (&c1)[8], int la, const uint32_t (&c2)[8],
                                           int lb, int lmax) {
  const int mm = a.mmin[la * (kLongLen + 1) + lb];
  const uint32_t none = (uint32_t)(a.krows - 1);
  const uint32_t lmask = lb >= 32 ? ~0u : ((1u << lb) - 1u);
  const int nw2 = (lb + 3) >> 2;
  uint32_t row[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    row[i] = 0;
    if (i < lmax && i < la) {
      const uint32_t code = (c1[i >> 2] >> (8 * (i & 3))) & 0xFFu;
      if (code != none) {
        const uint32_t rep = __builtin_amdgcn_perm(code, code, 0u);  // byte 0 of code in every byte
        uint32_t r = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (4 * k < lmax && k < nw2) r |= equal_bytes(c2[k], rep) << (4 * k);
        row[i] = r & lmask;
      }
    }
  }
  const int w = max(la, lb) / 2 - 1;
  uint32_t used = 0, flag1 = 0;
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    if (i < lmax && i < la) {
      const uint32_t avail = row[i] & (uint32_t)window_bits(i, w, lb) & ~used;
      if (avail) {
        used |= avail & (~avail + 1u);
        flag1 |= 1u << i;
      }
    }
  }
  const int m = __popc(used);
  if (m < mm) return 0;
  if (m == 0) return level_of(0.0f, a.lo, a.hi);  // only when a score of 0 reaches level 1
  int tt = 0;
  uint32_t u = used;
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    if (i < lmax && ((flag1 >> i) & 1u)) {
      const int j = __builtin_ctz(u);
      u &= u - 1u;
      if (!((row[i] >> j) & 1u)) ++tt;
    }
  }
  int pre = 0;
  const int lim = min(min(la, lb), 4);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i < lim && pre == i && ((row[i] >> i) & 1u)) ++pre;
  }
  return level_of(jw_formula(la, lb, m, tt, pre, a.p), a.lo, a.hi);
}

// Largest v over the wave's lanes for 0 <= v <= 63 (five ballots; every
// lane of the wave must call it).
__device__ __forceinline__ int wave_max_small(int v) {
  int m = 0;
#pragma unroll
  for (int b = 32; b >= 1; b >>= 1)
    if (__ballot(v >= m + b) != 0) m += b;
  return __builtin_amdgcn_readfirstlane(m);
}

// Row codes of sorted value si (words 0-7) of a code array.
__device__ __forceinline__ void load_codes(const uint32_t* codes, int64_t si, uint32_t (&cw)[8]) {
  const uint32_t* cp = codes + (size_t)si * kPrefixCodeWords;
#pragma unroll
  for (int k = 0; k < 8; ++k) cw[k] = cp[k];
}

// Byte order of two values of at most 32 bytes from their row codes: in
// deduplication both sides are one set, so every byte present has its own
// code, assigned in byte order; the first differing position decides, then
// the length (lex_greater's order).
__device__ __forceinline__ bool codes_greater(const uint32_t (&c1)[8], int l1, const uint32_t (&c2)[8], int l2) {
  int d = 32;
#pragma unroll
  for (int k = 7; k >= 0; --k) {
    const uint32_t x = c1[k] ^ c2[k];
    if (x) d = 4 * k + (int)(__builtin_ctz(x) >> 3);
  }
  if (d < min(l1, l2)) {
    const int k = d >> 2, sh = 8 * (d & 3);
    uint32_t b1 = 0, b2 = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j == k) {
        b1 = (c1[j] >> sh) & 0xFFu;
        b2 = (c2[j] >> sh) & 0xFFu;
      }
    return b1 > b2;
  }
  return l1 > l2;
}

// Level of survivor (sa, sb) of the prefix form, string 1 = sa, by
// codes_level; string-1 values over 32 bytes go to candidate_level.
// Deduplication pairs whose byte order is reversed are scored with the roles
// exchanged (string 1 = sb), as the reference scores the larger value as
// string 1.
// la, lb: the length classes of sa and sb; lmax: wave-uniform, at least
// la and lb of every lane with la <= 32.
__device__ __forceinline__ int prefix_level(const JwArgs& a, int64_t sa, int64_t sb, int la, int lb, int lmax,
                                            uint64_t& x, uint64_t& y) {
  const uint64_t ea = a.sorted_a[sa], eb = a.sorted_b[sb];
  x = ea;
  y = eb;
  if (la > 32) return candidate_level(a, sa, sb, x, y);
  uint32_t cw[8], cw2[8];
  load_codes(a.codes_a, sa, cw);
  load_codes(a.codes_b, sb, cw2);
  if (a.mode == FH_MODE_DEDUP && !codes_greater(cw, la, cw2, lb)) {
    x = eb;
    y = ea;
    return codes_level(a, cw2, lb, cw, la, lmax);
  }
  return codes_level(a, cw, la, cw2, lb, lmax);
}

// Appends the wave's buffered level keys to a.hits with one atomic on the
// hit counter (entries past a.cap are counted only).
__device__ __forceinline__ void eval_flush(const JwArgs& a, const uint64_t* buf, int n) {
  if (n == 0) return;
  unsigned long long base = 0;
  if (__lane_id() == 0) base = atomicAdd(a.counter, (unsigned long long)n);
  base = __shfl(base, 0, 64);
  for (int k = (int)__lane_id(); k < n; k += 64)
    if ((int64_t)(base + (unsigned long long)k) < a.cap) a.hits[base + k] = buf[k];
}

// Level keys leave through a per-wave LDS buffer of kEvalBuf entries, one
// atomic on the shared hit counter per flush: one returning atomic per wave
// and iteration on that single address serialised the scorer.
constexpr int kEvalBuf = 256;

#ifndef FH_EVAL_WAVES_PER_EU
#define FH_EVAL_WAVES_PER_EU 8
#endif

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FH_EVAL_WAVES_PER_EU))) k_eval_prefix(const JwArgs* __restrict__ ap, int64_t n) {
  const JwArgs& a = *ap;
  __shared__ int64_t starts[2][kNumClasses];
  __shared__ uint64_t hbuf[256 / 64][kEvalBuf];
  for (int i = threadIdx.x; i < kNumClasses; i += blockDim.x) {
    starts[0][i] = a.cls_start_a[i];
    starts[1][i] = a.cls_start[i];
  }
  __syncthreads();
  uint64_t* buf = hbuf[threadIdx.x >> 6];
  int nbuf = 0;  // wave-uniform
  const uint64_t lt = fhd::lanemask_lt();
  // Wave-uniform trip count, so every lane reaches the per-wave emission.
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e0 = (int64_t)blockIdx.x * blockDim.x; e0 < n; e0 += stride) {
    const int64_t e = e0 + threadIdx.x;
    int lvl = 0;
    uint64_t x = 0, y = 0;
    int64_t sa = 0, sb = 0;
    int la = 0, lb = 0;
    if (e < n) {
      const uint64_t v = a.cand[e];
      sa = (int64_t)(v >> 32);
      sb = (int64_t)(v & 0xFFFFFFFFull);
      la = class_of(starts[0], sa);
      lb = class_of(starts[1], sb);
    }
    // Longest value of the wave's pairs scored from row codes (string 1 of
    // at most 32 bytes; string 2 always is).
    const int lmax = wave_max_small(la <= 32 ? max(la, lb) : 0);
    if (e < n) lvl = prefix_level(a, sa, sb, la, lb, lmax, x, y);
    if (a.hkeys) {
      emit_level(a, x, y, lvl);
      continue;
    }
    const uint64_t mask = __ballot(lvl > 0);
    if (mask == 0) continue;
    const int cnt = __popcll(mask);
    if (nbuf + cnt > kEvalBuf) {
      eval_flush(a, buf, nbuf);
      nbuf = 0;
    }
    if (lvl > 0) buf[nbuf + __popcll(mask & lt)] = ((uint64_t)x * (uint64_t)a.n_b + y) * 4ull + (uint64_t)lvl;
    nbuf += cnt;
  }
  eval_flush(a, buf, nbuf);
}

template <int LMIN, int LMAX, int KIND>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(FH_JW_PREFIX_WAVES_PER_EU)))
k_jw_prefix(const JwArgs* __restrict__ ap, int64_t wave0, int64_t nwaves, int64_t tile0, int64_t ntiles,
            int tiles_per_chunk) {
  extern __shared__ uint32_t lds[];
  const int64_t t_begin = tile0 + (int64_t)blockIdx.y * tiles_per_chunk;
  const int64_t t_end = min(t_begin + tiles_per_chunk, tile0 + ntiles);
  run_item_prefix<LMIN, LMAX, KIND, (LMIN > 32)>(*ap, lds, wave0 + (int64_t)blockIdx.x * kWaves, wave0 + nwaves,
                                                 t_begin, t_end);
}

template <int LMIN, int LMAX, int KIND>
static int launch_prefix(const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0, int64_t nt,
                         hipStream_t s) {
  if (nw <= 0 || nt <= 0) return FH_OK;
  const int64_t gx = ceil_div(nw, kWaves);
  int tpc = 16;
  while (tpc > 1 && gx * ceil_div(nt, tpc) < 2048) tpc /= 2;
  int64_t gy = ceil_div(nt, tpc);
  while (gy > 65535) {
    tpc *= 2;
    gy = ceil_div(nt, tpc);
  }
  FH_REQUIRE(gx < (1ll << 31), "too many string-1 waves");
  hipLaunchKernelGGL((k_jw_prefix<LMIN, LMAX, KIND>), dim3((unsigned)gx, (unsigned)gy), dim3(kThreads), lds_bytes, s,
                     args, w0, nw, t0, nt, tpc);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

template <int LMIN, int LMAX>
int launch_prefix_kind(int kind, const JwArgs* args, size_t lds, int64_t w0, int64_t nw, int64_t t0, int64_t nt,
                       hipStream_t s) {
  switch (kind) {
    case kP5: return launch_prefix<LMIN, LMAX, kP5>(args, lds, w0, nw, t0, nt, s);
    case kP4: return launch_prefix<LMIN, LMAX, kP4>(args, lds, w0, nw, t0, nt, s);
    case kP3: return launch_prefix<LMIN, LMAX, kP3>(args, lds, w0, nw, t0, nt, s);
    case kP2: return launch_prefix<LMIN, LMAX, kP2>(args, lds, w0, nw, t0, nt, s);
    case kW32: return launch_prefix<LMIN, LMAX, kW32>(args, lds, w0, nw, t0, nt, s);
    default: set_error("launch_prefix_kind: kind %d has no prefix form", kind); return FH_ERR_INTERNAL;
  }
}

}  // namespace
}  // namespace fh
