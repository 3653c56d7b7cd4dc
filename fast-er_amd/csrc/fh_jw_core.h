// Device code of the all-pairs Jaro-Winkler comparison, shared by the
// translation units that instantiate it (fh_jw_r0/r1/r2.hip: the plain
// comparison per string-1 length range; fh_jw_blocked.hip: the persistent
// blocked form; fh_jw.hip: plan and entry point).
//
// All-pairs Jaro-Winkler levels on gfx950.
//
// Replaces the reference's comparison kernel (faster/comparison.py:133-186,
// deduplication copy faster/deduplication.py:132-196) and its thresholding
// and compaction (faster/comparison.py:321-338).
//
// Design (one GPU, 64-wide wavefronts, VALU-bound):
//   * Strings of 1..64 bytes are sorted by length. A wavefront holds 64
//     strings of one length ("string 1", one per lane); every length-derived
//     quantity of a pair (match window, pruning bound) is then uniform across
//     the wave and lives in scalar registers.
//   * Strings of side B ("string 2") are grouped into tiles of one length and
//     pre-encoded as complemented per-character position masks:
//     nmask[c][b] has bit j clear iff byte j of string b has alphabet code c.
//     A workgroup stages one tile in LDS (row stride = tile width + 1 dwords,
//     so the 64 lanes' rows fall in distinct banks) and streams it against its
//     8 waves of string 1.
//   * The reference's greedy first-unused match (faster/comparison.py:48-64)
//     becomes, per character of string 1, three VALU ops on a mask of still
//     unmatched positions of string 2 ("navail", complemented):
//         nc = nmask | nwin_i | navail        (v_bitop3: candidates = zeros)
//         t  = nc + 1                          (lowest candidate -> 1)
//         navail = (nc & navail) | (~nc & t)   (v_bitop3: mark it used)
//     Short string-2 masks are packed several to a dword, each lane's
//     positions preceded by a guard bit that is always set in nc: lanes of 8
//     bits (strings of <= 7 bytes, 4 per dword), 11 bits (<= 10 bytes, 3 per
//     dword, no guard under the lowest lane) or 16 bits (<= 15 bytes, 2 per
//     dword). One 32-bit add of 1 per lane then advances every lane at once:
//     a lane whose bits are all set carries into the next lane's guard, which
//     absorbs it (the guard is set in nc, so the select never copies it), and
//     no carry reaches a position bit of another lane.
//   * The match count m = popcount of used positions is an exact count for the
//     greedy match. A host-built table gives, for every (len1, len2), the
//     smallest m for which any transposition count and prefix can reach level
//     1 in the reference's fp32 arithmetic; pairs below it are level 0 and
//     cost nothing more. Survivors are queued per wave in LDS and evaluated
//     with all 64 lanes by an exact restatement of the reference function.
//   * Pairs with an empty string (only when a score of 0 reaches level 1) or a
//     string longer than 64 bytes go through a generic kernel.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "fh_common.h"

namespace fh {

constexpr int kWaves = 8;             // waves per workgroup
constexpr int kThreads = kWaves * 64;
#ifndef FH_JW_KU
#define FH_JW_KU 8
#endif
// Waves per SIMD the comparison kernel is compiled for. A workgroup is
// 8 waves (2 per SIMD), so only even values change residency; 6 (3 workgroups
// per CU, 80 VGPRs) measured fastest on every workload, ahead of 4 (the
// compiler's own choice) and 8 (heavy spills).
#ifndef FH_JW_WAVES_PER_EU
#define FH_JW_WAVES_PER_EU 6
#endif
constexpr int kU = FH_JW_KU;          // dword columns per inner step (4 or 8)
constexpr int kULog2 = kU == 8 ? 3 : 2;
static_assert(kU == 4 || kU == 8, "kU must be 4 or 8");
#ifndef FH_JW_PREFETCH
#define FH_JW_PREFETCH 1
#endif
constexpr int kQCap = 128;            // survivor queue entries per wave
constexpr int kLongLen = 64;          // longest string on the bit-parallel path
constexpr int kMaxGenericLen = 1024;  // longest string scored with its flags in registers
constexpr int kClassLong = kLongLen + 1;
constexpr int kNumClasses = kLongLen + 2;  // 0 = empty, 1..64, 65 = long

// String-2 mask layouts ("kinds"), chosen by string-2 length. The packed
// kinds hold L strings per dword in lanes of S bits: lane l's positions
// start at bit kind_lane_base + l * S, and the bit below them is the lane's
// guard (always set in nc; it absorbs a carry out of the lane below and is
// where the lane's +1 enters). Lane 0 of kP3 has no guard: nothing carries
// into it, and its +1 enters at position 0 itself, which lets three strings
// of 10 bytes share a dword.
//   kind  lengths  L  S   guard of lane 0
//   kP5   1-5      5  6   yes   (prefix form only; kP4 elsewhere)
//   kP4   1-7      4  8   yes
//   kP3   8-10     3  11  no
//   kP2   11-15    2  16  yes
// The per-lane form keeps one survivor bit per (dword, lane) of an 8-dword
// step in a 32-bit mask, which five lanes would overflow, so kP5 tiles are
// planned only for the prefix form (kind_for).
enum Kind { kP4 = 0, kP3 = 1, kP2 = 2, kW32 = 3, kW64 = 4, kP5 = 5 };
constexpr int kNumKinds = 6;

__host__ __device__ constexpr bool packed_kind(int k) { return k <= kP2 || k == kP5; }
// Strings per dword of a packed kind, the stride of its lanes and the bit of
// lane 0's first position.
__host__ __device__ constexpr int kind_lanes(int k) { return k == kP5 ? 5 : (k == kP4 ? 4 : (k == kP3 ? 3 : 2)); }
__host__ __device__ constexpr int kind_lane_bits(int k) { return k == kP5 ? 6 : (k == kP4 ? 8 : (k == kP3 ? 11 : 16)); }
__host__ __device__ constexpr int kind_lane_base(int k) { return k == kP3 ? 0 : 1; }
// First bit of string position 0 in lane l.
__host__ __device__ constexpr int kind_lane_shift(int k, int l) { return kind_lane_base(k) + l * kind_lane_bits(k); }
// The bit each lane's +1 enters at: its guard, or position 0 of a lane with none.
__host__ __device__ constexpr uint32_t kind_ones(int k) {
  uint32_t v = 0;
  for (int l = 0; l < kind_lanes(k); ++l) v |= 1u << (kind_lane_shift(k, l) > 0 ? kind_lane_shift(k, l) - 1 : 0);
  return v;
}
__host__ __device__ constexpr int kind_of_len(int len) {
  return len <= 7 ? kP4 : (len <= 10 ? kP3 : (len <= 15 ? kP2 : (len <= 32 ? kW32 : kW64)));
}
// The kind of a string-2 length in a plan that may use kP5 (p5: the prefix
// form's plans).
__host__ __device__ constexpr int kind_for(int len, bool p5) { return p5 && len <= 5 ? kP5 : kind_of_len(len); }
// Every lane's positions fit the dword, and the lanes fill it.
static_assert(kind_ones(kP4) == 0x01010101u && kind_lane_shift(kP4, 3) + 7 == 32, "kP4 layout");
static_assert(kind_ones(kP3) == 0x00200401u && kind_lane_shift(kP3, 2) + 10 == 32, "kP3 layout");
static_assert(kind_ones(kP2) == 0x00010001u && kind_lane_shift(kP2, 1) + 15 == 32, "kP2 layout");
static_assert(kind_ones(kP5) == 0x01041041u && kind_lane_shift(kP5, 4) + 5 == 30, "kP5 layout");
// Strings of one tile of `tw` dword columns.
__host__ __device__ constexpr int kind_tile_strings(int k, int tw) {
  return packed_kind(k) ? kind_lanes(k) * tw : (k == kW32 ? tw : tw / 2);
}
// log2 of the dword columns of a tile of `count` strings that the passes
// read, rounded up to a power of two >= kU and capped at the tile width.
__host__ __device__ inline int tile_cols_log2(int kind, int count, int tw) {
  const int used = packed_kind(kind) ? (count + kind_lanes(kind) - 1) / kind_lanes(kind)
                                     : (kind == kW32 ? count : 2 * count);
  int sh = kULog2;
  while ((1 << sh) < used && (1 << sh) < tw) ++sh;
  return sh;
}

// Algorithmic VALU lane-ops per pair of string-1 length la: three ops per
// character and dword (or3, add, select) plus the survivor test (extract and
// count each packed string, one compare per dword), divided by the strings
// sharing the dword; 64-bit masks take two dwords per string.
inline double kind_pair_ops(int k, int la) {
  if (packed_kind(k)) return (3.0 * la + 2.0 * kind_lanes(k) + 1.0) / kind_lanes(k);
  return k == kW32 ? 3.0 * la + 2.0 : 6.0 * la + 4.0;
}

struct WaveMeta {
  int32_t len;    // string-1 length of every lane
  int32_t count;  // valid lanes (1..64)
  int64_t s0;     // first sorted index
};

struct TileMeta {
  int32_t len;    // string-2 length of every string in the tile
  int32_t count;  // strings in the tile
  int64_t s0;     // first sorted index
  int32_t kind;   // mask layout of the tile's strings (Kind)
};

struct JwArgs {
  const uint8_t* chars_a;
  const int64_t* off_a;
  const uint64_t* sorted_a;
  const uint8_t* chars_b;
  const int64_t* off_b;
  const uint64_t* sorted_b;
  const WaveMeta* waves;
  const TileMeta* tiles;
  const uint32_t* table;  // [tile][krows][tw]
  const uint8_t* amap;    // byte -> alphabet code (krows - 1 = no match)
  const uint8_t* mmin;    // [65][65]
  int krows;
  int tw;
  int mode;
  int p5;       // 1: string-2 values of 1..5 bytes are in kP5 tiles (kind_for)
  int64_t n_b;  // id stride (n_b for linkage, n for deduplication)
  float p, lo, hi;
  uint64_t* hits;
  int64_t cap;
  unsigned long long* counter;
  int krows_max;  // LDS rows reserved (>= krows)
  // Hash emission (blocked linkage): when hkeys is set, a hit adds
  // level * code_stride to the code of its pair id in an open-addressing
  // table instead of being appended to `hits`.
  unsigned long long* hkeys;
  uint32_t* hcodes;
  uint64_t hmask;
  uint32_t code_stride;
  unsigned int* overflow;
  unsigned long long* n_candidates;  // pairs scored exactly (may be null)
  const int64_t* cls_start;          // [kNumClasses] first string-2 sorted index of each length
  const int64_t* cls_tile0;          // [kNumClasses] first bit-parallel tile of each length
  uint64_t* cand;                    // candidate buffer of the prefix form's rounds
  int64_t cand_cap;
  unsigned long long* cand_count;
  // Prefix form (fh_jw_prefix.h): row codes of every string-1 value of at
  // most 64 bytes, four codes per dword (layout at kPrefixCodeWords), and the
  // first string-1 sorted index of each length class.
  const uint32_t* codes_a;
  int64_t codes_hi0;      // first sorted index of a value over 32 bytes
  int64_t codes_hi_base;  // words 8-15 of those values start at codes_a + codes_hi_base
  const int64_t* cls_start_a;  // [kNumClasses]
  // Row codes of every string-2 value of at most 32 bytes in its sorted
  // order, same layout (words 0-7); in deduplication the string-1 codes.
  const uint32_t* codes_b;
};

constexpr unsigned long long kEmptyKey = ~0ull;
constexpr int kMaxProbes = 4096;

// Host side: the pruning bound.
inline float host_formula(int l1, int l2, int m, int tt, int pre, float p, bool fma) {
  if (m == 0) return 0.0f;
  float match = (float)m;
  float t = (float)tt;
  t /= 2;
  float dist = (float)((double)((match / (float)l1) + (match / (float)l2) + ((match - t) / match)) / 3.0);
  float prefix = (float)pre;
  if (fma) return std::fmaf(p * prefix, 1 - dist, dist);
  volatile float q = (p * prefix) * (1 - dist);
  return dist + q;
}

inline bool host_reaches(float s, float lo, float hi) { return (s >= hi) || (s >= lo && s < hi); }

// Smallest match count for which some transposition count and prefix reach
// level >= 1 (255: unreachable). Exhaustive over (tt, prefix) per m, both
// with and without the fused final update, so the bound is conservative.
inline uint8_t host_mmin(int l1, int l2, float p, float lo, float hi) {
  if (host_reaches(0.0f, lo, hi)) return 0;
  const int mmax = std::min(l1, l2);
  const int pmax = std::min(4, mmax);
  for (int m = 1; m <= mmax; ++m)
    for (int tt = 0; tt <= m; ++tt)
      for (int pre = 0; pre <= pmax; ++pre)
        if (host_reaches(host_formula(l1, l2, m, tt, pre, p, true), lo, hi) ||
            host_reaches(host_formula(l1, l2, m, tt, pre, p, false), lo, hi))
          return (uint8_t)m;
  return 255;
}

// host_mmin for every length pair 1..kLongLen (entry la * (kLongLen + 1) +
// lb), computed once per (p, lo, hi) and kept for the process (the table
// depends on nothing else; computing it takes milliseconds of host time).
const std::vector<uint8_t>& mmin_table(float p, float lo, float hi);


// Launches k_jw_tiles over waves [w0, w0 + nw) of one string-1 length range
// (r0: 1..16 bytes, r1: 17..32, r2: 33..64) against tiles [t0, t0 + nt) of
// one kind. Each range lives in its own translation unit.
int launch_jw_tiles_r0(int kind, const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0, int64_t nt,
                       hipStream_t s);
int launch_jw_tiles_r1(int kind, const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0, int64_t nt,
                       hipStream_t s);
int launch_jw_tiles_r2(int kind, const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0, int64_t nt,
                       hipStream_t s);

// Device code: internal to every translation unit that includes this header.
namespace {

// ---------------------------------------------------------------------------
// The reference's score, restated for the device (faster/comparison.py:11-131)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float jw_formula(int l1, int l2, int m, int tt, int pre, float p) {
  if (m == 0) return 0.0f;
  float match = (float)m;
  float t = (float)tt;
  t /= 2;
  float dist = (float)((double)((match / (float)l1) + (match / (float)l2) + ((match - t) / match)) / 3.0);
  float prefix = (float)pre;
  return fmaf(p * prefix, 1 - dist, dist);
}

__device__ __forceinline__ int level_of(float s, float lo, float hi) {
  if (s >= hi) return 2;
  if (s >= lo && s < hi) return 1;
  return 0;
}

// Greedy match with per-position "used" bits held in WORDS 64-bit words
// (strings up to WORDS * 64 bytes), transpositions, prefix and formula.
template <int WORDS>
__device__ float jw_exact(const uint8_t* __restrict__ s1, int l1, const uint8_t* __restrict__ s2, int l2,
                          float p) {
  if (l1 == 0 || l2 == 0) return 0.0f;
  const int w = max(l1, l2) / 2 - 1;
  uint64_t used2[WORDS];
  uint64_t flag1[WORDS];
#pragma unroll
  for (int k = 0; k < WORDS; ++k) {
    used2[k] = 0;
    flag1[k] = 0;
  }
  int m = 0;
  for (int i = 0; i < l1; ++i) {
    const uint8_t c = s1[i];
    const int jlo = max(0, i - w);
    const int jhi = min(l2, i + w + 1);
    for (int j = jlo; j < jhi; ++j) {
      if (s2[j] == c && !((used2[j >> 6] >> (j & 63)) & 1ull)) {
        used2[j >> 6] |= 1ull << (j & 63);
        flag1[i >> 6] |= 1ull << (i & 63);
        ++m;
        break;
      }
    }
  }
  if (m == 0) return 0.0f;
  int tt = 0;
  int wi = 0, wj = 0;
  uint64_t f = flag1[0], u = used2[0];
  for (int k = 0; k < m; ++k) {
    while (f == 0) f = flag1[++wi];
    while (u == 0) u = used2[++wj];
    const int i = wi * 64 + __builtin_ctzll(f);
    const int j = wj * 64 + __builtin_ctzll(u);
    if (s1[i] != s2[j]) ++tt;
    f &= f - 1;
    u &= u - 1;
  }
  int pre = 0;
  const int lim = min(min(l1, l2), 4);
  for (int i = 0; i < lim; ++i) {
    if (s1[i] == s2[i])
      ++pre;
    else
      break;
  }
  return jw_formula(l1, l2, m, tt, pre, p);
}

// The same score for values of any length (the reference sizes its work
// arrays from the offsets, faster/comparison.py:304-307): the used
// positions of string 2 and the matched positions of string 1 are bit flags
// in global memory, word k of this thread's flags at flags[k * stride]
// (string 2's ceil(l2 / 64) words first, then string 1's), so that threads
// sharing a scratch buffer interleave their words.
__device__ __noinline__ float jw_exact_mem(const uint8_t* __restrict__ s1, int l1, const uint8_t* __restrict__ s2,
                                           int l2, float p, uint64_t* __restrict__ flags, int64_t stride) {
  if (l1 == 0 || l2 == 0) return 0.0f;
  const int w = max(l1, l2) / 2 - 1;
  const int w2 = (l2 + 63) >> 6, w1 = (l1 + 63) >> 6;
  uint64_t* used2 = flags;
  uint64_t* flag1 = flags + (int64_t)w2 * stride;
  for (int k = 0; k < w2; ++k) used2[(int64_t)k * stride] = 0;
  for (int k = 0; k < w1; ++k) flag1[(int64_t)k * stride] = 0;
  int m = 0;
  for (int i = 0; i < l1; ++i) {
    const uint8_t c = s1[i];
    const int jlo = max(0, i - w);
    const int jhi = min(l2, i + w + 1);
    for (int j = jlo; j < jhi; ++j) {
      if (s2[j] != c) continue;
      uint64_t* u = used2 + (int64_t)(j >> 6) * stride;
      if ((*u >> (j & 63)) & 1ull) continue;
      *u |= 1ull << (j & 63);
      flag1[(int64_t)(i >> 6) * stride] |= 1ull << (i & 63);
      ++m;
      break;
    }
  }
  if (m == 0) return 0.0f;
  int tt = 0;
  int wi = 0, wj = 0;
  uint64_t f = flag1[0], u = used2[0];
  for (int k = 0; k < m; ++k) {
    while (f == 0) f = flag1[(int64_t)(++wi) * stride];
    while (u == 0) u = used2[(int64_t)(++wj) * stride];
    const int i = wi * 64 + __builtin_ctzll(f);
    const int j = wj * 64 + __builtin_ctzll(u);
    if (s1[i] != s2[j]) ++tt;
    f &= f - 1;
    u &= u - 1;
  }
  int pre = 0;
  const int lim = min(min(l1, l2), 4);
  for (int i = 0; i < lim; ++i) {
    if (s1[i] == s2[i])
      ++pre;
    else
      break;
  }
  return jw_formula(l1, l2, m, tt, pre, p);
}

// Byte-wise lexicographic order (UTF-8 byte order equals code-point order);
// returns true when (s1, l1) > (s2, l2).
__device__ __forceinline__ bool lex_greater(const uint8_t* s1, int l1, const uint8_t* s2, int l2) {
  const int n = min(l1, l2);
  for (int i = 0; i < n; ++i) {
    if (s1[i] != s2[i]) return s1[i] > s2[i];
  }
  return l1 > l2;
}


__device__ __forceinline__ uint64_t hash_id(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

// Adds `add` to the code of pair `id` (linear probing, bounded; a full
// neighbourhood raises the overflow flag and drops the update).
__device__ __forceinline__ void hash_add(const JwArgs& a, uint64_t id, uint32_t add) {
  uint64_t slot = hash_id(id) & a.hmask;
  for (int probe = 0; probe < kMaxProbes; ++probe) {
    const unsigned long long prev = atomicCAS(a.hkeys + slot, kEmptyKey, (unsigned long long)id);
    if (prev == kEmptyKey || prev == id) {
      atomicAdd(a.hcodes + slot, add);
      return;
    }
    slot = (slot + 1) & a.hmask;
  }
  atomicOr(a.overflow, 1u);
}

// Emits the level key of pair (x = string 1, y = string 2) when its level is > 0.
__device__ __forceinline__ void emit_level(const JwArgs& a, uint64_t x, uint64_t y, int lvl) {
  if (lvl <= 0) return;
  const uint64_t id = (uint64_t)x * (uint64_t)a.n_b + y;
  if (a.hkeys) {
    hash_add(a, id, (uint32_t)lvl * a.code_stride);
  } else {
    const unsigned long long pos = atomicAdd(a.counter, 1ull);
    if ((int64_t)pos < a.cap) a.hits[pos] = id * 4ull + (uint64_t)lvl;
  }
}

// The same for every active lane of a wave at once (all of them must call
// it): the appended keys take one atomic per wave instead of one per key.
__device__ __forceinline__ void emit_levels_wave(const JwArgs& a, uint64_t x, uint64_t y, int lvl) {
  if (a.hkeys) {
    emit_level(a, x, y, lvl);
    return;
  }
  const uint64_t mask = __ballot(lvl > 0);
  if (mask == 0) return;
  unsigned long long base = 0;
  const int first = __builtin_ctzll(mask);
  if ((int)__lane_id() == first) base = atomicAdd(a.counter, (unsigned long long)__popcll(mask));
  base = __shfl(base, first, 64);
  if (lvl > 0) {
    const unsigned long long pos = base + (unsigned long long)__popcll(mask & fhd::lanemask_lt());
    if ((int64_t)pos < a.cap) a.hits[pos] = ((uint64_t)x * (uint64_t)a.n_b + y) * 4ull + (uint64_t)lvl;
  }
}

// Window of string-2 positions [i - w, i + w] clipped to [0, lenB).
__device__ __forceinline__ uint64_t window_bits(int i, int w, int lenB) {
  const int lo_i = max(0, i - w);
  const int hi_i = min(lenB - 1, i + w);
  if (hi_i < lo_i) return 0ull;
  const int width = hi_i - lo_i + 1;
  return (width >= 64 ? ~0ull : ((1ull << width) - 1ull)) << lo_i;
}

// Level of one pair by the byte-wise restatement (jw_exact), plus 4 when
// string 1 is the B-side value eb (in deduplication the lexicographically
// larger value is string 1, faster/deduplication.py).
constexpr int kSwapped = 4;
__device__ __noinline__ int pair_level(const JwArgs& a, uint64_t ea, uint64_t eb) {
  const int64_t oa = a.off_a[ea];
  const int la = (int)(a.off_a[ea + 1] - oa);
  const int64_t ob = a.off_b[eb];
  const int lb = (int)(a.off_b[eb + 1] - ob);
  const uint8_t* sa = a.chars_a + oa;
  const uint8_t* sb = a.chars_b + ob;
  if (a.mode == FH_MODE_DEDUP && !lex_greater(sa, la, sb, lb))
    return level_of(jw_exact<1>(sb, lb, sa, la, a.p), a.lo, a.hi) + kSwapped;
  return level_of(jw_exact<1>(sa, la, sb, lb, a.p), a.lo, a.hi);
}

// Exact score of one pair (sorted indices sa, sb), from string 2's
// position masks in the bit-parallel table (the complemented word of row
// code(c), at the string's column, has bit j clear iff byte j is c):
//   * greedy match (faster/comparison.py:48-64): for character i of string 1
//     the lowest in-window, unused position of its row is taken;
//   * pairs with m < mmin stop here (level 0);
//   * transpositions (:73-99): the k-th matched character of string 1 is
//     compared with the k-th matched position of string 2 by testing that
//     position in the character's row;
//   * prefix (:106-123): position i of string 2 holds string-1 byte i iff bit
//     i of that byte's row is set; then the reference's formula.
// Bytes of string 2 are never read. In deduplication the lexicographically
// larger value is string 1 (faster/deduplication.py); when that is the
// string-2 side the pair goes to the byte-wise restatement (pair_level).
__device__ __forceinline__ uint64_t row_mask(const uint32_t* base, int tw, int code, int kind, int col, int shift,
                                             uint64_t lmask) {
  const uint32_t* row = base + (size_t)code * tw + col;
  const uint64_t w = kind == kW64 ? ~(((uint64_t)row[1] << 32) | row[0]) : (uint64_t)(~row[0] >> shift);
  return w & lmask;
}

//
// The pair already passed the greedy count (a comparison kernel's survivor),
// so in deduplication the string order is settled first and a swapped pair
// goes to the byte-wise restatement without the table's greedy pass.
__device__ __forceinline__ int candidate_level(const JwArgs& a, int64_t sa, int64_t sb, uint64_t& x, uint64_t& y) {
  const uint64_t ea = a.sorted_a[sa], eb = a.sorted_b[sb];
  x = ea;
  y = eb;
  const int64_t oa = a.off_a[ea];
  const int la = (int)(a.off_a[ea + 1] - oa);
  const int64_t ob = a.off_b[eb];
  const int lb = (int)(a.off_b[eb + 1] - ob);
  if (a.mode == FH_MODE_DEDUP && !lex_greater(a.chars_a + oa, la, a.chars_b + ob, lb)) {
    const int code = pair_level(a, ea, eb);
    x = eb;
    y = ea;
    return code - kSwapped;
  }
  const int mm = a.mmin[la * (kLongLen + 1) + lb];
  const int kind = kind_for(lb, a.p5 != 0);
  const int64_t pos = sb - a.cls_start[lb];
  const int per = kind_tile_strings(kind, a.tw);
  const int64_t tile = a.cls_tile0[lb] + pos / per;
  const int p = (int)(pos % per);
  const uint32_t* base = a.table + (size_t)tile * a.krows * a.tw;
  int col = p, shift = 0;
  if (packed_kind(kind)) {
    col = p / kind_lanes(kind);
    shift = kind_lane_shift(kind, p % kind_lanes(kind));
  } else if (kind == kW64) {
    col = 2 * p;
  }
  const uint8_t* s1 = a.chars_a + oa;
  const int w = max(la, lb) / 2 - 1;
  const uint64_t lmask = lb >= 64 ? ~0ull : ((1ull << lb) - 1ull);
  uint64_t used = 0, flag1 = 0;
#pragma unroll 4
  for (int i = 0; i < la; ++i) {
    uint64_t avail = row_mask(base, a.tw, a.amap[s1[i]], kind, col, shift, lmask);
    avail &= window_bits(i, w, lb) & ~used;
    if (avail) {
      used |= avail & (~avail + 1ull);
      flag1 |= 1ull << i;
    }
  }
  const int m = __popcll(used);
  if (m < mm) return 0;
  float s = 0.0f;
  if (m > 0) {
    int tt = 0;
    uint64_t f = flag1, u = used;
    for (int k = 0; k < m; ++k) {
      const int i = __builtin_ctzll(f), j = __builtin_ctzll(u);
      if (!((row_mask(base, a.tw, a.amap[s1[i]], kind, col, shift, lmask) >> j) & 1ull)) ++tt;
      f &= f - 1;
      u &= u - 1;
    }
    int pre = 0;
    const int lim = min(min(la, lb), 4);
    for (int i = 0; i < lim; ++i) {
      if ((row_mask(base, a.tw, a.amap[s1[i]], kind, col, shift, lmask) >> i) & 1ull)
        ++pre;
      else
        break;
    }
    s = jw_formula(la, lb, m, tt, pre, a.p);
  }
  return level_of(s, a.lo, a.hi);
}

// Processes the wave's queue of (string-1 lane, string-2 sorted index) pairs.
// With a candidate buffer (a.cand), the queue is appended to it as
// (string-1 sorted index << 32) | string-2 sorted index, with one atomic per
// flush, for a separate scoring kernel; entries past its capacity are
// counted only. Without one, the pairs are scored here from the string-2
// mask table (candidate_level: one row read per string-1 character instead
// of a byte loop over both strings; 35% faster on the longest real-data
// variable) and their keys appended with one atomic per wave.
// The persistent blocked kernel's form: hits accumulate in the pair table
// (hash_add, one slot per pair), so each lane scores and adds its own pairs;
// kept apart so the blocked kernel does not carry the list form's registers.
__device__ __noinline__ void score_and_add(const JwArgs& a, uint64_t ea, uint64_t eb) {
  const int code = pair_level(a, ea, eb);
  if (code >= kSwapped)
    emit_level(a, eb, ea, code & 3);
  else
    emit_level(a, ea, eb, code);
}

__device__ __noinline__ void flush_queue_hash(const JwArgs& a, const uint64_t* q, int qn, uint64_t s0a) {
  for (int k = (int)__lane_id(); k < qn; k += 64) {
    const uint64_t v = q[k];
    score_and_add(a, a.sorted_a[s0a + (v & 63ull)], a.sorted_b[v >> 6]);
  }
}

__device__ __noinline__ void flush_queue(const JwArgs& a, const uint64_t* q, int qn, uint64_t s0a) {
  if (a.cand) {
    unsigned long long base = 0;
    if (__lane_id() == 0) base = atomicAdd(a.cand_count, (unsigned long long)qn);
    base = __shfl(base, 0, 64);
    for (int k = (int)__lane_id(); k < qn; k += 64) {
      const uint64_t v = q[k];
      if (base + (unsigned long long)k < (unsigned long long)a.cand_cap)
        a.cand[base + k] = ((s0a + (v & 63ull)) << 32) | (v >> 6);
    }
    return;
  }
  if (a.n_candidates && __lane_id() == 0) atomicAdd(a.n_candidates, (unsigned long long)qn);
  for (int k0 = 0; k0 < qn; k0 += 64) {
    const int k = k0 + (int)__lane_id();
    int lvl = 0;
    uint64_t x = 0, y = 0;
    if (k < qn) {
      const uint64_t v = q[k];
      lvl = candidate_level(a, (int64_t)(s0a + (v & 63ull)), (int64_t)(v >> 6), x, y);
    }
    emit_levels_wave(a, x, y, lvl);
  }
}

// The persistent blocked kernel, compiled in a translation unit of its own
// with FH_JW_PAIR_TABLE set, accumulates hits in the pair table and uses
// only flush_queue_hash; the plain passes use only flush_queue. Each kernel
// then carries the registers of its own form alone.
#ifndef FH_JW_PAIR_TABLE
#define FH_JW_PAIR_TABLE 0
#endif
__device__ __forceinline__ void flush_pairs(const JwArgs& a, const uint64_t* q, int qn, uint64_t s0a) {
#if FH_JW_PAIR_TABLE
  flush_queue_hash(a, q, qn, s0a);
#else
  flush_queue(a, q, qn, s0a);
#endif
}

// Pushes one survivor per lane where `pred` holds; flushes first when full.
__device__ __forceinline__ void queue_push(const JwArgs& a, uint64_t* q, int& qn, bool pred, uint64_t sb,
                                           uint64_t s0a) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  if (qn + 64 > kQCap) {
    flush_pairs(a, q, qn, s0a);
    qn = 0;
  }
  if (pred) {
    const int pos = qn + __popcll(mask & fhd::lanemask_lt());
    q[pos] = (sb << 6) | (uint64_t)__lane_id();
  }
  qn += __popcll(mask);
}

// gfx950 v_bitop3_b32 forms used by the match recurrence. The truth-table
// index of bit k is (a_k << 2) | (b_k << 1) | c_k.
__device__ __forceinline__ uint32_t or3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0xFE);  // a | b | c
}
__device__ __forceinline__ uint32_t select_bits(uint32_t sel, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(sel, b, c, 0xCA);  // sel ? b : c, bitwise
}
// Per-wave state for one (wave, tile) pass.
struct PassCtx {
  const uint32_t* lds;  // staged tile, row stride tw + 1 dwords
  uint64_t* q;          // this wave's survivor queue
  int qn;
  int lenB, cntB, mm, w;
  int64_t s0a, s0b;
  bool lane_ok, diag;
};

__device__ __forceinline__ const uint32_t* row_ptr(const uint32_t* lds, uint32_t off_bytes, int col) {
  return reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(lds) + off_bytes) + col;
}

// Row of string-1 character I at byte column colb of the staged tile. The
// row byte offsets are packed two per dword (a workgroup's LDS is at most
// 64 KB), and one SDWA add reads the 16-bit half in place: as plain shifts
// the compiler would unpack every offset into a register of its own before
// the column loop, which costs the occupancy this halving is for.
// colb is the LDS address of the column (lds_addr of the staged tile plus
// 4 * col), so the sum is the row's LDS address itself. The callers' loops
// over i are fully unrolled, so i is a constant here.
typedef __attribute__((address_space(3))) const uint32_t lds_u32;
__device__ __forceinline__ uint32_t lds_addr(const uint32_t* p) {
  return (uint32_t)(uintptr_t)(lds_u32*)p;
}
template <int N>
__device__ __forceinline__ lds_u32* lds_row(const uint32_t (&o)[N], int i, uint32_t colb) {
  uint32_t r;
  if (i & 1)
    asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_1 src1_sel:DWORD"
        : "=v"(r)
        : "v"(o[i >> 1]), "v"(colb));
  else
    asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_0 src1_sel:DWORD"
        : "=v"(r)
        : "v"(o[i >> 1]), "v"(colb));
  return (lds_u32*)(uintptr_t)r;
}
// The same for a run-time i (rolled passes): a select chain, no indexing.
template <int N>
__device__ __forceinline__ uint32_t row_off_rt(const uint32_t (&o)[N], int i) {
  uint32_t w = 0;
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (k == (i >> 1)) w = o[k];
  return (i & 1) ? (w >> 16) : (w & 0xFFFFu);
}


// Complemented window of string-1 character i in every lane of a packed
// kind: the lanes' positions inside the window are clear; the guards, the
// positions outside the window or past lenB and the bits above the last lane
// are set.
template <int KIND>
__device__ __forceinline__ uint32_t packed_nwin(int i, int w, int lenB) {
  const uint32_t win = (uint32_t)window_bits(i, w, lenB);
  uint32_t v = 0;
#pragma unroll
  for (int l = 0; l < kind_lanes(KIND); ++l) v |= win << kind_lane_shift(KIND, l);
  return ~v;
}

// Match count of lane l of a packed dword. The guards and the bits above the
// last lane are never set in `na` (they are set in nc, so the select keeps
// them), so the last lane needs no mask.
template <int KIND>
__device__ __forceinline__ int lane_count(uint32_t na, int l) {
  constexpr int LB = kind_lane_bits(KIND);
  if (l == kind_lanes(KIND) - 1) return __popc(na >> kind_lane_shift(KIND, l));
  return __popc((na >> kind_lane_shift(KIND, l)) & ((1u << (LB - 1)) - 1u));
}

// Bits [0, n) of a dword, n clipped to [0, 32].
__device__ __forceinline__ uint32_t low_bits(int64_t n) {
  return n <= 0 ? 0u : (n >= 32 ? ~0u : ((1u << n) - 1u));
}

// Queues the candidates whose bit is set in `smask` (bit k = string-2 tile
// index first + k, k < n), after the validity and triangle checks. Rare
// path: each lane walks its own set bits, so the wave loops as often as the
// busiest lane has candidates (almost always once).
__device__ __forceinline__ void push_candidates(const JwArgs& a, PassCtx& c, uint32_t smask, int first, int n) {
  uint32_t m = c.lane_ok ? smask & low_bits(min(n, c.cntB - first)) : 0u;
  // Triangle: string-2 index s0b + first + k below string-1 index s0a + lane.
  if (c.diag) m &= low_bits(c.s0a + (int64_t)__lane_id() - c.s0b - first);
#pragma unroll 1
  while (__ballot(m != 0u) != 0) {
    const bool pred = m != 0u;
    const int k = pred ? __builtin_ctz(m) : 0;
    queue_push(a, c.q, c.qn, pred, (uint64_t)(c.s0b + first + k), (uint64_t)c.s0a);
    m &= m - 1u;
  }
}

// One string-1 length LA (<= 32) against a staged tile of 16- or 32-bit
// masks, fully unrolled over the LA characters so the row offsets stay in
// fixed VGPRs and the windows in SGPRs.
template <int LA, int KIND, int LMAX>
__device__ __forceinline__ void tile_pass(const JwArgs& a, PassCtx& c, const uint32_t (&off)[(LMAX + 1) / 2]) {
  const int lenB = c.lenB;
  if constexpr (packed_kind(KIND)) {
    constexpr int NL = kind_lanes(KIND);
    constexpr uint32_t ONES = kind_ones(KIND);
    constexpr int KU = kU;
    static_assert(NL * KU <= 32, "one survivor mask bit per pair of a step");
    uint32_t nwin[LA];
#pragma unroll
    for (int i = 0; i < LA; ++i) nwin[i] = __builtin_amdgcn_readfirstlane(packed_nwin<KIND>(i, c.w, lenB));
    // navail starts at 0: positions past lenB never become candidates because
    // the tile masks and the windows block them, so the lane counts are m.
    const uint32_t init = 0u;
    const int mm = c.mm;
    const int ncols = (c.cntB + NL - 1) / NL;
    for (int col = 0; col < ncols; col += KU) {
      const uint32_t colb = lds_addr(c.lds) + (uint32_t)col * 4u;
      uint32_t na[KU], cur[KU];
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        na[u] = init;
        cur[u] = lds_row(off, 0, colb)[u];
      }
#pragma unroll
      for (int i = 0; i < LA; ++i) {
        // With FH_JW_PREFETCH, rows of character i + 1 are read while
        // character i is processed (the scheduling fences keep the reads
        // ahead of this step's ops); without, other waves hide the latency.
        uint32_t nxt[KU];
        if (FH_JW_PREFETCH && i + 1 < LA) {
#pragma unroll
          for (int u = 0; u < KU; ++u) nxt[u] = lds_row(off, i + 1, colb)[u];
        }
        if (!FH_JW_PREFETCH && i > 0) {
#pragma unroll
          for (int u = 0; u < KU; ++u) cur[u] = lds_row(off, i, colb)[u];
        }
        __builtin_amdgcn_sched_barrier(0);
        uint32_t nc[KU], t[KU];
#pragma unroll
        for (int u = 0; u < KU; ++u) nc[u] = or3(cur[u], nwin[i], na[u]);
#pragma unroll
        for (int u = 0; u < KU; ++u) t[u] = nc[u] + ONES;
#pragma unroll
        for (int u = 0; u < KU; ++u) na[u] = select_bits(nc[u], na[u], t[u]);
        __builtin_amdgcn_sched_barrier(0);
        if (FH_JW_PREFETCH && i + 1 < LA) {
#pragma unroll
          for (int u = 0; u < KU; ++u) cur[u] = nxt[u];
        }
      }
      // Largest match count over the NL * KU pairs.
      uint32_t mx = 0;
#pragma unroll
      for (int u = 0; u < KU; ++u) {
#pragma unroll
        for (int k = 0; k < NL; ++k) mx = max(mx, (uint32_t)lane_count<KIND>(na[u], k));
      }
      if (__ballot((int)mx >= mm && c.lane_ok) == 0) continue;
      // Rare path (the counts above stay live for it). Bit NL * u + k set
      // when pair (u, k) reaches mm; only the dwords in which some lane has
      // such a pair (usually one) build their bits.
      uint32_t pass = 0;
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        uint32_t mu = 0;
#pragma unroll
        for (int k = 0; k < NL; ++k) mu = max(mu, (uint32_t)lane_count<KIND>(na[u], k));
        if (__ballot((int)mu >= mm && c.lane_ok) == 0) continue;
#pragma unroll
        for (int k = 0; k < NL; ++k)
          pass |= (uint32_t)(lane_count<KIND>(na[u], k) >= mm) << (NL * u + k);
      }
      push_candidates(a, c, pass, NL * col, NL * KU);
    }
  } else {
    uint32_t nwin[LA];
#pragma unroll
    for (int i = 0; i < LA; ++i) nwin[i] = __builtin_amdgcn_readfirstlane(~(uint32_t)window_bits(i, c.w, lenB));
    const uint32_t init = 0u;
    const int mm = c.mm;
    for (int col = 0; col < c.cntB; col += kU) {
      const uint32_t colb = lds_addr(c.lds) + (uint32_t)col * 4u;
      uint32_t na[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) na[u] = init;
#pragma unroll
      for (int i = 0; i < LA; ++i) {
        lds_u32* row = lds_row(off, i, colb);
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const uint32_t nc = or3(row[u], nwin[i], na[u]);
          na[u] = select_bits(nc, na[u], nc + 1u);
        }
      }
      uint32_t smask = 0;
#pragma unroll
      for (int u = 0; u < kU; ++u) smask |= (uint32_t)((int)__popc(na[u]) >= mm) << u;
      if (__ballot(smask != 0 && c.lane_ok) == 0) continue;
      push_candidates(a, c, smask, col, kU);
    }
  }
}

// Any string-1 length (rolled loop over its characters) against any mask
// width; used for the rare long strings (33..64 bytes on either side).
template <int KIND, int LMAX>
__device__ __forceinline__ void tile_pass_rolled(const JwArgs& a, PassCtx& c, const uint32_t (&off)[(LMAX + 1) / 2], int lenA) {
  const int lenB = c.lenB;
  constexpr int per = packed_kind(KIND) ? kind_lanes(KIND) : 1;  // strings per dword
  constexpr int dws = KIND == kW64 ? 2 : 1;                      // dwords per string
  const int ncols = packed_kind(KIND) ? (c.cntB + per - 1) / per : c.cntB * dws;
#pragma unroll 1
  for (int col = 0; col < ncols; col += dws) {
    uint64_t na = 0;  // lane counts = m (see tile_pass)
#pragma unroll 1
    for (int i = 0; i < lenA; ++i) {
      const uint32_t* row = row_ptr(c.lds, row_off_rt(off, i), col);
      if constexpr (packed_kind(KIND)) {
        const uint32_t nc = row[0] | packed_nwin<KIND>(i, c.w, lenB) | (uint32_t)na;
        na = select_bits(nc, (uint32_t)na, nc + kind_ones(KIND));
      } else if constexpr (KIND == kW32) {
        const uint32_t nc = row[0] | ~(uint32_t)window_bits(i, c.w, lenB) | (uint32_t)na;
        na = (nc & (uint32_t)na) | (~nc & (nc + 1u));
      } else {
        const uint64_t nm = ((uint64_t)row[1] << 32) | row[0];
        const uint64_t nc = nm | ~window_bits(i, c.w, lenB) | na;
        na = (nc & na) | (~nc & (nc + 1ull));
      }
    }
    uint32_t smask = 0;
    if constexpr (packed_kind(KIND)) {
#pragma unroll
      for (int k = 0; k < per; ++k) smask |= (uint32_t)(lane_count<KIND>((uint32_t)na, k) >= c.mm) << k;
    } else if constexpr (KIND == kW32) {
      smask = (uint32_t)((int)__popc((uint32_t)na) >= c.mm);
    } else {
      smask = (uint32_t)((int)__popcll(na) >= c.mm);
    }
    if (__ballot(smask != 0 && c.lane_ok) == 0) continue;
    push_candidates(a, c, smask, col * per / dws, per);
  }
}

// Uniform dispatch from the wave's string-1 length to its unrolled pass.
template <int LO, int HI, int KIND, int LMAX>
struct PassDispatch {
  __device__ static __forceinline__ void run(int lenA, const JwArgs& a, PassCtx& c, const uint32_t (&off)[(LMAX + 1) / 2]) {
    if (lenA == LO)
      tile_pass<LO, KIND, LMAX>(a, c, off);
    else
      PassDispatch<LO + 1, HI, KIND, LMAX>::run(lenA, a, c, off);
  }
};
template <int HI, int KIND, int LMAX>
struct PassDispatch<HI, HI, KIND, LMAX> {
  __device__ static __forceinline__ void run(int lenA, const JwArgs& a, PassCtx& c, const uint32_t (&off)[(LMAX + 1) / 2]) {
    if (lenA == HI) tile_pass<HI, KIND, LMAX>(a, c, off);
  }
};

// Compares this workgroup's waves [wave_first, wave_first + kWaves) ∩
// [.., wave_end) against the string-2 tiles [t_begin, t_end).
// KIND >= 0: every tile has that mask kind. KIND == kMixed: tiles of any
// kind (per-tile uniform dispatch), used by the persistent blocked kernel.
constexpr int kMixed = -1;

template <int LMIN, int LMAX, int KIND>
__device__ __forceinline__ void run_item(const JwArgs& a, uint32_t* lds, int64_t wave_first, int64_t wave_end,
                                         int64_t t_begin, int64_t t_end) {
  const int tw = a.tw;
  const int stride = tw + 1;
  uint64_t* queue = reinterpret_cast<uint64_t*>(lds + a.krows_max * stride + (a.krows_max * stride & 1));
  const int wave = threadIdx.x >> 6;
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
  c.s0a = wm.s0;

  // Byte offsets of each character's LDS row, two per dword (lds_row).
  uint32_t off[(LMAX + 1) / 2];
  {
    int64_t base = 0;
    if (c.lane_ok) base = a.off_a[a.sorted_a[wm.s0 + lane]];
    const uint32_t none = (uint32_t)(a.krows - 1);
#pragma unroll
    for (int k = 0; k < (LMAX + 1) / 2; ++k) off[k] = 0;
#pragma unroll
    for (int i = 0; i < LMAX; ++i) {
      uint32_t code = none;
      if (c.lane_ok && i < lenA) code = a.amap[a.chars_a[base + i]];
      off[i >> 1] |= (code * (uint32_t)stride * 4u) << (16 * (i & 1));
    }
  }

  for (int64_t t = t_begin; t < t_end; ++t) {
    const TileMeta tm = a.tiles[t];
    __syncthreads();
    {
      // Only the columns the passes read: the tile's used columns rounded up
      // to a power of two >= kU (columns past the tile's strings hold stale
      // data, whose candidates the passes drop by index).
      const int shift = tile_cols_log2(tm.kind, tm.count, tw);
      const int cols = 1 << shift;
      const uint32_t* src = a.table + (size_t)t * a.krows * tw;
      const int total = a.krows << shift;
      for (int idx = threadIdx.x; idx < total; idx += kThreads) {
        const int r = idx >> shift;
        const int col = idx & (cols - 1);
        lds[r * stride + col] = src[r * tw + col];
      }
    }
    __syncthreads();
    if (!wave_ok) continue;
    c.lenB = __builtin_amdgcn_readfirstlane(tm.len);
    c.cntB = __builtin_amdgcn_readfirstlane(tm.count);
    c.s0b = tm.s0;
    c.diag = false;
    if (a.mode == FH_MODE_DEDUP) {
      if (lenA < c.lenB) continue;
      if (lenA == c.lenB) {
        if (wm.s0 + cntA - 1 <= c.s0b) continue;  // no string-1 index above any string-2 index
        c.diag = true;
      }
    }
    c.mm = a.mmin[lenA * (kLongLen + 1) + c.lenB];
    if (c.mm > min(lenA, c.lenB)) continue;
    c.w = max(lenA, c.lenB) / 2 - 1;
    if constexpr (KIND == kMixed) {
      const int kind = tm.kind;
      if constexpr (LMAX <= 32) {
        if (kind == kP4)
          PassDispatch<LMIN, LMAX, kP4, LMAX>::run(lenA, a, c, off);
        else if (kind == kP3)
          PassDispatch<LMIN, LMAX, kP3, LMAX>::run(lenA, a, c, off);
        else if (kind == kP2)
          PassDispatch<LMIN, LMAX, kP2, LMAX>::run(lenA, a, c, off);
        else if (kind == kW32)
          PassDispatch<LMIN, LMAX, kW32, LMAX>::run(lenA, a, c, off);
        else
          tile_pass_rolled<kW64, LMAX>(a, c, off, lenA);
      } else {
        if (kind == kP4)
          tile_pass_rolled<kP4, LMAX>(a, c, off, lenA);
        else if (kind == kP3)
          tile_pass_rolled<kP3, LMAX>(a, c, off, lenA);
        else if (kind == kP2)
          tile_pass_rolled<kP2, LMAX>(a, c, off, lenA);
        else if (kind == kW32)
          tile_pass_rolled<kW32, LMAX>(a, c, off, lenA);
        else
          tile_pass_rolled<kW64, LMAX>(a, c, off, lenA);
      }
    } else if constexpr (LMAX <= 32 && KIND != kW64) {
      PassDispatch<LMIN, LMAX, KIND, LMAX>::run(lenA, a, c, off);
    } else {
      tile_pass_rolled<KIND, LMAX>(a, c, off, lenA);
    }
  }
  if (wave_ok && c.qn > 0) flush_pairs(a, c.q, c.qn, (uint64_t)wm.s0);
}

// LMIN..LMAX: the string-1 lengths this instance handles.
// `ap` points to the arguments in device memory, so the rarely called
// exact-scoring functions take them by reference without a private copy.
template <int LMIN, int LMAX, int KIND>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(FH_JW_WAVES_PER_EU))) k_jw_tiles(const JwArgs* __restrict__ ap, int64_t wave0, int64_t nwaves,
                                                      int64_t tile0, int64_t ntiles, int tiles_per_chunk) {
  extern __shared__ uint32_t lds[];
  const int64_t t_begin = tile0 + (int64_t)blockIdx.y * tiles_per_chunk;
  const int64_t t_end = min(t_begin + tiles_per_chunk, tile0 + ntiles);
  run_item<LMIN, LMAX, KIND>(*ap, lds, wave0 + (int64_t)blockIdx.x * kWaves, wave0 + nwaves, t_begin, t_end);
}

// Complemented position masks of every string-2 tile, built in LDS: one
// workgroup per tile sets its krows x tw words to all ones, clears one bit
// per string byte with an LDS atomic (bit j of string s in the lane, column
// and kind its length gives), and writes the tile out with coalesced stores.
// Rows of bytes absent from string 1's side and columns past the tile's
// strings stay all ones. LDS: krows * tw * 4 bytes.
__global__ void __launch_bounds__(256) k_build_table(const uint8_t* __restrict__ chars,
                                                     const int64_t* __restrict__ off,
                                                     const uint64_t* __restrict__ sorted,
                                                     const TileMeta* __restrict__ tiles, int64_t ntiles,
                                                     const uint8_t* __restrict__ amap, int krows, int tw,
                                                     uint32_t* __restrict__ table) {
  extern __shared__ uint32_t tl[];
  const int total = krows * tw;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const TileMeta tm = tiles[t];
    __syncthreads();
    for (int i = threadIdx.x; i < total; i += blockDim.x) tl[i] = 0xFFFFFFFFu;
    __syncthreads();
    const int kind = tm.kind;
    const int n = tm.count * tm.len;
    for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
      const int sidx = idx / tm.len;
      const int j = idx - sidx * tm.len;
      const int64_t e = (int64_t)sorted[tm.s0 + sidx];
      const int code = amap[chars[off[e] + j]];
      if (code >= krows - 1) continue;
      int col, bit;
      if (packed_kind(kind)) {
        col = sidx / kind_lanes(kind);
        bit = kind_lane_shift(kind, sidx % kind_lanes(kind)) + j;
      } else if (kind == kW32) {
        col = sidx;
        bit = j;
      } else {
        col = 2 * sidx + (j >> 5);
        bit = j & 31;
      }
      atomicAnd(&tl[code * tw + col], ~(1u << bit));
    }
    __syncthreads();
    uint32_t* dst = table + (size_t)t * total;
    for (int i = threadIdx.x; i < total; i += blockDim.x) dst[i] = tl[i];
  }
}

// Launch of k_build_table over tiles [0, ntiles).
inline int launch_build_table(const uint8_t* chars, const int64_t* off, const uint64_t* sorted, const TileMeta* tiles,
                              int64_t ntiles, const uint8_t* amap, int krows, int tw, uint32_t* table, hipStream_t s) {
  if (ntiles <= 0) return FH_OK;
  const unsigned grid = (unsigned)std::min<int64_t>(ntiles, 1 << 16);
  hipLaunchKernelGGL(k_build_table, dim3(grid), dim3(256), (size_t)krows * tw * 4, s, chars, off, sorted, tiles,
                     ntiles, amap, krows, tw, table);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

template <int LMIN, int LMAX, int KIND>
static int launch_tiles(const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0, int64_t nt,
                        hipStream_t s) {
  if (nw <= 0 || nt <= 0) return FH_OK;
  const int64_t gx = ceil_div(nw, kWaves);
  // 16 tiles per workgroup, fewer when there are few string-1 waves (the
  // lengths with few strings): a launch of a handful of workgroups would run
  // each one's tiles one after another on an otherwise idle GPU.
  int tpc = 16;
  while (tpc > 1 && gx * ceil_div(nt, tpc) < 2048) tpc /= 2;
  int64_t gy = ceil_div(nt, tpc);
  while (gy > 65535) {
    tpc *= 2;
    gy = ceil_div(nt, tpc);
  }
  FH_REQUIRE(gx < (1ll << 31), "too many string-1 waves");
  hipLaunchKernelGGL((k_jw_tiles<LMIN, LMAX, KIND>), dim3((unsigned)gx, (unsigned)gy), dim3(kThreads), lds_bytes, s, args,
                     w0, nw, t0, nt, tpc);
  FH_HIP(hipGetLastError());
  return FH_OK;
}


// Dispatch from a runtime kind to the instance of one length range.
template <int LMIN, int LMAX>
int launch_kind(int kind, const JwArgs* args, size_t lds, int64_t w0, int64_t nw, int64_t t0, int64_t nt,
                hipStream_t s) {
  switch (kind) {
    case kP4: return launch_tiles<LMIN, LMAX, kP4>(args, lds, w0, nw, t0, nt, s);
    case kP3: return launch_tiles<LMIN, LMAX, kP3>(args, lds, w0, nw, t0, nt, s);
    case kP2: return launch_tiles<LMIN, LMAX, kP2>(args, lds, w0, nw, t0, nt, s);
    case kW32: return launch_tiles<LMIN, LMAX, kW32>(args, lds, w0, nw, t0, nt, s);
    case kW64: return launch_tiles<LMIN, LMAX, kW64>(args, lds, w0, nw, t0, nt, s);
    default: set_error("launch_kind: bad kind %d", kind); return FH_ERR_INTERNAL;
  }
}

}  // namespace
}  // namespace fh
