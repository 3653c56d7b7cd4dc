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
//     Strings of <= 16 bytes pack two string-2 masks per dword and the add is
//     a packed 16-bit add, so one op sequence advances two pairs.
//   * The match count m = popcount of used positions is an exact count for the
//     greedy match. A host-built table gives, for every (len1, len2), the
//     smallest m for which any transposition count and prefix can reach level
//     1 in the reference's fp32 arithmetic; pairs below it are level 0 and
//     cost nothing more. Survivors are queued per wave in LDS and evaluated
//     with all 64 lanes by an exact restatement of the reference function.
//   * Pairs with an empty string (only when a score of 0 reaches level 1) or a
//     string longer than 64 bytes go through a generic kernel.
#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <vector>

#include "fh_common.h"

namespace fh {
namespace {

constexpr int kWaves = 8;             // waves per workgroup
constexpr int kThreads = kWaves * 64;
constexpr int kU = 8;                 // dword columns per inner step
constexpr int kQCap = 128;            // survivor queue entries per wave
constexpr int kLongLen = 64;          // longest string on the bit-parallel path
constexpr int kMaxGenericLen = 1024;  // longest string supported at all
constexpr int kClassLong = kLongLen + 1;
constexpr int kNumClasses = kLongLen + 2;  // 0 = empty, 1..64, 65 = long

enum Kind { kP16 = 0, kW32 = 1, kW64 = 2 };

struct WaveMeta {
  int32_t len;    // string-1 length of every lane
  int32_t count;  // valid lanes (1..64)
  int64_t s0;     // first sorted index
};

struct TileMeta {
  int32_t len;    // string-2 length of every string in the tile
  int32_t count;  // strings in the tile
  int64_t s0;     // first sorted index
};

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

// Byte-wise lexicographic order (UTF-8 byte order equals code-point order);
// returns true when (s1, l1) > (s2, l2).
__device__ __forceinline__ bool lex_greater(const uint8_t* s1, int l1, const uint8_t* s2, int l2) {
  const int n = min(l1, l2);
  for (int i = 0; i < n; ++i) {
    if (s1[i] != s2[i]) return s1[i] > s2[i];
  }
  return l1 > l2;
}

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
};

constexpr unsigned long long kEmptyKey = ~0ull;
constexpr int kMaxProbes = 4096;

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

// Evaluates one queued pair exactly and emits its key when its level is > 0.
__device__ __noinline__ void eval_pair(const JwArgs& a, uint64_t ea, uint64_t eb) {
  const int64_t oa = a.off_a[ea];
  const int la = (int)(a.off_a[ea + 1] - oa);
  const int64_t ob = a.off_b[eb];
  const int lb = (int)(a.off_b[eb + 1] - ob);
  const uint8_t* sa = a.chars_a + oa;
  const uint8_t* sb = a.chars_b + ob;
  uint64_t x = ea, y = eb;
  float s;
  if (a.mode == FH_MODE_DEDUP && !lex_greater(sa, la, sb, lb)) {
    s = jw_exact<1>(sb, lb, sa, la, a.p);
    x = eb;
    y = ea;
  } else {
    s = jw_exact<1>(sa, la, sb, lb, a.p);
  }
  const int lvl = level_of(s, a.lo, a.hi);
  if (lvl > 0) {
    const uint64_t id = (uint64_t)x * (uint64_t)a.n_b + y;
    if (a.hkeys) {
      hash_add(a, id, (uint32_t)lvl * a.code_stride);
    } else {
      const unsigned long long pos = atomicAdd(a.counter, 1ull);
      if ((int64_t)pos < a.cap) a.hits[pos] = id * 4ull + (uint64_t)lvl;
    }
  }
}

// Processes the wave's queue of (string-1 lane, string-2 sorted index) pairs.
__device__ __noinline__ void flush_queue(const JwArgs& a, const uint64_t* q, int qn, uint64_t s0a) {
  for (int k = (int)__lane_id(); k < qn; k += 64) {
    const uint64_t v = q[k];
    const uint64_t ea = a.sorted_a[s0a + (v & 63ull)];
    const uint64_t eb = a.sorted_b[v >> 6];
    eval_pair(a, ea, eb);
  }
}

// Pushes one survivor per lane where `pred` holds; flushes first when full.
__device__ __forceinline__ void queue_push(const JwArgs& a, uint64_t* q, int& qn, bool pred, uint64_t sb,
                                           uint64_t s0a) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  if (qn + 64 > kQCap) {
    flush_queue(a, q, qn, s0a);
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
// Two independent 16-bit increments (v_pk_add_u16: no carry between halves).
__device__ __forceinline__ uint32_t add1_u16x2(uint32_t x) {
  typedef unsigned short us2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(us2, x) + (us2){1, 1});
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

// Window of string-2 positions [i - w, i + w] clipped to [0, lenB).
__device__ __forceinline__ uint64_t window_bits(int i, int w, int lenB) {
  const int lo_i = max(0, i - w);
  const int hi_i = min(lenB - 1, i + w);
  if (hi_i < lo_i) return 0ull;
  const int width = hi_i - lo_i + 1;
  return (width >= 64 ? ~0ull : ((1ull << width) - 1ull)) << lo_i;
}

// Queues the candidates whose bit is set in `smask` (bit k = string-2 tile
// index first + k), after the validity and triangle checks. Rare path.
__device__ __forceinline__ void push_candidates(const JwArgs& a, PassCtx& c, uint32_t smask, int first, int n) {
  const int lane = __lane_id();
#pragma unroll 1
  for (int k = 0; k < n; ++k) {
    const int sidx = first + k;
    bool pred = c.lane_ok && ((smask >> k) & 1u) && sidx < c.cntB;
    if (c.diag) pred = pred && (c.s0a + lane > c.s0b + sidx);
    queue_push(a, c.q, c.qn, pred, (uint64_t)(c.s0b + sidx), (uint64_t)c.s0a);
  }
}

// One string-1 length LA (<= 32) against a staged tile of 16- or 32-bit
// masks, fully unrolled over the LA characters so the row offsets stay in
// fixed VGPRs and the windows in SGPRs.
template <int LA, int KIND, int LMAX>
__device__ __forceinline__ void tile_pass(const JwArgs& a, PassCtx& c, const uint32_t (&off)[LMAX]) {
  const int lenB = c.lenB;
  if (KIND == kP16) {
    uint32_t nwin[LA];
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const uint32_t h = (~(uint32_t)window_bits(i, c.w, lenB)) & 0xFFFFu;
      nwin[i] = __builtin_amdgcn_readfirstlane(h | (h << 16));
    }
    // navail starts at 0: positions past lenB never become candidates because
    // the tile masks and the windows block them, so popcount(navail) = m.
    const uint32_t init = 0u;
    const int mm = c.mm;
    const int ncols = (c.cntB + 1) >> 1;
    for (int col = 0; col < ncols; col += kU) {
      uint32_t na[kU], cur[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        na[u] = init;
        cur[u] = row_ptr(c.lds, off[0], col)[u];
      }
#pragma unroll
      for (int i = 0; i < LA; ++i) {
        // Rows of character i + 1 are read while character i is processed
        // (the scheduling fences keep the reads ahead of this step's ops).
        uint32_t nxt[kU];
        if (i + 1 < LA) {
#pragma unroll
          for (int u = 0; u < kU; ++u) nxt[u] = row_ptr(c.lds, off[i + 1], col)[u];
        }
        __builtin_amdgcn_sched_barrier(0);
        uint32_t nc[kU], t[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) nc[u] = or3(cur[u], nwin[i], na[u]);
#pragma unroll
        for (int u = 0; u < kU; ++u) t[u] = add1_u16x2(nc[u]);
#pragma unroll
        for (int u = 0; u < kU; ++u) na[u] = select_bits(nc[u], na[u], t[u]);
        __builtin_amdgcn_sched_barrier(0);
        if (i + 1 < LA) {
#pragma unroll
          for (int u = 0; u < kU; ++u) cur[u] = nxt[u];
        }
      }
      // Largest match count over the 2 * kU pairs: popcount of each half.
      uint32_t mx = 0;
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const uint32_t lo = (uint32_t)__popc(na[u] & 0xFFFFu);
        const uint32_t hi = (uint32_t)__popc(na[u] >> 16);
        mx = max(mx, max(lo, hi));
      }
      if (__ballot((int)mx >= mm && c.lane_ok) == 0) continue;
      uint32_t smask = 0;
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        smask |= (uint32_t)((int)__popc(na[u] & 0xFFFFu) >= mm) << (2 * u);
        smask |= (uint32_t)((int)__popc(na[u] >> 16) >= mm) << (2 * u + 1);
      }
      push_candidates(a, c, smask, 2 * col, 2 * kU);
    }
  } else {
    uint32_t nwin[LA];
#pragma unroll
    for (int i = 0; i < LA; ++i) nwin[i] = __builtin_amdgcn_readfirstlane(~(uint32_t)window_bits(i, c.w, lenB));
    const uint32_t init = 0u;
    const int mm = c.mm;
    for (int col = 0; col < c.cntB; col += kU) {
      uint32_t na[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) na[u] = init;
#pragma unroll
      for (int i = 0; i < LA; ++i) {
        const uint32_t* row = row_ptr(c.lds, off[i], col);
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
__device__ __forceinline__ void tile_pass_rolled(const JwArgs& a, PassCtx& c, const uint32_t (&off)[LMAX], int lenA) {
  const int lenB = c.lenB;
  const int per = KIND == kP16 ? 2 : 1;       // strings per dword
  const int dws = KIND == kW64 ? 2 : 1;       // dwords per string
  const int ncols = KIND == kP16 ? (c.cntB + 1) >> 1 : c.cntB * dws;
  const int step = KIND == kW64 ? 2 : 1;
#pragma unroll 1
  for (int col = 0; col < ncols; col += step) {
    uint64_t na = 0;  // popcount(na) = m (see tile_pass)
#pragma unroll 1
    for (int i = 0; i < lenA; ++i) {
      uint32_t oi = 0;
#pragma unroll
      for (int k = 0; k < LMAX; ++k)
        if (k == i) oi = off[k];
      const uint32_t* row = row_ptr(c.lds, oi, col);
      const uint64_t win = window_bits(i, c.w, lenB);
      if (KIND == kP16) {
        const uint32_t h = (~(uint32_t)win) & 0xFFFFu;
        const uint32_t nc = row[0] | h | (h << 16) | (uint32_t)na;
        typedef unsigned short us2 __attribute__((ext_vector_type(2)));
        us2 v = __builtin_bit_cast(us2, nc);
        v = v + (us2){1, 1};
        const uint32_t t = __builtin_bit_cast(uint32_t, v);
        na = (nc & (uint32_t)na) | (~nc & t);
      } else if (KIND == kW32) {
        const uint32_t nc = row[0] | ~(uint32_t)win | (uint32_t)na;
        na = (nc & (uint32_t)na) | (~nc & (nc + 1u));
      } else {
        const uint64_t nm = ((uint64_t)row[1] << 32) | row[0];
        const uint64_t nc = nm | ~win | na;
        na = (nc & na) | (~nc & (nc + 1ull));
      }
    }
    uint32_t smask = 0;
    if (KIND == kP16) {
      smask = (uint32_t)((int)__popc((uint32_t)na & 0xFFFFu) >= c.mm) |
              ((uint32_t)((int)__popc((uint32_t)na >> 16) >= c.mm) << 1);
    } else if (KIND == kW32) {
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
  __device__ static __forceinline__ void run(int lenA, const JwArgs& a, PassCtx& c, const uint32_t (&off)[LMAX]) {
    if (lenA == LO)
      tile_pass<LO, KIND, LMAX>(a, c, off);
    else
      PassDispatch<LO + 1, HI, KIND, LMAX>::run(lenA, a, c, off);
  }
};
template <int HI, int KIND, int LMAX>
struct PassDispatch<HI, HI, KIND, LMAX> {
  __device__ static __forceinline__ void run(int lenA, const JwArgs& a, PassCtx& c, const uint32_t (&off)[LMAX]) {
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

  // Byte offsets of each character's LDS row.
  uint32_t off[LMAX];
  {
    int64_t base = 0;
    if (c.lane_ok) base = a.off_a[a.sorted_a[wm.s0 + lane]];
    const uint32_t none = (uint32_t)(a.krows - 1);
#pragma unroll
    for (int i = 0; i < LMAX; ++i) {
      uint32_t code = none;
      if (c.lane_ok && i < lenA) code = a.amap[a.chars_a[base + i]];
      off[i] = code * (uint32_t)stride * 4u;
    }
  }

  for (int64_t t = t_begin; t < t_end; ++t) {
    __syncthreads();
    {
      const uint32_t* src = a.table + (size_t)t * a.krows * tw;
      const int total = a.krows * tw;
      const int shift = __builtin_ctz(tw);
      for (int idx = threadIdx.x; idx < total; idx += kThreads) {
        const int r = idx >> shift;
        const int col = idx & (tw - 1);
        lds[r * stride + col] = src[idx];
      }
    }
    __syncthreads();
    if (!wave_ok) continue;
    const TileMeta tm = a.tiles[t];
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
      const int kind = c.lenB <= 16 ? kP16 : (c.lenB <= 32 ? kW32 : kW64);
      if constexpr (LMAX <= 32) {
        if (kind == kP16)
          PassDispatch<LMIN, LMAX, kP16, LMAX>::run(lenA, a, c, off);
        else if (kind == kW32)
          PassDispatch<LMIN, LMAX, kW32, LMAX>::run(lenA, a, c, off);
        else
          tile_pass_rolled<kW64, LMAX>(a, c, off, lenA);
      } else {
        if (kind == kP16)
          tile_pass_rolled<kP16, LMAX>(a, c, off, lenA);
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
  if (wave_ok && c.qn > 0) flush_queue(a, c.q, c.qn, (uint64_t)wm.s0);
}

// LMIN..LMAX: the string-1 lengths this instance handles.
// `ap` points to the arguments in device memory, so the rarely called
// exact-scoring functions take them by reference without a private copy.
template <int LMIN, int LMAX, int KIND>
__global__ void __launch_bounds__(kThreads) k_jw_tiles(const JwArgs* __restrict__ ap, int64_t wave0, int64_t nwaves,
                                                      int64_t tile0, int64_t ntiles, int tiles_per_chunk) {
  extern __shared__ uint32_t lds[];
  const int64_t t_begin = tile0 + (int64_t)blockIdx.y * tiles_per_chunk;
  const int64_t t_end = min(t_begin + tiles_per_chunk, tile0 + ntiles);
  run_item<LMIN, LMAX, KIND>(*ap, lds, wave0 + (int64_t)blockIdx.x * kWaves, wave0 + nwaves, t_begin, t_end);
}

// Pairs with a special string: rows = special string-1 entries (sorted
// indices [sp0, sp0 + nsp)) against every string 2 (x), or, with
// `swap`, special string-2 entries against every non-special string 1.
__global__ void __launch_bounds__(256) k_jw_generic(JwArgs a, int64_t sp0, int64_t nsp, int64_t n_other,
                                                    int swap, int empty_special, int64_t row0) {
  const int64_t r = row0 + blockIdx.y;
  if (r >= nsp) return;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n_other;
       x += (int64_t)gridDim.x * blockDim.x) {
    uint64_t ea, eb;
    if (!swap) {
      ea = a.sorted_a[sp0 + r];
      eb = (uint64_t)x;
    } else {
      eb = a.sorted_b[sp0 + r];
      ea = (uint64_t)x;
    }
    const int64_t oa = a.off_a[ea];
    const int la = (int)(a.off_a[ea + 1] - oa);
    const int64_t ob = a.off_b[eb];
    const int lb = (int)(a.off_b[eb + 1] - ob);
    const bool sa_special = la > kLongLen || (empty_special && la == 0);
    const bool sb_special = lb > kLongLen || (empty_special && lb == 0);
    if (a.mode == FH_MODE_DEDUP) {
      // Special x against every y != x; a pair of two specials once (x > y).
      if (ea == eb) continue;
      if (sb_special && ea < eb) continue;
    } else if (swap) {
      // Special string 2 against non-special string 1 only (the rest ran
      // in the unswapped launch).
      if (sa_special) continue;
    }
    const uint8_t* sa = a.chars_a + oa;
    const uint8_t* sb = a.chars_b + ob;
    uint64_t xx = ea, yy = eb;
    float s;
    if (a.mode == FH_MODE_DEDUP && !lex_greater(sa, la, sb, lb)) {
      s = jw_exact<kMaxGenericLen / 64>(sb, lb, sa, la, a.p);
      xx = eb;
      yy = ea;
    } else {
      s = jw_exact<kMaxGenericLen / 64>(sa, la, sb, lb, a.p);
    }
    const int lvl = level_of(s, a.lo, a.hi);
    if (lvl > 0) {
      const unsigned long long pos = atomicAdd(a.counter, 1ull);
      if ((int64_t)pos < a.cap) a.hits[pos] = ((uint64_t)xx * (uint64_t)a.n_b + yy) * 4ull + (uint64_t)lvl;
    }
  }
}

// Length class and alphabet presence of every string.
__global__ void __launch_bounds__(256) k_classify(const uint8_t* __restrict__ chars,
                                                  const int64_t* __restrict__ off, int64_t n,
                                                  uint32_t* __restrict__ cls, uint64_t* __restrict__ ent,
                                                  uint32_t* __restrict__ hist, uint32_t* __restrict__ present,
                                                  uint32_t* __restrict__ maxlen) {
  __shared__ uint32_t h[kNumClasses];
  __shared__ uint32_t pr[8];
  for (int i = threadIdx.x; i < kNumClasses; i += blockDim.x) h[i] = 0;
  if (threadIdx.x < 8) pr[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = off[e];
    const int64_t len = off[e + 1] - o;
    const uint32_t c = len > kLongLen ? (uint32_t)kClassLong : (uint32_t)len;
    cls[e] = c;
    ent[e] = (uint64_t)e;
    if (len > kLongLen) atomicMax(maxlen, len > 0xFFFFFFFFll ? 0xFFFFFFFFu : (uint32_t)len);
    atomicAdd(&h[c], 1u);
    if (len >= 1 && len <= kLongLen) {
      for (int64_t i = 0; i < len; ++i) {
        const uint32_t b = chars[o + i];
        atomicOr(&pr[b >> 5], 1u << (b & 31));
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kNumClasses; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
  if (threadIdx.x < 8 && pr[threadIdx.x]) atomicOr(&present[threadIdx.x], pr[threadIdx.x]);
}

// Complemented position masks of every string-2 tile: one thread per
// (tile column, alphabet row).
__global__ void __launch_bounds__(256) k_build_table(const uint8_t* __restrict__ chars,
                                                     const int64_t* __restrict__ off,
                                                     const uint64_t* __restrict__ sorted,
                                                     const TileMeta* __restrict__ tiles, int64_t ntiles,
                                                     const uint8_t* __restrict__ amap, int krows, int tw,
                                                     uint32_t* __restrict__ table) {
  const int64_t total = ntiles * (int64_t)tw;
  const int r = blockIdx.y;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = g / tw;
    const int col = (int)(g - t * tw);
    const TileMeta tm = tiles[t];
    uint32_t word = 0xFFFFFFFFu;
    if (r < krows - 1) {
      auto mask_of = [&](int sidx) -> uint64_t {
        if (sidx >= tm.count) return 0ull;
        const int64_t e = (int64_t)sorted[tm.s0 + sidx];
        const int64_t o = off[e];
        uint64_t m = 0;
        for (int j = 0; j < tm.len; ++j)
          if (amap[chars[o + j]] == r) m |= 1ull << j;
        return m;
      };
      if (tm.len <= 16) {
        const uint64_t m0 = mask_of(2 * col);
        const uint64_t m1 = mask_of(2 * col + 1);
        word = ~(uint32_t)(m0 | (m1 << 16));
      } else if (tm.len <= 32) {
        word = ~(uint32_t)mask_of(col);
      } else {
        const uint64_t m = mask_of(col >> 1);
        word = ~(uint32_t)((col & 1) ? (m >> 32) : m);
      }
    }
    table[((size_t)t * krows + r) * tw + col] = word;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static float host_formula(int l1, int l2, int m, int tt, int pre, float p, bool fma) {
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

static bool host_reaches(float s, float lo, float hi) { return (s >= hi) || (s >= lo && s < hi); }

// Smallest match count for which some transposition count and prefix reach
// level >= 1 (255: unreachable). Exhaustive over (tt, prefix) per m, both
// with and without the fused final update, so the bound is conservative.
static uint8_t host_mmin(int l1, int l2, float p, float lo, float hi) {
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

struct Plan {
  std::vector<WaveMeta> waves;
  std::vector<TileMeta> tiles;
  // Wave ranges per A length range (1..16, 17..32, 33..64) and tile ranges per kind.
  int64_t wave_lo[3], wave_hi[3];
  int64_t tile_lo[3], tile_hi[3];
};

static int lrange_of(int len) { return len <= 16 ? 0 : (len <= 32 ? 1 : 2); }

template <int LMIN, int LMAX, int KIND>
static int launch_tiles(const JwArgs* args, size_t lds_bytes, int64_t w0, int64_t nw, int64_t t0, int64_t nt,
                        hipStream_t s) {
  if (nw <= 0 || nt <= 0) return FH_OK;
  const int64_t gx = ceil_div(nw, kWaves);
  int tpc = 16;
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

template <int LMIN, int LMAX>
static int launch_kinds(const JwArgs* args, size_t lds, const Plan& pl, int r, hipStream_t s, int* launches) {
  const int64_t w0 = pl.wave_lo[r], nw = pl.wave_hi[r] - pl.wave_lo[r];
  if (nw <= 0) return FH_OK;
  for (int k = 0; k < 3; ++k) {
    const int64_t t0 = pl.tile_lo[k], nt = pl.tile_hi[k] - pl.tile_lo[k];
    if (nt <= 0) continue;
    ++*launches;
    if (k == kP16) FH_TRY((launch_tiles<LMIN, LMAX, kP16>(args, lds, w0, nw, t0, nt, s)));
    if (k == kW32) FH_TRY((launch_tiles<LMIN, LMAX, kW32>(args, lds, w0, nw, t0, nt, s)));
    if (k == kW64) FH_TRY((launch_tiles<LMIN, LMAX, kW64>(args, lds, w0, nw, t0, nt, s)));
  }
  return FH_OK;
}

}  // namespace
}  // namespace fh

extern "C" int fh_jw_levels(const uint8_t* chars_a, const int64_t* offsets_a, int64_t n_a, const uint8_t* chars_b,
                            const int64_t* offsets_b, int64_t n_b, int mode, float p, float lower_thr,
                            float upper_thr, uint64_t* hits, int64_t capacity, int64_t* n_hits,
                            fh_jw_stats* stats, fh_stream_t stream) {
  using namespace fh;
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  FH_REQUIRE(mode == FH_MODE_LINK || mode == FH_MODE_DEDUP, "fh_jw_levels: bad mode %d", mode);
  FH_REQUIRE(n_a >= 0 && n_b >= 0, "fh_jw_levels: negative sizes");
  FH_REQUIRE(n_hits != nullptr, "fh_jw_levels: n_hits is null");
  FH_REQUIRE(capacity >= 0 && (capacity == 0 || hits != nullptr), "fh_jw_levels: bad output arena");
  if (mode == FH_MODE_DEDUP) {
    chars_b = chars_a;
    offsets_b = offsets_a;
    n_b = n_a;
  }
  FH_REQUIRE(n_a < (1ll << 31) && n_b < (1ll << 31), "fh_jw_levels: more than 2^31 strings per side");
  *n_hits = 0;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  const double logical = mode == FH_MODE_LINK ? (double)n_a * (double)n_b : (double)n_a * (double)(n_a - 1) / 2.0;
  if (stats) stats->pairs_logical = (int64_t)logical;
  if (n_a == 0 || n_b == 0) return FH_OK;

  Scratch scr(s);
  const bool two = mode == FH_MODE_LINK;
  uint32_t *cls_a = nullptr, *cls_tmp = nullptr, *meta = nullptr;
  uint64_t *ent_a = nullptr, *ent_tmp = nullptr;
  uint32_t *cls_b = nullptr;
  uint64_t *ent_b = nullptr;
  const int64_t nmax = std::max(n_a, n_b);
  FH_TRY(scr.alloc(&cls_a, n_a));
  FH_TRY(scr.alloc(&ent_a, n_a));
  FH_TRY(scr.alloc(&cls_tmp, nmax));
  FH_TRY(scr.alloc(&ent_tmp, nmax));
  // meta: hist A [66], hist B [66], present A [8], present B [8]
  constexpr int kMeta = 2 * kNumClasses + 16 + 2;
  FH_TRY(scr.alloc(&meta, kMeta));
  FH_TRY(fill_u32(meta, 0, kMeta, s));
  uint32_t* hist_a = meta;
  uint32_t* hist_b = meta + kNumClasses;
  uint32_t* pres_a = meta + 2 * kNumClasses;
  uint32_t* pres_b = pres_a + 8;
  uint32_t* maxl_a = pres_b + 8;
  uint32_t* maxl_b = maxl_a + 1;
  auto classify = [&](const uint8_t* ch, const int64_t* off, int64_t n, uint32_t* cls, uint64_t* ent, uint32_t* hist,
                      uint32_t* pres, uint32_t* maxl) -> int {
    int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 4096);
    hipLaunchKernelGGL(k_classify, dim3((unsigned)blocks), dim3(256), 0, s, ch, off, n, cls, ent, hist, pres, maxl);
    FH_HIP(hipGetLastError());
    return radix_sort_u32(cls, cls_tmp, ent, ent_tmp, n, 0, 7, s);
  };
  FH_TRY(classify(chars_a, offsets_a, n_a, cls_a, ent_a, hist_a, pres_a, maxl_a));
  if (two) {
    FH_TRY(scr.alloc(&cls_b, n_b));
    FH_TRY(scr.alloc(&ent_b, n_b));
    FH_TRY(classify(chars_b, offsets_b, n_b, cls_b, ent_b, hist_b, pres_b, maxl_b));
  } else {
    ent_b = ent_a;
  }
  std::vector<uint32_t> hmeta(kMeta);
  FH_HIP(hipMemcpyAsync(hmeta.data(), meta, kMeta * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  const uint32_t* ha = hmeta.data();
  const uint32_t* hb = two ? hmeta.data() + kNumClasses : ha;
  const uint32_t* pa = hmeta.data() + 2 * kNumClasses;
  const uint32_t* pb = two ? pa + 8 : pa;

  {
    const uint32_t ml = std::max(hmeta[2 * kNumClasses + 16], two ? hmeta[2 * kNumClasses + 17] : 0u);
    if (ml > (uint32_t)kMaxGenericLen) {
      set_error("fh_jw_levels: strings of %u bytes exceed the supported maximum of %d", ml, kMaxGenericLen);
      return FH_ERR_UNSUPPORTED;
    }
  }
  const float lo = lower_thr, hi = upper_thr;
  const bool empty_special = host_reaches(0.0f, lo, hi);

  // Alphabet: bytes present on both sides get codes 0..K-1; row K is the
  // all-blocked row used by bytes that cannot match.
  std::vector<uint8_t> amap(256);
  int K = 0;
  for (int b = 0; b < 256; ++b) {
    const bool both = ((pa[b >> 5] >> (b & 31)) & 1u) && ((pb[b >> 5] >> (b & 31)) & 1u);
    if (both) ++K;
  }
  const int krows = K + 1;
  {
    int k = 0;
    for (int b = 0; b < 256; ++b) {
      const bool both = ((pa[b >> 5] >> (b & 31)) & 1u) && ((pb[b >> 5] >> (b & 31)) & 1u);
      amap[b] = both ? (uint8_t)k++ : (uint8_t)K;
    }
  }
  // Tile width: the widest of 256/128/64/32 dwords whose LDS tile fits.
  int tw = 256;
  const size_t qbytes = (size_t)kWaves * kQCap * sizeof(uint64_t);
  auto lds_for = [&](int w) { return ((size_t)krows * (w + 1) + 1) * 4 + qbytes; };
  while (tw > 32 && lds_for(tw) > 64 * 1024) tw /= 2;
  const size_t lds_bytes = lds_for(tw);
  FH_REQUIRE(lds_bytes <= 64 * 1024, "fh_jw_levels: alphabet of %d symbols too large", K);

  // Class starts in the sorted order.
  int64_t start_a[kNumClasses + 1], start_b[kNumClasses + 1];
  start_a[0] = start_b[0] = 0;
  for (int c = 0; c < kNumClasses; ++c) {
    start_a[c + 1] = start_a[c] + ha[c];
    start_b[c + 1] = start_b[c] + hb[c];
  }
  // Pruning table over the lengths present.
  std::vector<uint8_t> mmin((kLongLen + 1) * (kLongLen + 1), 255);
  for (int la = 1; la <= kLongLen; ++la) {
    if (!ha[la]) continue;
    for (int lb = 1; lb <= kLongLen; ++lb) {
      if (!hb[lb]) continue;
      mmin[la * (kLongLen + 1) + lb] = host_mmin(la, lb, p, lo, hi);
    }
  }
  // Plan: waves per A length class, tiles per B length class.
  Plan pl;
  for (int r = 0; r < 3; ++r) pl.wave_lo[r] = pl.wave_hi[r] = pl.tile_lo[r] = pl.tile_hi[r] = 0;
  for (int r = 0; r < 3; ++r) {
    pl.wave_lo[r] = (int64_t)pl.waves.size();
    for (int c = 1; c <= kLongLen; ++c) {
      if (lrange_of(c) != r) continue;
      for (int64_t s0 = start_a[c]; s0 < start_a[c + 1]; s0 += 64) {
        WaveMeta wm;
        wm.len = c;
        wm.count = (int32_t)std::min<int64_t>(64, start_a[c + 1] - s0);
        wm.s0 = s0;
        pl.waves.push_back(wm);
      }
    }
    pl.wave_hi[r] = (int64_t)pl.waves.size();
  }
  for (int k = 0; k < 3; ++k) {
    pl.tile_lo[k] = (int64_t)pl.tiles.size();
    for (int c = 1; c <= kLongLen; ++c) {
      if (lrange_of(c) != k) continue;
      const int per = k == kP16 ? 2 * tw : (k == kW32 ? tw : tw / 2);
      for (int64_t s0 = start_b[c]; s0 < start_b[c + 1]; s0 += per) {
        TileMeta tm;
        tm.len = c;
        tm.count = (int32_t)std::min<int64_t>(per, start_b[c + 1] - s0);
        tm.s0 = s0;
        pl.tiles.push_back(tm);
      }
    }
    pl.tile_hi[k] = (int64_t)pl.tiles.size();
  }
  // Work accounting (host arithmetic over the length histograms).
  if (stats) {
    stats->alphabet = K;
    double kern = 0, pruned = 0, ops = 0;
    for (int la = 1; la <= kLongLen; ++la)
      for (int lb = 1; lb <= kLongLen; ++lb) {
        double pairs;
        if (two)
          pairs = (double)ha[la] * (double)hb[lb];
        else if (la > lb)
          pairs = (double)ha[la] * (double)ha[lb];
        else if (la == lb)
          pairs = (double)ha[la] * (double)(ha[la] - 1) / 2.0;
        else
          pairs = 0;
        if (pairs == 0) continue;
        const int mm = mmin[la * (kLongLen + 1) + lb];
        if (mm > std::min(la, lb)) {
          pruned += pairs;
          continue;
        }
        kern += pairs;
        const double per_pair = lb <= 16 ? (3.0 * la + 3.0) / 2.0 : (lb <= 32 ? 3.0 * la + 3.0 : 6.0 * la + 4.0);
        ops += pairs * per_pair;
      }
    stats->pairs_kernel = (int64_t)kern;
    stats->pairs_pruned = (int64_t)pruned;
    stats->alg_lane_ops = ops;
  }

  // Upload plan, tables.
  WaveMeta* d_waves = nullptr;
  TileMeta* d_tiles = nullptr;
  uint8_t* d_small = nullptr;  // amap (256) + mmin (65 * 65)
  FH_TRY(scr.alloc(&d_waves, std::max<size_t>(1, pl.waves.size())));
  FH_TRY(scr.alloc(&d_tiles, std::max<size_t>(1, pl.tiles.size())));
  FH_TRY(scr.alloc(&d_small, 256 + mmin.size()));
  if (!pl.waves.empty())
    FH_HIP(hipMemcpyAsync(d_waves, pl.waves.data(), pl.waves.size() * sizeof(WaveMeta), hipMemcpyHostToDevice, s));
  if (!pl.tiles.empty())
    FH_HIP(hipMemcpyAsync(d_tiles, pl.tiles.data(), pl.tiles.size() * sizeof(TileMeta), hipMemcpyHostToDevice, s));
  FH_HIP(hipMemcpyAsync(d_small, amap.data(), 256, hipMemcpyHostToDevice, s));
  FH_HIP(hipMemcpyAsync(d_small + 256, mmin.data(), mmin.size(), hipMemcpyHostToDevice, s));
  uint32_t* d_table = nullptr;
  const int64_t ntiles = (int64_t)pl.tiles.size();
  FH_TRY(scr.alloc(&d_table, (size_t)std::max<int64_t>(1, ntiles) * krows * tw));
  if (ntiles > 0) {
    const int64_t total = ntiles * tw;
    const int64_t bx = std::min<int64_t>(ceil_div(total, 256), 65535);
    hipLaunchKernelGGL(k_build_table, dim3((unsigned)bx, (unsigned)krows), dim3(256), 0, s, chars_b, offsets_b,
                       ent_b, d_tiles, ntiles, d_small, krows, tw, d_table);
    FH_HIP(hipGetLastError());
  }
  unsigned long long* counter = nullptr;
  FH_TRY(scr.alloc(&counter, 1));
  FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(counter), 0, 1, s));

  JwArgs args;
  args.chars_a = chars_a;
  args.off_a = offsets_a;
  args.sorted_a = ent_a;
  args.chars_b = chars_b;
  args.off_b = offsets_b;
  args.sorted_b = ent_b;
  args.waves = d_waves;
  args.tiles = d_tiles;
  args.table = d_table;
  args.amap = d_small;
  args.mmin = d_small + 256;
  args.krows = krows;
  args.tw = tw;
  args.mode = mode;
  args.n_b = n_b;
  args.p = p;
  args.lo = lo;
  args.hi = hi;
  args.hits = hits;
  args.cap = capacity;
  args.counter = counter;
  args.krows_max = krows;
  args.hkeys = nullptr;
  args.hcodes = nullptr;
  args.hmask = 0;
  args.code_stride = 0;
  args.overflow = nullptr;

  JwArgs* d_args = nullptr;
  FH_TRY(scr.alloc(&d_args, 1));
  FH_HIP(hipMemcpyAsync(d_args, &args, sizeof(JwArgs), hipMemcpyHostToDevice, s));
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (stats) {
    FH_HIP(hipEventCreate(&ev0));
    FH_HIP(hipEventCreate(&ev1));
    FH_HIP(hipEventRecord(ev0, s));
  }
  int launches = 0;
  FH_TRY((launch_kinds<1, 16>(d_args, lds_bytes, pl, 0, s, &launches)));
  FH_TRY((launch_kinds<17, 32>(d_args, lds_bytes, pl, 1, s, &launches)));
  FH_TRY((launch_kinds<33, 64>(d_args, lds_bytes, pl, 2, s, &launches)));

  // Generic pairs: special strings (longer than 64 bytes, and empty strings
  // when a score of 0 reaches level 1).
  {
    const int64_t nsa_long = start_a[kClassLong + 1] - start_a[kClassLong];
    const int64_t nsb_long = start_b[kClassLong + 1] - start_b[kClassLong];
    auto gen = [&](int64_t sp0, int64_t nsp, int64_t n_other, int swap) -> int {
      if (nsp <= 0 || n_other <= 0) return FH_OK;
      const int64_t gx = std::min<int64_t>(ceil_div(n_other, 256), 4096);
      for (int64_t r0 = 0; r0 < nsp; r0 += 65535) {
        const int64_t gy = std::min<int64_t>(65535, nsp - r0);
        hipLaunchKernelGGL(k_jw_generic, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, args, sp0, nsp, n_other,
                           swap, (int)empty_special, r0);
        FH_HIP(hipGetLastError());
        ++launches;
      }
      return FH_OK;
    };
    double gpairs = 0;
    // Empty class first (sorted index range [0, hist[0])), then the long class.
    if (empty_special) {
      FH_TRY(gen(start_a[0], ha[0], n_b, 0));
      gpairs += (double)ha[0] * n_b;
    }
    FH_TRY(gen(start_a[kClassLong], nsa_long, n_b, 0));
    gpairs += (double)nsa_long * n_b;
    if (two) {
      if (empty_special) {
        FH_TRY(gen(start_b[0], hb[0], n_a, 1));
        gpairs += (double)hb[0] * n_a;
      }
      FH_TRY(gen(start_b[kClassLong], nsb_long, n_a, 1));
      gpairs += (double)nsb_long * n_a;
    }
    if (stats) stats->pairs_generic = (int64_t)gpairs;
  }
  if (stats) {
    FH_HIP(hipEventRecord(ev1, s));
  }
  uint64_t h_count = 0;
  FH_HIP(hipMemcpyAsync(&h_count, counter, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  if (stats) {
    float ms = 0;
    FH_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    stats->kernel_ms = ms;
    stats->launches = launches;
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
  }
  *n_hits = (int64_t)h_count;
  return FH_OK;
}

// ===========================================================================
// Batched blocked linkage: every (variable, block, wave group, tile chunk)
// work item of every block in one persistent launch.
//
// The reference links blocks by running one Comparison per block and summing
// the Counts (docs/source/usage.rst:34-76). Here rows carry a block id, the
// plan sorts each side by (block, length), builds waves and tiles inside
// blocks, and lists work items largest first; a persistent kernel pops them
// from an atomic queue. Hits go straight into an open-addressing table keyed
// by pair id whose value accumulates the pattern code over the variables; a
// second kernel adds the exact-variable digits (row-id lookups) and counts
// (block, pattern). The sequence is captured once into a hipGraph and
// replayed.
// ===========================================================================
namespace fh {
namespace {

constexpr int kMaxBlkVars = 8;
constexpr int kMaxBlkExact = 8;

struct WorkItem {
  int32_t var;
  int32_t nwaves;
  int64_t wave0;
  int64_t tile0;
  int64_t ntiles;
  double cost;
};

template <int LMIN, int LMAX>
__global__ void __launch_bounds__(kThreads) k_jw_blocked(const JwArgs* __restrict__ vars,
                                                        const WorkItem* __restrict__ items, int64_t n_items,
                                                        unsigned int* head) {
  extern __shared__ uint32_t lds[];
  __shared__ int64_t cur;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) cur = (int64_t)atomicAdd(head, 1u);
    __syncthreads();
    const int64_t it = cur;
    if (it >= n_items) break;  // same value in every wave: the whole workgroup exits
    const WorkItem w = items[it];
    run_item<LMIN, LMAX, kMixed>(vars[w.var], lds, w.wave0, w.wave0 + w.nwaves, w.tile0, w.tile0 + w.ntiles);
  }
}

struct ExactRowIds {
  const int64_t* a[kMaxBlkExact];
  const int64_t* b[kMaxBlkExact];
  uint32_t stride[kMaxBlkExact];
  int n;
};

__global__ void __launch_bounds__(256) k_blocked_hist(const unsigned long long* __restrict__ hkeys,
                                                      const uint32_t* __restrict__ hcodes, uint64_t nslots,
                                                      uint64_t n_b, const int32_t* __restrict__ block_a,
                                                      ExactRowIds ex, uint64_t n_patterns,
                                                      unsigned long long* __restrict__ hist,
                                                      unsigned long long* __restrict__ n_union) {
  unsigned long long mine = 0;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nslots; s += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long key = hkeys[s];
    if (key == kEmptyKey) continue;
    const uint64_t i = key / n_b, j = key % n_b;
    uint32_t code = hcodes[s];
    for (int q = 0; q < ex.n; ++q)
      if (ex.a[q][i] == ex.b[q][j]) code += ex.stride[q];
    atomicAdd(hist + (uint64_t)block_a[i] * n_patterns + code, 1ull);
    ++mine;
  }
  for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o, 64);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(n_union, mine);
}

// Sort keys for one side: block * kNumClasses + length class; presence bits.
__global__ void __launch_bounds__(256) k_blk_classify(const uint8_t* __restrict__ chars, const int64_t* __restrict__ off,
                                                      int64_t n, const int32_t* __restrict__ block,
                                                      uint32_t* __restrict__ key, uint64_t* __restrict__ ent,
                                                      uint32_t* __restrict__ hist, uint32_t* __restrict__ present,
                                                      uint32_t* __restrict__ maxlen) {
  __shared__ uint32_t pr[8];
  if (threadIdx.x < 8) pr[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = off[e];
    const int64_t len = off[e + 1] - o;
    const uint32_t c = len > kLongLen ? (uint32_t)kClassLong : (uint32_t)len;
    const uint32_t k = (uint32_t)block[e] * kNumClasses + c;
    key[e] = k;
    ent[e] = (uint64_t)e;
    atomicAdd(hist + k, 1u);
    if (len > kLongLen) atomicMax(maxlen, len > 0xFFFFFFFFll ? 0xFFFFFFFFu : (uint32_t)len);
    if (len >= 1 && len <= kLongLen)
      for (int64_t i = 0; i < len; ++i) {
        const uint32_t b = chars[o + i];
        atomicOr(&pr[b >> 5], 1u << (b & 31));
      }
  }
  __syncthreads();
  if (threadIdx.x < 8 && pr[threadIdx.x]) atomicOr(&present[threadIdx.x], pr[threadIdx.x]);
}

struct BlockedPlan {
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int n_vars = 0;
  int64_t n_a = 0, n_b = 0;
  int32_t n_blocks = 0;
  uint64_t n_patterns = 0;
  int tw = 256;
  size_t lds_bytes = 0;
  int grid = 0;
  // per variable
  std::vector<void*> bufs;  // every device allocation of the plan
  std::vector<JwArgs> args;
  std::vector<const uint8_t*> chars_b;
  std::vector<const int64_t*> off_b;
  std::vector<uint64_t*> sorted_b;
  std::vector<TileMeta*> tiles_dev;
  std::vector<int64_t> ntiles;
  std::vector<uint8_t*> amap_dev;
  std::vector<int> krows;
  std::vector<uint32_t*> table;
  JwArgs* args_dev = nullptr;
  WorkItem* items_dev[2] = {nullptr, nullptr};
  int64_t n_items[2] = {0, 0};
  unsigned int* heads = nullptr;
  unsigned long long* hkeys = nullptr;
  uint32_t* hcodes = nullptr;
  uint64_t hslots = 0;
  unsigned int* overflow = nullptr;
  unsigned long long* hist = nullptr;
  unsigned long long* n_union = nullptr;
  const int32_t* block_a = nullptr;
  ExactRowIds ex{};
  double pairs_kernel = 0, alg_ops = 0;

  template <typename T>
  int alloc(T** out, size_t count) {
    size_t bytes = std::max<size_t>(256, count * sizeof(T));
    void* ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess) {
      set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
      return FH_ERR_HIP;
    }
    bufs.push_back(ptr);
    *out = static_cast<T*>(ptr);
    return FH_OK;
  }

  ~BlockedPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* ptr : bufs) (void)hipFree(ptr);
    bufs.clear();
    for (hipEvent_t e : {ev_in, ev_out, ev_t0, ev_t1})
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // The device work of one run, enqueued on `stream` (capturable: fill
  // kernels only, no allocation, no host synchronisation).
  int enqueue() {
    FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(hkeys), kEmptyKey, (int64_t)hslots, stream));
    FH_TRY(fill_u32(hcodes, 0, (int64_t)hslots, stream));
    FH_TRY(fill_u32(heads, 0, 2, stream));
    FH_TRY(fill_u32(overflow, 0, 1, stream));
    FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(hist), 0, (int64_t)(n_patterns * (uint64_t)n_blocks), stream));
    FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(n_union), 0, 1, stream));
    for (int v = 0; v < n_vars; ++v) {
      if (ntiles[v] == 0) continue;
      const int64_t total = ntiles[v] * tw;
      const int64_t bx = std::min<int64_t>(ceil_div(total, 256), 65535);
      hipLaunchKernelGGL(k_build_table, dim3((unsigned)bx, (unsigned)krows[v]), dim3(256), 0, stream, chars_b[v],
                         off_b[v], sorted_b[v], tiles_dev[v], ntiles[v], amap_dev[v], krows[v], tw, table[v]);
      FH_HIP(hipGetLastError());
    }
    if (n_items[0] > 0) {
      hipLaunchKernelGGL((k_jw_blocked<1, 32>), dim3((unsigned)grid), dim3(kThreads), lds_bytes, stream, args_dev,
                         items_dev[0], n_items[0], heads);
      FH_HIP(hipGetLastError());
    }
    if (n_items[1] > 0) {
      hipLaunchKernelGGL((k_jw_blocked<33, 64>), dim3((unsigned)grid), dim3(kThreads), lds_bytes, stream, args_dev,
                         items_dev[1], n_items[1], heads + 1);
      FH_HIP(hipGetLastError());
    }
    const int64_t hb = std::min<int64_t>(ceil_div((int64_t)hslots, 256), 16384);
    hipLaunchKernelGGL(k_blocked_hist, dim3((unsigned)hb), dim3(256), 0, stream, hkeys, hcodes, hslots, (uint64_t)n_b,
                       block_a, ex, n_patterns, hist, n_union);
    FH_HIP(hipGetLastError());
    return FH_OK;
  }
};

}  // namespace
}  // namespace fh

extern "C" int fh_blocked_plan_create(int n_fuzzy, const uint8_t* const* chars_a, const int64_t* const* offsets_a,
                                      int64_t n_a, const uint8_t* const* chars_b, const int64_t* const* offsets_b,
                                      int64_t n_b, const int32_t* block_a, const int32_t* block_b, int32_t n_blocks,
                                      int n_exact, const int64_t* const* ids_a, const int64_t* const* ids_b, float p,
                                      float lower_thr, float upper_thr, const fh_blocked_tuning* tuning, void** plan_out,
                                      fh_stream_t stream) {
  using namespace fh;
  clear_error();
  FH_REQUIRE(plan_out != nullptr, "fh_blocked_plan_create: plan is null");
  *plan_out = nullptr;
  FH_REQUIRE(n_fuzzy >= 1 && n_fuzzy <= kMaxBlkVars, "fh_blocked_plan_create: 1..%d fuzzy variables", kMaxBlkVars);
  FH_REQUIRE(n_exact >= 0 && n_exact <= kMaxBlkExact, "fh_blocked_plan_create: 0..%d exact variables", kMaxBlkExact);
  FH_REQUIRE(n_a > 0 && n_b > 0 && n_a < (1ll << 31) && n_b < (1ll << 31), "fh_blocked_plan_create: bad sizes");
  FH_REQUIRE(n_blocks >= 1 && (int64_t)n_blocks * kNumClasses < (1ll << 31), "fh_blocked_plan_create: bad block count");
  fh_blocked_tuning tu;
  tu.tile_width = 256;
  tu.tiles_per_item = 4;
  tu.waves_per_item = kWaves;
  tu.wgs_per_cu = 2;
  tu.hash_slots = 0;
  if (tuning) tu = *tuning;
  FH_REQUIRE(tu.tile_width == 32 || tu.tile_width == 64 || tu.tile_width == 128 || tu.tile_width == 256,
             "fh_blocked_plan_create: tile_width must be 32, 64, 128 or 256");
  FH_REQUIRE(tu.tiles_per_item >= 1 && tu.waves_per_item >= 1 && tu.waves_per_item <= kWaves && tu.wgs_per_cu >= 1,
             "fh_blocked_plan_create: bad tuning");
  hipStream_t caller = (hipStream_t)stream;
  auto pl = std::make_unique<BlockedPlan>();
  pl->n_vars = n_fuzzy;
  pl->n_a = n_a;
  pl->n_b = n_b;
  pl->n_blocks = n_blocks;
  pl->tw = tu.tile_width;
  FH_HIP(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking));
  FH_HIP(hipEventCreateWithFlags(&pl->ev_in, hipEventDisableTiming));
  FH_HIP(hipEventCreateWithFlags(&pl->ev_out, hipEventDisableTiming));
  FH_HIP(hipEventCreate(&pl->ev_t0));
  FH_HIP(hipEventCreate(&pl->ev_t1));
  FH_HIP(hipEventRecord(pl->ev_in, caller));
  FH_HIP(hipStreamWaitEvent(pl->stream, pl->ev_in, 0));
  hipStream_t s = pl->stream;

  const float lo = lower_thr, hi = upper_thr;
  if (host_reaches(0.0f, lo, hi)) {
    set_error("fh_blocked_plan_create: thresholds at or below 0 make every pair a hit; use the plain path");
    return FH_ERR_UNSUPPORTED;
  }
  uint64_t n_patterns = 1;
  for (int v = 0; v < n_fuzzy; ++v) n_patterns *= 3;
  n_patterns <<= n_exact;
  pl->n_patterns = n_patterns;
  {
    uint32_t st = 1;
    for (int q = n_exact - 1; q >= 0; --q) {
      pl->ex.a[q] = ids_a[q];
      pl->ex.b[q] = ids_b[q];
      pl->ex.stride[q] = st;
      st *= 2;
    }
    pl->ex.n = n_exact;
  }
  pl->block_a = block_a;
  const int nkeys_bits = bits_for((uint64_t)n_blocks * kNumClasses);
  const int64_t nk = (int64_t)n_blocks * kNumClasses;
  std::vector<uint8_t> mmin((kLongLen + 1) * (kLongLen + 1), 255);
  std::vector<bool> mm_done((kLongLen + 1) * (kLongLen + 1), false);
  std::vector<WorkItem> items[2];
  int max_krows = 1;
  pl->args.resize(n_fuzzy);

  // Per-variable scratch for the sort (temporary) and per-variable persistent
  // buffers (sorted rows, waves, tiles, tables).
  for (int v = 0; v < n_fuzzy; ++v) {
    Scratch scr(s);
    uint32_t *key_a, *key_b, *key_tmp, *meta;
    uint64_t *ent_a, *ent_b, *ent_tmp;
    uint32_t *hist_a, *hist_b;
    FH_TRY(pl->alloc(&ent_a, (size_t)n_a));
    FH_TRY(pl->alloc(&ent_b, (size_t)n_b));
    FH_TRY(scr.alloc(&key_a, (size_t)n_a));
    FH_TRY(scr.alloc(&key_b, (size_t)n_b));
    FH_TRY(scr.alloc(&key_tmp, (size_t)std::max(n_a, n_b)));
    FH_TRY(scr.alloc(&ent_tmp, (size_t)std::max(n_a, n_b)));
    FH_TRY(scr.alloc(&hist_a, (size_t)nk));
    FH_TRY(scr.alloc(&hist_b, (size_t)nk));
    FH_TRY(scr.alloc(&meta, 32));
    FH_TRY(fill_u32(hist_a, 0, nk, s));
    FH_TRY(fill_u32(hist_b, 0, nk, s));
    FH_TRY(fill_u32(meta, 0, 32, s));
    const int64_t ga = std::min<int64_t>(ceil_div(n_a, 256), 4096), gb = std::min<int64_t>(ceil_div(n_b, 256), 4096);
    hipLaunchKernelGGL(k_blk_classify, dim3((unsigned)ga), dim3(256), 0, s, chars_a[v], offsets_a[v], n_a, block_a,
                       key_a, ent_a, hist_a, meta, meta + 16);
    FH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_blk_classify, dim3((unsigned)gb), dim3(256), 0, s, chars_b[v], offsets_b[v], n_b, block_b,
                       key_b, ent_b, hist_b, meta + 8, meta + 17);
    FH_HIP(hipGetLastError());
    FH_TRY(radix_sort_u32(key_a, key_tmp, ent_a, ent_tmp, n_a, 0, nkeys_bits, s));
    FH_TRY(radix_sort_u32(key_b, key_tmp, ent_b, ent_tmp, n_b, 0, nkeys_bits, s));
    std::vector<uint32_t> ha(nk), hb(nk), hm(32);
    FH_HIP(hipMemcpyAsync(ha.data(), hist_a, nk * 4, hipMemcpyDeviceToHost, s));
    FH_HIP(hipMemcpyAsync(hb.data(), hist_b, nk * 4, hipMemcpyDeviceToHost, s));
    FH_HIP(hipMemcpyAsync(hm.data(), meta, 32 * 4, hipMemcpyDeviceToHost, s));
    FH_HIP(hipStreamSynchronize(s));
    if (hm[16] > 0 || hm[17] > 0) {
      set_error("fh_blocked_plan_create: strings longer than %d bytes need the plain path", kLongLen);
      return FH_ERR_UNSUPPORTED;
    }
    // Alphabet of this variable.
    std::vector<uint8_t> amap(256);
    int K = 0;
    for (int b = 0; b < 256; ++b)
      if (((hm[b >> 5] >> (b & 31)) & 1u) && ((hm[8 + (b >> 5)] >> (b & 31)) & 1u)) ++K;
    const int krows = K + 1;
    for (int b = 0, k = 0; b < 256; ++b) {
      const bool both = ((hm[b >> 5] >> (b & 31)) & 1u) && ((hm[8 + (b >> 5)] >> (b & 31)) & 1u);
      amap[b] = both ? (uint8_t)k++ : (uint8_t)K;
    }
    max_krows = std::max(max_krows, krows);
    // Waves and tiles per (block, length class), in sorted order.
    std::vector<WaveMeta> waves;
    std::vector<TileMeta> tiles;
    std::vector<int64_t> wave_lo(n_blocks + 1), tile_lo(n_blocks + 1);
    std::vector<int64_t> wave_short_end(n_blocks);
    int64_t sa = 0, sb = 0;
    for (int32_t b = 0; b < n_blocks; ++b) {
      wave_lo[b] = (int64_t)waves.size();
      tile_lo[b] = (int64_t)tiles.size();
      for (int c = 0; c < kNumClasses; ++c) {
        const int64_t ca = ha[(int64_t)b * kNumClasses + c], cb = hb[(int64_t)b * kNumClasses + c];
        if (c == 33) wave_short_end[b] = (int64_t)waves.size();
        if (c >= 1 && c <= kLongLen) {
          for (int64_t s0 = sa; s0 < sa + ca; s0 += 64) {
            WaveMeta wm;
            wm.len = c;
            wm.count = (int32_t)std::min<int64_t>(64, sa + ca - s0);
            wm.s0 = s0;
            waves.push_back(wm);
          }
          const int per = c <= 16 ? 2 * pl->tw : (c <= 32 ? pl->tw : pl->tw / 2);
          for (int64_t s0 = sb; s0 < sb + cb; s0 += per) {
            TileMeta tm;
            tm.len = c;
            tm.count = (int32_t)std::min<int64_t>(per, sb + cb - s0);
            tm.s0 = s0;
            tiles.push_back(tm);
          }
          for (int lb = 1; lb <= kLongLen && ca > 0; ++lb) {
            if (hb[(int64_t)b * kNumClasses + lb] == 0) continue;
            const int idx = c * (kLongLen + 1) + lb;
            if (!mm_done[idx]) {
              mmin[idx] = host_mmin(c, lb, p, lo, hi);
              mm_done[idx] = true;
            }
            const double pairs = (double)ca * (double)hb[(int64_t)b * kNumClasses + lb];
            if (mmin[idx] <= std::min(c, lb)) {
              pl->pairs_kernel += pairs;
              pl->alg_ops += pairs * (lb <= 16 ? (3.0 * c + 3.0) / 2.0 : (lb <= 32 ? 3.0 * c + 3.0 : 6.0 * c + 4.0));
            }
          }
        }
        sa += ca;
        sb += cb;
      }
      if (wave_short_end[b] < wave_lo[b]) wave_short_end[b] = wave_lo[b];
    }
    wave_lo[n_blocks] = (int64_t)waves.size();
    tile_lo[n_blocks] = (int64_t)tiles.size();
    // Work items: per block, groups of waves (short: 1..32 bytes; long:
    // 33..64) x chunks of tiles; cost ~ lanes x strings x length.
    for (int32_t b = 0; b < n_blocks; ++b) {
      const int64_t t0 = tile_lo[b], t1 = tile_lo[b + 1];
      if (t1 == t0) continue;
      double tile_cost = 0;
      for (int64_t t = t0; t < t1; ++t) tile_cost += tiles[t].count;
      for (int part = 0; part < 2; ++part) {
        const int64_t w0 = part == 0 ? wave_lo[b] : wave_short_end[b];
        const int64_t w1 = part == 0 ? wave_short_end[b] : wave_lo[b + 1];
        for (int64_t w = w0; w < w1; w += tu.waves_per_item) {
          const int nw = (int)std::min<int64_t>(tu.waves_per_item, w1 - w);
          double lanes = 0;
          for (int k = 0; k < nw; ++k) lanes += (double)waves[w + k].len;
          for (int64_t t = t0; t < t1; t += tu.tiles_per_item) {
            WorkItem it;
            it.var = v;
            it.nwaves = nw;
            it.wave0 = w;
            it.tile0 = t;
            it.ntiles = std::min<int64_t>(tu.tiles_per_item, t1 - t);
            it.cost = lanes * tile_cost * (double)it.ntiles / (double)(t1 - t0);
            items[part].push_back(it);
          }
        }
      }
    }
    // Upload this variable's plan.
    WaveMeta* d_waves;
    TileMeta* d_tiles;
    uint8_t* d_small;
    uint32_t* d_table;
    FH_TRY(pl->alloc(&d_waves, std::max<size_t>(1, waves.size())));
    FH_TRY(pl->alloc(&d_tiles, std::max<size_t>(1, tiles.size())));
    FH_TRY(pl->alloc(&d_small, 256 + mmin.size()));
    FH_TRY(pl->alloc(&d_table, std::max<size_t>(1, tiles.size()) * krows * pl->tw));
    if (!waves.empty())
      FH_HIP(hipMemcpyAsync(d_waves, waves.data(), waves.size() * sizeof(WaveMeta), hipMemcpyHostToDevice, s));
    if (!tiles.empty())
      FH_HIP(hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(TileMeta), hipMemcpyHostToDevice, s));
    FH_HIP(hipMemcpyAsync(d_small, amap.data(), 256, hipMemcpyHostToDevice, s));
    FH_HIP(hipStreamSynchronize(s));
    JwArgs& a = pl->args[v];
    a.chars_a = chars_a[v];
    a.off_a = offsets_a[v];
    a.sorted_a = ent_a;
    a.chars_b = chars_b[v];
    a.off_b = offsets_b[v];
    a.sorted_b = ent_b;
    a.waves = d_waves;
    a.tiles = d_tiles;
    a.table = d_table;
    a.amap = d_small;
    a.mmin = d_small + 256;  // filled once every variable's lengths are known
    a.krows = krows;
    a.tw = pl->tw;
    a.mode = FH_MODE_LINK;
    a.n_b = n_b;
    a.p = p;
    a.lo = lo;
    a.hi = hi;
    a.hits = nullptr;
    a.cap = 0;
    a.counter = nullptr;
    a.code_stride = 1;
    for (int q = v + 1; q < n_fuzzy; ++q) a.code_stride *= 3;
    a.code_stride <<= n_exact;
    pl->chars_b.push_back(chars_b[v]);
    pl->off_b.push_back(offsets_b[v]);
    pl->sorted_b.push_back(ent_b);
    pl->tiles_dev.push_back(d_tiles);
    pl->ntiles.push_back((int64_t)tiles.size());
    pl->amap_dev.push_back(d_small);
    pl->krows.push_back(krows);
    pl->table.push_back(d_table);
  }
  // Pruning table (shared by the variables), hash table, counters.
  for (int v = 0; v < n_fuzzy; ++v)
    FH_HIP(hipMemcpyAsync(pl->amap_dev[v] + 256, mmin.data(), mmin.size(), hipMemcpyHostToDevice, s));
  uint64_t want = tu.hash_slots > 0 ? (uint64_t)tu.hash_slots : (uint64_t)std::max<int64_t>(1 << 22, 2 * (n_a + n_b));
  uint64_t slots = 1;
  while (slots < want) slots <<= 1;
  pl->hslots = slots;
  FH_TRY(pl->alloc(&pl->hkeys, slots));
  FH_TRY(pl->alloc(&pl->hcodes, slots));
  FH_TRY(pl->alloc(&pl->heads, 2));
  FH_TRY(pl->alloc(&pl->overflow, 1));
  FH_TRY(pl->alloc(&pl->hist, (size_t)(n_patterns * (uint64_t)n_blocks)));
  FH_TRY(pl->alloc(&pl->n_union, 1));
  for (auto& a : pl->args) {
    a.hkeys = pl->hkeys;
    a.hcodes = pl->hcodes;
    a.hmask = slots - 1;
    a.overflow = pl->overflow;
    a.krows_max = max_krows;
  }
  FH_TRY(pl->alloc(&pl->args_dev, pl->args.size()));
  FH_HIP(hipMemcpyAsync(pl->args_dev, pl->args.data(), pl->args.size() * sizeof(JwArgs), hipMemcpyHostToDevice, s));
  for (int part = 0; part < 2; ++part) {
    std::stable_sort(items[part].begin(), items[part].end(),
                     [](const WorkItem& x, const WorkItem& y) { return x.cost > y.cost; });
    pl->n_items[part] = (int64_t)items[part].size();
    FH_TRY(pl->alloc(&pl->items_dev[part], std::max<size_t>(1, items[part].size())));
    if (!items[part].empty())
      FH_HIP(hipMemcpyAsync(pl->items_dev[part], items[part].data(), items[part].size() * sizeof(WorkItem),
                            hipMemcpyHostToDevice, s));
  }
  const size_t qbytes = (size_t)kWaves * kQCap * sizeof(uint64_t);
  pl->lds_bytes = ((size_t)max_krows * (pl->tw + 1) + 1) * 4 + qbytes;
  FH_REQUIRE(pl->lds_bytes <= 64 * 1024, "fh_blocked_plan_create: alphabet too large for tile width %d", pl->tw);
  int dev = 0, cus = 0;
  FH_HIP(hipGetDevice(&dev));
  FH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  pl->grid = std::max(1, cus * tu.wgs_per_cu);
  FH_HIP(hipStreamSynchronize(s));
  *plan_out = pl.release();
  return FH_OK;
}

extern "C" int fh_blocked_run(void* plan, int use_graph, fh_stream_t stream) {
  using namespace fh;
  clear_error();
  FH_REQUIRE(plan != nullptr, "fh_blocked_run: plan is null");
  BlockedPlan* pl = static_cast<BlockedPlan*>(plan);
  hipStream_t caller = (hipStream_t)stream;
  FH_HIP(hipEventRecord(pl->ev_in, caller));
  FH_HIP(hipStreamWaitEvent(pl->stream, pl->ev_in, 0));
  FH_HIP(hipEventRecord(pl->ev_t0, pl->stream));
  if (use_graph) {
    if (!pl->exec) {
      FH_HIP(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal));
      const int st = pl->enqueue();
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(pl->stream, &g);
      if (st != FH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      FH_HIP(e);
      pl->graph = g;
      FH_HIP(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
    }
    FH_HIP(hipGraphLaunch(pl->exec, pl->stream));
  } else {
    FH_TRY(pl->enqueue());
  }
  FH_HIP(hipEventRecord(pl->ev_t1, pl->stream));
  FH_HIP(hipEventRecord(pl->ev_out, pl->stream));
  FH_HIP(hipStreamWaitEvent(caller, pl->ev_out, 0));
  return FH_OK;
}

extern "C" int fh_blocked_counts(void* plan, int64_t* hist, int64_t* n_union, int* overflow, double* ms) {
  using namespace fh;
  clear_error();
  FH_REQUIRE(plan != nullptr, "fh_blocked_counts: plan is null");
  BlockedPlan* pl = static_cast<BlockedPlan*>(plan);
  FH_HIP(hipStreamSynchronize(pl->stream));
  if (hist)
    FH_HIP(hipMemcpy(hist, pl->hist, pl->n_patterns * (uint64_t)pl->n_blocks * sizeof(int64_t), hipMemcpyDeviceToHost));
  uint64_t nu = 0;
  unsigned int ov = 0;
  FH_HIP(hipMemcpy(&nu, pl->n_union, sizeof(nu), hipMemcpyDeviceToHost));
  FH_HIP(hipMemcpy(&ov, pl->overflow, sizeof(ov), hipMemcpyDeviceToHost));
  if (n_union) *n_union = (int64_t)nu;
  if (overflow) *overflow = (int)ov;
  if (ms) {
    float t = 0;
    FH_HIP(hipEventElapsedTime(&t, pl->ev_t0, pl->ev_t1));
    *ms = t;
  }
  return FH_OK;
}

extern "C" int fh_blocked_info(void* plan, double* info) {
  using namespace fh;
  clear_error();
  FH_REQUIRE(plan != nullptr && info != nullptr, "fh_blocked_info: null argument");
  BlockedPlan* pl = static_cast<BlockedPlan*>(plan);
  info[0] = (double)pl->n_items[0];
  info[1] = (double)pl->n_items[1];
  info[2] = (double)pl->n_patterns;
  info[3] = (double)pl->hslots;
  info[4] = pl->pairs_kernel;
  info[5] = pl->alg_ops;
  info[6] = (double)pl->grid;
  info[7] = (double)pl->lds_bytes;
  return FH_OK;
}

extern "C" int fh_blocked_plan_destroy(void* plan) {
  fh::clear_error();
  delete static_cast<fh::BlockedPlan*>(plan);
  return FH_OK;
}
