// All-pairs Jaro-Winkler levels: plan and entry point (fh_jw_levels).
//
// Replaces the reference's comparison kernel (faster/comparison.py:133-186,
// deduplication copy faster/deduplication.py:132-196) and its thresholding
// and compaction (faster/comparison.py:321-338). The comparison kernel and
// its design are in fh_jw_core.h.
#include <array>
#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "fh_jw_core.h"
#include "fh_jw_prefix.h"

namespace fh {
namespace {
std::atomic<int> g_jw_method{FH_JW_METHOD_PREFIX};
// Survivors per kernel pair of the prefix form: the largest rate seen, decayed
// by 5% per call (sizes the next call's candidate buffer; any value is
// correct, a large enough one avoids repeated rounds).
std::atomic<double> g_prefix_rate{4e-4};


// Candidate buffers of the prefix form, kept between calls: a call leases
// one (concurrent calls on other streams lease others), grows it when its
// estimate or its count needs more, and returns it. They hold device memory
// until the process ends; growing one frees the old allocation.
struct CandBuffer {
  uint64_t* ptr = nullptr;
  int64_t cap = 0;
  int dev = -1;
};
std::mutex g_cand_mu;
std::vector<CandBuffer> g_cand_pool;
constexpr int64_t kCandMax = (int64_t)1 << 30;  // entries (8 GiB)

class CandLease {
 public:
  CandLease() = default;
  CandLease(const CandLease&) = delete;
  CandLease& operator=(const CandLease&) = delete;
  ~CandLease() {
    if (!held_) return;
    std::lock_guard<std::mutex> lock(g_cand_mu);
    g_cand_pool.push_back(buf_);
  }
  // Takes a buffer of this device from the pool (the smallest that holds
  // `need` entries, else the largest, else an empty one) and makes it hold at
  // least `need` entries. Taking the best fit keeps concurrent callers with
  // different needs from trading buffers and reallocating them (hipFree
  // waits for the whole device).
  int acquire(int64_t need, hipStream_t s) {
    int dev = 0;
    FH_HIP(hipGetDevice(&dev));
    {
      std::lock_guard<std::mutex> lock(g_cand_mu);
      int best = -1;
      for (size_t i = 0; i < g_cand_pool.size(); ++i) {
        const CandBuffer& b = g_cand_pool[i];
        if (b.dev != dev) continue;
        if (best < 0) {
          best = (int)i;
          continue;
        }
        const CandBuffer& c = g_cand_pool[(size_t)best];
        const bool b_fits = b.cap >= need, c_fits = c.cap >= need;
        if (b_fits != c_fits ? b_fits : (b_fits ? b.cap < c.cap : b.cap > c.cap)) best = (int)i;
      }
      if (best >= 0) {
        buf_ = g_cand_pool[(size_t)best];
        g_cand_pool.erase(g_cand_pool.begin() + best);
      }
    }
    buf_.dev = dev;
    held_ = true;
    return grow(need, s);
  }
  // At least `need` entries (at most kCandMax); earlier work on s that used
  // the buffer is waited for before it is freed.
  int grow(int64_t need, hipStream_t s) {
    need = std::min(std::max<int64_t>(need, 1 << 20), kCandMax);
    if (buf_.cap >= need) return FH_OK;
    const int64_t cap = std::min(std::max(need, 2 * buf_.cap), kCandMax);
    if (buf_.ptr) {
      FH_HIP(hipStreamSynchronize(s));
      FH_HIP(hipFree(buf_.ptr));
      buf_.ptr = nullptr;
      buf_.cap = 0;
    }
    hipError_t e = hipMalloc(&buf_.ptr, (size_t)cap * sizeof(uint64_t));
    if (e != hipSuccess) {
      buf_.ptr = nullptr;
      set_error("fh_jw_levels: hipMalloc of the candidate buffer (%lld entries) failed: %s", (long long)cap,
                hipGetErrorString(e));
      return FH_ERR_HIP;
    }
    buf_.cap = cap;
    return FH_OK;
  }
  uint64_t* ptr() const { return buf_.ptr; }
  int64_t cap() const { return buf_.cap; }

 private:
  CandBuffer buf_;
  bool held_ = false;
};
}  // namespace
const std::vector<uint8_t>& mmin_table(float p, float lo, float hi) {
  static std::mutex mu;
  static std::map<std::array<uint32_t, 3>, std::unique_ptr<std::vector<uint8_t>>> cache;
  std::array<uint32_t, 3> key;
  std::memcpy(&key[0], &p, 4);
  std::memcpy(&key[1], &lo, 4);
  std::memcpy(&key[2], &hi, 4);
  std::lock_guard<std::mutex> lock(mu);
  auto& slot = cache[key];
  if (!slot) {
    auto t = std::make_unique<std::vector<uint8_t>>((kLongLen + 1) * (kLongLen + 1), 255);
    for (int la = 1; la <= kLongLen; ++la)
      for (int lb = 1; lb <= kLongLen; ++lb) (*t)[la * (kLongLen + 1) + lb] = host_mmin(la, lb, p, lo, hi);
    slot = std::move(t);
  }
  return *slot;
}
}  // namespace fh

extern "C" int fh_jw_set_method(int method) {
  fh::clear_error();
  FH_REQUIRE(method == FH_JW_METHOD_TILES || method == FH_JW_METHOD_PREFIX,
             "fh_jw_set_method: bad method %d", method);
  fh::g_jw_method.store(method);
  return FH_OK;
}

extern "C" int fh_jw_get_method(void) { return fh::g_jw_method.load(); }

namespace fh {
namespace {

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
    if (la > kMaxGenericLen || lb > kMaxGenericLen) continue;  // k_jw_huge
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

// The pairs of k_jw_generic's row sets that have a value over
// kMaxGenericLen bytes, scored with their flags in global scratch
// (jw_exact_mem): a grid-stride loop over the (row, other) pairs, every
// thread owning words [k * nthreads + tid] of `scratch`.
__global__ void __launch_bounds__(256) k_jw_huge(JwArgs a, int64_t sp0, int64_t nsp, int64_t n_other, int swap,
                                                 uint64_t* __restrict__ scratch) {
  const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t* flags = scratch + tid;
  const int64_t total = nsp * n_other;
  for (int64_t q = tid; q < total; q += nthr) {
    const int64_t r = q / n_other, x = q - r * n_other;
    uint64_t ea, eb;
    if (!swap) {
      ea = a.sorted_a[sp0 + r];
      eb = (uint64_t)x;
    } else {
      eb = a.sorted_b[sp0 + r];
      ea = (uint64_t)x;
    }
    const int64_t oa = a.off_a[ea];
    const int64_t la = a.off_a[ea + 1] - oa;
    const int64_t ob = a.off_b[eb];
    const int64_t lb = a.off_b[eb + 1] - ob;
    if (la <= kMaxGenericLen && lb <= kMaxGenericLen) continue;
    // The same pair ownership as k_jw_generic (empty strings never reach
    // here: a pair with a value this long has a non-empty member).
    if (a.mode == FH_MODE_DEDUP) {
      if (ea == eb) continue;
      if (lb > kLongLen && ea < eb) continue;
    } else if (swap) {
      if (la > kLongLen) continue;
    }
    const uint8_t* sa = a.chars_a + oa;
    const uint8_t* sb = a.chars_b + ob;
    uint64_t xx = ea, yy = eb;
    float s;
    if (a.mode == FH_MODE_DEDUP && !lex_greater(sa, (int)la, sb, (int)lb)) {
      s = jw_exact_mem(sb, (int)lb, sa, (int)la, a.p, flags, nthr);
      xx = eb;
      yy = ea;
    } else {
      s = jw_exact_mem(sa, (int)la, sb, (int)lb, a.p, flags, nthr);
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

struct Plan {
  std::vector<WaveMeta> waves;
  std::vector<TileMeta> tiles;
  // Wave ranges per A length range (1..16, 17..32, 33..64) and tile ranges per kind.
  int64_t wave_lo[3], wave_hi[3];
  int64_t tile_lo[kNumKinds], tile_hi[kNumKinds];
};

static int lrange_of(int len) { return len <= 16 ? 0 : (len <= 32 ? 1 : 2); }

// Which (string-1 length range, tile kind) launches launch_kinds makes: all
// with the per-lane kernel; only those the prefix form covers (string-1
// values of up to 64 bytes against the packed and 32-bit kinds), with it;
// or the rest, with the per-lane kernel.
enum LaunchSet { kLaneAll, kPrefixOnly, kLaneRest };

// Launches the waves of length range r that lie in [u0, u1) against the
// tile kinds of `set` (LDS bytes: lds for the per-lane kernel, lds_prefix
// for the prefix form).
static int launch_kinds(const JwArgs* args, size_t lds, const Plan& pl, int r, int64_t u0, int64_t u1, hipStream_t s,
                        int* launches, LaunchSet set = kLaneAll, size_t lds_prefix = 0) {
  const int64_t w0 = std::max(pl.wave_lo[r], u0), nw = std::min(pl.wave_hi[r], u1) - w0;
  if (nw <= 0) return FH_OK;
  for (int k = 0; k < kNumKinds; ++k) {
    const int64_t t0 = pl.tile_lo[k], nt = pl.tile_hi[k] - pl.tile_lo[k];
    if (nt <= 0) continue;
    const bool covered = k != kW64;
    if (set == kPrefixOnly) {
      if (!covered) continue;
      ++*launches;
      if (r == 0) FH_TRY(launch_jw_prefix_r0(k, args, lds_prefix, w0, nw, t0, nt, s));
      if (r == 1) FH_TRY(launch_jw_prefix_r1(k, args, lds_prefix, w0, nw, t0, nt, s));
      if (r == 2) FH_TRY(launch_jw_prefix_r2(k, args, lds_prefix, w0, nw, t0, nt, s));
      continue;
    }
    if (set == kLaneRest && covered) continue;
    ++*launches;
    if (r == 0) FH_TRY(launch_jw_tiles_r0(k, args, lds, w0, nw, t0, nt, s));
    if (r == 1) FH_TRY(launch_jw_tiles_r1(k, args, lds, w0, nw, t0, nt, s));
    if (r == 2) FH_TRY(launch_jw_tiles_r2(k, args, lds, w0, nw, t0, nt, s));
  }
  return FH_OK;
}

// Row codes of the sorted string-1 values [s_begin, s_end) (all of at most
// 64 bytes) for the prefix form.
static int launch_prefix_codes(const uint8_t* chars, const int64_t* off, const uint64_t* sorted, int64_t s_begin,
                               int64_t s_end, const uint8_t* amap, int krows, int64_t hi0, int64_t hi_base,
                               uint32_t* codes, hipStream_t s) {
  if (s_end <= s_begin) return FH_OK;
  const int64_t blocks = std::min<int64_t>(ceil_div((s_end - s_begin) * 16, 256), 8192);
  hipLaunchKernelGGL(k_prefix_codes, dim3((unsigned)blocks), dim3(256), 0, s, chars, off, sorted, s_begin, s_end, amap,
                     krows, hi0, hi_base, codes);
  FH_HIP(hipGetLastError());
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

  // Longest value (values over kMaxGenericLen bytes are scored by k_jw_huge).
  const uint32_t max_len = std::max(hmeta[2 * kNumClasses + 16], two ? hmeta[2 * kNumClasses + 17] : 0u);
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
  // Pruning table (every length pair; cached per thresholds).
  const std::vector<uint8_t>& mmin = mmin_table(p, lo, hi);
  // Plan: waves per A length class, tiles per B length class.
  Plan pl;
  for (int r = 0; r < 3; ++r) pl.wave_lo[r] = pl.wave_hi[r] = 0;
  for (int k = 0; k < kNumKinds; ++k) pl.tile_lo[k] = pl.tile_hi[k] = 0;
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
  const int method = g_jw_method.load();
  // The prefix form needs 256-dword tiles (alphabets of up to about 55 symbols).
  const bool prefix = method == FH_JW_METHOD_PREFIX && tw == kPrefixTileWidth;
  const size_t lds_prefix = jw_prefix_lds_bytes(krows);
  // The prefix form's plans put string-2 values of 1..5 bytes in kP5 tiles.
  for (int k = 0; k < kNumKinds; ++k) {
    pl.tile_lo[k] = (int64_t)pl.tiles.size();
    for (int c = 1; c <= kLongLen; ++c) {
      if (kind_for(c, prefix) != k) continue;
      const int per = kind_tile_strings(k, tw);
      for (int64_t s0 = start_b[c]; s0 < start_b[c + 1]; s0 += per) {
        TileMeta tm;
        tm.len = c;
        tm.count = (int32_t)std::min<int64_t>(per, start_b[c + 1] - s0);
        tm.s0 = s0;
        tm.kind = k;
        pl.tiles.push_back(tm);
      }
    }
    pl.tile_hi[k] = (int64_t)pl.tiles.size();
  }
  // Work accounting (host arithmetic over the length histograms).
  if (stats) {
    stats->alphabet = K;
    double kern = 0, pruned = 0, ops = 0, frozen = 0;
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
        ops += pairs * kind_pair_ops(kind_for(lb, prefix), la);
        frozen += pairs * (8.0 * la + 24.0);
      }
    stats->pairs_kernel = (int64_t)kern;
    stats->pairs_pruned = (int64_t)pruned;
    stats->alg_lane_ops = ops;
    stats->frozen_model_ops = frozen;
    stats->method = prefix ? FH_JW_METHOD_PREFIX : FH_JW_METHOD_TILES;  // the form the plan runs
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
  uint64_t* d_cand = nullptr;  // candidate pairs of one round (prefix form)
  CandLease cand_lease;        // the prefix form's buffer, kept between calls
  unsigned long long* cand_count = nullptr;
  int64_t cand_cap = (int64_t)1 << 27;
  int64_t* d_cls = nullptr;
  // The prefix form's candidate pairs go to one buffer per round and are
  // scored by a separate kernel from the row codes of both values (the
  // lengths from the class starts); the per-lane form scores its survivors
  // in the kernel and finds a string-2 value's table column from its sorted
  // index through the class starts and first tile of each class.
  //
  // Rounds run over string-1 units (waves), each sized so that its
  // candidates fit the buffer at the candidate rate seen so far; a round that
  // overflows is split in two and redone (its comparison pass is repeated,
  // nothing is lost). The prefix form sizes its buffer from the survivor rate
  // of the previous calls.
  int64_t unit_lo = 0, unit_hi = (int64_t)pl.waves.size();
  std::vector<double> unit_pairs;  // prefix sums of kernel pairs per unit
  double rate = 4e-3;              // candidates per kernel pair, updated after every round
  const bool rounds = prefix;
  if (rounds) {
    std::vector<double> per_len(kLongLen + 1, 0.0);
    for (int la = 1; la <= kLongLen; ++la)
      for (int lb = 1; lb <= 32; ++lb) {  // the prefix form covers string-2 values of up to 32 bytes
        if (!hb[lb] || mmin[la * (kLongLen + 1) + lb] > std::min(la, lb)) continue;
        if (!two && lb > la) continue;  // deduplication: string 2 never longer than string 1
        per_len[la] += (!two && lb == la) ? 0.5 * (double)hb[lb] : (double)hb[lb];
      }
    unit_lo = pl.wave_lo[0];
    unit_hi = pl.wave_hi[2];
    unit_pairs.assign((size_t)(unit_hi - unit_lo) + 1, 0.0);
    for (int64_t u = unit_lo; u < unit_hi; ++u)
      unit_pairs[u - unit_lo + 1] = unit_pairs[u - unit_lo] + (double)pl.waves[u].count * per_len[pl.waves[u].len];
    rate = g_prefix_rate.load();
    // The first round's budget (0.7 of the buffer at this rate) then covers all units.
    FH_TRY(cand_lease.acquire((int64_t)std::min(1.5 * rate * unit_pairs.back(), (double)kCandMax), s));
    d_cand = cand_lease.ptr();
    cand_cap = cand_lease.cap();
    FH_TRY(scr.alloc(&cand_count, 1));
  }
  {
    std::vector<int64_t> cls(3 * kNumClasses, 0);
    for (int c = 0; c < kNumClasses; ++c) {
      cls[c] = start_b[c];
      cls[2 * kNumClasses + c] = start_a[c];
    }
    for (int64_t t = (int64_t)pl.tiles.size() - 1; t >= 0; --t) cls[kNumClasses + pl.tiles[t].len] = t;
    FH_TRY(scr.alloc(&d_cls, cls.size()));
    FH_HIP(hipMemcpyAsync(d_cls, cls.data(), cls.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
  }
  FH_TRY(launch_build_table(chars_b, offsets_b, ent_b, d_tiles, ntiles, d_small, krows, tw, d_table, s));
  uint32_t* d_codes = nullptr;
  uint32_t* d_codes_b = nullptr;
  int64_t codes_hi0 = 0, codes_hi_base = 0;
  unsigned long long* d_skipped = nullptr;  // per string-1 length: steps skipped per tile pass (stats)
  if (prefix) {
    // Values of 1..64 bytes occupy sorted indices [start_a[1], start_a[65]).
    const int64_t c1 = start_a[1], c2 = start_a[kPrefixMaxLen + 1];
    const int64_t hi0 = start_a[33], hi_base = std::max<int64_t>(1, n_a) * kPrefixCodeWords;
    FH_TRY(scr.alloc(&d_codes, (size_t)(hi_base + std::max<int64_t>(0, c2 - hi0) * kPrefixCodeWords)));
    FH_TRY(launch_prefix_codes(chars_a, offsets_a, ent_a, c1, c2, d_small, krows, hi0, hi_base, d_codes, s));
    codes_hi0 = hi0;
    codes_hi_base = hi_base;
    // String 2's codes for the survivor scoring (values of 1..32 bytes; in
    // deduplication both sides are one sorted set and share codes_a).
    if (mode == FH_MODE_DEDUP) {
      d_codes_b = d_codes;
    } else {
      const int64_t b1 = start_b[1], b2 = start_b[33];
      FH_TRY(scr.alloc(&d_codes_b, (size_t)std::max<int64_t>(1, n_b) * kPrefixCodeWords));
      FH_TRY(launch_prefix_codes(chars_b, offsets_b, ent_b, b1, b2, d_small, krows, b2,
                                 std::max<int64_t>(1, n_b) * kPrefixCodeWords, d_codes_b, s));
    }
    if (stats && c2 > c1) {
      FH_TRY(scr.alloc(&d_skipped, kPrefixMaxLen + 1));
      FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(d_skipped), 0, kPrefixMaxLen + 1, s));
      const int64_t blocks = std::min<int64_t>(ceil_div(c2 - c1, 256), 4096);
      hipLaunchKernelGGL(k_prefix_restarts, dim3((unsigned)blocks), dim3(256), 0, s, d_codes,
                         (const int64_t*)(d_cls + 2 * kNumClasses), c1, c2, d_skipped);
      FH_HIP(hipGetLastError());
    }
  }
  unsigned long long* counter = nullptr;  // [0] hits, [1] candidates
  FH_TRY(scr.alloc(&counter, 2));
  FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(counter), 0, 2, s));

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
  args.p5 = prefix ? 1 : 0;
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
  args.n_candidates = counter + 1;  // survivors the per-lane pass scores
  args.cls_start = d_cls;
  args.cls_tile0 = d_cls ? d_cls + kNumClasses : nullptr;
  args.cls_start_a = d_cls ? d_cls + 2 * kNumClasses : nullptr;
  args.cand = rounds ? d_cand : nullptr;
  args.cand_cap = cand_cap;
  args.cand_count = cand_count;
  args.codes_a = d_codes;
  args.codes_hi0 = codes_hi0;
  args.codes_hi_base = codes_hi_base;
  args.codes_b = d_codes_b;

  // d_args: the arguments of the prefix passes and of the scoring
  // kernel; d_args_lane (no candidate buffer): the per-lane kernel, which
  // scores its survivors itself.
  JwArgs* d_args = nullptr;
  FH_TRY(scr.alloc(&d_args, 2));
  JwArgs* d_args_lane = d_args + 1;
  FH_HIP(hipMemcpyAsync(d_args, &args, sizeof(JwArgs), hipMemcpyHostToDevice, s));
  JwArgs args_lane = args;
  args_lane.cand = nullptr;
  FH_HIP(hipMemcpyAsync(d_args_lane, &args_lane, sizeof(JwArgs), hipMemcpyHostToDevice, s));
  // Timings for stats: events recorded around the launches and read only
  // after the final synchronisation, so that asking for stats adds no wait.
  struct Events {
    std::vector<hipEvent_t> v;
    ~Events() {
      for (hipEvent_t e : v) (void)hipEventDestroy(e);
    }
  } events;
  std::vector<std::pair<std::pair<size_t, size_t>, double*>> spans;  // (start, end) -> accumulator
  auto mark = [&]() -> size_t {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return SIZE_MAX;
    events.v.push_back(e);
    (void)hipEventRecord(e, s);
    return events.v.size() - 1;
  };
  size_t ev0 = SIZE_MAX, ev1 = SIZE_MAX;
  if (stats) ev0 = mark();
  int launches = 0;
  int64_t n_candidates = 0;
  {
    // The per-lane kernel: everything, or with the prefix form what it does
    // not cover (string-2 values longer than 32 bytes: kW64 tiles).
    const size_t e0 = stats ? mark() : SIZE_MAX;
    const LaunchSet set = prefix ? kLaneRest : kLaneAll;
    for (int r = 0; r < 3; ++r)
      FH_TRY(launch_kinds(d_args_lane, lds_bytes, pl, r, 0, (int64_t)pl.waves.size(), s, &launches, set,
                          lds_prefix));
    if (stats) spans.push_back({{e0, mark()}, &stats->pass_ms});
  }
  if (rounds) {
    const int64_t nunits = unit_hi - unit_lo;
    std::vector<std::pair<int64_t, int64_t>> todo;  // unit ranges relative to unit_lo
    auto next_round = [&](int64_t u0) {
      const double budget = 0.7 * (double)cand_cap / rate;
      int64_t u1 = u0 + 1;
      while (u1 < nunits && unit_pairs[u1 + 1] - unit_pairs[u0] <= budget) ++u1;
      return u1;
    };
    double prefix_pairs = 0, prefix_got = 0;
    int64_t u0 = 0;
    while (u0 < nunits || !todo.empty()) {
      int64_t a0, a1;
      if (!todo.empty()) {
        a0 = todo.back().first;
        a1 = todo.back().second;
        todo.pop_back();
      } else {
        a0 = u0;
        a1 = next_round(u0);
        u0 = a1;
      }
      FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(cand_count), 0, 1, s));
      const size_t e0 = stats ? mark() : SIZE_MAX;
      for (int r = 0; r < 3; ++r)
        FH_TRY(launch_kinds(d_args, lds_bytes, pl, r, unit_lo + a0, unit_lo + a1, s, &launches, kPrefixOnly,
                            lds_prefix));
      if (stats) spans.push_back({{e0, mark()}, &stats->pass_ms});
      uint64_t got = 0;
      FH_HIP(hipMemcpyAsync(&got, cand_count, sizeof(got), hipMemcpyDeviceToHost, s));
      FH_HIP(hipStreamSynchronize(s));
      const double pairs = unit_pairs[a1] - unit_pairs[a0];
      if (pairs > 0) rate = std::max(1e-6, 1.25 * (double)got / pairs);
      if ((int64_t)got > cand_cap) {
        // A larger buffer (kept for later calls) and the same round again,
        // or the round split in two when the buffer is at its maximum.
        FH_TRY(cand_lease.grow((int64_t)((double)got * 1.1) + 1, s));
        FH_REQUIRE(cand_lease.cap() >= (int64_t)got || a1 - a0 > 1,
                   "fh_jw_levels: %llu candidates in one string-1 unit exceed the buffer", (unsigned long long)got);
        if (cand_lease.cap() < (int64_t)got) {
          const int64_t mid = a0 + (a1 - a0) / 2;
          todo.push_back({mid, a1});
          todo.push_back({a0, mid});
        } else {
          todo.push_back({a0, a1});
        }
        d_cand = cand_lease.ptr();
        cand_cap = cand_lease.cap();
        args.cand = d_cand;
        args.cand_cap = cand_cap;
        FH_HIP(hipMemcpyAsync(d_args, &args, sizeof(JwArgs), hipMemcpyHostToDevice, s));
        continue;
      }
      prefix_pairs += pairs;
      prefix_got += (double)got;
      n_candidates += (int64_t)got;
      const size_t e2 = stats ? mark() : SIZE_MAX;
      FH_TRY(launch_eval_prefix(d_args, (int64_t)got, s));
      ++launches;
      if (stats) spans.push_back({{e2, mark()}, &stats->eval_ms});
    }
    if (prefix_pairs > 0)
      g_prefix_rate.store(std::max({1e-6, prefix_got / prefix_pairs, 0.95 * g_prefix_rate.load()}));
  }

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
    // Values over kMaxGenericLen bytes: their flags in scratch, about 1 GiB
    // at most (fewer threads for longer values), 256 threads at least.
    if (max_len > (uint32_t)kMaxGenericLen) {
      const int64_t words = 2 * (((int64_t)max_len + 63) / 64);
      int64_t threads = std::min<int64_t>(65536, ((int64_t)1 << 27) / words);
      threads = std::max<int64_t>(256, threads / 256 * 256);
      uint64_t* scratch = nullptr;
      FH_TRY(scr.alloc(&scratch, (size_t)(threads * words)));
      auto huge = [&](int64_t sp0, int64_t nsp, int64_t n_other, int swap) -> int {
        if (nsp <= 0 || n_other <= 0) return FH_OK;
        hipLaunchKernelGGL(k_jw_huge, dim3((unsigned)(threads / 256)), dim3(256), 0, s, args, sp0, nsp, n_other, swap,
                           scratch);
        FH_HIP(hipGetLastError());
        ++launches;
        return FH_OK;
      };
      FH_TRY(huge(start_a[kClassLong], nsa_long, n_b, 0));
      if (two) FH_TRY(huge(start_b[kClassLong], nsb_long, n_a, 1));
    }
  }
  if (stats) ev1 = mark();
  uint64_t h_count[2] = {0, 0};
  FH_HIP(hipMemcpyAsync(h_count, counter, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  std::vector<unsigned long long> h_skipped(kPrefixMaxLen + 1, 0);
  if (d_skipped)
    FH_HIP(hipMemcpyAsync(h_skipped.data(), d_skipped, h_skipped.size() * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  if (stats && d_skipped) {
    // Each skipped step saves three ops on every string-2 dword the value
    // meets: 3 / L lane-ops per pair of a kind with L strings per dword.
    double skipped_ops = 0;
    for (int la = 1; la <= kPrefixMaxLen; ++la) {
      if (!h_skipped[la]) continue;
      double per_step = 0;
      for (int lb = 1; lb <= 32; ++lb) {
        if (!hb[lb] || mmin[la * (kLongLen + 1) + lb] > std::min(la, lb)) continue;
        if (!two && lb > la) continue;
        const int k = kind_for(lb, prefix);
        const double nb = (!two && lb == la) ? 0.5 * (double)hb[lb] : (double)hb[lb];
        per_step += nb * 3.0 / (packed_kind(k) ? kind_lanes(k) : 1);
      }
      skipped_ops += (double)h_skipped[la] * per_step;
    }
    stats->prefix_skipped_ops = skipped_ops;
  }
  if (stats) stats->pairs_candidates = n_candidates + (int64_t)h_count[1];
  if (stats) {
    // Every event is complete after the synchronisation above.
    auto elapsed = [&](size_t a, size_t b) -> double {
      float ms = 0;
      if (a == SIZE_MAX || b == SIZE_MAX || hipEventElapsedTime(&ms, events.v[a], events.v[b]) != hipSuccess)
        return 0.0;
      return (double)ms;
    };
    for (const auto& sp : spans) *sp.second += elapsed(sp.first.first, sp.first.second);
    stats->kernel_ms = elapsed(ev0, ev1);
    stats->launches = launches;
  }
  *n_hits = (int64_t)h_count[0];
  return FH_OK;
}

