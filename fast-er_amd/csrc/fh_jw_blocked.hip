// Batched blocked linkage (fh_blocked_*): the comparison of every block in
// one persistent, graph-captured launch. The comparison kernel is in
// fh_jw_core.h.
#include <memory>

#define FH_JW_PAIR_TABLE 1
#include "fh_jw_core.h"

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
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(FH_JW_WAVES_PER_EU))) k_jw_blocked(const JwArgs* __restrict__ vars,
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

// Pairs of one block with a value over kLongLen bytes, which the
// bit-parallel passes do not cover: sorted side-A indices [i0, i0 + ni)
// against sorted side-B indices [j0, j0 + nj) of variable `var` (A's long
// values against all of the block's B values, and A's other values against
// B's long ones).
struct LongItem {
  int32_t var;
  int32_t ni;
  int64_t i0;
  int64_t j0;
  int64_t nj;
};

// Scores the long-value pairs of every item (the reference's function, with
// its flags in registers up to kMaxGenericLen bytes and in the scratch
// words [k * nthreads + tid] beyond) and adds their levels to the pair table.
// Empty values pair at level 0 (the blocked plan requires thresholds above
// 0).
__global__ void __launch_bounds__(256) k_blk_long(const JwArgs* __restrict__ vars, const LongItem* __restrict__ items,
                                                  int64_t n_items, uint64_t* __restrict__ scratch) {
  const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
  uint64_t* flags = scratch + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
    const LongItem li = items[it];
    const JwArgs& a = vars[li.var];
    const int64_t total = (int64_t)li.ni * li.nj;
    for (int64_t q = threadIdx.x; q < total; q += blockDim.x) {
      const int64_t r = q / li.nj;
      const uint64_t ea = a.sorted_a[li.i0 + r], eb = a.sorted_b[li.j0 + (q - r * li.nj)];
      const int64_t oa = a.off_a[ea], ob = a.off_b[eb];
      const int64_t la = a.off_a[ea + 1] - oa, lb = a.off_b[eb + 1] - ob;
      if (la == 0 || lb == 0) continue;
      const float sc = (la <= kMaxGenericLen && lb <= kMaxGenericLen)
                           ? jw_exact<kMaxGenericLen / 64>(a.chars_a + oa, (int)la, a.chars_b + ob, (int)lb, a.p)
                           : jw_exact_mem(a.chars_a + oa, (int)la, a.chars_b + ob, (int)lb, a.p, flags, nthr);
      emit_level(a, ea, eb, level_of(sc, a.lo, a.hi));
    }
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
  mine = fhd::block_reduce(mine, fhd::OpSum());
  if (threadIdx.x == 0 && mine) atomicAdd(n_union, mine);
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
  // The same work items grouped by block, for the per-block launch loop:
  // block b's items of part p are items_blk_dev[p][items_blk_lo[p][b] ..
  // items_blk_lo[p][b + 1]); its tiles of variable v are tile_blk_lo[v][b ..
  // b + 1]; heads_blk holds one queue head per (block, part).
  WorkItem* items_blk_dev[2] = {nullptr, nullptr};
  std::vector<int64_t> items_blk_lo[2];
  std::vector<std::vector<int64_t>> tile_blk_lo;
  unsigned int* heads_blk = nullptr;
  // Long-value pairs (k_blk_long): items, a persistent grid and its scratch.
  LongItem* long_dev = nullptr;
  int64_t n_long = 0;
  int long_grid = 0;
  uint64_t* long_scratch = nullptr;
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
      FH_TRY(launch_build_table(chars_b[v], off_b[v], sorted_b[v], tiles_dev[v], ntiles[v], amap_dev[v], krows[v], tw,
                                table[v], stream));
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
    FH_TRY(enqueue_long());
    return enqueue_hist();
  }

  int enqueue_long() {
    if (n_long == 0) return FH_OK;
    hipLaunchKernelGGL(k_blk_long, dim3((unsigned)long_grid), dim3(256), 0, stream, args_dev, long_dev, n_long,
                       long_scratch);
    FH_HIP(hipGetLastError());
    return FH_OK;
  }

  int enqueue_hist() {
    const int64_t hb = std::min<int64_t>(ceil_div((int64_t)hslots, 256), 2048);
    hipLaunchKernelGGL(k_blocked_hist, dim3((unsigned)hb), dim3(256), 0, stream, hkeys, hcodes, hslots, (uint64_t)n_b,
                       block_a, ex, n_patterns, hist, n_union);
    FH_HIP(hipGetLastError());
    return FH_OK;
  }

  // The plain per-block loop over the same tables and kernels: block after
  // block, one table build per variable over the block's tiles and one
  // comparison launch per string-length part over the block's work items
  // (grid capped at the block's item count), all on the plan's stream with
  // no host synchronisation; one histogram pass at the end.
  int enqueue_per_block() {
    FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(hkeys), kEmptyKey, (int64_t)hslots, stream));
    FH_TRY(fill_u32(hcodes, 0, (int64_t)hslots, stream));
    FH_TRY(fill_u32(heads_blk, 0, 2 * (int64_t)n_blocks, stream));
    FH_TRY(fill_u32(overflow, 0, 1, stream));
    FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(hist), 0, (int64_t)(n_patterns * (uint64_t)n_blocks), stream));
    FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(n_union), 0, 1, stream));
    for (int32_t b = 0; b < n_blocks; ++b) {
      for (int v = 0; v < n_vars; ++v) {
        const int64_t t0 = tile_blk_lo[v][b], t1 = tile_blk_lo[v][b + 1];
        FH_TRY(launch_build_table(chars_b[v], off_b[v], sorted_b[v], tiles_dev[v] + t0, t1 - t0, amap_dev[v], krows[v],
                                  tw, table[v] + (size_t)t0 * (size_t)krows[v] * (size_t)tw, stream));
      }
      for (int part = 0; part < 2; ++part) {
        const int64_t i0 = items_blk_lo[part][b], ni = items_blk_lo[part][b + 1] - i0;
        if (ni <= 0) continue;
        const unsigned g = (unsigned)std::min<int64_t>(grid, ni);
        if (part == 0)
          hipLaunchKernelGGL((k_jw_blocked<1, 32>), dim3(g), dim3(kThreads), lds_bytes, stream, args_dev,
                             items_blk_dev[0] + i0, ni, heads_blk + 2 * (int64_t)b);
        else
          hipLaunchKernelGGL((k_jw_blocked<33, 64>), dim3(g), dim3(kThreads), lds_bytes, stream, args_dev,
                             items_blk_dev[1] + i0, ni, heads_blk + 2 * (int64_t)b + 1);
        FH_HIP(hipGetLastError());
      }
    }
    // The long-value pairs of every block in one launch.
    FH_TRY(enqueue_long());
    return enqueue_hist();
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
  const std::vector<uint8_t>& mmin = mmin_table(p, lo, hi);  // pruning table, cached per thresholds
  std::vector<WorkItem> items[2];
  std::vector<LongItem> long_items;
  uint32_t max_len = 0;
  std::vector<int32_t> item_keys[2];  // tile kind * 256 + string-1 length of each item
  std::vector<int32_t> item_block[2];
  int max_krows = 1;
  pl->args.resize(n_fuzzy);

  // First, per variable: both sides sorted by (block, length class) into
  // persistent row orders, their class histograms and the variable's
  // alphabet. The tile width follows from the largest alphabet; then, per
  // variable, the waves, tiles, work items and tables.
  struct VarSort {
    uint64_t* ent_a = nullptr;
    uint64_t* ent_b = nullptr;
    std::vector<uint32_t> ha, hb;
    std::vector<uint8_t> amap;
    int krows = 1;
  };
  std::vector<VarSort> sorts(n_fuzzy);
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
    max_len = std::max({max_len, hm[16], hm[17]});
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
    VarSort& vs = sorts[v];
    vs.ent_a = ent_a;
    vs.ent_b = ent_b;
    vs.ha.swap(ha);
    vs.hb.swap(hb);
    vs.amap.swap(amap);
    vs.krows = krows;
  }
  // Tile width: the widest power of two up to the requested one whose tile
  // rows (every variable's alphabet) and survivor queues fit 64 KB of LDS.
  const size_t qbytes = (size_t)kWaves * kQCap * sizeof(uint64_t);
  auto lds_for = [&](int w) { return ((size_t)max_krows * (w + 1) + 1) * 4 + qbytes; };
  while (pl->tw > 32 && lds_for(pl->tw) > 64 * 1024) pl->tw /= 2;
  pl->lds_bytes = lds_for(pl->tw);
  FH_REQUIRE(pl->lds_bytes <= 64 * 1024, "fh_blocked_plan_create: alphabet of %d symbols too large", max_krows - 1);
  for (int v = 0; v < n_fuzzy; ++v) {
    uint64_t* const ent_a = sorts[v].ent_a;
    uint64_t* const ent_b = sorts[v].ent_b;
    const std::vector<uint32_t>& ha = sorts[v].ha;
    const std::vector<uint32_t>& hb = sorts[v].hb;
    const std::vector<uint8_t>& amap = sorts[v].amap;
    const int krows = sorts[v].krows;
    // Waves and tiles per (block, length class), in sorted order.
    std::vector<WaveMeta> waves;
    std::vector<TileMeta> tiles;
    std::vector<int64_t> wave_lo(n_blocks + 1), tile_lo(n_blocks + 1);
    std::vector<int64_t> wave_short_end(n_blocks);
    int64_t sa = 0, sb = 0;
    for (int32_t b = 0; b < n_blocks; ++b) {
      wave_lo[b] = (int64_t)waves.size();
      tile_lo[b] = (int64_t)tiles.size();
      {
        // Pairs with a long value (class kClassLong, the last of the block):
        // A's long values x all B values, A's other values x B's long ones.
        int64_t na_blk = 0, nb_blk = 0;
        for (int c = 0; c < kNumClasses; ++c) {
          na_blk += ha[(int64_t)b * kNumClasses + c];
          nb_blk += hb[(int64_t)b * kNumClasses + c];
        }
        const int64_t na_long = ha[(int64_t)b * kNumClasses + kClassLong];
        const int64_t nb_long = hb[(int64_t)b * kNumClasses + kClassLong];
        auto add_long = [&](int64_t i0, int64_t ni, int64_t j0, int64_t nj) {
          if (ni <= 0 || nj <= 0) return;
          for (int64_t k = 0; k < ni; k += 64) {
            LongItem li;
            li.var = v;
            li.ni = (int32_t)std::min<int64_t>(64, ni - k);
            li.i0 = i0 + k;
            li.j0 = j0;
            li.nj = nj;
            long_items.push_back(li);
          }
        };
        add_long(sa + na_blk - na_long, na_long, sb, nb_blk);
        add_long(sa, na_blk - na_long, sb + nb_blk - nb_long, nb_long);
      }
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
          const int per = kind_tile_strings(kind_of_len(c), pl->tw);
          for (int64_t s0 = sb; s0 < sb + cb; s0 += per) {
            TileMeta tm;
            tm.len = c;
            tm.count = (int32_t)std::min<int64_t>(per, sb + cb - s0);
            tm.s0 = s0;
            tm.kind = kind_of_len(c);
            tiles.push_back(tm);
          }
          for (int lb = 1; lb <= kLongLen && ca > 0; ++lb) {
            if (hb[(int64_t)b * kNumClasses + lb] == 0) continue;
            const int idx = c * (kLongLen + 1) + lb;
            const double pairs = (double)ca * (double)hb[(int64_t)b * kNumClasses + lb];
            if (mmin[idx] <= std::min(c, lb)) {
              pl->pairs_kernel += pairs;
              pl->alg_ops += pairs * kind_pair_ops(kind_of_len(lb), c);
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
    // 33..64) x chunks of tiles; cost ~ lanes x strings x length. Items are
    // later ordered by (first tile kind, first string-1 length), so
    // co-resident workgroups mostly run the same unrolled pass, then
    // largest first.
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
            item_keys[part].push_back(kind_of_len(tiles[t].len) * 256 + waves[w].len);
            item_block[part].push_back(b);
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
    a.p5 = 0;  // the blocked kernel is the per-lane form: no kP5 tiles
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
    pl->tile_blk_lo.push_back(tile_lo);
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
    a.n_candidates = nullptr;
    a.cand = nullptr;
    a.cls_start = nullptr;  // tiles are per block: survivors are scored byte-wise
    a.cls_tile0 = nullptr;
    a.hcodes = pl->hcodes;
    a.hmask = slots - 1;
    a.overflow = pl->overflow;
    a.krows_max = max_krows;
  }
  FH_TRY(pl->alloc(&pl->args_dev, pl->args.size()));
  FH_HIP(hipMemcpyAsync(pl->args_dev, pl->args.data(), pl->args.size() * sizeof(JwArgs), hipMemcpyHostToDevice, s));
  for (int part = 0; part < 2; ++part) {
    std::vector<size_t> order(items[part].size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = k;
    const auto& keys = item_keys[part];
    const auto& its = items[part];
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
      if (keys[x] != keys[y]) return keys[x] < keys[y];
      return its[x].cost > its[y].cost;
    });
    std::vector<WorkItem> sorted_items(order.size());
    for (size_t k = 0; k < order.size(); ++k) sorted_items[k] = its[order[k]];
    items[part].swap(sorted_items);
    pl->n_items[part] = (int64_t)items[part].size();
    FH_TRY(pl->alloc(&pl->items_dev[part], std::max<size_t>(1, items[part].size())));
    if (!items[part].empty())
      FH_HIP(hipMemcpyAsync(pl->items_dev[part], items[part].data(), items[part].size() * sizeof(WorkItem),
                            hipMemcpyHostToDevice, s));
    // Grouped by block (the largest-first order kept inside a block).
    std::vector<int32_t> blk_sorted(order.size());
    for (size_t k = 0; k < order.size(); ++k) blk_sorted[k] = item_block[part][order[k]];
    std::vector<size_t> by_block(order.size());
    for (size_t k = 0; k < by_block.size(); ++k) by_block[k] = k;
    std::stable_sort(by_block.begin(), by_block.end(),
                     [&](size_t x, size_t y) { return blk_sorted[x] < blk_sorted[y]; });
    std::vector<WorkItem> blk_items(by_block.size());
    auto& lo = pl->items_blk_lo[part];
    lo.assign((size_t)n_blocks + 1, 0);
    for (size_t k = 0; k < by_block.size(); ++k) {
      blk_items[k] = items[part][by_block[k]];
      lo[(size_t)blk_sorted[by_block[k]] + 1]++;
    }
    for (int32_t b = 0; b < n_blocks; ++b) lo[(size_t)b + 1] += lo[(size_t)b];
    FH_TRY(pl->alloc(&pl->items_blk_dev[part], std::max<size_t>(1, blk_items.size())));
    if (!blk_items.empty())
      FH_HIP(hipMemcpyAsync(pl->items_blk_dev[part], blk_items.data(), blk_items.size() * sizeof(WorkItem),
                            hipMemcpyHostToDevice, s));
  }
  FH_TRY(pl->alloc(&pl->heads_blk, 2 * (size_t)n_blocks));
  // Long-value pairs: a persistent grid of up to 256 workgroups, fewer when
  // values over kMaxGenericLen bytes need more than about 1 GiB of flags.
  pl->n_long = (int64_t)long_items.size();
  if (pl->n_long > 0) {
    FH_TRY(pl->alloc(&pl->long_dev, long_items.size()));
    FH_HIP(hipMemcpyAsync(pl->long_dev, long_items.data(), long_items.size() * sizeof(LongItem), hipMemcpyHostToDevice,
                          s));
    int64_t wgs = std::min<int64_t>(256, pl->n_long);
    int64_t words = 1;
    if (max_len > (uint32_t)kMaxGenericLen) {
      words = 2 * (((int64_t)max_len + 63) / 64);
      wgs = std::max<int64_t>(1, std::min<int64_t>(wgs, ((int64_t)1 << 27) / (words * 256)));
    }
    pl->long_grid = (int)wgs;
    FH_TRY(pl->alloc(&pl->long_scratch, (size_t)(wgs * 256 * words)));
  }
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

extern "C" int fh_blocked_run_per_block(void* plan, fh_stream_t stream) {
  using namespace fh;
  clear_error();
  FH_REQUIRE(plan != nullptr, "fh_blocked_run_per_block: plan is null");
  BlockedPlan* pl = static_cast<BlockedPlan*>(plan);
  hipStream_t caller = (hipStream_t)stream;
  FH_HIP(hipEventRecord(pl->ev_in, caller));
  FH_HIP(hipStreamWaitEvent(pl->stream, pl->ev_in, 0));
  FH_HIP(hipEventRecord(pl->ev_t0, pl->stream));
  FH_TRY(pl->enqueue_per_block());
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
