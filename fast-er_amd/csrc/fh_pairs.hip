// Pair-list stages of the hot path: unique-value expansion, exact-equality
// pair generation, level split, sorted-set membership and the pattern merge.
//
// All lists are uint64 "level keys" (pair id * 4 + level). The expansion and
// exact-pair kernels are load-balanced over their outputs (each output finds
// its input in the scanned counts; see for_each_output), so a value shared by
// thousands of rows does not serialise one thread the way the reference's
// per-input loops do (faster/comparison.py:236-241,
// faster/deduplication.py:295-332).
#include <algorithm>
#include <vector>

#include "fh_common.h"

namespace fh {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ int64_t ceil_div_dev(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Largest index i in [0, n) with offs[i] <= o (offs ascending, offs[0] = 0).
template <typename T>
__device__ __forceinline__ int64_t upper_index(const T* __restrict__ offs, int64_t n, uint64_t o) {
  int64_t lo = 0, hi = n;  // invariant: offs[lo] <= o < offs[hi] (offs[n] = +inf)
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if ((uint64_t)offs[mid] <= o)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

inline unsigned grid_for(int64_t n, int64_t cap = 1 << 20) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kThreads), cap));
}

// Load-balanced expansion over scanned counts: a workgroup owns a tile of
// kExpTile consecutive outputs. The inputs ("rows") owning them form one
// contiguous range, found by two binary searches over the scanned offsets;
// the range's offsets are staged in LDS, so each output searches a few LDS
// entries instead of the whole offset array.
constexpr int kExpItems = 8;
constexpr int kExpTile = kThreads * kExpItems;
constexpr int kExpLds = 1024;

struct ExpTile {
  int64_t o0, o1;   // outputs [o0, o1)
  int64_t lo, cnt;  // owning rows [lo, lo + cnt)
  bool staged;      // offs[lo .. lo + cnt) copied to LDS
};

__device__ __forceinline__ ExpTile exp_tile_begin(const uint64_t* __restrict__ offs, int64_t n, int64_t tile,
                                                  int64_t total, uint64_t* s_offs, int64_t* s_rng) {
  ExpTile t;
  t.o0 = tile * (int64_t)kExpTile;
  t.o1 = min(t.o0 + (int64_t)kExpTile, total);
  __syncthreads();  // the previous tile's searches are done with s_offs
  if (threadIdx.x == 0) s_rng[0] = upper_index(offs, n, (uint64_t)t.o0);
  if (threadIdx.x == 64) s_rng[1] = upper_index(offs, n, (uint64_t)(t.o1 - 1));
  __syncthreads();
  t.lo = s_rng[0];
  t.cnt = s_rng[1] - t.lo + 1;
  t.staged = t.cnt <= kExpLds;
  if (t.staged)
    for (int k = threadIdx.x; k < t.cnt; k += blockDim.x) s_offs[k] = offs[t.lo + k];
  __syncthreads();
  return t;
}

// Row owning output o of the tile: offs[row] <= o < offs[row + 1].
__device__ __forceinline__ int64_t exp_row(const ExpTile& t, const uint64_t* __restrict__ offs,
                                           const uint64_t* s_offs, uint64_t o) {
  if (t.staged) {
    int64_t lo = 0, hi = t.cnt;  // s_offs[lo] <= o < s_offs[hi] (s_offs[cnt] = +inf)
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (s_offs[mid] <= o)
        lo = mid;
      else
        hi = mid;
    }
    return t.lo + lo;
  }
  return t.lo + upper_index(offs + t.lo, t.cnt, o);
}

inline unsigned grid_tiles(int64_t total, int64_t cap = 1 << 18) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, kExpTile), cap));
}

// Calls f(o, row) for every output o in [0, total), row being its owner.
template <typename F>
__device__ __forceinline__ void for_each_output(const uint64_t* __restrict__ offs, int64_t n_rows, int64_t total,
                                                F&& f) {
  __shared__ uint64_t s_offs[kExpLds];
  __shared__ int64_t s_rng[2];
  const int64_t n_tiles = ceil_div_dev(total, (int64_t)kExpTile);
  for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const ExpTile t = exp_tile_begin(offs, n_rows, tile, total, s_offs, s_rng);
    for (int k = 0; k < kExpItems; ++k) {
      const int64_t o = t.o0 + (int64_t)k * blockDim.x + threadIdx.x;
      if (o >= t.o1) break;
      f(o, exp_row(t, offs, s_offs, (uint64_t)o));
    }
  }
}

// --------------------------- linkage expansion -----------------------------
__global__ void __launch_bounds__(kThreads) k_expand_link_count(const uint64_t* __restrict__ hits, int64_t n_hits,
                                                                int64_t n_ub, const int64_t* __restrict__ starts_a,
                                                                const int64_t* __restrict__ starts_b,
                                                                uint64_t* __restrict__ cnt) {
  for (int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h < n_hits; h += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t id = hits[h] >> 2;
    const uint64_t ua = id / (uint64_t)n_ub, ub = id % (uint64_t)n_ub;
    cnt[h] = (uint64_t)(starts_a[ua + 1] - starts_a[ua]) * (uint64_t)(starts_b[ub + 1] - starts_b[ub]);
  }
}

__global__ void __launch_bounds__(kThreads) k_expand_link(const uint64_t* __restrict__ hits, int64_t n_hits,
                                                          int64_t n_ub, const int64_t* __restrict__ starts_a,
                                                          const int64_t* __restrict__ rows_a,
                                                          const int64_t* __restrict__ starts_b,
                                                          const int64_t* __restrict__ rows_b, int64_t n_b,
                                                          const uint64_t* __restrict__ offs, int64_t total,
                                                          uint64_t* __restrict__ out) {
  for_each_output(offs, n_hits, total, [&](int64_t o, int64_t h) {
    const uint64_t key = hits[h];
    const uint64_t id = key >> 2;
    const uint64_t ua = id / (uint64_t)n_ub, ub = id % (uint64_t)n_ub;
    const uint64_t k = (uint64_t)o - offs[h];
    const uint64_t cb = (uint64_t)(starts_b[ub + 1] - starts_b[ub]);
    const int64_t ra = rows_a[starts_a[ua] + (int64_t)(k / cb)];
    const int64_t rb = rows_b[starts_b[ub] + (int64_t)(k % cb)];
    out[o] = ((uint64_t)ra * (uint64_t)n_b + (uint64_t)rb) * 4ull + (key & 3ull);
  });
}

// ---------------------- linkage expansion in id order ----------------------
// When values repeat, the record pairs of a unique A value ua are the same
// (B row, level) list for each of its rows: the partner list P(ua), the rows
// of every B value ub it hits, sorted. So the sorted record pairs are built
// from the sorted partner lists (|rows_a(ua)| times smaller than the record
// pairs) and one ordered copy per A row, instead of a radix sort of every
// record pair.

// Per hit: record pairs (|rows_a(ua)| * |rows_b(ub)|) and partner entries
// (|rows_b(ub)|).
__global__ void __launch_bounds__(kThreads) k_expand_link_count2(const uint64_t* __restrict__ hits, int64_t n_hits,
                                                                 int64_t n_ub, const int64_t* __restrict__ starts_a,
                                                                 const int64_t* __restrict__ starts_b,
                                                                 uint64_t* __restrict__ cnt_full,
                                                                 uint64_t* __restrict__ cnt_part) {
  for (int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h < n_hits; h += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t id = hits[h] >> 2;
    const uint64_t ua = id / (uint64_t)n_ub, ub = id % (uint64_t)n_ub;
    const uint64_t cb = (uint64_t)(starts_b[ub + 1] - starts_b[ub]);
    cnt_full[h] = (uint64_t)(starts_a[ua + 1] - starts_a[ua]) * cb;
    cnt_part[h] = cb;
  }
}

// Partner entries ((ua * n_b + rb) * 4 + level), unordered.
__global__ void __launch_bounds__(kThreads) k_expand_partners(const uint64_t* __restrict__ hits, int64_t n_hits,
                                                              int64_t n_ub, const int64_t* __restrict__ starts_b,
                                                              const int64_t* __restrict__ rows_b, int64_t n_b,
                                                              const uint64_t* __restrict__ offs, int64_t total,
                                                              uint64_t* __restrict__ out) {
  for_each_output(offs, n_hits, total, [&](int64_t o, int64_t h) {
    const uint64_t key = hits[h];
    const uint64_t id = key >> 2;
    const uint64_t ua = id / (uint64_t)n_ub, ub = id % (uint64_t)n_ub;
    const int64_t rb = rows_b[starts_b[ub] + (o - (int64_t)offs[h])];
    out[o] = ((ua * (uint64_t)n_b + (uint64_t)rb) << 2) | (key & 3ull);
  });
}

// Unique value of every A row (rows of value g: rows_a[starts_a[g] ..
// starts_a[g + 1])).
// One thread per listed row (starts_a[n_ua] <= n_a of them), which finds its
// value by a search of the starts: a value shared by many rows does not
// leave one thread writing all of them.
__global__ void __launch_bounds__(kThreads) k_row_values(const int64_t* __restrict__ starts_a,
                                                         const int64_t* __restrict__ rows_a, int64_t n_ua,
                                                         int64_t n_a, int64_t* __restrict__ value_of) {
  const int64_t np = min(starts_a[n_ua], n_a);
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x)
    value_of[rows_a[p]] = upper_index(starts_a, n_ua, (uint64_t)p);
}

// First entry of each value's partner list in the sorted partners
// (value g's entries are [first[g], first[g + 1]); first[n_ua] = n_part).
__global__ void __launch_bounds__(kThreads) k_partner_bounds(const uint64_t* __restrict__ part, int64_t n_part,
                                                             int64_t n_ua, int64_t n_b, uint64_t* __restrict__ first) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g <= n_ua; g += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t key = ((uint64_t)g * (uint64_t)n_b) << 2;
    int64_t lo = 0, hi = n_part;  // first index with part[i] >= key
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (part[mid] < key)
        lo = mid + 1;
      else
        hi = mid;
    }
    first[g] = (uint64_t)lo;
  }
}

// Record pairs per A row: its value's partner count.
__global__ void __launch_bounds__(kThreads) k_row_pair_counts(const int64_t* __restrict__ value_of, int64_t n_a,
                                                              const uint64_t* __restrict__ first,
                                                              uint64_t* __restrict__ cnt) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_a; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = value_of[r];
    cnt[r] = g >= 0 ? first[g + 1] - first[g] : 0ull;
  }
}

// The sorted record pairs: A rows in order, each followed by its value's
// partner list with the row put in place of the value.
__global__ void __launch_bounds__(kThreads) k_expand_rows(const int64_t* __restrict__ value_of, int64_t n_a,
                                                          const uint64_t* __restrict__ first,
                                                          const uint64_t* __restrict__ part, int64_t n_b,
                                                          const uint64_t* __restrict__ offs, int64_t total,
                                                          uint64_t* __restrict__ out) {
  for_each_output(offs, n_a, total, [&](int64_t o, int64_t r) {
    const uint64_t g = (uint64_t)value_of[r];
    const uint64_t v = part[first[g] + (uint64_t)(o - (int64_t)offs[r])];
    out[o] = v + (((uint64_t)r - g) * (uint64_t)n_b << 2);  // (r * n_b + rb) * 4 + level, mod 2^64
  });
}

// -------------------------- deduplication expansion ------------------------
// Hit index h < n_hits: a value pair; h >= n_hits: value u = h - n_hits paired
// with itself (the reference's forced diagonal) at diag_level.
__device__ __forceinline__ void dedup_hit(const uint64_t* hits, int64_t n_hits, int64_t n_u, int diag_level,
                                          int64_t h, uint64_t* ux, uint64_t* uy, uint64_t* lvl) {
  if (h < n_hits) {
    const uint64_t key = hits[h];
    const uint64_t id = key >> 2;
    *ux = id / (uint64_t)n_u;
    *uy = id % (uint64_t)n_u;
    *lvl = key & 3ull;
  } else {
    *ux = *uy = (uint64_t)(h - n_hits);
    *lvl = (uint64_t)diag_level;
  }
}

__global__ void __launch_bounds__(kThreads) k_expand_dedup_count(const uint64_t* __restrict__ hits, int64_t n_hits,
                                                                 int64_t n_virtual, int64_t n_u, int diag_level,
                                                                 const int64_t* __restrict__ starts,
                                                                 uint64_t* __restrict__ cnt) {
  for (int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h < n_virtual; h += (int64_t)gridDim.x * blockDim.x) {
    uint64_t ux, uy, lvl;
    dedup_hit(hits, n_hits, n_u, diag_level, h, &ux, &uy, &lvl);
    const uint64_t cx = (uint64_t)(starts[ux + 1] - starts[ux]);
    const uint64_t cy = (uint64_t)(starts[uy + 1] - starts[uy]);
    cnt[h] = ux != uy ? cx * cy : cx * (cx - 1) / 2;
  }
}

// k-th pair (i > j) of a group in row-major order of the strict lower
// triangle: i = the largest with i(i-1)/2 <= k.
__device__ __forceinline__ void tri_decode(uint64_t k, uint64_t* i, uint64_t* j) {
  uint64_t r = (uint64_t)((1.0 + sqrt(1.0 + 8.0 * (double)k)) * 0.5);
  while (r > 1 && r * (r - 1) / 2 > k) --r;
  while ((r + 1) * r / 2 <= k) ++r;
  *i = r;
  *j = k - r * (r - 1) / 2;
}

__global__ void __launch_bounds__(kThreads) k_expand_dedup(const uint64_t* __restrict__ hits, int64_t n_hits,
                                                           int64_t n_virtual, int64_t n_u, int diag_level,
                                                           const int64_t* __restrict__ starts,
                                                           const int64_t* __restrict__ rows, int64_t n,
                                                           const uint64_t* __restrict__ offs, int64_t total,
                                                           uint64_t* __restrict__ out) {
  for_each_output(offs, n_virtual, total, [&](int64_t o, int64_t h) {
    uint64_t ux, uy, lvl;
    dedup_hit(hits, n_hits, n_u, diag_level, h, &ux, &uy, &lvl);
    const uint64_t k = (uint64_t)o - offs[h];
    uint64_t r1, r2;
    if (ux != uy) {
      const uint64_t cy = (uint64_t)(starts[uy + 1] - starts[uy]);
      r1 = (uint64_t)rows[starts[ux] + (int64_t)(k / cy)];
      r2 = (uint64_t)rows[starts[uy] + (int64_t)(k % cy)];
    } else {
      uint64_t i, j;
      tri_decode(k, &i, &j);
      r1 = (uint64_t)rows[starts[ux] + (int64_t)i];
      r2 = (uint64_t)rows[starts[ux] + (int64_t)j];
    }
    const uint64_t hi = r1 > r2 ? r1 : r2, lo = r1 > r2 ? r2 : r1;
    out[o] = (hi * (uint64_t)n + lo) * 4ull + lvl;
  });
}

// ------------------------------ exact pairs --------------------------------
__global__ void __launch_bounds__(kThreads) k_exact_link_count(const int64_t* __restrict__ group_a, int64_t n_a,
                                                               const int64_t* __restrict__ starts_b,
                                                               uint64_t* __restrict__ cnt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_a; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = group_a[i];
    cnt[i] = g >= 0 ? (uint64_t)(starts_b[g + 1] - starts_b[g]) : 0ull;
  }
}

__global__ void __launch_bounds__(kThreads) k_exact_link(const int64_t* __restrict__ group_a, int64_t n_a,
                                                         const int64_t* __restrict__ starts_b,
                                                         const int64_t* __restrict__ rows_b, int64_t n_b,
                                                         const uint64_t* __restrict__ offs, int64_t total,
                                                         uint64_t* __restrict__ out) {
  for_each_output(offs, n_a, total, [&](int64_t o, int64_t i) {
    const int64_t g = group_a[i];
    const int64_t j = rows_b[starts_b[g] + (o - (int64_t)offs[i])];
    out[o] = ((uint64_t)i * (uint64_t)n_b + (uint64_t)j) * 4ull + 1ull;
  });
}

// Position of every row inside its value group (rows ascending per group);
// 0 for a row without a value (negative group), which pairs with nothing.
// rows lists the starts[n_groups] rows that have a value.
__global__ void __launch_bounds__(kThreads) k_group_rank(const int64_t* __restrict__ group,
                                                         const int64_t* __restrict__ starts,
                                                         const int64_t* __restrict__ rows, int64_t n,
                                                         int64_t n_groups, uint64_t* __restrict__ rank) {
  const int64_t listed = min(starts[n_groups], n);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    if (group[k] < 0) rank[k] = 0;
    if (k < listed) {
      const int64_t r = rows[k];
      rank[r] = (uint64_t)(k - starts[group[r]]);
    }
  }
}

__global__ void __launch_bounds__(kThreads) k_exact_dedup(const int64_t* __restrict__ group, int64_t n,
                                                          const int64_t* __restrict__ starts,
                                                          const int64_t* __restrict__ rows,
                                                          const uint64_t* __restrict__ offs, int64_t total,
                                                          uint64_t* __restrict__ out) {
  for_each_output(offs, n, total, [&](int64_t o, int64_t i) {
    const int64_t j = rows[starts[group[i]] + (o - (int64_t)offs[i])];
    out[o] = ((uint64_t)i * (uint64_t)n + (uint64_t)j) * 4ull + 1ull;
  });
}

// ----------------------------- level split ---------------------------------
__global__ void __launch_bounds__(kThreads) k_level_flags(const uint64_t* __restrict__ keys, int64_t n, int level,
                                                          uint32_t* __restrict__ flags) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    flags[i] = (int)(keys[i] & 3ull) == level ? 1u : 0u;
}

// Level l's ids go after those of levels 1 .. l - 1, whose counts are
// tot[0 .. l - 2] (device), so the split needs no host readback per level.
__global__ void __launch_bounds__(kThreads) k_level_scatter(const uint64_t* __restrict__ keys, int64_t n, int level,
                                                            const uint32_t* __restrict__ pos,
                                                            const uint32_t* __restrict__ tot,
                                                            int64_t* __restrict__ ids) {
  int64_t base = 0;
  for (int l = 1; l < level; ++l) base += tot[l - 1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if ((int)(keys[i] & 3ull) == level) ids[base + pos[i]] = (int64_t)(keys[i] >> 2);
}

// ------------------------ sorted-set membership ----------------------------
__global__ void __launch_bounds__(kThreads) k_member_mask(const int64_t* __restrict__ a1, int64_t n1,
                                                          const int64_t* __restrict__ a2, int64_t n2, int inside,
                                                          uint8_t* __restrict__ mask) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n1; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = a1[i];
    int64_t lo = 0, hi = n2;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (a2[mid] < x)
        lo = mid + 1;
      else
        hi = mid;
    }
    const bool found = lo < n2 && a2[lo] == x;
    mask[i] = (uint8_t)(inside ? found : !found);
  }
}

// ----------------------------- pattern merge -------------------------------
// For each entry of the source list of variable `v`, finds the pair's level in
// every other variable's list (0 when absent). An entry is the owner of its
// pair when no earlier variable lists the pair; owners get their pattern code,
// the others 0xFFFFFFFF.
//
// Each block handles a tile of consecutive source entries, whose ids it keeps
// in LDS, and one range of every other list: the lists are split at the ids
// of the tiles' first entries by a separate search kernel. Entries of a
// range of up to kMergeInPlace entries binary-search the tile's ids; the
// tile's entries binary-search a longer range in place. Codes accumulate in
// LDS with atomic adds, as several lists' matches land on one entry.
constexpr int kMergeTile = 1024;
constexpr int kMergeInPlace = 8192;  // longer ranges are searched in place
constexpr int kMergeBatch = kMergeTile / kThreads;  // tile entries per thread
constexpr int kMaxVars = 16;

struct MergeArgs {
  const uint64_t* lists[kMaxVars];
  int64_t sizes[kMaxVars];
  uint32_t stride[kMaxVars];
  int n_vars;
  // Exact variables given as per-row value ids (level = ids equal): pair id
  // -> (id / id_div, id % id_div); row values imp_a[k][row1], imp_b[k][row2].
  const int64_t* imp_a[kMaxVars];
  const int64_t* imp_b[kMaxVars];
  uint32_t imp_stride[kMaxVars];
  int n_imp;
  uint64_t id_div;
  // Level count of every listed variable; a key whose level is 0 or not
  // below it sets *err (the merge then fails with FH_ERR_ARG).
  uint32_t nlev[kMaxVars];
  uint32_t* err;
  // Listed exact variables whose list is exactly their equal pairs (level 1)
  // of the per-row value ids lid_a[w][row1], lid_b[w][row2] (negative: no
  // value): other lists' entries take their level from the two ids instead
  // of a search of the list (null: searched).
  const int64_t* lid_a[kMaxVars];
  const int64_t* lid_b[kMaxVars];
  int n_lid;
};

__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t* a, int64_t lo, int64_t hi, uint64_t x) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// First index p[k] in [0, n) with a[p[k]] >= x[k] (a ascending, in LDS), or
// n, for K keys at once: the K searches take every halving together, so
// their LDS reads overlap instead of running one search after another.
template <int K>
__device__ __forceinline__ void lds_lower_bound(const uint64_t* a, int n, const uint64_t (&x)[K], int (&p)[K]) {
  int base[K];
#pragma unroll
  for (int k = 0; k < K; ++k) base[k] = 0;
  for (int cnt = n; cnt > 1;) {
    const int half = cnt >> 1;
#pragma unroll
    for (int k = 0; k < K; ++k) base[k] = a[base[k] + half] < x[k] ? base[k] + half : base[k];
    cnt -= half;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) p[k] = n > 0 && a[base[k]] < x[k] ? base[k] + 1 : base[k];
}

// bounds[w * (ntiles + 1) + t]: the first entry of list w whose pair id is at
// least the id of the first entry of tile t of list v, and list w's size for
// t = ntiles. Consecutive bounds split every other list into one range per
// tile of v. One thread per search: the dependent loads of many searches
// overlap, where a search inside the tile kernel would stall its workgroup.
__global__ void __launch_bounds__(kThreads) k_tile_bounds(MergeArgs m, int v, int64_t ntiles,
                                                          int64_t* __restrict__ bounds) {
  const int64_t per = ntiles + 1;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= per * m.n_vars) return;
  const int w = (int)(g / per);
  const int64_t t = g - (int64_t)w * per;
  int64_t b = 0;  // empty ranges for the source list and lists looked up by row ids
  if (w != v && m.lid_a[w] == nullptr) {
    b = t == ntiles ? m.sizes[w]
                    : lower_bound_u64(m.lists[w], 0, m.sizes[w], (m.lists[v][t * kMergeTile] >> 2) * 4ull);
  }
  bounds[g] = b;
}

// Lookups of the pattern-code kernel. Entries [e0, e0 + K * kThreads) of the
// sequence of short ranges (list w's range holds rshort[w] entries, in list
// order) find their pair among the tile's ids; a match marks the tile entry
// as not owned (an earlier list) or adds the level times the list's stride.
template <int K>
__device__ __forceinline__ void short_lookups(int e0, int nshort, int nv, int v, int tn, const int* rshort,
                                              const uint64_t* const* rptr, const uint32_t* rstride,
                                              const uint64_t* tids, uint32_t* tcode, uint8_t* tgone) {
  uint64_t x[K], xid[K];
  int xw[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    int e = e0 + k * kThreads + (int)threadIdx.x;
    xw[k] = -1;
    x[k] = 0;
    if (e < nshort) {
      int w = 0;
      for (; w < nv; ++w) {
        if (e < rshort[w]) break;
        e -= rshort[w];
      }
      xw[k] = w;
      x[k] = rptr[w][e];
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) xid[k] = x[k] >> 2;
  int p[K];
  lds_lower_bound(tids, tn, xid, p);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (xw[k] >= 0 && p[k] < tn && tids[p[k]] == xid[k]) {
      if (xw[k] < v)
        tgone[p[k]] = 1;
      else
        atomicAdd(&tcode[p[k]], (uint32_t)(x[k] & 3ull) * rstride[xw[k]]);
    }
  }
}

// The thread's tile entries (kMergeBatch of them) binary-search the ranges of
// the NL lists ws in place, all searches advancing together: one round trip
// per halving. The first index with range[index] >= key is base + (range[base]
// < key) after the halvings, and every read stays inside the range.
template <int NL>
__device__ __forceinline__ void inplace_lookups(const int (&ws)[NL], int v, int tn, const int64_t* rlen,
                                                const uint64_t* const* rptr, const uint32_t* rstride,
                                                const uint64_t* tids, uint32_t* tcode, uint8_t* tgone) {
  const int tid = (int)threadIdx.x;
  int64_t len[NL], cnt[NL], base[NL][kMergeBatch];
  const uint64_t* tl[NL];
  uint64_t key[kMergeBatch];
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    len[l] = rlen[ws[l]];
    cnt[l] = len[l];
    tl[l] = rptr[ws[l]];
#pragma unroll
    for (int k = 0; k < kMergeBatch; ++k) base[l][k] = 0;
  }
#pragma unroll
  for (int k = 0; k < kMergeBatch; ++k) {
    const int i = k * kThreads + tid;
    key[k] = i < tn ? tids[i] * 4ull : ~0ull;
  }
  for (bool more = true; more;) {
    uint64_t probe[NL][kMergeBatch];
#pragma unroll
    for (int l = 0; l < NL; ++l)
#pragma unroll
      for (int k = 0; k < kMergeBatch; ++k) probe[l][k] = cnt[l] > 1 ? tl[l][base[l][k] + (cnt[l] >> 1)] : 0ull;
    more = false;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      if (cnt[l] <= 1) continue;
      const int64_t half = cnt[l] >> 1;
#pragma unroll
      for (int k = 0; k < kMergeBatch; ++k) base[l][k] = probe[l][k] < key[k] ? base[l][k] + half : base[l][k];
      cnt[l] -= half;
      more = more || cnt[l] > 1;
    }
  }
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    uint64_t at[kMergeBatch], next[kMergeBatch];
#pragma unroll
    for (int k = 0; k < kMergeBatch; ++k) {
      at[k] = tl[l][base[l][k]];
      next[k] = base[l][k] + 1 < len[l] ? tl[l][base[l][k] + 1] : ~0ull;
    }
#pragma unroll
    for (int k = 0; k < kMergeBatch; ++k) {
      const int i = k * kThreads + tid;
      if (i >= tn) continue;
      const bool ge = at[k] >= key[k];
      const uint64_t xv = ge ? at[k] : next[k];
      if ((ge || base[l][k] + 1 < len[l]) && (xv >> 2) == tids[i]) {
        if (ws[l] < v)
          tgone[i] = 1;
        else
          atomicAdd(&tcode[i], (uint32_t)(xv & 3ull) * rstride[ws[l]]);
      }
    }
  }
}

// With `tile_hist`, each tile also counts its owners per code:
// tile_hist[(code - code_lo) * gridDim.x + tile], for k_code_scatter.
//
// The kernel is bound by the latency of its dependent global reads, so the
// tile's keys and every range's bounds are read in one round trip, all
// entries of the short ranges a thread looks up are loaded before the first
// lookup, and the in-place searches of a thread's four tile entries advance
// in lockstep (one round trip per halving for all four).
__global__ void __launch_bounds__(kThreads) k_pattern_codes(MergeArgs m, int v, const int64_t* __restrict__ bounds,
                                                            uint32_t* __restrict__ codes, uint32_t code_lo,
                                                            int ncodes, uint32_t* __restrict__ tile_hist) {
  static_assert(kThreads == 256, "one owner-count bin per thread");
  static_assert(kMergeTile == kMergeBatch * kThreads, "kMergeBatch tile entries per thread");
  __shared__ uint32_t th[256];
  __shared__ uint64_t tids[kMergeTile];   // the tile's pair ids, ascending
  __shared__ uint32_t tcode[kMergeTile];  // pattern code so far of each tile entry
  __shared__ uint8_t tgone[kMergeTile];   // 1: an earlier variable lists the pair
  __shared__ const uint64_t* rptr[kMaxVars];  // range of list w: rptr[w][0 .. rlen[w])
  __shared__ int64_t rlen[kMaxVars];          // 0 for the source list
  __shared__ uint32_t rstride[kMaxVars];
  __shared__ int rshort[kMaxVars];            // rlen[w] when it is not searched in place, else 0
  const uint64_t* src = m.lists[v];
  const int64_t n = m.sizes[v];
  const int64_t t0 = (int64_t)blockIdx.x * kMergeTile;
  const int64_t t1 = min(t0 + (int64_t)kMergeTile, n);
  const int tn = (int)(t1 - t0);
  const int64_t per = (int64_t)gridDim.x + 1;
  const int tid = threadIdx.x;
  const int nv = m.n_vars;
  th[tid] = 0;
  if (tid < nv) {
    const int64_t b0 = bounds[tid * per + blockIdx.x];
    const int64_t len = tid == v ? 0 : bounds[tid * per + blockIdx.x + 1] - b0;
    rptr[tid] = m.lists[tid] + b0;
    rlen[tid] = len > 0 ? len : 0;  // descending bounds only come from unsorted lists
    rshort[tid] = len > 0 && len <= kMergeInPlace ? (int)len : 0;
    rstride[tid] = m.stride[tid];
  }
  {
    // All four keys are loaded before any is used: the error flag's atomic
    // could alias the lists, so loads after it would wait for it.
    uint64_t key[kMergeBatch];
#pragma unroll
    for (int k = 0; k < kMergeBatch; ++k) {
      const int i = k * kThreads + tid;
      key[k] = i < tn ? src[t0 + i] : 0ull;
    }
    bool bad = false;
#pragma unroll
    for (int k = 0; k < kMergeBatch; ++k) {
      const int i = k * kThreads + tid;
      if (i >= tn) continue;
      const uint32_t lvl = (uint32_t)(key[k] & 3ull);
      bad = bad || lvl == 0 || lvl >= m.nlev[v];
      tids[i] = key[k] >> 2;
      tcode[i] = lvl * m.stride[v];
      tgone[i] = 0;
    }
    if (bad) atomicOr(m.err, 1u);
  }
  __syncthreads();
  // Ranges of up to kMergeInPlace entries, taken as one sequence in list
  // order: each of their entries looks itself up among the tile's ids, eight
  // (or, for the last 1,024 or fewer, four) per thread at a time, all loaded
  // before the lookups.
  int nshort = 0;
  for (int w = 0; w < nv; ++w) nshort += rshort[w];
  int e0 = 0;
  for (; nshort - e0 > kMergeBatch * kThreads; e0 += 2 * kMergeBatch * kThreads)
    short_lookups<2 * kMergeBatch>(e0, nshort, nv, v, tn, rshort, rptr, rstride, tids, tcode, tgone);
  if (e0 < nshort) short_lookups<kMergeBatch>(e0, nshort, nv, v, tn, rshort, rptr, rstride, tids, tcode, tgone);
  // Longer ranges: each tile entry binary-searches them in place, which
  // touches about log2 of a range's length entries instead of reading all
  // of it; two lists at a time, so that eight searches per thread share
  // every round trip.
  int pending = -1;
  for (int w = 0; w < nv; ++w) {
    if (rlen[w] <= kMergeInPlace) continue;
    if (pending < 0) {
      pending = w;
      continue;
    }
    const int ws[2] = {pending, w};
    inplace_lookups<2>(ws, v, tn, rlen, rptr, rstride, tids, tcode, tgone);
    pending = -1;
  }
  if (pending >= 0) {
    const int ws[1] = {pending};
    inplace_lookups<1>(ws, v, tn, rlen, rptr, rstride, tids, tcode, tgone);
  }
  __syncthreads();
  // Codes out; each wave adds its owners per code to the tile's counts with
  // one LDS atomic per distinct code (lanes of one code found by eight
  // ballots), as a list's owners often share a few codes.
  const uint64_t lt = fhd::lanemask_lt();
  bool lid_other = false;  // a list other than the source is looked up by row ids
  for (int w = 0; w < nv; ++w) lid_other = lid_other || (w != v && m.lid_a[w] != nullptr);
#pragma unroll
  for (int k = 0; k < kMergeBatch; ++k) {
    const int i = k * kThreads + tid;
    uint32_t code = 0xFFFFFFFFu;
    if (i < tn) {
      code = tcode[i];
      bool gone = tgone[i] != 0;
      if (!gone && (m.n_imp > 0 || lid_other)) {
        const uint64_t id = tids[i];
        const uint64_t r1 = id / m.id_div, r2 = id % m.id_div;
        for (int w = 0; w < m.n_vars; ++w) {
          if (w == v || m.lid_a[w] == nullptr) continue;
          const int64_t x = m.lid_a[w][r1];
          if (x >= 0 && x == m.lid_b[w][r2]) {
            if (w < v)
              gone = true;  // the earlier list w holds the pair
            else
              code += m.stride[w];  // level 1 of w
          }
        }
        for (int q = 0; q < m.n_imp; ++q)
          if (m.imp_a[q][r1] == m.imp_b[q][r2]) code += m.imp_stride[q];
      }
      if (gone) code = 0xFFFFFFFFu;
      codes[t0 + i] = code;
    }
    if (tile_hist) {
      bool valid = code != 0xFFFFFFFFu;
      const uint32_t d = valid ? code - code_lo : 0u;
      if (valid && d >= (uint32_t)ncodes) {
        atomicOr(m.err, 1u);  // only reachable through an invalid level key
        valid = false;
      }
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
      }
      if (valid && (peers & lt) == 0) atomicAdd(&th[d], (uint32_t)__popcll(peers));
    }
  }
  if (tile_hist) {
    __syncthreads();
    if ((int)threadIdx.x < ncodes) tile_hist[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = th[threadIdx.x];
  }
}

// Owners of one list, ordered by code (stable: list order within a code),
// their pair ids written at their final positions: the one counting-sort
// pass over the digit code - code_lo (fewer than 256 codes) that replaces
// flagging, compacting and sorting the owners. A tile is the kMergeTile
// entries of one k_pattern_codes tile, whose counts the scanned tile_hist
// holds; wave w ranks entries [w * 256, (w + 1) * 256) of it, 64 at a time,
// by eight ballots and an LDS atomic per code group, the tile's owners are
// staged in LDS in code order and each code leaves as one run.
//
// The list's output starts at state[0] (the owners of the variables already
// placed, kept on the device); nothing is written when the owners would
// pass `capacity` (the count still advances, so the caller sees the size).
__global__ void __launch_bounds__(kThreads) k_code_scatter(const uint32_t* __restrict__ codes,
                                                           const uint64_t* __restrict__ keys, int64_t n,
                                                           uint32_t code_lo, int ncodes,
                                                           const uint32_t* __restrict__ hist_scanned,
                                                           const uint32_t* __restrict__ list_owners,
                                                           const uint64_t* __restrict__ state, int64_t capacity,
                                                           int64_t* __restrict__ ids) {
  if ((int64_t)state[0] + (int64_t)*list_owners > capacity) return;
  uint64_t* __restrict__ out_ids = reinterpret_cast<uint64_t*>(ids + state[0]);
  static_assert(kThreads == 256 && kMergeTile % kThreads == 0, "one code bin per thread, four waves");
  constexpr int kWaveItems = kMergeTile / (kThreads / 64);
  __shared__ uint32_t goff[256];
  __shared__ uint32_t loff[256];
  __shared__ uint32_t wcnt[kThreads / 64][256];
  __shared__ uint32_t wtot[kThreads / 64];
  __shared__ uint64_t sid[kMergeTile];
  __shared__ uint32_t scode[kMergeTile];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t wbase = (int64_t)blockIdx.x * kMergeTile + (int64_t)wave * kWaveItems;
  // Codes and keys of the wave's entries are read up front, in one round trip.
  uint32_t code[kWaveItems / 64], rnk[kWaveItems / 64];
  uint64_t key[kWaveItems / 64];
#pragma unroll
  for (int r = 0; r < kWaveItems / 64; ++r) {
    const int64_t e = wbase + r * 64 + lane;
    code[r] = e < n ? codes[e] : 0xFFFFFFFFu;
    key[r] = e < n ? keys[e] : 0ull;
  }
  goff[tid] = tid < ncodes ? hist_scanned[(int64_t)tid * gridDim.x + blockIdx.x] : 0u;
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) wcnt[w][tid] = 0;
  __syncthreads();
  const uint64_t lt = fhd::lanemask_lt();
#pragma unroll
  for (int r = 0; r < kWaveItems / 64; ++r) {
    // Codes outside the variable's range come only from invalid level keys
    // (flagged by the code kernel); they are dropped, not indexed.
    if (code[r] != 0xFFFFFFFFu && code[r] - code_lo >= (uint32_t)ncodes) code[r] = 0xFFFFFFFFu;
    const bool valid = code[r] != 0xFFFFFFFFu;
    const uint32_t d = valid ? code[r] - code_lo : 0u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    uint32_t old = 0;
    if (valid && rank == 0) old = atomicAdd(&wcnt[wave][d], (uint32_t)__popcll(peers));
    const int leader = valid ? __ffsll((unsigned long long)peers) - 1 : lane;
    rnk[r] = (uint32_t)__shfl(old, leader, 64) + rank;
  }
  __syncthreads();
  {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
      const uint32_t t = wcnt[w][tid];
      wcnt[w][tid] = run;
      run += t;
    }
    uint32_t inc = run;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t o = __shfl_up(inc, off, 64);
      if (lane >= off) inc += o;
    }
    if (lane == 63) wtot[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += wtot[w];
    loff[tid] = before + inc - run;
  }
  __syncthreads();
  const int owners = (int)(wtot[0] + wtot[1] + wtot[2] + wtot[3]);
#pragma unroll
  for (int r = 0; r < kWaveItems / 64; ++r) {
    if (code[r] != 0xFFFFFFFFu) {
      const uint32_t d = code[r] - code_lo;
      const uint32_t slot = loff[d] + wcnt[wave][d] + rnk[r];
      sid[slot] = key[r] >> 2;
      scode[slot] = code[r];
    }
  }
  __syncthreads();
  for (int j = tid; j < owners; j += kThreads) {
    const uint32_t c = scode[j];
    const uint32_t d = c - code_lo;
    out_ids[goff[d] + ((uint32_t)j - loff[d])] = sid[j];
  }
}

__global__ void __launch_bounds__(kThreads) k_owner_flags(const uint32_t* __restrict__ codes, int64_t n,
                                                          uint32_t* __restrict__ flags) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    flags[i] = codes[i] != 0xFFFFFFFFu ? 1u : 0u;
}

__global__ void __launch_bounds__(kThreads) k_owner_scatter(const uint64_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ codes, int64_t n,
                                                            const uint32_t* __restrict__ pos,
                                                            uint32_t* __restrict__ out_codes,
                                                            uint64_t* __restrict__ out_ids) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = codes[i];
    if (c != 0xFFFFFFFFu) {
      out_codes[pos[i]] = c;
      out_ids[pos[i]] = keys[i] >> 2;
    }
  }
}

// offsets[c - code_lo] = first index of code c in the sorted codes (c in
// [code_lo, code_hi]).
__global__ void __launch_bounds__(kThreads) k_code_bounds(const uint32_t* __restrict__ codes, int64_t n,
                                                          uint32_t code_lo, uint32_t code_hi,
                                                          int64_t* __restrict__ offsets) {
  const int64_t c = (int64_t)code_lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > (int64_t)code_hi) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)codes[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  offsets[c - code_lo] = lo;
}

// offsets[c] = state[0] + first index of code c among the list's owners,
// for c in [code_lo, code_hi]: the scanned count of code c's first tile
// (hist_scanned[(c - code_lo) * ntiles]); state[0] alone when the list is
// empty or its owners did not fit the output.
__global__ void __launch_bounds__(kThreads) k_code_offsets(const uint32_t* __restrict__ hist_scanned,
                                                           int64_t ntiles, const uint32_t* __restrict__ list_owners,
                                                           const uint64_t* __restrict__ state, int64_t capacity,
                                                           uint32_t code_lo, uint32_t code_hi,
                                                           int64_t* __restrict__ offsets) {
  const int64_t c = (int64_t)code_lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > (int64_t)code_hi) return;
  const int64_t base = (int64_t)state[0];
  const int64_t n = (int64_t)*list_owners;
  int64_t first = 0;
  if (hist_scanned != nullptr && base + n <= capacity) first = hist_scanned[(c - code_lo) * ntiles];
  offsets[c] = base + first;
}

__global__ void k_advance_base(uint64_t* __restrict__ state, const uint32_t* __restrict__ list_owners) {
  if (threadIdx.x == 0 && blockIdx.x == 0) state[0] += *list_owners;
}

// ----------------------------- posteriors ----------------------------------
__global__ void __launch_bounds__(kThreads) k_pair_posteriors(const int64_t* __restrict__ offs, int n_patterns,
                                                              const double* __restrict__ ksi,
                                                              double* __restrict__ post, int64_t n_pairs) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_pairs; i += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = n_patterns;  // offs[lo] <= i < offs[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (offs[mid] <= i)
        lo = mid;
      else
        hi = mid;
    }
    post[i] = ksi[lo];
  }
}

// ---------------------------- numeric distance ------------------------------
// Two values x, y are within distance d when max(x, y) <= fl(min(x, y) + d)
// in float64 (fl: round to nearest); level 2 within d2, level 1 within d1
// (d2 <= d1) but not d2. NaN is a missing value and is compared with nothing.
// One side is sorted by value; for a value a, its partners form one window of
// the sorted values, found by two binary searches: below a the predicate
// fl(b + d) >= a, above it b <= fl(a + d), both monotone in b.

// Keys of the values to sort; NaNs are counted (they sort last).
__global__ void __launch_bounds__(kThreads) k_num_keys(const double* __restrict__ v, int64_t n,
                                                       uint64_t* __restrict__ keys, uint64_t* __restrict__ perm,
                                                       unsigned long long* __restrict__ small) {
  unsigned long long lo = ~0ull, hi = 0ull, nan = 0ull;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = fhd::value_key(__double_as_longlong(v[i]), true);
    keys[i] = k;
    perm[i] = (uint64_t)i;
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
    nan += k == fhd::kNanKey ? 1ull : 0ull;
  }
  lo = fhd::block_reduce(lo, fhd::OpMin());
  hi = fhd::block_reduce(hi, fhd::OpMax());
  nan = fhd::block_reduce(nan, fhd::OpSum());
  if (threadIdx.x == 0) {
    atomicMin(small, lo);
    atomicMax(small + 1, hi);
    if (nan) atomicAdd(small + 2, nan);
  }
}

__global__ void __launch_bounds__(kThreads) k_num_gather(const double* __restrict__ v, const uint64_t* __restrict__ perm,
                                                         int64_t n, double* __restrict__ vs) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    vs[k] = v[perm[k]];
}

// First k in [lo, hi) with fl(vs[k] + d) >= a (vs ascending), else hi.
__device__ __forceinline__ int64_t num_lower(const double* __restrict__ vs, int64_t lo, int64_t hi, double a, double d) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (__dadd_rn(vs[mid], d) >= a)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// First k in [lo, hi) with vs[k] > t, else hi.
__device__ __forceinline__ int64_t num_upper(const double* __restrict__ vs, int64_t lo, int64_t hi, double t) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (vs[mid] > t)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Linkage: per A row its window in the sorted B values (nbv non-NaN values):
// win[4 i .. 4 i + 3] = lo1, lo2, hi2, hi1.
__global__ void __launch_bounds__(kThreads) k_num_link_count(const double* __restrict__ va, int64_t n_a,
                                                             const double* __restrict__ vs, int64_t nbv, double d2,
                                                             double d1, int64_t* __restrict__ win,
                                                             uint64_t* __restrict__ cnt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_a; i += (int64_t)gridDim.x * blockDim.x) {
    const double a = va[i];
    if (a != a) {
      cnt[i] = 0;
      win[4 * i] = win[4 * i + 1] = win[4 * i + 2] = win[4 * i + 3] = 0;
      continue;
    }
    const int64_t lo1 = num_lower(vs, 0, nbv, a, d1);
    const int64_t hi1 = num_upper(vs, lo1, nbv, __dadd_rn(a, d1));
    const int64_t lo2 = num_lower(vs, lo1, hi1, a, d2);
    const int64_t hi2 = num_upper(vs, lo2, hi1, __dadd_rn(a, d2));
    win[4 * i] = lo1;
    win[4 * i + 1] = lo2;
    win[4 * i + 2] = hi2;
    win[4 * i + 3] = hi1;
    cnt[i] = (uint64_t)(hi1 - lo1);
  }
}

__global__ void __launch_bounds__(kThreads) k_num_link(const int64_t* __restrict__ win, int64_t n_a,
                                                       const uint64_t* __restrict__ perm, int64_t n_b,
                                                       const uint64_t* __restrict__ offs, int64_t total,
                                                       uint64_t* __restrict__ out) {
  for_each_output(offs, n_a, total, [&](int64_t o, int64_t i) {
    const int64_t k = win[4 * i] + (o - (int64_t)offs[i]);
    const uint64_t lvl = (k >= win[4 * i + 1] && k < win[4 * i + 2]) ? 2ull : 1ull;
    out[o] = ((uint64_t)i * (uint64_t)n_b + perm[k]) * 4ull + lvl;
  });
}

// Deduplication: for sorted position p (non-NaN), the partners q > p with
// vs[q] <= fl(vs[p] + d): win[2 p] = hi2, win[2 p + 1] = hi1.
__global__ void __launch_bounds__(kThreads) k_num_dedup_count(const double* __restrict__ vs, int64_t nv, double d2,
                                                              double d1, int64_t* __restrict__ win,
                                                              uint64_t* __restrict__ cnt) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nv; p += (int64_t)gridDim.x * blockDim.x) {
    const double x = vs[p];
    const int64_t hi1 = num_upper(vs, p + 1, nv, __dadd_rn(x, d1));
    const int64_t hi2 = num_upper(vs, p + 1, hi1, __dadd_rn(x, d2));
    win[2 * p] = hi2;
    win[2 * p + 1] = hi1;
    cnt[p] = (uint64_t)(hi1 - p - 1);
  }
}

__global__ void __launch_bounds__(kThreads) k_num_dedup(const int64_t* __restrict__ win, int64_t nv,
                                                        const uint64_t* __restrict__ perm, int64_t n,
                                                        const uint64_t* __restrict__ offs, int64_t total,
                                                        uint64_t* __restrict__ out) {
  for_each_output(offs, nv, total, [&](int64_t o, int64_t p) {
    const int64_t q = p + 1 + (o - (int64_t)offs[p]);
    const uint64_t lvl = q < win[2 * p] ? 2ull : 1ull;
    const uint64_t r1 = perm[p], r2 = perm[q];
    const uint64_t hi = r1 > r2 ? r1 : r2, lo = r1 > r2 ? r2 : r1;
    out[o] = (hi * (uint64_t)n + lo) * 4ull + lvl;
  });
}

// ----------------------------- missing levels -------------------------------
// Re-codes the pairs of a pattern-major list with the missing digits of their
// rows: code += sum of add[v] over the variables v missing on either row, and
// packs (code, id) into one key so that one radix sort restores pattern-major
// order with ids ascending inside each pattern.
constexpr int kMaxMissVars = 32;
struct MissingAdd {
  uint32_t add[kMaxMissVars];
  int n_vars;
};

__global__ void __launch_bounds__(kThreads) k_missing_recode(const int64_t* __restrict__ ids, int64_t n,
                                                             const int64_t* __restrict__ offs, int n_patterns,
                                                             const uint32_t* __restrict__ miss_a,
                                                             const uint32_t* __restrict__ miss_b, uint64_t id_div,
                                                             MissingAdd ma, int id_bits, uint64_t* __restrict__ keys) {
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = n_patterns;  // offs[lo] <= o < offs[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (offs[mid] <= o)
        lo = mid;
      else
        hi = mid;
    }
    const uint64_t id = (uint64_t)ids[o];
    const uint32_t m = miss_a[id / id_div] | miss_b[id % id_div];
    uint64_t code = (uint64_t)lo;
    for (int v = 0; v < ma.n_vars; ++v)
      if ((m >> v) & 1u) code += ma.add[v];
    keys[o] = (code << id_bits) | id;
  }
}

__global__ void __launch_bounds__(kThreads) k_missing_split(const uint64_t* __restrict__ keys, int64_t n, int id_bits,
                                                            int64_t* __restrict__ ids) {
  const uint64_t mask = (1ull << id_bits) - 1ull;
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (int64_t)gridDim.x * blockDim.x)
    ids[o] = (int64_t)(keys[o] & mask);
}

// offsets[c] = first sorted key with code >= c, c in [0, n_patterns].
__global__ void __launch_bounds__(kThreads) k_missing_offsets(const uint64_t* __restrict__ keys, int64_t n,
                                                              int n_patterns, int id_bits,
                                                              int64_t* __restrict__ offsets) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > n_patterns) return;
  const uint64_t x = (uint64_t)c << id_bits;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  offsets[c] = lo;
}

// Scans `cnt` (n entries) into offsets and returns the total on the host.
int scan_total(uint64_t* cnt, int64_t n, uint64_t* total_dev, uint64_t* total, hipStream_t s) {
  FH_TRY(exclusive_scan_u64(cnt, cnt, n, total_dev, s));
  FH_HIP(hipMemcpyAsync(total, total_dev, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  return FH_OK;
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" int fh_expand_link(const uint64_t* hits, int64_t n_hits, int64_t n_ub, const int64_t* starts_a,
                              const int64_t* rows_a, const int64_t* starts_b, const int64_t* rows_b, int64_t n_b,
                              uint64_t* out, int64_t capacity, int64_t* n_out, fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n_out != nullptr, "fh_expand_link: n_out is null");
  FH_REQUIRE(n_hits >= 0 && n_ub > 0 && n_b >= 0, "fh_expand_link: bad sizes");
  *n_out = 0;
  if (n_hits == 0) return FH_OK;
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  uint64_t *cnt = nullptr, *tot = nullptr;
  FH_TRY(scr.alloc(&cnt, (size_t)n_hits));
  FH_TRY(scr.alloc(&tot, 1));
  hipLaunchKernelGGL(k_expand_link_count, dim3(grid_for(n_hits)), dim3(kThreads), 0, s, hits, n_hits, n_ub, starts_a,
                     starts_b, cnt);
  FH_HIP(hipGetLastError());
  uint64_t total = 0;
  FH_TRY(scan_total(cnt, n_hits, tot, &total, s));
  *n_out = (int64_t)total;
  if ((int64_t)total > capacity || total == 0) return FH_OK;
  hipLaunchKernelGGL(k_expand_link, dim3(grid_tiles((int64_t)total)), dim3(kThreads), 0, s, hits, n_hits, n_ub,
                     starts_a, rows_a, starts_b, rows_b, n_b, cnt, (int64_t)total, out);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

extern "C" int fh_expand_link_sorted(const uint64_t* hits, int64_t n_hits, int64_t n_ua, int64_t n_ub,
                                     const int64_t* starts_a, const int64_t* rows_a, int64_t n_a,
                                     const int64_t* starts_b, const int64_t* rows_b, int64_t n_b, uint64_t* out,
                                     int64_t capacity, int64_t* n_out, fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n_out != nullptr, "fh_expand_link_sorted: n_out is null");
  FH_REQUIRE(n_hits >= 0 && n_ua > 0 && n_ub > 0 && n_a > 0 && n_b > 0, "fh_expand_link_sorted: bad sizes");
  *n_out = 0;
  if (n_hits == 0) return FH_OK;
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  uint64_t *cnt_full = nullptr, *cnt_part = nullptr, *tot = nullptr;
  FH_TRY(scr.alloc(&cnt_full, (size_t)n_hits));
  FH_TRY(scr.alloc(&cnt_part, (size_t)n_hits));
  FH_TRY(scr.alloc(&tot, 2));
  hipLaunchKernelGGL(k_expand_link_count2, dim3(grid_for(n_hits)), dim3(kThreads), 0, s, hits, n_hits, n_ub, starts_a,
                     starts_b, cnt_full, cnt_part);
  FH_HIP(hipGetLastError());
  FH_TRY(exclusive_scan_u64(cnt_full, cnt_full, n_hits, tot, s));
  FH_TRY(exclusive_scan_u64(cnt_part, cnt_part, n_hits, tot + 1, s));
  uint64_t h_tot[2] = {0, 0};
  FH_HIP(hipMemcpyAsync(h_tot, tot, sizeof(h_tot), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  const uint64_t total = h_tot[0], n_part = h_tot[1];
  *n_out = (int64_t)total;
  if ((int64_t)total > capacity || total == 0) return FH_OK;
  const int bits = bits_for(((uint64_t)n_a * (uint64_t)n_b - 1) * 4 + 3);
  if (total < 2 * n_part) {
    // Few repeated A values: expand every record pair and sort them.
    uint64_t* tmp = nullptr;
    FH_TRY(scr.alloc(&tmp, (size_t)total));
    hipLaunchKernelGGL(k_expand_link, dim3(grid_tiles((int64_t)total)), dim3(kThreads), 0, s, hits, n_hits, n_ub,
                       starts_a, rows_a, starts_b, rows_b, n_b, cnt_full, (int64_t)total, tmp);
    FH_HIP(hipGetLastError());
    return radix_sort_u64_into(tmp, out, (int64_t)total, 0, bits, s);
  }
  uint64_t *part = nullptr, *part_tmp = nullptr, *first = nullptr, *rcnt = nullptr;
  int64_t* value_of = nullptr;
  FH_TRY(scr.alloc(&part_tmp, (size_t)n_part));
  FH_TRY(scr.alloc(&part, (size_t)n_part));
  FH_TRY(scr.alloc(&first, (size_t)n_ua + 1));
  FH_TRY(scr.alloc(&value_of, (size_t)n_a));
  FH_TRY(scr.alloc(&rcnt, (size_t)n_a));
  hipLaunchKernelGGL(k_expand_partners, dim3(grid_tiles((int64_t)n_part)), dim3(kThreads), 0, s, hits, n_hits, n_ub,
                     starts_b, rows_b, n_b, cnt_part, (int64_t)n_part, part_tmp);
  FH_HIP(hipGetLastError());
  FH_TRY(radix_sort_u64_into(part_tmp, part, (int64_t)n_part, 0,
                             bits_for(((uint64_t)n_ua * (uint64_t)n_b - 1) * 4 + 3), s));
  // Rows without a value in the groups (not listed in rows_a) pair with nothing.
  FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(value_of), ~0ull, n_a, s));
  hipLaunchKernelGGL(k_row_values, dim3(grid_for(n_a)), dim3(kThreads), 0, s, starts_a, rows_a, n_ua, n_a, value_of);
  FH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_partner_bounds, dim3(grid_for(n_ua + 1)), dim3(kThreads), 0, s, part, (int64_t)n_part, n_ua,
                     n_b, first);
  FH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_row_pair_counts, dim3(grid_for(n_a)), dim3(kThreads), 0, s, value_of, n_a, first, rcnt);
  FH_HIP(hipGetLastError());
  FH_TRY(exclusive_scan_u64(rcnt, rcnt, n_a, nullptr, s));
  hipLaunchKernelGGL(k_expand_rows, dim3(grid_tiles((int64_t)total)), dim3(kThreads), 0, s, value_of, n_a, first,
                     part, n_b, rcnt, (int64_t)total, out);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

extern "C" int fh_expand_dedup(const uint64_t* hits, int64_t n_hits, int64_t n_u, const int64_t* starts,
                               const int64_t* rows, int64_t n, int diag_level, uint64_t* out, int64_t capacity,
                               int64_t* n_out, fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n_out != nullptr, "fh_expand_dedup: n_out is null");
  FH_REQUIRE(n_hits >= 0 && n_u >= 0 && n >= 0, "fh_expand_dedup: bad sizes");
  *n_out = 0;
  const int64_t n_virtual = n_hits + (diag_level > 0 ? n_u : 0);
  if (n_virtual == 0) return FH_OK;
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  uint64_t *cnt = nullptr, *tot = nullptr;
  FH_TRY(scr.alloc(&cnt, (size_t)n_virtual));
  FH_TRY(scr.alloc(&tot, 1));
  hipLaunchKernelGGL(k_expand_dedup_count, dim3(grid_for(n_virtual)), dim3(kThreads), 0, s, hits, n_hits, n_virtual,
                     n_u, diag_level, starts, cnt);
  FH_HIP(hipGetLastError());
  uint64_t total = 0;
  FH_TRY(scan_total(cnt, n_virtual, tot, &total, s));
  *n_out = (int64_t)total;
  if ((int64_t)total > capacity || total == 0) return FH_OK;
  hipLaunchKernelGGL(k_expand_dedup, dim3(grid_tiles((int64_t)total)), dim3(kThreads), 0, s, hits, n_hits, n_virtual,
                     n_u, diag_level, starts, rows, n, cnt, (int64_t)total, out);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

extern "C" int fh_exact_link(const int64_t* group_a, int64_t n_a, const int64_t* starts_b, const int64_t* rows_b,
                             int64_t n_groups, int64_t n_b, uint64_t* out, int64_t capacity, int64_t* n_out,
                             fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n_out != nullptr, "fh_exact_link: n_out is null");
  FH_REQUIRE(n_a >= 0 && n_groups >= 0 && n_b >= 0, "fh_exact_link: bad sizes");
  (void)n_groups;
  *n_out = 0;
  if (n_a == 0) return FH_OK;
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  uint64_t *cnt = nullptr, *tot = nullptr;
  FH_TRY(scr.alloc(&cnt, (size_t)n_a));
  FH_TRY(scr.alloc(&tot, 1));
  hipLaunchKernelGGL(k_exact_link_count, dim3(grid_for(n_a)), dim3(kThreads), 0, s, group_a, n_a, starts_b, cnt);
  FH_HIP(hipGetLastError());
  uint64_t total = 0;
  FH_TRY(scan_total(cnt, n_a, tot, &total, s));
  *n_out = (int64_t)total;
  if ((int64_t)total > capacity || total == 0) return FH_OK;
  hipLaunchKernelGGL(k_exact_link, dim3(grid_tiles((int64_t)total)), dim3(kThreads), 0, s, group_a, n_a, starts_b,
                     rows_b, n_b, cnt, (int64_t)total, out);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

extern "C" int fh_exact_dedup(const int64_t* group, int64_t n, const int64_t* starts, const int64_t* rows,
                              int64_t n_groups, uint64_t* out, int64_t capacity, int64_t* n_out,
                              fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n_out != nullptr, "fh_exact_dedup: n_out is null");
  FH_REQUIRE(n >= 0 && n_groups >= 0, "fh_exact_dedup: bad sizes");
  *n_out = 0;
  if (n == 0) return FH_OK;
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  uint64_t *cnt = nullptr, *tot = nullptr;
  FH_TRY(scr.alloc(&cnt, (size_t)n));
  FH_TRY(scr.alloc(&tot, 1));
  hipLaunchKernelGGL(k_group_rank, dim3(grid_for(n)), dim3(kThreads), 0, s, group, starts, rows, n, n_groups, cnt);
  FH_HIP(hipGetLastError());
  uint64_t total = 0;
  FH_TRY(scan_total(cnt, n, tot, &total, s));
  *n_out = (int64_t)total;
  if ((int64_t)total > capacity || total == 0) return FH_OK;
  hipLaunchKernelGGL(k_exact_dedup, dim3(grid_tiles((int64_t)total)), dim3(kThreads), 0, s, group, n, starts, rows,
                     cnt, (int64_t)total, out);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

extern "C" int fh_split_levels(const uint64_t* keys, int64_t n, int n_levels, int64_t* ids, int64_t* level_offsets,
                               fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n >= 0 && n_levels >= 1 && n_levels <= 4, "fh_split_levels: bad arguments");
  FH_REQUIRE(n < (int64_t)0xFFFFFFFF, "fh_split_levels: list too long");
  FH_REQUIRE(level_offsets != nullptr, "fh_split_levels: level_offsets is null");
  hipStream_t s = (hipStream_t)stream;
  level_offsets[0] = 0;
  level_offsets[1] = 0;
  if (n == 0) {
    for (int l = 1; l <= n_levels; ++l) level_offsets[l] = 0;
    return FH_OK;
  }
  Scratch scr(s);
  uint32_t *flags = nullptr, *tot = nullptr;
  FH_TRY(scr.alloc(&flags, (size_t)n));
  FH_TRY(scr.alloc(&tot, 4));
  for (int l = 1; l < n_levels; ++l) {
    hipLaunchKernelGGL(k_level_flags, dim3(grid_for(n)), dim3(kThreads), 0, s, keys, n, l, flags);
    FH_HIP(hipGetLastError());
    FH_TRY(exclusive_scan_u32(flags, flags, n, tot + (l - 1), s));
    hipLaunchKernelGGL(k_level_scatter, dim3(grid_for(n)), dim3(kThreads), 0, s, keys, n, l, flags,
                       (const uint32_t*)tot, ids);
    FH_HIP(hipGetLastError());
  }
  uint32_t c[4] = {0, 0, 0, 0};
  FH_HIP(hipMemcpyAsync(c, tot, sizeof(c), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  int64_t base = 0;
  for (int l = 1; l < n_levels; ++l) {
    base += c[l - 1];
    level_offsets[l + 1] = base;
  }
  return FH_OK;
}

extern "C" int fh_sorted_member_mask(const int64_t* arr1, int64_t n1, const int64_t* arr2, int64_t n2, int inside,
                                     uint8_t* mask, fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n1 >= 0 && n2 >= 0, "fh_sorted_member_mask: bad sizes");
  if (n1 == 0) return FH_OK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_member_mask, dim3(grid_for(n1)), dim3(kThreads), 0, s, arr1, n1, arr2, n2, inside, mask);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

static int pattern_merge_impl(int n_vars, const uint64_t* const* var_keys, const int64_t* var_sizes,
                              const int32_t* var_levels, int n_imp, const int64_t* const* imp_a,
                              const int64_t* const* imp_b, int64_t id_div, int64_t* ids, int64_t capacity,
                              int64_t* pattern_offsets, int64_t* n_union, hipStream_t s,
                              const int64_t* const* lid_a = nullptr, const int64_t* const* lid_b = nullptr) {
  FH_REQUIRE(n_vars >= 1 && n_vars <= kMaxVars, "fh_pattern_merge: 1..%d listed variables supported", kMaxVars);
  FH_REQUIRE(n_imp >= 0 && n_imp <= kMaxVars, "fh_pattern_merge: 0..%d row-id variables supported", kMaxVars);
  FH_REQUIRE(pattern_offsets != nullptr && n_union != nullptr, "fh_pattern_merge: null host outputs");
  MergeArgs m;
  m.n_vars = n_vars;
  m.n_imp = n_imp;
  m.n_lid = 0;
  for (int v = 0; v < kMaxVars; ++v) {
    m.lid_a[v] = nullptr;
    m.lid_b[v] = nullptr;
  }
  for (int v = 0; v < n_vars && lid_a && lid_b; ++v) {
    if (lid_a[v] == nullptr || lid_b[v] == nullptr) continue;
    FH_REQUIRE(var_levels[v] == 2, "fh_pattern_merge_listed: variable %d with row ids must have 2 levels", v);
    m.lid_a[v] = lid_a[v];
    m.lid_b[v] = lid_b[v];
    ++m.n_lid;
  }
  m.id_div = (uint64_t)(id_div > 0 ? id_div : 1);
  uint64_t n_patterns = 1;
  for (int v = 0; v < n_vars; ++v) {
    FH_REQUIRE(var_levels[v] >= 2 && var_levels[v] <= 4, "fh_pattern_merge: levels must be 2..4");
    FH_REQUIRE(var_sizes[v] >= 0 && var_sizes[v] < (int64_t)0xFFFFFFFF, "fh_pattern_merge: list too long");
    n_patterns *= (uint64_t)var_levels[v];
    FH_REQUIRE(n_patterns < (1ull << 31), "fh_pattern_merge: too many patterns");
    m.nlev[v] = (uint32_t)var_levels[v];
  }
  n_patterns <<= n_imp;
  FH_REQUIRE(n_patterns < (1ull << 31), "fh_pattern_merge: too many patterns");
  {
    uint32_t st = 1;
    for (int q = n_imp - 1; q >= 0; --q) {
      m.imp_stride[q] = st;
      m.imp_a[q] = imp_a[q];
      m.imp_b[q] = imp_b[q];
      st *= 2;
    }
    for (int v = n_vars - 1; v >= 0; --v) {
      m.stride[v] = st;
      m.lists[v] = var_keys[v];
      m.sizes[v] = var_sizes[v];
      st *= (uint32_t)var_levels[v];
    }
  }
  for (uint64_t c = 0; c <= n_patterns; ++c) pattern_offsets[c] = 0;
  Scratch scr(s);
  int64_t nmax = 0;
  for (int v = 0; v < n_vars; ++v) nmax = std::max(nmax, var_sizes[v]);
  auto code_range = [&](int v, uint32_t* lo, uint32_t* hi) {
    *lo = m.stride[v];
    *hi = (v == 0 ? (uint32_t)n_patterns : m.stride[v - 1]) - 1;
  };
  // Per variable: up to 256 owned codes take the one-pass path (the owners'
  // counts per tile come out of the code kernel and one ranked scatter places
  // them at the running output base, which stays on the device); more codes
  // take the general path (flag, compact, radix sort by code), which reads
  // the base and its owner count back.
  std::vector<char> one_pass(n_vars);
  bool any_one = false, any_general = false;
  for (int v = 0; v < n_vars; ++v) {
    uint32_t lo, hi;
    code_range(v, &lo, &hi);
    one_pass[v] = hi - lo + 1 <= 256;
    any_one = any_one || one_pass[v];
    any_general = any_general || !one_pass[v];
  }
  uint32_t *codes = nullptr, *flags = nullptr, *ocodes = nullptr, *ocodes_tmp = nullptr, *tot = nullptr,
           *err = nullptr;
  uint64_t *oids_tmp = nullptr, *state = nullptr;  // state[0]: owners placed so far
  int64_t *bounds = nullptr, *tile_bounds = nullptr, *d_offsets = nullptr;
  uint32_t* tile_hist = nullptr;  // owners per (code, tile), for the one-pass code sort
  FH_TRY(scr.alloc(&codes, (size_t)std::max<int64_t>(1, nmax)));
  FH_TRY(scr.alloc(&tot, 1));
  FH_TRY(scr.alloc(&err, 1));
  FH_TRY(scr.alloc(&state, 1));
  FH_TRY(scr.alloc(&d_offsets, (size_t)n_patterns + 1));
  FH_TRY(scr.alloc(&tile_bounds, (size_t)((ceil_div(nmax, kMergeTile) + 1) * n_vars)));
  if (any_one) FH_TRY(scr.alloc(&tile_hist, (size_t)(256 * ceil_div(std::max<int64_t>(1, nmax), kMergeTile))));
  if (any_general) {
    FH_TRY(scr.alloc(&flags, (size_t)std::max<int64_t>(1, nmax)));
    FH_TRY(scr.alloc(&ocodes_tmp, (size_t)std::max<int64_t>(1, nmax)));
    FH_TRY(scr.alloc(&ocodes, (size_t)std::max<int64_t>(1, nmax)));
    FH_TRY(scr.alloc(&oids_tmp, (size_t)std::max<int64_t>(1, nmax)));
    FH_TRY(scr.alloc(&bounds, (size_t)n_patterns + 1));
  }
  m.err = err;
  FH_TRY(fill_u64(state, 0, 1, s));
  FH_TRY(fill_u32(err, 0, 1, s));
  const int code_bits = bits_for(n_patterns - 1);
  // Codes owned by variable v lie in [stride_v, stride_{v-1}); the output is
  // pattern-major, so variables are processed from the last to the first.
  for (int v = n_vars - 1; v >= 0; --v) {
    const int64_t n = var_sizes[v];
    uint32_t code_lo, code_hi;
    code_range(v, &code_lo, &code_hi);
    const uint32_t ncodes = code_hi - code_lo + 1;
    const int64_t ntiles = std::max<int64_t>(1, ceil_div(n, kMergeTile));
    if (n > 0) {
      hipLaunchKernelGGL(k_tile_bounds, dim3((unsigned)ceil_div((ntiles + 1) * n_vars, kThreads)), dim3(kThreads), 0,
                         s, m, v, ntiles, tile_bounds);
      FH_HIP(hipGetLastError());
    }
    if (one_pass[v]) {
      if (n > 0) {
        hipLaunchKernelGGL(k_pattern_codes, dim3((unsigned)ntiles), dim3(kThreads), 0, s, m, v,
                           (const int64_t*)tile_bounds, codes, code_lo, (int)ncodes, tile_hist);
        FH_HIP(hipGetLastError());
        FH_TRY(exclusive_scan_u32(tile_hist, tile_hist, (int64_t)ncodes * ntiles, tot, s));
        hipLaunchKernelGGL(k_code_scatter, dim3((unsigned)ntiles), dim3(kThreads), 0, s, codes, m.lists[v], n,
                           code_lo, (int)ncodes, (const uint32_t*)tile_hist, (const uint32_t*)tot,
                           (const uint64_t*)state, capacity, ids);
        FH_HIP(hipGetLastError());
      } else {
        FH_TRY(fill_u32(tot, 0, 1, s));
      }
      hipLaunchKernelGGL(k_code_offsets, dim3((unsigned)ceil_div(ncodes, kThreads)), dim3(kThreads), 0, s,
                         n > 0 ? (const uint32_t*)tile_hist : (const uint32_t*)nullptr, ntiles,
                         (const uint32_t*)tot, (const uint64_t*)state, capacity, code_lo, code_hi, d_offsets);
      FH_HIP(hipGetLastError());
    } else {
      uint64_t h_base = 0;
      FH_HIP(hipMemcpyAsync(&h_base, state, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
      FH_HIP(hipStreamSynchronize(s));
      const int64_t base = (int64_t)h_base;
      uint32_t owners = 0;
      if (n > 0) {
        hipLaunchKernelGGL(k_pattern_codes, dim3((unsigned)ntiles), dim3(kThreads), 0, s, m, v,
                           (const int64_t*)tile_bounds, codes, code_lo, 0, (uint32_t*)nullptr);
        FH_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_owner_flags, dim3(grid_for(n)), dim3(kThreads), 0, s, codes, n, flags);
        FH_HIP(hipGetLastError());
        FH_TRY(exclusive_scan_u32(flags, flags, n, tot, s));
        FH_HIP(hipMemcpyAsync(&owners, tot, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        FH_HIP(hipStreamSynchronize(s));
      } else {
        FH_TRY(fill_u32(tot, 0, 1, s));
      }
      if (base + (int64_t)owners <= capacity && owners > 0) {
        uint64_t* oids = reinterpret_cast<uint64_t*>(ids + base);
        hipLaunchKernelGGL(k_owner_scatter, dim3(grid_for(n)), dim3(kThreads), 0, s, m.lists[v], codes, n, flags,
                           ocodes_tmp, oids_tmp);
        FH_HIP(hipGetLastError());
        FH_TRY(radix_sort_u32_into(ocodes_tmp, ocodes, oids_tmp, oids, owners, 0, code_bits, s));
        hipLaunchKernelGGL(k_code_bounds, dim3((unsigned)ceil_div(ncodes, kThreads)), dim3(kThreads), 0, s, ocodes,
                           (int64_t)owners, code_lo, code_hi, bounds);
        FH_HIP(hipGetLastError());
        std::vector<int64_t> hb(ncodes);
        FH_HIP(hipMemcpyAsync(hb.data(), bounds, ncodes * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        FH_HIP(hipStreamSynchronize(s));
        for (uint32_t c = 0; c < ncodes; ++c) pattern_offsets[code_lo + c] = base + hb[c];
      } else {
        for (uint32_t c = code_lo; c <= code_hi; ++c) pattern_offsets[c] = base;
      }
    }
    hipLaunchKernelGGL(k_advance_base, dim3(1), dim3(64), 0, s, state, (const uint32_t*)tot);
    FH_HIP(hipGetLastError());
  }
  // One readback: the final base, the error flag and the one-pass offsets.
  uint64_t h_base = 0;
  uint32_t h_err = 0;
  std::vector<int64_t> d_off_host;
  if (any_one) {
    d_off_host.resize((size_t)n_patterns + 1);
    FH_HIP(hipMemcpyAsync(d_off_host.data(), d_offsets, ((size_t)n_patterns + 1) * sizeof(int64_t),
                          hipMemcpyDeviceToHost, s));
  }
  FH_HIP(hipMemcpyAsync(&h_base, state, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipMemcpyAsync(&h_err, err, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  if (h_err) {
    set_error("fh_pattern_merge: a level key has level 0 or a level not below its variable's level count");
    return FH_ERR_ARG;
  }
  for (int v = 0; v < n_vars; ++v) {
    if (!one_pass[v]) continue;
    uint32_t lo, hi;
    code_range(v, &lo, &hi);
    for (uint32_t c = lo; c <= hi; ++c) pattern_offsets[c] = d_off_host[c];
  }
  pattern_offsets[0] = 0;
  pattern_offsets[n_patterns] = (int64_t)h_base;
  *n_union = (int64_t)h_base;
  return FH_OK;
}

extern "C" int fh_pattern_merge(int n_vars, const uint64_t* const* var_keys, const int64_t* var_sizes,
                                const int32_t* var_levels, int64_t* ids, int64_t capacity, int64_t* pattern_offsets,
                                int64_t* n_union, fh_stream_t stream) {
  clear_error();
  return pattern_merge_impl(n_vars, var_keys, var_sizes, var_levels, 0, nullptr, nullptr, 1, ids, capacity,
                            pattern_offsets, n_union, (hipStream_t)stream);
}

extern "C" int fh_pattern_merge_rowids(int n_vars, const uint64_t* const* var_keys, const int64_t* var_sizes,
                                       const int32_t* var_levels, int n_exact, const int64_t* const* ids_a,
                                       const int64_t* const* ids_b, int64_t id_div, int64_t* ids,
                                       int64_t capacity, int64_t* pattern_offsets, int64_t* n_union,
                                       fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(id_div > 0, "fh_pattern_merge_rowids: id_div must be positive");
  return pattern_merge_impl(n_vars, var_keys, var_sizes, var_levels, n_exact, ids_a, ids_b, id_div, ids, capacity,
                            pattern_offsets, n_union, (hipStream_t)stream);
}

extern "C" int fh_pattern_merge_listed(int n_vars, const uint64_t* const* var_keys, const int64_t* var_sizes,
                                       const int32_t* var_levels, const int64_t* const* ids_a,
                                       const int64_t* const* ids_b, int64_t id_div, int64_t* ids, int64_t capacity,
                                       int64_t* pattern_offsets, int64_t* n_union, fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(id_div > 0, "fh_pattern_merge_listed: id_div must be positive");
  FH_REQUIRE(ids_a != nullptr && ids_b != nullptr, "fh_pattern_merge_listed: null row-id arrays");
  return pattern_merge_impl(n_vars, var_keys, var_sizes, var_levels, 0, nullptr, nullptr, id_div, ids, capacity,
                            pattern_offsets, n_union, (hipStream_t)stream, ids_a, ids_b);
}

extern "C" int fh_pair_posteriors(const int64_t* pattern_offsets, int n_patterns, const double* ksi, double* post,
                                  int64_t n_pairs, fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n_patterns >= 1 && n_pairs >= 0, "fh_pair_posteriors: bad sizes");
  if (n_pairs == 0) return FH_OK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pair_posteriors, dim3(grid_for(n_pairs)), dim3(kThreads), 0, s, pattern_offsets, n_patterns,
                     ksi, post, n_pairs);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

// Sorts n float64 values (NaNs last) into vs with their positions in perm;
// returns the number of non-NaN values. Synchronises once.
static int sort_values(const double* v, int64_t n, Scratch& scr, double** vs, uint64_t** perm, int64_t* n_valid,
                       hipStream_t s) {
  uint64_t *keys = nullptr, *keys_tmp = nullptr, *pm = nullptr, *pm_tmp = nullptr, *small = nullptr;
  double* sorted = nullptr;
  FH_TRY(scr.alloc(&keys, (size_t)n));
  FH_TRY(scr.alloc(&keys_tmp, (size_t)n));
  FH_TRY(scr.alloc(&pm, (size_t)n));
  FH_TRY(scr.alloc(&pm_tmp, (size_t)n));
  FH_TRY(scr.alloc(&sorted, (size_t)n));
  FH_TRY(scr.alloc(&small, 4));
  FH_TRY(fill_u64(small, ~0ull, 1, s));
  FH_TRY(fill_u64(small + 1, 0, 3, s));
  hipLaunchKernelGGL(k_num_keys, dim3(grid_for(n, 1024)), dim3(kThreads), 0, s, v, n, keys, pm,
                     reinterpret_cast<unsigned long long*>(small));
  FH_HIP(hipGetLastError());
  uint64_t h[3] = {0, 0, 0};
  FH_HIP(hipMemcpyAsync(h, small, sizeof(h), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  const int bits = bits_for(h[0] ^ h[1]);
  if (bits > 0) FH_TRY(radix_sort_u64(keys, keys_tmp, pm, pm_tmp, n, 0, bits, s));
  const int64_t nv = n - (int64_t)h[2];
  if (nv > 0) {
    hipLaunchKernelGGL(k_num_gather, dim3(grid_for(nv)), dim3(kThreads), 0, s, v, pm, nv, sorted);
    FH_HIP(hipGetLastError());
  }
  *vs = sorted;
  *perm = pm;
  *n_valid = nv;
  return FH_OK;
}

extern "C" int fh_numeric_link(const double* values_a, int64_t n_a, const double* values_b, int64_t n_b, double d2,
                               double d1, uint64_t* out, int64_t capacity, int64_t* n_out, fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n_out != nullptr, "fh_numeric_link: n_out is null");
  FH_REQUIRE(n_a >= 0 && n_b >= 0, "fh_numeric_link: bad sizes");
  FH_REQUIRE(d2 >= 0.0 && d1 >= d2, "fh_numeric_link: need 0 <= d2 <= d1");
  *n_out = 0;
  if (n_a == 0 || n_b == 0) return FH_OK;
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  double* vs = nullptr;
  uint64_t* perm = nullptr;
  int64_t nbv = 0;
  FH_TRY(sort_values(values_b, n_b, scr, &vs, &perm, &nbv, s));
  uint64_t *cnt = nullptr, *tot = nullptr;
  int64_t* win = nullptr;
  FH_TRY(scr.alloc(&cnt, (size_t)n_a));
  FH_TRY(scr.alloc(&tot, 1));
  FH_TRY(scr.alloc(&win, 4 * (size_t)n_a));
  hipLaunchKernelGGL(k_num_link_count, dim3(grid_for(n_a)), dim3(kThreads), 0, s, values_a, n_a, (const double*)vs, nbv,
                     d2, d1, win, cnt);
  FH_HIP(hipGetLastError());
  uint64_t total = 0;
  FH_TRY(scan_total(cnt, n_a, tot, &total, s));
  *n_out = (int64_t)total;
  if ((int64_t)total > capacity || total == 0) return FH_OK;
  hipLaunchKernelGGL(k_num_link, dim3(grid_tiles((int64_t)total)), dim3(kThreads), 0, s, (const int64_t*)win, n_a,
                     (const uint64_t*)perm, n_b, cnt, (int64_t)total, out);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

extern "C" int fh_numeric_dedup(const double* values, int64_t n, double d2, double d1, uint64_t* out,
                                int64_t capacity, int64_t* n_out, fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n_out != nullptr, "fh_numeric_dedup: n_out is null");
  FH_REQUIRE(n >= 0, "fh_numeric_dedup: bad size");
  FH_REQUIRE(d2 >= 0.0 && d1 >= d2, "fh_numeric_dedup: need 0 <= d2 <= d1");
  *n_out = 0;
  if (n < 2) return FH_OK;
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  double* vs = nullptr;
  uint64_t* perm = nullptr;
  int64_t nv = 0;
  FH_TRY(sort_values(values, n, scr, &vs, &perm, &nv, s));
  if (nv < 2) return FH_OK;
  uint64_t *cnt = nullptr, *tot = nullptr;
  int64_t* win = nullptr;
  FH_TRY(scr.alloc(&cnt, (size_t)nv));
  FH_TRY(scr.alloc(&tot, 1));
  FH_TRY(scr.alloc(&win, 2 * (size_t)nv));
  hipLaunchKernelGGL(k_num_dedup_count, dim3(grid_for(nv)), dim3(kThreads), 0, s, (const double*)vs, nv, d2, d1, win,
                     cnt);
  FH_HIP(hipGetLastError());
  uint64_t total = 0;
  FH_TRY(scan_total(cnt, nv, tot, &total, s));
  *n_out = (int64_t)total;
  if ((int64_t)total > capacity || total == 0) return FH_OK;
  hipLaunchKernelGGL(k_num_dedup, dim3(grid_tiles((int64_t)total)), dim3(kThreads), 0, s, (const int64_t*)win, nv,
                     (const uint64_t*)perm, n, cnt, (int64_t)total, out);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

extern "C" int fh_pattern_add_missing(const int64_t* ids, const int64_t* pattern_offsets, int n_patterns, int64_t n,
                                      const uint32_t* miss_a, const uint32_t* miss_b, int64_t id_div, int64_t max_id,
                                      int n_vars, const int64_t* add, int64_t* ids_out, int64_t* offsets_out,
                                      fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n_patterns >= 1 && n >= 0 && id_div > 0 && max_id >= 0, "fh_pattern_add_missing: bad sizes");
  FH_REQUIRE(n_vars >= 1 && n_vars <= kMaxMissVars, "fh_pattern_add_missing: 1..%d variables", kMaxMissVars);
  FH_REQUIRE(offsets_out != nullptr && add != nullptr, "fh_pattern_add_missing: null host array");
  MissingAdd ma;
  ma.n_vars = n_vars;
  int64_t max_add = 0;
  for (int v = 0; v < n_vars; ++v) {
    FH_REQUIRE(add[v] >= 0 && add[v] < n_patterns, "fh_pattern_add_missing: digit weight out of range");
    ma.add[v] = (uint32_t)add[v];
    max_add += add[v];
  }
  FH_REQUIRE(max_add < n_patterns, "fh_pattern_add_missing: digit weights exceed the pattern count");
  const int id_bits = std::max(1, bits_for((uint64_t)max_id));
  const int code_bits = bits_for((uint64_t)n_patterns);
  FH_REQUIRE(id_bits + code_bits <= 64, "fh_pattern_add_missing: ids of %d bits and %d patterns do not fit 64 bits",
             id_bits, n_patterns);
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    for (int c = 0; c <= n_patterns; ++c) offsets_out[c] = 0;
    return FH_OK;
  }
  Scratch scr(s);
  uint64_t *keys = nullptr, *keys_tmp = nullptr;
  int64_t* d_offs = nullptr;
  FH_TRY(scr.alloc(&keys, (size_t)n));
  FH_TRY(scr.alloc(&keys_tmp, (size_t)n));
  FH_TRY(scr.alloc(&d_offs, (size_t)n_patterns + 1));
  hipLaunchKernelGGL(k_missing_recode, dim3(grid_for(n)), dim3(kThreads), 0, s, ids, n, pattern_offsets, n_patterns,
                     miss_a, miss_b, (uint64_t)id_div, ma, id_bits, keys);
  FH_HIP(hipGetLastError());
  FH_TRY(radix_sort_u64(keys, keys_tmp, nullptr, nullptr, n, 0, id_bits + code_bits, s));
  hipLaunchKernelGGL(k_missing_split, dim3(grid_for(n)), dim3(kThreads), 0, s, (const uint64_t*)keys, n, id_bits,
                     ids_out);
  FH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_missing_offsets, dim3((unsigned)ceil_div((int64_t)n_patterns + 1, kThreads)), dim3(kThreads), 0,
                     s, (const uint64_t*)keys, n, n_patterns, id_bits, d_offs);
  FH_HIP(hipGetLastError());
  FH_HIP(hipMemcpyAsync(offsets_out, d_offs, ((size_t)n_patterns + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  return FH_OK;
}
