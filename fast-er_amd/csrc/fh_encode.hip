// Device grouping of a string column by value (unique values + their rows).
//
// The reference takes the unique values of every column on the host with
// np.unique over Python strings (faster/comparison.py:377,395,
// faster/deduplication.py:420) and orders the rows of each value with a
// stable argsort of the inverse (faster/comparison.py:380-391). Here the rows
// are sorted on the device by their bytes: a stable LSD radix sort over the
// big-endian 8-byte words of the zero-padded strings, least significant word
// first, after a first pass on the length (so a string sorts before its
// extensions by NUL bytes). The order is byte-lexicographic with shorter
// prefixes first, which for UTF-8 is the code-point order np.unique gives,
// and rows of one value stay ascending because every pass is stable and the
// sort starts from the identity permutation.
#include <algorithm>
#include <vector>

#include "fh_common.h"

namespace fh {
namespace {

constexpr int kEncThreads = 256;

// Strings per word count min(ceil(len / 8), kWordBins - 1), for the suffix
// sorts of the word passes (see word_suffixes).
constexpr int kWordBins = 65;

__device__ __forceinline__ void count_words(unsigned int* hw, uint64_t len) {
  const uint64_t words = (len + 7) / 8;
  atomicAdd(&hw[words < (uint64_t)(kWordBins - 1) ? words : (uint64_t)(kWordBins - 1)], 1u);
}

__device__ __forceinline__ void flush_words(const unsigned int* hw, unsigned long long* wcount) {
  for (int j = threadIdx.x; j < kWordBins; j += blockDim.x)
    if (hw[j]) atomicAdd(wcount + j, (unsigned long long)hw[j]);
}

__global__ void __launch_bounds__(kEncThreads) k_enc_init(const int64_t* __restrict__ off, int64_t n,
                                                          uint64_t* __restrict__ perm, uint64_t* __restrict__ lens,
                                                          unsigned long long* __restrict__ max_len,
                                                          unsigned long long* __restrict__ wcount) {
  __shared__ unsigned int hw[kWordBins];
  for (int j = threadIdx.x; j < kWordBins; j += blockDim.x) hw[j] = 0;
  __syncthreads();
  unsigned long long m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t len = (uint64_t)(off[i + 1] - off[i]);
    perm[i] = (uint64_t)i;
    lens[i] = len;
    m = m > len ? m : len;
    count_words(hw, len);
  }
  m = fhd::block_reduce(m, fhd::OpMax());  // its barriers also order the LDS counts
  if (threadIdx.x == 0) atomicMax(max_len, m);
  flush_words(hw, wcount);
}

// keys[i] = bytes [8 * word, 8 * word + 8) of row perm[i], zero-padded,
// first byte most significant.
__global__ void __launch_bounds__(kEncThreads) k_enc_word_keys(const uint8_t* __restrict__ chars,
                                                               const int64_t* __restrict__ off,
                                                               const uint64_t* __restrict__ perm, int64_t n, int word,
                                                               uint64_t* __restrict__ keys) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t r = perm[i];
    const int64_t o = off[r];
    const int64_t len = off[r + 1] - o;
    const int64_t b0 = (int64_t)word * 8;
    uint64_t w = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint64_t byte = b0 + k < len ? (uint64_t)chars[o + b0 + k] : 0ull;
      w = (w << 8) | byte;
    }
    keys[i] = w;
  }
}

__device__ __forceinline__ bool same_string(const uint8_t* __restrict__ chars, const int64_t* __restrict__ off,
                                            uint64_t x, uint64_t y) {
  const int64_t ox = off[x], oy = off[y];
  const int64_t len = off[x + 1] - ox;
  if (len != off[y + 1] - oy) return false;
  for (int64_t k = 0; k < len; ++k)
    if (chars[ox + k] != chars[oy + k]) return false;
  return true;
}

// heads[i] = 1 where sorted position i starts a new value; rows[i] = perm[i].
__global__ void __launch_bounds__(kEncThreads) k_enc_heads(const uint8_t* __restrict__ chars,
                                                           const int64_t* __restrict__ off,
                                                           const uint64_t* __restrict__ perm, int64_t n,
                                                           uint64_t* __restrict__ heads, int64_t* __restrict__ rows) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t r = perm[i];
    rows[i] = (int64_t)r;
    heads[i] = (i == 0 || !same_string(chars, off, r, perm[i - 1])) ? 1ull : 0ull;
  }
}

// For every head: starts[g] = i and ulen[g] = its length (g = exclusive scan
// of the heads). The closing entry starts[n_unique] = n.
__global__ void __launch_bounds__(kEncThreads) k_enc_emit(const int64_t* __restrict__ off,
                                                          const uint64_t* __restrict__ perm,
                                                          const uint64_t* __restrict__ heads,
                                                          const uint64_t* __restrict__ gid, int64_t n,
                                                          const uint64_t* __restrict__ n_unique,
                                                          int64_t* __restrict__ starts, uint64_t* __restrict__ ulen) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (heads[i]) {
      const uint64_t g = gid[i];
      const uint64_t r = perm[i];
      starts[g] = i;
      ulen[g] = (uint64_t)(off[r + 1] - off[r]);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) starts[*n_unique] = n;
}

// Copies the bytes of every unique value (one thread per value).
__global__ void __launch_bounds__(kEncThreads) k_enc_copy(const uint8_t* __restrict__ chars,
                                                          const int64_t* __restrict__ off,
                                                          const int64_t* __restrict__ rows,
                                                          const int64_t* __restrict__ starts,
                                                          const int64_t* __restrict__ uoff, int64_t n_unique,
                                                          uint8_t* __restrict__ uchars) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n_unique;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = rows[starts[g]];
    const int64_t o = off[r];
    const int64_t len = off[r + 1] - o;
    uint8_t* dst = uchars + uoff[g];
    for (int64_t k = 0; k < len; ++k) dst[k] = chars[o + k];
  }
}

// Grid of the kernels that fold one global extreme or count (one atomic per
// workgroup).
constexpr unsigned kReduceGrid = 1024;

// Rows whose strings are longer than 8 * w bytes, for w = 0 .. words - 1,
// from the word counts (wcount[j]: strings of min(ceil(len / 8), 64) words).
// After the stable sort by length, the strings of at most 8 * w bytes form a
// prefix of the order, and their word-w key is 0 in every byte: a stable
// sort of the whole order by word w leaves that prefix first and sorts the
// rest, so each word pass sorts only the suffix of longer strings (a larger
// suffix, as for w >= 64, is equally correct).
std::vector<int64_t> word_suffixes(const unsigned long long* wcount, int words) {
  std::vector<int64_t> r((size_t)std::max(words, 0), 0);
  for (int w = 0; w < words; ++w) {
    int64_t c = 0;
    for (int j = std::min(w + 1, kWordBins - 1); j < kWordBins; ++j) c += (int64_t)wcount[j];
    r[w] = c;
  }
  return r;
}

unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kEncThreads), 8192)); }

// ------------------------- numeric value grouping --------------------------
__global__ void __launch_bounds__(kEncThreads) k_val_keys(const int64_t* __restrict__ values, int64_t n, int kind,
                                                          uint64_t* __restrict__ keys, uint64_t* __restrict__ perm,
                                                          unsigned long long* __restrict__ minmax) {
  unsigned long long lo = ~0ull, hi = 0ull;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = fhd::value_key(values[i], kind == FH_VALUES_F64);
    keys[i] = k;
    perm[i] = (uint64_t)i;
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
  lo = fhd::block_reduce(lo, fhd::OpMin());
  hi = fhd::block_reduce(hi, fhd::OpMax());
  if (threadIdx.x == 0) {
    atomicMin(minmax, lo);
    atomicMax(minmax + 1, hi);
  }
}

// heads[i] = 1 where sorted position i starts a new value; rows[i] = perm[i].
__global__ void __launch_bounds__(kEncThreads) k_val_heads(const uint64_t* __restrict__ keys,
                                                           const uint64_t* __restrict__ perm, int64_t n,
                                                           uint64_t* __restrict__ heads, int64_t* __restrict__ rows) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    rows[i] = (int64_t)perm[i];
    heads[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1ull : 0ull;
  }
}

// group[row] and starts from the scanned heads (gid = exclusive scan).
__global__ void __launch_bounds__(kEncThreads) k_val_emit(const int64_t* __restrict__ rows,
                                                          const uint64_t* __restrict__ heads,
                                                          const uint64_t* __restrict__ gid, int64_t n,
                                                          const uint64_t* __restrict__ n_groups,
                                                          int64_t* __restrict__ group, int64_t* __restrict__ starts) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = (int64_t)(gid[i] + heads[i]) - 1;
    group[rows[i]] = g;
    if (heads[i]) starts[g] = i;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) starts[*n_groups] = n;
}

// B side of a grouping of the concatenation [A; B]: flag[k] = 1 for the
// grouped entries that are B rows.
__global__ void __launch_bounds__(kEncThreads) k_side_flags(const int64_t* __restrict__ rows, int64_t n, int64_t n_a,
                                                            uint64_t* __restrict__ flag) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    flag[k] = rows[k] >= n_a ? 1ull : 0ull;
}

__global__ void __launch_bounds__(kEncThreads) k_side_emit(const int64_t* __restrict__ rows, int64_t n, int64_t n_a,
                                                           const uint64_t* __restrict__ pos,
                                                           int64_t* __restrict__ rows_b) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = rows[k];
    if (r >= n_a) rows_b[pos[k]] = r - n_a;
  }
}

// starts_b[g] = B entries before group g (pos at the group's first entry);
// with `group`, also the group of every row.
__global__ void __launch_bounds__(kEncThreads) k_side_starts(const int64_t* __restrict__ starts,
                                                             const int64_t* __restrict__ rows, int64_t n_groups,
                                                             int64_t n, const uint64_t* __restrict__ pos,
                                                             const uint64_t* __restrict__ total_b,
                                                             int64_t* __restrict__ starts_b,
                                                             int64_t* __restrict__ group) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n_groups; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k0 = starts[g], k1 = starts[g + 1];
    starts_b[g] = (int64_t)pos[k0];
    if (group)
      for (int64_t k = k0; k < k1; ++k) group[rows[k]] = g;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) starts_b[n_groups] = (int64_t)*total_b;
}

// ------------------- several string columns in one sort --------------------
// The rows of k columns form one index space (column c's row r is global
// index base[c] + r). One LSD sort over (column, string) orders every
// column's rows at once, so k columns take the launches of one.
constexpr int kMaxCols = 16;
struct ColSet {
  const uint8_t* chars[kMaxCols];
  const int64_t* off[kMaxCols];
  int64_t base[kMaxCols + 1];  // global index of each column's first row; base[k] = total rows
  int k;
};
struct ColOut {
  uint8_t* uchars[kMaxCols];
  int64_t* uoff[kMaxCols];
  int64_t* starts[kMaxCols];
  int64_t* rows[kMaxCols];
};

__device__ __forceinline__ int col_of(const ColSet& cs, int64_t g) {
  int c = 0;
  while (c + 1 < cs.k && g >= cs.base[c + 1]) ++c;
  return c;
}

__global__ void __launch_bounds__(kEncThreads) k_encm_init(ColSet cs, uint64_t* __restrict__ perm,
                                                           uint64_t* __restrict__ lens,
                                                           unsigned long long* __restrict__ max_len,
                                                           unsigned long long* __restrict__ wcount) {
  __shared__ unsigned int hw[kWordBins];
  for (int j = threadIdx.x; j < kWordBins; j += blockDim.x) hw[j] = 0;
  __syncthreads();
  unsigned long long m = 0;
  const int64_t n = cs.base[cs.k];
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int c = col_of(cs, g);
    const int64_t r = g - cs.base[c];
    const uint64_t len = (uint64_t)(cs.off[c][r + 1] - cs.off[c][r]);
    perm[g] = (uint64_t)g;
    lens[g] = len;
    m = m > len ? m : len;
    count_words(hw, len);
  }
  m = fhd::block_reduce(m, fhd::OpMax());  // its barriers also order the LDS counts
  if (threadIdx.x == 0) atomicMax(max_len, m);
  flush_words(hw, wcount);
}

// keys[i] = bytes [8 * word, 8 * word + 8) of global row perm[i] (zero-padded,
// first byte most significant), or its column when word < 0; i in [0, n).
__global__ void __launch_bounds__(kEncThreads) k_encm_keys(ColSet cs, const uint64_t* __restrict__ perm, int64_t n,
                                                           int word, uint64_t* __restrict__ keys) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = (int64_t)perm[i];
    const int c = col_of(cs, g);
    if (word < 0) {
      keys[i] = (uint64_t)c;
      continue;
    }
    const int64_t r = g - cs.base[c];
    const int64_t o = cs.off[c][r];
    const int64_t len = cs.off[c][r + 1] - o;
    const int64_t b0 = (int64_t)word * 8;
    uint64_t w = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint64_t byte = b0 + k < len ? (uint64_t)cs.chars[c][o + b0 + k] : 0ull;
      w = (w << 8) | byte;
    }
    keys[i] = w;
  }
}

// After the sort, column c's rows occupy sorted positions [base[c],
// base[c + 1]). heads[i] = 1 where position i starts a new value (of its
// column); rows of column c get their local ids.
__global__ void __launch_bounds__(kEncThreads) k_encm_heads(ColSet cs, ColOut out, const uint64_t* __restrict__ perm,
                                                            uint64_t* __restrict__ heads) {
  const int64_t n = cs.base[cs.k];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = col_of(cs, i);
    const int64_t r = (int64_t)perm[i] - cs.base[c];
    out.rows[c][i - cs.base[c]] = r;
    bool head = i == cs.base[c];
    if (!head) {
      const int64_t q = (int64_t)perm[i - 1] - cs.base[c];
      const int64_t* off = cs.off[c];
      const uint8_t* ch = cs.chars[c];
      const int64_t ox = off[r], oy = off[q];
      const int64_t len = off[r + 1] - ox;
      head = len != off[q + 1] - oy;
      for (int64_t k = 0; !head && k < len; ++k) head = ch[ox + k] != ch[oy + k];
    }
    heads[i] = head ? 1ull : 0ull;
  }
}

// First value id of every column (gid at its first sorted position; the
// value count for an empty trailing column).
__device__ __forceinline__ uint64_t col_g0(const ColSet& cs, const uint64_t* gid, const uint64_t* total, int c) {
  return cs.base[c] < cs.base[cs.k] ? gid[cs.base[c]] : *total;
}

// starts and the unique values' lengths (global value id g = gid[i]).
__global__ void __launch_bounds__(kEncThreads) k_encm_emit(ColSet cs, ColOut out, const uint64_t* __restrict__ perm,
                                                           const uint64_t* __restrict__ heads,
                                                           const uint64_t* __restrict__ gid,
                                                           const uint64_t* __restrict__ n_values,
                                                           uint64_t* __restrict__ ulen) {
  const int64_t n = cs.base[cs.k];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (!heads[i]) continue;
    const int c = col_of(cs, i);
    const uint64_t g = gid[i];
    const int64_t r = (int64_t)perm[i] - cs.base[c];
    out.starts[c][g - col_g0(cs, gid, n_values, c)] = i - cs.base[c];
    ulen[g] = (uint64_t)(cs.off[c][r + 1] - cs.off[c][r]);
  }
  if (blockIdx.x == 0 && (int)threadIdx.x < cs.k) {
    const int c = (int)threadIdx.x;
    const uint64_t g0 = col_g0(cs, gid, n_values, c);
    const uint64_t g1 = c + 1 < cs.k ? col_g0(cs, gid, n_values, c + 1) : *n_values;
    out.starts[c][g1 - g0] = cs.base[c + 1] - cs.base[c];
  }
}

// Per value: its column's offsets (uoff relative to the column's first
// value) and bytes; per column (first k threads): closing offset, value
// count and byte count into counts[2c], counts[2c + 1].
__global__ void __launch_bounds__(kEncThreads) k_encm_values(ColSet cs, ColOut out, const uint64_t* __restrict__ perm,
                                                             const uint64_t* __restrict__ gid,
                                                             const uint64_t* __restrict__ n_values,
                                                             const uint64_t* __restrict__ goff,
                                                             const uint64_t* __restrict__ total_bytes,
                                                             int64_t* __restrict__ counts) {
  const uint64_t nv = *n_values;
  __shared__ uint64_t g0[kMaxCols + 1];
  if ((int)threadIdx.x <= cs.k)
    g0[threadIdx.x] = (int)threadIdx.x < cs.k ? col_g0(cs, gid, n_values, (int)threadIdx.x) : nv;
  __syncthreads();
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nv; g += (uint64_t)gridDim.x * blockDim.x) {
    int c = 0;
    while (c + 1 < cs.k && g >= g0[c + 1]) ++c;
    const uint64_t gl = g - g0[c];
    const uint64_t o = goff[g] - goff[g0[c]];
    out.uoff[c][gl] = (int64_t)o;
    const int64_t pos = out.starts[c][gl] + cs.base[c];
    const int64_t r = (int64_t)perm[pos] - cs.base[c];
    const int64_t so = cs.off[c][r];
    const int64_t len = cs.off[c][r + 1] - so;
    for (int64_t k = 0; k < len; ++k) out.uchars[c][o + k] = cs.chars[c][so + k];
  }
  if (blockIdx.x == 0 && (int)threadIdx.x < cs.k) {
    const int c = (int)threadIdx.x;
    const uint64_t e = g0[c + 1] < nv ? goff[g0[c + 1]] : *total_bytes;
    const uint64_t b = g0[c] < nv ? goff[g0[c]] : *total_bytes;
    out.uoff[c][g0[c + 1] - g0[c]] = (int64_t)(e - b);
    counts[2 * c] = (int64_t)(g0[c + 1] - g0[c]);
    counts[2 * c + 1] = (int64_t)(e - b);
  }
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" int fh_group_strings(const uint8_t* chars, const int64_t* offsets, int64_t n, uint8_t* uchars,
                                int64_t* uoffsets, int64_t* starts, int64_t* rows, int64_t* n_unique,
                                int64_t* unique_bytes, fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n >= 0, "fh_group_strings: n < 0");
  FH_REQUIRE(n_unique && unique_bytes, "fh_group_strings: null output count");
  FH_REQUIRE(n < (int64_t)0xFFFFFFFF, "fh_group_strings: n = %lld exceeds 2^32 - 1", (long long)n);
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    const int64_t zero = 0;
    FH_HIP(hipMemcpyAsync(uoffsets, &zero, sizeof(int64_t), hipMemcpyHostToDevice, s));
    FH_HIP(hipMemcpyAsync(starts, &zero, sizeof(int64_t), hipMemcpyHostToDevice, s));
    FH_HIP(hipStreamSynchronize(s));
    *n_unique = 0;
    *unique_bytes = 0;
    return FH_OK;
  }
  Scratch scr(s);
  uint64_t *perm = nullptr, *perm_tmp = nullptr, *keys = nullptr, *keys_tmp = nullptr, *small = nullptr;
  FH_TRY(scr.alloc(&perm, (size_t)n));
  FH_TRY(scr.alloc(&perm_tmp, (size_t)n));
  FH_TRY(scr.alloc(&keys, (size_t)n + 1));
  FH_TRY(scr.alloc(&keys_tmp, (size_t)n + 1));
  // [0] max length, [1] unique count, [2] unique bytes, [4 ..] word counts
  FH_TRY(scr.alloc(&small, 4 + kWordBins));
  FH_TRY(fill_u64(small, 0, 4 + kWordBins, s));
  const unsigned grid = grid_for(n);

  // Pass 0: lengths (also the maximum length, which sets the word count,
  // and the strings per word count).
  hipLaunchKernelGGL(k_enc_init, dim3(std::min(grid, kReduceGrid)), dim3(kEncThreads), 0, s, offsets, n, perm, keys,
                     reinterpret_cast<unsigned long long*>(small), reinterpret_cast<unsigned long long*>(small + 4));
  FH_HIP(hipGetLastError());
  std::vector<unsigned long long> h_small(4 + kWordBins);
  FH_HIP(hipMemcpyAsync(h_small.data(), small, h_small.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  const uint64_t max_len = h_small[0];
  FH_TRY(radix_sort_u64(keys, keys_tmp, perm, perm_tmp, n, 0, std::max(1, bits_for(max_len)), s));

  // Words from the last to the first, each over the suffix of strings longer
  // than its first byte (word_suffixes); bytes past max_len are zero in
  // every key, so the low bits of a partial last word need no pass.
  const int words = (int)ceil_div((int64_t)max_len, 8);
  const std::vector<int64_t> suffix = word_suffixes(h_small.data() + 4, words);
  for (int w = words - 1; w >= 0; --w) {
    const int64_t m = suffix[w], s0 = n - m;
    if (m <= 0) continue;
    hipLaunchKernelGGL(k_enc_word_keys, dim3(grid_for(m)), dim3(kEncThreads), 0, s, chars, offsets, perm + s0, m, w,
                       keys + s0);
    FH_HIP(hipGetLastError());
    const int valid_bytes = (int)std::min<int64_t>(8, (int64_t)max_len - 8 * (int64_t)w);
    FH_TRY(radix_sort_u64(keys + s0, keys_tmp + s0, perm + s0, perm_tmp + s0, m, 8 * (8 - valid_bytes), 64, s));
  }

  // Value boundaries, value ids, starts and packed unique bytes.
  uint64_t* heads = keys;
  uint64_t* gid = keys_tmp;
  hipLaunchKernelGGL(k_enc_heads, dim3(grid), dim3(kEncThreads), 0, s, chars, offsets, perm, n, heads, rows);
  FH_HIP(hipGetLastError());
  FH_TRY(exclusive_scan_u64(heads, gid, n, small + 1, s));
  uint64_t* ulen = perm_tmp;
  hipLaunchKernelGGL(k_enc_emit, dim3(grid), dim3(kEncThreads), 0, s, offsets, perm, heads, gid, n, small + 1,
                     starts, ulen);
  FH_HIP(hipGetLastError());
  uint64_t nu = 0;
  FH_HIP(hipMemcpyAsync(&nu, small + 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  FH_TRY(exclusive_scan_u64(ulen, reinterpret_cast<uint64_t*>(uoffsets), (int64_t)nu, small + 2, s));
  FH_HIP(hipMemcpyAsync(uoffsets + nu, small + 2, sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_enc_copy, dim3(grid_for((int64_t)nu)), dim3(kEncThreads), 0, s, chars, offsets, rows, starts,
                     uoffsets, (int64_t)nu, uchars);
  FH_HIP(hipGetLastError());
  uint64_t ub = 0;
  FH_HIP(hipMemcpyAsync(&ub, small + 2, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  *n_unique = (int64_t)nu;
  *unique_bytes = (int64_t)ub;
  return FH_OK;
}

extern "C" int fh_group_strings_multi(int k, const uint8_t* const* chars, const int64_t* const* offsets,
                                      const int64_t* n, uint8_t* const* uchars, int64_t* const* uoffsets,
                                      int64_t* const* starts, int64_t* const* rows, int64_t* n_unique,
                                      int64_t* unique_bytes, fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(k >= 1 && k <= kMaxCols, "fh_group_strings_multi: 1..%d columns", kMaxCols);
  FH_REQUIRE(n_unique && unique_bytes, "fh_group_strings_multi: null output count");
  hipStream_t s = (hipStream_t)stream;
  ColSet cs{};
  ColOut out{};
  cs.k = k;
  cs.base[0] = 0;
  for (int c = 0; c < k; ++c) {
    FH_REQUIRE(n[c] >= 0, "fh_group_strings_multi: column %d has n < 0", c);
    cs.chars[c] = chars[c];
    cs.off[c] = offsets[c];
    cs.base[c + 1] = cs.base[c] + n[c];
    out.uchars[c] = uchars[c];
    out.uoff[c] = uoffsets[c];
    out.starts[c] = starts[c];
    out.rows[c] = rows[c];
  }
  const int64_t N = cs.base[k];
  FH_REQUIRE(N < (int64_t)0xFFFFFFFF, "fh_group_strings_multi: %lld rows exceed 2^32 - 1", (long long)N);
  if (N == 0) {
    const int64_t zero = 0;
    for (int c = 0; c < k; ++c) {
      FH_HIP(hipMemcpyAsync(uoffsets[c], &zero, sizeof(int64_t), hipMemcpyHostToDevice, s));
      FH_HIP(hipMemcpyAsync(starts[c], &zero, sizeof(int64_t), hipMemcpyHostToDevice, s));
      n_unique[c] = 0;
      unique_bytes[c] = 0;
    }
    FH_HIP(hipStreamSynchronize(s));
    return FH_OK;
  }
  Scratch scr(s);
  uint64_t *perm = nullptr, *perm_tmp = nullptr, *keys = nullptr, *keys_tmp = nullptr, *small = nullptr;
  int64_t* counts = nullptr;
  FH_TRY(scr.alloc(&perm, (size_t)N));
  FH_TRY(scr.alloc(&perm_tmp, (size_t)N));
  FH_TRY(scr.alloc(&keys, (size_t)N + 1));
  FH_TRY(scr.alloc(&keys_tmp, (size_t)N + 1));
  // [0] max length, [1] value count, [2] value bytes, [4 ..] word counts
  FH_TRY(scr.alloc(&small, 4 + kWordBins));
  FH_TRY(scr.alloc(&counts, 2 * (size_t)k));
  FH_TRY(fill_u64(small, 0, 4 + kWordBins, s));
  const unsigned grid = grid_for(N);
  hipLaunchKernelGGL(k_encm_init, dim3(std::min(grid, kReduceGrid)), dim3(kEncThreads), 0, s, cs, perm, keys,
                     reinterpret_cast<unsigned long long*>(small), reinterpret_cast<unsigned long long*>(small + 4));
  FH_HIP(hipGetLastError());
  std::vector<unsigned long long> h_small(4 + kWordBins);
  FH_HIP(hipMemcpyAsync(h_small.data(), small, h_small.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  const uint64_t max_len = h_small[0];
  // Length, then the words from the last to the first (each over the suffix
  // of strings longer than its first byte, word_suffixes), then the column.
  FH_TRY(radix_sort_u64(keys, keys_tmp, perm, perm_tmp, N, 0, std::max(1, bits_for(max_len)), s));
  const int words = (int)ceil_div((int64_t)max_len, 8);
  const std::vector<int64_t> suffix = word_suffixes(h_small.data() + 4, words);
  for (int w = words - 1; w >= 0; --w) {
    const int64_t m = suffix[w], s0 = N - m;
    if (m <= 0) continue;
    hipLaunchKernelGGL(k_encm_keys, dim3(grid_for(m)), dim3(kEncThreads), 0, s, cs, perm + s0, m, w, keys + s0);
    FH_HIP(hipGetLastError());
    const int valid_bytes = (int)std::min<int64_t>(8, (int64_t)max_len - 8 * (int64_t)w);
    FH_TRY(radix_sort_u64(keys + s0, keys_tmp + s0, perm + s0, perm_tmp + s0, m, 8 * (8 - valid_bytes), 64, s));
  }
  if (k > 1) {
    hipLaunchKernelGGL(k_encm_keys, dim3(grid), dim3(kEncThreads), 0, s, cs, perm, N, -1, keys);
    FH_HIP(hipGetLastError());
    FH_TRY(radix_sort_u64(keys, keys_tmp, perm, perm_tmp, N, 0, bits_for((uint64_t)k - 1), s));
  }
  uint64_t* heads = keys;
  uint64_t* gid = keys_tmp;
  hipLaunchKernelGGL(k_encm_heads, dim3(grid), dim3(kEncThreads), 0, s, cs, out, perm, heads);
  FH_HIP(hipGetLastError());
  FH_TRY(exclusive_scan_u64(heads, gid, N, small + 1, s));
  uint64_t* ulen = perm_tmp;  // value lengths, then their offsets in place
  FH_TRY(fill_u64(ulen, 0, N, s));
  hipLaunchKernelGGL(k_encm_emit, dim3(grid), dim3(kEncThreads), 0, s, cs, out, perm, heads, gid, small + 1, ulen);
  FH_HIP(hipGetLastError());
  // Up to N values: the entries past the value count are 0, so the scan's
  // total is the unique bytes.
  FH_TRY(exclusive_scan_u64(ulen, ulen, N, small + 2, s));
  hipLaunchKernelGGL(k_encm_values, dim3(grid), dim3(kEncThreads), 0, s, cs, out, perm, gid, small + 1, ulen,
                     small + 2, counts);
  FH_HIP(hipGetLastError());
  std::vector<int64_t> h(2 * (size_t)k);
  FH_HIP(hipMemcpyAsync(h.data(), counts, h.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  for (int c = 0; c < k; ++c) {
    n_unique[c] = h[2 * c];
    unique_bytes[c] = h[2 * c + 1];
  }
  return FH_OK;
}

extern "C" int fh_group_values(const int64_t* values, int64_t n, int kind, int64_t* group, int64_t* starts,
                               int64_t* rows, int64_t* n_groups, fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n >= 0 && n_groups != nullptr, "fh_group_values: bad arguments");
  FH_REQUIRE(kind == FH_VALUES_I64 || kind == FH_VALUES_F64, "fh_group_values: unknown value kind %d", kind);
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    const int64_t zero = 0;
    FH_HIP(hipMemcpyAsync(starts, &zero, sizeof(int64_t), hipMemcpyHostToDevice, s));
    FH_HIP(hipStreamSynchronize(s));
    *n_groups = 0;
    return FH_OK;
  }
  Scratch scr(s);
  uint64_t *keys = nullptr, *keys_tmp = nullptr, *perm = nullptr, *perm_tmp = nullptr, *small = nullptr;
  FH_TRY(scr.alloc(&keys, (size_t)n));
  FH_TRY(scr.alloc(&keys_tmp, (size_t)n));
  FH_TRY(scr.alloc(&perm, (size_t)n));
  FH_TRY(scr.alloc(&perm_tmp, (size_t)n));
  FH_TRY(scr.alloc(&small, 4));  // [0] min key, [1] max key, [2] group count
  FH_TRY(fill_u64(small, ~0ull, 1, s));
  FH_TRY(fill_u64(small + 1, 0, 3, s));
  const unsigned grid = grid_for(n);
  hipLaunchKernelGGL(k_val_keys, dim3(std::min(grid, kReduceGrid)), dim3(kEncThreads), 0, s, values, n, kind, keys, perm,
                     reinterpret_cast<unsigned long long*>(small));
  FH_HIP(hipGetLastError());
  uint64_t mm[2] = {0, 0};
  FH_HIP(hipMemcpyAsync(mm, small, sizeof(mm), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  // Keys between min and max share the bits above their highest differing bit.
  const int bits = bits_for(mm[0] ^ mm[1]);
  if (bits > 0) FH_TRY(radix_sort_u64(keys, keys_tmp, perm, perm_tmp, n, 0, bits, s));
  uint64_t* heads = keys_tmp;
  uint64_t* gid = perm_tmp;
  hipLaunchKernelGGL(k_val_heads, dim3(grid), dim3(kEncThreads), 0, s, keys, perm, n, heads, rows);
  FH_HIP(hipGetLastError());
  FH_TRY(exclusive_scan_u64(heads, gid, n, small + 2, s));
  hipLaunchKernelGGL(k_val_emit, dim3(grid), dim3(kEncThreads), 0, s, rows, heads, gid, n, small + 2, group, starts);
  FH_HIP(hipGetLastError());
  uint64_t ng = 0;
  FH_HIP(hipMemcpyAsync(&ng, small + 2, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  *n_groups = (int64_t)ng;
  return FH_OK;
}

extern "C" int fh_groups_side_b(const int64_t* starts, const int64_t* rows, int64_t n_groups, int64_t n_a,
                                int64_t n_b, int64_t* starts_b, int64_t* rows_b, int64_t* group,
                                fh_stream_t stream) {
  clear_error();
  FH_REQUIRE(n_groups >= 0 && n_a >= 0 && n_b >= 0, "fh_groups_side_b: bad sizes");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = n_a + n_b;
  Scratch scr(s);
  uint64_t *pos = nullptr, *total = nullptr;
  FH_TRY(scr.alloc(&pos, (size_t)std::max<int64_t>(1, n)));
  FH_TRY(scr.alloc(&total, 1));
  FH_TRY(fill_u64(total, 0, 1, s));
  if (n > 0) {
    hipLaunchKernelGGL(k_side_flags, dim3(grid_for(n)), dim3(kEncThreads), 0, s, rows, n, n_a, pos);
    FH_HIP(hipGetLastError());
    FH_TRY(exclusive_scan_u64(pos, pos, n, total, s));
    hipLaunchKernelGGL(k_side_emit, dim3(grid_for(n)), dim3(kEncThreads), 0, s, rows, n, n_a, pos, rows_b);
    FH_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_side_starts, dim3(grid_for(std::max<int64_t>(1, n_groups))), dim3(kEncThreads), 0, s, starts,
                     rows, n_groups, n, pos, total, starts_b, group);
  FH_HIP(hipGetLastError());
  return FH_OK;
}
