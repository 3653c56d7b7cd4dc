// Device-wide exclusive scan and stable LSD radix sort (8-bit digits).
//
// These take over the roles of CuPy's cumsum and sort in the reference
// (e.g. faster/comparison.py:443,456). Both are written for 64-wide
// wavefronts: the block scan combines wave-level shuffles with one LDS
// exchange, and the radix scatter ranks keys inside a wave with eight ballots
// (a match-any over the digit bits) before combining waves through LDS, which
// keeps the sort stable without per-thread digit counters; it then stages the
// digit-sorted tile in LDS so that each digit's keys are written as one run.
#include "fh_common.h"

namespace fh {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  const int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    T o = __shfl_up(v, off, 64);
    if (lane >= off) v += o;
  }
  return v;
}

// Exclusive scan of one value per thread across a block of THREADS threads;
// returns the exclusive prefix and writes the block total to *total.
template <typename T, int THREADS = kScanThreads>
__device__ __forceinline__ T block_exclusive_scan(T v, T* total) {
  __shared__ T wave_tot[THREADS / 64];
  const int lane = __lane_id();
  const int wave = threadIdx.x / 64;
  T inc = wave_inclusive_scan(v);
  if (lane == 63) wave_tot[wave] = inc;
  __syncthreads();
  T base = 0, all = 0;
#pragma unroll
  for (int w = 0; w < THREADS / 64; ++w) {
    T t = wave_tot[w];
    if (w < wave) base += t;
    all += t;
  }
  __syncthreads();
  *total = all;
  return base + inc - v;
}

template <typename T>
__global__ void __launch_bounds__(kScanThreads) k_scan_partials(const T* __restrict__ in, int64_t n,
                                                               T* __restrict__ partial) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  T s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    if (i < n) s += in[i];
  }
  T tot;
  block_exclusive_scan<T>(s, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

template <typename T>
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const T* in, int64_t n, const T* __restrict__ prefix,
                                                            T* out, T* total) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  T v[kScanItems];
  T s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    v[k] = (i < n) ? in[i] : T(0);
    s += v[k];
  }
  T tot;
  T run = block_exclusive_scan<T>(s, &tot);
  const T blk = prefix ? prefix[blockIdx.x] : T(0);
  run += blk;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = blk + tot;
}

// The whole scan in one workgroup, a tile at a time with a running carry:
// one launch instead of three for the short scans (radix-sort digit counts
// of small inputs, per-variable flags) whose cost is launch latency. Each
// tile is a dependent round trip, so longer scans take the three-launch
// path (a 1,024-thread form of this kernel over 50k counts took 47 us
// against about 10 us of kernel time for the three launches).
constexpr int kScanSingleThreads = kScanThreads;
constexpr int64_t kScanSingleTile = kScanTile;
constexpr int64_t kScanSingleMax = 16 * kScanSingleTile;

template <typename T>
__global__ void __launch_bounds__(kScanSingleThreads) k_scan_single(const T* in, int64_t n, T* out, T* total) {
  T carry = 0;
  for (int64_t t0 = 0; t0 < n; t0 += kScanSingleTile) {
    const int64_t base = t0 + (int64_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    T sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const int64_t i = base + k;
      v[k] = (i < n) ? in[i] : T(0);
      sum += v[k];
    }
    T tot;
    T run = block_exclusive_scan<T, kScanSingleThreads>(sum, &tot) + carry;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const int64_t i = base + k;
      if (i < n) out[i] = run;
      run += v[k];
    }
    carry += tot;
  }
  if (total && threadIdx.x == 0) *total = carry;
}

template <typename T>
int exclusive_scan_impl(const T* in, T* out, int64_t n, T* total, hipStream_t s) {
  if (n <= 0) {
    if (total) {
      if (sizeof(T) == 8)
        FH_TRY(fill_u64(reinterpret_cast<uint64_t*>(total), 0, 1, s));
      else
        FH_TRY(fill_u32(reinterpret_cast<uint32_t*>(total), 0, 1, s));
    }
    return FH_OK;
  }
  const int64_t nb = ceil_div(n, kScanTile);
  if (n <= kScanSingleMax) {
    hipLaunchKernelGGL(k_scan_single<T>, dim3(1), dim3(kScanSingleThreads), 0, s, in, n, out, total);
    FH_HIP(hipGetLastError());
    return FH_OK;
  }
  Scratch scr(s);
  T* partial = nullptr;
  FH_TRY(scr.alloc(&partial, (size_t)nb));
  hipLaunchKernelGGL(k_scan_partials<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, partial);
  FH_HIP(hipGetLastError());
  FH_TRY(exclusive_scan_impl<T>(partial, partial, nb, (T*)nullptr, s));
  hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n,
                     (const T*)partial, out, total);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

// ---------------------------------------------------------------------------
// Radix sort
// ---------------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096 keys per tile

template <typename K>
__global__ void __launch_bounds__(kSortThreads) k_rs_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                         int64_t ntiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int r = 0; r < kSortItems; ++r) {
    int64_t e = base + (int64_t)r * kSortThreads + threadIdx.x;
    if (e < n) {
      uint32_t d = (uint32_t)((keys[e] >> shift) & 255u);
      atomicAdd(&h[d], 1u);
    }
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// One tile of kSortTile keys per workgroup. Wave w owns the tile's elements
// [w * kWaveItems, (w + 1) * kWaveItems) and reads them 64 at a time, so the
// tile's element order is (wave, round, lane) and every read is one
// contiguous 512-byte (u64) or 256-byte (u32) run. Each wave ranks its keys
// within their digit with eight ballots and one LDS atomic per digit group on
// a wave-private digit counter; the waves' counts then combine into each
// key's slot in the digit-sorted tile, which is staged in LDS and written
// out in slot order: keys of one digit leave as one contiguous run instead of
// one scattered write per key. Values are not staged: each slot keeps the
// key's position in the tile (16 bits), and the value is gathered from the
// tile's 32 KB of values, which the L2 holds, as it is written out.
constexpr int kWaveItems = kSortTile / kSortWaves;  // 1024
constexpr int kWaveRounds = kWaveItems / 64;       // 16

template <typename K, bool kVals>
__global__ void __launch_bounds__(kSortThreads) k_rs_scatter(const K* __restrict__ keys_in,
                                                            const uint64_t* __restrict__ vals_in, int64_t n,
                                                            int shift, int64_t ntiles,
                                                            const uint32_t* __restrict__ hist_scanned,
                                                            K* __restrict__ keys_out,
                                                            uint64_t* __restrict__ vals_out) {
  static_assert(kSortThreads == kScanThreads, "the digit scan uses the block scan");
  __shared__ uint32_t goff[256];              // first output index of digit d for this tile
  __shared__ uint32_t loff[256];              // first tile slot of digit d
  __shared__ uint32_t wcnt[kSortWaves][256];  // per-wave digit counts, then per-wave starts
  __shared__ K skeys[kSortTile];
  __shared__ uint16_t sidx[kVals ? kSortTile : 1];  // tile position of the key in each slot
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  goff[tid] = hist_scanned[(int64_t)tid * ntiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < kSortWaves; ++w) wcnt[w][tid] = 0;
  __syncthreads();

  const uint64_t lt = fhd::lanemask_lt();
  const int64_t tile0 = (int64_t)blockIdx.x * kSortTile;
  const int64_t wbase = tile0 + (int64_t)wave * kWaveItems;
  uint32_t* cnt = wcnt[wave];
  K key[kWaveRounds];
  uint32_t rnk[kWaveRounds];
#pragma unroll
  for (int r = 0; r < kWaveRounds; ++r) {
    const int64_t e = wbase + r * 64 + lane;
    const bool valid = e < n;
    key[r] = valid ? keys_in[e] : K(0);
    const uint32_t d = (uint32_t)((key[r] >> shift) & 255u);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    // The first lane of each digit group adds the group's size to the wave's
    // counter and hands the old value to its group. The wave's LDS atomics
    // complete in issue order, so the rounds need no wait between them.
    uint32_t old = 0;
    if (valid && rank == 0) old = atomicAdd(&cnt[d], (uint32_t)__popcll(peers));
    const int leader = valid ? __ffsll((unsigned long long)peers) - 1 : lane;
    rnk[r] = (uint32_t)__shfl(old, leader, 64) + rank;
  }
  __syncthreads();
  {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
      const uint32_t t = wcnt[w][tid];
      wcnt[w][tid] = run;
      run += t;
    }
    uint32_t all;
    loff[tid] = block_exclusive_scan<uint32_t>(run, &all);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kWaveRounds; ++r) {
    const int64_t e = wbase + r * 64 + lane;
    if (e < n) {
      const uint32_t d = (uint32_t)((key[r] >> shift) & 255u);
      const uint32_t slot = loff[d] + wcnt[wave][d] + rnk[r];
      skeys[slot] = key[r];
      if (kVals) sidx[slot] = (uint16_t)(wave * kWaveItems + r * 64 + lane);
    }
  }
  __syncthreads();
  const int tile_n = (int)min((int64_t)kSortTile, n - tile0);
#pragma unroll 4
  for (int j = tid; j < tile_n; j += kSortThreads) {
    const K k = skeys[j];
    const uint32_t d = (uint32_t)((k >> shift) & 255u);
    const uint32_t pos = goff[d] + ((uint32_t)j - loff[d]);
    keys_out[pos] = k;
    if (kVals) vals_out[pos] = vals_in[tile0 + sidx[j]];
  }
}

template <typename K>
int radix_sort_impl(K* keys, K* keys_tmp, uint64_t* vals, uint64_t* vals_tmp, int64_t n,
                    int begin_bit, int end_bit, hipStream_t s) {
  if (n <= 1 || end_bit <= begin_bit) return FH_OK;
  FH_REQUIRE(n < (int64_t)0xFFFFFFFF, "radix sort: n = %lld exceeds 2^32 - 1", (long long)n);
  const int64_t ntiles = ceil_div(n, kSortTile);
  Scratch scr(s);
  uint32_t* hist = nullptr;
  FH_TRY(scr.alloc(&hist, (size_t)(256 * ntiles)));
  K* src = keys;
  K* dst = keys_tmp;
  uint64_t* vsrc = vals;
  uint64_t* vdst = vals_tmp;
  int passes = 0;
  for (int shift = begin_bit; shift < end_bit; shift += 8, ++passes) {
    hipLaunchKernelGGL(k_rs_hist<K>, dim3((unsigned)ntiles), dim3(kSortThreads), 0, s, src, n, shift,
                       ntiles, hist);
    FH_HIP(hipGetLastError());
    FH_TRY(exclusive_scan_impl<uint32_t>(hist, hist, 256 * ntiles, (uint32_t*)nullptr, s));
    if (vals) {
      hipLaunchKernelGGL((k_rs_scatter<K, true>), dim3((unsigned)ntiles), dim3(kSortThreads), 0, s, src,
                         vsrc, n, shift, ntiles, hist, dst, vdst);
    } else {
      hipLaunchKernelGGL((k_rs_scatter<K, false>), dim3((unsigned)ntiles), dim3(kSortThreads), 0, s, src,
                         (const uint64_t*)nullptr, n, shift, ntiles, hist, dst, (uint64_t*)nullptr);
    }
    FH_HIP(hipGetLastError());
    K* t = src;
    src = dst;
    dst = t;
    uint64_t* vt = vsrc;
    vsrc = vdst;
    vdst = vt;
  }
  if (passes & 1) {
    FH_HIP(hipMemcpyAsync(keys, src, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, s));
    if (vals) FH_HIP(hipMemcpyAsync(vals, vsrc, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
  }
  return FH_OK;
}

// Sorts (in, vin) into (out, vout), using in and vin as the second buffer of
// the passes: with an odd number of 8-bit passes the last one lands in out
// and nothing is copied; with an even number the result is copied over.
template <typename K>
int radix_sort_into_impl(K* in, K* out, uint64_t* vin, uint64_t* vout, int64_t n, int begin_bit, int end_bit,
                         hipStream_t s) {
  if (n <= 0) return FH_OK;
  const int passes = end_bit > begin_bit ? (int)ceil_div(end_bit - begin_bit, 8) : 0;
  if (n == 1 || (passes & 1) == 0) {
    if (n > 1) FH_TRY(radix_sort_impl<K>(in, out, vin, vout, n, begin_bit, end_bit, s));
    FH_HIP(hipMemcpyAsync(out, in, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, s));
    if (vin) FH_HIP(hipMemcpyAsync(vout, vin, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    return FH_OK;
  }
  // An odd pass count: the passes alternate in -> out -> in -> ..., so the
  // last one writes out.
  const int64_t ntiles = ceil_div(n, kSortTile);
  FH_REQUIRE(n < (int64_t)0xFFFFFFFF, "radix sort: n = %lld exceeds 2^32 - 1", (long long)n);
  Scratch scr(s);
  uint32_t* hist = nullptr;
  FH_TRY(scr.alloc(&hist, (size_t)(256 * ntiles)));
  K* src = in;
  K* dst = out;
  uint64_t* vsrc = vin;
  uint64_t* vdst = vout;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist<K>, dim3((unsigned)ntiles), dim3(kSortThreads), 0, s, src, n, shift,
                       ntiles, hist);
    FH_HIP(hipGetLastError());
    FH_TRY(exclusive_scan_impl<uint32_t>(hist, hist, 256 * ntiles, (uint32_t*)nullptr, s));
    if (vin) {
      hipLaunchKernelGGL((k_rs_scatter<K, true>), dim3((unsigned)ntiles), dim3(kSortThreads), 0, s, src,
                         vsrc, n, shift, ntiles, hist, dst, vdst);
    } else {
      hipLaunchKernelGGL((k_rs_scatter<K, false>), dim3((unsigned)ntiles), dim3(kSortThreads), 0, s, src,
                         (const uint64_t*)nullptr, n, shift, ntiles, hist, dst, (uint64_t*)nullptr);
    }
    FH_HIP(hipGetLastError());
    K* t = src;
    src = dst;
    dst = t;
    uint64_t* vt = vsrc;
    vsrc = vdst;
    vdst = vt;
  }
  return FH_OK;
}

}  // namespace

int exclusive_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, uint64_t* total, hipStream_t s) {
  return exclusive_scan_impl<uint64_t>(in, out, n, total, s);
}
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, hipStream_t s) {
  return exclusive_scan_impl<uint32_t>(in, out, n, total, s);
}

int radix_sort_u64(uint64_t* keys, uint64_t* keys_tmp, uint64_t* vals, uint64_t* vals_tmp, int64_t n,
                   int begin_bit, int end_bit, hipStream_t s) {
  return radix_sort_impl<uint64_t>(keys, keys_tmp, vals, vals_tmp, n, begin_bit, end_bit, s);
}
int radix_sort_u32(uint32_t* keys, uint32_t* keys_tmp, uint64_t* vals, uint64_t* vals_tmp, int64_t n,
                   int begin_bit, int end_bit, hipStream_t s) {
  return radix_sort_impl<uint32_t>(keys, keys_tmp, vals, vals_tmp, n, begin_bit, end_bit, s);
}
int radix_sort_u32_into(uint32_t* in, uint32_t* out, uint64_t* vin, uint64_t* vout, int64_t n, int begin_bit,
                        int end_bit, hipStream_t s) {
  return radix_sort_into_impl<uint32_t>(in, out, vin, vout, n, begin_bit, end_bit, s);
}

int radix_sort_u64_into(uint64_t* in, uint64_t* out, int64_t n, int begin_bit, int end_bit, hipStream_t s) {
  return radix_sort_into_impl<uint64_t>(in, out, nullptr, nullptr, n, begin_bit, end_bit, s);
}

}  // namespace fh

extern "C" int fh_sort_u64(uint64_t* keys, int64_t n, int begin_bit, int end_bit, fh_stream_t stream) {
  fh::clear_error();
  FH_REQUIRE(n >= 0, "fh_sort_u64: n < 0");
  FH_REQUIRE(begin_bit >= 0 && end_bit <= 64, "fh_sort_u64: bit range [%d, %d)", begin_bit, end_bit);
  if (n <= 1) return FH_OK;
  hipStream_t s = (hipStream_t)stream;
  fh::Scratch scr(s);
  uint64_t* tmp = nullptr;
  FH_TRY(scr.alloc(&tmp, (size_t)n));
  return fh::radix_sort_u64(keys, tmp, nullptr, nullptr, n, begin_bit, end_bit, s);
}

extern "C" int fh_sort_u64_into(uint64_t* in, uint64_t* out, int64_t n, int begin_bit, int end_bit,
                                fh_stream_t stream) {
  fh::clear_error();
  FH_REQUIRE(n >= 0, "fh_sort_u64_into: n < 0");
  FH_REQUIRE(begin_bit >= 0 && end_bit <= 64, "fh_sort_u64_into: bit range [%d, %d)", begin_bit, end_bit);
  FH_REQUIRE(n == 0 || (in && out && in != out), "fh_sort_u64_into: in and out must be distinct buffers");
  return fh::radix_sort_into_impl<uint64_t>(in, out, nullptr, nullptr, n, begin_bit, end_bit,
                                            (hipStream_t)stream);
}
