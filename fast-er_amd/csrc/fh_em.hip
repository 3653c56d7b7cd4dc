// Fellegi-Sunter EM over the pattern histogram, FP64, on the device.
//
// Replaces Estimation.fit and _match_probability
// (faster/estimation.py:72-158). The histogram has at most a few thousand
// bins, so one workgroup runs every iteration without returning to the host:
// the E-step gives each thread a strided set of patterns, the M-step sums are
// wave reductions (64-lane shuffles) combined through LDS, and the stopping
// test (max |delta pi| < tol, faster/estimation.py:146) is evaluated in LDS.
#include <mutex>
#include <vector>

#include "fh_common.h"

namespace fh {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxVars = 16;
constexpr int kMaxParams = 1 + 2 * 64;  // lambda + pi_0 + pi_1, at most 64 levels in total
constexpr int kMaxPatterns = 1 << 24;

struct EmShape {
  int n_vars;
  int n_patterns;
  int levels[kMaxVars];
  int stride[kMaxVars];
  int pi_off[kMaxVars];  // offset of variable v's levels inside pi_0 / pi_1
  int n_lv;              // total levels over all variables
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// E-step for pattern p: match probability under (lam, pi0, pi1) held in LDS,
// with the conditional probabilities formed as exp(sum log pi) as the
// reference does (faster/estimation.py:80-94), in FP64.
__device__ __forceinline__ double e_step(const EmShape& sh, const double* par, int p) {
  const double lam = par[0];
  const double* pi0 = par + 1;
  const double* pi1 = par + 1 + sh.n_lv;
  double l0 = 0.0, l1 = 0.0;
  for (int v = 0; v < sh.n_vars; ++v) {
    const int d = (p / sh.stride[v]) % sh.levels[v];
    l0 += log(pi0[sh.pi_off[v] + d]);
    l1 += log(pi1[sh.pi_off[v] + d]);
  }
  const double c0 = exp(l0), c1 = exp(l1);
  return (lam * c1) / (lam * c1 + (1.0 - lam) * c0);
}

__global__ void __launch_bounds__(kThreads) k_em(EmShape sh, const int64_t* __restrict__ counts, double* params,
                                                 double tol, int max_iter, double* __restrict__ ksi_out,
                                                 double* __restrict__ wk, double* __restrict__ wn,
                                                 int* __restrict__ status) {
  __shared__ double par[kMaxParams];
  __shared__ double nxt[kMaxParams];
  __shared__ double red[kWaves][4];
  __shared__ int conv;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int npar = 1 + 2 * sh.n_lv;
  for (int i = tid; i < npar; i += kThreads) par[i] = params[i];
  if (tid == 0) conv = 0;
  __syncthreads();

  int it = 1;
  while (!conv && it <= max_iter) {
    // E-step and the weighted histogram.
    double sk = 0.0, sn = 0.0, sall = 0.0;
    for (int p = tid; p < sh.n_patterns; p += kThreads) {
      const double n = (double)counts[p];
      const double k = e_step(sh, par, p);
      wk[p] = n * k;
      wn[p] = n * (1.0 - k);
      sk += n * k;
      sn += n * (1.0 - k);
      sall += n;
    }
    sk = wave_sum(sk);
    sn = wave_sum(sn);
    sall = wave_sum(sall);
    if (lane == 0) {
      red[wave][0] = sk;
      red[wave][1] = sn;
      red[wave][2] = sall;
    }
    __syncthreads();
    // M-step: one (variable, level, class) sum per wave round.
    for (int q = wave; q < 2 * sh.n_lv; q += kWaves) {
      const int cls = q / sh.n_lv;  // 0: non-match sums, 1: match sums
      const int lv = q % sh.n_lv;
      int v = 0;
      while (v + 1 < sh.n_vars && sh.pi_off[v + 1] <= lv) ++v;
      const int l = lv - sh.pi_off[v];
      const double* src = cls ? wk : wn;
      double acc = 0.0;
      for (int p = lane; p < sh.n_patterns; p += 64)
        if ((p / sh.stride[v]) % sh.levels[v] == l) acc += src[p];
      acc = wave_sum(acc);
      if (lane == 0) nxt[1 + cls * sh.n_lv + lv] = acc;
    }
    __syncthreads();
    if (tid == 0) {
      double tk = 0.0, tn = 0.0, ta = 0.0;
      for (int w = 0; w < kWaves; ++w) {
        tk += red[w][0];
        tn += red[w][1];
        ta += red[w][2];
      }
      nxt[0] = tk / ta;
      double delta = 0.0;
      for (int i = 0; i < sh.n_lv; ++i) {
        nxt[1 + i] /= tn;
        nxt[1 + sh.n_lv + i] /= tk;
        delta = fmax(delta, fabs(nxt[1 + i] - par[1 + i]));
        delta = fmax(delta, fabs(nxt[1 + sh.n_lv + i] - par[1 + sh.n_lv + i]));
      }
      if (delta < tol) conv = 1;
    }
    __syncthreads();
    for (int i = tid; i < npar; i += kThreads) par[i] = nxt[i];
    ++it;
    __syncthreads();
  }
  for (int i = tid; i < npar; i += kThreads) params[i] = par[i];
  if (ksi_out)
    for (int p = tid; p < sh.n_patterns; p += kThreads) ksi_out[p] = e_step(sh, par, p);
  if (tid == 0) {
    status[0] = it - 1;
    status[1] = conv;
  }
}

__global__ void __launch_bounds__(kThreads) k_match_probability(EmShape sh, const double* __restrict__ params,
                                                                double* __restrict__ ksi) {
  __shared__ double par[kMaxParams];
  const int npar = 1 + 2 * sh.n_lv;
  for (int i = threadIdx.x; i < npar; i += blockDim.x) par[i] = params[i];
  __syncthreads();
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < sh.n_patterns; p += gridDim.x * blockDim.x)
    ksi[p] = e_step(sh, par, p);
}

int make_shape(int n_vars, const int32_t* levels, EmShape* sh) {
  FH_REQUIRE(n_vars >= 1 && n_vars <= kMaxVars, "EM: 1..%d variables supported", kMaxVars);
  sh->n_vars = n_vars;
  int64_t np = 1;
  int nl = 0;
  for (int v = 0; v < n_vars; ++v) {
    FH_REQUIRE(levels[v] >= 1, "EM: variable %d has %d levels", v, levels[v]);
    sh->levels[v] = levels[v];
    sh->pi_off[v] = nl;
    nl += levels[v];
    np *= levels[v];
    FH_REQUIRE(np <= kMaxPatterns, "EM: more than %d patterns", kMaxPatterns);
  }
  FH_REQUIRE(nl <= 64, "EM: more than 64 levels in total");
  int st = 1;
  for (int v = n_vars - 1; v >= 0; --v) {
    sh->stride[v] = st;
    st *= levels[v];
  }
  sh->n_patterns = (int)np;
  sh->n_lv = nl;
  return FH_OK;
}

}  // namespace
}  // namespace fh

extern "C" int fh_em_fit(int n_vars, const int32_t* levels, const int64_t* counts, double* params, double tol,
                         int max_iter, double* ksi, int* iterations, int* converged, fh_stream_t stream) {
  using namespace fh;
  clear_error();
  EmShape sh;
  FH_TRY(make_shape(n_vars, levels, &sh));
  FH_REQUIRE(max_iter >= 0, "fh_em_fit: max_iter < 0");
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  int* status = nullptr;
  double *wk = nullptr, *wn = nullptr;
  FH_TRY(scr.alloc(&status, 2));
  FH_TRY(scr.alloc(&wk, (size_t)sh.n_patterns));
  FH_TRY(scr.alloc(&wn, (size_t)sh.n_patterns));
  hipLaunchKernelGGL(k_em, dim3(1), dim3(kThreads), 0, s, sh, counts, params, tol, max_iter, ksi, wk, wn, status);
  FH_HIP(hipGetLastError());
  int hs[2] = {0, 0};
  FH_HIP(hipMemcpyAsync(hs, status, sizeof(hs), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  if (iterations) *iterations = hs[0];
  if (converged) *converged = hs[1];
  return FH_OK;
}

// ---------------------------------------------------------------------------
// EM without allocation: the per-pattern weights and the status words live in
// a file-scope device array. Its host-side name is only a stand-in for the
// symbol, so the device address is looked up once per device with
// hipGetSymbolAddress and kept here; kernels get that address.
// ---------------------------------------------------------------------------
namespace fh {
namespace {

constexpr int kNoAllocMaxPatterns = 1 << 16;
constexpr int kNoAllocMaxDevices = 64;
// wk[kNoAllocMaxPatterns], wn[kNoAllocMaxPatterns], then two int status words.
constexpr size_t kNoAllocDoubles = 2 * (size_t)kNoAllocMaxPatterns + 1;

__device__ __align__(256) double scratch[kNoAllocDoubles];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};

int scratch_for_current_device(double** out) {
  int dev = -1;
  FH_HIP(hipGetDevice(&dev));
  FH_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "fh_em_fit_noalloc: device %d beyond %d", dev,
             kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    FH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FH_REQUIRE(p != nullptr, "fh_em_fit_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return FH_OK;
}

}  // namespace
}  // namespace fh

extern "C" int fh_em_fit_noalloc(int n_vars, const int32_t* levels, const int64_t* counts, double* params,
                                 double tol, int max_iter, double* ksi, int* iterations, int* converged,
                                 fh_stream_t stream) {
  using namespace fh;
  clear_error();
  EmShape sh;
  FH_TRY(make_shape(n_vars, levels, &sh));
  FH_REQUIRE(max_iter >= 0, "fh_em_fit_noalloc: max_iter < 0");
  FH_REQUIRE(sh.n_patterns <= kNoAllocMaxPatterns,
             "fh_em_fit_noalloc: %d patterns exceed the scratch limit of %d (use fh_em_fit)", sh.n_patterns,
             kNoAllocMaxPatterns);
  double* base = nullptr;
  FH_TRY(scratch_for_current_device(&base));
  double* wk = base;
  double* wn = base + kNoAllocMaxPatterns;
  int* status = reinterpret_cast<int*>(base + 2 * (size_t)kNoAllocMaxPatterns);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_em, dim3(1), dim3(kThreads), 0, s, sh, counts, params, tol, max_iter, ksi, wk, wn, status);
  FH_HIP(hipGetLastError());
  if (iterations == nullptr && converged == nullptr) return FH_OK;
  int hs[2] = {0, 0};
  FH_HIP(hipMemcpyAsync(hs, status, sizeof(hs), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  if (iterations) *iterations = hs[0];
  if (converged) *converged = hs[1];
  return FH_OK;
}

extern "C" int fh_match_probability(int n_vars, const int32_t* levels, const double* params, double* ksi,
                                    fh_stream_t stream) {
  using namespace fh;
  clear_error();
  EmShape sh;
  FH_TRY(make_shape(n_vars, levels, &sh));
  hipStream_t s = (hipStream_t)stream;
  const int blocks = (int)ceil_div(sh.n_patterns, kThreads);
  hipLaunchKernelGGL(k_match_probability, dim3(blocks), dim3(kThreads), 0, s, sh, params, ksi);
  FH_HIP(hipGetLastError());
  return FH_OK;
}
