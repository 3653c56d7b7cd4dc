// Differentiable Fellegi-Sunter log-likelihood over a pattern histogram, FP64.
//
// No counterpart in the reference (which only runs EM, faster/estimation.py:
// 98-158). Parameters are logits theta = [a, beta (non-match class, per
// variable), alpha (match class, per variable)] with lambda = sigmoid(a),
// pi_0[k] = softmax(beta_k), pi_1[k] = softmax(alpha_k); the layout follows
// the EM parameter layout [lambda, pi_0, pi_1] of fh_em_fit. For pattern p with counts N_p:
//   M_p = lambda * prod_k pi_1[k][g_pk] + (1 - lambda) * prod_k pi_0[k][g_pk]
//   LL  = sum_p N_p log M_p,   r_p = lambda * P1_p / M_p (posterior).
// Gradient (hand-written backward):
//   dLL/da        = sum_p N_p (r_p - lambda)
//   dLL/dalpha_kl = sum_p N_p r_p       (1[g_pk = l] - pi_1kl)
//   dLL/dbeta_kl  = sum_p N_p (1 - r_p) (1[g_pk = l] - pi_0kl)
// Hessian (observed information is its negative): with
// w_p = (1; -(e_p - pi_0); e_p - pi_1) in theta order,
//   H = sum_p N_p r_p (1 - r_p) w_p w_p^T
//       - (sum_p N_p) lambda (1 - lambda) at (a, a)
//       - (sum_p N_p r_p) C_1k on each alpha_k block
//       - (sum_p N_p (1 - r_p)) C_0k on each beta_k block,
// C_k[l, l'] = pi_kl (delta_ll' - pi_kl').
#include "fh_common.h"

namespace fh {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxVars = 16;
constexpr int kMaxLv = 64;

struct LikShape {
  int n_vars;
  int n_patterns;
  int levels[kMaxVars];
  int stride[kMaxVars];
  int off[kMaxVars];
  int n_lv;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum of one value per thread (all threads get the result).
__device__ double block_sum(double v, double* scratch) {
  v = wave_sum(v);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) scratch[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kWaves; ++w) t += scratch[w];
  __syncthreads();
  return t;
}

__device__ __forceinline__ double log1pexp(double x) { return x > 0 ? x + log1p(exp(-x)) : log1p(exp(x)); }

// Per-variable log-softmax of both classes into lp1 / lp0 (shared), plus
// log lambda and log(1 - lambda).
__device__ void prepare(const LikShape& sh, const double* __restrict__ theta, double* lp1, double* lp0, double* ll) {
  if (threadIdx.x < 2 * sh.n_vars) {
    const int cls = threadIdx.x / sh.n_vars;
    const int k = threadIdx.x % sh.n_vars;
    const double* x = theta + 1 + cls * sh.n_lv + sh.off[k];
    double mx = x[0];
    for (int l = 1; l < sh.levels[k]; ++l) mx = fmax(mx, x[l]);
    double s = 0.0;
    for (int l = 0; l < sh.levels[k]; ++l) s += exp(x[l] - mx);
    const double lse = mx + log(s);
    double* out = cls ? lp1 : lp0;  // theta order: non-match, then match
    for (int l = 0; l < sh.levels[k]; ++l) out[sh.off[k] + l] = x[l] - lse;
  }
  if (threadIdx.x == 0) {
    const double a = theta[0];
    ll[0] = -log1pexp(-a);  // log sigmoid(a)
    ll[1] = -log1pexp(a);   // log(1 - sigmoid(a))
  }
  __syncthreads();
}

// log M_p and r_p of pattern p, in the log domain.
__device__ __forceinline__ void pattern_terms(const LikShape& sh, const double* lp1, const double* lp0,
                                              const double* ll, int p, double* logm, double* r) {
  double s1 = ll[0], s0 = ll[1];
  for (int k = 0; k < sh.n_vars; ++k) {
    const int g = (p / sh.stride[k]) % sh.levels[k];
    s1 += lp1[sh.off[k] + g];
    s0 += lp0[sh.off[k] + g];
  }
  const double mx = fmax(s1, s0);
  const double lm = mx + log(exp(s1 - mx) + exp(s0 - mx));
  *logm = lm;
  *r = exp(s1 - lm);
}

__global__ void __launch_bounds__(kThreads) k_fslik_fwd(LikShape sh, const int64_t* __restrict__ counts,
                                                        const double* __restrict__ theta, double* __restrict__ out,
                                                        double* __restrict__ post) {
  __shared__ double lp1[kMaxLv], lp0[kMaxLv], ll[2], scratch[kWaves];
  prepare(sh, theta, lp1, lp0, ll);
  double acc = 0.0;
  for (int p = threadIdx.x; p < sh.n_patterns; p += kThreads) {
    double lm, r;
    pattern_terms(sh, lp1, lp0, ll, p, &lm, &r);
    const double n = (double)counts[p];
    acc += n == 0.0 ? 0.0 : n * lm;
    if (post) post[p] = r;
  }
  const double tot = block_sum(acc, scratch);
  if (threadIdx.x == 0) out[0] = tot;
}

// grad_theta = g * dLL/dtheta. wr / wn: scratch of n_patterns doubles each
// (N_p r_p and N_p (1 - r_p)); the per-level sums are wave reductions in a
// fixed order, so the gradient is deterministic.
__global__ void __launch_bounds__(kThreads) k_fslik_bwd(LikShape sh, const int64_t* __restrict__ counts,
                                                        const double* __restrict__ theta,
                                                        const double* __restrict__ grad_out,
                                                        double* __restrict__ grad, double* __restrict__ wr,
                                                        double* __restrict__ wn) {
  __shared__ double lp1[kMaxLv], lp0[kMaxLv], ll[2], scratch[kWaves];
  __shared__ double s[2 * kMaxLv];
  prepare(sh, theta, lp1, lp0, ll);
  double tn = 0.0, tr = 0.0;
  for (int p = threadIdx.x; p < sh.n_patterns; p += kThreads) {
    double lm, r;
    pattern_terms(sh, lp1, lp0, ll, p, &lm, &r);
    const double n = (double)counts[p];
    tn += n;
    tr += n * r;
    wr[p] = n * r;
    wn[p] = n * (1.0 - r);
  }
  tn = block_sum(tn, scratch);
  tr = block_sum(tr, scratch);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int q = wave; q < 2 * sh.n_lv; q += kWaves) {
    const int cls = q / sh.n_lv;
    const int lv = q % sh.n_lv;
    int k = 0;
    while (k + 1 < sh.n_vars && sh.off[k + 1] <= lv) ++k;
    const int l = lv - sh.off[k];
    const double* src = cls ? wr : wn;  // cls 0: non-match sums, 1: match sums
    double acc = 0.0;
    for (int p = lane; p < sh.n_patterns; p += 64)
      if ((p / sh.stride[k]) % sh.levels[k] == l) acc += src[p];
    acc = wave_sum(acc);
    if (lane == 0) s[q] = acc;
  }
  __syncthreads();
  const double g = grad_out[0];
  const double lam = exp(ll[0]);
  if (threadIdx.x == 0) grad[0] = g * (tr - tn * lam);
  for (int i = threadIdx.x; i < sh.n_lv; i += kThreads) {
    grad[1 + i] = g * (s[i] - (tn - tr) * exp(lp0[i]));
    grad[1 + sh.n_lv + i] = g * (s[sh.n_lv + i] - tr * exp(lp1[i]));
  }
}

// Hessian (dim x dim, dim = 1 + 2 n_lv), one thread per upper-triangle
// entry; post = r_p from the forward kernel.
__global__ void __launch_bounds__(kThreads) k_fslik_hess(LikShape sh, const int64_t* __restrict__ counts,
                                                         const double* __restrict__ theta,
                                                         const double* __restrict__ post, double* __restrict__ hess) {
  __shared__ double lp1[kMaxLv], lp0[kMaxLv], ll[2];
  prepare(sh, theta, lp1, lp0, ll);
  const int dim = 1 + 2 * sh.n_lv;
  const int n_entries = dim * (dim + 1) / 2;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_entries) return;
  // entry e -> (i, j), i <= j
  int i = 0, rem = e;
  while (rem >= dim - i) {
    rem -= dim - i;
    ++i;
  }
  const int j = i + rem;
  auto coord = [&](int c, int* cls, int* k, int* l) {
    if (c == 0) {
      *cls = -1;
      return;
    }
    const int q = c - 1;
    *cls = q / sh.n_lv;  // 0: beta (non-match), 1: alpha (match)
    const int lv = q % sh.n_lv;
    int kk = 0;
    while (kk + 1 < sh.n_vars && sh.off[kk + 1] <= lv) ++kk;
    *k = kk;
    *l = lv - sh.off[kk];
  };
  int ci, ki = 0, li = 0, cj, kj = 0, lj = 0;
  coord(i, &ci, &ki, &li);
  coord(j, &cj, &kj, &lj);
  auto w_of = [&](int cls, int k, int l, int p) -> double {
    if (cls < 0) return 1.0;
    const int g = (p / sh.stride[k]) % sh.levels[k];
    const double e1 = g == l ? 1.0 : 0.0;
    return cls == 1 ? e1 - exp(lp1[sh.off[k] + l]) : -(e1 - exp(lp0[sh.off[k] + l]));
  };
  double acc = 0.0, tn = 0.0, tr = 0.0;
  for (int p = 0; p < sh.n_patterns; ++p) {
    const double n = (double)counts[p];
    if (n == 0.0) continue;
    const double r = post[p];
    tn += n;
    tr += n * r;
    acc += n * r * (1.0 - r) * w_of(ci, ki, li, p) * w_of(cj, kj, lj, p);
  }
  if (ci < 0 && cj < 0) {
    const double lam = exp(ll[0]);
    acc -= tn * lam * (1.0 - lam);
  } else if (ci >= 0 && ci == cj && ki == kj) {
    const double* lp = ci == 1 ? lp1 : lp0;
    const double pi_l = exp(lp[sh.off[ki] + li]);
    const double pi_m = exp(lp[sh.off[kj] + lj]);
    const double c = pi_l * ((li == lj ? 1.0 : 0.0) - pi_m);
    acc -= (ci == 1 ? tr : tn - tr) * c;
  }
  hess[i * dim + j] = acc;
  hess[j * dim + i] = acc;
}

int make_shape(int n_vars, const int32_t* levels, LikShape* sh) {
  FH_REQUIRE(n_vars >= 1 && n_vars <= kMaxVars, "likelihood: 1..%d variables supported", kMaxVars);
  sh->n_vars = n_vars;
  int64_t np = 1;
  int nl = 0;
  for (int v = 0; v < n_vars; ++v) {
    FH_REQUIRE(levels[v] >= 1, "likelihood: variable %d has %d levels", v, levels[v]);
    sh->levels[v] = levels[v];
    sh->off[v] = nl;
    nl += levels[v];
    np *= levels[v];
    FH_REQUIRE(np <= (1 << 24), "likelihood: too many patterns");
  }
  FH_REQUIRE(nl <= kMaxLv, "likelihood: more than %d levels in total", kMaxLv);
  FH_REQUIRE(2 * n_vars <= kThreads, "likelihood: too many variables");
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

extern "C" int fh_fs_loglik_fwd(int n_vars, const int32_t* levels, const int64_t* counts, const double* theta,
                                double* ll, double* post, fh_stream_t stream) {
  using namespace fh;
  clear_error();
  LikShape sh;
  FH_TRY(make_shape(n_vars, levels, &sh));
  hipLaunchKernelGGL(k_fslik_fwd, dim3(1), dim3(kThreads), 0, (hipStream_t)stream, sh, counts, theta, ll, post);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

extern "C" int fh_fs_loglik_bwd(int n_vars, const int32_t* levels, const int64_t* counts, const double* theta,
                                const double* grad_out, double* grad_theta, fh_stream_t stream) {
  using namespace fh;
  clear_error();
  LikShape sh;
  FH_TRY(make_shape(n_vars, levels, &sh));
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  double *wr = nullptr, *wn = nullptr;
  FH_TRY(scr.alloc(&wr, (size_t)sh.n_patterns));
  FH_TRY(scr.alloc(&wn, (size_t)sh.n_patterns));
  hipLaunchKernelGGL(k_fslik_bwd, dim3(1), dim3(kThreads), 0, s, sh, counts, theta, grad_out, grad_theta, wr, wn);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

extern "C" int fh_fs_loglik_hess(int n_vars, const int32_t* levels, const int64_t* counts, const double* theta,
                                 double* hess, fh_stream_t stream) {
  using namespace fh;
  clear_error();
  LikShape sh;
  FH_TRY(make_shape(n_vars, levels, &sh));
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(s);
  double *post = nullptr, *ll = nullptr;
  FH_TRY(scr.alloc(&post, (size_t)sh.n_patterns));
  FH_TRY(scr.alloc(&ll, 1));
  hipLaunchKernelGGL(k_fslik_fwd, dim3(1), dim3(kThreads), 0, s, sh, counts, theta, ll, post);
  FH_HIP(hipGetLastError());
  const int dim = 1 + 2 * sh.n_lv;
  const int n_entries = dim * (dim + 1) / 2;
  hipLaunchKernelGGL(k_fslik_hess, dim3((unsigned)ceil_div(n_entries, kThreads)), dim3(kThreads), 0, s, sh, counts,
                     theta, (const double*)post, hess);
  FH_HIP(hipGetLastError());
  return FH_OK;
}
