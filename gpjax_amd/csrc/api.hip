// C ABI of libgpk and the host-side orchestration of the exact-GP path.
//
// Forward (gpk_mll_forward), replacing conjugate_mll (gpjax/objectives.py:93-107):
//   1. Sigma = K + (jitter + obs_sd^2) I into L (lower NB-tiles), identity pad.
//   2. Recursive Cholesky:  potrf(A11); L21 = A21 L11^-T (recursive TRSM whose
//      leaves multiply by the float64 leaf inverses); A22 -= L21 L21^T (SYRK on
//      the matrix cores); potrf(A22).  Leaves factor NB x NB blocks in float64
//      and also emit their inverses and log-diagonal sums.  Single problems of
//      at least four 2,048-wide panels take the look-ahead form (factor.h).
//   3. Recursive triangular inverse X = L^-1:  X21 = -X22 (L21 X11).
//   4. alpha = X^T X r, then float64 iterative refinement against Sigma in
//      float64 (a stored copy in the mixed path, recomputed from the inputs in
//      the FP64 path): alpha += X^T X (r - Sigma alpha), stopped on the device
//      at 1e-12; the mixed path's a-posteriori probes ride along.
//   5. logdet = sum of leaf log-diagonals, quad = r . alpha, mll.
// Backward (gpk_mll_backward): the fused LAUUM + adjoint tile pass and a
// deterministic reduction to the gradient.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "factor.h"

namespace gpk {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

namespace {

constexpr size_t ALIGN = 256;
// Rademacher probes of the mixed path's a-posteriori inverse check.
constexpr int NPROBE = 2;
inline size_t al(size_t b) { return (b + ALIGN - 1) / ALIGN * ALIGN; }

// ---- phase profiling: HIP events recorded on the launch stream ----
struct PhaseRec {
  int phase;
  hipEvent_t a, b;
};
std::mutex g_prof_mu;
std::vector<PhaseRec> g_prof;
bool g_prof_on = false;

struct PhaseScope {
  int phase;
  hipStream_t s;
  hipEvent_t a = nullptr;
  PhaseScope(int ph, hipStream_t st) : phase(ph), s(st) {
    if (!g_prof_on) return;
    if (hipEventCreate(&a) != hipSuccess) { a = nullptr; return; }
    if (hipEventRecord(a, s) != hipSuccess) { (void)hipEventDestroy(a); a = nullptr; }
  }
  ~PhaseScope() {
    if (!a) return;
    hipEvent_t b = nullptr;
    if (hipEventCreate(&b) != hipSuccess || hipEventRecord(b, s) != hipSuccess) {
      (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
      return;
    }
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_prof.push_back({phase, a, b});
  }
};

// Arithmetic of the three product stages of one call: float64 products for
// the FP64 path; for the mixed path the defaults (6, 3, 6) or the stages that
// GPK_PREC_MIXED_ARITH(p, t, l) names.  Per call, so no process-wide state.
struct Arith3 {
  int potrf = AR_F64, trtri = AR_F64, lauum = AR_F64;
};

inline bool arith_ok(int v) {
  return v == AR_F64 || v == AR_BF16X3 || v == AR_FP16X3 || v == AR_BF16X6;
}

// Splits a `prec` argument into the storage precision (GPK_PREC_FP64 or
// GPK_PREC_MIXED) and the stage arithmetic; -1 if it is not a valid code.
int decode_prec(int prec, Arith3* ar) {
  const int base = prec & 0xFF;
  Arith3 a;
  if (base == GPK_PREC_FP64) {
    if (prec != GPK_PREC_FP64) return -1;
  } else if (base == GPK_PREC_MIXED) {
    a.potrf = AR_BF16X6;
    a.trtri = AR_BF16X3;
    a.lauum = AR_BF16X6;
    if (prec & GPK_PREC_ARITH_FLAG) {
      a.potrf = (prec >> 12) & 0xF;
      a.trtri = (prec >> 16) & 0xF;
      a.lauum = (prec >> 20) & 0xF;
      if ((prec & ~0xFFF1FF) != 0 || !arith_ok(a.potrf) || !arith_ok(a.trtri) ||
          !arith_ok(a.lauum))
        return -1;
    } else if (prec != GPK_PREC_MIXED) {
      return -1;
    }
  } else {
    return -1;
  }
  if (ar) *ar = a;
  return base;
}

struct Layout {
  int64_t np = 0;
  int64_t htop = 0;
  size_t ts = 0;
  size_t off_L = 0, off_X = 0, off_T = 0, off_leafinv = 0, off_leafld = 0, off_leafinfo = 0;
  // r, alpha, u and v each lead a block of nvec vectors (np apart): with the
  // mixed path's probes, [r z1 z2], [alpha w1 w2], [u u1 u2], [v v1 v2].
  int nvec = 1;
  size_t off_r = 0, off_alpha = 0, off_u = 0, off_v = 0, off_res = 0, off_delta = 0;
  size_t off_skip = 0;
  size_t off_dblk = 0;
  // float32 storage: Sigma's lower NB-tiles in float64, packed, and the
  // partial sums of their matvec (the refinement's Sigma products)
  bool has_spk = false;
  size_t off_spk = 0, off_spart = 0;
  size_t off_part = 0, off_gpart = 0, off_scal = 0, off_probe = 0, total = 0;
};

// spk: room for Sigma's packed float64 lower tiles (stored_sigma); by default
// in the mixed path only.
Layout make_layout(int64_t n, int64_t d, int prec, int spk = -1) {
  Layout L;
  L.has_spk = spk < 0 ? prec != GPK_PREC_FP64 : spk != 0;
  L.np = round_up(n, NB);
  L.ts = prec == GPK_PREC_FP64 ? sizeof(double) : sizeof(float);
  const int64_t np = L.np;
  L.htop = np > NB ? split_point(np) : 0;
  const int64_t nleaf = np / NB;
  const int64_t ntile = nleaf * (nleaf + 1) / 2;
  const int64_t nchunk = (np + 1023) / 1024;
  size_t o = 0;
  L.off_L = o; o += al((size_t)np * np * L.ts);
  L.off_X = o; o += al((size_t)np * np * L.ts);
  L.off_T = o; o += al((size_t)(np - L.htop) * (L.htop > 0 ? L.htop : 1) * L.ts);
  L.off_leafinv = o; o += al((size_t)np * NB * sizeof(double));
  L.off_leafld = o; o += al((size_t)nleaf * sizeof(double));
  L.off_leafinfo = o; o += al((size_t)nleaf * sizeof(int));
  L.nvec = prec == GPK_PREC_FP64 ? 1 : 1 + NPROBE;
  const size_t VB = al((size_t)L.nvec * np * sizeof(double));
  L.off_r = o; o += VB;
  L.off_alpha = o; o += VB;
  L.off_u = o; o += VB;
  L.off_v = o; o += VB;
  L.off_res = o; o += al((size_t)np * sizeof(double));
  L.off_delta = o; o += al((size_t)np * sizeof(double));
  L.off_dblk = o; o += al(prec == GPK_PREC_FP64 ? 0 : (size_t)np * NB * sizeof(double));
  L.off_spk = o; o += al(L.has_spk ? (size_t)ntile * NB * NB * sizeof(double) : 0);
  L.off_spart = o;
  o += al(L.has_spk ? (size_t)2 * L.nvec * ntile * NB * sizeof(double) : 0);
  L.off_part = o; o += al((size_t)L.nvec * nchunk * np * sizeof(double));
  L.off_skip = o; o += al(sizeof(int));
  L.off_gpart = o; o += al((size_t)ntile * (d + 2) * sizeof(double));
  L.off_scal = o; o += al(16 * sizeof(double));
  L.off_probe = o; o += al(3 * NPROBE * sizeof(double));
  L.total = o;
  return L;
}

// ---- small kernels ----
__global__ void prep_r_kernel(const double* __restrict__ y, const double* __restrict__ theta,
                              int nls, int64_t n, int64_t np, double* __restrict__ r, Batch bt) {
  if (bt.gp) {
    const int64_t p = bt.gp[blockIdx.z];
    y += bt.xoff[p];
    n = bt.xoff[p + 1] - bt.xoff[p];
    theta += p * bt.theta_stride;
  }
  r = bws(r, bt);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) r[i] = i < n ? y[i] - theta[nls + 2] : 0.0;
}

__global__ void sub_kernel(const double* __restrict__ a, const double* __restrict__ b,
                           int64_t np, double* __restrict__ out, Batch bt) {
  if (bskip(bt)) return;
  a = bws(a, bt);
  b = bws(b, bt);
  out = bws(out, bt);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) out[i] = a[i] - b[i];
}

__global__ void add_kernel(double* __restrict__ a, const double* __restrict__ b, int64_t np,
                           Batch bt) {
  if (bskip(bt)) return;
  a = bws(a, bt);
  b = bws(b, bt);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) a[i] += b[i];
}

// Refinement stop flag of each member: cleared at the start of a forward (a
// kernel, not a memset, so that the forward stays capturable).
__global__ void skip_clear_kernel(int* skip, Batch bt) {
  if (threadIdx.x == 0) *bws(skip, bt) = 0;
}

__device__ double block_sum(double v, double* red);

// Sets the member's stop flag once ||r - Sigma alpha|| <= REFINE_TOL ||r||.
constexpr double REFINE_TOL = 1e-12;
__global__ __launch_bounds__(256) void refine_check_kernel(const double* __restrict__ r,
                                                           const double* __restrict__ res,
                                                           int64_t np, int* skip, Batch bt) {
  __shared__ double red[4];
  if (bskip(bt)) return;
  r = bws(r, bt);
  res = bws(res, bt);
  double rr = 0.0, ee = 0.0;
  for (int64_t i = threadIdx.x; i < np; i += 256) {
    rr += r[i] * r[i];
    ee += res[i] * res[i];
  }
  rr = block_sum(rr, red);
  ee = block_sum(ee, red);
  if (threadIdx.x == 0 && ee <= REFINE_TOL * REFINE_TOL * rr) *bws(skip, bt) = 1;
}

__device__ double block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = blockDim.x >> 6;
  for (int q = 0; q < nw; ++q) s += red[q];
  return s;
}

__global__ __launch_bounds__(256) void finalize_fwd_kernel(
    const double* __restrict__ leafld, const int* __restrict__ leafinfo, int64_t nleaf,
    const double* __restrict__ r, const double* __restrict__ alpha,
    const double* __restrict__ res, int64_t n, int64_t np, double* __restrict__ result,
    Batch bt) {
  __shared__ double red[4];
  leafld = bws(leafld, bt);
  leafinfo = bws(leafinfo, bt);
  r = bws(r, bt);
  alpha = bws(alpha, bt);
  res = bws(res, bt);
  if (bt.gp) {
    const int64_t p = bt.gp[blockIdx.z];
    n = bt.xoff[p + 1] - bt.xoff[p];
    result += p * GPK_RESULT_LEN;
  }
  double q = 0.0, rr = 0.0, ee = 0.0;
  for (int64_t i = threadIdx.x; i < np; i += 256) {
    if (i < n) q += r[i] * alpha[i];
    rr += r[i] * r[i];
    ee += res[i] * res[i];
  }
  q = block_sum(q, red);
  rr = block_sum(rr, red);
  ee = block_sum(ee, red);
  if (threadIdx.x == 0) {
    double logdet = 0.0;
    int info = 0;
    for (int64_t b = 0; b < nleaf; ++b) {
      logdet += leafld[b];
      if (info == 0 && leafinfo[b] != 0) info = leafinfo[b];
    }
    const double log2pi = 1.8378770664093453;
    double mll = -0.5 * ((double)n * log2pi + logdet + q);
    if (info != 0) mll = __builtin_nan("");
    result[GPK_RES_MLL] = mll;
    result[GPK_RES_LOGDET] = logdet;
    result[GPK_RES_QUAD] = q;
    result[GPK_RES_RESID] = rr > 0.0 ? sqrt(ee / rr) : sqrt(ee);
    result[GPK_RES_INFO] = (double)info;
    for (int k = 5; k < GPK_RESULT_LEN; ++k) result[k] = 0.0;
  }
}

// ---- a-posteriori check of the mixed inverse ----
// With X ~ L^-1 from the mixed factorisation, E = I - Sigma X^T X measures how
// far X^T X is from Sigma^-1 (Sigma recomputed in float64 from the inputs).
// Rademacher probes z give ||E z|| (E||E z||^2 = ||E||_F^2) and z^T E z
// (E[z^T E z] = tr E, which to first order is the error of log|Sigma| when X is
// the inverse of the computed factor: E ~ dSigma Sigma^-1).
__device__ __forceinline__ uint64_t splitmix64(uint64_t v) {
  v += 0x9E3779B97F4A7C15ull;
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
  return v ^ (v >> 31);
}

__global__ void probe_fill_kernel(double* __restrict__ z, int64_t n, int64_t np, int probe,
                                  Batch bt) {
  z = bws(z, bt);
  if (bt.gp) {
    const int64_t p = bt.gp[blockIdx.z];
    n = bt.xoff[p + 1] - bt.xoff[p];
  }
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const uint64_t h = splitmix64(((uint64_t)probe << 40) ^ (uint64_t)i);
  z[i] = i < n ? ((h >> 63) ? -1.0 : 1.0) : 0.0;
}

// For probe z with w = X^T X z and v = Sigma w: stats[3 probe] = ||z - v||^2,
// stats[3 probe + 1] = z . (z - v), stats[3 probe + 2] = ||w||^2; one workgroup.
__global__ __launch_bounds__(256) void probe_stats_kernel(const double* __restrict__ z,
                                                          const double* __restrict__ w,
                                                          const double* __restrict__ v,
                                                          int64_t np, int probe,
                                                          double* __restrict__ stats,
                                                          Batch bt) {
  __shared__ double red[4];
  z = bws(z, bt);
  w = bws(w, bt);
  v = bws(v, bt);
  stats = bws(stats, bt);
  double ee = 0.0, ze = 0.0, ww = 0.0;
  for (int64_t i = threadIdx.x; i < np; i += 256) {
    const double e = z[i] - v[i];
    ee += e * e;
    ze += z[i] * e;
    ww += w[i] * w[i];
  }
  ee = block_sum(ee, red);
  ze = block_sum(ze, red);
  ww = block_sum(ww, red);
  if (threadIdx.x == 0) {
    stats[3 * probe] = ee;
    stats[3 * probe + 1] = ze;
    stats[3 * probe + 2] = ww;
  }
}

// result[GPK_RES_PROBE] = sqrt(mean ||E z||^2 / n), result[GPK_RES_TRACE] = mean z^T E z,
// result[GPK_RES_SINV] ~ ||Sigma^-1||_F, result[GPK_RES_MLL_ERR] = the MLL error
// model below, and result[GPK_RES_FLAG] = 1 when any of these holds:
//  * the estimate of ||E||_F = sqrt(mean ||E z||^2) exceeds PROBE_FLAG_FROB.
//    To first order the trace terms of the gradient carry an error
//    tr(Sigma^-1 E dSigma) <= ||E||_F ||Sigma^-1 dSigma||_F: with ||E||_F near 1
//    they are not controlled.  Calibrated on 36 problems (N = 333 .. 65,536,
//    D = 1 .. 16, obs_sd 1e-3 .. 0.3, RBF and Matern-5/2): every one whose
//    gradient was outside 1e-4 of the float64 path read ||E||_F >= 0.98; the
//    N = 65,536, D = 16 benchmark problem reads 0.40.
//  * MLL_ERR_SAFETY * MLL_ERR > 1e-6 |MLL|.  The log-determinant of the
//    float32-stored factor is off by tr(Sigma^-1 D) for its backward error D,
//    |D_ij| ~ 2^-24 max|Sigma| with signs that do not line up, growing with the
//    recursion depth; the model 2^-25 max|Sigma| ||Sigma^-1||_F sqrt(max(1,
//    n / 1024)) was fitted to 4,108 problems (the 4,096 of the many-small-GPs
//    recipe, N = 32 .. 1,024 and noise down to 1e-3, and 12 single problems up
//    to N = 8,000): where the ||E||_F test above passes, the measured |dMLL|
//    was at most 8.1 times the model (5.1 times at the N = 65,536 benchmark).
//  * the refinement of alpha stopped above a relative residual of 1e-10;
//  * the Cholesky updates or the fused Sigma^-1 pass use fp16 or three-product
//    bf16 splits (GPK_PREC_MIXED_ARITH): the tests above were calibrated with
//    six-product or float64 products there, and do not see those stages'
//    product errors -- (4, 3, 4) read ||E||_F = 0.56 and a small MLL model with
//    its gradient 1.3e-4 from the float64 path at N = 65,536.
constexpr double PROBE_FLAG_FROB = 0.75;
constexpr double MLL_ERR_SAFETY = 10.0;

__global__ void probe_finish_kernel(const double* __restrict__ stats, int64_t n,
                                    const double* __restrict__ theta, int nls, double jitter,
                                    int uncalibrated, double* __restrict__ result, Batch bt) {
  if (threadIdx.x != 0) return;
  stats = bws(stats, bt);
  if (bt.gp) {
    const int64_t p = bt.gp[blockIdx.z];
    n = bt.xoff[p + 1] - bt.xoff[p];
    result += p * GPK_RESULT_LEN;
    theta += p * bt.theta_stride;
  }
  double ee = 0.0, ze = 0.0, ww = 0.0;
  for (int p = 0; p < NPROBE; ++p) {
    ee += stats[3 * p];
    ze += stats[3 * p + 1];
    ww += stats[3 * p + 2];
  }
  const double frob = sqrt(ee / NPROBE);
  result[GPK_RES_PROBE] = frob / sqrt((double)n);
  result[GPK_RES_TRACE] = ze / NPROBE;
  // ||Sigma^-1||_F ~ sqrt(E ||Sigma^-1 z||^2) with Sigma^-1 z ~ X^T X z
  const double sinv = sqrt(ww / NPROBE);
  const double smax = theta[nls] + jitter + theta[nls + 1] * theta[nls + 1];
  const double mll_err = 0x1p-25 * smax * sinv * sqrt(fmax(1.0, (double)n / 1024.0));
  result[GPK_RES_SINV] = sinv;
  result[GPK_RES_MLL_ERR] = mll_err;
  const bool bad = uncalibrated || frob > PROBE_FLAG_FROB || !(frob == frob) ||
                   !(MLL_ERR_SAFETY * mll_err <= 1e-6 * fabs(result[GPK_RES_MLL])) ||
                   !(result[GPK_RES_RESID] <= 1e-10);
  result[GPK_RES_FLAG] = bad ? 1.0 : 0.0;
}

// One workgroup per gradient slot: sum the tile partials in a fixed order.
__global__ __launch_bounds__(256) void finalize_bwd_kernel(
    const double* __restrict__ gpart, int64_t ntile, int nls, const double* __restrict__ theta,
    const double* __restrict__ alpha, int64_t n, double gbar, double* __restrict__ grad,
    Batch bt) {
  __shared__ double red[4];
  gpart = bws(gpart, bt);
  alpha = bws(alpha, bt);
  if (bt.gp) {
    const int64_t p = bt.gp[blockIdx.z];
    n = bt.xoff[p + 1] - bt.xoff[p];
    theta += p * bt.theta_stride;
    grad += p * (nls + 3);
  }
  const int q = blockIdx.x;
  const int nv = nls + 2;
  double s = 0.0;
  if (q < nv) {
    for (int64_t t = threadIdx.x; t < ntile; t += 256) s += gpart[t * nv + q];
  } else {
    for (int64_t i = threadIdx.x; i < n; i += 256) s += alpha[i];
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) {
    const double var = theta[nls], sd = theta[nls + 1];
    double g;
    if (q < nls) g = 0.5 * var * s / theta[q];
    else if (q == nls) g = 0.5 * s;
    else if (q == nls + 1) g = sd * s;
    else g = s;
    grad[q] = gbar * g;
  }
}

inline unsigned nblk(int64_t np) { return (unsigned)((np + 255) / 256); }

// Power-of-two fp16 operand scales [sL, sL, sX, sT] (AR_FP16X3): each maps its
// bound b to at most 2^14.  |L_ij| <= sqrt(Sigma_ii) = sqrt(var + sd^2 + jitter)
// (also bounds every trailing Schur complement); |X_ij| <= ||Sigma^-1||^(1/2) <=
// 1 / sqrt(sd^2 + jitter); |(L21 X11)_ij| <= bL bX.  Bounds are at least 1 (the
// identity pad) and at most 2^14.
__global__ void fp16_scale_kernel(const double* __restrict__ theta, int nls, double jitter,
                                  float* __restrict__ fs) {
  if (threadIdx.x != 0) return;
  const double var = theta[nls], sd = theta[nls + 1];
  auto clampb = [](double b) { return fmin(fmax(b, 1.0), 16384.0); };
  const double bl = clampb(sqrt(var + sd * sd + jitter));
  const double bx = clampb(1.0 / sqrt(sd * sd + jitter));
  const double bt = clampb(bl * bx);
  auto sc = [](double b) { return (float)ldexp(1.0, (int)floor(log2(16384.0 / b))); };
  fs[0] = sc(bl);
  fs[1] = sc(bl);
  fs[2] = sc(bx);
  fs[3] = sc(bt);
}

int check_common(int kind, int prec, const void* x, int64_t n, int64_t d, const void* theta,
                 int nls) {
  if (kind < GPK_KERNEL_RBF || kind > GPK_KERNEL_MATERN52) {
    set_error("unknown kernel kind %d", kind);
    return GPK_ERR_BAD_ARG;
  }
  if (decode_prec(prec, nullptr) < 0) {
    set_error("unknown precision code %d", prec);
    return GPK_ERR_BAD_ARG;
  }
  if (!x || !theta) {
    set_error("null input pointer");
    return GPK_ERR_BAD_ARG;
  }
  if (n < 1 || d < 1 || d > MAX_D) {
    set_error("need n >= 1 and 1 <= d <= %d (got n=%lld, d=%lld)", MAX_D, (long long)n,
              (long long)d);
    return GPK_ERR_BAD_ARG;
  }
  if (nls != 1 && nls != d) {
    set_error("nls must be 1 or d (got %d, d=%lld)", nls, (long long)d);
    return GPK_ERR_BAD_ARG;
  }
  return GPK_OK;
}

// Workspace size and alignment, as every entry point that takes one checks them.
int check_ws(const char* fn, const void* ws, size_t have, size_t need) {
  if (have < need) {
    set_error("%s: workspace %zu bytes < required %zu", fn, have, need);
    return GPK_ERR_WORKSPACE;
  }
  if (reinterpret_cast<uintptr_t>(ws) % ALIGN) {
    set_error("%s: workspace must be %zu-byte aligned", fn, ALIGN);
    return GPK_ERR_BAD_ARG;
  }
  return GPK_OK;
}

// Launch variants of one exact-GP evaluation, both exact-arithmetic choices:
// stored_sigma: the refinement's Sigma products read the packed float64 copy
//   that the Gram kernel writes (when the layout has room for it) instead of
//   evaluating the kernel again;
// adj_transposed: the fused Sigma^-1 pass reads U = X^T along k (one transpose
//   first) instead of X down its columns.
// The single-problem path runs the defaults; the batched path takes them per
// size group from the autotuner's table.
struct Variant {
  bool stored_sigma = true;
  bool adj_transposed = true;
};

template <typename TS>
int mll_forward_t(int kind, const Arith3& ar, const double* x, const double* y, int64_t n, int d,
                  const double* theta, int nls, double jitter, int iters, void* ws,
                  const Layout& Lo, double* result, hipStream_t s,
                  const Batch& bt = Batch(), const Variant& vr = Variant()) {
  const int64_t np = Lo.np;
  const dim3 gv(nblk(np), 1, bt.count);
  TS* L = at<TS>(ws, Lo.off_L);
  TS* X = at<TS>(ws, Lo.off_X);
  double* r = at<double>(ws, Lo.off_r);
  double* alpha = at<double>(ws, Lo.off_alpha);
  double* u = at<double>(ws, Lo.off_u);
  double* v = at<double>(ws, Lo.off_v);
  double* res = at<double>(ws, Lo.off_res);
  double* delta = at<double>(ws, Lo.off_delta);
  double* part = at<double>(ws, Lo.off_part);
  const bool f64 = std::is_same<TS, double>::value;
  Factor<TS> F{L,
               X,
               np,
               at<TS>(ws, Lo.off_T),
               at<double>(ws, Lo.off_leafinv),
               at<double>(ws, Lo.off_leafld),
               at<int>(ws, Lo.off_leafinfo),
               f64 ? (int)AR_F64 : ar.potrf,
               f64 ? (int)AR_F64 : ar.trtri,
               s};
  // float32 storage: the diagonal NB-blocks live in float64 on the side, and
  // so does all of Sigma's lower triangle (packed tiles), for the float64
  // products of the refinement
  double* dblk = f64 ? nullptr : at<double>(ws, Lo.off_dblk);
  double* spk = (Lo.has_spk && vr.stored_sigma) ? at<double>(ws, Lo.off_spk) : nullptr;
  // v = Sigma a (nvv vectors): read from the float64 copy when there is one,
  // otherwise recomputed from the inputs in float64
  auto sigma_mv = [&](const Batch& b, int nvv) -> int {
    if (spk)
      return launch_packed_symv(spk, np, alpha, v, at<double>(ws, Lo.off_spart), s, b, nvv);
    return launch_sigma_matvec(kind, x, n, np, d, theta, nls, jitter, alpha, v, part, s, b, nvv);
  };
  F.dblk = dblk;
  F.bt = bt;
  if (!f64) {
    float* fs = at<float>(ws, Lo.off_scal);
    hipLaunchKernelGGL(fp16_scale_kernel, dim3(1), dim3(64), 0, s, theta, nls, jitter, fs);
    GPK_LAUNCH_CHECK();
    F.fscale = fs;
  }

  hipLaunchKernelGGL(prep_r_kernel, gv, dim3(256), 0, s, y, theta, nls, n, np, r, bt);
  GPK_LAUNCH_CHECK();
  const bool probes = !f64;
  if (probes) {
    // Rademacher probes z_p of the a-posteriori check, behind r
    for (int p = 0; p < NPROBE; ++p) {
      hipLaunchKernelGGL(probe_fill_kernel, gv, dim3(256), 0, s, r + (p + 1) * np, n, np, p, bt);
      GPK_LAUNCH_CHECK();
    }
  }
  int* skip = at<int>(ws, Lo.off_skip);
  hipLaunchKernelGGL(skip_clear_kernel, dim3(1, 1, bt.count), dim3(64), 0, s, skip, bt);
  GPK_LAUNCH_CHECK();
  Batch bts = bt;  // the refinement's launches, which stop once it has converged
  bts.skip = skip;
  {
    PhaseScope ps(GPK_PHASE_GRAM, s);
    GPK_TRY(launch_kernel_matrix<TS>(kind, x, n, x, n, d, theta, nls, jitter, 2, L, np, np, np,
                                     true, s, dblk, bt, spk));
  }
  {
    PhaseScope ps(GPK_PHASE_POTRF, s);
    const int64_t la = bt.gp ? 0 : lookahead_width(np, std::is_same<TS, float>::value);
    if (la > 0)
      GPK_TRY(F.potrf_lookahead(np, la));
    else
      GPK_TRY(F.potrf(0, np));
  }
  {
    PhaseScope ps(GPK_PHASE_TRTRI, s);
    GPK_TRY(F.trtri(0, np));
  }
  PhaseScope ps_solve(GPK_PHASE_SOLVE, s);
  // alpha0 = X^T X r; the probes ride along: w_p = X^T X z_p
  const int nv = probes ? 1 + NPROBE : 1;
  GPK_TRY(launch_trmv<TS>(X, np, np, r, u, false, part, s, bt, nv));
  GPK_TRY(launch_trmv<TS>(X, np, np, u, alpha, true, part, s, bt, nv));
  // Refinement: alpha += X^T X (r - Sigma alpha) until ||r - Sigma alpha|| <=
  // 1e-12 ||r|| or `iters` steps; the first Sigma product also forms Sigma w_p.
  for (int it = 0; it < iters; ++it) {
    GPK_TRY(sigma_mv(it == 0 ? bt : bts, it == 0 ? nv : 1));
    hipLaunchKernelGGL(sub_kernel, gv, dim3(256), 0, s, r, v, np, res, bts);
    GPK_LAUNCH_CHECK();
    if (it > 0) {
      hipLaunchKernelGGL(refine_check_kernel, dim3(1, 1, bt.count), dim3(256), 0, s, r, res, np,
                         skip, bts);
      GPK_LAUNCH_CHECK();
    }
    GPK_TRY(launch_trmv<TS>(X, np, np, res, u, false, part, s, bts));
    GPK_TRY(launch_trmv<TS>(X, np, np, u, delta, true, part, s, bts));
    hipLaunchKernelGGL(add_kernel, gv, dim3(256), 0, s, alpha, delta, np, bts);
    GPK_LAUNCH_CHECK();
  }
  if (probes && iters == 0) {
    GPK_TRY(sigma_mv(bt, nv));
  }
  if (probes) {
    // a-posteriori check of X ~ L^-1: e_p = z_p - Sigma X^T X z_p
    double* stats = at<double>(ws, Lo.off_probe);
    for (int p = 0; p < NPROBE; ++p) {
      hipLaunchKernelGGL(probe_stats_kernel, dim3(1, 1, bt.count), dim3(256), 0, s,
                         r + (p + 1) * np, alpha + (p + 1) * np, v + (p + 1) * np, np, p,
                         stats, bt);
      GPK_LAUNCH_CHECK();
    }
  }
  // The final residual r - Sigma alpha.  A refinement that stopped early left it
  // in `res` already (alpha has not changed since), so these launches return
  // at once then; otherwise they measure the last correction.
  GPK_TRY(sigma_mv(bts, 1));
  hipLaunchKernelGGL(sub_kernel, gv, dim3(256), 0, s, r, v, np, res, bts);
  GPK_LAUNCH_CHECK();
  hipLaunchKernelGGL(finalize_fwd_kernel, dim3(1, 1, bt.count), dim3(256), 0, s,
                     at<double>(ws, Lo.off_leafld), at<int>(ws, Lo.off_leafinfo), np / NB, r,
                     alpha, res, n, np, result, bt);
  GPK_LAUNCH_CHECK();
  if (probes) {
    const bool calibrated = (ar.potrf == AR_BF16X6 || ar.potrf == AR_F64) &&
                            (ar.lauum == AR_BF16X6 || ar.lauum == AR_F64);
    hipLaunchKernelGGL(probe_finish_kernel, dim3(1, 1, bt.count), dim3(64), 0, s,
                       at<double>(ws, Lo.off_probe), n, theta, nls, jitter, calibrated ? 0 : 1,
                       result, bt);
    GPK_LAUNCH_CHECK();
  }
  return GPK_OK;
}

template <typename TS>
int mll_backward_t(int kind, const Arith3& ar, const double* x, int64_t n, int d,
                   const double* theta, int nls, double gbar, void* ws, const Layout& Lo,
                   double* grad, hipStream_t s, const Batch& bt = Batch(),
                   const Variant& vr = Variant()) {
  const int64_t np = Lo.np;
  const bool f64 = std::is_same<TS, double>::value;
  double* gpart = at<double>(ws, Lo.off_gpart);
  const float* fs = f64 ? nullptr : at<float>(ws, Lo.off_scal) + 2;
  if (vr.adj_transposed) {
    // The factor L is no longer needed: its buffer takes U = X^T, whose panels
    // the fused pass reads along k (16-byte loads) instead of down columns.
    // (The transpose, HBM-bound, is outside the adjoint phase's timing.)
    TS* U = at<TS>(ws, Lo.off_L);
    GPK_TRY(launch_lower_to_upper<TS>(at<TS>(ws, Lo.off_X), U, np, np, s, bt));
    PhaseScope ps(GPK_PHASE_ADJOINT, s);
    GPK_TRY(launch_adjoint<TS>(kind, f64 ? (int)AR_F64 : ar.lauum, U, np, np, x, n, d, theta,
                               nls, at<double>(ws, Lo.off_alpha), gpart, s, bt, nullptr, false,
                               fs, true));
  } else {
    PhaseScope ps(GPK_PHASE_ADJOINT, s);
    GPK_TRY(launch_adjoint<TS>(kind, f64 ? (int)AR_F64 : ar.lauum, at<TS>(ws, Lo.off_X), np,
                               np, x, n, d, theta, nls, at<double>(ws, Lo.off_alpha), gpart, s,
                               bt, nullptr, false, fs));
  }
  const int64_t nleaf = np / NB;
  hipLaunchKernelGGL(finalize_bwd_kernel, dim3((unsigned)(nls + 3), 1, bt.count), dim3(256), 0, s,
                     gpart, nleaf * (nleaf + 1) / 2, nls, theta, at<double>(ws, Lo.off_alpha), n,
                     gbar, grad, bt);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

// ---- predict: ConjugatePosterior.predict (gpjax/gps.py:494-526) ----
// tc: test points per pass.  The full covariance needs all of them at once (its
// T x T block is the output anyway); the diagonal mode streams them in passes
// whose two np x tc buffers (Kxt and A) hold at most PRED_ELEMS doubles each,
// so T is bounded only by the caller's mean and variance vectors.
constexpr int64_t PRED_ELEMS = (int64_t)1 << 26;  // 512 MB per buffer

struct PredLayout {
  int64_t tp = 0, tc = 0;
  size_t off_kxt = 0, off_a = 0, off_ktt = 0, off_part = 0, off_vec = 0, off_res = 0, total = 0;
};

PredLayout make_pred_layout(const Layout& Lo, int64_t t, int full_cov) {
  PredLayout P;
  P.tp = round_up(t, NB);
  const int64_t np = Lo.np, tp = P.tp;
  int64_t tc = tp;
  if (!full_cov) {
    tc = (PRED_ELEMS / np) / NB * NB;
    if (tc < NB) tc = NB;
    if (tc > tp) tc = tp;
  }
  P.tc = tc;
  size_t o = Lo.total;
  P.off_kxt = o; o += al((size_t)np * tc * sizeof(double));
  P.off_a = o; o += al((size_t)np * tc * sizeof(double));
  P.off_ktt = o; o += al((size_t)(full_cov ? tp * tp : 1) * sizeof(double));
  P.off_part = o; o += al((size_t)((np + 1023) / 1024) * tc * sizeof(double));
  P.off_vec = o; o += al((size_t)tc * sizeof(double));
  P.off_res = o; o += al((size_t)GPK_RESULT_LEN * sizeof(double));
  P.total = o;
  return P;
}

__global__ void pred_mean_kernel(const double* __restrict__ v, const double* __restrict__ theta,
                                 int nls, int64_t t, double* __restrict__ mean) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < t) mean[j] = theta[nls + 2] + v[j];
}

__global__ void pred_var_kernel(const double* __restrict__ colsq,
                                const double* __restrict__ theta, int nls, double jitter_prior,
                                int64_t t, double* __restrict__ var_out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < t) var_out[j] = (theta[nls] + jitter_prior) - colsq[j];
}

template <typename TS>
int predict_t(int kind, const Arith3& ar, const double* x, const double* y, int64_t n, int d,
              const double* xt, int64_t t, const double* theta, int nls, double jitter_post,
              double jitter_prior, int full_cov, int iters, void* ws, const Layout& Lo,
              const PredLayout& P, double* mean, double* cov, double* result_out,
              hipStream_t s) {
  const int64_t np = Lo.np, tp = P.tp;
  double* result = result_out ? result_out : at<double>(ws, P.off_res);
  // L, X = L^-1 and alpha = Sigma^-1 (y - m), Sigma with the posterior's jitter (gps.py:506-511)
  GPK_TRY(mll_forward_t<TS>(kind, ar, x, y, n, d, theta, nls, jitter_post, iters, ws, Lo,
                            result, s));
  double* kxt = at<double>(ws, P.off_kxt);
  double* A = at<double>(ws, P.off_a);
  double* part = at<double>(ws, P.off_part);
  double* vec = at<double>(ws, P.off_vec);
  const double* alpha = at<double>(ws, Lo.off_alpha);
  const TS* X = at<TS>(ws, Lo.off_X);
  if (!full_cov) {
    // Diagonal only, the test points in passes of tc:
    //   mean_j = m(t) + Kxt^T alpha,  var_j = k(0) + jitter_prior - sum_i A_ij^2.
    // With float32 storage (mixed) A = X Kxt runs as six-product bf16 splits of
    // X and a float32 copy of Kxt, written in float64, like the Cholesky updates;
    // the mean takes the float64 Kxt first, and the copy then reuses its buffer.
    const bool f64 = std::is_same<TS, double>::value;
    float* kxtf = reinterpret_cast<float*>(kxt);
    for (int64_t t0 = 0; t0 < t; t0 += P.tc) {
      const int64_t tn = (t - t0 < P.tc) ? t - t0 : P.tc;
      const int64_t tnp = round_up(tn, NB);
      GPK_TRY(launch_cross<double>(kind, x, n, xt + t0 * d, tn, d, theta, nls, false, kxt, tnp,
                                   np, tnp, s));
      GPK_TRY(launch_colreduce(kxt, tnp, np, tnp, alpha, vec, part, s));
      hipLaunchKernelGGL(pred_mean_kernel, dim3(nblk(tn)), dim3(256), 0, s, vec, theta, nls, tn,
                         mean + t0);
      GPK_LAUNCH_CHECK();
      if (f64) {
        GPK_TRY(launch_gemm<TS, double, double>(AR_F64, false, true, X, np, kxt, tnp, A, tnp, np,
                                                tnp, np, 1.0, false, false, KR_LE_M, s));
        GPK_TRY(launch_colreduce(A, tnp, np, tnp, nullptr, vec, part, s));
      } else {
        GPK_TRY(launch_cross<float>(kind, x, n, xt + t0 * d, tn, d, theta, nls, false, kxtf,
                                    tnp, np, tnp, s));
        GPK_TRY(launch_gemm<float, float, double>(AR_BF16X6, false, true,
                                                  reinterpret_cast<const float*>(X), np, kxtf,
                                                  tnp, A, tnp, np, tnp, np, 1.0, false, false,
                                                  KR_LE_M, s));
        GPK_TRY(launch_colreduce(A, tnp, np, tnp, nullptr, vec, part, s));
      }
      hipLaunchKernelGGL(pred_var_kernel, dim3(nblk(tn)), dim3(256), 0, s, vec, theta, nls,
                         jitter_prior, tn, cov + t0);
      GPK_LAUNCH_CHECK();
    }
    return GPK_OK;
  }
  // Kxt (gps.py:515), zero outside n x t
  GPK_TRY(launch_cross<double>(kind, x, n, xt, t, d, theta, nls, false, kxt, tp, np, tp, s));
  // mean = m(t) + Kxt^T alpha  (= m(t) + (L^-1 Kxt)^T L^-1 (y - m), gps.py:517-520)
  GPK_TRY(launch_colreduce(kxt, tp, np, tp, alpha, vec, part, s));
  hipLaunchKernelGGL(pred_mean_kernel, dim3(nblk(t)), dim3(256), 0, s, vec, theta, nls, t, mean);
  GPK_LAUNCH_CHECK();
  // A = L^-1 Kxt  (X lower: k <= row tile)
  GPK_TRY(launch_gemm<TS, double, double>(AR_F64, false, true, X, np, kxt, tp, A, tp, np, tp, np,
                                          1.0, false, false, KR_LE_M, s));
  {
    // cov = Ktt + jitter_prior I - A^T A  (gps.py:514,522-523)
    double* ktt = at<double>(ws, P.off_ktt);
    GPK_TRY(launch_kernel_matrix<double>(kind, xt, t, xt, t, d, theta, nls, jitter_prior,
                                         1 | KM_ZERO_PAD, ktt, tp, tp, tp, false, s));
    GPK_TRY(launch_gemm<double, double, double>(AR_F64, true, true, A, tp, A, tp, ktt, tp, tp, tp,
                                                np, -1.0, true, false, KR_FULL, s));
    GPK_HIP_TRY(hipMemcpy2DAsync(cov, (size_t)t * sizeof(double), ktt, (size_t)tp * sizeof(double),
                                 (size_t)t * sizeof(double), (size_t)t, hipMemcpyDeviceToDevice, s));
  }
  return GPK_OK;
}

// ---- leave-one-out predictive probability: conjugate_loocv (objectives.py:156-178) ----
// With alpha = Sigma^-1 r and d_i = [Sigma^-1]_ii (column sums of squares of
// X = L^-1; the reference inverts Sigma densely, objectives.py:171):
//   LOOCV = sum_{i<n} -1/2 log 2pi + 1/2 log d_i - 1/2 alpha_i^2 / d_i.
// Its derivative is 1/2 tr(W dSigma) + sum(w) dc with v = alpha / d,
// w = Sigma^-1 v, c_i^2 = 1/d_i + alpha_i^2/d_i^2 and
//   W = alpha w^T + w alpha^T - Y^T Y,   Y = diag(c) Sigma^-1,
// which the adjoint pass reduces against dSigma/dtheta with Y dense.
// Sigma^-1 = X^T X is formed (lower tiles, mirrored and row-scaled in place)
// in the buffer that held L, which the backward no longer needs.
struct LooLayout {
  size_t off_d = 0, off_v = 0, off_w = 0, off_c = 0, off_tmp = 0, off_mres = 0, total = 0;
};

LooLayout make_loo_layout(const Layout& Lo) {
  LooLayout P;
  const size_t V = al((size_t)Lo.np * sizeof(double));
  size_t o = Lo.total;
  P.off_d = o; o += V;
  P.off_v = o; o += V;
  P.off_w = o; o += V;
  P.off_c = o; o += V;
  P.off_tmp = o; o += V;
  P.off_mres = o; o += al((size_t)GPK_RESULT_LEN * sizeof(double));
  P.total = o;
  return P;
}

__global__ __launch_bounds__(256) void loo_fwd_kernel(const double* __restrict__ dg,
                                                      const double* __restrict__ alpha,
                                                      const double* __restrict__ mres, int64_t n,
                                                      int64_t np, double* __restrict__ v,
                                                      double* __restrict__ c,
                                                      double* __restrict__ result) {
  __shared__ double red[4];
  double sl = 0.0, sq = 0.0;
  for (int64_t i = threadIdx.x; i < np; i += 256) {
    if (i < n) {
      const double di = dg[i], a = alpha[i];
      const double q = a * a / di;
      sl += log(di);
      sq += q;
      v[i] = a / di;
      c[i] = sqrt((1.0 + q) / di);
    } else {
      v[i] = 0.0;
      c[i] = 0.0;
    }
  }
  sl = block_sum(sl, red);
  sq = block_sum(sq, red);
  if (threadIdx.x == 0) {
    const double log2pi = 1.8378770664093453;
    const double info = mres[GPK_RES_INFO];
    const double val = -0.5 * ((double)n * log2pi - sl + sq);
    result[GPK_RES_MLL] = info != 0.0 ? __builtin_nan("") : val;
    result[GPK_RES_LOGDET] = sl;
    result[GPK_RES_QUAD] = sq;
    result[GPK_RES_RESID] = mres[GPK_RES_RESID];
    result[GPK_RES_INFO] = info;
    for (int k = 5; k < GPK_RESULT_LEN; ++k) result[k] = mres[k];
  }
}

// Y <- diag(c) S with S symmetric held in its lower 64-tiles (upper tiles are
// overwritten).  One workgroup per lower tile (I, J): it reads only tile (I, J)
// and writes tiles (I, J) and (J, I), so the update in place has no race.
constexpr int MT = 64;
// Lower-triangle tile t -> (I, J), I >= J, row-major over I.
__device__ __forceinline__ void tri_tile(int64_t t, int64_t& I, int64_t& J) {
  int64_t a = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (a * (a + 1) / 2 > t) --a;
  while ((a + 1) * (a + 2) / 2 <= t) ++a;
  I = a;
  J = t - a * (a + 1) / 2;
}

template <typename TS>
__global__ __launch_bounds__(256) void loo_scale_mirror_kernel(TS* __restrict__ Y, int64_t ld,
                                                               const double* __restrict__ c) {
  __shared__ TS tl[MT][MT + 1];
  int64_t I, J;
  tri_tile(blockIdx.x, I, J);
  const int64_t r0 = I * MT, c0 = J * MT;
  for (int e = threadIdx.x; e < MT * MT; e += 256) {
    const int r = e / MT, q = e % MT;
    tl[r][q] = Y[(r0 + r) * ld + c0 + q];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < MT * MT; e += 256) {
    const int r = e / MT, q = e % MT;
    if (I == J) {
      const TS sv = r >= q ? tl[r][q] : tl[q][r];
      Y[(r0 + r) * ld + c0 + q] = (TS)(c[r0 + r] * (double)sv);
    } else {
      Y[(r0 + r) * ld + c0 + q] = (TS)(c[r0 + r] * (double)tl[r][q]);
      Y[(c0 + r) * ld + r0 + q] = (TS)(c[c0 + r] * (double)tl[q][r]);
    }
  }
}

template <typename TS>
int loocv_forward_t(int kind, const Arith3& ar, const double* x, const double* y, int64_t n, int d,
                    const double* theta, int nls, double jitter, int iters, void* ws,
                    const Layout& Lo, const LooLayout& P, double* result, hipStream_t s) {
  const int64_t np = Lo.np;
  GPK_TRY(mll_forward_t<TS>(kind, ar, x, y, n, d, theta, nls, jitter, iters, ws, Lo,
                            at<double>(ws, P.off_mres), s));
  double* dg = at<double>(ws, P.off_d);
  GPK_TRY(launch_colreduce<TS>(at<TS>(ws, Lo.off_X), np, np, np, nullptr, dg,
                               at<double>(ws, Lo.off_part), s, true));
  hipLaunchKernelGGL(loo_fwd_kernel, dim3(1), dim3(256), 0, s, dg, at<double>(ws, Lo.off_alpha),
                     at<double>(ws, P.off_mres), n, np, at<double>(ws, P.off_v),
                     at<double>(ws, P.off_c), result);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

template <typename TS>
int loocv_backward_t(int kind, const Arith3& ar3, const double* x, int64_t n, int d,
                     const double* theta, int nls, double gbar, void* ws, const Layout& Lo,
                     const LooLayout& P, double* grad, hipStream_t s) {
  const int64_t np = Lo.np;
  const bool f64 = std::is_same<TS, double>::value;
  // Y = diag(c) Sigma^-1 has no a-priori bound for an fp16 scale: BF16x6 there.
  const int ar = f64 ? (int)AR_F64 : (ar3.lauum == AR_FP16X3 ? (int)AR_BF16X6 : ar3.lauum);
  const TS* X = at<TS>(ws, Lo.off_X);
  TS* Y = at<TS>(ws, Lo.off_L);
  double* part = at<double>(ws, Lo.off_part);
  double* w = at<double>(ws, P.off_w);
  // w = Sigma^-1 v = X^T (X v)
  GPK_TRY(launch_trmv<TS>(X, np, np, at<double>(ws, P.off_v), at<double>(ws, P.off_tmp), false,
                          part, s));
  GPK_TRY(launch_trmv<TS>(X, np, np, at<double>(ws, P.off_tmp), w, true, part, s));
  // Y = diag(c) X^T X
  GPK_TRY(launch_gemm<TS, TS, TS>(ar, true, true, X, np, X, np, Y, np, np, np, np, 1.0, false,
                                  true, KR_GE_M, s));
  const int64_t mt = np / MT;
  hipLaunchKernelGGL(loo_scale_mirror_kernel<TS>, dim3((unsigned)(mt * (mt + 1) / 2)), dim3(256),
                     0, s, Y, np, at<double>(ws, P.off_c));
  GPK_LAUNCH_CHECK();
  double* gpart = at<double>(ws, Lo.off_gpart);
  GPK_TRY(launch_adjoint<TS>(kind, ar, Y, np, np, x, n, d, theta, nls,
                             at<double>(ws, Lo.off_alpha), gpart, s, Batch(), w, true));
  const int64_t nleaf = np / NB;
  hipLaunchKernelGGL(finalize_bwd_kernel, dim3((unsigned)(nls + 3)), dim3(256), 0, s, gpart,
                     nleaf * (nleaf + 1) / 2, nls, theta, w, n, gbar, grad, Batch());
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

// ---- batched exact GPs ----
// Problems are grouped by padded size np = round_up(n, 128) (stable in the
// problem index); each group runs the exact-GP forward and backward once, with
// one member per blockIdx.z.  Workspace: the member -> problem map and the row
// offsets (uploaded by gpk_batched_plan), then per group `count` member
// workspaces of one Layout each.
struct BatchGroup {
  int64_t np = 0, first = 0, count = 0;
  size_t off = 0, stride = 0;
  Layout lo;
};
struct BatchPlan {
  std::vector<int> order;  // member slot -> problem index
  std::vector<BatchGroup> groups;
  size_t off_gp = 0, off_xoff = 0, total = 0;
};

int check_offsets(const int64_t* off, int64_t count, int64_t d) {
  if (!off || count < 1) {
    set_error("batched: need offsets and count >= 1");
    return GPK_ERR_BAD_ARG;
  }
  if (d < 1 || d > MAX_D) {
    set_error("batched: need 1 <= d <= %d (got %lld)", MAX_D, (long long)d);
    return GPK_ERR_BAD_ARG;
  }
  if (off[0] != 0) {
    set_error("batched: offsets[0] must be 0");
    return GPK_ERR_BAD_ARG;
  }
  for (int64_t p = 0; p < count; ++p) {
    if (off[p + 1] - off[p] < 1) {
      set_error("batched: problem %lld has no rows", (long long)p);
      return GPK_ERR_BAD_ARG;
    }
  }
  return GPK_OK;
}

// The batched path's `prec`: as gpk_mll_forward's, except the fp16 split (its
// operand scales come from one problem's hyperparameters).  -1 if refused.
int batched_prec(int prec, Arith3* ar) {
  Arith3 a;
  const int base = decode_prec(prec, &a);
  if (base < 0 || a.potrf == AR_FP16X3 || a.trtri == AR_FP16X3 || a.lauum == AR_FP16X3) {
    set_error("batched: prec 0x%x refused (FP64, or mixed with bf16 splits: the fp16 split's "
              "operand scales are one problem's)", prec);
    return -1;
  }
  if (ar) *ar = a;
  return base;
}

BatchPlan make_batch_plan(const int64_t* off, int64_t count, int64_t d, int base) {
  BatchPlan P;
  P.order.resize((size_t)count);
  for (int64_t p = 0; p < count; ++p) P.order[(size_t)p] = (int)p;
  auto padded = [&](int p) { return round_up(off[p + 1] - off[p], NB); };
  std::stable_sort(P.order.begin(), P.order.end(),
                   [&](int a, int b) { return padded(a) < padded(b); });
  size_t o = 0;
  P.off_gp = o; o += al((size_t)count * sizeof(int));
  P.off_xoff = o; o += al((size_t)(count + 1) * sizeof(int64_t));
  for (int64_t i = 0; i < count;) {
    BatchGroup g;
    g.np = padded(P.order[(size_t)i]);
    g.first = i;
    while (i < count && padded(P.order[(size_t)i]) == g.np) ++i;
    g.count = i - g.first;
    g.lo = make_layout(g.np, d, base, 1);  // room for either Sigma-product variant
    g.stride = al(g.lo.total);
    g.off = o;
    o += g.stride * (size_t)g.count;
    P.groups.push_back(g);
  }
  P.total = o;
  return P;
}

// What gpk_batched_plan wrote into a workspace, keyed by its address, so that
// gpk_batched_mll_grad can refuse (on the host, without a sync) a workspace
// whose device-side problem map belongs to other offsets: the kernels index x
// and y through that map.
struct PlanRecord {
  int64_t count = 0, d = 0;
  int base = GPK_PREC_FP64;
  uint64_t hash = 0;
};
std::mutex g_plan_mu;
std::unordered_map<const void*, PlanRecord> g_plans;

uint64_t hash_offsets(const int64_t* off, int64_t count) {
  uint64_t h = 1469598103934665603ull;  // FNV-1a over the offsets' bytes
  const unsigned char* b = reinterpret_cast<const unsigned char*>(off);
  for (size_t i = 0; i < (size_t)(count + 1) * sizeof(int64_t); ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}

#ifndef GPK_BATCH_STREAMS
#define GPK_BATCH_STREAMS 4
#endif
constexpr int BATCH_STREAMS = GPK_BATCH_STREAMS;

// ---- autotuning of the batched path ----
// gpk_batched_tune times the launch variants (Variant: 4 combinations) of each
// size group on the caller's own problems and keeps the fastest in a
// process-wide table keyed by (device architecture, kernel kind, prec, padded
// size, member-count class = ceil(log2 members), d), guarded by a mutex;
// gpk_batched_mll_grad reads it per group.  Groups without an entry run the
// untuned default (recomputed Sigma products, column reads in the fused pass).
struct TuneKey {
  std::string arch;
  int kind = 0, prec = 0, d = 0, members_class = 0;
  int64_t np = 0;
  bool operator<(const TuneKey& o) const {
    return std::tie(arch, kind, prec, d, members_class, np) <
           std::tie(o.arch, o.kind, o.prec, o.d, o.members_class, o.np);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, Variant> g_tune;
constexpr int NVARIANT = 4;

Variant variant_of(int id) {
  Variant v;
  v.stored_sigma = (id & 1) != 0;
  v.adj_transposed = (id & 2) != 0;
  return v;
}

// gcnArchName of the current device (cached per device ordinal).
int current_arch(std::string* arch) {
  static std::mutex mu;
  static std::map<int, std::string> names;
  int dev = 0;
  GPK_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(mu);
  auto it = names.find(dev);
  if (it == names.end()) {
    hipDeviceProp_t prop;
    GPK_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    it = names.emplace(dev, std::string(prop.gcnArchName)).first;
  }
  *arch = it->second;
  return GPK_OK;
}

TuneKey tune_key(const std::string& arch, int kind, int prec, const BatchGroup& g, int d) {
  TuneKey k;
  k.arch = arch;
  k.kind = kind;
  k.prec = prec;
  k.d = d;
  k.np = g.np;
  int c = 0;
  while (((int64_t)1 << c) < g.count) ++c;
  k.members_class = c;
  return k;
}

Batch group_batch(const BatchPlan& P, const BatchGroup& g, void* ws, int nls) {
  Batch bt;
  bt.ws_stride = (int64_t)g.stride;
  bt.gp = at<int>(ws, P.off_gp) + g.first;
  bt.xoff = at<int64_t>(ws, P.off_xoff);
  bt.theta_stride = nls + 3;
  bt.count = (unsigned)g.count;
  return bt;
}

template <typename TS>
int run_group(int kind, const Arith3& ar, const double* x, const double* y, int d,
              const double* theta, int nls, double jitter, int iters, void* ws,
              const BatchPlan& P, const BatchGroup& g, const Variant& vr, double* result,
              double* grad, hipStream_t s) {
  if (g.count > 65535) {
    set_error("batched: %lld problems of padded size %lld exceed one launch",
              (long long)g.count, (long long)g.np);
    return GPK_ERR_BAD_ARG;
  }
  const Batch bt = group_batch(P, g, ws, nls);
  void* wsg = static_cast<char*>(ws) + g.off;
  // n is taken per member inside the kernels; np fixes the shapes.
  GPK_TRY(mll_forward_t<TS>(kind, ar, x, y, g.np, d, theta, nls, jitter, iters, wsg, g.lo,
                            result, s, bt, vr));
  return mll_backward_t<TS>(kind, ar, x, g.np, d, theta, nls, 1.0, wsg, g.lo, grad, s, bt, vr);
}

template <typename TS>
int batched_mll_grad(int kind, int prec, const Arith3& ar, const double* x, const double* y,
                     const int64_t* off, int64_t count, int d, const double* theta, int nls,
                     double jitter, int iters, void* ws, double* result, double* grad,
                     hipStream_t s) {
  const int base = std::is_same<TS, double>::value ? GPK_PREC_FP64 : GPK_PREC_MIXED;
  const BatchPlan P = make_batch_plan(off, count, d, base);
  std::string arch;
  GPK_TRY(current_arch(&arch));
  // The size groups are independent (disjoint workspaces and outputs), and so
  // are the members of one group: the work is cut into pieces -- whole groups,
  // or contiguous member ranges of a group heavier than 1/BATCH_STREAMS of the
  // batch (count * np^3) -- dealt largest first to the least loaded of up to
  // BATCH_STREAMS streams, the caller's and library-owned ones, forked from and
  // joined back to the caller's by events, so one piece's latency-bound launches
  // overlap another's large ones.
  struct Piece {
    int group;
    BatchGroup g;
    double cost;
  };
  std::vector<Piece> pieces;
  double total = 0.0;
  for (const BatchGroup& g : P.groups) {
    const double np = (double)g.np;
    total += (double)g.count * np * np * np;
  }
  const int ns_max = std::min(BATCH_STREAMS, 1 + AUX_STREAMS);
  for (int gi = 0; gi < (int)P.groups.size(); ++gi) {
    const BatchGroup& g = P.groups[gi];
    const double np = (double)g.np;
    const double cost = (double)g.count * np * np * np;
    int64_t k = ns_max > 1 ? (int64_t)std::ceil(cost / (total / ns_max) - 1e-9) : 1;
    k = std::max<int64_t>(1, std::min<int64_t>(k, g.count));
    for (int64_t q = 0; q < k; ++q) {
      const int64_t a = g.count * q / k, b = g.count * (q + 1) / k;
      Piece pc{gi, g, 0.0};
      pc.g.first = g.first + a;
      pc.g.count = b - a;
      pc.g.off = g.off + (size_t)a * g.stride;
      pc.cost = (double)pc.g.count * np * np * np;
      pieces.push_back(pc);
    }
  }
  const int np_ = (int)pieces.size();
  const int ns = std::min(ns_max, np_);
  hipStream_t st[1 + AUX_STREAMS] = {s};
  hipEvent_t ev[AUX_EVENTS] = {};
  std::vector<int> on(np_, 0);
  if (ns > 1) {
    GPK_TRY(aux_streams(st + 1, ev));
    std::vector<int> order(np_);
    for (int i = 0; i < np_; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return pieces[a].cost > pieces[b].cost; });
    std::vector<double> load(ns, 0.0);
    for (int i : order) {
      const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      on[i] = k;
      load[k] += pieces[i].cost;
    }
    GPK_HIP_TRY(hipEventRecord(ev[3], s));
    for (int k = 1; k < ns; ++k) GPK_HIP_TRY(hipStreamWaitEvent(st[k], ev[3], 0));
  }
  // joins every stream used back into the caller's (an early error return included)
  struct JoinAll {
    hipStream_t* st;
    hipEvent_t* ev;
    int ns;
    ~JoinAll() {
      for (int k = 1; k < ns; ++k)
        if (hipEventRecord(ev[3 + k], st[k]) == hipSuccess)
          (void)hipStreamWaitEvent(st[0], ev[3 + k], 0);
    }
  } join{st, ev, ns};
  for (int i = 0; i < np_; ++i) {
    Variant vr;
    vr.stored_sigma = false;
    vr.adj_transposed = false;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune.find(tune_key(arch, kind, prec, P.groups[pieces[i].group], d));
      if (it != g_tune.end()) vr = it->second;
    }
    GPK_TRY(run_group<TS>(kind, ar, x, y, d, theta, nls, jitter, iters, ws, P, pieces[i].g, vr,
                          result, grad, st[on[i]]));
  }
  return GPK_OK;
}

// Times every variant of every group (one untimed run, then the mean of
// `reps` runs between two events on the stream) and records the fastest.
template <typename TS>
int batched_tune(int kind, int prec, const Arith3& ar, const double* x, const double* y,
                 const int64_t* off, int64_t count, int d, const double* theta, int nls,
                 double jitter, int iters, void* ws, double* result, double* grad,
                 float* report, int report_len, hipStream_t s) {
  const int base = std::is_same<TS, double>::value ? GPK_PREC_FP64 : GPK_PREC_MIXED;
  const BatchPlan P = make_batch_plan(off, count, d, base);
  std::string arch;
  GPK_TRY(current_arch(&arch));
  constexpr int reps = 3;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  GPK_HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_error("tune: hipEventCreate failed");
    return GPK_ERR_HIP;
  }
  int rc = GPK_OK;
  int gi = 0;
  for (const BatchGroup& g : P.groups) {
    float ms[NVARIANT] = {};
    int best = 0;
    for (int v = 0; v < NVARIANT && rc == GPK_OK; ++v) {
      const Variant vr = variant_of(v);
      rc = run_group<TS>(kind, ar, x, y, d, theta, nls, jitter, iters, ws, P, g, vr, result,
                         grad, s);
      if (rc == GPK_OK && hipEventRecord(e0, s) != hipSuccess) {
        set_error("tune: hipEventRecord failed");
        rc = GPK_ERR_HIP;
      }
      for (int r = 0; r < reps && rc == GPK_OK; ++r)
        rc = run_group<TS>(kind, ar, x, y, d, theta, nls, jitter, iters, ws, P, g, vr, result,
                           grad, s);
      if (rc == GPK_OK && (hipEventRecord(e1, s) != hipSuccess ||
                           hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&ms[v], e0, e1) != hipSuccess)) {
        set_error("tune: event timing failed");
        rc = GPK_ERR_HIP;
      }
      ms[v] /= reps;
      if (ms[v] < ms[best]) best = v;
    }
    if (rc != GPK_OK) break;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune[tune_key(arch, kind, prec, g, d)] = variant_of(best);
    }
    if (report && (gi + 1) * 7 <= report_len) {
      float* q = report + gi * 7;
      q[0] = (float)g.np;
      q[1] = (float)g.count;
      q[2] = (float)best;
      for (int v = 0; v < NVARIANT; ++v) q[3 + v] = ms[v];
    }
    // leave this group's outputs as the chosen variant computes them
    rc = run_group<TS>(kind, ar, x, y, d, theta, nls, jitter, iters, ws, P, g, variant_of(best),
                       result, grad, s);
    if (rc != GPK_OK) break;
    ++gi;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// Argument, workspace and plan checks shared by gpk_batched_mll_grad and
// gpk_batched_tune (host only, no sync).
int check_batched_call(const char* fn, int kind, int prec, const double* x, const double* y,
                       const int64_t* off, int64_t count, int64_t d, const double* theta,
                       int nls, const void* ws, size_t ws_bytes, const double* result,
                       const double* grad, Arith3* ar, int* base_out) {
  int rc = check_offsets(off, count, d);
  if (rc) return rc;
  const int base = batched_prec(prec, ar);
  if (base < 0) return GPK_ERR_BAD_ARG;
  rc = check_common(kind, prec, x, off[count], d, theta, nls);
  if (rc) return rc;
  if (!y || !ws || !result || !grad) {
    set_error("%s: null y/ws/result/grad", fn);
    return GPK_ERR_BAD_ARG;
  }
  const size_t need = make_batch_plan(off, count, d, base).total;
  if (ws_bytes < need) {
    set_error("%s: workspace %zu bytes < required %zu", fn, ws_bytes, need);
    return GPK_ERR_WORKSPACE;
  }
  if (reinterpret_cast<uintptr_t>(ws) % ALIGN) {
    set_error("%s: workspace must be %zu-byte aligned", fn, ALIGN);
    return GPK_ERR_BAD_ARG;
  }
  std::lock_guard<std::mutex> g(g_plan_mu);
  auto it = g_plans.find(ws);
  if (it == g_plans.end()) {
    set_error("%s: no gpk_batched_plan has run on this workspace", fn);
    return GPK_ERR_BAD_ARG;
  }
  const PlanRecord& r = it->second;
  if (r.count != count || r.d != d || r.base != base || r.hash != hash_offsets(off, count)) {
    set_error("%s: the workspace was planned for other offsets/count/d/prec", fn);
    return GPK_ERR_BAD_ARG;
  }
  *base_out = base;
  return GPK_OK;
}

// ---- allocation-free exact-GP evaluation ----
// For callers that cannot allocate device memory (code under HIP-graph
// capture, serving loops, harnesses that forbid hipMalloc): the workspace is a
// file-scope device array instead of a caller-provided buffer.  Each device
// holds its own copy for the life of the process; host code reaches it only
// through the address hipGetSymbolAddress returns on that device (the array's
// name is not a device address in host code).
constexpr int64_t NOALLOC_MAX_N = 1024;
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)20 << 20;  // 20 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Address of this device's scratch array, looked up once per device (with
// that device current) and kept.
int device_scratch(void** out) {
  int dev = -1;
  GPK_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_error("noalloc: device %d outside the %d devices supported", dev, NOALLOC_MAX_DEVICES);
    return GPK_ERR_NO_DEVICE;
  }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    GPK_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("noalloc: hipGetSymbolAddress gave no address for the scratch array");
      return GPK_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return GPK_OK;
}

}  // namespace

struct AuxRes {
  hipStream_t s[AUX_STREAMS] = {};
  hipEvent_t ev[AUX_EVENTS] = {};
};

int aux_resources(AuxRes** out) {
  thread_local std::map<int, AuxRes> res;
  int dev = 0;
  GPK_HIP_TRY(hipGetDevice(&dev));
  auto it = res.find(dev);
  if (it == res.end()) {
    AuxRes r;
    for (int i = 0; i < AUX_STREAMS; ++i)
      GPK_HIP_TRY(hipStreamCreateWithFlags(&r.s[i], hipStreamNonBlocking));
    for (int i = 0; i < AUX_EVENTS; ++i)
      GPK_HIP_TRY(hipEventCreateWithFlags(&r.ev[i], hipEventDisableTiming));
    it = res.emplace(dev, r).first;
  }
  *out = &it->second;
  return GPK_OK;
}

int aux_streams(hipStream_t* streams, hipEvent_t* events) {
  AuxRes* r = nullptr;
  GPK_TRY(aux_resources(&r));
  for (int i = 0; i < AUX_STREAMS; ++i) streams[i] = r->s[i];
  for (int i = 0; i < AUX_EVENTS; ++i) events[i] = r->ev[i];
  return GPK_OK;
}

// The look-ahead Cholesky uses aux stream 0 and events 0..2.
int lookahead_resources(hipStream_t* s2, hipEvent_t* ev3) {
  AuxRes* r = nullptr;
  GPK_TRY(aux_resources(&r));
  *s2 = r->s[0];
  for (int i = 0; i < 3; ++i) ev3[i] = r->ev[i];
  return GPK_OK;
}

}  // namespace gpk

using namespace gpk;

extern "C" {

int gpk_mll_forward_noalloc(int kind, int prec, const double* x, const double* y, int64_t n,
                            int64_t d, const double* theta, int nls, double jitter,
                            int refine_iters, double* result, double* grad, void* stream) {
  int rc = check_common(kind, prec, x, n, d, theta, nls);
  if (rc) return rc;
  if (!y || !result) {
    set_error("mll_forward_noalloc: null y/result");
    return GPK_ERR_BAD_ARG;
  }
  if (n > NOALLOC_MAX_N) {
    set_error("mll_forward_noalloc: n = %lld exceeds the %lld rows its fixed scratch holds",
              (long long)n, (long long)NOALLOC_MAX_N);
    return GPK_ERR_BAD_ARG;
  }
  Arith3 ar;
  const int base = decode_prec(prec, &ar);
  const Layout Lo = make_layout(n, d, base);
  if (Lo.total > NOALLOC_SCRATCH_BYTES) {
    set_error("mll_forward_noalloc: %zu bytes of scratch needed, %zu available", Lo.total,
              NOALLOC_SCRATCH_BYTES);
    return GPK_ERR_WORKSPACE;
  }
  void* ws = nullptr;
  if ((rc = device_scratch(&ws))) return rc;
  const int iters = refine_iters >= 0 ? refine_iters : (base == GPK_PREC_FP64 ? 1 : 6);
  hipStream_t s = (hipStream_t)stream;
  if (base == GPK_PREC_FP64) {
    GPK_TRY(mll_forward_t<double>(kind, ar, x, y, n, (int)d, theta, nls, jitter, iters, ws, Lo,
                                  result, s));
    if (grad)
      GPK_TRY(mll_backward_t<double>(kind, ar, x, n, (int)d, theta, nls, 1.0, ws, Lo, grad, s));
    return GPK_OK;
  }
  GPK_TRY(mll_forward_t<float>(kind, ar, x, y, n, (int)d, theta, nls, jitter, iters, ws, Lo,
                               result, s));
  if (grad)
    GPK_TRY(mll_backward_t<float>(kind, ar, x, n, (int)d, theta, nls, 1.0, ws, Lo, grad, s));
  return GPK_OK;
}

const char* gpk_version(void) { return "gpk 0.1.0 (gfx950)"; }

const char* gpk_last_error(void) { return g_err; }

int gpk_cross_covariance(int kind, const double* x, int64_t n, const double* z, int64_t m,
                         int64_t d, const double* theta, int nls, double* out, int64_t ldo,
                         void* stream) {
  int rc = check_common(kind, GPK_PREC_FP64, x, n, d, theta, nls);
  if (rc) return rc;
  if (!z || !out || m < 1 || ldo < m) {
    set_error("cross_covariance: bad z/out/m/ldo");
    return GPK_ERR_BAD_ARG;
  }
  return launch_kernel_matrix<double>(kind, x, n, z, m, (int)d, theta, nls, 0.0, 0, out, ldo, n,
                                      m, false, (hipStream_t)stream);
}

int gpk_gram(int kind, const double* x, int64_t n, int64_t d, const double* theta, int nls,
             double jitter, int add_noise, double* out, int64_t ldo, void* stream) {
  int rc = check_common(kind, GPK_PREC_FP64, x, n, d, theta, nls);
  if (rc) return rc;
  if (!out || ldo < n) {
    set_error("gram: bad out/ldo");
    return GPK_ERR_BAD_ARG;
  }
  return launch_kernel_matrix<double>(kind, x, n, x, n, (int)d, theta, nls, jitter,
                                      add_noise ? 2 : 1, out, ldo, n, n, false,
                                      (hipStream_t)stream);
}

size_t gpk_mll_workspace_bytes(int64_t n, int64_t d, int prec) {
  const int base = decode_prec(prec, nullptr);
  if (n < 1 || d < 1 || base < 0) return 0;
  return make_layout(n, d, base).total;
}

int gpk_mll_forward(int kind, int prec, const double* x, const double* y, int64_t n, int64_t d,
                    const double* theta, int nls, double jitter, int refine_iters, void* ws,
                    size_t ws_bytes, double* result, void* stream) {
  int rc = check_common(kind, prec, x, n, d, theta, nls);
  if (rc) return rc;
  if (!y || !ws || !result) {
    set_error("mll_forward: null y/ws/result");
    return GPK_ERR_BAD_ARG;
  }
  Arith3 ar;
  const int base = decode_prec(prec, &ar);
  const Layout Lo = make_layout(n, d, base);
  if ((rc = check_ws("mll_forward", ws, ws_bytes, Lo.total))) return rc;
  const int iters = refine_iters >= 0 ? refine_iters : (base == GPK_PREC_FP64 ? 1 : 6);
  hipStream_t s = (hipStream_t)stream;
  if (base == GPK_PREC_FP64)
    return mll_forward_t<double>(kind, ar, x, y, n, (int)d, theta, nls, jitter, iters, ws, Lo,
                                 result, s);
  return mll_forward_t<float>(kind, ar, x, y, n, (int)d, theta, nls, jitter, iters, ws, Lo,
                              result, s);
}

int gpk_mll_backward(int kind, int prec, const double* x, int64_t n, int64_t d,
                     const double* theta, int nls, double gbar_host, void* ws, size_t ws_bytes,
                     double* grad, void* stream) {
  int rc = check_common(kind, prec, x, n, d, theta, nls);
  if (rc) return rc;
  if (!ws || !grad) {
    set_error("mll_backward: null ws/grad");
    return GPK_ERR_BAD_ARG;
  }
  Arith3 ar;
  const int base = decode_prec(prec, &ar);
  const Layout Lo = make_layout(n, d, base);
  if ((rc = check_ws("mll_backward", ws, ws_bytes, Lo.total))) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (base == GPK_PREC_FP64)
    return mll_backward_t<double>(kind, ar, x, n, (int)d, theta, nls, gbar_host, ws, Lo, grad,
                                  s);
  return mll_backward_t<float>(kind, ar, x, n, (int)d, theta, nls, gbar_host, ws, Lo, grad, s);
}

size_t gpk_predict_workspace_bytes(int64_t n, int64_t d, int64_t t, int prec, int full_cov) {
  const int base = decode_prec(prec, nullptr);
  if (n < 1 || d < 1 || t < 1 || base < 0) return 0;
  return make_pred_layout(make_layout(n, d, base), t, full_cov).total;
}

int gpk_predict(int kind, int prec, const double* x, const double* y, int64_t n, int64_t d,
                const double* xt, int64_t t, const double* theta, int nls, double jitter_post,
                double jitter_prior, int full_cov, int refine_iters, void* ws, size_t ws_bytes,
                double* mean, double* cov, double* result, void* stream) {
  int rc = check_common(kind, prec, x, n, d, theta, nls);
  if (rc) return rc;
  if (!y || !xt || t < 1 || !ws || !mean || !cov) {
    set_error("predict: null y/xt/ws/mean/cov or t < 1");
    return GPK_ERR_BAD_ARG;
  }
  Arith3 ar;
  const int base = decode_prec(prec, &ar);
  // The predictive variance is k(t, t) - ||X k_t||^2 with X ~ L^-1 itself, so
  // plain GPK_PREC_MIXED takes six-product splits for the triangular inverse
  // here (three-product ones left 2e-5 of the largest variance at N = 16,384,
  // T = 1e6; the inverse is a small share of a predict call).
  if (prec == GPK_PREC_MIXED) ar.trtri = AR_BF16X6;
  const Layout Lo = make_layout(n, d, base);
  const PredLayout P = make_pred_layout(Lo, t, full_cov);
  if ((rc = check_ws("predict", ws, ws_bytes, P.total))) return rc;
  const int iters = refine_iters >= 0 ? refine_iters : (base == GPK_PREC_FP64 ? 1 : 6);
  hipStream_t s = (hipStream_t)stream;
  if (base == GPK_PREC_FP64)
    return predict_t<double>(kind, ar, x, y, n, (int)d, xt, t, theta, nls, jitter_post,
                             jitter_prior, full_cov, iters, ws, Lo, P, mean, cov, result, s);
  return predict_t<float>(kind, ar, x, y, n, (int)d, xt, t, theta, nls, jitter_post,
                          jitter_prior, full_cov, iters, ws, Lo, P, mean, cov, result, s);
}

size_t gpk_loocv_workspace_bytes(int64_t n, int64_t d, int prec) {
  const int base = decode_prec(prec, nullptr);
  if (n < 1 || d < 1 || base < 0) return 0;
  return make_loo_layout(make_layout(n, d, base)).total;
}

int gpk_loocv_forward(int kind, int prec, const double* x, const double* y, int64_t n, int64_t d,
                      const double* theta, int nls, double jitter, int refine_iters, void* ws,
                      size_t ws_bytes, double* result, void* stream) {
  int rc = check_common(kind, prec, x, n, d, theta, nls);
  if (rc) return rc;
  if (!y || !ws || !result) {
    set_error("loocv_forward: null y/ws/result");
    return GPK_ERR_BAD_ARG;
  }
  Arith3 ar;
  const int base = decode_prec(prec, &ar);
  const Layout Lo = make_layout(n, d, base);
  const LooLayout P = make_loo_layout(Lo);
  if ((rc = check_ws("loocv_forward", ws, ws_bytes, P.total))) return rc;
  const int iters = refine_iters >= 0 ? refine_iters : (base == GPK_PREC_FP64 ? 1 : 6);
  hipStream_t s = (hipStream_t)stream;
  if (base == GPK_PREC_FP64)
    return loocv_forward_t<double>(kind, ar, x, y, n, (int)d, theta, nls, jitter, iters, ws, Lo,
                                   P, result, s);
  return loocv_forward_t<float>(kind, ar, x, y, n, (int)d, theta, nls, jitter, iters, ws, Lo, P,
                                result, s);
}

int gpk_loocv_backward(int kind, int prec, const double* x, int64_t n, int64_t d,
                       const double* theta, int nls, double gbar_host, void* ws, size_t ws_bytes,
                       double* grad, void* stream) {
  int rc = check_common(kind, prec, x, n, d, theta, nls);
  if (rc) return rc;
  if (!ws || !grad) {
    set_error("loocv_backward: null ws/grad");
    return GPK_ERR_BAD_ARG;
  }
  Arith3 ar;
  const int base = decode_prec(prec, &ar);
  const Layout Lo = make_layout(n, d, base);
  const LooLayout P = make_loo_layout(Lo);
  if ((rc = check_ws("loocv_backward", ws, ws_bytes, P.total))) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (base == GPK_PREC_FP64)
    return loocv_backward_t<double>(kind, ar, x, n, (int)d, theta, nls, gbar_host, ws, Lo, P,
                                    grad, s);
  return loocv_backward_t<float>(kind, ar, x, n, (int)d, theta, nls, gbar_host, ws, Lo, P, grad,
                                 s);
}

size_t gpk_batched_workspace_bytes(const int64_t* offsets_host, int64_t count, int64_t d,
                                   int prec) {
  const int base = batched_prec(prec, nullptr);
  if (base < 0 || check_offsets(offsets_host, count, d) != GPK_OK) return 0;
  return make_batch_plan(offsets_host, count, d, base).total;
}

int gpk_batched_plan(const int64_t* offsets_host, int64_t count, int64_t d, int prec, void* ws,
                     size_t ws_bytes) {
  int rc = check_offsets(offsets_host, count, d);
  if (rc) return rc;
  const int base = batched_prec(prec, nullptr);
  if (base < 0) return GPK_ERR_BAD_ARG;
  const BatchPlan P = make_batch_plan(offsets_host, count, d, base);
  if (!ws || ws_bytes < P.total) {
    set_error("batched_plan: workspace %zu bytes < required %zu", ws_bytes, P.total);
    return GPK_ERR_WORKSPACE;
  }
  if (reinterpret_cast<uintptr_t>(ws) % ALIGN) {
    set_error("batched_plan: workspace must be %zu-byte aligned", ALIGN);
    return GPK_ERR_BAD_ARG;
  }
  GPK_HIP_TRY(hipMemcpy(at<int>(ws, P.off_gp), P.order.data(), (size_t)count * sizeof(int),
                        hipMemcpyHostToDevice));
  GPK_HIP_TRY(hipMemcpy(at<int64_t>(ws, P.off_xoff), offsets_host,
                        (size_t)(count + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  std::lock_guard<std::mutex> g(g_plan_mu);
  g_plans[ws] = PlanRecord{count, d, base, hash_offsets(offsets_host, count)};
  return GPK_OK;
}

int gpk_batched_mll_grad(int kind, int prec, const double* x, const double* y,
                         const int64_t* offsets_host, int64_t count, int64_t d,
                         const double* theta, int nls, double jitter, int refine_iters, void* ws,
                         size_t ws_bytes, double* result, double* grad, void* stream) {
  Arith3 ar;
  int base = 0;
  int rc = check_batched_call("batched_mll_grad", kind, prec, x, y, offsets_host, count, d,
                              theta, nls, ws, ws_bytes, result, grad, &ar, &base);
  if (rc) return rc;
  const int iters = refine_iters >= 0 ? refine_iters : (base == GPK_PREC_FP64 ? 1 : 6);
  if (base == GPK_PREC_FP64)
    return batched_mll_grad<double>(kind, prec, ar, x, y, offsets_host, count, (int)d, theta,
                                    nls, jitter, iters, ws, result, grad, (hipStream_t)stream);
  return batched_mll_grad<float>(kind, prec, ar, x, y, offsets_host, count, (int)d, theta, nls,
                                 jitter, iters, ws, result, grad, (hipStream_t)stream);
}

int gpk_batched_tune(int kind, int prec, const double* x, const double* y,
                     const int64_t* offsets_host, int64_t count, int64_t d, const double* theta,
                     int nls, double jitter, int refine_iters, void* ws, size_t ws_bytes,
                     double* result, double* grad, float* report_host, int report_len,
                     void* stream) {
  Arith3 ar;
  int base = 0;
  int rc = check_batched_call("batched_tune", kind, prec, x, y, offsets_host, count, d, theta,
                              nls, ws, ws_bytes, result, grad, &ar, &base);
  if (rc) return rc;
  const int iters = refine_iters >= 0 ? refine_iters : (base == GPK_PREC_FP64 ? 1 : 6);
  if (base == GPK_PREC_FP64)
    return batched_tune<double>(kind, prec, ar, x, y, offsets_host, count, (int)d, theta, nls,
                                jitter, iters, ws, result, grad, report_host, report_len,
                                (hipStream_t)stream);
  return batched_tune<float>(kind, prec, ar, x, y, offsets_host, count, (int)d, theta, nls,
                             jitter, iters, ws, result, grad, report_host, report_len,
                             (hipStream_t)stream);
}

int gpk_profile_enable(int on) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof_on = on != 0;
  return GPK_OK;
}

int gpk_profile_collect(double* ms_host, int* counts_host) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  for (int p = 0; p < GPK_PHASE_COUNT; ++p) {
    if (ms_host) ms_host[p] = 0.0;
    if (counts_host) counts_host[p] = 0;
  }
  int rc = GPK_OK;
  for (auto& rec : g_prof) {
    float ms = 0.f;
    if (rc == GPK_OK) {
      hipError_t e = hipEventSynchronize(rec.b);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, rec.a, rec.b);
      if (e != hipSuccess) {
        set_error("profile: %s", hipGetErrorString(e));
        rc = GPK_ERR_HIP;
      }
    }
    if (rc == GPK_OK && rec.phase >= 0 && rec.phase < GPK_PHASE_COUNT) {
      if (ms_host) ms_host[rec.phase] += ms;
      if (counts_host) counts_host[rec.phase] += 1;
    }
    (void)hipEventDestroy(rec.a);
    (void)hipEventDestroy(rec.b);
  }
  g_prof.clear();
  return rc;
}

int gpk_mfma_selftest(double* out_f64, float* out_bf16, void* stream) {
  if (!out_f64 || !out_bf16) {
    set_error("selftest: null output");
    return GPK_ERR_BAD_ARG;
  }
  return launch_mfma_selftest(out_f64, out_bf16, (hipStream_t)stream);
}

}  // extern "C"
