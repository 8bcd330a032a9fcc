// Internal declarations shared by the libgpk HIP sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <type_traits>

#include "../../include/gpk.h"

namespace gpk {

// Block size of the recursive factorisation: leaves are NB x NB diagonal
// blocks factored in LDS, and every GEMM tile is NB x NB.  Internal matrices
// are padded to a multiple of NB with an identity tail.
constexpr int NB = 128;
// Largest input dimension of the exact-GP kernels (the Sigma products keep a
// point's coordinates in registers; the gradient epilogue stages them in LDS in
// chunks of 32).
constexpr int MAX_D = 64;

void set_error(const char* fmt, ...);

#define GPK_HIP_TRY(call)                                                            \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      ::gpk::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                       __LINE__);                                                    \
      return GPK_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

#define GPK_LAUNCH_CHECK()                                                           \
  do {                                                                               \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess) {                                                          \
      ::gpk::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_),     \
                       __FILE__, __LINE__);                                          \
      return GPK_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// ---------------------------------------------------------------------------
// Batches of independent problems of one padded size, one member per
// blockIdx.z.  The workspace of member z sits ws_stride bytes after that of
// member z - 1, so every workspace pointer of a kernel is shifted by
// z * ws_stride.  Inputs and outputs are addressed through the member's
// problem index p = gp[z]: rows xoff[p] .. xoff[p+1]-1 of the concatenated
// x (and y), hyperparameter row theta + p * theta_stride, output row p.  The
// default (gp == nullptr, ws_stride == 0, one member) is a single problem.
// ---------------------------------------------------------------------------
struct Batch {
  int64_t ws_stride = 0;
  const int* gp = nullptr;
  const int64_t* xoff = nullptr;
  int64_t theta_stride = 0;
  unsigned count = 1;
  // Optional device flag (member z's at skip + z * ws_stride bytes): the
  // vector kernels of an iterative refinement step return at once when it is
  // set, so a converged refinement skips its remaining steps without a host
  // sync.
  const int* skip = nullptr;
};

// True when this member's refinement has converged (see Batch::skip).
__device__ __forceinline__ bool bskip(const Batch& b) {
  if (!b.skip) return false;
  const int* p = b.skip;
  if (b.ws_stride)
    p = reinterpret_cast<const int*>(reinterpret_cast<const char*>(p) +
                                     (int64_t)blockIdx.z * b.ws_stride);
  return *p != 0;
}

// The shift is byte arithmetic on the pointer itself (not through an integer),
// so the compiler keeps the global address space and emits global (not flat)
// loads and stores.
template <typename T>
__device__ __forceinline__ T* bws(T* p, const Batch& b) {
  if (!p || !b.ws_stride) return p;
  typedef typename std::conditional<std::is_const<T>::value, const char, char>::type Byte;
  return reinterpret_cast<T*>(reinterpret_cast<Byte*>(p) + (int64_t)blockIdx.z * b.ws_stride);
}
// Problem index of this member (0 for a single problem).
__device__ __forceinline__ int64_t bprob(const Batch& b) {
  return b.gp ? (int64_t)b.gp[blockIdx.z] : 0;
}
// Rebase this member's inputs: x rows, n and the hyperparameter row.
template <typename PX, typename PT>
__device__ __forceinline__ void bmember(const Batch& b, int d, PX& x, int64_t& n, PT& theta) {
  if (!b.gp) return;
  const int64_t p = b.gp[blockIdx.z];
  const int64_t r0 = b.xoff[p];
  x += r0 * d;
  n = b.xoff[p + 1] - r0;
  theta += p * b.theta_stride;
}

// ---------------------------------------------------------------------------
// Stationary kernel formulas (unit variance), from rho^2 of the
// lengthscale-scaled inputs.  kval() is k(rho); kfac() is F(rho) with
// dK/dl_d = var * F(rho) * u_d / l_d, u_d = (x_d - z_d)^2 / l_d^2.
//   RBF        gpjax/kernels/stationary/rbf.py:43          k = exp(-rho^2/2)
//   Matern-1/2 gpjax/kernels/stationary/matern12.py:47     k = exp(-rho)
//   Matern-3/2 gpjax/kernels/stationary/matern32.py:48-53  k = (1+s3 rho) exp(-s3 rho)
//   Matern-5/2 gpjax/kernels/stationary/matern52.py:47-52  k = (1+s5 rho+5/3 rho^2) exp(-s5 rho)
// Matern distances are sqrt(max(rho^2, 1e-36)) (stationary/utils.py:67).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double kval(int kind, double d2) {
  if (kind == GPK_KERNEL_RBF) return exp(-0.5 * d2);
  const double t = sqrt(fmax(d2, 1e-36));
  if (kind == GPK_KERNEL_MATERN12) return exp(-t);
  if (kind == GPK_KERNEL_MATERN32) {
    const double s3 = 1.7320508075688772;
    return (1.0 + s3 * t) * exp(-s3 * t);
  }
  const double s5 = 2.23606797749979;
  return (1.0 + s5 * t + (5.0 / 3.0) * t * t) * exp(-s5 * t);
}

// kval in float32 arithmetic (rho^2 formed in float64 by the caller), for
// entries that are stored in float32 anyway: a few float32 ulps of relative
// error instead of the half ulp of rounding a float64 value.
__device__ __forceinline__ float kvalf(int kind, float d2) {
  if (kind == GPK_KERNEL_RBF) return expf(-0.5f * d2);
  const float t = sqrtf(fmaxf(d2, 1e-36f));
  if (kind == GPK_KERNEL_MATERN12) return expf(-t);
  if (kind == GPK_KERNEL_MATERN32) {
    const float s3 = 1.7320508f;
    return (1.0f + s3 * t) * expf(-s3 * t);
  }
  const float s5 = 2.2360680f;
  return (1.0f + s5 * t + (5.0f / 3.0f) * t * t) * expf(-s5 * t);
}

// Both k(rho) and F(rho) with one exponential.
__device__ __forceinline__ void kval_fac(int kind, double d2, double& k, double& f) {
  if (kind == GPK_KERNEL_RBF) {
    k = exp(-0.5 * d2);
    f = k;
    return;
  }
  const double t = sqrt(fmax(d2, 1e-36));
  if (kind == GPK_KERNEL_MATERN12) {
    const double e = exp(-t);
    k = e;
    f = d2 > 1e-36 ? e / t : 0.0;
    return;
  }
  if (kind == GPK_KERNEL_MATERN32) {
    const double s3 = 1.7320508075688772;
    const double e = exp(-s3 * t);
    k = (1.0 + s3 * t) * e;
    f = 3.0 * e;
    return;
  }
  const double s5 = 2.23606797749979;
  const double e = exp(-s5 * t);
  k = (1.0 + s5 * t + (5.0 / 3.0) * t * t) * e;
  f = (5.0 / 3.0) * (1.0 + s5 * t) * e;
}

// kval_fac in float32 arithmetic (the mixed collapsed ELBO's gradient pass).
__device__ __forceinline__ void kval_facf(int kind, float d2, float& k, float& f) {
  if (kind == GPK_KERNEL_RBF) {
    k = expf(-0.5f * d2);
    f = k;
    return;
  }
  const float t = sqrtf(fmaxf(d2, 1e-36f));
  if (kind == GPK_KERNEL_MATERN12) {
    const float e = expf(-t);
    k = e;
    f = d2 > 1e-36f ? e / t : 0.0f;
    return;
  }
  if (kind == GPK_KERNEL_MATERN32) {
    const float s3 = 1.7320508f;
    const float e = expf(-s3 * t);
    k = (1.0f + s3 * t) * e;
    f = 3.0f * e;
    return;
  }
  const float s5 = 2.2360680f;
  const float e = expf(-s5 * t);
  k = (1.0f + s5 * t + (5.0f / 3.0f) * t * t) * e;
  f = (5.0f / 3.0f) * (1.0f + s5 * t) * e;
}

// ---------------------------------------------------------------------------
// Host-side launchers (each defined next to its kernels).
// ---------------------------------------------------------------------------

// Gram / cross-covariance tiles.  Rows i < rows_total, cols j < cols_total are
// written; entries with i >= n or j >= m are the identity pad (1 on i == j,
// else 0).  diag_mode & KM_DIAG_MASK: 0 none, 1 add jitter, 2 add jitter +
// obs_sd^2 (i == j < n); diag_mode | KM_ZERO_PAD makes the pad all zeros.
constexpr int KM_DIAG_MASK = 3;
constexpr int KM_ZERO_PAD = 4;
// lower_only: only 128-tiles on or below the diagonal (rows_total == cols_total).
// dblk (optional, lower_only mode): every diagonal NB x NB block is also
// written in float64 to dblk[b][NB][NB].
// spk (optional, lower_only mode): every lower NB-tile (I >= J) is also written
// in float64, packed: tile I(I+1)/2 + J at spk + (I(I+1)/2 + J) * NB * NB.
template <typename TS>
int launch_kernel_matrix(int kind, const double* x, int64_t n, const double* z, int64_t m,
                         int d, const double* theta, int nls, double jitter, int diag_mode,
                         TS* out, int64_t ldo, int64_t rows_total, int64_t cols_total,
                         bool lower_only, hipStream_t s, double* dblk = nullptr,
                         const Batch& bt = Batch(), double* spk = nullptr);

// Cross-covariance K[i][j] = var k(x_i, z_j) (i < n, j < m; zero elsewhere up
// to rows_total x cols_total) for long point streams (the collapsed ELBO's Kzx
// chunks: x = inducing inputs, z = the chunk's points; the predictive's test
// points).  Entries equal launch_kernel_matrix's (d > 32 runs it); f32k
// (float32 output only, d <= 32) evaluates the kernel function in float32
// arithmetic from the float64 squared distances (kvalf).
template <typename TS>
int launch_cross(int kind, const double* x, int64_t n, const double* z, int64_t m, int d,
                 const double* theta, int nls, bool f32k, TS* out, int64_t ldo,
                 int64_t rows_total, int64_t cols_total, hipStream_t s);

// Leaf: factor the NB x NB diagonal block at (o, o) of A in float64 in LDS,
// write L (lower, zero upper) to A, L^-1 to X (lower, zero upper) and to
// leafinv (float64), 2*sum(log diag) to *logdet, first bad pivot to *info.
// dsrc (optional): read the block from this float64 NB x NB array instead of A.
template <typename TS>
int launch_leaf(TS* A, TS* X, int64_t ld, int64_t o, double* leafinv, double* logdet, int* info,
                const double* dsrc, hipStream_t s, const Batch& bt = Batch());

// k-range modes of the tile GEMM.
enum KRange { KR_FULL = 0, KR_GE_N = 1, KR_LE_M = 2, KR_GE_M = 3 };
// Product arithmetic of the tile GEMM.  AR_FP16X3: float32 operands scaled by a
// power of two (so that every entry sits well inside the fp16 range), split
// into fp16 hi + lo planes, products hi*hi + hi*lo + lo*hi.
enum Arith { AR_F64 = 0, AR_BF16X3 = 3, AR_FP16X3 = 4, AR_BF16X6 = 6 };

// C[M x N] = alpha * sum_k A(m, k) B(n, k) (+ C if beta_one), M, N, K multiples
// of NB.  A(m, k) is A[m*lda + k] (a_km = false) or A[k*lda + m] (a_km = true);
// B(n, k) is B[n*ldb + k] (b_kn = false) or B[k*ldb + n] (b_kn = true).
// lower_tiles: only tiles with mt >= nt.  krange restricts k per tile.
// fscale (AR_FP16X3 only): device pointer to the power-of-two scales of A and B.
template <typename TA, typename TB, typename TC>
int launch_gemm(int arith, bool a_km, bool b_kn, const TA* A, int64_t lda, const TB* B,
                int64_t ldb, TC* C, int64_t ldc, int64_t M, int64_t N, int64_t K, double alpha,
                bool beta_one, bool lower_tiles, int krange, hipStream_t s,
                double* cdiag = nullptr, const Batch& bt = Batch(),
                const float* fscale = nullptr);

// Fused LAUUM + adjoint: for every lower NB-tile of Sinv = X^T X (X = L^-1,
// lower, Np x Np), reduce 1/2 tr(W dSigma/dtheta) contributions with
// W = alpha alpha^T - Sinv against kernel tiles recomputed from x.
// partial: [num_tiles][nls + 2] float64.
// beta (optional): W = alpha beta^T + beta alpha^T - X^T X instead; dense: X is
// a dense Np x Np matrix (every k) rather than lower triangular (k >= row).
// fscale (AR_FP16X3 only): device pointer to the power-of-two scale of X.
template <typename TS>
int launch_adjoint(int kind, int arith, const TS* X, int64_t ld, int64_t np, const double* x,
                   int64_t n, int d, const double* theta, int nls, const double* alpha,
                   double* partial, hipStream_t s, const Batch& bt = Batch(),
                   const double* beta = nullptr, bool dense = false,
                   const float* fscale = nullptr, bool xt = false);
// xt: the operand passed is U = X^T (upper, row-major) from launch_lower_to_upper.

// U = X^T for lower-triangular X (np x np, np a multiple of NB): upper U with
// zeros below the diagonal inside its diagonal NB-tiles.
template <typename TS>
int launch_lower_to_upper(const TS* X, TS* U, int64_t ld, int64_t np, hipStream_t s,
                          const Batch& bt = Batch());

// y = X r (trans = false) or y = X^T r (trans = true), X lower triangular
// np x np (only k <= i is read); float64 accumulation through `partial`
// (nv x ceil(np / 1024) x np doubles).  nv = 1 or 3 right-hand sides, stored
// np apart in r and in y, share each pass over X.
template <typename TS>
int launch_trmv(const TS* X, int64_t ld, int64_t np, const double* r, double* y, bool trans,
                double* partial, hipStream_t s, const Batch& bt = Batch(), int nv = 1);

// v = Sigma a, Sigma recomputed from x in float64 (identity pad for i >= n);
// nv = 1 or 3 vectors np apart share each kernel evaluation.
int launch_sigma_matvec(int kind, const double* x, int64_t n, int64_t np, int d,
                        const double* theta, int nls, double jitter, const double* a, double* v,
                        double* partial, hipStream_t s, const Batch& bt = Batch(), int nv = 1);

// Library-owned streams (default priority, non-blocking) and events for the
// calling thread on the current device, created on first use and kept for the
// process.  The look-ahead Cholesky takes stream 0 and events 0..2 (never
// inside a batched call); the batched driver takes the streams and events
// 3..3+AUX_STREAMS; the collapsed ELBO takes stream 0 and events 6 and 7 (it
// never runs the look-ahead Cholesky).
constexpr int AUX_STREAMS = 3;
constexpr int AUX_EVENTS = 8;
int aux_streams(hipStream_t* streams, hipEvent_t* events);
int lookahead_resources(hipStream_t* s2, hipEvent_t* ev3);

// Joins a forked library stream back into the caller's: at scope exit (an
// early error return included) `s` waits for everything issued to `s2` so far,
// so no work of a call is left unjoined.
struct StreamJoin {
  hipStream_t s, s2;
  hipEvent_t e;
  ~StreamJoin() {
    if (hipEventRecord(e, s2) == hipSuccess) (void)hipStreamWaitEvent(s, e, 0);
  }
};

// v = S a for a symmetric S held as its packed lower NB-tiles in float64 (the
// spk layout of launch_kernel_matrix), np x np; nv = 1 or 3 vectors np apart
// share each read of S.  partial holds 2 * nv * T(T+1)/2 * NB doubles, T = np/NB.
int launch_packed_symv(const double* S, int64_t np, const double* a, double* v,
                       double* partial, hipStream_t s, const Batch& bt = Batch(), int nv = 1);

// y_j = sum_i M[i][j] v_i (v non-null) or sum_i M[i][j]^2 (v null), dense
// rows x cols (float64 or float32 entries, float64 sums); partial holds
// ceil(rows/1024) x cols doubles.  lower: only rows i >= j enter column j.
template <typename TM>
int launch_colreduce(const TM* M, int64_t ld, int64_t rows, int64_t cols, const double* v,
                     double* y, double* partial, hipStream_t s, bool lower = false);

int launch_mfma_selftest(double* out64, float* outbf, hipStream_t s);

}  // namespace gpk
