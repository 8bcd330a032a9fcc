/*
 * gpk.h -- C ABI of libgpk, the MI355X (gfx950) exact-GP / collapsed-sparse-GP
 * hot path.
 *
 * Conventions
 *   - Every pointer argument is a DEVICE pointer unless the name ends in _host.
 *     Buffers are caller-owned; nothing is freed or retained by the library.
 *   - Every call is stream-ordered on `stream` (a hipStream_t passed as void*;
 *     NULL = the legacy default stream) and returns without a host sync, except
 *     the set-up calls gpk_batched_plan and gpk_batched_tune and
 *     gpk_profile_collect, which synchronise and say so.
 *   - Process-wide state: the batched plan registry and the autotune table
 *     (both mutex-guarded host tables), the profiler's event list, the
 *     per-device scratch of gpk_mll_forward_noalloc, and three library-owned
 *     streams (default priority) with a few events per calling thread and
 *     device.  Nothing else outlives a call; the arithmetic is chosen per call
 *     through `prec`.
 *   - Some calls fork work onto those library streams and join it back into
 *     `stream` with events before they return (the look-ahead Cholesky for
 *     n > 1,920, and for the collapsed ELBO's M x M factor when m > 1,920, the
 *     batched size buckets, the collapsed ELBO's Kzz-side work during its
 *     streamed chunks): to the caller the call stays
 *     ordered on `stream`, and under HIP-graph capture the fork and join become
 *     graph edges.  The streams and events are created by a thread's first
 *     such call on a device: a thread that captures should make one such call
 *     uncaptured first (a warm-up step), so that no stream is created inside
 *     the capture.
 *   - Matrices are row-major float64 with an explicit leading dimension.
 *   - Return value: GPK_OK or an error code; gpk_last_error() gives a
 *     thread-local message for the last failing call.
 *   - Numerical failure (a non-positive pivot in the Cholesky) is detected on
 *     the device and reported through the `result` buffer (info word and a NaN
 *     objective), never by a host sync.
 *
 * Hyperparameter vector `theta` (device, float64, length nls + 3):
 *   theta[0 .. nls-1]  lengthscale(s) l_d   (nls == 1: isotropic, nls == d: ARD)
 *   theta[nls]         kernel variance sigma_f^2
 *   theta[nls + 1]     likelihood obs_stddev sigma_n (a standard deviation)
 *   theta[nls + 2]     constant mean c (0 for the Zero mean)
 * Gradients use the same layout, in constrained coordinates.
 */
#ifndef GPK_H
#define GPK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define GPK_OK 0
#define GPK_ERR_BAD_ARG 1
/* 2 is unused: a non-positive pivot is found on the device, so it is reported
 * through result[GPK_RES_INFO] (1 + pivot index) and a NaN objective, never by
 * a status code (which would need a host sync). */
#define GPK_ERR_NO_DEVICE 3
#define GPK_ERR_HIP 4
#define GPK_ERR_WORKSPACE 5

/* stationary kernel families (gpjax/kernels/stationary/{rbf,matern12,matern32,matern52}.py) */
#define GPK_KERNEL_RBF 0
#define GPK_KERNEL_MATERN12 1
#define GPK_KERNEL_MATERN32 2
#define GPK_KERNEL_MATERN52 3

/* arithmetic of the factorisation path (the `prec` argument) */
#define GPK_PREC_FP64 0  /* float64 storage, FP64 MFMA products                    */
#define GPK_PREC_MIXED 1 /* float32 storage, BF16 split products, FP64 panels      */
/*
 * The mixed path with the products of each stage named per call: P for the
 * Cholesky updates, T for the triangular inverse, L for the fused Sigma^-1
 * pass, each 6 (bf16 hi/mid/lo split, six products: about float32-product
 * accuracy), 3 (bf16 hi/lo split, three products), 4 (fp16 hi/lo split of
 * power-of-two scaled operands) or 0 (float64 products of the float32
 * operands).  GPK_PREC_MIXED is GPK_PREC_MIXED_ARITH(6, 3, 6).  No counterpart
 * in the reference (which has one float64 path); it trades speed against the
 * accuracy figures the benchmark reports.
 */
#define GPK_PREC_ARITH_FLAG 0x100
#define GPK_PREC_MIXED_ARITH(p, t, l) \
  (GPK_PREC_MIXED | GPK_PREC_ARITH_FLAG | ((p) << 12) | ((t) << 16) | ((l) << 20))

/* result[] slots written by gpk_mll_forward */
#define GPK_RES_MLL 0      /* marginal log-likelihood (NaN if not PD)             */
#define GPK_RES_LOGDET 1   /* log|Sigma|                                          */
#define GPK_RES_QUAD 2     /* r^T Sigma^-1 r                                      */
#define GPK_RES_RESID 3    /* ||r - Sigma alpha|| / ||r|| after refinement        */
#define GPK_RES_INFO 4     /* 0, or 1 + index of the first non-positive pivot     */
/* Mixed path only (0 in the FP64 path): a-posteriori check of the computed
 * inverse factor X ~ L^-1 with E = I - Sigma X^T X, Sigma in float64, over
 * Rademacher probes z:                                                        */
#define GPK_RES_PROBE 5    /* sqrt(mean ||E z||^2 / n)  (~ ||E||_F / sqrt(n))     */
#define GPK_RES_TRACE 6    /* mean z^T E z  (estimates tr E ~ error of log|Sigma|) */
#define GPK_RES_FLAG 7     /* 1 when the check says the mixed result may be       */
                           /* outside 1e-6 (MLL) / 1e-4 (gradient): recompute in  */
                           /* GPK_PREC_FP64 (see GPK_RES_MLL_ERR)                  */
#define GPK_RES_SINV 8     /* sqrt(mean ||X^T X z||^2) (~ ||Sigma^-1||_F)          */
#define GPK_RES_MLL_ERR 9  /* model of |error| of the MLL from float32 storage:  */
                           /* 2^-25 max|Sigma| ||Sigma^-1||_F sqrt(max(1, n/1024)) */
                           /* (measured errors <= 8.1x it); flagged when 10x it   */
                           /* exceeds 1e-6 |MLL|, or ||E||_F > 0.75, or the       */
                           /* refinement stopped above 1e-10, and always when the */
                           /* Cholesky updates or the fused Sigma^-1 pass use     */
                           /* fp16 or three-product splits (the check is not      */
                           /* calibrated for their product errors)                */
#define GPK_RESULT_LEN 10

const char* gpk_version(void);
const char* gpk_last_error(void);

/*
 * Cross-covariance K[i, j] = var * k(x_i / l, z_j / l), i < n, j < m.
 * Replaces DenseKernelComputation._cross_covariance
 * (gpjax/kernels/computations/dense.py:32-36) behind AbstractKernel.cross_covariance
 * (gpjax/kernels/base.py:101-113).  Direct pairwise differences of the
 * lengthscale-scaled inputs, as gpjax/kernels/stationary/utils.py:42-67.
 */
int gpk_cross_covariance(int kind, const double* x, int64_t n, const double* z, int64_t m,
                         int64_t d, const double* theta, int nls, double* out, int64_t ldo,
                         void* stream);

/*
 * Gram K[i, j] = var * k(x_i, x_j) (+ jitter on the diagonal, + obs_stddev^2 when
 * add_noise != 0), full n x n.  Replaces AbstractKernelComputation.gram
 * (gpjax/kernels/computations/base.py:56-72) and the Sigma assembly of
 * gpjax/objectives.py:100-103 / gpjax/linalg/utils.py:39-65.
 */
int gpk_gram(int kind, const double* x, int64_t n, int64_t d, const double* theta, int nls,
             double jitter, int add_noise, double* out, int64_t ldo, void* stream);

/* Bytes of workspace that gpk_mll_forward / gpk_mll_backward need for (n, d, prec). */
size_t gpk_mll_workspace_bytes(int64_t n, int64_t d, int prec);

/*
 * Conjugate marginal log-likelihood, gpjax/objectives.py:36-107
 * (GaussianDistribution.log_prob, gpjax/distributions.py:115-134, whose
 * logdet/solve are gpjax/linalg/operations.py:109-111,163-165):
 *   mll = -1/2 [ n log 2pi + log|Sigma| + r^T Sigma^-1 r ],
 *   Sigma = K + (jitter + obs_stddev^2) I,  r = y - c.
 * Computes L = chol(Sigma) and L^-1 in the workspace, alpha = Sigma^-1 r with
 * at most `refine_iters` steps of float64 iterative refinement (< 0: 1 for
 * FP64, 6 for mixed; the steps stop once ||r - Sigma alpha|| <= 1e-12 ||r||,
 * decided on the device), and writes
 * result[GPK_RESULT_LEN] (slots GPK_RES_*).  The workspace then holds the state
 * that gpk_mll_backward consumes.  1 <= d <= 64 (the exact-GP entry points;
 * GPJax takes any d, gpjax/kernels/stationary/base.py:120-146).  For n > 8,064
 * the Cholesky forks onto a library stream and joins back (see Conventions).
 */
int gpk_mll_forward(int kind, int prec, const double* x, const double* y, int64_t n, int64_t d,
                    const double* theta, int nls, double jitter, int refine_iters, void* ws,
                    size_t ws_bytes, double* result, void* stream);

/*
 * Gradient of the MLL (what jax.grad of conjugate_mll returns,
 * tests/test_objectives.py:83-85, gpjax/fit.py:160), scaled by the host scalar
 * `gbar_host` (the upstream adjoint): grad = gbar * 1/2 tr((alpha alpha^T -
 * Sigma^-1) dSigma/dtheta), layout of theta.  Must follow gpk_mll_forward on the
 * same workspace and stream.  Sigma^-1 is never stored: each tile of
 * L^-T L^-1 is formed on the matrix cores and reduced against the kernel
 * derivative tile, recomputed from x, in the same pass.
 * The workspace is an in/out buffer of this call: the factor L in it is
 * replaced by L^-T (read along rows by that pass); L^-1 and alpha stay, so the
 * backward may be repeated.  A caller that treats its buffers as immutable
 * (an XLA FFI handler) declares ws as aliased input/output
 * (input_output_aliases) or passes a copy.
 */
int gpk_mll_backward(int kind, int prec, const double* x, int64_t n, int64_t d,
                     const double* theta, int nls, double gbar_host, void* ws, size_t ws_bytes,
                     double* grad, void* stream);

/*
 * gpk_mll_forward, and with grad != NULL also gpk_mll_backward (gbar = 1), with
 * no workspace argument and no allocation: the library's own fixed device
 * scratch array (20 MiB per device, held until the process ends) takes the
 * workspace's place, so n <= 1024 (larger n is refused with GPK_ERR_BAD_ARG).
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP-graph capture, tight serving loops, harnesses
 * that forbid hipMalloc), e.g. many small GPs in a Bayesian-optimisation loop.
 * Results equal gpk_mll_forward / gpk_mll_backward's.  Every call on a device
 * shares that device's scratch: calls that overlap on one device (concurrent
 * streams or threads) race.
 */
int gpk_mll_forward_noalloc(int kind, int prec, const double* x, const double* y, int64_t n,
                            int64_t d, const double* theta, int nls, double jitter,
                            int refine_iters, double* result, double* grad, void* stream);

/* Bytes of workspace gpk_elbo_forward / gpk_elbo_backward need. */
size_t gpk_elbo_workspace_bytes(int64_t n, int64_t m, int64_t d, int prec);

/*
 * Collapsed (Titsias) evidence lower bound, gpjax/objectives.py:321-416, for
 * n data points x [n, d], y [n] and m inducing inputs z [m, d], with Kzz
 * jitter `jitter` (CollapsedVariationalGaussian, variational_families.py:767-784).
 * The data enter through P = Kzx Kxz, b = Kzx (y - c) and q = Kzx 1, which are
 * accumulated on the matrix cores chunk by chunk (Kzx evaluated once per chunk
 * of 131,072 points into the workspace): no n x m matrix is stored.  result
 * slots: [GPK_RES_MLL] = ELBO, [1] = log|B|, [2] = the quadratic term, [3] = the
 * trace term, [GPK_RES_INFO] = pivot status.  d <= 32.
 * prec: GPK_PREC_FP64, or GPK_PREC_MIXED: the Kzx chunks in float32 (squared
 * distances in float64, the kernel function in float32), P = Kzx Kxz as
 * six-product bf16 splits and the backward's Kxz G_P as the fp16 split, both
 * with float64 sums; the m x m algebra stays float64.
 * GPK_PREC_MIXED_ARITH(p, t, l) names the two products instead: p for P, t for
 * the backward's, each 0, 3, 4 or 6 (l unused).
 * Mixed only: [GPK_RES_MLL_ERR] = 2^-24 ||G_P o P||_F / sqrt(2), a model of
 * the ELBO's error from the split products (G_P = dELBO / dP, measured errors
 * <= 2.9x it), and [GPK_RES_FLAG] = 1 when 5x it exceeds 1e-6 |ELBO|, the
 * factorisation broke down, or the products are ones the model is not
 * calibrated for (p other than 6 or 0, t = 3): recompute in GPK_PREC_FP64.
 * The same workspace size serves both.
 */
int gpk_elbo_forward(int kind, int prec, const double* x, const double* y, int64_t n,
                     const double* z, int64_t m, int64_t d, const double* theta, int nls,
                     double jitter, void* ws, size_t ws_bytes, double* result, void* stream);

/*
 * Gradient of the ELBO (scaled by gbar_host) w.r.t. theta (layout above) into
 * grad [nls + 3] and w.r.t. the inducing inputs into grad_z [m, d].  Must follow
 * gpk_elbo_forward on the same workspace and stream; Kzx is evaluated again,
 * chunk by chunk.
 */
int gpk_elbo_backward(int kind, int prec, const double* x, const double* y, int64_t n,
                      const double* z, int64_t m, int64_t d, const double* theta, int nls,
                      double gbar_host, void* ws, size_t ws_bytes, double* grad, double* grad_z,
                      void* stream);

/* Bytes of workspace gpk_elbo_predict needs for (n, m, d, t test points, full_cov). */
size_t gpk_elbo_predict_workspace_bytes(int64_t n, int64_t m, int64_t d, int64_t t,
                                        int full_cov);

/*
 * Predictive of the collapsed posterior at t test inputs xt [t, d]
 * (CollapsedVariationalGaussian.predict, gpjax/variational_families.py:786-870):
 *   mean = c + Kzt^T C^-1 Kzx (y - c) / s,   C = Kzz + jitter I + Kzx Kxz / s,
 *   cov  = Ktt - (Lz^-1 Kzt)^T (Lz^-1 Kzt) + (L_C^-1 Kzt)^T (L_C^-1 Kzt) + jitter I,
 * s = obs_sd^2, Lz = chol(Kzz + jitter I), L_C = chol(C) = Lz chol(I + A A^T).
 * Runs the ELBO forward first (same statistics, no n x m matrix).  mean [t];
 * cov [t, t] (full_cov) or its diagonal [t].  Float64 throughout.
 */
int gpk_elbo_predict(int kind, int prec, const double* x, const double* y, int64_t n,
                     const double* z, int64_t m, int64_t d, const double* xt, int64_t t,
                     const double* theta, int nls, double jitter, int full_cov, void* ws,
                     size_t ws_bytes, double* mean, double* cov, void* stream);

/* Bytes of workspace gpk_predict needs for (n, d, t test points, prec, full_cov). */
size_t gpk_predict_workspace_bytes(int64_t n, int64_t d, int64_t t, int prec, int full_cov);

/*
 * Conjugate posterior predictive at t test inputs xt [t, d]
 * (ConjugatePosterior.predict, gpjax/gps.py:443-526):
 *   mean = c + Kxt^T Sigma^-1 (y - c),  Sigma = K + (jitter_post + obs_sd^2) I,
 *   cov  = Ktt - (L^-1 Kxt)^T (L^-1 Kxt) + jitter_prior I.
 * full_cov != 0 writes the dense t x t covariance (row-major, ld = t) to cov;
 * full_cov == 0 writes only its diagonal (t entries).  The factorisation runs
 * in `prec` (plain GPK_PREC_MIXED takes six-product splits for the triangular
 * inverse here too: X ~ L^-1 enters the variance directly); in the mixed
 * path's diagonal mode L^-1 Kxt is a six-product bf16-split product of float32
 * operands with float64 sums; everything else runs in float64.  result
 * (optional, device, GPK_RESULT_LEN): the factorisation's slots as
 * gpk_mll_forward writes them (pivot status, refinement residual, the mixed
 * path's accuracy check).
 */
int gpk_predict(int kind, int prec, const double* x, const double* y, int64_t n, int64_t d,
                const double* xt, int64_t t, const double* theta, int nls, double jitter_post,
                double jitter_prior, int full_cov, int refine_iters, void* ws, size_t ws_bytes,
                double* mean, double* cov, double* result, void* stream);

/* Bytes of workspace gpk_loocv_forward / gpk_loocv_backward need. */
size_t gpk_loocv_workspace_bytes(int64_t n, int64_t d, int prec);

/*
 * Leave-one-out log predictive probability, gpjax/objectives.py:110-178:
 *   sum_i log N(y_i; y_i - alpha_i / d_i, 1 / d_i),
 * alpha = Sigma^-1 (y - c), d = diag(Sigma^-1), Sigma = K + (jitter + obs_sd^2) I
 * (the reference inverts Sigma densely, objectives.py:171; here d is the column
 * sums of squares of L^-1).  Same factorisation, refinement and precisions as
 * gpk_mll_forward.  result slots: [GPK_RES_MLL] = LOOCV, [1] = sum log d_i,
 * [2] = sum alpha_i^2 / d_i, [GPK_RES_RESID] = relative residual of alpha,
 * [GPK_RES_INFO] = pivot status.
 */
int gpk_loocv_forward(int kind, int prec, const double* x, const double* y, int64_t n, int64_t d,
                      const double* theta, int nls, double jitter, int refine_iters, void* ws,
                      size_t ws_bytes, double* result, void* stream);

/*
 * Gradient of the LOOCV objective (scaled by gbar_host) w.r.t. theta into grad
 * [nls + 3] (jax.grad of conjugate_loocv in the reference,
 * tests/test_objectives.py:123-126).  Must follow gpk_loocv_forward on the same
 * workspace and stream: forms Sigma^-1 = L^-T L^-1 in the buffer that held L and
 * reduces W = alpha w^T + w alpha^T - Sigma^-1 diag(c^2) Sigma^-1 against
 * dSigma/dtheta (w = Sigma^-1 (alpha / d), c^2 = 1/d + alpha^2/d^2).  Like
 * gpk_mll_backward, the workspace is an in/out buffer (L is overwritten; L^-1,
 * alpha and d stay, so the backward may be repeated).
 */
int gpk_loocv_backward(int kind, int prec, const double* x, int64_t n, int64_t d,
                       const double* theta, int nls, double gbar_host, void* ws, size_t ws_bytes,
                       double* grad, void* stream);

/*
 * Many independent exact GPs in one call: `count` problems sharing the kernel
 * kind, the input dimension d and the lengthscale layout nls, each with its own
 * rows, targets and hyperparameters.  Each problem's MLL and gradient are those
 * of gpk_mll_forward / gpk_mll_backward: conjugate_mll and its jax.grad
 * (gpjax/objectives.py:36-107; tests/test_objectives.py:83-85) applied per
 * problem, as a vmap over problems would.
 *
 * prec: GPK_PREC_FP64, or GPK_PREC_MIXED (float32 storage, bf16-split products,
 *   float64 diagonal blocks / refinement / log-determinant, and each problem's
 *   own a-posteriori probe check in result[GPK_RES_PROBE..GPK_RES_FLAG]), or
 *   GPK_PREC_MIXED_ARITH(p, t, l) with bf16 stages only (the fp16 split scales
 *   its operands by one problem's hyperparameters, so it is refused here).
 *   The workspace, the plan and the call must name the same prec.
 *
 * offsets_host: HOST array of count + 1 row offsets (offsets[0] = 0): problem p
 *   owns rows offsets[p] .. offsets[p+1]-1 (at least one) of the concatenated
 *   x [offsets[count], d] and y [offsets[count]] (device).
 * theta: device [count][nls + 3], each row in gpk_mll_forward's layout.
 * Problems are grouped by padded size; one group runs as one set of launches
 * with one problem per blockIdx.z.
 */
size_t gpk_batched_workspace_bytes(const int64_t* offsets_host, int64_t count, int64_t d,
                                   int prec);

/*
 * Writes the batch's problem map into ws (synchronous host-to-device copies).
 * Call once for a given (offsets, ws) before gpk_batched_mll_grad, outside any
 * HIP-graph capture.  The library remembers, per workspace address, which
 * (offsets, count, d, prec) it planned (a mutex-guarded host table), and
 * gpk_batched_mll_grad refuses a workspace planned for anything else.
 */
int gpk_batched_plan(const int64_t* offsets_host, int64_t count, int64_t d, int prec, void* ws,
                     size_t ws_bytes);

/*
 * MLL diagnostics and gradient of every problem: result [count][GPK_RESULT_LEN]
 * (slots as gpk_mll_forward), grad [count][nls + 3] (d MLL / d theta in theta's
 * layout), both device.  Launches only (stream-ordered, no host sync, no
 * allocation), so it may be captured in a HIP graph once gpk_batched_plan has
 * run on ws.  refine_iters < 0 selects the default (1 for FP64, 6 for mixed;
 * each problem stops at 1e-12 as in gpk_mll_forward).
 */
int gpk_batched_mll_grad(int kind, int prec, const double* x, const double* y,
                         const int64_t* offsets_host, int64_t count, int64_t d,
                         const double* theta, int nls, double jitter, int refine_iters, void* ws,
                         size_t ws_bytes, double* result, double* grad, void* stream);

/*
 * Autotuner of the batched path.  For every size group of the batch it times
 * the four launch variants of one evaluation -- the refinement's Sigma products
 * from a stored float64 copy of Sigma or recomputed from the inputs, and the
 * fused Sigma^-1 pass reading X^T along k (after one transpose) or X down its
 * columns; all four compute the same quantities in float64-exact order-free
 * ways -- on these very inputs (one untimed run, then three timed), and keeps
 * the fastest in a process-wide table keyed by (device architecture, kernel
 * kind, prec, padded size, ceil(log2 members), d), guarded by a mutex.
 * gpk_batched_mll_grad then runs each group's entry; untuned groups run the
 * recomputed / column-read variant.  Host-synchronous (not for graph capture);
 * same arguments and outputs as gpk_batched_mll_grad, whose result and grad it
 * leaves as the chosen variants compute them.  report_host (optional): 7 floats
 * per group, up to report_len floats -- padded size, members, chosen variant
 * (bit 0: stored Sigma, bit 1: transposed pass), then the four variants' mean
 * milliseconds.  No counterpart in the reference (XLA compiles one program per
 * shape).
 */
int gpk_batched_tune(int kind, int prec, const double* x, const double* y,
                     const int64_t* offsets_host, int64_t count, int64_t d, const double* theta,
                     int nls, double jitter, int refine_iters, void* ws, size_t ws_bytes,
                     double* result, double* grad, float* report_host, int report_len,
                     void* stream);

/*
 * Dense linear algebra behind gpjax/linalg/operations.py, on the same
 * recursive Cholesky / triangular-inverse kernels as the objectives.  float64
 * row-major matrices with a leading dimension; workspace from
 * gpk_linalg_workspace_bytes(n, m, prec) (m = right-hand-side columns of
 * gpk_trsm, else 0), 256-byte aligned.
 */
size_t gpk_linalg_workspace_bytes(int64_t n, int64_t m, int prec);

/*
 * L = chol(A), lower with zeros above (lower_cholesky of a Dense operator,
 * gpjax/linalg/operations.py:54-55, i.e. jnp.linalg.cholesky with its
 * symmetrized input (A + A^T)/2).  logdet (device double, optional) =
 * log|A| = 2 sum log L_ii, the Dense logdet for a PD A (operations.py:163-165);
 * info (device int, optional) = 0, or 1 + the first non-positive pivot, in
 * which case logdet is NaN and L is NaN from that pivot on.  prec
 * GPK_PREC_MIXED factors in float32 storage with BF16x6 products and float64
 * diagonal blocks (L to about float32 accuracy).
 */
int gpk_potrf(int prec, const double* A, int64_t n, int64_t lda, double* L, int64_t ldl,
              double* logdet, int* info, void* ws, size_t ws_bytes, void* stream);

/*
 * B <- L^-1 B for a lower-triangular L (n x n) and B (n x m): solve of a
 * Triangular operator, jax.scipy.linalg.solve_triangular(L, b, lower=True)
 * (gpjax/linalg/operations.py:105-107).  Computed as L^-1 (triangular inverse
 * on the matrix cores, float64) times B.
 */
int gpk_trsm(const double* L, int64_t n, int64_t ldl, double* B, int64_t m, int64_t ldb,
             void* ws, size_t ws_bytes, void* stream);

/*
 * Ainv = (L L^T)^-1 = L^-T L^-1, full symmetric (n x n): the dense inverse of
 * a PD matrix from its factor, which GPJax forms with jnp.linalg.inv
 * (gpjax/objectives.py:171) or solve(A, I).
 */
int gpk_potri(const double* L, int64_t n, int64_t ldl, double* Ainv, int64_t lda, void* ws,
              size_t ws_bytes, void* stream);

/* out (device double) = sum_i log L_ii: logdet of a Triangular operator
 * (gpjax/linalg/operations.py:145-147). */
int gpk_logdet_tri(const double* L, int64_t n, int64_t ldl, double* out, void* stream);

/* Bytes of workspace gpk_mvn_log_prob needs (diag != 0: a diagonal covariance). */
size_t gpk_mvn_workspace_bytes(int64_t n, int diag);

/*
 * GaussianDistribution.log_prob and entropy (gpjax/distributions.py:115-134,
 * 82-86): out[0] = -1/2 [n log 2pi + log|S| + (y - loc)^T S^-1 (y - loc)],
 * out[1] = log|S|, out[2] = the quadratic term, out[3] = the entropy
 * 1/2 [n (1 + log 2pi) + log|S|].  S is dense n x n (lds) or, with diag != 0,
 * the n variances of a diagonal covariance.  The dense path factors S by
 * Cholesky (a NaN result if S is not PD, where the reference's LU slogdet
 * drops the sign).
 */
int gpk_mvn_log_prob(const double* loc, const double* S, int64_t n, int64_t lds, int diag,
                     const double* y, double* out, void* ws, size_t ws_bytes, void* stream);

/*
 * One optimiser step of gpjax.fit (gpjax/fit.py:133-180) for `count` float64
 * coordinates, in one launch: with f = softplus where softplus_mask[i] != 0
 * (DEFAULT_BIJECTION "positive" / "non_negative", gpjax/parameters.py:140-146)
 * and the identity elsewhere, and grad the loss gradient w.r.t. the
 * constrained values f(raw):
 *   g = grad * f'(raw);  m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2;  step += 1;
 *   raw -= lr (m / (1 - b1^step)) / (sqrt(v / (1 - b2^step)) + eps);
 *   value = f(raw);  history[step - 1] = *loss  (when history is non-null and
 *   step <= history_len).
 * This is optax.adam (fit.py:161-162) with eps_root = 0.  All pointers are
 * device pointers; `step` is one float64 (0 before the first step).  Launch
 * only, so a whole training step may be captured in a HIP graph.
 */
int gpk_adam_step(int64_t count, const double* grad, const uint8_t* softplus_mask, double* raw,
                  double* m, double* v, double* step, double lr, double b1, double b2, double eps,
                  double* value, const double* loss, double* history, int64_t history_len,
                  void* stream);

/*
 * Phase profiling (no counterpart in the reference).  While enabled, each
 * phase of gpk_mll_forward / gpk_mll_backward is bracketed by HIP events
 * recorded on the call's stream.  gpk_profile_collect synchronises on those
 * events, writes the summed milliseconds and call counts per phase into
 * host arrays of GPK_PHASE_COUNT entries, and clears the record.
 */
#define GPK_PHASE_GRAM 0    /* Sigma assembly                                  */
#define GPK_PHASE_POTRF 1   /* recursive Cholesky (leaves, TRSM, SYRK)          */
#define GPK_PHASE_TRTRI 2   /* recursive triangular inverse                    */
#define GPK_PHASE_SOLVE 3   /* alpha, float64 refinement, log-det, objective    */
#define GPK_PHASE_ADJOINT 4 /* fused Sigma^-1 + gradient pass (backward)        */
#define GPK_PHASE_COUNT 5
int gpk_profile_enable(int on);
int gpk_profile_collect(double* ms_host, int* counts_host);

/* Self-test of the matrix-core lane layouts (writes 64*4 doubles + 64*4 floats). */
int gpk_mfma_selftest(double* out_f64, float* out_bf16, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GPK_H */
