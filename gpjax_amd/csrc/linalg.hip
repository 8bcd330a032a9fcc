// Dense linear-algebra entry points behind gpjax/linalg/operations.py, so that
// GPJax's own call sites of lower_cholesky / solve / logdet / inverse can run
// on the same recursive Cholesky and triangular-inverse machinery as the
// objectives (factor.h):
//   gpk_potrf      lower_cholesky(Dense A)          operations.py:54-55
//                  logdet(Dense A) for a PD A       operations.py:163-165
//   gpk_trsm       solve(Triangular L, b), lower    operations.py:105-107
//   gpk_potri      inverse of L L^T (the dense inverse GPJax forms with
//                  jnp.linalg.inv, objectives.py:171)
//   gpk_logdet_tri logdet(Triangular L)             operations.py:145-147
// Matrices are float64, row-major with a leading dimension, as given; the
// workspace holds the padded working copies (n rounded up to 256 with an
// identity tail, as in the objectives).
#include "factor.h"
#include "gemm_core.h"

namespace gpk {
namespace {

inline size_t al256(size_t b) { return (b + 255) / 256 * 256; }

struct LinLayout {
  int64_t np = 0, mp = 0;
  size_t ts = sizeof(double);
  size_t off_L = 0, off_X = 0, off_T = 0, off_leafinv = 0, off_leafld = 0, off_leafinfo = 0;
  size_t off_dblk = 0, off_B = 0, off_C = 0, off_r = 0, off_u = 0, off_part = 0, off_scal = 0;
  size_t total = 0;
};

LinLayout make_lin_layout(int64_t n, int64_t m, int prec) {
  LinLayout P;
  P.np = round_up(n, NB);
  P.mp = m > 0 ? round_up(m, NB) : 0;
  P.ts = prec == GPK_PREC_FP64 ? sizeof(double) : sizeof(float);
  const int64_t np = P.np, nleaf = np / NB;
  size_t o = 0;
  P.off_L = o; o += al256((size_t)np * np * P.ts);
  P.off_X = o; o += al256((size_t)np * np * P.ts);
  P.off_T = o; o += al256((size_t)factor_scratch_elems(np) * P.ts);
  P.off_leafinv = o; o += al256((size_t)np * NB * sizeof(double));
  P.off_leafld = o; o += al256((size_t)nleaf * sizeof(double));
  P.off_leafinfo = o; o += al256((size_t)nleaf * sizeof(int));
  P.off_dblk = o; o += al256(prec == GPK_PREC_FP64 ? 0 : (size_t)np * NB * sizeof(double));
  P.off_B = o; o += al256((size_t)np * P.mp * sizeof(double));
  P.off_C = o; o += al256((size_t)np * P.mp * sizeof(double));
  P.off_r = o; o += al256((size_t)np * sizeof(double));
  P.off_u = o; o += al256((size_t)np * sizeof(double));
  P.off_part = o; o += al256((size_t)((np + 1023) / 1024) * np * sizeof(double));
  P.off_scal = o; o += al256(4 * sizeof(double));
  P.total = o;
  return P;
}

// Lower NB-tiles of the padded np x np working matrix from A (n x n, lda):
// entry (i, j), j <= i < n, is (A_ij + A_ji) / 2 (jnp.linalg.cholesky's
// symmetrize_input), the tail is the identity; entries above the diagonal
// inside diagonal tiles are zero.  With `tri` the strict upper triangle of A
// is not read (a triangular operand: A_ij as given).  dblk (optional): the
// diagonal tiles again in float64, [tile][NB][NB].
template <typename TS>
__global__ __launch_bounds__(256) void load_lower_kernel(const double* __restrict__ A,
                                                         int64_t lda, int64_t n, int tri,
                                                         TS* __restrict__ L, int64_t ld,
                                                         double* __restrict__ dblk) {
  int64_t I, J;
  lower_tile(blockIdx.x, I, J);
  for (int e = threadIdx.x; e < NB * NB; e += 256) {
    const int r = e / NB, c = e % NB;
    const int64_t i = I * NB + r, j = J * NB + c;
    double v;
    if (i < n && j < n) {
      if (j > i) v = 0.0;
      else if (j == i || tri) v = A[i * lda + j];
      else v = 0.5 * (A[i * lda + j] + A[j * lda + i]);
    } else {
      v = (i == j) ? 1.0 : 0.0;
    }
    L[i * ld + j] = (TS)v;
    if (dblk && I == J) dblk[I * NB * NB + e] = v;
  }
}

// out (n x n, ldo) = lower triangle of the working matrix (zeros above), or
// with `sym` its symmetric completion from the lower tiles.
template <typename TS>
__global__ __launch_bounds__(256) void store_kernel(const TS* __restrict__ L, int64_t ld,
                                                    int64_t n, int sym, double* __restrict__ out,
                                                    int64_t ldo) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  for (int64_t i = blockIdx.y; i < n; i += gridDim.y) {
    double v;
    if (j <= i) v = (double)L[i * ld + j];
    else v = sym ? (double)L[j * ld + i] : 0.0;
    out[i * ldo + j] = v;
  }
}

__global__ void potrf_finish_kernel(const double* __restrict__ leafld,
                                    const int* __restrict__ leafinfo, int64_t nleaf,
                                    double* __restrict__ logdet, int* __restrict__ info) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  int bad = 0;
  for (int64_t b = 0; b < nleaf; ++b) {
    s += leafld[b];
    if (bad == 0 && leafinfo[b] != 0) bad = leafinfo[b];
  }
  if (logdet) *logdet = bad ? __builtin_nan("") : s;
  if (info) *info = bad;
}

// Inverse of one NB x NB lower-triangular diagonal block of the working L, in
// float64, by right-looking forward substitution on L X = I: thread j owns
// column j of X (held in LDS); at step k it finishes X_kj = X_kj / L_kk and
// removes L_ik X_kj from the rows below.  L_ik is the same for every thread
// (a broadcast read).  Writes X's diagonal block and the float64 leaf inverse.
__global__ __launch_bounds__(NB) void leaf_trinv_kernel(const double* __restrict__ L, int64_t ld,
                                                        double* __restrict__ X,
                                                        double* __restrict__ leafinv) {
  __shared__ double xs[NB][NB];
  const int64_t o = (int64_t)blockIdx.x * NB;
  const int j = threadIdx.x;
  for (int i = 0; i < NB; ++i) xs[i][j] = (i == j) ? 1.0 : 0.0;
  const double* Lb = L + o * ld + o;
  for (int k = 0; k < NB; ++k) {
    if (j <= k) {
      const double x = xs[k][j] / Lb[k * ld + k];
      xs[k][j] = x;
      for (int i = k + 1; i < NB; ++i) xs[i][j] -= Lb[i * ld + k] * x;
    }
  }
  for (int i = 0; i < NB; ++i) {
    X[(o + i) * ld + o + j] = xs[i][j];
    leafinv[(int64_t)blockIdx.x * NB * NB + i * NB + j] = xs[i][j];
  }
}

// Bp (np x mp, zero padded) <- B (n x m, ldb); and back.
__global__ __launch_bounds__(256) void pad_rhs_kernel(const double* __restrict__ B, int64_t ldb,
                                                      int64_t n, int64_t m, int64_t np,
                                                      double* __restrict__ Bp, int64_t mp) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= mp) return;
  for (int64_t i = blockIdx.y; i < np; i += gridDim.y)
    Bp[i * mp + j] = (i < n && j < m) ? B[i * ldb + j] : 0.0;
}

__global__ __launch_bounds__(256) void unpad_rhs_kernel(const double* __restrict__ Cp,
                                                        int64_t mp, int64_t n, int64_t m,
                                                        double* __restrict__ B, int64_t ldb) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= m) return;
  for (int64_t i = blockIdx.y; i < n; i += gridDim.y) B[i * ldb + j] = Cp[i * mp + j];
}

// Rows of a row-looped 2-D launch: at most 4,096 workgroups along y.
inline unsigned rows_grid(int64_t rows) { return (unsigned)(rows < 4096 ? rows : 4096); }

__global__ __launch_bounds__(256) void logdet_tri_kernel(const double* __restrict__ L,
                                                         int64_t ld, int64_t n,
                                                         double* __restrict__ out) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += log(L[i * ld + i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = (red[0] + red[1]) + (red[2] + red[3]);
}

// r = y - loc on the padded vector (zeros on the pad).
__global__ void mvn_resid_kernel(const double* __restrict__ y, const double* __restrict__ loc,
                                 int64_t n, int64_t np, double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) r[i] = i < n ? y[i] - loc[i] : 0.0;
}

// out = [log_prob, log|S|, r^T S^-1 r, entropy]: dense S from the factor's
// log-determinant and u = L^-1 r; diagonal S (s = variances) directly.
__global__ __launch_bounds__(256) void mvn_finish_kernel(const double* __restrict__ u,
                                                         const double* __restrict__ s_diag,
                                                         const double* __restrict__ logdet_in,
                                                         int64_t n, double* __restrict__ out) {
  __shared__ double red[2][4];
  double q = 0.0, ld = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    if (s_diag) {
      q += u[i] * u[i] / s_diag[i];
      ld += log(s_diag[i]);
    } else {
      q += u[i] * u[i];
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    q += __shfl_xor(q, off);
    ld += __shfl_xor(ld, off);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][w] = q;
    red[1][w] = ld;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    q = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    ld = s_diag ? (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]) : *logdet_in;
    const double log2pi = 1.8378770664093453;
    out[0] = -0.5 * ((double)n * log2pi + ld + q);
    out[1] = ld;
    out[2] = q;
    out[3] = 0.5 * ((double)n * (1.0 + log2pi) + ld);
  }
}

int check_mat(const char* fn, const void* A, int64_t n, int64_t lda) {
  if (!A || n < 1 || lda < n) {
    set_error("%s: need a matrix with n >= 1 and leading dimension >= n", fn);
    return GPK_ERR_BAD_ARG;
  }
  if (n > (int64_t)65535 * NB) {
    set_error("%s: n = %lld too large", fn, (long long)n);
    return GPK_ERR_BAD_ARG;
  }
  return GPK_OK;
}

int check_lin_ws(const char* fn, const void* ws, size_t have, size_t need) {
  if (!ws || have < need) {
    set_error("%s: workspace %zu bytes < required %zu", fn, have, need);
    return GPK_ERR_WORKSPACE;
  }
  if (reinterpret_cast<uintptr_t>(ws) % 256) {
    set_error("%s: workspace must be 256-byte aligned", fn);
    return GPK_ERR_BAD_ARG;
  }
  return GPK_OK;
}

template <typename TS>
int potrf_t(int prec, const double* A, int64_t n, int64_t lda, double* Lout, int64_t ldl,
            double* logdet, int* info, void* ws, const LinLayout& P, hipStream_t s) {
  const int64_t np = P.np, nt = np / NB;
  const bool f64 = std::is_same<TS, double>::value;
  TS* L = at<TS>(ws, P.off_L);
  double* dblk = f64 ? nullptr : at<double>(ws, P.off_dblk);
  hipLaunchKernelGGL(load_lower_kernel<TS>, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0,
                     s, A, lda, n, 0, L, np, dblk);
  GPK_LAUNCH_CHECK();
  Factor<TS> F{L,
               at<TS>(ws, P.off_X),
               np,
               at<TS>(ws, P.off_T),
               at<double>(ws, P.off_leafinv),
               at<double>(ws, P.off_leafld),
               at<int>(ws, P.off_leafinfo),
               f64 ? (int)AR_F64 : (int)AR_BF16X6,
               f64 ? (int)AR_F64 : (int)AR_BF16X3,
               s};
  F.dblk = dblk;
  (void)prec;
  GPK_TRY(F.potrf(0, np));
  hipLaunchKernelGGL(store_kernel<TS>, dim3((unsigned)((n + 255) / 256), rows_grid(n)),
                     dim3(256), 0, s, L, np, n, 0, Lout, ldl);
  GPK_LAUNCH_CHECK();
  if (logdet || info) {
    hipLaunchKernelGGL(potrf_finish_kernel, dim3(1), dim3(64), 0, s,
                       at<double>(ws, P.off_leafld), at<int>(ws, P.off_leafinfo), nt, logdet,
                       info);
    GPK_LAUNCH_CHECK();
  }
  return GPK_OK;
}

// X = L^-1 in the workspace from a caller's lower-triangular L (float64).
int inverse_factor(const double* Lin, int64_t n, int64_t ldl, void* ws, const LinLayout& P,
                   hipStream_t s) {
  const int64_t np = P.np, nt = np / NB;
  double* L = at<double>(ws, P.off_L);
  double* X = at<double>(ws, P.off_X);
  hipLaunchKernelGGL(load_lower_kernel<double>, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256),
                     0, s, Lin, ldl, n, 1, L, np, nullptr);
  GPK_LAUNCH_CHECK();
  hipLaunchKernelGGL(leaf_trinv_kernel, dim3((unsigned)nt), dim3(NB), 0, s, L, np, X,
                     at<double>(ws, P.off_leafinv));
  GPK_LAUNCH_CHECK();
  Factor<double> F{L,
                   X,
                   np,
                   at<double>(ws, P.off_T),
                   at<double>(ws, P.off_leafinv),
                   at<double>(ws, P.off_leafld),
                   at<int>(ws, P.off_leafinfo),
                   (int)AR_F64,
                   (int)AR_F64,
                   s};
  return F.trtri(0, np);
}

}  // namespace
}  // namespace gpk

using namespace gpk;

extern "C" {

size_t gpk_mvn_workspace_bytes(int64_t n, int diag) {
  if (n < 1) return 0;
  if (diag) return al256((size_t)n * sizeof(double));
  return make_lin_layout(n, 0, GPK_PREC_FP64).total;
}

size_t gpk_linalg_workspace_bytes(int64_t n, int64_t m, int prec) {
  if (n < 1 || m < 0 || (prec != GPK_PREC_FP64 && prec != GPK_PREC_MIXED)) return 0;
  return make_lin_layout(n, m, prec).total;
}

int gpk_potrf(int prec, const double* A, int64_t n, int64_t lda, double* L, int64_t ldl,
              double* logdet, int* info, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_mat("potrf", A, n, lda);
  if (rc) return rc;
  if (!L || ldl < n) {
    set_error("potrf: need an output L with leading dimension >= n");
    return GPK_ERR_BAD_ARG;
  }
  if (prec != GPK_PREC_FP64 && prec != GPK_PREC_MIXED) {
    set_error("potrf: unknown precision code %d", prec);
    return GPK_ERR_BAD_ARG;
  }
  const LinLayout P = make_lin_layout(n, 0, prec);
  if ((rc = check_lin_ws("potrf", ws, ws_bytes, P.total))) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (prec == GPK_PREC_FP64)
    return potrf_t<double>(prec, A, n, lda, L, ldl, logdet, info, ws, P, s);
  return potrf_t<float>(prec, A, n, lda, L, ldl, logdet, info, ws, P, s);
}

int gpk_trsm(const double* L, int64_t n, int64_t ldl, double* B, int64_t m, int64_t ldb,
             void* ws, size_t ws_bytes, void* stream) {
  int rc = check_mat("trsm", L, n, ldl);
  if (rc) return rc;
  if (!B || m < 1 || ldb < m) {
    set_error("trsm: need a right-hand side with m >= 1 columns and leading dimension >= m");
    return GPK_ERR_BAD_ARG;
  }
  const LinLayout P = make_lin_layout(n, m, GPK_PREC_FP64);
  if ((rc = check_lin_ws("trsm", ws, ws_bytes, P.total))) return rc;
  if (P.mp / NB > 65535) {
    set_error("trsm: m = %lld too large", (long long)m);
    return GPK_ERR_BAD_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  GPK_TRY(inverse_factor(L, n, ldl, ws, P, s));
  const int64_t np = P.np, mp = P.mp;
  double* Bp = at<double>(ws, P.off_B);
  double* Cp = at<double>(ws, P.off_C);
  hipLaunchKernelGGL(pad_rhs_kernel, dim3((unsigned)((mp + 255) / 256), rows_grid(np)),
                     dim3(256), 0, s, B, ldb, n, m, np, Bp, mp);
  GPK_LAUNCH_CHECK();
  // C = X Bp, X lower: k <= row tile
  GPK_TRY(launch_gemm<double, double, double>(AR_F64, false, true, at<double>(ws, P.off_X), np,
                                              Bp, mp, Cp, mp, np, mp, np, 1.0, false, false,
                                              KR_LE_M, s));
  hipLaunchKernelGGL(unpad_rhs_kernel, dim3((unsigned)((m + 255) / 256), rows_grid(n)),
                     dim3(256), 0, s, Cp, mp, n, m, B, ldb);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

int gpk_potri(const double* L, int64_t n, int64_t ldl, double* Ainv, int64_t lda, void* ws,
              size_t ws_bytes, void* stream) {
  int rc = check_mat("potri", L, n, ldl);
  if (rc) return rc;
  if (!Ainv || lda < n) {
    set_error("potri: need an output with leading dimension >= n");
    return GPK_ERR_BAD_ARG;
  }
  const LinLayout P = make_lin_layout(n, 0, GPK_PREC_FP64);
  if ((rc = check_lin_ws("potri", ws, ws_bytes, P.total))) return rc;
  hipStream_t s = (hipStream_t)stream;
  GPK_TRY(inverse_factor(L, n, ldl, ws, P, s));
  const int64_t np = P.np;
  double* S = at<double>(ws, P.off_L);
  const double* X = at<double>(ws, P.off_X);
  // S = X^T X over the lower tiles (X lower: k >= row tile)
  GPK_TRY(launch_gemm<double, double, double>(AR_F64, true, true, X, np, X, np, S, np, np, np,
                                              np, 1.0, false, true, KR_GE_M, s));
  hipLaunchKernelGGL(store_kernel<double>, dim3((unsigned)((n + 255) / 256), rows_grid(n)),
                     dim3(256), 0, s, S, np, n, 1, Ainv, lda);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

int gpk_mvn_log_prob(const double* loc, const double* S, int64_t n, int64_t lds, int diag,
                     const double* y, double* out, void* ws, size_t ws_bytes, void* stream) {
  if (!loc || !S || !y || !out || n < 1) {
    set_error("mvn_log_prob: null loc/S/y/out or n < 1");
    return GPK_ERR_BAD_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  if (diag) {
    int rc = check_lin_ws("mvn_log_prob", ws, ws_bytes, al256((size_t)n * sizeof(double)));
    if (rc) return rc;
    hipLaunchKernelGGL(mvn_resid_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, y,
                       loc, n, n, static_cast<double*>(ws));
    GPK_LAUNCH_CHECK();
    hipLaunchKernelGGL(mvn_finish_kernel, dim3(1), dim3(256), 0, s,
                       static_cast<const double*>(ws), S, nullptr, n, out);
    GPK_LAUNCH_CHECK();
    return GPK_OK;
  }
  int rc = check_mat("mvn_log_prob", S, n, lds);
  if (rc) return rc;
  const LinLayout P = make_lin_layout(n, 0, GPK_PREC_FP64);
  if ((rc = check_lin_ws("mvn_log_prob", ws, ws_bytes, P.total))) return rc;
  const int64_t np = P.np, nt = np / NB;
  double* L = at<double>(ws, P.off_L);
  double* X = at<double>(ws, P.off_X);
  double* r = at<double>(ws, P.off_r);
  double* u = at<double>(ws, P.off_u);
  double* ldv = at<double>(ws, P.off_scal);
  hipLaunchKernelGGL(load_lower_kernel<double>, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256),
                     0, s, S, lds, n, 0, L, np, nullptr);
  GPK_LAUNCH_CHECK();
  Factor<double> F{L,
                   X,
                   np,
                   at<double>(ws, P.off_T),
                   at<double>(ws, P.off_leafinv),
                   at<double>(ws, P.off_leafld),
                   at<int>(ws, P.off_leafinfo),
                   (int)AR_F64,
                   (int)AR_F64,
                   s};
  GPK_TRY(F.potrf(0, np));
  GPK_TRY(F.trtri(0, np));
  hipLaunchKernelGGL(potrf_finish_kernel, dim3(1), dim3(64), 0, s, at<double>(ws, P.off_leafld),
                     at<int>(ws, P.off_leafinfo), nt, ldv, nullptr);
  GPK_LAUNCH_CHECK();
  hipLaunchKernelGGL(mvn_resid_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, y,
                     loc, n, np, r);
  GPK_LAUNCH_CHECK();
  GPK_TRY(launch_trmv<double>(X, np, np, r, u, false, at<double>(ws, P.off_part), s));
  hipLaunchKernelGGL(mvn_finish_kernel, dim3(1), dim3(256), 0, s, u, nullptr, ldv, n, out);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

int gpk_logdet_tri(const double* L, int64_t n, int64_t ldl, double* out, void* stream) {
  int rc = check_mat("logdet_tri", L, n, ldl);
  if (rc) return rc;
  if (!out) {
    set_error("logdet_tri: null output");
    return GPK_ERR_BAD_ARG;
  }
  hipLaunchKernelGGL(logdet_tri_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, L, ldl, n,
                     out);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

}  // extern "C"
