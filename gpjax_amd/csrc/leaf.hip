// Leaf of the recursive Cholesky: factor one NB x NB diagonal block in float64
// and form its inverse in the same sweep.
//
// Replaces the diagonal work of jnp.linalg.cholesky / slogdet
// (gpjax/linalg/operations.py:54-55,163-165).  Algorithm: right-looking
// Cholesky by columns; alongside it the numerators of X = L^-1 are eliminated
// row by row (forward substitution of L X = I, right-looking form).  Storage is
// one NB x NB float64 array held in registers across the workgroup: the lower
// triangle holds the trailing matrix / L, the strict upper triangle holds
// X^T (entry (c, i), c < i, is the numerator of X[i][c]).  At step k column k
// of that array is exactly [X-row-k numerators ; pivot ; L-column-k
// numerators], so one broadcast of column k per step drives both updates.
//
// Thread t of 1024 owns the 4 x 4 block rows 4*(t/32).., cols 4*(t%32)..;
// column k+1 is published to a double-buffered LDS vector by its owners
// right after their update, so each step costs one barrier.
//
// Outputs: L (lower, zero strict upper) into A, X = L^-1 (lower, zero strict
// upper) into X, X in float64 into leafinv, 2*sum(log L_kk) into *logdet and
// 1 + global index of the first non-positive pivot (or 0) into *info.  A
// non-positive pivot turns the rest of the block into NaN, which propagates
// to the objective.
#include "gpk_internal.h"

namespace gpk {
namespace {

// 1024 threads, thread t owning the EB x EB block at rows EB*(t/32), cols
// EB*(t%32): four waves per SIMD hide the per-column LDS and barrier latency.
constexpr int LEAF_THREADS = 1024;
constexpr int EB = 4;
constexpr int BPR = NB / EB;

template <typename TS>
__global__ __launch_bounds__(LEAF_THREADS) void leaf_kernel(TS* A, TS* X, int64_t ld, int64_t o,
                                                            double* __restrict__ leafinv,
                                                            double* __restrict__ logdet,
                                                            int* __restrict__ info,
                                                            const double* __restrict__ dsrc,
                                                            Batch bt) {
  __shared__ double col_s[2][NB];
  __shared__ double dinv_s[NB];
  __shared__ double dk_s[NB];
  __shared__ double tr_s[NB][NB + 1];
  A = bws(A, bt);
  X = bws(X, bt);
  leafinv = bws(leafinv, bt);
  logdet = bws(logdet, bt);
  info = bws(info, bt);
  dsrc = bws(dsrc, bt);

  const int t = threadIdx.x;
  const int rb = (t / BPR) * EB;
  const int cb = (t % BPR) * EB;
  double s[EB][EB];
#pragma unroll
  for (int a = 0; a < EB; ++a) {
    const TS* row = A + (o + rb + a) * ld + o + cb;
#pragma unroll
    for (int b = 0; b < EB; ++b) {
      const int i = rb + a, j = cb + b;
      const double v = dsrc ? dsrc[i * NB + j] : (double)row[b];
      s[a][b] = (j <= i) ? v : 0.0;
    }
  }
  if (cb == 0) {
#pragma unroll
    for (int a = 0; a < EB; ++a) col_s[0][rb + a] = s[a][0];
  }
  __syncthreads();

  int bad = -1;
  for (int k = 0; k < NB; ++k) {
    const double* col = col_s[k & 1];
    const double piv = col[k];
    const double dk = sqrt(piv);
    const double inv = 1.0 / dk;
    // The logarithms are taken after the sweep, off the per-column critical path.
    if (t == 0) {
      if (bad < 0 && !(piv > 0.0)) bad = k;
      dk_s[k] = dk;
      dinv_s[k] = inv;
    }
    // Every update at step k touches column k or a column > k.
    if (cb + EB - 1 >= k) {
      double sc_c[EB];
#pragma unroll
      for (int b = 0; b < EB; ++b) {
        const int j = cb + b;
        sc_c[b] = (j == k) ? inv : col[j] * inv;
      }
#pragma unroll
      for (int a = 0; a < EB; ++a) {
        const int i = rb + a;
        const double sci = (i == k) ? inv : col[i] * inv;
#pragma unroll
        for (int b = 0; b < EB; ++b) {
          const int j = cb + b;
          if (j == k) {
            // column k: L[i][k] below the diagonal, X[k][i] (stored at (i, k)) above
            s[a][b] = (i == k) ? dk : sci;
          } else if (j > k && (i >= j || i <= k)) {
            // lower part: trailing update L[i][k] L[j][k];
            // strict upper (i, j), i <= k: numerator of X[j][i] -= L[j][k] X[k][i]
            s[a][b] -= sci * sc_c[b];
          }
        }
      }
    }
    if (k + 1 < NB && cb <= k + 1 && k + 1 < cb + EB) {
#pragma unroll
      for (int b = 0; b < EB; ++b) {
        if (cb + b == k + 1) {
#pragma unroll
          for (int a = 0; a < EB; ++a) col_s[(k + 1) & 1][rb + a] = s[a][b];
        }
      }
    }
    __syncthreads();
  }

  // L out (lower incl. diagonal), and X^T staged in LDS for a coalesced write.
#pragma unroll
  for (int a = 0; a < EB; ++a) {
    TS* row = A + (o + rb + a) * ld + o + cb;
#pragma unroll
    for (int b = 0; b < EB; ++b) {
      const int i = rb + a, j = cb + b;
      row[b] = (TS)((j <= i) ? s[a][b] : 0.0);
      if (j > i) tr_s[j][i] = s[a][b];
    }
  }
  if (t < NB) dk_s[t] = log(dk_s[t]);
  __syncthreads();
  for (int e = t; e < NB * NB; e += LEAF_THREADS) {
    const int i = e / NB, j = e % NB;
    const double v = (j < i) ? tr_s[i][j] : (j == i ? dinv_s[i] : 0.0);
    X[(o + i) * ld + o + j] = (TS)v;
    leafinv[e] = v;
  }
  if (t == 0) {
    double ld_acc = 0.0;  // in column order, as the sweep produced the pivots
    for (int k = 0; k < NB; ++k) ld_acc += dk_s[k];
    *logdet = 2.0 * ld_acc;
    *info = bad >= 0 ? (int)(o + bad + 1) : 0;
  }
}

}  // namespace

template <typename TS>
int launch_leaf(TS* A, TS* X, int64_t ld, int64_t o, double* leafinv, double* logdet, int* info,
                const double* dsrc, hipStream_t s, const Batch& bt) {
  hipLaunchKernelGGL(leaf_kernel<TS>, dim3(1, 1, bt.count), dim3(LEAF_THREADS), 0, s, A, X, ld,
                     o, leafinv, logdet, info, dsrc, bt);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

template int launch_leaf<double>(double*, double*, int64_t, int64_t, double*, double*, int*,
                                 const double*, hipStream_t, const Batch&);
template int launch_leaf<float>(float*, float*, int64_t, int64_t, double*, double*, int*,
                                const double*, hipStream_t, const Batch&);

}  // namespace gpk
