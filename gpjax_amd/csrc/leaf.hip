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
// Thread t of 1024 owns the 4 x 4 block rows 4*(t%32).., cols 4*(t/32)..,
// so one wave holds two whole column blocks: at step k only the waves holding
// a column >= k have work, and the others go straight to the barrier.  Column
// k+1 is published to a double-buffered LDS vector by its owners right after
// their update, together with sqrt and reciprocal of the new pivot (computed
// once, by the owner of the diagonal entry), so each step costs one barrier.
// The update is one FMA per entry, the masks folded into its factors.
//
// Outputs: L (lower, zero strict upper) into A, X = L^-1 (lower, zero strict
// upper) into X, X in float64 into leafinv, 2*sum(log L_kk) into *logdet and
// 1 + global index of the first non-positive pivot (or 0) into *info.  A
// non-positive pivot turns the rest of the block into NaN, which propagates
// to the objective.
#include "gpk_internal.h"

namespace gpk {
namespace {

// 1024 threads, thread t owning the EB x EB block at rows EB*(t%32), cols
// EB*(t/32): four waves per SIMD hide the per-column LDS and barrier latency.
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
  __shared__ int bad_s;
  __shared__ double tr_s[NB][NB + 1];
  A = bws(A, bt);
  X = bws(X, bt);
  leafinv = bws(leafinv, bt);
  logdet = bws(logdet, bt);
  info = bws(info, bt);
  dsrc = bws(dsrc, bt);

  const int t = threadIdx.x;
  const int rb = (t % BPR) * EB;
  const int cb = (t / BPR) * EB;
  // Stage the block through LDS so that global reads stay row-contiguous.
  for (int e = t; e < NB * NB; e += LEAF_THREADS) {
    const int i = e / NB, j = e % NB;
    tr_s[i][j] = dsrc ? dsrc[e] : (double)A[(o + i) * ld + o + j];
  }
  if (t == 0) bad_s = NB;
  __syncthreads();
  double s[EB][EB];
#pragma unroll
  for (int a = 0; a < EB; ++a)
#pragma unroll
    for (int b = 0; b < EB; ++b) s[a][b] = (cb + b <= rb + a) ? tr_s[rb + a][cb + b] : 0.0;
  // pivot 0 and column 0
  if (cb == 0) {
#pragma unroll
    for (int a = 0; a < EB; ++a) col_s[0][rb + a] = s[a][0];
    if (rb == 0) {
      const double piv = s[0][0];
      const double dk = sqrt(piv);
      dk_s[0] = dk;
      dinv_s[0] = 1.0 / dk;
      if (!(piv > 0.0)) bad_s = 0;
    }
  }
  __syncthreads();

  for (int k = 0; k < NB; ++k) {
    // Every update at step k touches column k or a column > k; the test is
    // uniform over each half-wave (one column block).
    if (cb + EB - 1 >= k) {
      const double* col = col_s[k & 1];
      const double dk = dk_s[k];
      const double inv = dinv_s[k];
      // Entry (i, j) changes at step k only if j > k and either i >= j (lower
      // trailing update L[i][k] L[j][k]) or i <= k (strict upper: numerator of
      // X[j][i] -= L[j][k] X[k][i]).  The j > k mask goes into the column
      // factors, the i <= k mask into the row factors, and whether i >= j is
      // fixed per thread: always (rb > cb), never (rb < cb), or a >= b on the
      // diagonal block, which is known at compile time.  So every entry takes
      // one FMA, with a zero factor where it must not change.
      double sci[EB], alo[EB], ahi[EB], scc[EB];
#pragma unroll
      for (int a = 0; a < EB; ++a) {
        const double v = col[rb + a] * inv;
        sci[a] = (rb + a == k) ? inv : v;
        const double sx = (rb + a <= k) ? sci[a] : 0.0;
        alo[a] = (rb >= cb) ? sci[a] : sx;  // entries with a >= b
        ahi[a] = (rb > cb) ? sci[a] : sx;   // entries with a < b
      }
#pragma unroll
      for (int b = 0; b < EB; ++b) scc[b] = (cb + b > k) ? col[cb + b] * inv : 0.0;
#pragma unroll
      for (int a = 0; a < EB; ++a)
#pragma unroll
        for (int b = 0; b < EB; ++b)
          s[a][b] = fma(-(a >= b ? alo[a] : ahi[a]), scc[b], s[a][b]);
      // Column k itself: L[i][k] below the diagonal, the pivot on it, X[k][i]
      // (stored at (i, k)) above; only the half-wave holding column k.
      if (k >= cb) {
#pragma unroll
        for (int b = 0; b < EB; ++b)
          if (cb + b == k) {
#pragma unroll
            for (int a = 0; a < EB; ++a) s[a][b] = (rb + a == k) ? dk : sci[a];
          }
      }
      // Publish column k+1 and, from the owner of its diagonal entry, its pivot.
      const int q = k + 1 - cb;
      if (k + 1 < NB && q >= 0 && q < EB) {
        double v[EB];
#pragma unroll
        for (int a = 0; a < EB; ++a) {
          v[a] = s[a][0];
#pragma unroll
          for (int b = 1; b < EB; ++b) v[a] = (q == b) ? s[a][b] : v[a];
          col_s[(k + 1) & 1][rb + a] = v[a];
        }
        const int p = k + 1 - rb;
        if (p >= 0 && p < EB) {
          double piv = v[0];
#pragma unroll
          for (int a = 1; a < EB; ++a) piv = (p == a) ? v[a] : piv;
          const double dkn = sqrt(piv);
          dk_s[k + 1] = dkn;
          dinv_s[k + 1] = 1.0 / dkn;
          if (!(piv > 0.0) && k + 1 < bad_s) bad_s = k + 1;
        }
      }
    }
    __syncthreads();
  }

  // L (lower incl. diagonal) and X (from the upper part, transposed) staged
  // in LDS for row-contiguous writes.
#pragma unroll
  for (int a = 0; a < EB; ++a)
#pragma unroll
    for (int b = 0; b < EB; ++b) tr_s[rb + a][cb + b] = s[a][b];
  if (t < NB) dk_s[t] = log(dk_s[t]);
  __syncthreads();
  for (int e = t; e < NB * NB; e += LEAF_THREADS) {
    const int i = e / NB, j = e % NB;
    A[(o + i) * ld + o + j] = (TS)((j <= i) ? tr_s[i][j] : 0.0);
    const double v = (j < i) ? tr_s[j][i] : (j == i ? dinv_s[i] : 0.0);
    X[(o + i) * ld + o + j] = (TS)v;
    leafinv[e] = v;
  }
  if (t == 0) {
    double ld_acc = 0.0;  // in column order, as the sweep produced the pivots
    for (int k = 0; k < NB; ++k) ld_acc += dk_s[k];
    *logdet = 2.0 * ld_acc;
    *info = bad_s < NB ? (int)(o + bad_s + 1) : 0;
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
