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
  // pub[buf][kk][i]: the step factors of the four columns of the last panel
  // (row factor of row i; for rows below the step's column also its column
  // factor).  Double-buffered across panels.
  __shared__ double pub[2][EB][NB];
  __shared__ double dk_s[NB];
  __shared__ double dinv_s[NB];
  __shared__ int bad_s;
  __shared__ double tr_s[NB][NB + 1];
  A = bws(A, bt);
  X = bws(X, bt);
  leafinv = bws(leafinv, bt);
  logdet = bws(logdet, bt);
  info = bws(info, bt);
  dsrc = bws(dsrc, bt);

  const int t = threadIdx.x;
  const int rb = (t % BPR) * EB;   // first row of this thread's block
  const int cb = (t / BPR) * EB;   // first column (the half-wave t / 32 owns columns cb .. cb+3)
  const int hl = t % BPR;          // lane within the half-wave: rows 4 hl .. 4 hl + 3
  const int hbase = (t & 63) & ~(BPR - 1);  // first wave lane of this half-wave
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

  for (int p = 0; p < NB / EB; ++p) {
    const int c0 = p * EB;  // the panel's columns c0 .. c0 + 3
    if (cb == c0) {
      // Factor the panel inside this half-wave.  Step k = c0 + kk: the pivot
      // and the panel's own entries of column k (rows c0 .. c0+3) sit in the
      // half-wave lane p; each is broadcast from there.  As in the unblocked
      // sweep, entry (i, j), j > k, takes one FMA with row factor
      // f_i = s_ik / d_k (1 / d_k on the diagonal) and column factor s_jk / d_k
      // when i >= j (trailing update of L) or i <= k (numerator of X[j][i]).
#pragma unroll
      for (int kk = 0; kk < EB; ++kk) {
        const int k = c0 + kk;
        double cv[EB];
#pragma unroll
        for (int a = 0; a < EB; ++a) cv[a] = __shfl(s[a][kk], hbase + p);
        const double piv = cv[kk];
        const double dk = sqrt(piv);
        const double inv = 1.0 / dk;
        double f[EB];
#pragma unroll
        for (int a = 0; a < EB; ++a) f[a] = (rb + a == k) ? inv : s[a][kk] * inv;
#pragma unroll
        for (int b = kk + 1; b < EB; ++b) {
          const double g = cv[b] * inv;  // column factor of column c0 + b
#pragma unroll
          for (int a = 0; a < EB; ++a) {
            const int i = rb + a, j = c0 + b;
            const double fa = (i >= j || i <= k) ? f[a] : 0.0;
            s[a][b] = fma(-fa, g, s[a][b]);
          }
        }
#pragma unroll
        for (int a = 0; a < EB; ++a) s[a][kk] = (rb + a == k) ? dk : f[a];
#pragma unroll
        for (int a = 0; a < EB; ++a) pub[p & 1][kk][rb + a] = f[a];
        if (hl == p) {
          dk_s[k] = dk;
          dinv_s[k] = inv;
          if (!(piv > 0.0) && k < bad_s) bad_s = k;
        }
      }
    }
    __syncthreads();
    if (cb > c0) {
      // Rank-4 update of the columns right of the panel, steps in order.
      const double* P = &pub[p & 1][0][0];
#pragma unroll
      for (int kk = 0; kk < EB; ++kk) {
        const int k = c0 + kk;
        double fr[EB], g[EB];
#pragma unroll
        for (int a = 0; a < EB; ++a) fr[a] = P[kk * NB + rb + a];
#pragma unroll
        for (int b = 0; b < EB; ++b) g[b] = P[kk * NB + cb + b];
#pragma unroll
        for (int a = 0; a < EB; ++a) {
          const int i = rb + a;
          // j = cb + b > k always here; the mask is i >= j or i <= k
          const double flo = fr[a];
          const double fhi = (i <= k) ? fr[a] : 0.0;
#pragma unroll
          for (int b = 0; b < EB; ++b) {
            const double fa = (i >= cb + b) ? flo : fhi;
            s[a][b] = fma(-fa, g[b], s[a][b]);
          }
        }
      }
    }
  }

  // L (lower incl. diagonal) and X (from the upper part, transposed) staged
  // in LDS for row-contiguous writes.
  __syncthreads();
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
