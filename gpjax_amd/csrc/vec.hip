// Vector kernels of the exact-GP path: triangular matvecs with X = L^-1,
// the Sigma * a matvec with Sigma recomputed from the inputs (float64, for
// iterative refinement of alpha = Sigma^-1 r), and the matrix-core lane-layout
// self-test.  Every reduction writes per-chunk partials and sums them in a
// fixed order, so results are bitwise reproducible run to run.
#include "gemm_core.h"

namespace gpk {
namespace {

constexpr int TRMV_CW = 2048;  // columns per chunk, non-transposed
constexpr int TRMV_RC = 1024;  // rows per chunk, transposed

// y_i partial over columns [c*CW, (c+1)*CW), k <= i.  One wave per row, 16 rows
// per workgroup; each lane reads 16 bytes per load (four float32 or two
// float64 entries, 1 KB per wave), the row's tail up to the diagonal entry by
// entry.  NV right-hand sides (stride np) share every load of X.
template <typename TS, int NV>
__global__ __launch_bounds__(256) void trmv_n_kernel(const TS* __restrict__ X, int64_t ld,
                                                     int64_t np, const double* __restrict__ r,
                                                     double* __restrict__ partial, Batch bt) {
  if (bskip(bt)) return;
  X = bws(X, bt);
  r = bws(r, bt);
  partial = bws(partial, bt);
  const int64_t c = blockIdx.y;
  const int64_t nc = gridDim.y;
  const int64_t row0 = (int64_t)blockIdx.x * 16;
  const int64_t k0 = c * TRMV_CW;
  if (k0 > row0 + 15) return;
  __shared__ double rs[NV][TRMV_CW];
  for (int e = threadIdx.x; e < TRMV_CW; e += 256)
#pragma unroll
    for (int q = 0; q < NV; ++q) rs[q][e] = (k0 + e < np) ? r[q * np + k0 + e] : 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int V = 16 / sizeof(TS);  // entries per 16-byte load
  typedef typename std::conditional<sizeof(TS) == 4, float4, double2>::type Vec;
  for (int qr = 0; qr < 4; ++qr) {
    const int64_t i = row0 + w * 4 + qr;
    const int64_t kend = (i + 1 < k0 + TRMV_CW) ? i + 1 : k0 + TRMV_CW;
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    const TS* xr = X + i * ld;
    for (int64_t kb = k0 + V * lane; kb < kend; kb += 64 * V) {
      const int64_t o = kb - k0;
      if (kb + V <= kend) {
        const Vec v = *reinterpret_cast<const Vec*>(xr + kb);
        const TS* e = reinterpret_cast<const TS*>(&v);
#pragma unroll
        for (int u = 0; u < V; ++u)
#pragma unroll
          for (int q = 0; q < NV; ++q) acc[q] += (double)e[u] * rs[q][o + u];
      } else {
        for (int u = 0; kb + u < kend; ++u)
#pragma unroll
          for (int q = 0; q < NV; ++q) acc[q] += (double)xr[kb + u] * rs[q][o + u];
      }
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double a = acc[q];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
      if (lane == 0) partial[(q * nc + c) * np + i] = a;
    }
  }
}

__global__ void trmv_n_reduce(const double* __restrict__ partial, int64_t np, int64_t nc,
                              int nv, double* __restrict__ y, Batch bt) {
  if (bskip(bt)) return;
  partial = bws(partial, bt);
  y = bws(y, bt);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int64_t last = i / TRMV_CW;
  for (int q = 0; q < nv; ++q) {
    double s = 0.0;
    for (int64_t c = 0; c <= last; ++c) s += partial[(q * nc + c) * np + i];
    y[q * np + i] = s;
  }
}

// y_k partial over rows [c*RC, (c+1)*RC), i >= k.  64 columns per workgroup;
// its four waves take every fourth row (each wave reads 64 consecutive
// columns of a row, 512 or 256 bytes), so four times as many row loads are in
// flight per column, and the four row sums are added in a fixed order.  NV
// right-hand sides (stride np) share every load of X.
constexpr int TRMV_TC = 64;  // columns per workgroup, transposed
template <typename TS, int NV>
__global__ __launch_bounds__(256) void trmv_t_kernel(const TS* __restrict__ X, int64_t ld,
                                                     int64_t np, const double* __restrict__ r,
                                                     double* __restrict__ partial, Batch bt) {
  if (bskip(bt)) return;
  X = bws(X, bt);
  r = bws(r, bt);
  partial = bws(partial, bt);
  const int64_t c = blockIdx.y;
  const int64_t nc = gridDim.y;
  const int64_t col0 = (int64_t)blockIdx.x * TRMV_TC;
  const int64_t i0 = c * TRMV_RC;
  const int64_t i1 = (i0 + TRMV_RC < np) ? i0 + TRMV_RC : np;
  if (i1 <= col0) return;
  __shared__ double rs[NV][TRMV_RC];
  __shared__ double red[NV][4][TRMV_TC];
  for (int e = threadIdx.x; e < TRMV_RC; e += 256)
#pragma unroll
    for (int q = 0; q < NV; ++q) rs[q][e] = (i0 + e < np) ? r[q * np + i0 + e] : 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t k = col0 + lane;
  const int64_t ib = (k > i0) ? k : i0;
  // Four independent sums per thread and vector (rows i, i+4, i+8, i+12) keep
  // four loads in flight; they are combined in a fixed order.
  double a0[NV], a1[NV], a2[NV], a3[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) a0[q] = a1[q] = a2[q] = a3[q] = 0.0;
  int64_t i = ib + w;
  for (; i + 12 < i1; i += 16) {
    const double x0 = (double)X[i * ld + k], x1 = (double)X[(i + 4) * ld + k];
    const double x2 = (double)X[(i + 8) * ld + k], x3 = (double)X[(i + 12) * ld + k];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      a0[q] += x0 * rs[q][i - i0];
      a1[q] += x1 * rs[q][i + 4 - i0];
      a2[q] += x2 * rs[q][i + 8 - i0];
      a3[q] += x3 * rs[q][i + 12 - i0];
    }
  }
  for (; i < i1; i += 4) {
    const double x0 = (double)X[i * ld + k];
#pragma unroll
    for (int q = 0; q < NV; ++q) a0[q] += x0 * rs[q][i - i0];
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) red[q][w][lane] = (a0[q] + a1[q]) + (a2[q] + a3[q]);
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q)
      partial[(q * nc + c) * np + k] =
          ((red[q][0][lane] + red[q][1][lane]) + red[q][2][lane]) + red[q][3][lane];
  }
}

__global__ void trmv_t_reduce(const double* __restrict__ partial, int64_t np, int64_t nchunks,
                              int nv, double* __restrict__ y, Batch bt) {
  if (bskip(bt)) return;
  partial = bws(partial, bt);
  y = bws(y, bt);
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  for (int q = 0; q < nv; ++q) {
    double s = 0.0;
    for (int64_t c = k / TRMV_RC; c < nchunks; ++c) s += partial[(q * nchunks + c) * np + k];
    y[q * np + k] = s;
  }
}

// v_i partial = sum_{j in chunk} Sigma_ij a_j, Sigma recomputed in float64.
constexpr int SMV_CW = 4096;  // columns per chunk
constexpr int SMV_JT = 256;   // columns staged per LDS pass

template <int DMAX, int NV>
__global__ __launch_bounds__(256) void sigma_mv_kernel(int kind, const double* __restrict__ x,
                                                       int64_t n, int64_t np, int d,
                                                       const double* __restrict__ theta,
                                                       int nls, double jitter,
                                                       const double* __restrict__ a,
                                                       double* __restrict__ partial, Batch bt) {
  if (bskip(bt)) return;
  __shared__ __align__(16) double xj_s[SMV_JT][DMAX];  // unpadded: 16-byte broadcast reads
  __shared__ double aj_s[NV][SMV_JT];
  bmember(bt, d, x, n, theta);
  a = bws(a, bt);
  partial = bws(partial, bt);
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t c = blockIdx.y;
  const int64_t nc = gridDim.y;
  const int64_t j0c = c * SMV_CW;
  const double var = theta[nls];
  const double sd = theta[nls + 1];
  double xi[DMAX];
#pragma unroll
  for (int q = 0; q < DMAX; ++q)
    xi[q] = (q < d && i < n) ? x[i * d + q] / theta[nls == 1 ? 0 : q] : 0.0;
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (int64_t j0 = j0c; j0 < j0c + SMV_CW && j0 < np; j0 += SMV_JT) {
    __syncthreads();
    for (int e = threadIdx.x; e < SMV_JT * DMAX; e += 256) {
      const int r = e / DMAX, q = e % DMAX;
      const int64_t gj = j0 + r;
      xj_s[r][q] = (q < d && gj < n) ? x[gj * d + q] / theta[nls == 1 ? 0 : q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NV; ++q)
      aj_s[q][threadIdx.x] = (j0 + threadIdx.x < np) ? a[q * np + j0 + threadIdx.x] : 0.0;
    __syncthreads();
    if (i < n) {
      const int64_t jend = (j0 + SMV_JT < n) ? SMV_JT : n - j0;
      for (int jj = 0; jj < jend; ++jj) {
        double d2 = 0.0;
#pragma unroll
        for (int q = 0; q < DMAX; ++q) {
          const double df = xi[q] - xj_s[jj][q];
          d2 += df * df;
        }
        const double kv = var * kval(kind, d2);
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] += kv * aj_s[q][jj];
      }
    }
  }
  if (i < np) {
    // diagonal additions (jitter + obs_sd^2) and the identity pad
    const bool mine = i >= j0c && i < j0c + SMV_CW;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double v = acc[q];
      if (mine) v += (i < n ? jitter + sd * sd : 1.0) * a[q * np + i];
      partial[(q * nc + c) * np + i] = v;
    }
  }
}

__global__ void sum_chunks(const double* __restrict__ partial, int64_t np, int64_t nchunks,
                           int nv, double* __restrict__ y, Batch bt) {
  if (bskip(bt)) return;
  partial = bws(partial, bt);
  y = bws(y, bt);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  for (int q = 0; q < nv; ++q) {
    double s = 0.0;
    for (int64_t c = 0; c < nchunks; ++c) s += partial[(q * nchunks + c) * np + i];
    y[q * np + i] = s;
  }
}

// ---- v = S a with S symmetric, stored as packed lower NB-tiles in float64 ----
// One workgroup per tile t = I(I+1)/2 + J (I >= J), read once: row sums
// S_IJ a_J go to prow[t] and, below the diagonal, column sums S_IJ^T a_I to
// pcol[t] (layout [NV][ntile][NB]).  Thread (ti, tj) holds rows 8ti .. 8ti+7
// and columns {32p + 2tj, 32p + 2tj + 1 : p < 4} of the tile, so each 16-byte
// load instruction of a wave reads four 256-byte row segments.
template <int NV>
__global__ __launch_bounds__(256) void packed_symv_tile_kernel(const double* __restrict__ S,
                                                               int64_t np,
                                                               const double* __restrict__ a,
                                                               double* __restrict__ prow,
                                                               double* __restrict__ pcol,
                                                               int64_t ntile, Batch bt) {
  if (bskip(bt)) return;
  S = bws(S, bt);
  a = bws(a, bt);
  prow = bws(prow, bt);
  pcol = bws(pcol, bt);
  const int64_t t = blockIdx.x;
  int64_t I = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (I * (I + 1) / 2 > t) --I;
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  const int64_t J = t - I * (I + 1) / 2;
  const bool diag = I == J;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int ti = tid >> 4, tj = tid & 15;
  const double* tile = S + t * NB * NB;
  double2 m[8][4];
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int p = 0; p < 4; ++p)
      m[r][p] = *reinterpret_cast<const double2*>(tile + (ti * 8 + r) * NB + 32 * p + 2 * tj);
  __shared__ double csum[NV][4][NB];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double* aq = a + q * np;
    double aj[4][2], ai[8];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      aj[p][0] = aq[J * NB + 32 * p + 2 * tj];
      aj[p][1] = aq[J * NB + 32 * p + 2 * tj + 1];
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) ai[r] = aq[I * NB + ti * 8 + r];
    // row sums over this thread's columns, then over the 16 threads of a row group
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      double v = 0.0;
#pragma unroll
      for (int p = 0; p < 4; ++p) v += m[r][p].x * aj[p][0] + m[r][p].y * aj[p][1];
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off);
      if (tj == 0) prow[((int64_t)q * ntile + t) * NB + ti * 8 + r] = v;
    }
    if (!diag) {
      // column sums over this thread's rows, then over the 4 row groups of a
      // wave (lanes 16 apart) and, through LDS, over the 4 waves
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        double c0 = 0.0, c1 = 0.0;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          c0 += m[r][p].x * ai[r];
          c1 += m[r][p].y * ai[r];
        }
        c0 += __shfl_xor(c0, 16);
        c0 += __shfl_xor(c0, 32);
        c1 += __shfl_xor(c1, 16);
        c1 += __shfl_xor(c1, 32);
        if (lane < 16) {
          csum[q][w][32 * p + 2 * tj] = c0;
          csum[q][w][32 * p + 2 * tj + 1] = c1;
        }
      }
    }
  }
  if (diag) return;
  __syncthreads();
  for (int e = tid; e < NV * NB; e += 256) {
    const int q = e / NB, j = e % NB;
    pcol[((int64_t)q * ntile + t) * NB + j] =
        ((csum[q][0][j] + csum[q][1][j]) + csum[q][2][j]) + csum[q][3][j];
  }
}

// v[I*NB + i] = sum_{J <= I} prow[t(I, J)][i] + sum_{K > I} pcol[t(K, I)][i]:
// one 512-thread workgroup per row block, four interleaved partial sums per
// entry combined in a fixed order.
__global__ __launch_bounds__(512) void packed_symv_reduce(const double* __restrict__ prow,
                                                          const double* __restrict__ pcol,
                                                          int64_t T, int64_t ntile, int nv,
                                                          double* __restrict__ v, Batch bt) {
  if (bskip(bt)) return;
  prow = bws(prow, bt);
  pcol = bws(pcol, bt);
  v = bws(v, bt);
  const int64_t I = blockIdx.x;
  const int g = threadIdx.x >> 7, i = threadIdx.x & (NB - 1);
  __shared__ double red[4][NB];
  const int64_t np = T * NB;
  for (int q = 0; q < nv; ++q) {
    double acc = 0.0;
#pragma unroll 4
    for (int64_t e = g; e < T; e += 4) {
      const int64_t t = e <= I ? I * (I + 1) / 2 + e : e * (e + 1) / 2 + I;
      const double* src = e <= I ? prow : pcol;
      acc += src[((int64_t)q * ntile + t) * NB + i];
    }
    red[g][i] = acc;
    __syncthreads();
    if (g == 0) v[q * np + I * NB + i] = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
    __syncthreads();
  }
}

// y_j partial over rows [c*RC, (c+1)*RC) of a dense rows x cols matrix:
// sum_i M[i][j] v_i, or sum_i M[i][j]^2 when v is null (rows i >= j only when
// lower), float64 sums.  A block takes 64 columns; its four waves take every
// fourth row of the chunk (wave w: rows c*RC + w, + 4, ...), each lane one
// column, and their four sums are added in wave order.  With one thread per
// column and 256 columns per block, the predictive's 16,384 x 4,096 passes ran
// 256 blocks of one serial 1,024-row chain each, at about 1.1 TB/s.
template <typename TM>
__global__ __launch_bounds__(256) void colreduce_kernel(const TM* __restrict__ M, int64_t ld,
                                                        int64_t rows, int64_t cols,
                                                        const double* __restrict__ v, int lower,
                                                        double* __restrict__ partial) {
  __shared__ double red[4][64];
  const int64_t c = blockIdx.y;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t j = (int64_t)blockIdx.x * 64 + lane;
  const int64_t i0 = c * TRMV_RC;
  const int64_t i1 = (i0 + TRMV_RC < rows) ? i0 + TRMV_RC : rows;
  double acc = 0.0;
  if (j < cols) {
    for (int64_t i = i0 + w; i < i1; i += 4) {
      if (lower && i < j) continue;
      const double m = (double)M[i * ld + j];
      acc += v ? m * v[i] : m * m;
    }
  }
  red[w][lane] = acc;
  __syncthreads();
  if (w == 0 && j < cols)
    partial[c * cols + j] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

__global__ void colreduce_sum(const double* __restrict__ partial, int64_t cols, int64_t nchunks,
                              double* __restrict__ y) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  for (int64_t c = 0; c < nchunks; ++c) s += partial[c * cols + j];
  y[j] = s;
}

// ---- matrix-core lane layout self-test ----
// A[i][k] = i + 100 k, B[k][j] = 1000 + j - 7 k: D = A B^T-as-B computed per lane.
__global__ void mfma_selftest_kernel(double* out64, float* outbf) {
  const int l = threadIdx.x;
  {
    const double a = (double)(l % 16) + 100.0 * (l / 16);
    const double b = 1000.0 + (l % 16) - 7.0 * (l / 16);
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) out64[l * 4 + r] = acc[r];
  }
  {
    bf16x8 av, bv;
    for (int q = 0; q < 8; ++q) {
      const int k = 8 * (l / 16) + q;
      av[q] = (__bf16)(float)((l % 16) + 2 * (k % 8) - (k / 8));
      bv[q] = (__bf16)(float)(3 * (l % 16) - k + 1);
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    for (int r = 0; r < 4; ++r) outbf[l * 4 + r] = acc[r];
  }
}

}  // namespace

template <typename TS>
int launch_trmv(const TS* X, int64_t ld, int64_t np, const double* r, double* y, bool trans,
                double* partial, hipStream_t s, const Batch& bt, int nv) {
  if (np % NB) {
    set_error("trmv: size %lld not a multiple of %d", (long long)np, NB);
    return GPK_ERR_BAD_ARG;
  }
  if (nv != 1 && nv != 3) {
    set_error("trmv: %d right-hand sides (1 or 3 supported)", nv);
    return GPK_ERR_BAD_ARG;
  }
  const dim3 gvec((unsigned)((np + 255) / 256), 1, bt.count);
  if (!trans) {
    const int64_t nc = (np + TRMV_CW - 1) / TRMV_CW;
    const dim3 grid((unsigned)(np / 16), (unsigned)nc, bt.count);
    if (nv == 1)
      hipLaunchKernelGGL((trmv_n_kernel<TS, 1>), grid, dim3(256), 0, s, X, ld, np, r, partial, bt);
    else
      hipLaunchKernelGGL((trmv_n_kernel<TS, 3>), grid, dim3(256), 0, s, X, ld, np, r, partial, bt);
    GPK_LAUNCH_CHECK();
    hipLaunchKernelGGL(trmv_n_reduce, gvec, dim3(256), 0, s, partial, np, nc, nv, y, bt);
  } else {
    const int64_t nc = (np + TRMV_RC - 1) / TRMV_RC;
    const dim3 grid((unsigned)(np / TRMV_TC), (unsigned)nc, bt.count);
    if (nv == 1)
      hipLaunchKernelGGL((trmv_t_kernel<TS, 1>), grid, dim3(256), 0, s, X, ld, np, r, partial, bt);
    else
      hipLaunchKernelGGL((trmv_t_kernel<TS, 3>), grid, dim3(256), 0, s, X, ld, np, r, partial, bt);
    GPK_LAUNCH_CHECK();
    hipLaunchKernelGGL(trmv_t_reduce, gvec, dim3(256), 0, s, partial, np, nc, nv, y, bt);
  }
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

template int launch_trmv<double>(const double*, int64_t, int64_t, const double*, double*, bool,
                                 double*, hipStream_t, const Batch&, int);
template int launch_trmv<float>(const float*, int64_t, int64_t, const double*, double*, bool,
                                double*, hipStream_t, const Batch&, int);

int launch_sigma_matvec(int kind, const double* x, int64_t n, int64_t np, int d,
                        const double* theta, int nls, double jitter, const double* a, double* v,
                        double* partial, hipStream_t s, const Batch& bt, int nv) {
  if (d < 1 || d > MAX_D) {
    set_error("sigma matvec: input dimension %d outside [1, %d]", d, MAX_D);
    return GPK_ERR_BAD_ARG;
  }
  if (nv != 1 && nv != 3) {
    set_error("sigma matvec: %d right-hand sides (1 or 3 supported)", nv);
    return GPK_ERR_BAD_ARG;
  }
  const int64_t nc = (np + SMV_CW - 1) / SMV_CW;
  dim3 grid((unsigned)((np + 255) / 256), (unsigned)nc, bt.count);
#define GPK_SMV(DM)                                                                           \
  do {                                                                                        \
    if (nv == 1)                                                                              \
      hipLaunchKernelGGL((sigma_mv_kernel<DM, 1>), grid, dim3(256), 0, s, kind, x, n, np, d,  \
                         theta, nls, jitter, a, partial, bt);                                 \
    else                                                                                      \
      hipLaunchKernelGGL((sigma_mv_kernel<DM, 3>), grid, dim3(256), 0, s, kind, x, n, np, d,  \
                         theta, nls, jitter, a, partial, bt);                                 \
  } while (0)
  if (d <= 1) GPK_SMV(1);
  else if (d <= 2) GPK_SMV(2);
  else if (d <= 4) GPK_SMV(4);
  else if (d <= 8) GPK_SMV(8);
  else if (d <= 16) GPK_SMV(16);
  else if (d <= 32) GPK_SMV(32);
  else GPK_SMV(64);
#undef GPK_SMV
  GPK_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_chunks, dim3((unsigned)((np + 255) / 256), 1, bt.count), dim3(256), 0, s,
                     partial, np, nc, nv, v, bt);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

int launch_packed_symv(const double* S, int64_t np, const double* a, double* v,
                       double* partial, hipStream_t s, const Batch& bt, int nv) {
  if (np % NB) {
    set_error("packed symv: size %lld not a multiple of %d", (long long)np, NB);
    return GPK_ERR_BAD_ARG;
  }
  if (nv != 1 && nv != 3) {
    set_error("packed symv: %d right-hand sides (1 or 3 supported)", nv);
    return GPK_ERR_BAD_ARG;
  }
  const int64_t T = np / NB, ntile = T * (T + 1) / 2;
  if (ntile > 0x7fffffff) {
    set_error("packed symv: %lld tiles exceed one launch", (long long)ntile);
    return GPK_ERR_BAD_ARG;
  }
  double* prow = partial;
  double* pcol = partial + (size_t)nv * ntile * NB;
  const dim3 grid((unsigned)ntile, 1, bt.count);
  if (nv == 1)
    hipLaunchKernelGGL(packed_symv_tile_kernel<1>, grid, dim3(256), 0, s, S, np, a, prow, pcol,
                       ntile, bt);
  else
    hipLaunchKernelGGL(packed_symv_tile_kernel<3>, grid, dim3(256), 0, s, S, np, a, prow, pcol,
                       ntile, bt);
  GPK_LAUNCH_CHECK();
  hipLaunchKernelGGL(packed_symv_reduce, dim3((unsigned)T, 1, bt.count), dim3(512), 0, s, prow,
                     pcol, T, ntile, nv, v, bt);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

template <typename TM>
int launch_colreduce(const TM* M, int64_t ld, int64_t rows, int64_t cols, const double* v,
                     double* y, double* partial, hipStream_t s, bool lower) {
  if (rows <= 0 || cols <= 0) return GPK_OK;
  const int64_t nc = (rows + TRMV_RC - 1) / TRMV_RC;
  hipLaunchKernelGGL(colreduce_kernel<TM>, dim3((unsigned)((cols + 63) / 64), (unsigned)nc),
                     dim3(256), 0, s, M, ld, rows, cols, v, lower ? 1 : 0, partial);
  GPK_LAUNCH_CHECK();
  hipLaunchKernelGGL(colreduce_sum, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, s,
                     partial, cols, nc, y);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

template int launch_colreduce<double>(const double*, int64_t, int64_t, int64_t, const double*,
                                      double*, double*, hipStream_t, bool);
template int launch_colreduce<float>(const float*, int64_t, int64_t, int64_t, const double*,
                                     double*, double*, hipStream_t, bool);

int launch_mfma_selftest(double* out64, float* outbf, hipStream_t s) {
  hipLaunchKernelGGL(mfma_selftest_kernel, dim3(1), dim3(64), 0, s, out64, outbf);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

}  // namespace gpk
