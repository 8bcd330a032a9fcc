// Gram and cross-covariance builder.
//
// K[i, j] = var * k(rho(x_i / l, z_j / l)) by direct pairwise differences of the
// lengthscale-scaled inputs, the arithmetic of
// gpjax/kernels/computations/dense.py:32-36 with gpjax/kernels/stationary/rbf.py:41-43
// and gpjax/kernels/stationary/utils.py:42-67 (x/l and z/l are formed first, then
// sum_d (x_d - z_d)^2).  The diagonal epilogue adds jitter and then obs_sd^2, the
// order of gpjax/objectives.py:101-102.
//
// The kernel is store-bound (about 3D + 20 float64 operations per 4- or 8-byte
// output element), so the distances stay on the float64 vector ALU: the
// ||x||^2 + ||z||^2 - 2 x.z expansion on the matrix cores would save nothing
// here and would lose digits to cancellation for nearby points.
//
// Work decomposition: one 256-thread workgroup per 64 x 64 output tile; thread
// (rg, jj) owns column jj and rows rg*16 .. rg*16+15, so each store
// instruction of a wave writes 64 consecutive elements of one row.  The scaled
// row inputs are read from LDS as wave-wide broadcasts, the column inputs from
// a transposed, conflict-free LDS tile.
#include "gpk_internal.h"

namespace gpk {
namespace {

constexpr int KT = 64;    // output tile edge
constexpr int DCH = 32;   // input dimensions staged per LDS pass

// F32K (float32 output only): the kernel function in float32 arithmetic (kvalf).
template <typename TS, bool F32K = false>
__global__ __launch_bounds__(256) void kmat_kernel(int kind, const double* __restrict__ x,
                                                   int64_t n, const double* __restrict__ z,
                                                   int64_t m, int d,
                                                   const double* __restrict__ theta, int nls,
                                                   double jitter, int diag_mode, TS* out,
                                                   int64_t ldo, int64_t rows_total,
                                                   int64_t cols_total, int lower_only,
                                                   double* __restrict__ dblk,
                                                   double* __restrict__ spk, Batch bt) {
  __shared__ double xi_s[KT][DCH + 1];
  __shared__ double zj_s[DCH][KT + 1];
  if (bt.gp) {  // batch member: its own inputs (a Gram matrix, z = x)
    bmember(bt, d, x, n, theta);
    z = x;
    m = n;
  }
  out = bws(out, bt);
  dblk = bws(dblk, bt);
  spk = bws(spk, bt);

  int64_t ti, tj;
  if (lower_only) {
    // blockIdx.x enumerates NB-tiles on/below the diagonal, blockIdx.y the
    // four 64-subtiles of that NB-tile.
    const int64_t t = blockIdx.x;
    int64_t a = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (a * (a + 1) / 2 > t) --a;
    while ((a + 1) * (a + 2) / 2 <= t) ++a;
    const int64_t b = t - a * (a + 1) / 2;
    ti = a * 2 + (blockIdx.y >> 1);
    tj = b * 2 + (blockIdx.y & 1);
  } else {
    tj = blockIdx.x;
    ti = blockIdx.y;
  }
  const int64_t i0 = ti * KT, j0 = tj * KT;
  const int tid = threadIdx.x;
  const int jj = tid & 63;
  const int rg = tid >> 6;

  double d2[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) d2[r] = 0.0;

  for (int dc = 0; dc < d; dc += DCH) {
    const int dn = min(DCH, d - dc);
    __syncthreads();
    for (int e = tid; e < KT * DCH; e += 256) {
      const int r = e / DCH, q = e % DCH;
      if (q < dn) {
        const int dd = dc + q;
        const double l = theta[nls == 1 ? 0 : dd];
        const int64_t gi = i0 + r;
        xi_s[r][q] = gi < n ? x[gi * d + dd] / l : 0.0;
        const int64_t gj = j0 + r;
        zj_s[q][r] = gj < m ? z[gj * d + dd] / l : 0.0;
      }
    }
    __syncthreads();
    for (int q = 0; q < dn; ++q) {
      const double zv = zj_s[q][jj];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const double df = xi_s[rg * 16 + r][q] - zv;
        d2[r] += df * df;
      }
    }
  }

  const double var = theta[nls];
  const double sd = theta[nls + 1];
  const int64_t j = j0 + jj;
  if (j >= cols_total) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int64_t i = i0 + rg * 16 + r;
    if (i >= rows_total) continue;
    double v;
    if constexpr (F32K) {
      out[i * ldo + j] = (i < n && j < m) ? (float)var * kvalf(kind, (float)d2[r]) : 0.0f;
      continue;
    }
    if (i < n && j < m) {
      v = var * kval(kind, d2[r]);
      const int dm = diag_mode & KM_DIAG_MASK;
      if (dm > 0 && i == j) {
        v = v + jitter;
        if (dm > 1) v = v + sd * sd;
      }
    } else {
      v = (i == j && !(diag_mode & KM_ZERO_PAD)) ? 1.0 : 0.0;
    }
    out[i * ldo + j] = (TS)v;
    // Diagonal NB-blocks also go to a float64 side buffer (lower_only mode):
    // the factorisation keeps its pivots' blocks in float64.
    if (dblk && (i / NB) == (j / NB)) {
      const int64_t bk = i / NB;
      dblk[bk * NB * NB + (i - bk * NB) * NB + (j - bk * NB)] = v;
    }
    // The whole lower triangle in float64, NB-tiles packed in blockIdx.x order
    // (lower_only mode): the matrix of the float64 refinement's products.
    if (spk) spk[(int64_t)blockIdx.x * NB * NB + (i % NB) * NB + (j % NB)] = v;
  }
}

}  // namespace

template <typename TS>
int launch_kernel_matrix(int kind, const double* x, int64_t n, const double* z, int64_t m,
                         int d, const double* theta, int nls, double jitter, int diag_mode,
                         TS* out, int64_t ldo, int64_t rows_total, int64_t cols_total,
                         bool lower_only, hipStream_t s, double* dblk, const Batch& bt,
                         double* spk) {
  if (rows_total <= 0 || cols_total <= 0) return GPK_OK;
  dim3 grid;
  if (lower_only) {
    if (rows_total != cols_total || rows_total % NB != 0) {
      set_error("kernel matrix: lower_only needs a square multiple of %d", NB);
      return GPK_ERR_BAD_ARG;
    }
    const int64_t nt = rows_total / NB;
    grid = dim3((unsigned)(nt * (nt + 1) / 2), 4, bt.count);
  } else {
    const int64_t gy = (rows_total + KT - 1) / KT;
    if (gy > 65535) {
      set_error("kernel matrix: %lld rows exceed one launch", (long long)rows_total);
      return GPK_ERR_BAD_ARG;
    }
    grid = dim3((unsigned)((cols_total + KT - 1) / KT), (unsigned)gy, bt.count);
  }
  hipLaunchKernelGGL(kmat_kernel<TS>, grid, dim3(256), 0, s, kind, x, n, z, m, d, theta, nls,
                     jitter, diag_mode, out, ldo, rows_total, cols_total, lower_only ? 1 : 0,
                     dblk, lower_only ? spk : nullptr, bt);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

int launch_cross_f32(int kind, const double* x, int64_t n, const double* z, int64_t m, int d,
                     const double* theta, int nls, float* out, int64_t ldo, int64_t rows_total,
                     int64_t cols_total, hipStream_t s) {
  if (rows_total <= 0 || cols_total <= 0) return GPK_OK;
  const int64_t gy = (rows_total + KT - 1) / KT;
  if (gy > 65535) {
    set_error("kernel matrix: %lld rows exceed one launch", (long long)rows_total);
    return GPK_ERR_BAD_ARG;
  }
  dim3 grid((unsigned)((cols_total + KT - 1) / KT), (unsigned)gy, 1);
  hipLaunchKernelGGL((kmat_kernel<float, true>), grid, dim3(256), 0, s, kind, x, n, z, m, d,
                     theta, nls, 0.0, KM_ZERO_PAD, out, ldo, rows_total, cols_total, 0, nullptr,
                     nullptr, Batch());
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

template int launch_kernel_matrix<double>(int, const double*, int64_t, const double*, int64_t,
                                          int, const double*, int, double, int, double*,
                                          int64_t, int64_t, int64_t, bool, hipStream_t, double*,
                                          const Batch&, double*);
template int launch_kernel_matrix<float>(int, const double*, int64_t, const double*, int64_t,
                                         int, const double*, int, double, int, float*, int64_t,
                                         int64_t, int64_t, bool, hipStream_t, double*,
                                         const Batch&, double*);

}  // namespace gpk
