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

template <typename TS>
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

// Cross-covariance over a long stream of points (the collapsed ELBO's Kzx
// chunks): out[i][j] = var k(x_i / l, z_j / l) for i < n, j < m, zero elsewhere
// up to rows_total x cols_total.  One 256-thread block owns 64 rows (x, scaled
// and staged once in LDS, read as wave-wide broadcasts) and walks CT column
// tiles of 64 (z) through two LDS buffers: the next tile's scaled inputs are
// loaded into registers before the current tile's arithmetic and stored after
// it, so their global-load latency hides behind a tile of work.  Thread
// (rg, jj) = (wave, lane) computes rows rg*16 .. +15 of column jj.  kmat_kernel's
// short-lived blocks (one tile, inputs staged per tile) ran at a third of the
// store and ALU bound on these shapes.  The dimension loop stays a runtime loop
// (unrolled, the compiler issues all 16 x d row reads at once and runs out of
// registers).  The distances are the same arithmetic in the same order as
// kmat_kernel's, so the float64 entries are identical; F32K evaluates the
// kernel function in float32 (kvalf).
constexpr int CT = 8;

template <typename TS, bool F32K, int DMAX>
__global__ __launch_bounds__(256) void cross_kernel(int kind, const double* __restrict__ x,
                                                    int64_t n, const double* __restrict__ z,
                                                    int64_t m, int d,
                                                    const double* __restrict__ theta, int nls,
                                                    TS* __restrict__ out, int64_t ldo,
                                                    int64_t rows_total, int64_t cols_total) {
  constexpr int NPRE = KT * DMAX / 256;  // staged inputs per thread and tile (DMAX >= 4)
  __shared__ double xs[DMAX][KT];  // [dimension][row]: a wave's 16 rows are contiguous
  __shared__ double zs[2][DMAX][KT];
  const int tid = threadIdx.x, jj = tid & 63, rg = tid >> 6;
  const int64_t i0 = (int64_t)blockIdx.y * KT;
  for (int e = tid; e < KT * DMAX; e += 256) {
    const int r = e / DMAX, q = e % DMAX;
    const int64_t gi = i0 + r;
    xs[q][r] = (q < d && gi < n) ? x[gi * d + q] / theta[nls == 1 ? 0 : q] : 0.0;
  }
  // staged element k of a tile: point r = (tid + 256 k) / DMAX, dimension
  // q = (tid + 256 k) % DMAX
  double pre[NPRE];
  auto load = [&](int64_t jt) {
#pragma unroll
    for (int k = 0; k < NPRE; ++k) {
      const int e = tid + 256 * k, r = e / DMAX, q = e % DMAX;
      const int64_t gj = jt * KT + r;
      pre[k] = (q < d && gj < m) ? z[gj * d + q] / theta[nls == 1 ? 0 : q] : 0.0;
    }
  };
  auto stage = [&](int b) {
#pragma unroll
    for (int k = 0; k < NPRE; ++k) {
      const int e = tid + 256 * k;
      zs[b][e % DMAX][e / DMAX] = pre[k];
    }
  };
  const int64_t ntiles = (cols_total + KT - 1) / KT;
  int64_t jt = (int64_t)blockIdx.x * CT;
  load(jt);
  stage(0);
  const double var = theta[nls];
  __syncthreads();
  for (int c = 0; c < CT && jt < ntiles; ++c, ++jt) {
    const bool more = c + 1 < CT && jt + 1 < ntiles;
    if (more) load(jt + 1);
    const int b = c & 1;
    double d2[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) d2[r] = 0.0;
    for (int q = 0; q < d; ++q) {
      const double zv = zs[b][q][jj];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const double df = xs[q][rg * 16 + r] - zv;
        d2[r] += df * df;
      }
    }
    const int64_t j = jt * KT + jj;
    if (j < cols_total) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t i = i0 + rg * 16 + r;
        if (i >= rows_total) break;
        const bool in = i < n && j < m;
        if constexpr (F32K) {
          out[i * ldo + j] = in ? (float)var * kvalf(kind, (float)d2[r]) : 0.0f;
        } else {
          out[i * ldo + j] = (TS)(in ? var * kval(kind, d2[r]) : 0.0);
        }
      }
    }
    if (more) stage(b ^ 1);
    __syncthreads();
  }
}

}  // namespace

template <typename TS>
int launch_cross(int kind, const double* x, int64_t n, const double* z, int64_t m, int d,
                 const double* theta, int nls, bool f32k, TS* out, int64_t ldo,
                 int64_t rows_total, int64_t cols_total, hipStream_t s) {
  if (rows_total <= 0 || cols_total <= 0) return GPK_OK;
  if (f32k && !std::is_same<TS, float>::value) {
    set_error("cross-covariance: the float32 kernel function needs float32 output");
    return GPK_ERR_BAD_ARG;
  }
  if (d > 32 && !f32k)  // wider inputs: the general builder (same entries)
    return launch_kernel_matrix<TS>(kind, x, n, z, m, d, theta, nls, 0.0, KM_ZERO_PAD, out, ldo,
                                    rows_total, cols_total, false, s);
  const int64_t gy = (rows_total + KT - 1) / KT;
  if (gy > 65535 || d < 1 || d > 32) {
    set_error("cross-covariance: %lld rows / d = %d outside one launch's range",
              (long long)rows_total, d);
    return GPK_ERR_BAD_ARG;
  }
  const int64_t ntiles = (cols_total + KT - 1) / KT;
  dim3 grid((unsigned)((ntiles + CT - 1) / CT), (unsigned)gy, 1);
#define GPK_CROSS(DM)                                                                          \
  do {                                                                                         \
    if (f32k)                                                                                  \
      hipLaunchKernelGGL((cross_kernel<TS, std::is_same<TS, float>::value, DM>), grid,         \
                         dim3(256), 0, s, kind, x, n, z, m, d, theta, nls, out, ldo,           \
                         rows_total, cols_total);                                              \
    else                                                                                       \
      hipLaunchKernelGGL((cross_kernel<TS, false, DM>), grid, dim3(256), 0, s, kind, x, n, z, \
                         m, d, theta, nls, out, ldo, rows_total, cols_total);                  \
  } while (0)
  if (d <= 4) GPK_CROSS(4);
  else if (d <= 8) GPK_CROSS(8);
  else if (d <= 16) GPK_CROSS(16);
  else GPK_CROSS(32);
#undef GPK_CROSS
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

template int launch_cross<double>(int, const double*, int64_t, const double*, int64_t, int,
                                  const double*, int, bool, double*, int64_t, int64_t, int64_t,
                                  hipStream_t);
template int launch_cross<float>(int, const double*, int64_t, const double*, int64_t, int,
                                 const double*, int, bool, float*, int64_t, int64_t, int64_t,
                                 hipStream_t);

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

template int launch_kernel_matrix<double>(int, const double*, int64_t, const double*, int64_t,
                                          int, const double*, int, double, int, double*,
                                          int64_t, int64_t, int64_t, bool, hipStream_t, double*,
                                          const Batch&, double*);
template int launch_kernel_matrix<float>(int, const double*, int64_t, const double*, int64_t,
                                         int, const double*, int, double, int, float*, int64_t,
                                         int64_t, int64_t, bool, hipStream_t, double*,
                                         const Batch&, double*);

}  // namespace gpk
