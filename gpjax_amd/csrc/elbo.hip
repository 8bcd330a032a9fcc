// Collapsed (Titsias) ELBO and its gradient, gpjax/objectives.py:321-416 with
// CollapsedVariationalGaussian (gpjax/variational_families.py:767-784).
//
// The reference forms A = Lz^-1 Kzx / sigma (an M x N matrix, materialised)
// and AA^T.  Here the data enter only through the N-summed statistics
//   P = Kzx Kxz (M x M),  b = Kzx r,  q = Kzx 1,  r = y - c,
// accumulated chunk by chunk: Kzx is evaluated once per chunk of 131,072
// points into an M x 131,072 buffer and P += Kc Kc^T runs on the matrix
// cores, so no N x M buffer exists.  With s = sigma^2 and C = Kzz + P / s the
// determinant and inversion lemmas that the reference's comments state
// (objectives.py:361-385) give
//   2 ELBO = -n log(2 pi s) - log|C| + log|Kzz| - r^T r / s + b^T C^-1 b / s^2
//            - n var / s + tr(Kzz^-1 P) / s.
// Gradient (derived by differentiating the line above; c = C^-1 b):
//   2 dELBO = 2 sum(W * dKzx) + tr(G_K dKzz) + g_s ds - (n / s) dvar,
//   W   = G_P Kzx + (c / s^2) r^T,  G_P = (Kzz^-1 - C^-1) / s - c c^T / s^3,
//   G_K = Kzz^-1 - C^-1 - c c^T / s^2 - Kzz^-1 P Kzz^-1 / s,
//   g_s = -n/s + tr(C^-1 P)/s^2 + r^T r/s^2 + c^T P c/s^4 - 2 b^T c/s^3
//         + n var/s^2 - tr(Kzz^-1 P)/s^2.
// The backward evaluates Kzx once per chunk of 131,072 points, forms W for the
// chunk with one GEMM on the matrix cores and reduces it against
// dKzx / d(lengthscale, variance, Z) elementwise.
// GPK_PREC_FP64 runs everything in float64.  GPK_PREC_MIXED keeps the Kzx
// chunks in float32 (squared distances in float64, the kernel function in
// float32 arithmetic) and runs the two N-streamed products, P += Kc Kc^T and
// Kc^T G_P, as bf16-split products whose fp32 chunk sums are flushed to
// float64; P, b, q, the M x M factorisations and inverses and every reduction
// stay float64.
#include <cmath>

#include "factor.h"
#include "gemm_core.h"

namespace gpk {
namespace {

constexpr int EDMAX = 32;   // largest input dimension on the ELBO path
// Points per Kzx chunk (statistics, data pass).  Against 65,536 it halves the
// chunk launches and lets the split-K statistics GEMM fill whole rounds of the
// chip (see S below): 231 -> 228 ms (FP64) and 83.3 -> 82.8 ms (mixed) at
// N = 1e6, M = 2,048, for 2 GB more workspace.
constexpr int64_t BW_CHUNK = 131072;
constexpr int64_t EW = 1024;         // points per block of the gradient data pass

__device__ __forceinline__ double lsq(const double* theta, int nls, int q) {
  return theta[nls == 1 ? 0 : q];
}

template <int DMAX>
__device__ __forceinline__ double d2_of(const double* a, const double* b) {
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < DMAX; ++q) {
    const double df = a[q] - b[q];
    s += df * df;
  }
  return s;
}

// ---- E1: streamed statistics b = Kzx r and q = Kzx 1 over one chunk ----
// Kc[i * ldk + (g - c0)] = k(z_i, x_g) for the chunk's points; one block per
// inducing row, fixed-order reduction; the first chunk writes, later ones add.
// (P = Kzx Kxz is accumulated from the same chunk by one matrix-core GEMM.)
template <typename TK>
__global__ __launch_bounds__(256) void chunk_rowsum_kernel(
    const TK* __restrict__ Kc, int64_t ldk, int64_t c0, int64_t ncol,
    const double* __restrict__ y, const double* __restrict__ theta, int nls, int accumulate,
    double* __restrict__ b, double* __restrict__ q) {
  __shared__ double red[2][4];
  const int64_t i = blockIdx.x;
  const double cmean = theta[nls + 2];
  const TK* row = Kc + i * ldk;
  double sb = 0.0, sq = 0.0;
  for (int64_t j = threadIdx.x; j < ncol; j += 256) {
    const double k = (double)row[j];
    sb += k * (y[c0 + j] - cmean);
    sq += k;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sb += __shfl_xor(sb, off);
    sq += __shfl_xor(sq, off);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = sb;
    red[1][threadIdx.x >> 6] = sq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double tb = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    const double tq = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    b[i] = (accumulate ? b[i] : 0.0) + tb;
    q[i] = (accumulate ? q[i] : 0.0) + tq;
  }
}

// sum r_i^2 and sum r_i over the data, r = y - c: per-block partials.
__global__ __launch_bounds__(256) void rstats_kernel(const double* __restrict__ y, int64_t n,
                                                     const double* __restrict__ theta, int nls,
                                                     double* __restrict__ rpart) {
  __shared__ double red[2][4];
  const double c = theta[nls + 2];
  double s2 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double r = y[i] - c;
    s2 += r * r;
    s1 += r;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s2 += __shfl_xor(s2, off);
    s1 += __shfl_xor(s1, off);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][w] = s2;
    red[1][w] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    rpart[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    rpart[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

// C += P / s on the lower NB-tiles.
__global__ void add_p_kernel(double* __restrict__ C, const double* __restrict__ P, int64_t ld,
                             int64_t mp, const double* __restrict__ theta, int nls) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= mp * mp) return;
  const int64_t i = e / mp, j = e % mp;
  if (i / NB < j / NB) return;
  const double sd = theta[nls + 1];
  C[i * ld + j] += P[i * ld + j] / (sd * sd);
}

// A[j][i] = A[i][j] for the strictly lower NB-tiles (diagonal tiles are full).
// P (full, symmetric) from the split-K partials Pp[i * ldp + z * kz + j]
// (z < S) of its lower tiles; off-diagonal tiles are mirrored.
__global__ void psum_kernel(const double* __restrict__ Pp, int64_t ldp, int64_t kz, int64_t S,
                            double* __restrict__ P, int64_t mp) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= mp * mp) return;
  const int64_t i = e / mp, j = e % mp;
  if (i / TILE < j / TILE) return;
  double v = 0.0;
  for (int64_t z = 0; z < S; ++z) v += Pp[i * ldp + z * kz + j];
  P[i * mp + j] = v;
  if (i / TILE > j / TILE) P[j * mp + i] = v;
}

__global__ void mirror_kernel(double* __restrict__ A, int64_t ld, int64_t mp) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= mp * mp) return;
  const int64_t i = e / mp, j = e % mp;
  if (i / NB > j / NB) A[j * ld + i] = A[i * ld + j];
}

// G_P = (Kinv - Cinv) / s - c c^T / s^3 over the dense mp x mp (ld = mp), and
// its float32 rounding in Gf (the operand of the mixed data pass) when Gf is
// given; gmax (optional) gets per-block partials of max |G_P| (the fp16 split's
// operand scale).  Grid-stride over 256-thread blocks.
__global__ __launch_bounds__(256) void gp_kernel(double* __restrict__ G, float* __restrict__ Gf,
                                                 const double* __restrict__ Kinv,
                                                 const double* __restrict__ Cinv,
                                                 const double* __restrict__ c, int64_t mp,
                                                 const double* __restrict__ theta, int nls,
                                                 double* __restrict__ gmax) {
  __shared__ double red[4];
  const double sd = theta[nls + 1], s = sd * sd;
  double mx = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < mp * mp;
       e += (int64_t)gridDim.x * 256) {
    const int64_t i = e / mp, j = e % mp;
    const double g = (Kinv[e] - Cinv[e]) / s - c[i] * c[j] / (s * s * s);
    G[e] = g;
    if (Gf) Gf[e] = (float)g;
    mx = fmax(mx, fabs(g));
  }
  if (!gmax) return;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) gmax[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Power-of-two fp16 operand scales of the mixed ELBO's products: each maps its
// bound b to at most 2^14 (b clamped to [2^-60, 2^60]).  |Kzx| <= var (every
// kernel here peaks at var); gmax: nf partials of max |G_P| (or null).
// fs[0] = fs[1] = scale of Kzx (statistics product), fs[2] = scale of Kzx and
// fs[3] = scale of G_P (data-pass product).
// One wave of 64 threads.
__global__ void elbo_fp16_scale_kernel(const double* __restrict__ theta, int nls,
                                       const double* __restrict__ gmax, int nf,
                                       float* __restrict__ fs) {
  auto sc = [](double b) {
    b = fmin(fmax(b, 0x1p-60), 0x1p60);
    return (float)ldexp(1.0, (int)floor(log2(16384.0 / b)));
  };
  double g = 0.0;
  if (gmax)
    for (int k = threadIdx.x; k < nf; k += 64) g = fmax(g, gmax[k]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) g = fmax(g, __shfl_xor(g, off));
  if (threadIdx.x != 0) return;
  const float sk = sc(theta[nls]);
  fs[0] = sk;
  fs[1] = sk;
  fs[2] = sk;
  if (gmax) fs[3] = sc(g);
}

// G_K = Kinv - Cinv - c c^T / s^2 - T2 / s,  T2 = Kinv P Kinv
__global__ void gk_kernel(double* __restrict__ G, const double* __restrict__ Kinv,
                          const double* __restrict__ Cinv, const double* __restrict__ T2,
                          const double* __restrict__ c, int64_t ld, int64_t mp,
                          const double* __restrict__ theta, int nls) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= mp * mp) return;
  const int64_t i = e / mp, j = e % mp;
  const double sd = theta[nls + 1], s = sd * sd;
  G[i * ld + j] = Kinv[i * ld + j] - Cinv[i * ld + j] - c[i] * c[j] / (s * s) -
                  T2[i * ld + j] / s;
}

// Per-block partials of sum_ij A_ij B_ij over a dense mp x mp pair.
__global__ __launch_bounds__(256) void frob_kernel(const double* __restrict__ A,
                                                   const double* __restrict__ B, int64_t cnt,
                                                   double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < cnt;
       e += (int64_t)gridDim.x * 256)
    s += A[e] * B[e];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Per-block partials of sum_ij (G_P,ij P_ij)^2, G_P = (Kinv - Cinv) / s - c c^T / s^3
// formed on the fly (the a-posteriori error model of the mixed statistics).
__global__ __launch_bounds__(256) void gpfrob_kernel(const double* __restrict__ Kinv,
                                                     const double* __restrict__ Cinv,
                                                     const double* __restrict__ c,
                                                     const double* __restrict__ P, int64_t mp,
                                                     const double* __restrict__ theta, int nls,
                                                     double* __restrict__ part) {
  __shared__ double red[4];
  const double sd = theta[nls + 1], s = sd * sd;
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < mp * mp;
       e += (int64_t)gridDim.x * 256) {
    const int64_t i = e / mp, j = e % mp;
    const double g = (Kinv[e] - Cinv[e]) / s - c[i] * c[j] / (s * s * s);
    const double v = g * P[e];
    acc += v * v;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// ---- E3: W = G_P Kzx + (c / s^2) r^T contracted against dKzx ----
// One launch per chunk of points [c0, c0 + ncol).  Wt holds the chunk's
// (G_P Kzx)^T, Wt[(g - c0) * mp + i], formed beforehand by one matrix-core GEMM
// Kc^T G_P (G_P is symmetric) from Kc = Kzx evaluated once for the chunk.
// Block (mt, cb): inducing rows mt*TILE .. +TILE (thread row li = t % TILE; the
// two halves of the block take alternate points) against points
// c0 + cb*EW .. +EW.  The kernel value and its derivative factor are
// recomputed from the inputs per element.  Partials, split = c0 / EW + cb:
// gzpart[(split * mp + i) * DMAX + q] and
// tpart[(split * mt_count + mt) * (DMAX + 1) + q] (lengthscales, then variance).
// Mixed path (TW = float): Wt in float32 and the kernel value and factor in
// float32 arithmetic; distances, products and sums stay float64.
template <int DMAX, typename TW>
__global__ __launch_bounds__(256) void elbo_wgrad_kernel(
    int kind, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, int64_t m, int d, const double* __restrict__ theta, int nls,
    const TW* __restrict__ Wt, int64_t mp, int64_t c0, int64_t ncol,
    const double* __restrict__ cvec, double* __restrict__ gzpart, double* __restrict__ tpart) {
  __shared__ double xc[TILE][DMAX + 1];
  __shared__ double rc[TILE];
  __shared__ double gzs[TILE][DMAX];
  __shared__ double red_s[4][DMAX + 1];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int li = t & (TILE - 1), half = t >> 7;
  const int64_t mt = blockIdx.x, cb = blockIdx.y;
  const int64_t gi = mt * TILE + li;
  const bool rowok = gi < m;
  const double sd = theta[nls + 1], s = sd * sd, cmean = theta[nls + 2];
  double zi[DMAX], s_ls[DMAX], gzr[DMAX];
#pragma unroll
  for (int q = 0; q < DMAX; ++q) {
    zi[q] = (rowok && q < d) ? z[gi * d + q] / lsq(theta, nls, q) : 0.0;
    s_ls[q] = 0.0;
    gzr[q] = 0.0;
  }
  const double gb = rowok ? cvec[gi] / (s * s) : 0.0;
  double s_var = 0.0;
  const int64_t p0 = c0 + cb * EW;
  const int64_t p1 = (p0 + EW < c0 + ncol) ? p0 + EW : c0 + ncol;
  for (int64_t b0 = p0; b0 < p1; b0 += TILE) {
    __syncthreads();
    for (int e = t; e < TILE * DMAX; e += 256) {
      const int r = e / DMAX, q = e % DMAX;
      const int64_t g = b0 + r;
      xc[r][q] = (q < d && g < p1) ? x[g * d + q] / lsq(theta, nls, q) : 0.0;
    }
    if (t < TILE) rc[t] = (b0 + t < p1) ? y[b0 + t] - cmean : 0.0;
    __syncthreads();
    const int jend = (b0 + TILE < p1) ? TILE : (int)(p1 - b0);
    if (rowok) {
      const TW* wcol = Wt + (b0 - c0) * mp + gi;
      for (int jj = half; jj < jend; jj += 2) {
        const double W = (double)wcol[(int64_t)jj * mp] + gb * rc[jj];
        const double d2 = d2_of<DMAX>(zi, xc[jj]);
        double kv, fv;
        if constexpr (std::is_same<TW, float>::value) {
          float k32, f32;
          kval_facf(kind, (float)d2, k32, f32);
          kv = k32;
          fv = f32;
        } else {
          kval_fac(kind, d2, kv, fv);
        }
        s_var += W * kv;
        const double wf = W * fv;
#pragma unroll
        for (int q = 0; q < DMAX; ++q) {
          const double df = zi[q] - xc[jj][q];
          s_ls[q] += wf * df * df;
          gzr[q] += wf * df;
        }
      }
    }
  }
  if (half == 1) {
#pragma unroll
    for (int q = 0; q < DMAX; ++q) gzs[li][q] = gzr[q];
  }
  __syncthreads();
  const int64_t split = c0 / EW + cb;
  if (half == 0) {
#pragma unroll
    for (int q = 0; q < DMAX; ++q) gzpart[(split * mp + gi) * DMAX + q] = gzr[q] + gzs[li][q];
  }
#pragma unroll
  for (int q = 0; q < DMAX + 1; ++q) {
    double v = q < DMAX ? s_ls[q] : s_var;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red_s[w][q] = v;
  }
  __syncthreads();
  if (t < DMAX + 1)
    tpart[(split * gridDim.x + mt) * (DMAX + 1) + t] =
        red_s[0][t] + red_s[1][t] + red_s[2][t] + red_s[3][t];
}

// ---- E4: tr(G_K dKzz) terms; one thread per inducing point (row), G_K symmetric ----
template <int DMAX>
__global__ __launch_bounds__(256) void elbo_kzz_adj_kernel(
    int kind, const double* __restrict__ z, int64_t m, int d, const double* __restrict__ theta,
    int nls, const double* __restrict__ GK, int64_t ld, double* __restrict__ gzkz,
    double* __restrict__ tpart) {
  __shared__ double zj[256][DMAX + 1];
  __shared__ double red_s[4][DMAX + 1];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t i = (int64_t)blockIdx.x * 256 + t;
  double zi[DMAX];
#pragma unroll
  for (int q = 0; q < DMAX; ++q) zi[q] = (q < d && i < m) ? z[i * d + q] / lsq(theta, nls, q) : 0.0;
  double s_var = 0.0, s_ls[DMAX], gzr[DMAX];
#pragma unroll
  for (int q = 0; q < DMAX; ++q) {
    s_ls[q] = 0.0;
    gzr[q] = 0.0;
  }
  for (int64_t j0 = 0; j0 < m; j0 += 256) {
    __syncthreads();
    for (int e = t; e < 256 * DMAX; e += 256) {
      const int r = e / DMAX, q = e % DMAX;
      const int64_t g = j0 + r;
      zj[r][q] = (q < d && g < m) ? z[g * d + q] / lsq(theta, nls, q) : 0.0;
    }
    __syncthreads();
    if (i < m) {
      const int64_t jn = (m - j0 < 256) ? m - j0 : 256;
      for (int jj = 0; jj < jn; ++jj) {
        const double g = GK[(j0 + jj) * ld + i];
        double kv, fv;
        kval_fac(kind, d2_of<DMAX>(zi, zj[jj]), kv, fv);
        s_var += g * kv;
        const double gf = g * fv;
#pragma unroll
        for (int q = 0; q < DMAX; ++q) {
          const double df = zi[q] - zj[jj][q];
          s_ls[q] += gf * df * df;
          gzr[q] += gf * df;
        }
      }
    }
  }
  if (i < m) {
#pragma unroll
    for (int q = 0; q < DMAX; ++q) gzkz[i * DMAX + q] = gzr[q];
  }
#pragma unroll
  for (int q = 0; q < DMAX + 1; ++q) {
    double v = q < DMAX ? s_ls[q] : s_var;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red_s[w][q] = v;
  }
  __syncthreads();
  if (t < DMAX + 1)
    tpart[blockIdx.x * (DMAX + 1) + t] = red_s[0][t] + red_s[1][t] + red_s[2][t] + red_s[3][t];
}

// ---- scalar slots ----
enum Scal {
  SC_RR = 0, SC_SR, SC_LDK, SC_LDC, SC_BC, SC_TRKP, SC_TRCP, SC_CPC, SC_INFO, SC_ELBO, SC_COUNT
};

__device__ double block_sum256(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void elbo_fwd_finalize(
    const double* __restrict__ rpart, int nr, const double* __restrict__ fpart, int nf,
    const double* __restrict__ leafK, const int* __restrict__ infoK,
    const double* __restrict__ leafC, const int* __restrict__ infoC, int64_t nleaf,
    const double* __restrict__ b, const double* __restrict__ c, int64_t mp, int64_t n,
    const double* __restrict__ theta, int nls, const double* __restrict__ epart,
    int uncalibrated, double* __restrict__ scal, double* __restrict__ result) {
  __shared__ double red[4];
  double bc = 0.0;
  for (int64_t i = threadIdx.x; i < mp; i += 256) bc += b[i] * c[i];
  bc = block_sum256(bc, red);
  double tr = 0.0;
  for (int k = threadIdx.x; k < nf; k += 256) tr += fpart[k];
  tr = block_sum256(tr, red);
  double gp2 = 0.0;
  if (epart) {
    for (int k = threadIdx.x; k < nf; k += 256) gp2 += epart[k];
    gp2 = block_sum256(gp2, red);
  }
  if (threadIdx.x == 0) {
    double rr = 0.0, sr = 0.0;
    for (int k = 0; k < nr; ++k) {
      rr += rpart[2 * k];
      sr += rpart[2 * k + 1];
    }
    double ldk = 0.0, ldc = 0.0;
    int info = 0;
    for (int64_t k = 0; k < nleaf; ++k) {
      ldk += leafK[k];
      ldc += leafC[k];
      if (info == 0 && infoK[k] != 0) info = infoK[k];
      if (info == 0 && infoC[k] != 0) info = infoC[k];
    }
    const double var = theta[nls], sd = theta[nls + 1], s = sd * sd;
    const double two_pi = 6.283185307179586;
    const double nn = (double)n;
    const double logB = ldc - ldk;
    const double quad = (rr - bc / s) / s;
    const double trace = nn * var / s - tr / s;
    double elbo = 0.5 * (-nn * log(two_pi * s) - logB - quad - trace);
    if (info != 0) elbo = __builtin_nan("");
    scal[SC_RR] = rr;
    scal[SC_SR] = sr;
    scal[SC_LDK] = ldk;
    scal[SC_LDC] = ldc;
    scal[SC_BC] = bc;
    scal[SC_TRKP] = tr;
    scal[SC_INFO] = info;
    scal[SC_ELBO] = elbo;
    result[GPK_RES_MLL] = elbo;
    result[GPK_RES_LOGDET] = logB;
    result[GPK_RES_QUAD] = quad;
    result[GPK_RES_RESID] = trace;
    result[GPK_RES_INFO] = (double)info;
    for (int k = 5; k < GPK_RESULT_LEN; ++k) result[k] = 0.0;
    if (epart) {
      // Mixed statistics: the bf16-split products and their fp32 chunk sums
      // leave P with entry errors R_ij of relative size ~ 2^-24 and random
      // sign, whose first-order effect on the ELBO is tr(G_P R) / 2.  Modelled
      // as 2^-24 |P_ij| with independent signs its spread is
      // 2^-24 sqrt(2) ||G_P o P||_F / 2 (measured |errors| <= 2.9x it).
      const double err = 0.5 * 0x1p-24 * sqrt(2.0 * gp2);
      result[GPK_RES_MLL_ERR] = err;
      const bool bad = info != 0 || !isfinite(elbo) || !(5.0 * err <= 1e-6 * fabs(elbo)) ||
                       uncalibrated;
      result[GPK_RES_FLAG] = bad ? 1.0 : 0.0;
    }
  }
}

// Backward scalars: tr(C^-1 P) from fpart, c^T P c from Pc.
__global__ __launch_bounds__(256) void elbo_bwd_scalars(const double* __restrict__ fpart, int nf,
                                                        const double* __restrict__ c,
                                                        const double* __restrict__ Pc,
                                                        int64_t mp, double* __restrict__ scal) {
  __shared__ double red[4];
  double tr = 0.0;
  for (int k = threadIdx.x; k < nf; k += 256) tr += fpart[k];
  tr = block_sum256(tr, red);
  double cpc = 0.0;
  for (int64_t i = threadIdx.x; i < mp; i += 256) cpc += c[i] * Pc[i];
  cpc = block_sum256(cpc, red);
  if (threadIdx.x == 0) {
    scal[SC_TRCP] = tr;
    scal[SC_CPC] = cpc;
  }
}

// Gradient assembly: block q < nls -> lengthscale q, nls -> variance, nls+1 ->
// obs_stddev, nls+2 -> mean; blocks beyond write g_Z rows.
__global__ __launch_bounds__(256) void elbo_grad_finalize(
    const double* __restrict__ tpart, int64_t nt3, const double* __restrict__ tpart2,
    int64_t nt4, const double* __restrict__ gzpart, int64_t S2, const double* __restrict__ gzkz,
    const double* __restrict__ c, const double* __restrict__ qv, const double* __restrict__ scal,
    int64_t m, int64_t mp, int d, int64_t n, const double* __restrict__ theta, int nls,
    double gbar, double* __restrict__ grad, double* __restrict__ grad_z, int dmax) {
  __shared__ double red[4];
  const int q = blockIdx.x;
  const double var = theta[nls], sd = theta[nls + 1], s = sd * sd;
  const double nn = (double)n;
  if (q < nls + 3) {
    double acc3 = 0.0, acc4 = 0.0;
    if (q < nls) {
      const int q0 = nls == 1 ? 0 : q, q1 = nls == 1 ? d : q + 1;
      for (int64_t k = threadIdx.x; k < nt3; k += 256)
        for (int qq = q0; qq < q1; ++qq) acc3 += tpart[k * (dmax + 1) + qq];
      for (int64_t k = threadIdx.x; k < nt4; k += 256)
        for (int qq = q0; qq < q1; ++qq) acc4 += tpart2[k * (dmax + 1) + qq];
    } else if (q == nls) {
      for (int64_t k = threadIdx.x; k < nt3; k += 256) acc3 += tpart[k * (dmax + 1) + dmax];
      for (int64_t k = threadIdx.x; k < nt4; k += 256) acc4 += tpart2[k * (dmax + 1) + dmax];
    } else if (q == nls + 2) {
      for (int64_t i = threadIdx.x; i < mp; i += 256) acc3 += c[i] * qv[i];
    }
    acc3 = block_sum256(acc3, red);
    acc4 = block_sum256(acc4, red);
    if (threadIdx.x == 0) {
      double g;
      if (q < nls) {
        g = var * (acc3 + 0.5 * acc4) / theta[q];
      } else if (q == nls) {
        g = acc3 + 0.5 * acc4 - 0.5 * nn / s;
      } else if (q == nls + 1) {
        const double gs = -nn / s + scal[SC_TRCP] / (s * s) + scal[SC_RR] / (s * s) +
                          scal[SC_CPC] / (s * s * s * s) - 2.0 * scal[SC_BC] / (s * s * s) +
                          nn * var / (s * s) - scal[SC_TRKP] / (s * s);
        g = gs * sd;
      } else {
        g = scal[SC_SR] / s - acc3 / (s * s);
      }
      grad[q] = gbar * g;
    }
    return;
  }
  const int64_t e = (int64_t)(q - (nls + 3)) * 256 + threadIdx.x;
  if (e >= m * d) return;
  const int64_t i = e / d;
  const int qq = (int)(e % d);
  double g = gzkz[i * dmax + qq];
  for (int64_t k = 0; k < S2; ++k) g += gzpart[(k * mp + i) * dmax + qq];
  grad_z[e] = gbar * (-var * g / lsq(theta, nls, qq));
}

// ---- workspace layout ----
struct ElboLayout {
  int64_t mp = 0, mt = 0, ntile = 0, S2 = 1, chunk2 = EW, nbw = TILE;
  int64_t nr = 256, nf = 1024;
  size_t off_K = 0, off_XK = 0, off_C = 0, off_XC = 0, off_P = 0, off_Kinv = 0, off_Cinv = 0;
  size_t off_G = 0, off_T1 = 0, off_T2 = 0, off_T = 0;
  size_t off_liK = 0, off_ldK = 0, off_ifK = 0, off_liC = 0, off_ldC = 0, off_ifC = 0;
  size_t off_b = 0, off_q = 0, off_c = 0, off_u = 0, off_Pc = 0;
  size_t off_gzpart = 0, off_tpart = 0;
  size_t off_tpart2 = 0, off_gzkz = 0, off_rpart = 0, off_fpart = 0, off_epart = 0, off_fs = 0;
  size_t off_vpart = 0, off_scal = 0;
  // one chunk of Kzx [mp][nbw]; (G_P Kzx)^T [nbw][mp] in the backward, the
  // split-K partials of P in the forward
  size_t off_kc = 0, off_wt = 0;
  size_t total = 0;
};

inline size_t al(size_t b) { return (b + 255) / 256 * 256; }

ElboLayout make_elbo_layout(int64_t n, int64_t m) {
  ElboLayout L;
  L.mp = round_up(m, 256);
  L.mt = L.mp / TILE;
  L.ntile = L.mt * (L.mt + 1) / 2;
  // chunk width in columns: the chunk's points rounded up, at least mp (the
  // forward's split-K partials of P are mp columns wide within a row of nbw)
  L.nbw = round_up(n, TILE) < BW_CHUNK ? round_up(n, TILE) : BW_CHUNK;
  if (L.nbw < L.mp) L.nbw = L.mp;
  L.chunk2 = EW;
  L.S2 = (n + EW - 1) / EW;
  const int64_t mp = L.mp;
  const size_t M2 = (size_t)mp * mp * sizeof(double);
  const size_t V = (size_t)mp * sizeof(double);
  const int64_t nleaf = mp / NB;
  size_t o = 0;
  L.off_K = o; o += al(M2);
  L.off_XK = o; o += al(M2);
  L.off_C = o; o += al(M2);
  L.off_XC = o; o += al(M2);
  L.off_P = o; o += al(M2);
  L.off_Kinv = o; o += al(M2);
  L.off_Cinv = o; o += al(M2);
  L.off_G = o; o += al(M2);
  L.off_T1 = o; o += al(M2);
  L.off_T2 = o; o += al(M2);
  L.off_T = o; o += al((size_t)factor_scratch_elems(mp) * sizeof(double));
  L.off_liK = o; o += al((size_t)mp * NB * sizeof(double));
  L.off_ldK = o; o += al((size_t)nleaf * sizeof(double));
  L.off_ifK = o; o += al((size_t)nleaf * sizeof(int));
  L.off_liC = o; o += al((size_t)mp * NB * sizeof(double));
  L.off_ldC = o; o += al((size_t)nleaf * sizeof(double));
  L.off_ifC = o; o += al((size_t)nleaf * sizeof(int));
  L.off_b = o; o += al(V);
  L.off_q = o; o += al(V);
  L.off_c = o; o += al(V);
  L.off_u = o; o += al(V);
  L.off_Pc = o; o += al(V);
  L.off_gzpart = o; o += al((size_t)L.S2 * mp * EDMAX * sizeof(double));
  L.off_tpart = o; o += al((size_t)L.S2 * L.mt * (EDMAX + 1) * sizeof(double));
  L.off_tpart2 = o; o += al((size_t)(mp / 256) * (EDMAX + 1) * sizeof(double));
  L.off_gzkz = o; o += al((size_t)mp * EDMAX * sizeof(double));
  L.off_rpart = o; o += al((size_t)L.nr * 2 * sizeof(double));
  L.off_fpart = o; o += al((size_t)L.nf * sizeof(double));
  L.off_epart = o; o += al((size_t)L.nf * sizeof(double));
  L.off_fs = o; o += al(4 * sizeof(float));
  L.off_vpart = o; o += al((size_t)((mp + 1023) / 1024) * mp * sizeof(double));
  L.off_scal = o; o += al(32 * sizeof(double));
  L.off_kc = o; o += al((size_t)mp * L.nbw * sizeof(double));
  L.off_wt = o; o += al((size_t)L.nbw * mp * sizeof(double));
  L.total = o;
  return L;
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// Storage and product arithmetic of one ELBO call.  GPK_PREC_FP64: Kzx chunks
// in float64, FP64 matrix-core products.  GPK_PREC_MIXED: float32 chunks,
// P += Kc Kc^T with six-product bf16 splits (the ELBO's accuracy rests on P)
// and Kc^T G_P with the fp16 split (it feeds only the gradient; measured
// within 1.5x of the six-product gradient error at twice the speed);
// GPK_PREC_MIXED_ARITH(p, t, l) names the two products' arithmetic instead
// (p: statistics product, t: data-pass product, each 0, 3, 4 or 6; l is not
// used by the ELBO).  The fp16 split (4) scales Kzx by its bound var and G_P by
// its largest entry, found on the device.
struct ElboArith {
  bool f32 = false;
  int fwd = AR_F64, bwd = AR_F64;
};

int decode_elbo_prec(int prec, ElboArith* ea) {
  ElboArith a;
  if (prec == GPK_PREC_FP64) {
    *ea = a;
    return GPK_OK;
  }
  if ((prec & 0xFF) != GPK_PREC_MIXED) return GPK_ERR_BAD_ARG;
  a.f32 = true;
  a.fwd = AR_BF16X6;
  a.bwd = AR_FP16X3;
  if (prec & GPK_PREC_ARITH_FLAG) {
    a.fwd = (prec >> 12) & 0xF;
    a.bwd = (prec >> 16) & 0xF;
    const int l = (prec >> 20) & 0xF;
    auto ok = [](int v) {
      return v == AR_F64 || v == AR_BF16X3 || v == AR_BF16X6 || v == AR_FP16X3;
    };
    if ((prec & ~0xFFF1FF) != 0 || !ok(a.fwd) || !ok(a.bwd) || !ok(l)) return GPK_ERR_BAD_ARG;
  } else if (prec != GPK_PREC_MIXED) {
    return GPK_ERR_BAD_ARG;
  }
  *ea = a;
  return GPK_OK;
}

#define GPK_DMAX_DISPATCH(d, MACRO) \
  do {                              \
    if ((d) <= 1) MACRO(1);         \
    else if ((d) <= 2) MACRO(2);    \
    else if ((d) <= 4) MACRO(4);    \
    else if ((d) <= 8) MACRO(8);    \
    else if ((d) <= 16) MACRO(16);  \
    else MACRO(32);                 \
  } while (0)

int elbo_forward_impl(int kind, const ElboArith& ea, const double* x, const double* y,
                      int64_t n, const double* z, int64_t m, int d, const double* theta, int nls,
                      double jitter, void* ws, const ElboLayout& Lo, double* result,
                      hipStream_t s) {
  const int64_t mp = Lo.mp;
  double* P = at<double>(ws, Lo.off_P);
  double* b = at<double>(ws, Lo.off_b);
  double* qv = at<double>(ws, Lo.off_q);
  double* cvec = at<double>(ws, Lo.off_c);
  double* u = at<double>(ws, Lo.off_u);
  double* vpart = at<double>(ws, Lo.off_vpart);
  // 1. streamed statistics, chunk by chunk: Kc = Kzx for the chunk's points
  //    (evaluated once, zero outside m x ncol), P (+)= Kc Kc^T on the matrix
  //    cores (lower tiles), b (+)= Kc r, q (+)= Kc 1.
  //    The lower tiles alone (136 at M = 2,048) would leave most of the chip
  //    idle, so the GEMM is split over k into S members of one launch (the
  //    batch descriptor shifts every operand by the same z * kz * sizeof(Kc
  //    entry) bytes): member z reads columns z*kz .. +kz of Kc and accumulates
  //    its float64 partial of P at column z*kp (kp = kz for float64 chunks,
  //    kz / 2 for float32 ones; kp >= mp keeps the partials apart) of a
  //    [mp][nbw] float64 buffer (the one that holds Wt in the backward).  S is
  //    the largest power of two up to 64 that keeps kp >= mp: at M = 2,048
  //    and 131,072-point chunks, S = 32 (float32 chunks, 4,352 workgroups, one
  //    resident per CU) or S = 64 (float64, 8,704, two per CU), 17 full rounds
  //    of the chip either way (65,536-point chunks left half a round idle).
  const size_t ke = ea.f32 ? sizeof(float) : sizeof(double);
  double* Pp = at<double>(ws, Lo.off_wt);
  const int64_t nbw = Lo.nbw;
  const int64_t kdiv = ea.f32 ? 2 : 1;
  int64_t S = 1;
  while (S < 64 && (nbw / (2 * S)) % TILE == 0 && nbw / (2 * S) >= kdiv * mp) S *= 2;
  const int64_t kz = nbw / S, kp = kz / kdiv;
  Batch sk;
  sk.ws_stride = kz * (int64_t)ke;
  sk.count = (unsigned)S;
  float* fs = at<float>(ws, Lo.off_fs);
  if (ea.f32 && ea.fwd == AR_FP16X3) {
    hipLaunchKernelGGL(elbo_fp16_scale_kernel, dim3(1), dim3(64), 0, s, theta, nls, nullptr, 0,
                       fs);
    GPK_LAUNCH_CHECK();
  }
  double* T = at<double>(ws, Lo.off_T);
  double* Kb = at<double>(ws, Lo.off_K);
  double* Kinv = at<double>(ws, Lo.off_Kinv);
  Factor<double> FK{Kb, at<double>(ws, Lo.off_XK), mp, T, at<double>(ws, Lo.off_liK),
                    at<double>(ws, Lo.off_ldK), at<int>(ws, Lo.off_ifK), AR_F64, AR_F64, s};
  // The Kzz branch (Kzz + jitter I with identity pad, its factor, inverse factor
  // and Kzz^-1 = X^T X, lower tiles then mirrored) needs nothing from the data:
  // it runs on a library stream (aux stream 0, events 6 and 7) while the chunks
  // stream through, and is joined back into s at the end of this scope (an
  // early error return included), before C is factored: the two factorisations
  // share the scratch T.
  {
    hipStream_t st[AUX_STREAMS];
    hipEvent_t ev[AUX_EVENTS];
    GPK_TRY(aux_streams(st, ev));
    hipStream_t s2 = st[0];
    GPK_HIP_TRY(hipEventRecord(ev[6], s));
    GPK_HIP_TRY(hipStreamWaitEvent(s2, ev[6], 0));
    StreamJoin join{s, s2, ev[7]};
    FK.s = s2;
    GPK_TRY(launch_kernel_matrix<double>(kind, z, m, z, m, d, theta, nls, jitter, 1, Kb, mp, mp,
                                         mp, true, s2));
    GPK_TRY(FK.potrf(0, mp));
    GPK_TRY(FK.trtri(0, mp));
    GPK_TRY(launch_gemm<double, double, double>(AR_F64, true, true, FK.X, mp, FK.X, mp, Kinv, mp,
                                                mp, mp, mp, 1.0, false, true, KR_GE_M, s2));
    hipLaunchKernelGGL(mirror_kernel, dim3(cdiv(mp * mp, 256)), dim3(256), 0, s2, Kinv, mp, mp);
    GPK_LAUNCH_CHECK();
    FK.s = s;
    for (int64_t c0 = 0; c0 < n; c0 += BW_CHUNK) {
      const int64_t ncol = (n - c0 < BW_CHUNK) ? n - c0 : BW_CHUNK;
      if (ea.f32) {
        float* Kc = at<float>(ws, Lo.off_kc);
        GPK_TRY(launch_cross<float>(kind, z, m, x + c0 * d, ncol, d, theta, nls, true, Kc, nbw,
                                    mp, nbw, s));
        GPK_TRY(launch_gemm<float, float, double>(ea.fwd, false, false, Kc, nbw, Kc, nbw, Pp, nbw,
                                                  mp, mp, kz, 1.0, c0 > 0, true, KR_FULL, s,
                                                  nullptr, sk, fs));
        hipLaunchKernelGGL(chunk_rowsum_kernel<float>, dim3((unsigned)mp), dim3(256), 0, s, Kc,
                           nbw, c0, ncol, y, theta, nls, c0 > 0 ? 1 : 0, b, qv);
      } else {
        double* Kc = at<double>(ws, Lo.off_kc);
        GPK_TRY(launch_cross<double>(kind, z, m, x + c0 * d, ncol, d, theta, nls, false, Kc, nbw,
                                     mp, nbw, s));
        GPK_TRY(launch_gemm<double, double, double>(AR_F64, false, false, Kc, nbw, Kc, nbw, Pp,
                                                    nbw, mp, mp, kz, 1.0, c0 > 0, true, KR_FULL, s,
                                                    nullptr, sk));
        hipLaunchKernelGGL(chunk_rowsum_kernel<double>, dim3((unsigned)mp), dim3(256), 0, s, Kc,
                           nbw, c0, ncol, y, theta, nls, c0 > 0 ? 1 : 0, b, qv);
      }
      GPK_LAUNCH_CHECK();
    }
  }  // joins the Kzz branch
  hipLaunchKernelGGL(psum_kernel, dim3(cdiv(mp * mp, 256)), dim3(256), 0, s, Pp, nbw, kp, S, P,
                     mp);
  GPK_LAUNCH_CHECK();
  hipLaunchKernelGGL(rstats_kernel, dim3((unsigned)Lo.nr), dim3(256), 0, s, y, n, theta, nls,
                     at<double>(ws, Lo.off_rpart));
  GPK_LAUNCH_CHECK();
  // 2. C = Kzz + jitter I + P / s with identity pad; factor + invert
  double* Cb = at<double>(ws, Lo.off_C);
  GPK_TRY(launch_kernel_matrix<double>(kind, z, m, z, m, d, theta, nls, jitter, 1, Cb, mp, mp,
                                       mp, true, s));
  hipLaunchKernelGGL(add_p_kernel, dim3(cdiv(mp * mp, 256)), dim3(256), 0, s, Cb, P, mp, mp,
                     theta, nls);
  GPK_LAUNCH_CHECK();
  Factor<double> FC{Cb, at<double>(ws, Lo.off_XC), mp, T, at<double>(ws, Lo.off_liC),
                    at<double>(ws, Lo.off_ldC), at<int>(ws, Lo.off_ifC), AR_F64, AR_F64, s};
  // The Kzz branch has joined, so the look-ahead's second stream is free here.
  const int64_t laC = lookahead_width(mp, false);
  if (laC > 0)
    GPK_TRY(FC.potrf_lookahead(mp, laC));
  else
    GPK_TRY(FC.potrf(0, mp));
  GPK_TRY(FC.trtri(0, mp));
  // 3. c = C^-1 b = X_C^T X_C b
  GPK_TRY(launch_trmv<double>(FC.X, mp, mp, b, u, false, vpart, s));
  GPK_TRY(launch_trmv<double>(FC.X, mp, mp, u, cvec, true, vpart, s));
  // 4. C^-1 = X^T X (lower tiles, then mirrored; Kzz^-1 came with the Kzz branch)
  double* Cinv = at<double>(ws, Lo.off_Cinv);
  GPK_TRY(launch_gemm<double, double, double>(AR_F64, true, true, FC.X, mp, FC.X, mp, Cinv, mp,
                                              mp, mp, mp, 1.0, false, true, KR_GE_M, s));
  hipLaunchKernelGGL(mirror_kernel, dim3(cdiv(mp * mp, 256)), dim3(256), 0, s, Cinv, mp, mp);
  GPK_LAUNCH_CHECK();
  // 5. tr(Kzz^-1 P) and the objective (mixed: with the error model's ||G_P o P||_F)
  double* fpart = at<double>(ws, Lo.off_fpart);
  hipLaunchKernelGGL(frob_kernel, dim3((unsigned)Lo.nf), dim3(256), 0, s, Kinv, P, mp * mp,
                     fpart);
  GPK_LAUNCH_CHECK();
  double* epart = nullptr;
  if (ea.f32) {
    epart = at<double>(ws, Lo.off_epart);
    hipLaunchKernelGGL(gpfrob_kernel, dim3((unsigned)Lo.nf), dim3(256), 0, s, Kinv, Cinv, cvec,
                       P, mp, theta, nls, epart);
    GPK_LAUNCH_CHECK();
  }
  // The error model is calibrated for six-product (or float64) statistics
  // products and six-product, fp16-split (or float64) data-pass products.
  const bool calf = ea.fwd == AR_BF16X6 || ea.fwd == AR_F64;
  const bool calb = ea.bwd == AR_BF16X6 || ea.bwd == AR_FP16X3 || ea.bwd == AR_F64;
  const int uncal = ea.f32 && !(calf && calb) ? 1 : 0;
  hipLaunchKernelGGL(elbo_fwd_finalize, dim3(1), dim3(256), 0, s, at<double>(ws, Lo.off_rpart),
                     (int)Lo.nr, fpart, (int)Lo.nf, at<double>(ws, Lo.off_ldK),
                     at<int>(ws, Lo.off_ifK), at<double>(ws, Lo.off_ldC), at<int>(ws, Lo.off_ifC),
                     mp / NB, b, cvec, mp, n, theta, nls, epart, uncal,
                     at<double>(ws, Lo.off_scal), result);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

int elbo_backward_impl(int kind, const ElboArith& ea, const double* x, const double* y,
                       int64_t n, const double* z, int64_t m, int d, const double* theta, int nls,
                       double gbar, void* ws, const ElboLayout& Lo, double* grad, double* grad_z,
                       hipStream_t s) {
  const int64_t mp = Lo.mp;
  double* P = at<double>(ws, Lo.off_P);
  double* Kinv = at<double>(ws, Lo.off_Kinv);
  double* Cinv = at<double>(ws, Lo.off_Cinv);
  double* cvec = at<double>(ws, Lo.off_c);
  double* G = at<double>(ws, Lo.off_G);
  double* T1 = at<double>(ws, Lo.off_T1);
  double* T2 = at<double>(ws, Lo.off_T2);
  double* scal = at<double>(ws, Lo.off_scal);
  double* fpart = at<double>(ws, Lo.off_fpart);
  double* Pc = at<double>(ws, Lo.off_Pc);
  // backward scalars: tr(C^-1 P), c^T P c
  hipLaunchKernelGGL(frob_kernel, dim3((unsigned)Lo.nf), dim3(256), 0, s, Cinv, P, mp * mp,
                     fpart);
  GPK_LAUNCH_CHECK();
  GPK_TRY(launch_colreduce(P, mp, mp, mp, cvec, Pc, at<double>(ws, Lo.off_vpart), s));
  hipLaunchKernelGGL(elbo_bwd_scalars, dim3(1), dim3(256), 0, s, fpart, (int)Lo.nf, cvec, Pc, mp,
                     scal);
  GPK_LAUNCH_CHECK();
  // The Kzz-side terms (G_K = Kinv - Cinv - c c^T / s^2 - Kinv P Kinv / s into
  // T1, then tr(G_K dKzz) for every theta and Z) need only the forward's
  // matrices: they run on a library stream (aux stream 0, events 6 and 7)
  // during the data pass and are joined back (an early error return included)
  // before the gradients are assembled.
  hipStream_t st[AUX_STREAMS];
  hipEvent_t ev[AUX_EVENTS];
  GPK_TRY(aux_streams(st, ev));
  hipStream_t s2 = st[0];
  GPK_HIP_TRY(hipEventRecord(ev[6], s));
  GPK_HIP_TRY(hipStreamWaitEvent(s2, ev[6], 0));
  {
    StreamJoin join{s, s2, ev[7]};
    GPK_TRY(launch_gemm<double, double, double>(AR_F64, false, false, Kinv, mp, P, mp, T1, mp,
                                                mp, mp, mp, 1.0, false, false, KR_FULL, s2));
    GPK_TRY(launch_gemm<double, double, double>(AR_F64, false, false, T1, mp, Kinv, mp, T2, mp,
                                                mp, mp, mp, 1.0, false, false, KR_FULL, s2));
    hipLaunchKernelGGL(gk_kernel, dim3(cdiv(mp * mp, 256)), dim3(256), 0, s2, T1, Kinv, Cinv, T2,
                       cvec, mp, mp, theta, nls);
    GPK_LAUNCH_CHECK();
#define GPK_E4(DM)                                                                              \
  hipLaunchKernelGGL(elbo_kzz_adj_kernel<DM>, dim3((unsigned)(mp / 256)), dim3(256), 0, s2, kind, \
                     z, m, d, theta, nls, T1, mp, at<double>(ws, Lo.off_gzkz),                   \
                     at<double>(ws, Lo.off_tpart2))
    GPK_DMAX_DISPATCH(d, GPK_E4);
#undef GPK_E4
    GPK_LAUNCH_CHECK();
    // G_P and the data pass
    // the float32 G_P of the mixed data pass lives in XK (the inverse factor of
    // Kzz, which no backward step reads); the fp16 split's max |G_P| partials in
    // epart (free after the forward)
    float* Gf = ea.f32 ? at<float>(ws, Lo.off_XK) : nullptr;
    const bool f16 = ea.f32 && ea.bwd == AR_FP16X3;
    double* gmax = f16 ? at<double>(ws, Lo.off_epart) : nullptr;
    float* fs = at<float>(ws, Lo.off_fs);
    hipLaunchKernelGGL(gp_kernel, dim3((unsigned)Lo.nf), dim3(256), 0, s, G, Gf, Kinv, Cinv, cvec,
                       mp, theta, nls, gmax);
    GPK_LAUNCH_CHECK();
    if (f16) {
      hipLaunchKernelGGL(elbo_fp16_scale_kernel, dim3(1), dim3(64), 0, s, theta, nls, gmax,
                         (int)Lo.nf, fs);
      GPK_LAUNCH_CHECK();
    }
    // Data pass, chunk by chunk: Kc = Kzx for the chunk's points (zero outside
    // m x ncol), Wt = Kc^T G_P on the matrix cores (float64 sums, stored in
    // float64, or float32 in the mixed path), then the elementwise contraction
    // against dKzx.
    double* Wt = at<double>(ws, Lo.off_wt);
    float* Wtf = at<float>(ws, Lo.off_wt);
    for (int64_t c0 = 0; c0 < n; c0 += BW_CHUNK) {
      const int64_t ncol = (n - c0 < BW_CHUNK) ? n - c0 : BW_CHUNK;
      const int64_t ncp = round_up(ncol, TILE);
      if (ea.f32) {
        float* Kc = at<float>(ws, Lo.off_kc);
        GPK_TRY(launch_cross<float>(kind, z, m, x + c0 * d, ncol, d, theta, nls, true, Kc, ncp,
                                    mp, ncp, s));
        GPK_TRY(launch_gemm<float, float, float>(ea.bwd, true, false, Kc, ncp, Gf, mp, Wtf, mp,
                                                 ncp, mp, mp, 1.0, false, false, KR_FULL, s,
                                                 nullptr, Batch(), fs + 2));
      } else {
        double* Kc = at<double>(ws, Lo.off_kc);
        GPK_TRY(launch_cross<double>(kind, z, m, x + c0 * d, ncol, d, theta, nls, false, Kc, ncp,
                                     mp, ncp, s));
        GPK_TRY(launch_gemm<double, double, double>(AR_F64, true, false, Kc, ncp, G, mp, Wt, mp,
                                                    ncp, mp, mp, 1.0, false, false, KR_FULL, s));
      }
      dim3 grid((unsigned)Lo.mt, (unsigned)((ncol + EW - 1) / EW));
#define GPK_E3(DM)                                                                         \
  do {                                                                                     \
    if (ea.f32)                                                                            \
      hipLaunchKernelGGL((elbo_wgrad_kernel<DM, float>), grid, dim3(256), 0, s, kind, x, y, \
                         z, m, d, theta, nls, Wtf, mp, c0, ncol, cvec,                     \
                         at<double>(ws, Lo.off_gzpart), at<double>(ws, Lo.off_tpart));     \
    else                                                                                   \
      hipLaunchKernelGGL((elbo_wgrad_kernel<DM, double>), grid, dim3(256), 0, s, kind, x,  \
                         y, z, m, d, theta, nls, Wt, mp, c0, ncol, cvec,                   \
                         at<double>(ws, Lo.off_gzpart), at<double>(ws, Lo.off_tpart));     \
  } while (0)
      GPK_DMAX_DISPATCH(d, GPK_E3);
#undef GPK_E3
      GPK_LAUNCH_CHECK();
    }
  }  // joins the Kzz-side terms
  int dmax = 32;
  if (d <= 1) dmax = 1;
  else if (d <= 2) dmax = 2;
  else if (d <= 4) dmax = 4;
  else if (d <= 8) dmax = 8;
  else if (d <= 16) dmax = 16;
  const unsigned gblocks = (unsigned)(nls + 3) + cdiv(m * d, 256);
  hipLaunchKernelGGL(elbo_grad_finalize, dim3(gblocks), dim3(256), 0, s,
                     at<double>(ws, Lo.off_tpart), Lo.S2 * Lo.mt, at<double>(ws, Lo.off_tpart2),
                     (int64_t)(mp / 256), at<double>(ws, Lo.off_gzpart), Lo.S2,
                     at<double>(ws, Lo.off_gzkz), cvec, at<double>(ws, Lo.off_q), scal, m, mp, d,
                     n, theta, nls, gbar, grad, grad_z, dmax);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

// ea == nullptr: the caller runs float64 only (the collapsed predictive).
int check_elbo(int kind, int prec, const double* x, const double* y, int64_t n, const double* z,
               int64_t m, int64_t d, const double* theta, int nls, ElboArith* ea) {
  if (kind < GPK_KERNEL_RBF || kind > GPK_KERNEL_MATERN52) {
    set_error("unknown kernel kind %d", kind);
    return GPK_ERR_BAD_ARG;
  }
  if (!ea && prec != GPK_PREC_FP64) {
    set_error("collapsed predict: only GPK_PREC_FP64 is implemented");
    return GPK_ERR_BAD_ARG;
  }
  if (ea && decode_elbo_prec(prec, ea) != GPK_OK) {
    set_error("collapsed ELBO: invalid prec 0x%x (GPK_PREC_FP64, GPK_PREC_MIXED or "
              "GPK_PREC_MIXED_ARITH(p, t, l) with p, t, l in {0, 3, 4, 6})", prec);
    return GPK_ERR_BAD_ARG;
  }
  if (!x || !y || !z || !theta) {
    set_error("collapsed ELBO: null input pointer");
    return GPK_ERR_BAD_ARG;
  }
  if (n < 1 || m < 1 || d < 1 || d > EDMAX) {
    set_error("collapsed ELBO: need n, m >= 1 and 1 <= d <= %d", EDMAX);
    return GPK_ERR_BAD_ARG;
  }
  if (nls != 1 && nls != d) {
    set_error("nls must be 1 or d");
    return GPK_ERR_BAD_ARG;
  }
  return GPK_OK;
}

// ---- predictive of the collapsed posterior (variational_families.py:786-870) ----
// With C = Kzz + Kzx Kxz / s = Lz (I + A A^T) Lz^T (A = Lz^-1 Kzx / sd, Kzz with
// jitter), the reference's quantities are
//   mean = c + Ktz C^-1 Kzx (y - c) / s                      (:840-856)
//   cov  = Ktt - |Lz^-1 Kzt|^2 + |L_C^-1 Kzt|^2 + jitter I   (:850-866)
// where |M|^2 = M^T M and L_C^-1 = L^-1 Lz^-1 (L = chol(I + A A^T)).  The
// forward leaves X_K = Lz^-1, X_C = L_C^-1 and C^-1 b (b = Kzx (y - c)) in the
// workspace; Kzt is formed once, zero-padded to mp x tp.
// tc: test points per pass (all of them for the full covariance; in the
// diagonal mode passes whose three mp x tc buffers hold at most 2^26 doubles
// each, so T is bounded only by the caller's output vectors).
constexpr int64_t EPRED_ELEMS = (int64_t)1 << 26;

struct ElboPredLayout {
  int64_t tp = 0, tc = 0;
  size_t off_kzt = 0, off_a1 = 0, off_a2 = 0, off_ktt = 0, off_part = 0, off_v1 = 0, off_v2 = 0;
  size_t off_res = 0, total = 0;
};

ElboPredLayout make_elbo_pred_layout(const ElboLayout& Lo, int64_t t, int full_cov) {
  ElboPredLayout P;
  P.tp = round_up(t, NB);
  const int64_t mp = Lo.mp, tp = P.tp;
  int64_t tc = tp;
  if (!full_cov) {
    tc = (EPRED_ELEMS / mp) / NB * NB;
    if (tc < NB) tc = NB;
    if (tc > tp) tc = tp;
  }
  P.tc = tc;
  size_t o = Lo.total;
  P.off_kzt = o; o += al((size_t)mp * tc * sizeof(double));
  P.off_a1 = o; o += al((size_t)mp * tc * sizeof(double));
  P.off_a2 = o; o += al((size_t)mp * tc * sizeof(double));
  P.off_ktt = o; o += al((size_t)(full_cov ? tp * tp : 1) * sizeof(double));
  P.off_part = o; o += al((size_t)((mp + 1023) / 1024) * tc * sizeof(double));
  P.off_v1 = o; o += al((size_t)tc * sizeof(double));
  P.off_v2 = o; o += al((size_t)tc * sizeof(double));
  P.off_res = o; o += al(GPK_RESULT_LEN * sizeof(double));
  P.total = o;
  return P;
}

__global__ void sgpr_mean_kernel(const double* __restrict__ v, const double* __restrict__ theta,
                                 int nls, int64_t t, double* __restrict__ mean) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double sd = theta[nls + 1];
  if (j < t) mean[j] = theta[nls + 2] + v[j] / (sd * sd);
}

// var_j = k(t_j, t_j) + jitter - |Lz^-1 Kzt|^2_j + |L_C^-1 Kzt|^2_j; k(0) = variance.
__global__ void sgpr_var_kernel(const double* __restrict__ v1, const double* __restrict__ v2,
                                const double* __restrict__ theta, int nls, double jitter,
                                int64_t t, double* __restrict__ var_out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < t) var_out[j] = (theta[nls] + jitter) - v1[j] + v2[j];
}

int elbo_predict_impl(int kind, const double* x, const double* y, int64_t n, const double* z,
                      int64_t m, int d, const double* xt, int64_t t, const double* theta, int nls,
                      double jitter, int full_cov, void* ws, const ElboLayout& Lo,
                      const ElboPredLayout& P, double* mean, double* cov, hipStream_t s) {
  const int64_t mp = Lo.mp, tp = P.tp;
  GPK_TRY(elbo_forward_impl(kind, ElboArith(), x, y, n, z, m, d, theta, nls, jitter, ws, Lo,
                            at<double>(ws, P.off_res), s));
  double* kzt = at<double>(ws, P.off_kzt);
  double* a1 = at<double>(ws, P.off_a1);
  double* a2 = at<double>(ws, P.off_a2);
  double* part = at<double>(ws, P.off_part);
  double* v1 = at<double>(ws, P.off_v1);
  double* v2 = at<double>(ws, P.off_v2);
  const double* XK = at<double>(ws, Lo.off_XK);
  const double* XC = at<double>(ws, Lo.off_XC);
  if (!full_cov) {
    // Diagonal only, the test points in passes of tc.
    for (int64_t t0 = 0; t0 < t; t0 += P.tc) {
      const int64_t tn = (t - t0 < P.tc) ? t - t0 : P.tc;
      const int64_t tnp = round_up(tn, NB);
      GPK_TRY(launch_cross<double>(kind, z, m, xt + t0 * d, tn, d, theta, nls, false, kzt, tnp,
                                   mp, tnp, s));
      GPK_TRY(launch_colreduce(kzt, tnp, mp, tnp, at<double>(ws, Lo.off_c), v1, part, s));
      hipLaunchKernelGGL(sgpr_mean_kernel, dim3(cdiv(tn, 256)), dim3(256), 0, s, v1, theta, nls,
                         tn, mean + t0);
      GPK_LAUNCH_CHECK();
      GPK_TRY(launch_gemm<double, double, double>(AR_F64, false, true, XK, mp, kzt, tnp, a1, tnp,
                                                  mp, tnp, mp, 1.0, false, false, KR_LE_M, s));
      GPK_TRY(launch_gemm<double, double, double>(AR_F64, false, true, XC, mp, kzt, tnp, a2, tnp,
                                                  mp, tnp, mp, 1.0, false, false, KR_LE_M, s));
      GPK_TRY(launch_colreduce(a1, tnp, mp, tnp, nullptr, v1, part, s));
      GPK_TRY(launch_colreduce(a2, tnp, mp, tnp, nullptr, v2, part, s));
      hipLaunchKernelGGL(sgpr_var_kernel, dim3(cdiv(tn, 256)), dim3(256), 0, s, v1, v2, theta,
                         nls, jitter, tn, cov + t0);
      GPK_LAUNCH_CHECK();
    }
    return GPK_OK;
  }
  GPK_TRY(launch_cross<double>(kind, z, m, xt, t, d, theta, nls, false, kzt, tp, mp, tp, s));
  // mean: c + Kzt^T (C^-1 b) / s
  GPK_TRY(launch_colreduce(kzt, tp, mp, tp, at<double>(ws, Lo.off_c), v1, part, s));
  hipLaunchKernelGGL(sgpr_mean_kernel, dim3(cdiv(t, 256)), dim3(256), 0, s, v1, theta, nls, t,
                     mean);
  GPK_LAUNCH_CHECK();
  // A1 = Lz^-1 Kzt, A2 = L_C^-1 Kzt (lower-triangular left factors: k <= row tile)
  GPK_TRY(launch_gemm<double, double, double>(AR_F64, false, true, XK, mp, kzt, tp, a1, tp, mp,
                                              tp, mp, 1.0, false, false, KR_LE_M, s));
  GPK_TRY(launch_gemm<double, double, double>(AR_F64, false, true, XC, mp, kzt, tp, a2, tp, mp,
                                              tp, mp, 1.0, false, false, KR_LE_M, s));
  {
    double* ktt = at<double>(ws, P.off_ktt);
    GPK_TRY(launch_kernel_matrix<double>(kind, xt, t, xt, t, d, theta, nls, jitter,
                                         1 | KM_ZERO_PAD, ktt, tp, tp, tp, false, s));
    GPK_TRY(launch_gemm<double, double, double>(AR_F64, true, true, a1, tp, a1, tp, ktt, tp, tp,
                                                tp, mp, -1.0, true, false, KR_FULL, s));
    GPK_TRY(launch_gemm<double, double, double>(AR_F64, true, true, a2, tp, a2, tp, ktt, tp, tp,
                                                tp, mp, 1.0, true, false, KR_FULL, s));
    GPK_HIP_TRY(hipMemcpy2DAsync(cov, (size_t)t * sizeof(double), ktt, (size_t)tp * sizeof(double),
                                 (size_t)t * sizeof(double), (size_t)t, hipMemcpyDeviceToDevice,
                                 s));
  }
  return GPK_OK;
}

}  // namespace
}  // namespace gpk

using namespace gpk;

extern "C" {

size_t gpk_elbo_workspace_bytes(int64_t n, int64_t m, int64_t d, int prec) {
  (void)d;
  (void)prec;
  if (n < 1 || m < 1) return 0;
  return make_elbo_layout(n, m).total;
}

int gpk_elbo_forward(int kind, int prec, const double* x, const double* y, int64_t n,
                     const double* z, int64_t m, int64_t d, const double* theta, int nls,
                     double jitter, void* ws, size_t ws_bytes, double* result, void* stream) {
  ElboArith ea;
  int rc = check_elbo(kind, prec, x, y, n, z, m, d, theta, nls, &ea);
  if (rc) return rc;
  if (!ws || !result) {
    set_error("collapsed ELBO: null workspace/result");
    return GPK_ERR_BAD_ARG;
  }
  const ElboLayout Lo = make_elbo_layout(n, m);
  if (ws_bytes < Lo.total) {
    set_error("collapsed ELBO: workspace %zu bytes < required %zu", ws_bytes, Lo.total);
    return GPK_ERR_WORKSPACE;
  }
  if (reinterpret_cast<uintptr_t>(ws) % 256) {
    set_error("collapsed ELBO: workspace must be 256-byte aligned");
    return GPK_ERR_BAD_ARG;
  }
  return elbo_forward_impl(kind, ea, x, y, n, z, m, (int)d, theta, nls, jitter, ws, Lo, result,
                           (hipStream_t)stream);
}

int gpk_elbo_backward(int kind, int prec, const double* x, const double* y, int64_t n,
                      const double* z, int64_t m, int64_t d, const double* theta, int nls,
                      double gbar_host, void* ws, size_t ws_bytes, double* grad, double* grad_z,
                      void* stream) {
  ElboArith ea;
  int rc = check_elbo(kind, prec, x, y, n, z, m, d, theta, nls, &ea);
  if (rc) return rc;
  if (!ws || !grad || !grad_z) {
    set_error("collapsed ELBO: null workspace/grad");
    return GPK_ERR_BAD_ARG;
  }
  const ElboLayout Lo = make_elbo_layout(n, m);
  if (ws_bytes < Lo.total) {
    set_error("collapsed ELBO: workspace %zu bytes < required %zu", ws_bytes, Lo.total);
    return GPK_ERR_WORKSPACE;
  }
  if (reinterpret_cast<uintptr_t>(ws) % 256) {
    set_error("collapsed ELBO: workspace must be 256-byte aligned");
    return GPK_ERR_BAD_ARG;
  }
  return elbo_backward_impl(kind, ea, x, y, n, z, m, (int)d, theta, nls, gbar_host, ws, Lo,
                            grad, grad_z, (hipStream_t)stream);
}

size_t gpk_elbo_predict_workspace_bytes(int64_t n, int64_t m, int64_t d, int64_t t,
                                        int full_cov) {
  (void)d;
  if (n < 1 || m < 1 || t < 1) return 0;
  return make_elbo_pred_layout(make_elbo_layout(n, m), t, full_cov).total;
}

int gpk_elbo_predict(int kind, int prec, const double* x, const double* y, int64_t n,
                     const double* z, int64_t m, int64_t d, const double* xt, int64_t t,
                     const double* theta, int nls, double jitter, int full_cov, void* ws,
                     size_t ws_bytes, double* mean, double* cov, void* stream) {
  int rc = check_elbo(kind, prec, x, y, n, z, m, d, theta, nls, nullptr);
  if (rc) return rc;
  if (!xt || t < 1 || !ws || !mean || !cov) {
    set_error("collapsed predict: null xt/ws/mean/cov or t < 1");
    return GPK_ERR_BAD_ARG;
  }
  const ElboLayout Lo = make_elbo_layout(n, m);
  const ElboPredLayout P = make_elbo_pred_layout(Lo, t, full_cov);
  if (ws_bytes < P.total) {
    set_error("collapsed predict: workspace %zu bytes < required %zu", ws_bytes, P.total);
    return GPK_ERR_WORKSPACE;
  }
  if (reinterpret_cast<uintptr_t>(ws) % 256) {
    set_error("collapsed predict: workspace must be 256-byte aligned");
    return GPK_ERR_BAD_ARG;
  }
  return elbo_predict_impl(kind, x, y, n, z, m, (int)d, xt, t, theta, nls, jitter, full_cov, ws,
                           Lo, P, mean, cov, (hipStream_t)stream);
}

}  // extern "C"
