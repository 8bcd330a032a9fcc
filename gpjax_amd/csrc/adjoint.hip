// Fused LAUUM + adjoint trace pass for d MLL / d theta.
//
// The reference obtains the gradient from jax.grad through the LU slogdet and
// solve (gpjax/objectives.py:96-107, gpjax/linalg/operations.py:111,164; used as
// in tests/test_objectives.py:83-85 and gpjax/fit.py:160).  That derivative is
//   d MLL / d theta = 1/2 tr(W dSigma/dtheta),   W = alpha alpha^T - Sigma^-1.
// Here every NB-tile (I, J), I >= J, of Sigma^-1 = X^T X (X = L^-1, lower) is
// accumulated on the matrix cores over k >= I*NB, and in the epilogue of the
// same workgroup W_ij is combined with the kernel tile recomputed from the
// inputs:
//   S_var  += wt * W_ij * k(rho_ij)
//   S_ls_d += wt * W_ij * F(rho_ij) * u_d,    u_d = (x_id - x_jd)^2 / l_d^2
//   S_tr   += W_ii
// with wt = 2 off the diagonal tile (symmetry) and 1 on it.  Sigma^-1 and
// dK/dtheta are never written to memory.  Each tile writes its nls + 2 sums to
// `partial`; a deterministic reduction follows.
//
// The same pass serves any objective whose derivative has the form
// 1/2 tr(W dSigma) with W = alpha beta^T + beta alpha^T - Y^T Y: `beta` set
// selects that form (beta null: alpha alpha^T), and `dense` makes Y a dense
// matrix (k over all rows) instead of the lower-triangular X.  The
// leave-one-out objective uses it with Y = diag(c) Sigma^-1.
#include "gemm_core.h"

namespace gpk {
namespace {

constexpr int ADJ_DCH = 32;  // input dimensions staged per pass in the epilogue

template <int AR, typename TS, int DMAX, bool XT>
__global__ __launch_bounds__(WaveLayout<AR>::NT) void adjoint_kernel(int kind,
                                                                     const TS* __restrict__ X,
                                                      int64_t ld, int64_t np,
                                                      const double* __restrict__ x, int64_t n,
                                                      int d, const double* __restrict__ theta,
                                                      int nls, const double* __restrict__ alpha,
                                                      const double* __restrict__ beta, int dense,
                                                      const float* __restrict__ fscale,
                                                      double* __restrict__ partial, Batch bt) {
  typedef typename AccT<AR>::scalar AS;
  typedef WaveLayout<AR> WL;
  constexpr int NS = WL::NJ * 4;  // accumulator slots per lane and 16-row block
  bmember(bt, d, x, n, theta);
  X = bws(X, bt);
  alpha = bws(alpha, bt);
  beta = bws(beta, bt);
  partial = bws(partial, bt);
  // Inputs wider than ADJ_DCH are staged in chunks of ADJ_DCH dimensions.
  constexpr int DS = DMAX > ADJ_DCH ? ADJ_DCH : DMAX;
  // The fused pass keeps the padded LDS rows for BF16x6 (measured faster here).
  constexpr bool SWZ = AR != AR_BF16X6;
  constexpr size_t ML = smem_bytes<AR, SWZ>();
  constexpr size_t XS = (size_t)2 * TILE * (DS + 1) * sizeof(double);
  constexpr size_t AB = (size_t)4 * TILE * sizeof(double);
  constexpr size_t WB = (size_t)WL::NW * NS * 64 * sizeof(AS) + DMAX * sizeof(double);
  constexpr size_t EP = XS + AB + WB;
  __shared__ __align__(16) char smem[ML > EP ? ML : EP];
  __shared__ double red_s[WL::NW][DS + 2];

  int64_t mt, nt;
  lower_tile_banded(blockIdx.x, np / TILE, mt, nt);
  const int64_t m0 = mt * TILE, n0 = nt * TILE;
  const int64_t kb = dense ? 0 : m0;
  typename AccT<AR>::vec acc[4][WL::NJ];
  // XT: the operand is U = X^T (upper, row-major), so both panels are read
  // along k with 16-byte loads; otherwise X itself, read down its columns.
  if constexpr (AR == AR_F64) {
    mainloop_f64<TS, TS, !XT, !XT>(X, ld, X, ld, m0, n0, kb, np, acc, smem);
  } else {
    float sx = 1.f;
    if constexpr (AR == AR_FP16X3) sx = fscale[0];
    mainloop_bf16<AR, !XT, !XT, true, false, SWZ>(X, ld, X, ld, m0, n0, kb, np, acc, smem, sx,
                                                  sx);
  }
  __syncthreads();

  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int wm = w / WL::WN, wn = w % WL::WN;
  const double wt = (mt == nt) ? 1.0 : 2.0;
  const bool ard = nls > 1;
  const int nv = nls + 2;
  // Epilogue staging: scaled inputs (DS dimensions at a time) and alpha of the
  // tile's rows (i) and columns (j), and a per-lane strip through which each
  // lane re-reads its own accumulators with a rolled loop.
  double* xm = reinterpret_cast<double*>(smem);        // [TILE][DS+1]
  double* xn = xm + TILE * (DS + 1);                   // [TILE][DS+1]
  double* am = xn + TILE * (DS + 1);                   // [TILE]
  double* an = am + TILE;                              // [TILE]
  double* bm = an + TILE;                              // [TILE]
  double* bn = bm + TILE;                              // [TILE]
  AS* wbuf = reinterpret_cast<AS*>(bn + TILE);         // [waves][NS][64 lanes]
  auto stage = [&](int dc) {
    for (int e = t; e < TILE * DS; e += WL::NT) {
      const int r = e / DS, q = e % DS;
      if (dc + q < d) {
        const double l = theta[nls == 1 ? 0 : dc + q];
        const int64_t gi = m0 + r, gj = n0 + r;
        xm[r * (DS + 1) + q] = gi < n ? x[gi * d + dc + q] / l : 0.0;
        xn[r * (DS + 1) + q] = gj < n ? x[gj * d + dc + q] / l : 0.0;
      }
    }
  };
  stage(0);
  if (t < TILE) {
    am[t] = alpha[m0 + t];
    an[t] = alpha[n0 + t];
    if (beta) {
      bm[t] = beta[m0 + t];
      bn[t] = beta[n0 + t];
    }
  }
  __syncthreads();
  AS* mine = wbuf + w * NS * 64 + lane;

  if constexpr (DMAX <= ADJ_DCH) {
    double s_var = 0.0, s_tr = 0.0;
    double s_ls[DMAX];
#pragma unroll
    for (int q = 0; q < DMAX; ++q) s_ls[q] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      // Only this lane reads these NS slots back, so no barrier is needed.
#pragma unroll
      for (int j = 0; j < WL::NJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) mine[(j * 4 + r) * 64] = acc[i][j][r];
#pragma unroll 1
      for (int e = 0; e < NS; ++e) {
        const int j = e >> 2, r = e & 3;
        const int li = wm * 64 + i * 16 + acc_row<AR>(lane, r);
        const int lj = wn * WL::WC + j * 16 + (lane & 15);
        const int64_t gi = m0 + li, gj = n0 + lj;
        if (gi >= n || gj >= n) continue;
        const double wij = (beta ? am[li] * bn[lj] + bm[li] * an[lj] : am[li] * an[lj]) -
                           (double)mine[e * 64];
        const double* pi = xm + li * (DMAX + 1);
        const double* pj = xn + lj * (DMAX + 1);
        double d2 = 0.0;
#pragma unroll
        for (int q = 0; q < DMAX; ++q) {
          if (q < d) {
            const double df = pi[q] - pj[q];
            d2 += df * df;
          }
        }
        double kv, fv;
        kval_fac(kind, d2, kv, fv);
        const double g = wt * wij;
        s_var += g * kv;
        if (gi == gj) s_tr += wij;
        const double gf = g * fv;
        if (ard) {
#pragma unroll
          for (int q = 0; q < DMAX; ++q) {
            if (q < d) {
              const double df = pi[q] - pj[q];
              s_ls[q] += gf * df * df;
            }
          }
        } else {
          s_ls[0] += gf * d2;
        }
      }
    }

    // Workgroup reduction: wave shuffles, then across the waves in order.
#pragma unroll
    for (int q = 0; q < DMAX + 2; ++q) {
      double v = q < DMAX ? s_ls[q] : (q == DMAX ? s_var : s_tr);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0) red_s[w][q] = v;
    }
    __syncthreads();
    if (t < nv) {
      const int q = t < nls ? t : (t == nls ? DMAX : DMAX + 1);
      double v = red_s[0][q];
#pragma unroll
      for (int u = 1; u < WL::NW; ++u) v += red_s[u][q];
      partial[(int64_t)blockIdx.x * nv + t] = v;
    }
  } else {
    // Inputs wider than ADJ_DCH: per 16-row block, the squared distances of the
    // lane's NS elements are summed over the dimension chunks (the chunk staged
    // in turn), then each element's derivative weight gf is formed, and a second
    // sweep over the chunks adds gf * (x_id - x_jd)^2 into per-dimension sums,
    // which are reduced across the workgroup chunk by chunk into LDS (acc_ls).
    double* acc_ls = reinterpret_cast<double*>(wbuf + WL::NW * NS * 64);  // [DMAX]
    if (t < DMAX) acc_ls[t] = 0.0;
    double s_var = 0.0, s_tr = 0.0, s_iso = 0.0;
    const int nch = (d + DS - 1) / DS;
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int j = 0; j < WL::NJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) mine[(j * 4 + r) * 64] = acc[i][j][r];
      double d2[NS], gf[NS];
#pragma unroll
      for (int e = 0; e < NS; ++e) d2[e] = 0.0;
      for (int c = 0; c < nch; ++c) {
        if (c > 0 || i > 0) {
          __syncthreads();
          stage(c * DS);
          __syncthreads();
        }
#pragma unroll
        for (int e = 0; e < NS; ++e) {
          const int li = wm * 64 + i * 16 + acc_row<AR>(lane, e & 3);
          const int lj = wn * WL::WC + (e >> 2) * 16 + (lane & 15);
          const double* pi = xm + li * (DS + 1);
          const double* pj = xn + lj * (DS + 1);
#pragma unroll
          for (int q = 0; q < DS; ++q) {
            if (c * DS + q < d) {
              const double df = pi[q] - pj[q];
              d2[e] += df * df;
            }
          }
        }
      }
#pragma unroll
      for (int e = 0; e < NS; ++e) {
        const int li = wm * 64 + i * 16 + acc_row<AR>(lane, e & 3);
        const int lj = wn * WL::WC + (e >> 2) * 16 + (lane & 15);
        const int64_t gi = m0 + li, gj = n0 + lj;
        gf[e] = 0.0;
        if (gi >= n || gj >= n) continue;
        const double wij = (beta ? am[li] * bn[lj] + bm[li] * an[lj] : am[li] * an[lj]) -
                           (double)mine[e * 64];
        double kv, fv;
        kval_fac(kind, d2[e], kv, fv);
        const double g = wt * wij;
        s_var += g * kv;
        if (gi == gj) s_tr += wij;
        gf[e] = g * fv;
        s_iso += gf[e] * d2[e];
      }
      if (!ard) continue;
      for (int c = nch - 1; c >= 0; --c) {
        // the last chunk staged is still in LDS; the others are staged again
        if (c != nch - 1) {
          __syncthreads();
          stage(c * DS);
          __syncthreads();
        }
        double s_ls[DS];
#pragma unroll
        for (int q = 0; q < DS; ++q) s_ls[q] = 0.0;
#pragma unroll
        for (int e = 0; e < NS; ++e) {
          const int li = wm * 64 + i * 16 + acc_row<AR>(lane, e & 3);
          const int lj = wn * WL::WC + (e >> 2) * 16 + (lane & 15);
          const double* pi = xm + li * (DS + 1);
          const double* pj = xn + lj * (DS + 1);
#pragma unroll
          for (int q = 0; q < DS; ++q) {
            const double df = pi[q] - pj[q];
            s_ls[q] += gf[e] * df * df;
          }
        }
#pragma unroll
        for (int q = 0; q < DS; ++q) {
          double v = s_ls[q];
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
          if (lane == 0) red_s[w][q] = v;
        }
        __syncthreads();
        if (t < DS && c * DS + t < d) {
          double v = red_s[0][t];
#pragma unroll
          for (int u = 1; u < WL::NW; ++u) v += red_s[u][t];
          acc_ls[c * DS + t] += v;
        }
      }
    }
    // variance and trace (and the isotropic lengthscale) sums
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      double v = q == 0 ? s_var : (q == 1 ? s_tr : s_iso);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0) red_s[w][q] = v;
    }
    __syncthreads();
    if (t < nv) {
      double v;
      if (t < nls) {
        if (ard) {
          v = acc_ls[t];
        } else {
          v = red_s[0][2];
#pragma unroll
          for (int u = 1; u < WL::NW; ++u) v += red_s[u][2];
        }
      } else {
        const int q = t == nls ? 0 : 1;
        v = red_s[0][q];
#pragma unroll
        for (int u = 1; u < WL::NW; ++u) v += red_s[u][q];
      }
      partial[(int64_t)blockIdx.x * nv + t] = v;
    }
  }
}

template <int AR, typename TS>
int launch_ar(int kind, const TS* X, int64_t ld, int64_t np, const double* x, int64_t n, int d,
              const double* theta, int nls, const double* alpha, double* partial,
              hipStream_t s, const Batch& bt, const double* beta, bool dense,
              const float* fscale, bool xt) {
  const int64_t nt = np / TILE;
  const dim3 grid((unsigned)(nt * (nt + 1) / 2), 1, bt.count);
#define GPK_ADJ(DM)                                                                        \
  do {                                                                                     \
    if (xt)                                                                                \
      hipLaunchKernelGGL((adjoint_kernel<AR, TS, DM, true>), grid, dim3(WaveLayout<AR>::NT), \
                         0, s, kind, X, ld, np, x, n, d, theta, nls, alpha, beta,          \
                         dense ? 1 : 0, fscale, partial, bt);                              \
    else                                                                                   \
      hipLaunchKernelGGL((adjoint_kernel<AR, TS, DM, false>), grid,                        \
                         dim3(WaveLayout<AR>::NT), 0, s, kind, X, ld, np, x, n, d, theta,  \
                         nls, alpha, beta, dense ? 1 : 0, fscale, partial, bt);            \
  } while (0)
  if (d <= 1) GPK_ADJ(1);
  else if (d <= 2) GPK_ADJ(2);
  else if (d <= 4) GPK_ADJ(4);
  else if (d <= 8) GPK_ADJ(8);
  else if (d <= 16) GPK_ADJ(16);
  else if (d <= 32) GPK_ADJ(32);
  else GPK_ADJ(64);
#undef GPK_ADJ
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

}  // namespace

template <typename TS>
int launch_adjoint(int kind, int arith, const TS* X, int64_t ld, int64_t np, const double* x,
                   int64_t n, int d, const double* theta, int nls, const double* alpha,
                   double* partial, hipStream_t s, const Batch& bt, const double* beta,
                   bool dense, const float* fscale, bool xt) {
  if (d < 1 || d > MAX_D) {
    set_error("adjoint: input dimension %d outside [1, %d]", d, MAX_D);
    return GPK_ERR_BAD_ARG;
  }
  if (np % TILE) {
    set_error("adjoint: padded size %lld not a multiple of %d", (long long)np, TILE);
    return GPK_ERR_BAD_ARG;
  }
  if (arith == AR_F64)
    return launch_ar<AR_F64, TS>(kind, X, ld, np, x, n, d, theta, nls, alpha, partial, s, bt,
                                 beta, dense, fscale, xt);
  if constexpr (std::is_same<TS, float>::value) {
    if (arith == AR_BF16X3)
      return launch_ar<AR_BF16X3, float>(kind, X, ld, np, x, n, d, theta, nls, alpha, partial, s,
                                         bt, beta, dense, fscale, xt);
    if (arith == AR_BF16X6)
      return launch_ar<AR_BF16X6, float>(kind, X, ld, np, x, n, d, theta, nls, alpha, partial, s,
                                         bt, beta, dense, fscale, xt);
    if (arith == AR_FP16X3) {
      if (!fscale) {
        set_error("adjoint: fp16 products need the operand scale");
        return GPK_ERR_BAD_ARG;
      }
      return launch_ar<AR_FP16X3, float>(kind, X, ld, np, x, n, d, theta, nls, alpha, partial, s,
                                         bt, beta, dense, fscale, xt);
    }
  }
  set_error("adjoint: unsupported arithmetic %d", arith);
  return GPK_ERR_BAD_ARG;
}

template int launch_adjoint<double>(int, int, const double*, int64_t, int64_t, const double*,
                                    int64_t, int, const double*, int, const double*, double*,
                                    hipStream_t, const Batch&, const double*, bool,
                                    const float*, bool);
template int launch_adjoint<float>(int, int, const float*, int64_t, int64_t, const double*,
                                   int64_t, int, const double*, int, const double*, double*,
                                   hipStream_t, const Batch&, const double*, bool,
                                   const float*, bool);

// U = X^T for a lower-triangular X (np x np, lower 128-tiles): every lower
// tile (I, J) of X, the diagonal ones in full (their strict upper part is zero),
// lands as tile (J, I) of U, so U is upper triangular with zeros below the
// diagonal inside its diagonal tiles.  One workgroup per 64 x 64 quarter of a
// tile, staged through LDS so that both the reads and the writes are
// row-contiguous.
namespace {
template <typename TS>
__global__ __launch_bounds__(256) void lower_to_upper_kernel(const TS* __restrict__ X,
                                                             TS* __restrict__ U, int64_t ld,
                                                             Batch bt) {
  __shared__ TS tl[64][65];
  X = bws(X, bt);
  U = bws(U, bt);
  int64_t I, J;
  lower_tile(blockIdx.x, I, J);
  const int64_t r0 = I * TILE + (blockIdx.y >> 1) * 64, c0 = J * TILE + (blockIdx.y & 1) * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, q = e & 63;
    tl[r][q] = X[(r0 + r) * ld + c0 + q];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, q = e & 63;
    U[(c0 + r) * ld + r0 + q] = tl[q][r];
  }
}
}  // namespace

template <typename TS>
int launch_lower_to_upper(const TS* X, TS* U, int64_t ld, int64_t np, hipStream_t s,
                          const Batch& bt) {
  if (np % TILE) {
    set_error("transpose: size %lld not a multiple of %d", (long long)np, TILE);
    return GPK_ERR_BAD_ARG;
  }
  const int64_t nt = np / TILE;
  hipLaunchKernelGGL(lower_to_upper_kernel<TS>, dim3((unsigned)(nt * (nt + 1) / 2), 4, bt.count),
                     dim3(256), 0, s, X, U, ld, bt);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}
template int launch_lower_to_upper<double>(const double*, double*, int64_t, int64_t,
                                           hipStream_t, const Batch&);
template int launch_lower_to_upper<float>(const float*, float*, int64_t, int64_t, hipStream_t,
                                          const Batch&);

}  // namespace gpk
