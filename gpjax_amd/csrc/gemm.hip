// Tile GEMM used by the recursive Cholesky, triangular solve and triangular
// inverse:  C = alpha * A B^T (+ C), with per-tile k-ranges for triangular
// operands and an optional lower-triangular tile set (SYRK).  These replace the
// LAPACK/BLAS work behind jnp.linalg.cholesky / solve / slogdet
// (gpjax/linalg/operations.py:54-55,109-111,163-165).
#include "gemm_core.h"

namespace gpk {
namespace {

struct GemmParams {
  const void* A;
  int64_t lda;
  const void* B;
  int64_t ldb;
  void* C;
  int64_t ldc;
  int64_t K;
  double alpha;
  int beta_one;
  int lower_tiles;
  int krange;
  double* cdiag;  // lower_tiles: diagonal tiles update this float64 [tile][NB][NB] instead of C
  Batch bt;       // batch of same-shape problems, member blockIdx.z (operands in workspaces)
  const float* fscale;  // AR_FP16X3: power-of-two scales of A and B (device)
  int64_t tiles_m;      // M / TILE (the lower-tile decode needs the side)
};

template <int AR, typename TA, typename TB, typename TC, bool A_KM, bool B_KN>
__global__ __launch_bounds__(WaveLayout<AR>::NT) void gemm_kernel(GemmParams p) {
  typedef WaveLayout<AR> WL;
  __shared__ __align__(16) char smem[smem_bytes<AR>()];
  p.A = bws(p.A, p.bt);
  p.B = bws(p.B, p.bt);
  p.C = bws(p.C, p.bt);
  p.cdiag = bws(p.cdiag, p.bt);
  // Banded tile orders (16 x 16 blocks of tiles resident together share their
  // operand panels in L2).
  int64_t mt, nt;
  if (p.lower_tiles) {
    lower_tile_banded(blockIdx.x, p.tiles_m, mt, nt);
  } else {
    const int64_t b = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    if (p.krange == KR_GE_N) {
      // k >= column tile: the first tile columns carry the most work, so band
      // by columns and dispatch them first.
      grid_tile_banded(b, gridDim.x, gridDim.y, nt, mt);
    } else {
      grid_tile_banded(b, gridDim.y, gridDim.x, mt, nt);
      // k <= row tile: the last tile rows carry the most work, so dispatch them first.
      if (p.krange == KR_LE_M) mt = gridDim.y - 1 - mt;
    }
  }
  int64_t kb, ke;
  k_range(p.krange, mt, nt, p.K, kb, ke);
  const int64_t m0 = mt * TILE, n0 = nt * TILE;
  typename AccT<AR>::vec acc[4][WL::NJ];
  if constexpr (AR == AR_F64) {
    mainloop_f64<TA, TB, A_KM, B_KN>(static_cast<const TA*>(p.A), p.lda,
                                     static_cast<const TB*>(p.B), p.ldb, m0, n0, kb, ke, acc,
                                     smem);
  } else {
    float sa = 1.f, sb = 1.f;
    if constexpr (AR == AR_FP16X3) {
      sa = p.fscale[0];
      sb = p.fscale[1];
    }
    mainloop_bf16<AR, A_KM, B_KN>(static_cast<const float*>(p.A), p.lda,
                                  static_cast<const float*>(p.B), p.ldb, m0, n0, kb, ke, acc,
                                  smem, sa, sb);
  }
  TC* C = static_cast<TC*>(p.C);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = w / WL::WN, wn = w % WL::WN;
  if (p.cdiag && p.lower_tiles && mt == nt) {
    double* cd = p.cdiag + mt * TILE * TILE;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < WL::NJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = wm * 64 + i * 16 + acc_row<AR>(lane, r);
          const int colx = wn * WL::WC + j * 16 + (lane & 15);
          double* c = cd + row * TILE + colx;
          double v = p.alpha * (double)acc[i][j][r];
          if (p.beta_one) v += *c;
          *c = v;
        }
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < WL::NJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 64 + i * 16 + acc_row<AR>(lane, r);
        const int64_t colx = n0 + wn * WL::WC + j * 16 + (lane & 15);
        TC* c = C + row * p.ldc + colx;
        double v = p.alpha * (double)acc[i][j][r];
        if (p.beta_one) v += (double)*c;
        *c = (TC)v;
      }
}

template <int AR, typename TA, typename TB, typename TC, bool A_KM, bool B_KN>
int launch_t(const GemmParams& p, int64_t M, int64_t N, hipStream_t s) {
  dim3 grid;
  if (p.lower_tiles) {
    const int64_t nt = M / TILE;
    grid = dim3((unsigned)(nt * (nt + 1) / 2), 1, p.bt.count);
  } else {
    grid = dim3((unsigned)(N / TILE), (unsigned)(M / TILE), p.bt.count);
  }
  hipLaunchKernelGGL((gemm_kernel<AR, TA, TB, TC, A_KM, B_KN>), grid,
                     dim3(WaveLayout<AR>::NT), 0, s, p);
  GPK_LAUNCH_CHECK();
  return GPK_OK;
}

template <int AR, typename TA, typename TB, typename TC>
int dispatch_layout(bool a_km, bool b_kn, const GemmParams& p, int64_t M, int64_t N,
                    hipStream_t s) {
  if (!a_km && !b_kn) return launch_t<AR, TA, TB, TC, false, false>(p, M, N, s);
  if (!a_km && b_kn) return launch_t<AR, TA, TB, TC, false, true>(p, M, N, s);
  if (a_km && !b_kn) return launch_t<AR, TA, TB, TC, true, false>(p, M, N, s);
  return launch_t<AR, TA, TB, TC, true, true>(p, M, N, s);
}

}  // namespace

template <typename TA, typename TB, typename TC>
int launch_gemm(int arith, bool a_km, bool b_kn, const TA* A, int64_t lda, const TB* B,
                int64_t ldb, TC* C, int64_t ldc, int64_t M, int64_t N, int64_t K, double alpha,
                bool beta_one, bool lower_tiles, int krange, hipStream_t s, double* cdiag,
                const Batch& bt, const float* fscale) {
  if (M <= 0 || N <= 0) return GPK_OK;
  if (M % TILE || N % TILE || K % TILE || K < 0) {
    set_error("gemm: extents %lld x %lld x %lld not multiples of %d", (long long)M,
              (long long)N, (long long)K, TILE);
    return GPK_ERR_BAD_ARG;
  }
  if (lower_tiles && M != N) {
    set_error("gemm: lower tile set needs M == N");
    return GPK_ERR_BAD_ARG;
  }
  if (!lower_tiles && M / TILE > 65535) {
    set_error("gemm: M too large for one launch");
    return GPK_ERR_BAD_ARG;
  }
  GemmParams p{A, lda, B, ldb, C, ldc, K, alpha, beta_one ? 1 : 0, lower_tiles ? 1 : 0, krange,
               cdiag, bt, fscale, M / TILE};
  if (arith == AR_F64) return dispatch_layout<AR_F64, TA, TB, TC>(a_km, b_kn, p, M, N, s);
  if constexpr (std::is_same<TA, float>::value && std::is_same<TB, float>::value) {
    if (arith == AR_BF16X3)
      return dispatch_layout<AR_BF16X3, float, float, TC>(a_km, b_kn, p, M, N, s);
    if (arith == AR_BF16X6)
      return dispatch_layout<AR_BF16X6, float, float, TC>(a_km, b_kn, p, M, N, s);
    if (arith == AR_FP16X3) {
      if (!fscale) {
        set_error("gemm: fp16 products need the operand scales");
        return GPK_ERR_BAD_ARG;
      }
      return dispatch_layout<AR_FP16X3, float, float, TC>(a_km, b_kn, p, M, N, s);
    }
  }
  set_error("gemm: unsupported arithmetic %d for these operand types", arith);
  return GPK_ERR_BAD_ARG;
}

#define GPK_GEMM_INST(TA, TB, TC)                                                            \
  template int launch_gemm<TA, TB, TC>(int, bool, bool, const TA*, int64_t, const TB*,       \
                                       int64_t, TC*, int64_t, int64_t, int64_t, int64_t, double, \
                                       bool, bool, int, hipStream_t, double*, const Batch&, \
                                       const float*);
GPK_GEMM_INST(double, double, double)
GPK_GEMM_INST(float, float, float)
GPK_GEMM_INST(float, double, float)
GPK_GEMM_INST(float, double, double)
GPK_GEMM_INST(float, float, double)
#undef GPK_GEMM_INST

}  // namespace gpk
