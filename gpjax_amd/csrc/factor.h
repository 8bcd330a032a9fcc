// Recursive Cholesky / triangular-solve / triangular-inverse driver on the tile
// GEMM, shared by the exact-GP and collapsed-ELBO paths.
//
//   potrf(A):   potrf(A11); L21 = A21 L11^-T (trsm); A22 -= L21 L21^T; potrf(A22)
//   trsm(B, L): Y L^T = B by recursion on the columns, leaves multiplying by the
//               float64 inverse of an NB x NB diagonal block
//   trtri(L):   X = L^-1 with X21 = -X22 (L21 X11), both halves recursively
// These replace jnp.linalg.cholesky, jsp.linalg.solve_triangular and the
// LU-based slogdet/solve behind gpjax/linalg/operations.py:54-55,105-111,163-165.
// Splits fall on multiples of NB; leaves (NB x NB) are factored and inverted
// in float64 in LDS (leaf.hip).
#pragma once

#include "gpk_internal.h"

namespace gpk {

#define GPK_TRY(...)               \
  do {                             \
    int rc_ = (__VA_ARGS__);       \
    if (rc_ != GPK_OK) return rc_; \
  } while (0)

inline int64_t split_point(int64_t n) { return (n / NB / 2) * NB; }

// Panel width of the look-ahead Cholesky (Factor::potrf_lookahead) for a single
// np x np problem, np >= 2,048: np/32 with float32 storage, np/64 in float64,
// rounded down to a multiple of NB, at least NB.  Wider panels put more of the
// update into large rank-B GEMMs; narrower ones shorten the second stream's
// chain per panel, which the first and last panels wait for.  Measured (ms per
// MLL + gradient, float32 storage / float64; * = the rule's choice):
//   np =  2,048: no look-ahead 2.96 / 6.1, B = 256 2.86 / 5.5, 128 2.76* / 5.1*
//   np =  4,096: no look-ahead 6.16 / 14.1, B = 256 5.53 / 11.7, 128 5.31* / 10.8*
//   np =  8,192: B = 2,048 15.8 / 35.1, 512 13.9 / 30.6, 256 12.9* / 28.4,
//                128 13.1 / 27.2*
//   np = 16,384: B = 2,048 50.3 / 123.5, 1,024 45.9 / 117.5, 512 44.4* / 113.1,
//                256 47.3 / 109.0*, 128 57.9 / 114.7
//   np = 32,768: B = 2,048 227.3 / 662.5, 1,024 225.1* / 655.7, 512 239.3 / 652.5*
//   np = 65,536: B = 4,096 1,521 / 4,602, 2,048 1,525* / 4,589, 1,024 1,576 / 4,592*
// GPK_LOOKAHEAD_B set at build time fixes the width (0 turns the look-ahead off).
#ifndef GPK_LOOKAHEAD_B
#define GPK_LOOKAHEAD_B -1
#endif
inline int64_t lookahead_width(int64_t np, bool f32) {
  if (GPK_LOOKAHEAD_B >= 0) return np >= 4 * GPK_LOOKAHEAD_B ? GPK_LOOKAHEAD_B : 0;
  if (np < 2048) return 0;
  const int64_t b = (np / (f32 ? 32 : 64)) / NB * NB;
  return b < NB ? NB : b;
}

template <typename T>
inline T* at(void* ws, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(ws) + off);
}

template <typename TS>
struct Factor {
  TS* L;             // matrix in, factor out (lower NB-tiles), leading dimension ld
  TS* X;             // L^-1 out (lower)
  int64_t ld;
  TS* T;             // scratch of (n - split(n)) * split(n) elements
  double* leafinv;   // n/NB blocks of NB x NB float64 leaf inverses
  double* leafld;    // n/NB leaf log-determinants
  int* leafinfo;     // n/NB leaf pivot status
  int ar_potrf, ar_trtri;
  hipStream_t s;
  // Optional float64 copy of the diagonal NB-blocks ([n/NB][NB][NB]): when set
  // (float32 storage), SYRK updates those blocks in float64 and the leaves
  // factor them from there, so no pivot block is ever rounded to float32.
  double* dblk = nullptr;
  // Optional batch: all pointers above are member 0's; member z's are
  // bt.ws_stride bytes further per z (same padded size for every member).
  Batch bt = Batch();
  // fp16 operand scales (device, AR_FP16X3 only): [sL, sL, sX, sT] with sL for
  // entries of L and the trailing matrix, sX for X = L^-1, sT for L21 X11; the
  // GEMMs read the (A, B) pair at fscale + 0 (POTRF), + 1 and + 2 (TRTRI).
  const float* fscale = nullptr;
  // Diagonal blocks of this width (at multiples of it) already hold their
  // inverse in X when trtri runs (potrf_lookahead inverts each panel's
  // diagonal block on its second stream); NB = only the leaves.
  int64_t tri_unit = NB;

  double* dblk_at(int64_t o) const { return dblk ? dblk + (o / NB) * NB * NB : nullptr; }
  int leaf(int64_t o) {
    return launch_leaf<TS>(L, X, ld, o, leafinv + (o / NB) * NB * NB, leafld + o / NB,
                           leafinfo + o / NB, dblk_at(o), s, bt);
  }
  // Solve Y L^T = B in place, B = L[row0 : row0+m, o : o+n], triangle L[o:o+n, o:o+n].
  int trsm(int64_t row0, int64_t m, int64_t o, int64_t n) {
    if (n == NB) {
      if (ar_potrf != AR_F64) {
        // float32 storage: the leaf also left L11^-1 in float32 in X's diagonal
        // block, so the product takes the Cholesky updates' split arithmetic.
        return launch_gemm<TS, TS, TS>(ar_potrf, false, false, L + row0 * ld + o, ld,
                                       X + o * ld + o, ld, L + row0 * ld + o, ld, m, NB, NB, 1.0,
                                       false, false, KR_FULL, s, nullptr, bt,
                                       fscale ? fscale + 1 : nullptr);
      }
      const double* inv = leafinv + (o / NB) * NB * NB;
      return launch_gemm<TS, double, TS>(AR_F64, false, false, L + row0 * ld + o, ld, inv, NB,
                                         L + row0 * ld + o, ld, m, NB, NB, 1.0, false, false,
                                         KR_FULL, s, nullptr, bt);
    }
    const int64_t h = split_point(n);
    GPK_TRY(trsm(row0, m, o, h));
    GPK_TRY(launch_gemm<TS, TS, TS>(ar_potrf, false, false, L + row0 * ld + o, ld,
                                    L + (o + h) * ld + o, ld, L + row0 * ld + o + h, ld, m,
                                    n - h, h, -1.0, true, false, KR_FULL, s, nullptr, bt,
                                    fscale));
    return trsm(row0, m, o + h, n - h);
  }
  int potrf(int64_t o, int64_t n) {
    if (n == NB) return leaf(o);
    const int64_t h = split_point(n);
    GPK_TRY(potrf(o, h));
    GPK_TRY(trsm(o + h, n - h, o, h));
    GPK_TRY(launch_gemm<TS, TS, TS>(ar_potrf, false, false, L + (o + h) * ld + o, ld,
                                    L + (o + h) * ld + o, ld, L + (o + h) * ld + o + h, ld,
                                    n - h, n - h, h, -1.0, true, true, KR_FULL, s,
                                    dblk_at(o + h), bt, fscale));
    return potrf(o + h, n - h);
  }
  // Right-looking over panels of width B with a look-ahead of one panel: while
  // this stream applies panel j's update to the trailing columns beyond panel
  // j+1, a second stream factors panel j+1's diagonal block (the recursion's
  // latency-bound bottom: leaves and small updates) and solves its column
  // block, so that chain runs under the large update instead of after it.
  //   stream s : (a) diag block j+1 -= P_j P_j^T   -> ev0
  //              (b) column block j+1 -= P_j P_j^T -> ev1
  //              (c) columns beyond j+1 -= ...      (overlaps s2)
  //              wait ev2
  //   stream s2: wait ev0; potrf(diag j+1); wait ev1; trsm(column block j+1); -> ev2
  // The operations and their arithmetic are the recursion's; only the order of
  // the trailing updates differs (rank-B steps instead of halving splits).
  int potrf_lookahead(int64_t n, int64_t B) {
    hipStream_t s2 = nullptr;
    hipEvent_t ev[3];
    GPK_TRY(lookahead_resources(&s2, ev));
    // Whatever happens below, the calling stream ends up waiting for all the
    // second stream's work (an early error return included), so no work of
    // this call is left unjoined.
    StreamJoin join{s, s2, ev[2]};
    Factor<TS> G = *this;
    G.s = s2;
    GPK_TRY(potrf(0, B));
    GPK_TRY(trtri(0, B));
    GPK_TRY(trsm(B, n - B, 0, B));
    for (int64_t o = 0; o + B < n; o += B) {
      const int64_t o1 = o + B;
      const int64_t b1 = n - o1 < B ? n - o1 : B;
      const int64_t o2 = o1 + b1;
      const int64_t rest = n - o2;
      TS* P = L + o1 * ld + o;  // panel j below its diagonal block: (n - o1) x B
      GPK_TRY(launch_gemm<TS, TS, TS>(ar_potrf, false, false, P, ld, P, ld, L + o1 * ld + o1, ld,
                                      b1, b1, B, -1.0, true, true, KR_FULL, s, dblk_at(o1), bt,
                                      fscale));
      GPK_HIP_TRY(hipEventRecord(ev[0], s));
      if (rest > 0)
        GPK_TRY(launch_gemm<TS, TS, TS>(ar_potrf, false, false, P + b1 * ld, ld, P, ld,
                                        L + o2 * ld + o1, ld, rest, b1, B, -1.0, true, false,
                                        KR_FULL, s, nullptr, bt, fscale));
      GPK_HIP_TRY(hipEventRecord(ev[1], s));
      GPK_HIP_TRY(hipStreamWaitEvent(s2, ev[0], 0));
      GPK_TRY(G.potrf(o1, b1));
      GPK_TRY(G.trtri(o1, b1));
      GPK_HIP_TRY(hipStreamWaitEvent(s2, ev[1], 0));
      if (rest > 0) GPK_TRY(G.trsm(o2, rest, o1, b1));
      GPK_HIP_TRY(hipEventRecord(ev[2], s2));
      if (rest > 0)
        GPK_TRY(launch_gemm<TS, TS, TS>(ar_potrf, false, false, P + b1 * ld, ld, P + b1 * ld,
                                        ld, L + o2 * ld + o2, ld, rest, rest, B, -1.0, true, true,
                                        KR_FULL, s, dblk_at(o2), bt, fscale));
      GPK_HIP_TRY(hipStreamWaitEvent(s, ev[2], 0));
    }
    tri_unit = B;
    return GPK_OK;
  }
  // X = L^-1 below the leaves, whose diagonal blocks the leaves already wrote.
  int trtri(int64_t o, int64_t n) {
    if (n <= tri_unit) return GPK_OK;
    const int64_t hu = n / tri_unit / 2;
    const int64_t h = (hu > 0 ? hu : 1) * tri_unit;
    GPK_TRY(trtri(o, h));
    GPK_TRY(trtri(o + h, n - h));
    // T = L21 X11   ((n-h) x h; X11 lower, so k >= column tile)
    GPK_TRY(launch_gemm<TS, TS, TS>(ar_trtri, false, true, L + (o + h) * ld + o, ld,
                                    X + o * ld + o, ld, T, h, n - h, h, h, 1.0, false, false,
                                    KR_GE_N, s, nullptr, bt, fscale ? fscale + 1 : nullptr));
    // X21 = -X22 T  (X22 lower, so k <= row tile)
    return launch_gemm<TS, TS, TS>(ar_trtri, false, true, X + (o + h) * ld + o + h, ld, T, h,
                                   X + (o + h) * ld + o, ld, n - h, h, n - h, -1.0, false,
                                   false, KR_LE_M, s, nullptr, bt,
                                   fscale ? fscale + 2 : nullptr);
  }
};

// Elements of the trtri scratch T for an n x n factorisation.
inline int64_t factor_scratch_elems(int64_t n) {
  const int64_t h = n > NB ? split_point(n) : 1;
  return (n - h) * h;
}

}  // namespace gpk
