// Tile main loop shared by the GEMM kernels and the fused adjoint kernel.
//
// One 512-thread workgroup computes a 128 x 128 tile of
//   acc(m, n) = sum_{k in [kb, ke)} A(m, k) * B(n, k)
// on the CDNA4 matrix cores with eight waves in 2 x 4, 64 x 32 per wave
// (WaveLayout), so two waves share each SIMD and one's staging work (loads,
// bf16 split, LDS stores) overlaps the other's matrix-core work.  The earlier
// 64 x 64-per-wave form needs ~450 registers (bf16) or ~310 (FP64) and leaves
// one wave per SIMD; the two-per-SIMD form measured 11-15 % faster (bf16) and
// 1-7 % faster (FP64) on the same tiles, with identical results.
//   AR_F64    : v_mfma_f64_16x16x4f64, operands staged in LDS as float64,
//               k-slabs of 16, double-buffered LDS (one barrier per slab).
//   AR_BF16X3 : fp32 operands split on load into bf16 hi + lo planes;
//               products hi*hi + hi*lo + lo*hi (Henry, Tang & Heinecke,
//               ARITH 2019) on v_mfma_f32_16x16x32_bf16.
//   AR_BF16X6 : three planes hi + mid + lo, six products
//               (hh, hm, mh, mm, hl, lh): about fp32-product accuracy.
//   AR_FP16X3 : operands scaled by a power of two, two fp16 planes (11 + 11
//               significant bits), products hh + hl + lh on
//               v_mfma_f32_16x16x32_f16: 2^-22 per product, half the matrix-core
//               work and two thirds of the LDS traffic of BF16x6.
// Operand layouts: A(m, k) = A[m*lda + k] (row-contiguous) or A[k*lda + m]
// (column-contiguous); likewise B(n, k).  All extents are multiples of 128.
//
// Matrix-core lane layouts used (gfx950, wave64):
//   16x16x4 f64 : lane l supplies A[l%16][l/16], B[l/16][l%16];
//                 acc element r of lane l is D[l/16 + 4*r][l%16].
//   16x16x32 bf16: lane l supplies A[l%16][8*(l/16) .. +7],
//                 B[8*(l/16) .. +7][l%16]; acc element r of lane l is
//                 D[4*(l/16) + r][l%16].
// gpk_mfma_selftest checks both layouts on the device.
#pragma once

#include "gpk_internal.h"

namespace gpk {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

constexpr int TILE = 128;

// Output accumulators of every main loop are float64.  In the bf16 loops the
// matrix cores accumulate in fp32 over short k-chunks only (see
// mainloop_bf16); each chunk is then added into float64.
template <int AR>
struct AccT {
  typedef f64x4 vec;
  typedef double scalar;
};

// Waves of one tile: NT threads, NW waves of which WN along n; each wave owns
// rows wm*64 .. +64 (four 16-row blocks) and columns wn*WC .. +WC (NJ 16-column
// blocks), wm = w / WN, wn = w % WN.
template <int AR>
struct WaveLayout {
  static constexpr int NT = 512;
  static constexpr int NW = NT / 64;
  static constexpr int WN = 4;
  static constexpr int WC = TILE / WN;
  static constexpr int NJ = WC / 16;
};

// ---- bf16 split: a = p0 + p1 (+ p2) + remainder, each plane rounded to
// nearest even by the hardware conversion (v_cvt_pk_bf16_f32, two values per
// instruction); the residual a - p is exact in fp32.  Eight values become NP
// planes of eight bf16, one uint4 per plane.
template <int NP>
__device__ __forceinline__ void split_pack8(const float (&v)[8], uint4 (&out)[NP]) {
  unsigned wq[NP][4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    f32x2 r = {v[2 * e], v[2 * e + 1]};
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const unsigned u = __builtin_bit_cast(unsigned, __builtin_convertvector(r, bf16x2));
      wq[q][e] = u;
      if (q + 1 < NP) {
        r.x -= __builtin_bit_cast(float, u << 16);
        r.y -= __builtin_bit_cast(float, u & 0xffff0000u);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) out[q] = make_uint4(wq[q][0], wq[q][1], wq[q][2], wq[q][3]);
}

// ---- fp16 split: s*a = h + l + remainder with h = fp16(s*a), l = fp16(s*a - h)
// (both round to nearest even; the residual s*a - h is exact in fp32).  The
// power-of-two scale s keeps |s*a| <= 2^14, so h never overflows, and l is a
// normal fp16 number for every entry within 2^-3 of that bound (smaller entries
// keep an absolute error below 2^-25 / s).
__device__ __forceinline__ void split_pack8_f16(const float (&v)[8], float s, uint4 (&out)[2]) {
  unsigned wq[2][4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    f32x2 r = {v[2 * e] * s, v[2 * e + 1] * s};
    const f16x2 h = __builtin_convertvector(r, f16x2);
    const f32x2 hf = __builtin_convertvector(h, f32x2);
    r.x -= hf.x;
    r.y -= hf.y;
    const f16x2 l = __builtin_convertvector(r, f16x2);
    wq[0][e] = __builtin_bit_cast(unsigned, h);
    wq[1][e] = __builtin_bit_cast(unsigned, l);
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) out[q] = make_uint4(wq[q][0], wq[q][1], wq[q][2], wq[q][3]);
}

// ---- LDS footprint of one main loop ----
constexpr int F64_BK = 16;
constexpr int F64_LDP = TILE + 1;  // padded row length (doubles)
constexpr size_t F64_SMEM = 2 /*A,B*/ * 2 /*buffers*/ * F64_BK * F64_LDP * sizeof(double);
constexpr int BF_BK = 32;
// LDS rows of the bf16 planes.  Swizzled (SWZ): unpadded 64-byte rows whose
// four 16-byte chunks are permuted by chunk ^ ((row >> 1) & 3).  Padded: 96-byte
// rows.  Measured at N = 65,536 (SQ_LDS_BANK_CONFLICT, one evaluation; times on
// one box, two runs each): the swizzled rows have no bank-conflict cycles in any
// of the loops (the padded rows: 27-29 % of LDS-active cycles; the permutation
// chunk ^ ((row >> 2) & 3): 36-44 %, and chunk ^ (row & 3), (chunk + row) & 3
// were no better), and they take the Cholesky updates from 587 to 571 ms and
// the triangular inverse from 335 to 321 ms; the fused Sigma^-1 pass, with its
// single-set pipeline and float64 epilogue, ran 553 -> 561 ms with them and keeps
// the padded rows.
template <bool SWZ>
constexpr int bf_ldk() { return SWZ ? BF_BK : BF_BK + 16; }  // row length (bf16 elements)
template <bool SWZ>
__device__ __forceinline__ int bf_off(int row, int chunk) {
  return row * bf_ldk<SWZ>() + 8 * (SWZ ? chunk ^ ((row >> 1) & 3) : chunk);
}
constexpr int BF_CHUNK = 4;        // slabs per fp32 chunk: 128 k
template <int AR>
constexpr int bf_planes() { return AR == AR_BF16X6 ? 3 : 2; }

// One 16 x 16 x 32 matrix-core product of two 8-element operand fragments held
// as 16 bytes: bf16 for the BF16 arithmetics, fp16 for AR_FP16X3.
template <int AR>
__device__ __forceinline__ f32x4 mfma32(uint4 a, uint4 b, f32x4 c) {
  if constexpr (AR == AR_FP16X3)
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
template <int AR, bool SWZ = true>
constexpr size_t smem_bytes() {
  return AR == AR_F64 ? F64_SMEM
                      : (size_t)2 /*buffers*/ * 2 /*A,B*/ * bf_planes<AR>() * TILE * bf_ldk<SWZ>() *
                            sizeof(unsigned short);
}

// One 16-deep k-slab of FP64 matrix-core products from LDS operands stored
// k-major: a[k * ldp + m], b[k * ldp + n]; the wave owns rows wm*64 .. +64 and
// columns wn*16*NJ .. +16*NJ.
template <int NJ = 4>
__device__ __forceinline__ void mfma_slab_f64(const double* a, const double* b, int ldp,
                                              int lane, int wm, int wn, f64x4 (&acc)[4][NJ]) {
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    const int kr = kk * 4 + (lane >> 4);
    double av[4], bv[NJ];
#pragma unroll
    for (int i = 0; i < 4; ++i) av[i] = a[kr * ldp + wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
    for (int j = 0; j < NJ; ++j) bv[j] = b[kr * ldp + wn * 16 * NJ + j * 16 + (lane & 15)];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NJ; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
  }
}

// ---------------------------------------------------------------------------
// float64 main loop
// ---------------------------------------------------------------------------
template <typename TA, typename TB, bool A_KM, bool B_KN>
__device__ __forceinline__ void mainloop_f64(const TA* __restrict__ A, int64_t lda,
                                             const TB* __restrict__ B, int64_t ldb, int64_t m0,
                                             int64_t n0, int64_t kb, int64_t ke,
                                             f64x4 (&acc)[4][WaveLayout<AR_F64>::NJ],
                                             char* smem) {
  typedef WaveLayout<AR_F64> WL;
  constexpr int NJ = WL::NJ;
  double* As = reinterpret_cast<double*>(smem);                     // [2][BK][LDP]
  double* Bs = As + 2 * F64_BK * F64_LDP;                           // [2][BK][LDP]
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w / WL::WN, wn = w % WL::WN;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
  if (kb >= ke) return;

  // Global -> register staging: 4 elements per thread per operand (512 threads).
  // Row-contiguous operand: row = t/4, k offset (t%4)*4.
  // Column-contiguous operand: k = t/32, row offset (t%32)*4.
  double ra[4], rb[4];
  auto load4 = [](const auto* __restrict__ p, double (&v)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = (double)p[q];
  };
  auto gload = [&](int64_t k0) {
    if (!A_KM) load4(A + (m0 + (t >> 2)) * lda + k0 + (t & 3) * 4, ra);
    else load4(A + (k0 + (t >> 5)) * lda + m0 + (t & 31) * 4, ra);
    if (!B_KN) load4(B + (n0 + (t >> 2)) * ldb + k0 + (t & 3) * 4, rb);
    else load4(B + (k0 + (t >> 5)) * ldb + n0 + (t & 31) * 4, rb);
  };
  auto sstore = [&](int buf) {
    double* a = As + buf * F64_BK * F64_LDP;
    double* b = Bs + buf * F64_BK * F64_LDP;
    if (!A_KM) {
#pragma unroll
      for (int q = 0; q < 4; ++q) a[((t & 3) * 4 + q) * F64_LDP + (t >> 2)] = ra[q];
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) a[(t >> 5) * F64_LDP + (t & 31) * 4 + q] = ra[q];
    }
    if (!B_KN) {
#pragma unroll
      for (int q = 0; q < 4; ++q) b[((t & 3) * 4 + q) * F64_LDP + (t >> 2)] = rb[q];
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) b[(t >> 5) * F64_LDP + (t & 31) * 4 + q] = rb[q];
    }
  };

  const int64_t nslab = (ke - kb) / F64_BK;
  gload(kb);
  sstore(0);
  __syncthreads();
  for (int64_t sl = 0; sl < nslab; ++sl) {
    const int buf = (int)(sl & 1);
    if (sl + 1 < nslab) gload(kb + (sl + 1) * F64_BK);
    mfma_slab_f64<NJ>(As + buf * F64_BK * F64_LDP, Bs + buf * F64_BK * F64_LDP, F64_LDP, lane,
                      wm, wn, acc);
    if (sl + 1 < nslab) sstore(buf ^ 1);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// bf16-split main loop (fp32 operands)
// ---------------------------------------------------------------------------
// Software pipeline over k-slabs of BF_BK with two LDS buffers and one
// barrier per slab: while the matrix cores work on slab s (buffer s&1), the
// registers holding slab s+1 are split into bf16 planes and stored into
// buffer (s+1)&1, and the same registers are refilled with slab s+3, so each
// global load has two slabs to land (DEEP: two register sets alternate;
// loading one slab ahead, with one set, measured 1-4 % slower for row-major
// operands and equal for column-major ones, where the fused Sigma^-1 pass
// needs the registers).  Eight waves
// (two per SIMD) interleave these phases in hardware; no scheduling hints.
//
// Accumulation: the leading product hi*hi goes into its own fp32
// accumulators (one fp32 rounding per slab), the correction products, 2^-8
// and less of it, into a second set.  The leading set restarts from zero every
// BF_CHUNK slabs (128 k; every k-range is a multiple of 128) and is then added
// into the float64 accumulators; the correction set runs over the whole
// k-range, or (CFLUSH) is flushed with it.  In the fused Sigma^-1 pass CFLUSH
// costs 1.6 % of the kernel's time and takes the gradient's relative error at
// N = 65,536 from 9.0e-5 to 7.2e-5; the Cholesky and inverse updates keep the
// single chain (the MLL error they leave, 1.5e-7, is well inside 1e-6).  With all six products in one fp32 accumulator
// over 256 k, the rounding of the running sums dominated the error of the
// whole mixed-precision factorisation (float64 products of the same float32
// operands are ~1000x closer to the float64 result).
template <int AR, bool A_KM, bool B_KN, bool CFLUSH = false, bool DEEP = true, bool SWZ = true>
__device__ __forceinline__ void mainloop_bf16(const float* __restrict__ A, int64_t lda,
                                              const float* __restrict__ B, int64_t ldb,
                                              int64_t m0, int64_t n0, int64_t kb, int64_t ke,
                                              f64x4 (&acc64)[4][WaveLayout<AR>::NJ],
                                              char* smem, float sa = 1.f, float sb = 1.f) {
  typedef WaveLayout<AR> WL;
  constexpr int NJ = WL::NJ;
  constexpr int NP = bf_planes<AR>();
  constexpr int LDK = bf_ldk<SWZ>();
  constexpr int OPS = NP * TILE * LDK;  // bf16 elements of one operand in one buffer
  unsigned short* S = reinterpret_cast<unsigned short*>(smem);  // [2][A, B][NP][TILE][LDK]
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w / WL::WN, wn = w % WL::WN;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc64[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
  if (kb >= ke) return;

  // Every thread stages 8 consecutive k of one row of each operand per slab.
  // Row-contiguous operand: row = t/4, k = (t%4)*8 .. +7 (two 16-byte loads).
  // Column-contiguous operand: row = t%128, k = (t/128)*8 .. +7 (eight
  // 4-byte loads; consecutive lanes read consecutive rows, 256 bytes per load).
  constexpr int NSET = DEEP ? 2 : 1;
  float ra[NSET][8], rb[NSET][8];
  auto gload_one = [&](const float* __restrict__ P, int64_t ld, int64_t r0, bool km,
                       int64_t k0, float (&r)[8]) {
    if (!km) {
      const float* p = P + (r0 + (t >> 2)) * ld + k0 + (t & 3) * 8;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const float4 v = *reinterpret_cast<const float4*>(p + 4 * q);
        r[4 * q] = v.x; r[4 * q + 1] = v.y; r[4 * q + 2] = v.z; r[4 * q + 3] = v.w;
      }
    } else {
      const float* p = P + (k0 + (t >> 7) * 8) * ld + r0 + (t & 127);
#pragma unroll
      for (int q = 0; q < 8; ++q) r[q] = p[q * ld];
    }
  };
  auto gload = [&](int64_t k0, int set) {
    gload_one(A, lda, m0, A_KM, k0, ra[DEEP ? set : 0]);
    gload_one(B, ldb, n0, B_KN, k0, rb[DEEP ? set : 0]);
  };
  // Split into bf16 planes and store one 16-byte write per plane into the
  // k-contiguous layout the matrix-core operands read.
  auto sstore_one = [&](unsigned short* base, const float (&r)[8], bool km, float sc) {
    const int row = km ? (t & 127) : (t >> 2);
    const int ko = km ? (t >> 7) * 8 : (t & 3) * 8;
    uint4 pk[NP];
    if constexpr (AR == AR_FP16X3) split_pack8_f16(r, sc, pk);
    else split_pack8<NP>(r, pk);
#pragma unroll
    for (int pl = 0; pl < NP; ++pl)
      *reinterpret_cast<uint4*>(base + pl * TILE * LDK + bf_off<SWZ>(row, ko / 8)) = pk[pl];
  };
  // Slab s sits in register set s & 1 (DEEP) and is stored into LDS buffer s & 1.
  auto sstore = [&](int buf) {
    sstore_one(S + buf * 2 * OPS, ra[DEEP ? buf : 0], A_KM, sa);
    sstore_one(S + buf * 2 * OPS + OPS, rb[DEEP ? buf : 0], B_KN, sb);
  };

  const int64_t nslab = (ke - kb) / BF_BK;  // a multiple of BF_CHUNK
  gload(kb, 0);
  sstore(0);
  gload(kb + BF_BK, 1);
  if (DEEP) gload(kb + 2 * BF_BK, 0);  // every k-range holds at least 4 slabs
  __syncthreads();
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  f32x4 accc[4][NJ];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) accc[i][j] = zero;
  for (int64_t c0 = 0; c0 < nslab; c0 += BF_CHUNK) {
    f32x4 acch[4][NJ];
#pragma unroll
    for (int u = 0; u < BF_CHUNK; ++u) {
      const unsigned short* As = S + (u & 1) * 2 * OPS;
      const unsigned short* Bs = As + OPS;
      // Fragments of planes 0 and 1; plane 2 (x6) later reuses plane 1's registers.
      uint4 a0[4], b0[NJ], a1[4], b1[NJ];
      auto frag = [&](const unsigned short* P, int pl, int base, int i) {
        return *reinterpret_cast<const uint4*>(
            P + pl * TILE * LDK + bf_off<SWZ>(base + i * 16 + (lane & 15), lane >> 4));
      };
#pragma unroll
      for (int i = 0; i < 4; ++i) a0[i] = frag(As, 0, wm * 64, i);
#pragma unroll
      for (int j = 0; j < NJ; ++j) b0[j] = frag(Bs, 0, wn * WL::WC, j);
#pragma unroll
      for (int i = 0; i < 4; ++i) a1[i] = frag(As, 1, wm * 64, i);
#pragma unroll
      for (int j = 0; j < NJ; ++j) b1[j] = frag(Bs, 1, wn * WL::WC, j);
      // Leading products first: they need only plane 0, and at the chunk's
      // last slab their float64 flush overlaps the correction products.
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          acch[i][j] = mfma32<AR>(a0[i], b0[j], u == 0 ? zero : acch[i][j]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          f32x4 c = accc[i][j];
          if (AR == AR_BF16X6) c = mfma32<AR>(a1[i], b1[j], c);
          c = mfma32<AR>(a1[i], b0[j], c);
          c = mfma32<AR>(a0[i], b1[j], c);
          accc[i][j] = c;
        }
      if (AR == AR_BF16X6) {
#pragma unroll
        for (int i = 0; i < 4; ++i) a1[i] = frag(As, 2, wm * 64, i);
#pragma unroll
        for (int j = 0; j < NJ; ++j) b1[j] = frag(Bs, 2, wn * WL::WC, j);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < NJ; ++j) {
            f32x4 c = accc[i][j];
            c = mfma32<AR>(a1[i], b0[j], c);
            c = mfma32<AR>(a0[i], b1[j], c);
            accc[i][j] = c;
          }
      }
      if (u == BF_CHUNK - 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc64[i][j][r] += (double)acch[i][j][r];
        if constexpr (CFLUSH) {
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
#pragma unroll
              for (int r = 0; r < 4; ++r) acc64[i][j][r] += (double)accc[i][j][r];
              accc[i][j] = zero;
            }
        }
      }
      // Stage slab s+1 (at the last slab a repeat of the final slab, stored and
      // never read) and refill its register set with slab s+3.
      sstore((u + 1) & 1);
      const int64_t ahead = c0 + u + (DEEP ? 3 : 2);
      gload(kb + (ahead < nslab ? ahead : nslab - 1) * BF_BK, (u + 1) & 1);
      __syncthreads();
    }
  }
  // The correction products are at most 2^-8 of the leading ones, so their
  // single fp32 chain over the whole k-range adds little error (zero here
  // under CFLUSH).
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[i][j][r] += (double)accc[i][j][r];
  if constexpr (AR == AR_FP16X3) {
    const double inv = 1.0 / ((double)sa * (double)sb);  // a power of two: exact
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[i][j][r] *= inv;
  }
}

// Row (within a 16 x 16 block) of accumulator element r held by `lane`.
template <int AR>
__device__ __forceinline__ int acc_row(int lane, int r) {
  return AR == AR_F64 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
}

// Tile decode for lower-triangular tile sets: t -> (mt, nt), mt >= nt, with
// tiles of small mt first.
__device__ __forceinline__ void lower_tile(int64_t t, int64_t& mt, int64_t& nt) {
  int64_t a = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (a * (a + 1) / 2 > t) --a;
  while ((a + 1) * (a + 2) / 2 <= t) ++a;
  mt = a;
  nt = t - a * (a + 1) / 2;
}

// Tile decode for a lower-triangular tile set of T x T tiles in bands: rows
// 16S .. 16S+15 form band S (bands from the top, so the tiles with the longest
// k >= row ranges still come first); inside a band the tiles go 16 columns at a
// time (row-major within each 16-column block, the diagonal block last).  About
// 256 consecutive tiles, the ones resident at a time, then share 16 row and 16
// column operand panels instead of one row panel and 256 column panels.
__device__ __forceinline__ void lower_tile_banded(int64_t t, int64_t T, int64_t& mt,
                                                  int64_t& nt) {
  constexpr int64_t H = 16;
  auto cum = [](int64_t S) { return 128 * S * (S - 1) + 136 * S; };  // tiles of full bands < S
  int64_t S = (int64_t)((-8.0 + sqrt(64.0 + 512.0 * (double)t)) / 256.0);
  while (S > 0 && cum(S) > t) --S;
  while (cum(S + 1) <= t && H * (S + 1) < T) ++S;
  const int64_t R0 = H * S;
  const int64_t h = (T - R0 < H) ? T - R0 : H;
  const int64_t u = t - cum(S);
  if (u < S * H * h) {
    const int64_t cb = u / (H * h), v = u % (H * h);
    mt = R0 + v / H;
    nt = H * cb + v % H;
  } else {
    int64_t a, b;
    lower_tile(u - S * H * h, a, b);
    mt = R0 + a;
    nt = R0 + b;
  }
}

// The same for a full Mt x Nt tile grid: b (row-major launch order) -> bands of
// 16 tile rows from the top, 16-column blocks within a band, row-major within
// a block (ragged last band and block included).
__device__ __forceinline__ void grid_tile_banded(int64_t b, int64_t Mt, int64_t Nt, int64_t& mt,
                                                 int64_t& nt) {
  constexpr int64_t H = 16;
  const int64_t S = b / (H * Nt);
  const int64_t u = b - S * H * Nt;
  const int64_t h = (Mt - S * H < H) ? Mt - S * H : H;
  const int64_t cb = u / (h * H);
  const int64_t v = u - cb * h * H;
  const int64_t w = (Nt - H * cb < H) ? Nt - H * cb : H;
  mt = S * H + v / w;
  nt = H * cb + v % w;
}

__device__ __forceinline__ void k_range(int krange, int64_t mt, int64_t nt, int64_t K,
                                        int64_t& kb, int64_t& ke) {
  kb = 0;
  ke = K;
  if (krange == KR_GE_N) kb = nt * TILE;
  else if (krange == KR_LE_M) ke = (mt + 1) * TILE < K ? (mt + 1) * TILE : K;
  else if (krange == KR_GE_M) kb = mt * TILE;
}

}  // namespace gpk
