// admx_mixed.hpp -- BF16 / matrix-core helpers shared by the mixed-precision
// passes (admx_mixed.hip) and the matrix-core log-likelihood and EM sums
// (admx_mfma_ll.hip).
#pragma once
#include <hip/hip_bf16.h>

#include "admx_common.hpp"

namespace admx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int K> struct Tiles32 {
  static constexpr int NP = K * (K + 1) / 2;
  static constexpr int NT = (NP + 31) / 32;  // 32-wide pair tiles
  static constexpr int NW = NT * 32;
};

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Two FP32 values rounded to BF16 (round to nearest even) in one register.
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}

// Genotype counts of four 2-bit codes held in the low bits of the four bytes of
// `codes`: byte x of the result is g (table 0,0,1,2 for codes 00,01,10,11) or
// 2-g on observed entries (2,0,1,0), one v_perm_b32 byte lookup each.
__device__ __forceinline__ uint32_t g_bytes(uint32_t codes) { return __builtin_amdgcn_perm(0u, 0x02010000u, codes); }
__device__ __forceinline__ uint32_t h_bytes(uint32_t codes) { return __builtin_amdgcn_perm(0u, 0x00010002u, codes); }
__device__ __forceinline__ float byte_f(uint32_t w, int x) { return (float)((w >> (8 * x)) & 0xffu); }

// 8 consecutive BF16 of row `n` of a [NW][ld] table starting at column c (16-byte aligned).
__device__ __forceinline__ bf16x8 load_b8(const __bf16* __restrict__ t, int64_t ld, int n, int64_t c) {
  const u32x4 v = *reinterpret_cast<const u32x4*>(t + (int64_t)n * ld + c);
  return __builtin_bit_cast(bf16x8, v);
}

// Swap the lanes' halves so that lo/hi (weights of SNPs 0-7 / 8-15 of the
// lane's own row) become the A operands of row tiles 0 (lanes 0-31) and 1.
__device__ __forceinline__ void to_tiles(const uint32_t (&lo)[4], const uint32_t (&hi)[4], bf16x8& t0, bf16x8& t1) {
  u32x4 a, b;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const auto r = __builtin_amdgcn_permlane32_swap(lo[e], hi[e], false, false);
    a[e] = r[0];
    b[e] = r[1];
  }
  t0 = __builtin_bit_cast(bf16x8, a);
  t1 = __builtin_bit_cast(bf16x8, b);
}

// Split tables: every FP64 value v becomes three BF16 terms v1 + v2 + v3
// (v1 = bf16(v), v2 = bf16(v - v1), v3 = bf16(v - v1 - v2), a residual of about
// 2^-24 |v|); the six products of order < 3 of two such values stack along an
// MFMA's reduction dimension, KD = 16 ceil(6K / 16) entries per row.
template <int K> struct SplitDim {
  static constexpr int KD = ((6 * K + 15) / 16) * 16;
  static constexpr int NM = KD / 16;
};

__device__ __forceinline__ void split3(double v, __bf16& h1, __bf16& h2, __bf16& h3) {
  h1 = (__bf16)(float)v;
  const double r1 = v - (double)(float)h1;
  h2 = (__bf16)(float)r1;
  const double r2 = r1 - (double)(float)h2;
  h3 = (__bf16)(float)r2;
}

// 1 - sum_k q_ik in FP64, rounded to FP32: q.(1-p) + (1 - sum q) = 1 - q.p is the
// reference's t also where the columns of q do not sum to 1 (after EM steps,
// whose q sums to the observed fraction, src/algorithms_inner.jl:76-89).
template <int K>
__device__ __forceinline__ float sum_complement(const double* __restrict__ q, int64_t i) {
  double sq = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) sq += q[(int64_t)K * i + k];
  return (float)(1.0 - sq);
}

// Leading and residual BF16 of two FP32 weights, packed.
__device__ __forceinline__ void split_pack(float w0, float w1, uint32_t& hi, uint32_t& lo) {
  hi = pack_bf16x2(w0, w1);
  const float r0 = w0 - __builtin_bit_cast(float, hi << 16), r1 = w1 - __builtin_bit_cast(float, hi & 0xffff0000u);
  lo = pack_bf16x2(r0, r1);
}

__device__ __forceinline__ bf16x8 b8(const uint32_t* w) {
  return __builtin_bit_cast(bf16x8, (u32x4){w[0], w[1], w[2], w[3]});
}

}  // namespace admx
