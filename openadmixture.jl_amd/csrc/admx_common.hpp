// admx_common.hpp -- device helpers shared by the CDNA4 kernels of libadmixhip.
//
// Data layout in HBM (identical to the reference's Julia arrays, column-major):
//   BED   : SNP-major 2-bit genotypes, SNP j occupies `pitch` bytes starting at
//           bed + pitch*j (pitch = ceil(I/4) rounded up to 64 bytes, so the
//           16-byte loads of a SNP row and the 64-byte P-side tiles stay
//           aligned and inside the row); J + 16 rows, the last 16 holding the
//           missing code; individual i sits in byte i>>2 at bit offset 2*(i&3)
//           (src/loops.jl:529-534).
//   q     : K x I  (q[k,i] at k + K*i),  p : K x J  (p[k,j] at k + K*j)
//   x     : K x (I+J), q = x[:, 0:I], p = x[:, I:I+J]  (src/structs.jl:120-137)
//   XtX   : K x K x n,  Xtz : K x n
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace admx {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// (x, y) tile of a 1-D launch of ntot = nx * ny workgroups, ordered so that each
// XCD receives a run of consecutive tiles: workgroup L is dispatched to XCD
// L % 8 as that XCD's (L / 8)-th, and tile lid = first(L % 8) + L / 8.  Tiles
// adjacent in x (neighbouring individuals, the same 128-byte BED lines) then
// share one XCD's L2 instead of being fetched by two.
struct Tile2 {
  int64_t x, y;
};
__device__ __forceinline__ Tile2 xcd_tile(int64_t nx) {
  const int64_t ntot = gridDim.x, L = blockIdx.x;
  const int64_t q = ntot >> 3, r = ntot & 7;
  const int64_t xcd = L & 7, slot = L >> 3;
  const int64_t lid = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
  return {lid % nx, lid / nx};
}

// PLINK-1 code -> (m*g, m*(2-g)) with m the nonmissing flag:
//   00 -> (0,2)  01 (missing) -> (0,0)  10 -> (1,1)  11 -> (2,0)
// i.e. g_map = [0,0,1,2], nonmissing_map = [1,0,1,1] of src/loops.jl:2-5.
__device__ __forceinline__ int code_g(uint32_t c) { return (0x02010000u >> (c * 8u)) & 0xffu; }
__device__ __forceinline__ int code_h(uint32_t c) { return (0x00010002u >> (c * 8u)) & 0xffu; }

// (m*g, m*(2-g)) of the four codes as one 8- or 16-byte LDS entry: a pass
// reads the pair for a code with one LDS load instead of two bit-field
// extracts and two integer-to-float conversions on the vector ALU.
template <typename T> struct __align__(2 * sizeof(T)) CodeGH { T g, h; };
template <typename T> __device__ __forceinline__ void code_table_init(CodeGH<T>* tab) {
  if (threadIdx.x < 4) tab[threadIdx.x] = {(T)code_g(threadIdx.x), (T)code_h(threadIdx.x)};
  __syncthreads();
}

// Factors of the product-form log-likelihood term s^g (1-s)^(2-g) of a code as
// x = ax s + bx, y = ay s + by (two FMAs, each rounding exactly as 1 - s or s):
//   00 -> (1-s)(1-s)   01 (missing) -> 1 * 1   10 -> s (1-s)   11 -> s s
struct __align__(16) CodeLL { double ax, bx, ay, by; };
__device__ __forceinline__ void code_table_init(CodeLL* tab) {
  const uint32_t c = threadIdx.x;
  if (c < 4)
    tab[c] = {c >= 2u ? 1.0 : (c == 1u ? 0.0 : -1.0), c >= 2u ? 0.0 : 1.0,
              c == 3u ? 1.0 : (c == 1u ? 0.0 : -1.0), c == 3u ? 0.0 : 1.0};
  __syncthreads();
}

// Reciprocal: hardware estimate refined to full precision.  FP64: the estimate
// r = (1 + d)/x carries |d| up to about 2^-24 (measured on gfx950: 2.6e8 ulp);
// one third-order step r (1 + e + e^2), e = 1 - x r = -d, leaves (1 + d^3)/x,
// i.e. the rounding of the last FMA only (measured: at most 0.5 ulp over 4M
// arguments in [2^-36, 2), the same as two Newton steps, in three FMAs instead
// of four).  FP32: one Newton step.
__device__ __forceinline__ double rcp(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  const double e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, __builtin_fma(e, e, e), r);
}
__device__ __forceinline__ float rcp(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  float e = __builtin_fmaf(-x, r, 1.0f);
  return __builtin_fmaf(r, e, r);
}

template <typename T> __device__ __forceinline__ T tlog(T x);
template <> __device__ __forceinline__ double tlog<double>(double x) { return log(x); }
template <> __device__ __forceinline__ float tlog<float>(float x) { return logf(x); }

// Number of upper-triangle entries of a K x K symmetric matrix.
template <int K> struct Tri { static constexpr int N = K * (K + 1) / 2; };

// Packed index of (a <= b) in row-major upper-triangle order.
__host__ __device__ constexpr int tri_idx(int a, int b, int K) {
  return a * K - a * (a - 1) / 2 + (b - a);
}

// Deterministic in-block tree sum of one double per thread (blockDim.x = NT).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

}  // namespace admx

// Host-side dispatch over K = 1..16: CALL is an expression that uses the
// compile-time constant KK.
#define ADMX_KCASE(k, CALL) \
  case k: {                 \
    constexpr int KK = k;   \
    CALL;                   \
  } break;
#define ADMX_DISPATCH_K(Kval, CALL)                                                   \
  switch (Kval) {                                                                     \
    ADMX_KCASE(1, CALL) ADMX_KCASE(2, CALL) ADMX_KCASE(3, CALL) ADMX_KCASE(4, CALL)     \
    ADMX_KCASE(5, CALL) ADMX_KCASE(6, CALL) ADMX_KCASE(7, CALL) ADMX_KCASE(8, CALL)     \
    ADMX_KCASE(9, CALL) ADMX_KCASE(10, CALL) ADMX_KCASE(11, CALL) ADMX_KCASE(12, CALL)  \
    ADMX_KCASE(13, CALL) ADMX_KCASE(14, CALL) ADMX_KCASE(15, CALL) ADMX_KCASE(16, CALL) \
    default: break;                                                                   \
  }
