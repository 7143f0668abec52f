// admx_mixed.hip -- mixed-precision gradient + Hessian passes: per-genotype
// weights and gradient in FP32 on the VALU, the K x K Hessian on the matrix
// cores in BF16 (v_mfma_f32_32x32x16_bf16) accumulating in FP32.  Everything
// downstream (chunk sums, QP, projections, log-likelihood, quasi-Newton,
// convergence test) stays FP64.
//
// Q side (update_q!, src/loops.jl:867-894), per individual i:
//   XtX_q = sum_j a_ij pp_j + b_ij mm_j,  pp_j = p_j p_j', mm_j = (1-p_j)(1-p_j)'
//   Xtz_q = sum_j c_ij p_j + d_ij (1-p_j)           (FP32 VALU)
// the direct two-term form: every term is non-negative, so BF16 rounding
// (2^-8 relative) cannot be amplified by cancellation.  One lane owns one
// individual and streams SNPs with p_j wave-uniform (scalar loads), exactly as
// the FP64 kernel; every 16 SNPs the lane's 16 weights a (and b) are packed to
// BF16 and one v_permlane32_swap per register pair turns the 64 lanes' rows
// into the A operands of two 32-row tiles:
//   A[i][k] lives in lane i + 32*(k/8), element k%8;  B[k][n] in lane n + 32*(k/8);
//   D[i][n] in lane n + 32*((i/4)%2), register (i%4) + 4*(i/8)   (checked on gfx950).
// The reduction dimension stacks the two terms: k in [0,16) pairs weight a with
// pp, k in [16,32) pairs b with mm, into the same accumulator.  B comes from
// BF16 pair tables PPt/MMt [pair][SNP] built once per pass.
// P side (update_p!, src/loops.jl:923-950): XtX_p = sum_i w_ij qq_i (w = a+b),
// Xtz_p = sum_i (c-d) q_i; one lane per SNP, individuals streamed 16 at a time
// (one 32-bit genotype word), q_i wave-uniform.
#include "admx_internal.hpp"
#include "admx_mixed.hpp"

namespace admx {

// ------------------------------------------------------------------ pair tables
// PP[n][c] = x_a x_b and (optionally) MM[n][c] = (1-x_a)(1-x_b) for the
// upper-triangle pair n = (a <= b) of column c; zero for padding columns/pairs.
template <int K>
__global__ __launch_bounds__(256) void k_pair_tables(const double* __restrict__ x, int64_t n, int64_t ld,
                                                     __bf16* __restrict__ PP, __bf16* __restrict__ MM) {
  using TL = Tiles32<K>;
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= ld) return;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = (c < n) ? x[(int64_t)K * c + k] : 0.0;
  int pn = 0;
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = a; b < K; ++b, ++pn) {
      PP[(int64_t)pn * ld + c] = (__bf16)(c < n ? (float)(v[a] * v[b]) : 0.0f);
      if (MM) MM[(int64_t)pn * ld + c] = (__bf16)(c < n ? (float)((1.0 - v[a]) * (1.0 - v[b])) : 0.0f);
    }
  for (; pn < TL::NW; ++pn) {
    PP[(int64_t)pn * ld + c] = (__bf16)0.0f;
    if (MM) MM[(int64_t)pn * ld + c] = (__bf16)0.0f;
  }
}

__global__ __launch_bounds__(256) void k_to_float(const double* __restrict__ x, int64_t n, float* __restrict__ y) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) y[e] = (float)x[e];
}

// Store the 16 accumulators of one 32 x 32 tile: rows are the 32 individuals
// (or SNPs) of row tile T, columns pairs 32t .. 32t+31.
template <int NP>
__device__ __forceinline__ void store_tile(float* __restrict__ part, int64_t n, int64_t row0, int t,
                                           const f32x16& acc) {
  const int lane = threadIdx.x & 63;
  const int pn = 32 * t + (lane & 31);
  if (pn >= NP) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int64_t row = row0 + (r & 3) + 4 * (lane >> 5) + 8 * (r >> 2);
    if (row < n) part[(int64_t)pn * n + row] = acc[r];
  }
}

// ------------------------------------------------------------------ Q side
// The gradient is accumulated as z_k = sum (c-d) p_k + sum d (the ADMIXTURE
// gradient plus the constant column), K+1 FMAs per genotype instead of 2K; all
// terms are O(1) and the two sums do not cancel.
//
// Genotype bytes: each step of 16 SNPs needs, per lane, the byte of its
// individual in 16 SNP rows (16 x 16 bytes per wave).  Lane l fetches one
// 32-bit word, bytes 4(l%4)..4(l%4)+3 of the wave's 16-byte span of row l/4, and
// the 16 x 16 byte block is transposed through 256 bytes of LDS per wave, so
// that every lane reads its 16 bytes (one per SNP) with one 128-bit load.  The
// device pipeline starts at individual 0, so every wave's span is 16-byte
// aligned and lies inside the padded BED pitch.
template <int K>
__global__ __launch_bounds__(256) void k_qside_bf16(PassArgs a, const float* __restrict__ fq,
                                                    const float* __restrict__ fp, const __bf16* __restrict__ PPt,
                                                    const __bf16* __restrict__ MMt, int64_t ld) {
  using TL = Tiles32<K>;
  constexpr int NP = TL::NP, NT = TL::NT;
  __shared__ __align__(16) uint8_t tbuf[4][256];
  uint8_t* __restrict__ tb = tbuf[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int64_t Ilen = a.i1 - a.i0;
  const Tile2 tile = xcd_tile((Ilen + 255) / 256);
  const int64_t il = tile.x * 256 + threadIdx.x;
  const int64_t wrow0 = tile.x * 256 + (threadIdx.x & ~63);  // first row of this wave
  const bool valid = il < Ilen;
  const int64_t i = a.i0 + (valid ? il : 0);
  const int64_t jb = a.j0 + tile.y * a.chunk;
  const int64_t je = (a.j1 < jb + a.chunk) ? a.j1 : jb + a.chunk;

  float qi[K];
#pragma unroll
  for (int k = 0; k < K; ++k) qi[k] = fq[k + (int64_t)K * i];
  const uint32_t sh = (uint32_t)(lane & 3) * 2u;  // i0 = 0: lane l is individual wrow0 + l
  float z[K], dsum = 0.f;
#pragma unroll
  for (int k = 0; k < K; ++k) z[k] = 0.f;
  f32x16 acc[2][NT];
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[T][t] = (f32x16){};
  const int kh = lane >> 5, cn = lane & 31;

  // The genotype word of step js+16 is in flight while step js is computed, and
  // the B tiles of step js are loaded at its start and consumed by the MFMAs
  // at its end; the P column of SNP j+1 is loaded while SNP j is computed.  Only
  // the cohort's last step can run past je = J: the BED's 16 padding rows hold
  // the missing code and the FP32 P copy's padding columns hold 0.5, so such
  // SNPs add nothing.  Lanes past the last individual compute on individual 0
  // and store nothing.
  const uint8_t* __restrict__ gbase = a.bed + ((a.i0 + wrow0) >> 2) + 4 * (lane & 3) + a.pitch * (lane >> 2);
  uint32_t gw_n = 0;
  float pn[K];
  if (jb < je) {
    gw_n = *reinterpret_cast<const uint32_t*>(gbase + a.pitch * jb);
#pragma unroll
    for (int k = 0; k < K; ++k) pn[k] = fp[(int64_t)K * jb + k];
  }
  for (int64_t js = jb; js < je; js += 16) {
    bf16x8 bpp[NT], bmm[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      bpp[t] = load_b8(PPt, ld, 32 * t + cn, js + 8 * kh);
      bmm[t] = load_b8(MMt, ld, 32 * t + cn, js + 8 * kh);
    }
    // byte e of this lane's word is byte 4(l%4)+e of SNP row l/4 -> tb[byte][row]
#pragma unroll
    for (int e = 0; e < 4; ++e) tb[(4 * (lane & 3) + e) * 16 + (lane >> 2)] = (uint8_t)(gw_n >> (8 * e));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const u32x4 cur = *reinterpret_cast<const u32x4*>(tb + 16 * (lane >> 2));  // 16 SNPs of this lane's byte
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (js + 16 < je) gw_n = *reinterpret_cast<const uint32_t*>(gbase + a.pitch * (js + 16));
    uint32_t gb[4], hb[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const uint32_t cd = (cur[w] >> sh) & 0x03030303u;
      gb[w] = g_bytes(cd);
      hb[w] = h_bytes(cd);
    }
    uint32_t wa[8], wb[8];
#pragma unroll
    for (int u = 0; u < 16; u += 2) {
      float av[2], bv[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        float pv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) pv[k] = pn[k];
        const float* __restrict__ pj1 = fp + (int64_t)K * (js + u + h + 1);  // padded past J
#pragma unroll
        for (int k = 0; k < K; ++k) pn[k] = pj1[k];
        float s = qi[0] * pv[0];
#pragma unroll
        for (int k = 1; k < K; ++k) s = __builtin_fmaf(qi[k], pv[k], s);
        const float t1 = 1.f - s;
        const float gm = byte_f(gb[(u + h) >> 2], (u + h) & 3), hm = byte_f(hb[(u + h) >> 2], (u + h) & 3);
        const float r = __builtin_amdgcn_rcpf(s * t1);  // 1 ulp: the weights are FP32-grade
        const float is = t1 * r, it = s * r;
        const float cc = gm * is, dd = hm * it;
        av[h] = cc * is;
        bv[h] = dd * it;
        const float uu = cc - dd;
        dsum += dd;
#pragma unroll
        for (int k = 0; k < K; ++k) z[k] = __builtin_fmaf(uu, pv[k], z[k]);
      }
      wa[u >> 1] = pack_bf16x2(av[0], av[1]);
      wb[u >> 1] = pack_bf16x2(bv[0], bv[1]);
    }
    bf16x8 a0, a1, b0, b1;
    to_tiles({wa[0], wa[1], wa[2], wa[3]}, {wa[4], wa[5], wa[6], wa[7]}, a0, a1);
    to_tiles({wb[0], wb[1], wb[2], wb[3]}, {wb[4], wb[5], wb[6], wb[7]}, b0, b1);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bpp[t], acc[0][t], 0, 0, 0);
      acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bpp[t], acc[1][t], 0, 0, 0);
      acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b0, bmm[t], acc[0][t], 0, 0, 0);
      acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b1, bmm[t], acc[1][t], 0, 0, 0);
    }
  }
  constexpr int NA = NP + K;
  float* __restrict__ part = static_cast<float*>(a.part) + tile.y * NA * Ilen;
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int t = 0; t < NT; ++t) store_tile<NP>(part, Ilen, wrow0 + 32 * T, t, acc[T][t]);
  if (valid) {
#pragma unroll
    for (int k = 0; k < K; ++k) part[(int64_t)(NP + k) * Ilen + il] = z[k] + dsum;
  }
}

// ------------------------------------------------------------------ P side
template <int K>
__global__ __launch_bounds__(256) void k_pside_bf16(PassArgs a, const float* __restrict__ fq,
                                                    const float* __restrict__ fp, const __bf16* __restrict__ QQt,
                                                    int64_t ld) {
  using TL = Tiles32<K>;
  constexpr int NP = TL::NP, NT = TL::NT;
  const int lane = threadIdx.x & 63;
  const int64_t Jlen = a.j1 - a.j0;
  const int64_t jl = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t wrow0 = (int64_t)blockIdx.x * 256 + (threadIdx.x & ~63);
  const bool valid = jl < Jlen;
  const int64_t j = a.j0 + (valid ? jl : 0);
  const int64_t ia0 = (int64_t)blockIdx.y * a.chunk;  // i0 = 0 in the device pipeline
  const int64_t ia1 = (a.i1 < ia0 + a.chunk) ? a.i1 : ia0 + a.chunk;

  float pj[K];
#pragma unroll
  for (int k = 0; k < K; ++k) pj[k] = fp[k + (int64_t)K * j];
  const uint8_t* __restrict__ col = a.bed + a.pitch * j;
  float z[K];
#pragma unroll
  for (int k = 0; k < K; ++k) z[k] = 0.f;
  f32x16 acc[2][NT];
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[T][t] = (f32x16){};
  const int kh = lane >> 5, cn = lane & 31;

  // Genotype tiles of 256 individuals: 64 bytes of the lane's SNP row with four
  // 16-byte loads issued together, parked in LDS as [word][lane] and read back
  // one word per step, as in the FP64 P side.  The tile is loaded when it is
  // reached (no registers held across the tile keep the occupancy at 4 waves;
  // the other waves cover the load).
  __shared__ uint32_t tw[16][256];
  for (int64_t ib = ia0; ib < ia1; ib += 16) {
    const int sw = (int)((ib - ia0) >> 4) & 15;
    if (sw == 0) {  // uniform: a new tile
      const u32x4* __restrict__ src = reinterpret_cast<const u32x4*>(col + (ib >> 2));
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const u32x4 v = src[c];
#pragma unroll
        for (int e = 0; e < 4; ++e) tw[4 * c + e][threadIdx.x] = v[e];
      }
    }
    bf16x8 bqq[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) bqq[t] = load_b8(QQt, ld, 32 * t + cn, ib + 8 * kh);
    uint32_t w = tw[sw][threadIdx.x];
    if (ib + 16 > ia1) {  // uniform: individuals past the chunk end are missing
      uint32_t keep = 0;
#pragma unroll
      for (int t = 0; t < 16; ++t)
        if (ib + t < ia1) keep |= 3u << (2 * t);
      w = (w & keep) | (0x55555555u & ~keep);
    }
    if (!valid) w = 0x55555555u;
    // plane t holds, in byte b, the code of individual ib + 4b + t
    uint32_t gb[4], hb[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint32_t cd = (w >> (2 * t)) & 0x03030303u;
      gb[t] = g_bytes(cd);
      hb[t] = h_bytes(cd);
    }
    uint32_t wv[8];
#pragma unroll
    for (int u = 0; u < 16; u += 2) {
      float wt[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int v = u + h;
        const float* __restrict__ qv = fq + (int64_t)K * (ib + v);  // padded by 16 individuals
        float qq[K];
#pragma unroll
        for (int k = 0; k < K; ++k) qq[k] = qv[k];
        float s = qq[0] * pj[0];
#pragma unroll
        for (int k = 1; k < K; ++k) s = __builtin_fmaf(qq[k], pj[k], s);
        const float t1 = 1.f - s;
        const float gm = byte_f(gb[v & 3], v >> 2), hm = byte_f(hb[v & 3], v >> 2);
        const float r = __builtin_amdgcn_rcpf(s * t1);  // 1 ulp: the weights are FP32-grade
        const float is = t1 * r, it = s * r;
        const float cc = gm * is, dd = hm * it;
        wt[h] = cc * is + dd * it;
        const float uu = cc - dd;
#pragma unroll
        for (int k = 0; k < K; ++k) z[k] = __builtin_fmaf(uu, qq[k], z[k]);
      }
      wv[u >> 1] = pack_bf16x2(wt[0], wt[1]);
    }
    bf16x8 a0, a1;
    to_tiles({wv[0], wv[1], wv[2], wv[3]}, {wv[4], wv[5], wv[6], wv[7]}, a0, a1);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bqq[t], acc[0][t], 0, 0, 0);
      acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bqq[t], acc[1][t], 0, 0, 0);
    }
  }
  constexpr int NA = NP + K;
  float* __restrict__ part = static_cast<float*>(a.part) + (int64_t)blockIdx.y * NA * Jlen;
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int t = 0; t < NT; ++t) store_tile<NP>(part, Jlen, wrow0 + 32 * T, t, acc[T][t]);
  if (valid) {
#pragma unroll
    for (int k = 0; k < K; ++k) part[(int64_t)(NP + k) * Jlen + jl] = z[k];
  }
}

// ------------------------------------------------------------------ launchers
static unsigned grid_cap(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

void launch_to_float(const double* x, int64_t n, float* y, hipStream_t s) {
  hipLaunchKernelGGL(k_to_float, dim3(grid_cap(n)), dim3(256), 0, s, x, n, y);
}

template <int K>
static void pair_tables_k(const double* x, int64_t n, int64_t ld, __bf16* PP, __bf16* MM, hipStream_t s) {
  hipLaunchKernelGGL((k_pair_tables<K>), dim3((unsigned)((ld + 255) / 256)), dim3(256), 0, s, x, n, ld, PP, MM);
}
void launch_pair_tables(int K, const double* x, int64_t n, int64_t ld, void* PP, void* MM, hipStream_t s) {
  ADMX_DISPATCH_K(K, (pair_tables_k<KK>(x, n, ld, (__bf16*)PP, (__bf16*)MM, s)));
}
int pair_table_rows(int K) {
  const int np = K * (K + 1) / 2;
  return ((np + 31) / 32) * 32;
}

template <int K>
static void qside_bf16_k(const PassArgs& a, int nchunk, const float* fq, const float* fp, const void* PPt,
                         const void* MMt, int64_t ld, hipStream_t s) {
  dim3 grid((unsigned)(((a.i1 - a.i0 + 255) / 256) * nchunk));  // 1-D, tiles ordered per XCD (xcd_tile)
  hipLaunchKernelGGL((k_qside_bf16<K>), grid, dim3(256), 0, s, a, fq, fp, (const __bf16*)PPt, (const __bf16*)MMt,
                     ld);
}
template <int K>
static void pside_bf16_k(const PassArgs& a, int nchunk, const float* fq, const float* fp, const void* QQt, int64_t ld,
                         hipStream_t s) {
  dim3 grid((unsigned)((a.j1 - a.j0 + 255) / 256), (unsigned)nchunk);
  hipLaunchKernelGGL((k_pside_bf16<K>), grid, dim3(256), 0, s, a, fq, fp, (const __bf16*)QQt, ld);
}
void launch_qside_bf16(int K, const PassArgs& a, int nchunk, const float* fq, const float* fp, const void* PPt,
                       const void* MMt, int64_t ld, hipStream_t s) {
  ADMX_DISPATCH_K(K, (qside_bf16_k<KK>(a, nchunk, fq, fp, PPt, MMt, ld, s)));
}
void launch_pside_bf16(int K, const PassArgs& a, int nchunk, const float* fq, const float* fp, const void* QQt,
                       int64_t ld, hipStream_t s) {
  ADMX_DISPATCH_K(K, (pside_bf16_k<KK>(a, nchunk, fq, fp, QQt, ld, s)));
}

}  // namespace admx
