// admx_mfma_ll.hip -- the BF16 form's log-likelihood and FRAPPE EM sums with
// s = q.p and t = q.(1-p) on the matrix cores (v_mfma_f32_32x32x16_bf16).
// Built with the MFMA accumulators in VGPRs (-amdgpu-mfma-vgpr-form): the
// per-genotype VALU work reads s and t straight from the accumulators.
#include "admx_internal.hpp"
#include "admx_mixed.hpp"

namespace admx {

// ------------------------------------------------------------------ log-likelihood terms on the matrix cores
// s = q.p and t = q.(1-p) for a 32-SNP x 32-individual tile are two products
// of BF16 matrices: every FP64 value v is split into three BF16 terms
// v1 + v2 + v3 (v1 = bf16(v), v2 = bf16(v - v1), v3 = bf16(v - v1 - v2), a
// residual of about 2^-24 |v|), and the six products of order < 3,
//   s = q1 p1 + q1 p2 + q2 p1 + q1 p3 + q2 p2 + q3 p1,
// stack along the MFMA's reduction dimension: row j of the SNP table holds
// [p1 p2 p1 p3 p2 p1] (K values each), row i of the individual table
// [q1 q1 q2 q1 q2 q3], zero-padded to KD = 16 ceil(6K / 16); a second SNP table
// holds the same for m = 1 - p, so t is accurate relative to itself even where
// s is close to 1; t then takes 1 - sum_k q_ik on top, so it is 1 - s also when
// q's column does not sum to 1.  The dropped products and the FP32
// accumulation leave s and t with relative errors of a few 1e-7.

// Individual table (role 0) or SNP tables (role 1: T1 from p, T2 from 1 - p),
// `rows` rows of KD BF16 each, zero past column n.
template <int K>
__global__ __launch_bounds__(256) void k_split_tables(const double* __restrict__ x, int64_t n, int64_t rows, int role,
                                                      __bf16* __restrict__ T1, __bf16* __restrict__ T2) {
  constexpr int KD = SplitDim<K>::KD;
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= rows) return;
  // order of the three terms in the six product slots
  constexpr int QO[6] = {0, 0, 1, 0, 1, 2};
  constexpr int PO[6] = {0, 1, 0, 2, 1, 0};
  __bf16 o1[KD], o2[KD];
#pragma unroll
  for (int e = 0; e < KD; ++e) o1[e] = o2[e] = (__bf16)0.0f;
  if (c < n) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double v = x[(int64_t)K * c + k];
      __bf16 h[3], g[3];
      split3(v, h[0], h[1], h[2]);
      split3(1.0 - v, g[0], g[1], g[2]);
#pragma unroll
      for (int slot = 0; slot < 6; ++slot) {
        o1[slot * K + k] = role ? h[PO[slot]] : h[QO[slot]];
        o2[slot * K + k] = g[PO[slot]];
      }
    }
  }
  u32x4* d1 = reinterpret_cast<u32x4*>(T1 + c * KD);
#pragma unroll
  for (int e = 0; e < KD / 8; ++e) {
    u32x4 v;
#pragma unroll
    for (int w = 0; w < 4; ++w) v[w] = __builtin_bit_cast(uint32_t, (bf16x2){o1[8 * e + 2 * w], o1[8 * e + 2 * w + 1]});
    d1[e] = v;
  }
  if (role) {
    u32x4* d2 = reinterpret_cast<u32x4*>(T2 + c * KD);
#pragma unroll
    for (int e = 0; e < KD / 8; ++e) {
      u32x4 v;
#pragma unroll
      for (int w = 0; w < 4; ++w)
        v[w] = __builtin_bit_cast(uint32_t, (bf16x2){o2[8 * e + 2 * w], o2[8 * e + 2 * w + 1]});
      d2[e] = v;
    }
  }
}

// Skipmissing log-likelihood with matrix-core s and t.  A wave owns 64
// individuals (two 32-column tiles) and walks its SNP chunk in tiles of 32:
//   D[j][i] = sum_k A[j][k] B[k][i] on v_mfma_f32_32x32x16_bf16, A = SNP table
//   rows (lane j + 32 (k/8)), B = individual table rows (lane i + 32 (k/8)),
//   D[j][i] in lane i + 32 ((j/4) % 2), register (j % 4) + 4 (j / 8);
// so lane l holds, for individual l % 32 of each tile, the 16 SNPs
// 8a + 4 (l/32) + b (register 4a + b).  The tile's genotype bytes (32 rows x
// 16 bytes) come with one 8-byte load per lane and are transposed through
// 512 bytes of LDS into [byte][row], from which a lane reads the four SNPs
// 8a + 4 (l/32) .. + 3 of its individual as one dword.  Per genotype the lane
// forms s^g t^(2-g) in FP32 (1 if missing) and multiplies it into an FP64
// product renormalised every 8 genotypes: one log per lane at the end.
template <int K>
__global__ __launch_bounds__(256) void k_ll_mfma(const uint8_t* __restrict__ bed, int64_t pitch, int64_t I, int64_t J,
                                                 int64_t jchunk, const double* __restrict__ q64,
                                                 const __bf16* __restrict__ QT, const __bf16* __restrict__ PT,
                                                 const __bf16* __restrict__ MT, double* __restrict__ part,
                                                 const int* __restrict__ skip) {
  if (skip && *skip) return;  // uniform: the state may give s = 0 or 1, the term-by-term FP64 pass runs instead
  constexpr int KD = SplitDim<K>::KD, NM = SplitDim<K>::NM;
  __shared__ __align__(16) uint8_t tbuf[4][512];
  __shared__ double sh_red[256];
  uint8_t* __restrict__ tb = tbuf[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63, h = lane >> 5, cn = lane & 31;
  const int64_t nbx = (I + 255) / 256;
  const Tile2 tile = xcd_tile(nbx);  // neighbouring 256-individual blocks share BED lines: keep them on one XCD
  const int64_t i0 = (tile.x * 4 + (threadIdx.x >> 6)) * 64;
  const int64_t jb = tile.y * jchunk;
  const int64_t je = (jb + jchunk < J) ? jb + jchunk : J;

  bf16x8 bq[2][NM];
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int m = 0; m < NM; ++m) bq[T][m] = load_b8(QT, KD, (int)0, (i0 + 32 * T + cn) * KD + 16 * m + 8 * h);
  const bool ok0 = i0 + cn < I, ok1 = i0 + 32 + cn < I;
  const uint32_t shift = 2u * (uint32_t)(cn & 3);
  const int bi0 = cn >> 2, bi1 = 8 + (cn >> 2);
  const float corr0 = ok0 ? sum_complement<K>(q64, i0 + cn) : 0.f;
  const float corr1 = ok1 ? sum_complement<K>(q64, i0 + 32 + cn) : 0.f;

  const uint8_t* __restrict__ gbase = bed + (i0 >> 2) + 8 * (lane & 1);
  auto gload = [&](int64_t j0) {
    const int64_t j = j0 + (lane >> 1);
    uint2 w = {0x55555555u, 0x55555555u};
    if (j < je) w = *reinterpret_cast<const uint2*>(gbase + pitch * j);
    return w;
  };
  double prod = 1.0;
  int esum = 0;
  uint2 wn = {0x55555555u, 0x55555555u};
  if (jb < je) wn = gload(jb);
  for (int64_t js = jb; js < je; js += 32) {
    bf16x8 as[NM], at[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      as[m] = load_b8(PT, KD, (int)0, (js + cn) * KD + 16 * m + 8 * h);
      at[m] = load_b8(MT, KD, (int)0, (js + cn) * KD + 16 * m + 8 * h);
    }
    // bytes e of this lane's 8 -> tb[byte 8 (l%2) + e][row l/2]
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      tb[(8 * (lane & 1) + e) * 32 + (lane >> 1)] = (uint8_t)(wn.x >> (8 * e));
      tb[(8 * (lane & 1) + 4 + e) * 32 + (lane >> 1)] = (uint8_t)(wn.y >> (8 * e));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint32_t rd[2][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      rd[0][a] = *reinterpret_cast<const uint32_t*>(tb + bi0 * 32 + 8 * a + 4 * h);
      rd[1][a] = *reinterpret_cast<const uint32_t*>(tb + bi1 * 32 + 8 * a + 4 * h);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (js + 32 < je) wn = gload(js + 32);
#pragma unroll
    for (int T = 0; T < 2; ++T) {
      f32x16 sa = (f32x16){}, ta = (f32x16){};
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        sa = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as[m], bq[T][m], sa, 0, 0, 0);
        ta = __builtin_amdgcn_mfma_f32_32x32x16_bf16(at[m], bq[T][m], ta, 0, 0, 0);
      }
      const bool ok = T ? ok1 : ok0;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const uint32_t code = (rd[T][a] >> (8 * b + shift)) & 3u;
          const float sv = sa[4 * a + b], tv = ta[4 * a + b] + (T ? corr1 : corr0);
          const float x = (code >= 2u) ? sv : tv;
          const float y = (code == 3u) ? sv : tv;
          float f = x * y;
          f = (code == 1u || !ok) ? 1.0f : f;
          prod *= (double)f;
        }
        if (a & 1) {
          int e;
          prod = __builtin_frexp(prod, &e);
          esum += e;
        }
      }
    }
  }
  const double acc = log(prod) + (double)esum * 0.69314718055994530942;
  const double tot = block_sum<256>(acc, sh_red);
  if (threadIdx.x == 0) part[tile.y * nbx + tile.x] = tot;
}

template <int K>
static void split_tables_k(const double* x, int64_t n, int64_t rows, int role, __bf16* T1, __bf16* T2, hipStream_t s) {
  hipLaunchKernelGGL((k_split_tables<K>), dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, x, n, rows, role, T1,
                     T2);
}
void launch_split_tables(int K, const double* x, int64_t n, int64_t rows, int role, void* T1, void* T2,
                         hipStream_t s) {
  ADMX_DISPATCH_K(K, (split_tables_k<KK>(x, n, rows, role, (__bf16*)T1, (__bf16*)T2, s)));
}
int split_table_cols(int K) { return ((6 * K + 15) / 16) * 16; }

template <int K>
static void ll_mfma_k(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk, int nchunk,
                      const double* q, const __bf16* QT, const __bf16* PT, const __bf16* MT, double* part,
                      const int* skip, hipStream_t s) {
  const dim3 grid((unsigned)((I + 255) / 256 * nchunk));  // 1-D, tiles ordered per XCD (xcd_tile)
  hipLaunchKernelGGL((k_ll_mfma<K>), grid, dim3(256), 0, s, bed, pitch, I, J, jchunk, q, QT, PT, MT, part, skip);
}
void launch_ll_mfma(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk, int nchunk,
                    const double* q, const void* QT, const void* PT, const void* MT, double* part, const int* skip,
                    hipStream_t s) {
  ADMX_DISPATCH_K(K, (ll_mfma_k<KK>(bed, pitch, I, J, jchunk, nchunk, q, (const __bf16*)QT, (const __bf16*)PT,
                                    (const __bf16*)MT, part, skip, s)));
}

// ------------------------------------------------------------------ FRAPPE EM sums on the matrix cores
// em_q! (src/loops.jl:610-632) sums, per individual i, c p_k and d (1 - p_k)
// over the SNPs (two products of non-negative terms: the K+1 form
// sum (c - d) p_k + sum d of the FP64 pass cancels badly in FP32 where p is
// close to 1 and d large); em_p!
// (src/loops.jl:700-721) sums, per SNP j, c q_k and d q_k over the individuals
// (c = g/s, d = (2-g)/(1-s) on observed entries).  s and t come from the
// matrix cores exactly as in k_ll_mfma, the weights c, d, u are FP32 on the
// VALU, and the sums are BF16 products on v_mfma_f32_32x32x16_bf16 with FP32
// accumulation: each weight w is split into w1 = bf16(w) and w2 = bf16(w - w1),
// each q or p into the rows of a split table (rows k: bf16(x_k), rows 16 + k:
// bf16(x_k - that)), so one product of the two-row table with w1 and one with
// w2 give all four cross terms (relative error about 2^-17 per term).  The
// reduction index of those products runs over the SNPs (individuals) the
// lane's D registers of s already hold: register r = 4a + b is SNP (or
// individual) 8a + 4h + b of the tile, h = lane / 32, so the B operand of the
// first product is registers 0..7 (positions 8a + 4h + b, a < 2) and of the
// second registers 8..15, and the A operand loads the table at the same
// positions.  Rows k and 16 + k of the result sit in registers r and r + 8 of
// one lane.  Missing entries and padding contribute 0.
template <int K>
__global__ __launch_bounds__(256) void k_em_split(const double* __restrict__ x, int64_t n, int64_t ld,
                                                  __bf16* __restrict__ S, __bf16* __restrict__ S2,
                                                  float* __restrict__ qc) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= ld) return;
  double sq = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    __bf16 hi = (__bf16)0.0f, lo = (__bf16)0.0f, mh = (__bf16)0.0f, ml = (__bf16)0.0f;
    if (k < K && c < n) {
      const double v = x[(int64_t)K * c + k], m = 1.0 - v;
      sq += v;
      hi = (__bf16)(float)v;
      lo = (__bf16)(float)(v - (double)(float)hi);
      mh = (__bf16)(float)m;
      ml = (__bf16)(float)(m - (double)(float)mh);
    }
    S[(int64_t)k * ld + c] = hi;
    S[(int64_t)(16 + k) * ld + c] = lo;
    if (S2) {
      S2[(int64_t)k * ld + c] = mh;
      S2[(int64_t)(16 + k) * ld + c] = ml;
    }
  }
  if (qc && c < n) qc[c] = (float)(1.0 - sq);
}

// The split-table A operand of one EM product: row `row` (zero unless its
// population index row % 16 is below K), positions p0 .. p0+3 and p0+8 .. p0+11.
template <int K>
__device__ __forceinline__ bf16x8 em_rows(const __bf16* __restrict__ row, int rix, int64_t p0) {
  u32x4 v = {0u, 0u, 0u, 0u};
  if ((rix & 15) < K) {
    const uint2 lo = *reinterpret_cast<const uint2*>(row + p0);
    const uint2 hi = *reinterpret_cast<const uint2*>(row + p0 + 8);
    v = (u32x4){lo.x, lo.y, hi.x, hi.y};
  }
  return __builtin_bit_cast(bf16x8, v);
}

// FP32 weights c = g/s, d = (2-g)/t of one genotype from gm = m g, hm = m (2-g)
// (0 for missing entries) and the matrix-core s, t.  s t is clamped at 1e-30,
// so that padding rows (s = 0, from the zero rows of the split tables) and
// missing entries give c = d = 0 without a select; real genotypes have
// s t >= 1e-30 by many orders of magnitude (q and p are at least 1e-5 after
// the projections).
__device__ __forceinline__ void em_weights(float gm, float hm, float sv, float tv, float& cc, float& dd) {
  const float r = __builtin_amdgcn_rcpf(fmaxf(sv * tv, 1e-30f));
  cc = gm * (tv * r);
  dd = hm * (sv * r);
}


// em_q: the wave and tile layout of k_ll_mfma (64 individuals, 32-SNP tiles);
// partials [chunk][K+1][I]: sum c p_k + d (1 - p_k) (k < K) and 0 (the FP64
// pass's separate sum d).  Both EM kernels are held to 256 registers (two waves
// per SIMD): left to the compiler they take 266-340 and run one wave per SIMD,
// which leaves the MFMA and table-load latencies exposed (13.9 / 12.0 ms
// instead of 6.6 / 6.4 ms at I = J = 100,000, K = 8).
template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_emq_mfma(const uint8_t* __restrict__ bed, int64_t pitch, int64_t I, int64_t J,
                                                  int64_t jchunk, const double* __restrict__ q64,
                                                  const __bf16* __restrict__ QT, const __bf16* __restrict__ PT,
                                                  const __bf16* __restrict__ MT, const __bf16* __restrict__ PS,
                                                  const __bf16* __restrict__ MS, int64_t ldps,
                                                  double* __restrict__ part) {
  constexpr int KD = SplitDim<K>::KD, NM = SplitDim<K>::NM;
  __shared__ __align__(16) uint8_t tbuf[4][512];
  uint8_t* __restrict__ tb = tbuf[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63, h = lane >> 5, cn = lane & 31;
  const Tile2 tile = xcd_tile((I + 255) / 256);  // as in k_ll_mfma
  const int64_t i0 = (tile.x * 4 + (threadIdx.x >> 6)) * 64;
  const int64_t jb = tile.y * jchunk;
  const int64_t je = (jb + jchunk < J) ? jb + jchunk : J;

  bf16x8 bq[2][NM];
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int m = 0; m < NM; ++m) bq[T][m] = load_b8(QT, KD, (int)0, (i0 + 32 * T + cn) * KD + 16 * m + 8 * h);
  const bool ok0 = i0 + cn < I, ok1 = i0 + 32 + cn < I;
  const uint32_t shift = 2u * (uint32_t)(cn & 3);
  const int bi0 = cn >> 2, bi1 = 8 + (cn >> 2);
  const float corr0 = ok0 ? sum_complement<K>(q64, i0 + cn) : 0.f;
  const float corr1 = ok1 ? sum_complement<K>(q64, i0 + 32 + cn) : 0.f;
  const __bf16* __restrict__ prow = PS + (int64_t)cn * ldps + 4 * h;
  const __bf16* __restrict__ mrow = MS + (int64_t)cn * ldps + 4 * h;

  const uint8_t* __restrict__ gbase = bed + (i0 >> 2) + 8 * (lane & 1);
  auto gload = [&](int64_t j0) {
    const int64_t j = j0 + (lane >> 1);
    uint2 w = {0x55555555u, 0x55555555u};
    if (j < je) w = *reinterpret_cast<const uint2*>(gbase + pitch * j);
    return w;
  };
  f32x16 zc[2] = {(f32x16){}, (f32x16){}}, zd[2] = {(f32x16){}, (f32x16){}};
  uint2 wn = {0x55555555u, 0x55555555u};
  if (jb < je) wn = gload(jb);
  for (int64_t js = jb; js < je; js += 32) {
    bf16x8 as[NM], at[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      as[m] = load_b8(PT, KD, (int)0, (js + cn) * KD + 16 * m + 8 * h);
      at[m] = load_b8(MT, KD, (int)0, (js + cn) * KD + 16 * m + 8 * h);
    }
    const bf16x8 az0 = em_rows<K>(prow, cn, js), az1 = em_rows<K>(prow, cn, js + 16);
    const bf16x8 am0 = em_rows<K>(mrow, cn, js), am1 = em_rows<K>(mrow, cn, js + 16);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      tb[(8 * (lane & 1) + e) * 32 + (lane >> 1)] = (uint8_t)(wn.x >> (8 * e));
      tb[(8 * (lane & 1) + 4 + e) * 32 + (lane >> 1)] = (uint8_t)(wn.y >> (8 * e));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint32_t rd[2][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      rd[0][a] = *reinterpret_cast<const uint32_t*>(tb + bi0 * 32 + 8 * a + 4 * h);
      rd[1][a] = *reinterpret_cast<const uint32_t*>(tb + bi1 * 32 + 8 * a + 4 * h);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (js + 32 < je) wn = gload(js + 32);
#pragma unroll
    for (int T = 0; T < 2; ++T) {
      f32x16 sa = (f32x16){}, ta = (f32x16){};
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        sa = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as[m], bq[T][m], sa, 0, 0, 0);
        ta = __builtin_amdgcn_mfma_f32_32x32x16_bf16(at[m], bq[T][m], ta, 0, 0, 0);
      }
      const float corr = T ? corr1 : corr0;
      // Individuals past I (zero split rows: s = 0; their sums are not written)
      // and SNP rows past the chunk (missing codes) contribute 0.
      uint32_t ch[8], cl[8], dh[8], dl[8];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const uint32_t codes = (rd[T][a] >> shift) & 0x03030303u;  // SNP 8a + 4h + b in byte b
        const uint32_t gb = g_bytes(codes), hb = h_bytes(codes);
#pragma unroll
        for (int b = 0; b < 4; b += 2) {
          float cv[2], dv[2];
#pragma unroll
          for (int x = 0; x < 2; ++x)
            em_weights(byte_f(gb, b + x), byte_f(hb, b + x), sa[4 * a + b + x], ta[4 * a + b + x] + corr, cv[x],
                       dv[x]);
          split_pack(cv[0], cv[1], ch[(4 * a + b) >> 1], cl[(4 * a + b) >> 1]);
          split_pack(dv[0], dv[1], dh[(4 * a + b) >> 1], dl[(4 * a + b) >> 1]);
        }
      }
      zc[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(az0, b8(ch), zc[T], 0, 0, 0);
      zc[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(az0, b8(cl), zc[T], 0, 0, 0);
      zc[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(az1, b8(ch + 4), zc[T], 0, 0, 0);
      zc[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(az1, b8(cl + 4), zc[T], 0, 0, 0);
      zd[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am0, b8(dh), zd[T], 0, 0, 0);
      zd[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am0, b8(dl), zd[T], 0, 0, 0);
      zd[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am1, b8(dh + 4), zd[T], 0, 0, 0);
      zd[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am1, b8(dl + 4), zd[T], 0, 0, 0);
    }
  }
#pragma unroll
  for (int T = 0; T < 2; ++T) {
    if (!(T ? ok1 : ok0)) continue;
    double* __restrict__ out = part + tile.y * (K + 1) * I + i0 + 32 * T + cn;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int row = 8 * (r >> 2) + 4 * h + (r & 3);
      if (row < K)
        out[(int64_t)row * I] = ((double)zc[T][r] + (double)zc[T][r + 8]) + ((double)zd[T][r] + (double)zd[T][r + 8]);
    }
    if (h == 0) out[(int64_t)K * I] = 0.0;
  }
}

// em_p: a wave owns 64 SNPs (two 32-SNP tiles) and walks its individual chunk
// in steps of 32 (D[i][j]: A = individual table, B = SNP tables, lane l holds
// SNP l % 32 of each tile and individuals 8a + 4 (l/32) + b).  The genotype
// bytes of 64 individuals come with one 16-byte load per SNP row, one block
// ahead; the t correction 1 - sum q of individual i comes from lane i % 32.
// Partials [chunk][2K][J]: sum c q_k, then sum d q_k.
template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_emp_mfma(const uint8_t* __restrict__ bed, int64_t pitch, int64_t I, int64_t J,
                                                  int64_t ichunk, const float* __restrict__ qc,
                                                  const __bf16* __restrict__ QT, const __bf16* __restrict__ PT,
                                                  const __bf16* __restrict__ MT, const __bf16* __restrict__ QS,
                                                  int64_t ldqs, double* __restrict__ part) {
  constexpr int KD = SplitDim<K>::KD, NM = SplitDim<K>::NM;
  const int lane = threadIdx.x & 63, h = lane >> 5, cn = lane & 31;
  const int64_t j0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
  const int64_t ib0 = (int64_t)blockIdx.y * ichunk;  // ichunk: multiple of 64
  const int64_t ib1 = (ib0 + ichunk < I) ? ib0 + ichunk : I;
  bf16x8 bp[2][NM], bm[2][NM];
  bool okj[2];
  const uint8_t* __restrict__ grow[2];
#pragma unroll
  for (int T = 0; T < 2; ++T) {
    const int64_t jt = j0 + 32 * T + cn;
    okj[T] = jt < J;
    const int64_t jr = okj[T] ? jt : J - 1;  // rows past J: read a real row, sums dropped
    grow[T] = bed + pitch * jr;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      bp[T][m] = load_b8(PT, KD, (int)0, jr * KD + 16 * m + 8 * h);
      bm[T][m] = load_b8(MT, KD, (int)0, jr * KD + 16 * m + 8 * h);
    }
  }
  const __bf16* __restrict__ qrow = QS + (int64_t)cn * ldqs + 4 * h;
  f32x16 e1[2] = {(f32x16){}, (f32x16){}}, e2[2] = {(f32x16){}, (f32x16){}};
  u32x4 gw[2], gn[2] = {(u32x4){0u, 0u, 0u, 0u}, (u32x4){0u, 0u, 0u, 0u}};
  auto gload = [&](int64_t ib) {
#pragma unroll
    for (int T = 0; T < 2; ++T) gn[T] = *reinterpret_cast<const u32x4*>(grow[T] + (ib >> 2));
  };
  if (ib0 < ib1) gload(ib0);
  for (int64_t ib = ib0; ib < ib1; ib += 64) {
    gw[0] = gn[0];
    gw[1] = gn[1];
    if (ib + 64 < ib1) gload(ib + 64);
#pragma unroll
    for (int sub = 0; sub < 2; ++sub) {
      const int64_t is = ib + 32 * sub;
      if (is >= ib1) break;  // uniform
      bf16x8 aq[NM];
#pragma unroll
      for (int m = 0; m < NM; ++m) aq[m] = load_b8(QT, KD, (int)0, (is + cn) * KD + 16 * m + 8 * h);
      const bf16x8 aw0 = em_rows<K>(qrow, cn, is), aw1 = em_rows<K>(qrow, cn, is + 16);
      const float qcl = (is + cn < I) ? qc[is + cn] : 0.f;
      float corr[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) corr[r] = __shfl(qcl, 8 * (r >> 2) + 4 * h + (r & 3));
#pragma unroll
      for (int T = 0; T < 2; ++T) {
        f32x16 sa = (f32x16){}, ta = (f32x16){};
#pragma unroll
        for (int m = 0; m < NM; ++m) {
          sa = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aq[m], bp[T][m], sa, 0, 0, 0);
          ta = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aq[m], bm[T][m], ta, 0, 0, 0);
        }
        // Individuals past I have zero split rows (s = 0) and t correction 0, so
        // they contribute 0 whatever their padding bits hold; chunks other than
        // the last end on a 64-individual block.
        uint32_t ch[8], cl[8], dh[8], dl[8];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int xb = 8 * sub + 2 * a + h;  // byte of the 64-individual block
          const uint32_t byte = (gw[T][xb >> 2] >> (8 * (xb & 3))) & 0xffu;
          // the four 2-bit codes of individuals 8a + 4h + b into byte b
          const uint32_t codes = (byte | (byte << 6) | (byte << 12) | (byte << 18)) & 0x03030303u;
          const uint32_t gb = g_bytes(codes), hb = h_bytes(codes);
#pragma unroll
          for (int b = 0; b < 4; b += 2) {
            float cv[2], dv[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
              const int r = 4 * a + b + x;
              em_weights(byte_f(gb, b + x), byte_f(hb, b + x), sa[r], ta[r] + corr[r], cv[x], dv[x]);
            }
            split_pack(cv[0], cv[1], ch[(4 * a + b) >> 1], cl[(4 * a + b) >> 1]);
            split_pack(dv[0], dv[1], dh[(4 * a + b) >> 1], dl[(4 * a + b) >> 1]);
          }
        }
        e1[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aw0, b8(ch), e1[T], 0, 0, 0);
        e1[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aw0, b8(cl), e1[T], 0, 0, 0);
        e1[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aw1, b8(ch + 4), e1[T], 0, 0, 0);
        e1[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aw1, b8(cl + 4), e1[T], 0, 0, 0);
        e2[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aw0, b8(dh), e2[T], 0, 0, 0);
        e2[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aw0, b8(dl), e2[T], 0, 0, 0);
        e2[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aw1, b8(dh + 4), e2[T], 0, 0, 0);
        e2[T] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aw1, b8(dl + 4), e2[T], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int T = 0; T < 2; ++T) {
    if (!okj[T]) continue;
    double* __restrict__ out = part + (int64_t)blockIdx.y * 2 * K * J + j0 + 32 * T + cn;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int row = 8 * (r >> 2) + 4 * h + (r & 3);
      if (row < K) {
        out[(int64_t)row * J] = (double)e1[T][r] + (double)e1[T][r + 8];
        out[(int64_t)(K + row) * J] = (double)e2[T][r] + (double)e2[T][r + 8];
      }
    }
  }
}

template <int K>
static void em_split_k(const double* x, int64_t n, int64_t ld, __bf16* S, __bf16* S2, float* qc, hipStream_t s) {
  hipLaunchKernelGGL((k_em_split<K>), dim3((unsigned)((ld + 255) / 256)), dim3(256), 0, s, x, n, ld, S, S2, qc);
}
void launch_em_split(int K, const double* x, int64_t n, int64_t ld, void* S, void* S2, float* qc, hipStream_t s) {
  ADMX_DISPATCH_K(K, (em_split_k<KK>(x, n, ld, (__bf16*)S, (__bf16*)S2, qc, s)));
}
template <int K>
static void emq_mfma_k(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk, int nchunk,
                       const double* q, const __bf16* QT, const __bf16* PT, const __bf16* MT, const __bf16* PS,
                       const __bf16* MS, int64_t ldps, double* part, hipStream_t s) {
  const dim3 grid((unsigned)((I + 255) / 256 * nchunk));  // 1-D, tiles ordered per XCD (xcd_tile)
  hipLaunchKernelGGL((k_emq_mfma<K>), grid, dim3(256), 0, s, bed, pitch, I, J, jchunk, q, QT, PT, MT, PS, MS, ldps,
                     part);
}
void launch_emq_mfma(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk, int nchunk,
                     const double* q, const void* QT, const void* PT, const void* MT, const void* PS, const void* MS,
                     int64_t ldps, double* part, hipStream_t s) {
  ADMX_DISPATCH_K(K, (emq_mfma_k<KK>(bed, pitch, I, J, jchunk, nchunk, q, (const __bf16*)QT, (const __bf16*)PT,
                                     (const __bf16*)MT, (const __bf16*)PS, (const __bf16*)MS, ldps, part, s)));
}
template <int K>
static void emp_mfma_k(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t ichunk, int nchunk,
                       const float* qc, const __bf16* QT, const __bf16* PT, const __bf16* MT, const __bf16* QS,
                       int64_t ldqs, double* part, hipStream_t s) {
  const dim3 grid((unsigned)((J + 255) / 256), (unsigned)nchunk);
  hipLaunchKernelGGL((k_emp_mfma<K>), grid, dim3(256), 0, s, bed, pitch, I, J, ichunk, qc, QT, PT, MT, QS, ldqs, part);
}
void launch_emp_mfma(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t ichunk, int nchunk,
                     const float* qc, const void* QT, const void* PT, const void* MT, const void* QS, int64_t ldqs,
                     double* part, hipStream_t s) {
  ADMX_DISPATCH_K(K, (emp_mfma_k<KK>(bed, pitch, I, J, ichunk, nchunk, qc, (const __bf16*)QT, (const __bf16*)PT,
                                     (const __bf16*)MT, (const __bf16*)QS, ldqs, part, s)));
}

}  // namespace admx
