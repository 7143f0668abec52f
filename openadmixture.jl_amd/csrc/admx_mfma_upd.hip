// admx_mfma_upd.hip -- the BF16 form's update_q! pass (src/loops.jl:867-894)
// with every per-SNP product on the matrix cores:
//
//   s = q.p                                    three-term BF16 splits, as in the
//                                              matrix-core log-likelihood
//                                              (v_mfma_f32_32x32x16_bf16), t = 1 - s
//   XtX_q = sum_j a_ij pp_j + b_ij mm_j        a = g/s^2, b = (2-g)/t^2 in BF16,
//                                              pp, mm the BF16 pair products
//   Xtz_q = sum_j c_ij p_j + d_ij (1-p_j)      c = g/s, d = (2-g)/t split into a
//                                              leading and a residual BF16 term,
//                                              p and 1-p into two-row tables
//
// (the Hessian and, for K <= 8, the gradient on v_mfma_f32_16x16x32_bf16),
// leaving the VALU the per-genotype weights only (about 20 instructions per
// genotype instead of 33 in k_qside_bf16, whose s and gradient are FP32 FMAs).
// t = 1 - s in FP32, as in k_qside_bf16: the reference's t is 1 - s.
//
// A wave owns 32 individuals and a workgroup (4 waves) 128; the workgroup walks
// its SNP chunk in blocks of 32.  Every block's SNP-side operands (split rows
// of p, the pair tables, the gradient tables) are one contiguous image in HBM,
// built per pass by k_qimg in exactly the LDS layout the MFMA operand loads
// read, so the workgroup copies it with global_load_lds_dwordx4 (no registers,
// no LDS writes) into one of two LDS buffers while it computes on the other;
// one barrier per block publishes the copy.  Each wave reads the block's
// operands from LDS (16 bytes per lane per operand), instead of every wave
// re-reading them from L2.
//
// Tile layouts (checked on gfx950):
//   32x32x16: A[m][k] in lane m + 32 (k/8), element k%8;  B[k][n] in lane
//             n + 32 (k/8);  D[m][n] in lane n + 32 ((m/4)%2), register
//             (m%4) + 4 (m/8).
//   16x16x32: A[m][k] in lane m + 16 (k/8), element k%8;  B[k][n] in lane
//             n + 16 (k/8);  D[m][n] in lane n + 16 (m/4), register m%4.
// s is D[j][i] = sum A[j][.] (SNP split rows) B[.][i] (individual split rows):
// lane i + 32h holds individual i and SNPs 8a + 4h + b in register 4a + b, and
// its packed weights of registers 8 half .. 8 half + 7 (SNPs 16 half ..
// 16 half + 15) are 8 BF16 per register quadruple, element e <-> SNP
// 16 half + 8 (e/4) + 4h + e%4.  v_permlane16_swap of two such quadruples X, Y
// (rows of 16 lanes: X' = [X0 Y0 X2 Y2], Y' = [X1 Y1 X3 Y3]) gives, for the
// individuals 0-15 (X') and 16-31 (Y') of the wave, the 16x16x32 operand whose
// reduction groups k/8 = 0..3 are (X, h = 0), (Y, h = 0), (X, h = 1), (Y, h = 1):
// X, Y = a, b for the Hessian, c_hi, c_lo (d_hi, d_lo) for the gradient.  The
// pair and gradient tables store their 32 SNP columns in that order, so each
// table operand is one 16-byte LDS read.
#include "admx_internal.hpp"
#include "admx_mixed.hpp"

namespace admx {

// 16 bytes per lane from global memory straight into LDS (global_load_lds_dwordx4):
// lane l's bytes land at dst + 16 l, dst uniform across the wave.
__device__ __forceinline__ void copy_to_lds16(const uint8_t* src, uint8_t* dst) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}

// Byte layout of one 32-SNP block image (and of each LDS buffer).
// EM (em_q!, src/loops.jl:610-632): the split rows of 1 - p as well (t from
// its own product, accurate where s is close to 1: FRAPPE EM states are not
// projected), no pair rows.
template <int K, bool EM = false> struct QImg {
  static constexpr int KD = SplitDim<K>::KD, NM = SplitDim<K>::NM;
  static constexpr int NP = K * (K + 1) / 2, NT16 = (NP + 15) / 16;  // 16-pair Hessian tiles
  // Rows are padded by 16 bytes so that the 16-byte reads of lanes 0-15 fall
  // on 16 different bank groups.
  static constexpr int RS = 2 * KD + 16;  // split row: KD BF16
  static constexpr int RT = 80;           // pair / gradient row: 32 BF16
  static constexpr int SP = 0;            // 32 split rows of p
  static constexpr int SM = SP + 32 * RS; // EM: 32 split rows of 1 - p
  static constexpr int PP = SM + (EM ? 32 * RS : 0);  // NP pair rows p_a p_b (update)
  static constexpr int MM = PP + (EM ? 0 : NP * RT);  // NP pair rows (1-p_a)(1-p_b) (update)
  static constexpr int GP = MM + (EM ? 0 : NP * RT);  // 2K rows: bf16(p_k), then the residual p_k - bf16(p_k)
  static constexpr int GM = GP + 2 * K * RT;  // the same for 1 - p_k
  static constexpr int ZR = GM + 2 * K * RT;  // one zero row (padding pairs and gradient rows)
  static constexpr int BYTES = (ZR + RT + 1023) / 1024 * 1024;
  static constexpr int NCH = BYTES / 1024;  // 1 KiB copies (one wave instruction each)
};

// SNP (within its block) of column `pos` of the pair and gradient tables.
__host__ __device__ constexpr int qimg_snp(int pos) {
  return 16 * (pos >> 4) + 8 * ((pos & 7) >> 2) + 4 * ((pos & 15) >> 3) + (pos & 3);
}

// ------------------------------------------------------------------ block images
// Slot b (0 <= b <= ceil(J / 32)) holds the split rows of SNP block b and the
// pair and gradient tables of block b - 1.  One workgroup per slot, each thread
// writing whole 16-byte chunks, so the image is written with coalesced stores.
// SNPs outside [0, J) take p = 0.5 in the split rows (s and t stay inside
// (0, 1), and their missing genotype codes give zero weights) and 0 in the pair
// and gradient tables.
template <int K, bool EM>
__global__ __launch_bounds__(256) void k_qimg(const double* __restrict__ p, int64_t J, uint8_t* __restrict__ img) {
  using G = QImg<K, EM>;
  const int64_t slot = blockIdx.x;
  uint8_t* __restrict__ out = img + slot * G::BYTES;
  const int64_t sblk = slot, tblk = slot - 1;  // blocks of the split rows and of the tables
  auto pv = [&](int64_t blk, int s, int k) -> double {  // p of SNP s of a block (0.5 outside [0, J))
    const int64_t j = blk * 32 + s;
    return (j >= 0 && j < J) ? p[(int64_t)K * j + k] : 0.5;
  };
  auto in_range = [&](int64_t blk, int s) { return blk >= 0 && blk * 32 + s < J; };
  for (int c = threadIdx.x; c < G::BYTES / 16; c += 256) {
    const int off = 16 * c;
    __bf16 v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (__bf16)0.0f;
    if (off < G::PP) {  // split rows (of p, then EM's of 1 - p): entry e of slot e / K is term PO[slot] of value k = e % K
      const bool m = EM && off >= G::SM;
      const int base = m ? G::SM : G::SP;
      const int r = (off - base) / G::RS, col = ((off - base) % G::RS) / 16;
      if (col < G::KD / 8) {
        constexpr int PO[6] = {0, 1, 0, 2, 1, 0};
        for (int e = 0; e < 8; ++e) {
          const int x = 8 * col + e, slot = x / K, k = x % K;
          if (slot < 6) {
            __bf16 t[3];
            const double pk = pv(sblk, r, k);
            split3(m ? 1.0 - pk : pk, t[0], t[1], t[2]);
            v[e] = t[PO[slot]];
          }
        }
      }
    } else if (off < G::GP) {  // pair rows, columns in operand order
      const bool m = off >= G::MM;
      const int r = (off - (m ? G::MM : G::PP)) / G::RT, col = ((off - (m ? G::MM : G::PP)) % G::RT) / 16;
      if (col < 4) {
        int a = 0, base = 0;
        while (r >= base + (K - a)) base += K - (a++);
        const int b = a + (r - base);
        for (int e = 0; e < 8; ++e) {
          const int s = qimg_snp(8 * col + e);
          if (in_range(tblk, s)) {
            const double xa = pv(tblk, s, a), xb = pv(tblk, s, b);
            v[e] = (__bf16)(m ? (float)((1.0 - xa) * (1.0 - xb)) : (float)(xa * xb));
          }
        }
      }
    } else if (off < G::ZR) {  // gradient rows: bf16(x_k), then x_k - bf16(x_k), x = p or 1 - p
      const bool m = off >= G::GM;
      const int r = (off - (m ? G::GM : G::GP)) / G::RT, col = ((off - (m ? G::GM : G::GP)) % G::RT) / 16;
      if (col < 4) {
        const int k = r % K;
        for (int e = 0; e < 8; ++e) {
          const int s = qimg_snp(8 * col + e);
          if (in_range(tblk, s)) {
            const double x = m ? 1.0 - pv(tblk, s, k) : pv(tblk, s, k);
            const __bf16 hi = (__bf16)(float)x;
            v[e] = r < K ? hi : (__bf16)(float)(x - (double)(float)hi);
          }
        }
      }
    }
    u32x4 w;
#pragma unroll
    for (int e = 0; e < 4; ++e) w[e] = __builtin_bit_cast(uint32_t, (bf16x2){v[2 * e], v[2 * e + 1]});
    *reinterpret_cast<u32x4*>(out + off) = w;
  }
}

__device__ __forceinline__ bf16x8 lds_b8(const uint8_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}

// ------------------------------------------------------------------ the pass
// Partials [chunk][NP + K][I] in FP32, the layout of k_qside_bf16, reduced in
// fixed order by the FP64 finalize.  jchunk is a multiple of 32.  Individuals
// past I have zero split rows (s = 0): their weights are not finite, but a
// weight only reaches its own individual's row (Hessian) or column (gradient)
// of the products, and those are not stored.
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void swap16(uint32_t (&x)[4], uint32_t (&y)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const auto r = __builtin_amdgcn_permlane16_swap(x[e], y[e], false, false);
    x[e] = r[0];
    y[e] = r[1];
  }
}

// EM: em_q!'s sums sum_j c p_k + d (1 - p_k) (the gradient's algebra) with
// t = q.(1-p) + (1 - sum_k q_k) from its own product, no Hessian, written as
// the FP64 partials [chunk][K+1][I] of the FP64 EM pass (row K: 0).
template <int K, bool EM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_qside_mfma(const uint8_t* __restrict__ bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk,
                  const __bf16* __restrict__ QT, const uint8_t* __restrict__ img, void* __restrict__ part,
                  const double* __restrict__ q64) {
  using G = QImg<K, EM>;
  constexpr int KD = G::KD, NM = G::NM, NP = G::NP, NT = G::NT16, NA = NP + K;
  constexpr bool G16 = K <= 8;  // gradient rows (k leading, k residual) fit one 16-row tile
  __shared__ __align__(16) uint8_t sbuf0[G::BYTES];
  __shared__ __align__(16) uint8_t sbuf1[G::BYTES];
  __shared__ __align__(16) uint8_t tbuf[4][256];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: the copies' LDS base goes to M0
  const int lane = threadIdx.x & 63, h = lane >> 5, cn = lane & 31;
  const int n16 = lane & 15, g16 = lane >> 4;  // 16x16x32 operand lane: row/column n16, reduction group g16
  uint8_t* __restrict__ tb = tbuf[wave];
  const int64_t nbx = (I + 127) / 128;
  const Tile2 tile = xcd_tile(nbx);  // neighbouring 128-individual blocks share BED lines: keep them on one XCD
  const int64_t i0 = tile.x * 128 + 32 * wave;
  const int64_t jb = tile.y * jchunk;
  const int64_t je = (jb + jchunk < J) ? jb + jchunk : J;

  bf16x8 bq[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) bq[m] = load_b8(QT, KD, (int)0, (i0 + cn) * KD + 16 * m + 8 * h);
  const uint32_t shift = 2u * (uint32_t)(cn & 3);
  const int bi = cn >> 2;
  float corr = 0.f;  // EM: t = q.(1-p) + 1 - sum_k q_k
  if constexpr (EM) corr = (i0 + cn < I) ? sum_complement<K>(q64, i0 + cn) : 0.f;

  // per-lane operand offsets inside a buffer
  const int o_sp = G::SP + cn * G::RS + 16 * h, o_sm = G::SM + cn * G::RS + 16 * h;
  int o_h[NT];  // Hessian B: pair row 16 t + n16 of pp (even groups) or mm (odd groups), SNPs of h = g16 / 2
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int n = 16 * t + n16;
    o_h[t] = (n < NP ? ((g16 & 1) ? G::MM : G::PP) + n * G::RT : G::ZR) + 16 * (g16 >> 1);
  }
  // gradient A: rows k (leading) and k + 8 (residual) of a 16-row tile (K <= 8),
  // else rows k and 16 + k of a 32-row tile (lanes as the 32x32x16 A operand)
  const int gm_ = G16 ? (n16 & 7) : (cn & 15);
  const int grow = (G16 ? (n16 < 8 ? 0 : K) : (cn < 16 ? 0 : K)) + gm_;
  const int gsh = G16 ? 16 * (g16 >> 1) : 16 * h;
  const int o_gp = (gm_ < K ? G::GP + grow * G::RT : G::ZR) + gsh;
  const int o_gm = (gm_ < K ? G::GM + grow * G::RT : G::ZR) + gsh;

  const uint8_t* __restrict__ gbase = bed + (i0 >> 2) + 4 * (lane & 1);
  auto gload = [&](int64_t j0) {
    const int64_t j = j0 + (lane >> 1);
    uint32_t w = 0x55555555u;
    if (j < je) w = *reinterpret_cast<const uint32_t*>(gbase + pitch * j);
    return w;
  };
  const uint8_t* __restrict__ isrc = img + lane * 16;
  auto copy_block = [&](int64_t blk, uint8_t* dst) {  // this wave's 1 KiB pieces of image slot blk
    for (int c = wave; c < G::NCH; c += 4)
      copy_to_lds16(isrc + blk * G::BYTES + c * 1024, dst + c * 1024);
  };
  auto st_mfma = [&](const uint8_t* __restrict__ buf, f32x16& sv, f32x16& tv) {  // s (and EM's t) from the split rows
    sv = (f32x16){};
    tv = (f32x16){};
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      sv = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_b8(buf + o_sp + 32 * m), bq[m], sv, 0, 0, 0);
      if constexpr (EM) tv = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_b8(buf + o_sm + 32 * m), bq[m], tv, 0, 0, 0);
    }
  };

  // Slot b of the image holds the split rows of block b and the pair and
  // gradient tables of block b - 1, so the buffer of block b's products also
  // holds what s of block b + 1 needs: those MFMAs are issued first and run
  // under block b's weights on the vector ALU.
  f32x4 acc[2][NT];  // Hessian, individuals 0-15 / 16-31 of the wave
  f32x4 z16[2] = {(f32x4){}, (f32x4){}};
  f32x16 z32 = (f32x16){};
#pragma unroll
  for (int gi = 0; gi < 2; ++gi)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[gi][t] = (f32x4){};
  copy_block(jb / 32, sbuf0);
  uint32_t wn = gload(jb);
  __syncthreads();  // waits for the copy (vmcnt) before the barrier
  f32x16 sa, ta;
  st_mfma(sbuf0, sa, ta);
  copy_block(jb / 32 + 1, sbuf1);
  __syncthreads();  // the second copy has landed and every wave is done with sbuf0
  int nb = 1;
  for (int64_t js = jb; js < je; js += 32, nb ^= 1) {
    const uint8_t* __restrict__ cur = nb ? sbuf1 : sbuf0;
    // The other buffer was last read in the previous block, before the barrier
    // that ended it.
    if (js + 32 < je) copy_block(js / 32 + 2, nb ? sbuf0 : sbuf1);
    f32x16 sn, tn;
    st_mfma(cur, sn, tn);  // block js + 32 (past the chunk on its last block: not used)
    // genotype bytes: byte e of this lane's word is byte 4(l%2)+e of SNP row l/2 -> tb[byte][row]
#pragma unroll
    for (int e = 0; e < 4; ++e) tb[(4 * (lane & 1) + e) * 32 + (lane >> 1)] = (uint8_t)(wn >> (8 * e));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint32_t rd[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) rd[a] = *reinterpret_cast<const uint32_t*>(tb + bi * 32 + 8 * a + 4 * h);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (js + 32 < je) wn = gload(js + 32);

#pragma unroll
    for (int half = 0; half < 2; ++half) {  // SNPs 16 half .. 16 half + 15: registers 8 half .. 8 half + 7
      uint32_t wa[4], wb[4], ch[4], cl[4], dh[4], dl[4];
#pragma unroll
      for (int a = 2 * half; a < 2 * half + 2; ++a) {
        const uint32_t codes = (rd[a] >> shift) & 0x03030303u;  // SNP 8a + 4h + b in byte b
        const uint32_t gb = g_bytes(codes), hb = h_bytes(codes);
#pragma unroll
        for (int b = 0; b < 4; b += 2) {
          float av[2], bv[2], cv[2], dv[2];
#pragma unroll
          for (int x = 0; x < 2; ++x) {
            const int r = 4 * a + b + x;
            const float sv = sa[r], tv = EM ? ta[r] + corr : 1.0f - sv;
            // 1-ulp reciprocals (the weights are FP32-grade): two of them cost fewer
            // vector instructions than one of s t and three products
            const float is = __builtin_amdgcn_rcpf(sv), it = __builtin_amdgcn_rcpf(tv);
            cv[x] = byte_f(gb, b + x) * is;
            dv[x] = byte_f(hb, b + x) * it;
            av[x] = cv[x] * is;
            bv[x] = dv[x] * it;
          }
          const int w = ((4 * a + b) >> 1) & 3;
          if constexpr (!EM) {
            wa[w] = pack_bf16x2(av[0], av[1]);
            wb[w] = pack_bf16x2(bv[0], bv[1]);
          }
          split_pack(cv[0], cv[1], ch[w], cl[w]);
          split_pack(dv[0], dv[1], dh[w], dl[w]);
        }
      }
      if constexpr (!G16) {  // gradient on 32x32x16 tiles, from the lanes' own weights
        const bf16x8 gp = lds_b8(cur + o_gp + 32 * half), gm = lds_b8(cur + o_gm + 32 * half);
        z32 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(gp, b8(ch), z32, 0, 0, 0);
        z32 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(gp, b8(cl), z32, 0, 0, 0);
        z32 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(gm, b8(dh), z32, 0, 0, 0);
        z32 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(gm, b8(dl), z32, 0, 0, 0);
      }
      if constexpr (!EM) {
        swap16(wa, wb);  // wa: individuals 0-15, wb: 16-31; groups (a h0) (b h0) (a h1) (b h1)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const bf16x8 bt = lds_b8(cur + o_h[t] + 32 * half);
          acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b8(wa), bt, acc[0][t], 0, 0, 0);
          acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b8(wb), bt, acc[1][t], 0, 0, 0);
        }
      }
      if constexpr (G16) {
        swap16(ch, cl);  // ch: individuals 0-15, cl: 16-31; groups (hi h0) (lo h0) (hi h1) (lo h1)
        swap16(dh, dl);
        const bf16x8 gp = lds_b8(cur + o_gp + 32 * half), gm = lds_b8(cur + o_gm + 32 * half);
        z16[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(gp, b8(ch), z16[0], 0, 0, 0);
        z16[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(gp, b8(cl), z16[1], 0, 0, 0);
        z16[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(gm, b8(dh), z16[0], 0, 0, 0);
        z16[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(gm, b8(dl), z16[1], 0, 0, 0);
      }
    }
    sa = sn;
    ta = tn;
    // The next block's image and genotype word have landed (the barrier waits
    // for them), and every wave is done reading `cur` before it is refilled.
    __syncthreads();
  }
  if constexpr (EM) {
    double* __restrict__ o64 = static_cast<double*>(part) + tile.y * (K + 1) * I;
    if constexpr (G16) {
#pragma unroll
      for (int gi = 0; gi < 2; ++gi) {
        const int64_t col = i0 + 16 * gi + n16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double v = (double)z16[gi][r] + (double)__shfl_xor(z16[gi][r], 32);
          const int k = 4 * g16 + r;
          if (g16 < 2 && k < K && col < I) o64[(int64_t)k * I + col] = v;
        }
        if (g16 == 2 && col < I) o64[(int64_t)K * I + col] = 0.0;
      }
    } else if (i0 + cn < I) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int row = 8 * (r >> 2) + 4 * h + (r & 3);
        if (row < K) o64[(int64_t)row * I + i0 + cn] = (double)z32[r] + (double)z32[r + 8];
      }
      if (h == 0) o64[(int64_t)K * I + i0 + cn] = 0.0;
    }
    return;
  }
  float* __restrict__ out = static_cast<float*>(part) + tile.y * NA * I;
#pragma unroll
  for (int gi = 0; gi < 2; ++gi)  // D[individual][pair]: lane (pair n16, g16) holds individuals 4 g16 + r
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int pn = 16 * t + n16;
      if (pn < NP) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = i0 + 16 * gi + 4 * g16 + r;
          if (row < I) out[(int64_t)pn * I + row] = acc[gi][t][r];
        }
      }
    }
  if constexpr (G16) {  // D[row][individual]: lane (individual n16, g16) holds rows 4 g16 + r
#pragma unroll
    for (int gi = 0; gi < 2; ++gi) {
      const int64_t col = i0 + 16 * gi + n16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float v = z16[gi][r] + __shfl_xor(z16[gi][r], 32);  // leading (rows < 8) + residual (rows 8..15)
        const int k = 4 * g16 + r;
        if (g16 < 2 && k < K && col < I) out[(int64_t)(NP + k) * I + col] = v;
      }
    }
  } else if (i0 + cn < I) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int row = 8 * (r >> 2) + 4 * h + (r & 3);
      if (row < K) out[(int64_t)(NP + row) * I + i0 + cn] = z32[r] + z32[r + 8];
    }
  }
}

// ------------------------------------------------------------------ P side
// update_p! (src/loops.jl:923-950) on the matrix cores, the mirror of
// k_qside_mfma with the roles of individuals and SNPs exchanged:
//   s = q.p (three-term splits; D[i][j], individuals as rows), t = 1 - s,
//   XtX_p = sum_i w_ij qq_i, w = a + b = g/s^2 + (2-g)/t^2 in BF16,
//   Xtz_p = sum_i u_ij q_i,  u = c - d = g/s - (2-g)/t split into a leading
//   and a residual FP16 term, q into two-row FP16 tables.
// A wave owns 32 SNPs (lanes as the columns of the s tile), a workgroup 128;
// the workgroup walks its individual chunk in blocks of 32, each block's
// individual-side operands one image slot in LDS layout (slot b: split rows of
// block b, tables of block b - 1).  With one Hessian weight per genotype, the
// 16x16x32 operand of one block comes from swapping the weights of its two
// halves: reduction groups (half 0, h 0) (half 1, h 0) (half 0, h 1)
// (half 1, h 1), the column order of the qq rows.  Each lane reads its own
// SNP row's 8 genotype bytes of a block (no transpose).
// The P-side gradient weight u = c - d has either sign, so its sum over the
// individuals cancels: it is split into FP16 terms (2^-22 relative, against
// 2^-17 for BF16) on v_mfma_f32_*_f16, with u scaled by 2^-8 and the q rows by
// 2^8 (|u| <= 2 / min(s, t) stays inside FP16's range for s, t >= 1e-5, and the
// smallest q stays a normal number).
constexpr double PGRAD_SCALE = 256.0;
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split_pack_f16(float w0, float w1, uint32_t& hi, uint32_t& lo) {
  const f16x2 h = __builtin_convertvector((f32x2){w0, w1}, f16x2);
  hi = __builtin_bit_cast(uint32_t, h);
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){w0 - (float)h[0], w1 - (float)h[1]}, f16x2));
}
__device__ __forceinline__ f16x8 h8(const uint32_t* w) {
  return __builtin_bit_cast(f16x8, (u32x4){w[0], w[1], w[2], w[3]});
}
__device__ __forceinline__ f16x8 lds_h8(const uint8_t* p) {
  return __builtin_bit_cast(f16x8, *reinterpret_cast<const u32x4*>(p));
}

template <int K> struct PImg {
  static constexpr int KD = SplitDim<K>::KD, NM = SplitDim<K>::NM;
  static constexpr int NP = K * (K + 1) / 2, NT16 = (NP + 15) / 16;
  static constexpr int RS = 2 * KD + 16, RT = 80;
  static constexpr int SQ = 0;                // 32 split rows of q
  static constexpr int QQ = SQ + 32 * RS;     // NP pair rows q_a q_b
  static constexpr int GQ = QQ + NP * RT;     // 2K FP16 rows: 2^8 q_k leading, then residual
  static constexpr int ZR = GQ + 2 * K * RT;  // one zero row
  static constexpr int BYTES = (ZR + RT + 1023) / 1024 * 1024;
  static constexpr int NCH = BYTES / 1024;
};

// Individual (within its block) of column `pos` of the qq rows: pos = 8 g + e
// holds half g % 2, h = g / 2, element e.
__host__ __device__ constexpr int pimg_hcol(int pos) {
  return 16 * ((pos >> 3) & 1) + 8 * ((pos & 7) >> 2) + 4 * (pos >> 4) + (pos & 3);
}

// Slot b (0 <= b <= ceil(I / 32)): split rows of individual block b (individuals
// past I take q_k = 1/K, so that s stays inside (0, 1) where p does; their
// genotype codes are masked to missing), the qq and gradient rows of block
// b - 1 (0 outside [0, I)).
template <int K>
__global__ __launch_bounds__(256) void k_pimg(const double* __restrict__ q, int64_t I, uint8_t* __restrict__ img) {
  using G = PImg<K>;
  const int64_t slot = blockIdx.x;
  uint8_t* __restrict__ out = img + slot * G::BYTES;
  const int64_t sblk = slot, tblk = slot - 1;
  auto in_range = [&](int64_t blk, int s) { return blk >= 0 && blk * 32 + s < I; };
  auto qv = [&](int64_t blk, int s, int k) -> double {
    return in_range(blk, s) ? q[(int64_t)K * (blk * 32 + s) + k] : 1.0 / K;
  };
  for (int c = threadIdx.x; c < G::BYTES / 16; c += 256) {
    const int off = 16 * c;
    __bf16 v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (__bf16)0.0f;
    if (off < G::QQ) {  // split rows
      const int r = (off - G::SQ) / G::RS, col = ((off - G::SQ) % G::RS) / 16;
      if (col < G::KD / 8) {
        constexpr int QO[6] = {0, 0, 1, 0, 1, 2};
        for (int e = 0; e < 8; ++e) {
          const int x = 8 * col + e, slot6 = x / K, k = x % K;
          if (slot6 < 6) {
            __bf16 t[3];
            split3(qv(sblk, r, k), t[0], t[1], t[2]);
            v[e] = t[QO[slot6]];
          }
        }
      }
    } else if (off < G::GQ) {  // qq rows, columns in the Hessian operand order
      const int r = (off - G::QQ) / G::RT, col = ((off - G::QQ) % G::RT) / 16;
      if (col < 4) {
        int a = 0, base = 0;
        while (r >= base + (K - a)) base += K - (a++);
        const int b = a + (r - base);
        for (int e = 0; e < 8; ++e) {
          const int i = pimg_hcol(8 * col + e);
          if (in_range(tblk, i)) v[e] = (__bf16)(float)(qv(tblk, i, a) * qv(tblk, i, b));
        }
      }
    } else if (off < G::ZR) {  // gradient rows in FP16, columns in the order of the weights' halves
      const int r = (off - G::GQ) / G::RT, col = ((off - G::GQ) % G::RT) / 16;
      if (col < 4) {
        const int k = r % K;
        _Float16 f[8];
        for (int e = 0; e < 8; ++e) {
          f[e] = (_Float16)0.0f;
          const int i = qimg_snp(8 * col + e);
          if (in_range(tblk, i)) {
            const double x = qv(tblk, i, k) * PGRAD_SCALE;
            const _Float16 hi = (_Float16)(float)x;
            f[e] = r < K ? hi : (_Float16)(float)(x - (double)(float)hi);
          }
        }
        u32x4 w;
#pragma unroll
        for (int e = 0; e < 4; ++e) w[e] = __builtin_bit_cast(uint32_t, (f16x2){f[2 * e], f[2 * e + 1]});
        *reinterpret_cast<u32x4*>(out + off) = w;
        continue;
      }
    }
    u32x4 w;
#pragma unroll
    for (int e = 0; e < 4; ++e) w[e] = __builtin_bit_cast(uint32_t, (bf16x2){v[2 * e], v[2 * e + 1]});
    *reinterpret_cast<u32x4*>(out + off) = w;
  }
}

// Partials [chunk][NP + K][J] in FP32, the layout of k_pside_bf16.  ichunk is a
// multiple of 32.  SNPs past J have zero split rows (s = 0): their weights are
// not finite but reach only their own SNP's row (Hessian) or column (gradient).
template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_pside_mfma(const uint8_t* __restrict__ bed, int64_t pitch, int64_t I, int64_t J, int64_t ichunk,
                  const __bf16* __restrict__ PT, const uint8_t* __restrict__ img, float* __restrict__ part) {
  using G = PImg<K>;
  constexpr int KD = G::KD, NM = G::NM, NP = G::NP, NT = G::NT16, NA = NP + K;
  constexpr bool G16 = K <= 8;
  __shared__ __align__(16) uint8_t sbuf0[G::BYTES];
  __shared__ __align__(16) uint8_t sbuf1[G::BYTES];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: the copies' LDS base goes to M0
  const int lane = threadIdx.x & 63, h = lane >> 5, cn = lane & 31;
  const int n16 = lane & 15, g16 = lane >> 4;
  const int64_t nbx = (J + 127) / 128;
  const Tile2 tile = xcd_tile(nbx);  // workgroups of one individual chunk share its image slots: one XCD
  const int64_t j0 = tile.x * 128 + 32 * wave;
  const int64_t ib = tile.y * ichunk;
  const int64_t ie = (ib + ichunk < I) ? ib + ichunk : I;
  const bool okj = j0 + cn < J;

  bf16x8 bp[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) bp[m] = load_b8(PT, KD, (int)0, (j0 + cn) * KD + 16 * m + 8 * h);

  const int o_sq = G::SQ + cn * G::RS + 16 * h;
  int o_h[NT];  // Hessian B: qq row 16 t + n16, columns of group g16
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int n = 16 * t + n16;
    o_h[t] = (n < NP ? G::QQ + n * G::RT : G::ZR) + 16 * g16;
  }
  const int gm_ = G16 ? (n16 & 7) : (cn & 15);
  const int grow = (G16 ? (n16 < 8 ? 0 : K) : (cn < 16 ? 0 : K)) + gm_;
  const int o_gq = (gm_ < K ? G::GQ + grow * G::RT : G::ZR) + (G16 ? 16 * (g16 >> 1) : 16 * h);

  // genotypes: this lane's SNP row, 8 bytes (32 individuals) per block
  const uint8_t* __restrict__ grow_p = bed + pitch * (okj ? j0 + cn : 0);
  auto gload = [&](int64_t i) {
    uint2 w = {0x55555555u, 0x55555555u};
    if (okj && i < ie) w = *reinterpret_cast<const uint2*>(grow_p + (i >> 2));
    return w;
  };
  const uint8_t* __restrict__ isrc = img + lane * 16;
  auto copy_block = [&](int64_t blk, uint8_t* dst) {
    for (int c = wave; c < G::NCH; c += 4)
      copy_to_lds16(isrc + blk * G::BYTES + c * 1024, dst + c * 1024);
  };
  auto s_mfma = [&](const uint8_t* __restrict__ buf) {  // s from a buffer's individual split rows
    f32x16 sv = (f32x16){};
#pragma unroll
    for (int m = 0; m < NM; ++m)
      sv = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds_b8(buf + o_sq + 32 * m), bp[m], sv, 0, 0, 0);
    return sv;
  };

  f32x4 acc[2][NT];  // Hessian, SNPs 0-15 / 16-31 of the wave
  f32x4 z16[2] = {(f32x4){}, (f32x4){}};
  f32x16 z32 = (f32x16){};
#pragma unroll
  for (int gi = 0; gi < 2; ++gi)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[gi][t] = (f32x4){};
  copy_block(ib / 32, sbuf0);
  uint2 wn = gload(ib);
  __syncthreads();
  f32x16 sa = s_mfma(sbuf0);
  copy_block(ib / 32 + 1, sbuf1);
  __syncthreads();
  int nb = 1;
  for (int64_t is = ib; is < ie; is += 32, nb ^= 1) {
    const uint8_t* __restrict__ cur = nb ? sbuf1 : sbuf0;
    if (is + 32 < ie) copy_block(is / 32 + 2, nb ? sbuf0 : sbuf1);
    const f32x16 sn = s_mfma(cur);
    uint2 w = wn;
    if (is + 32 < ie) wn = gload(is + 32);
    if (is + 32 > ie) {  // uniform: individuals past the chunk end are missing
      uint32_t keep[2] = {0u, 0u};
#pragma unroll
      for (int t = 0; t < 32; ++t)
        if (is + t < ie) keep[t >> 4] |= 3u << (2 * (t & 15));
      w.x = (w.x & keep[0]) | (0x55555555u & ~keep[0]);
      w.y = (w.y & keep[1]) | (0x55555555u & ~keep[1]);
    }
    uint32_t wh[2][4], uh[2][4], ul[2][4];  // per half: packed weights w, u leading, u residual
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int xb = 2 * a + h;  // byte of individuals 8a + 4h .. 8a + 4h + 3
      const uint32_t byte = ((xb < 4 ? w.x : w.y) >> (8 * (xb & 3))) & 0xffu;
      const uint32_t codes = (byte | (byte << 6) | (byte << 12) | (byte << 18)) & 0x03030303u;
      const uint32_t gb = g_bytes(codes), hb = h_bytes(codes);
#pragma unroll
      for (int b = 0; b < 4; b += 2) {
        float wv[2], uv[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          const int r = 4 * a + b + x;
          const float sv = sa[r], tv = 1.0f - sv;
          const float is_ = __builtin_amdgcn_rcpf(sv), it = __builtin_amdgcn_rcpf(tv);  // as on the Q side
          const float cc = byte_f(gb, b + x) * is_, dd = byte_f(hb, b + x) * it;
          wv[x] = cc * is_ + dd * it;
          uv[x] = (cc - dd) * (float)(1.0 / PGRAD_SCALE);
        }
        const int half = a >> 1, e = ((4 * a + b) >> 1) & 3;
        wh[half][e] = pack_bf16x2(wv[0], wv[1]);
        split_pack_f16(uv[0], uv[1], uh[half][e], ul[half][e]);
      }
    }
    if constexpr (!G16) {
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const f16x8 gq = lds_h8(cur + o_gq + 32 * half);
        z32 = __builtin_amdgcn_mfma_f32_32x32x16_f16(gq, h8(uh[half]), z32, 0, 0, 0);
        z32 = __builtin_amdgcn_mfma_f32_32x32x16_f16(gq, h8(ul[half]), z32, 0, 0, 0);
      }
    }
    swap16(wh[0], wh[1]);  // wh[0]: SNPs 0-15, wh[1]: 16-31; groups (half 0, h 0) (half 1, h 0) (half 0, h 1) (half 1, h 1)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const bf16x8 bt = lds_b8(cur + o_h[t]);
      acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b8(wh[0]), bt, acc[0][t], 0, 0, 0);
      acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b8(wh[1]), bt, acc[1][t], 0, 0, 0);
    }
    if constexpr (G16) {
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        swap16(uh[half], ul[half]);  // uh: SNPs 0-15, ul: 16-31; groups (hi h0) (lo h0) (hi h1) (lo h1)
        const f16x8 gq = lds_h8(cur + o_gq + 32 * half);
        z16[0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(gq, h8(uh[half]), z16[0], 0, 0, 0);
        z16[1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(gq, h8(ul[half]), z16[1], 0, 0, 0);
      }
    }
    sa = sn;
    __syncthreads();
  }
  float* __restrict__ out = part + tile.y * NA * J;
#pragma unroll
  for (int gi = 0; gi < 2; ++gi)  // D[SNP][pair]: lane (pair n16, g16) holds SNPs 4 g16 + r
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int pn = 16 * t + n16;
      if (pn < NP) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t col = j0 + 16 * gi + 4 * g16 + r;
          if (col < J) out[(int64_t)pn * J + col] = acc[gi][t][r];
        }
      }
    }
  if constexpr (G16) {
#pragma unroll
    for (int gi = 0; gi < 2; ++gi) {
      const int64_t col = j0 + 16 * gi + n16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float v = z16[gi][r] + __shfl_xor(z16[gi][r], 32);
        const int k = 4 * g16 + r;
        if (g16 < 2 && k < K && col < J) out[(int64_t)(NP + k) * J + col] = v;
      }
    }
  } else if (okj) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int row = 8 * (r >> 2) + 4 * h + (r & 3);
      if (row < K) out[(int64_t)(NP + row) * J + j0 + cn] = z32[r] + z32[r + 8];
    }
  }
}

// ------------------------------------------------------------------ launchers
template <int K, bool EM>
static void qimg_k(const double* p, int64_t J, void* img, hipStream_t s) {
  hipLaunchKernelGGL((k_qimg<K, EM>), dim3((unsigned)((J + 31) / 32 + 1)), dim3(256), 0, s, p, J, (uint8_t*)img);
}
template <int K>
static void qside_mfma_k(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk, int nchunk,
                         const void* QT, const void* img, void* part, const double* q64, hipStream_t s) {
  const dim3 grid((unsigned)((I + 127) / 128 * nchunk));  // 1-D, tiles ordered per XCD (xcd_tile)
  if (q64)
    hipLaunchKernelGGL((k_qside_mfma<K, true>), grid, dim3(256), 0, s, bed, pitch, I, J, jchunk, (const __bf16*)QT,
                       (const uint8_t*)img, part, q64);
  else
    hipLaunchKernelGGL((k_qside_mfma<K, false>), grid, dim3(256), 0, s, bed, pitch, I, J, jchunk, (const __bf16*)QT,
                       (const uint8_t*)img, part, q64);
}

template <int K>
static void pimg_k(const double* q, int64_t I, void* img, hipStream_t s) {
  hipLaunchKernelGGL((k_pimg<K>), dim3((unsigned)((I + 31) / 32 + 1)), dim3(256), 0, s, q, I, (uint8_t*)img);
}
template <int K>
static void pside_mfma_k(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t ichunk, int nchunk,
                         const void* PT, const void* img, float* part, hipStream_t s) {
  const dim3 grid((unsigned)((J + 127) / 128 * nchunk));  // 1-D, tiles ordered per XCD (xcd_tile)
  hipLaunchKernelGGL((k_pside_mfma<K>), grid, dim3(256), 0, s, bed, pitch, I, J, ichunk, (const __bf16*)PT,
                     (const uint8_t*)img, part);
}

int pimg_bytes(int K) {
  int b = 0;
  ADMX_DISPATCH_K(K, (b = PImg<KK>::BYTES));
  return b;
}
void launch_pimg(int K, const double* q, int64_t I, void* img, hipStream_t s) {
  ADMX_DISPATCH_K(K, (pimg_k<KK>(q, I, img, s)));
}
void launch_pside_mfma(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t ichunk, int nchunk,
                       const void* PT, const void* img, float* part, hipStream_t s) {
  ADMX_DISPATCH_K(K, (pside_mfma_k<KK>(bed, pitch, I, J, ichunk, nchunk, PT, img, part, s)));
}

int qimg_bytes(int K, bool em) {
  int b = 0;
  ADMX_DISPATCH_K(K, (b = em ? QImg<KK, true>::BYTES : QImg<KK, false>::BYTES));
  return b;
}
void launch_qimg(int K, const double* p, int64_t J, void* img, bool em, hipStream_t s) {
  if (em) {
    ADMX_DISPATCH_K(K, (qimg_k<KK, true>(p, J, img, s)));
  } else {
    ADMX_DISPATCH_K(K, (qimg_k<KK, false>(p, J, img, s)));
  }
}
void launch_qside_mfma(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk, int nchunk,
                       const void* QT, const void* img, float* part, hipStream_t s) {
  ADMX_DISPATCH_K(K, (qside_mfma_k<KK>(bed, pitch, I, J, jchunk, nchunk, QT, img, part, nullptr, s)));
}
void launch_emq_img(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk, int nchunk,
                    const double* q, const void* QT, const void* img, double* part, hipStream_t s) {
  ADMX_DISPATCH_K(K, (qside_mfma_k<KK>(bed, pitch, I, J, jchunk, nchunk, QT, img, part, q, s)));
}

}  // namespace admx
