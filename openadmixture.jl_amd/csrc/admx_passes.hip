// admx_passes.hip -- the genotype passes: log-likelihood, FRAPPE-EM and the
// per-genotype gradient + K x K Hessian accumulation, decoded straight from
// the 2-bit SNP-major BED in HBM.
//
// Q side (log-likelihood, em_q, update_q; src/loops.jl:517-539, 610-632,
// 867-894): one lane per individual i, a workgroup streams a chunk of SNPs.
// For each SNP every lane needs the same p[:,j] -- a wave-uniform address --
// so P arrives through scalar loads into SGPRs and feeds the FMAs as the
// scalar operand; the lane's own q[:,i] and all accumulators live in VGPRs.
// Genotypes arrive 16 SNPs at a time through a per-wave LDS transpose.
// P side (em_p, update_p; src/loops.jl:700-721, 923-950): one lane per SNP j,
// streaming individuals; q[:,i] is wave-uniform (SGPRs, loaded one genotype
// ahead) and each lane's SNP row arrives in 64-byte tiles parked in LDS.
//
// IMP selects the mean-imputation loops (loops.jl:493-514, 589-608, 679-698,
// 840-865, 896-921): a missing genotype counts as g = mu_j (the SNP's mean,
// gm = mu_j, hm = 2 - mu_j) instead of being skipped.
//
// Every chunk writes its partial sums to HBM; a second kernel reduces them in
// a fixed order, so results are bitwise reproducible run to run (the
// reference's CUDA path uses float atomics).
#include <type_traits>
#include "admx_common.hpp"
#include "admx_internal.hpp"

namespace admx {

// ------------------------------------------------------------------ Q side
// Per-genotype algebra, with s = q_i . p_j, t = 1 - s, gm = m*g, hm = m*(2-g):
//   c = gm/s, d = hm/t, a = c/s, b = d/t
//   Xtz_q[k,i]    += c p_k + d (1 - p_k)
//   XtX_q[k2,k,i] += a p_k p_k2 + b (1 - p_k)(1 - p_k2)
// FP64 accumulates W = sum (a+b) p p', hb = sum b p, sb = sum b and sd = sum d,
// from which the finalize forms the equivalent rank-1 form
//   XtX = W + sb 11' - 1 hb' - hb 1'
//   Xtz = W q - hb + sd 1
// (one FMA per Hessian entry instead of two).  The gradient needs no
// accumulators of its own: c = a s, d = b t and s = q.p, t = 1 - s give
// sum c p = sum a p (p.q) and sum d (1-p) = sum d - sum b p + sum b s p, so
// sum c p + d (1-p) = W q - hb + sd exactly, whatever sum(q) is.  Both forms
// cancel at most a factor 1/min(p,1-p)^2 <= 1e10 of the largest b term on a
// 1e-16 unit roundoff.  FP32 keeps the direct two-term form, where that factor
// would matter.
// PART (update mode): 0 accumulates everything; at large K the pass runs twice,
// PART 1 with the first half of the Hessian triangle plus the vector sums and
// PART 2 with the second half, so that each launch holds half the accumulators
// (two waves per SIMD instead of one at K = 12).
// STRICT (FP64 log-likelihood): the term-by-term form.  With a state-check
// flag (a.strict) the product-form and term-by-term instances are launched one
// after the other and the one that does not match the flag returns at once, so
// the product form keeps its registers and occupancy.
template <typename T, int K, int MODE, bool TRICK, bool IMP, int PART = 0, bool STRICT = false>
__global__ __launch_bounds__(256) void k_qside(PassArgs a) {
  if constexpr (MODE == MODE_LL) {
    if (a.strict && ((*a.strict != 0) != STRICT)) return;  // uniform: the other instance runs
    // product form next to k_qside_ll2: runs only where that pass stepped aside
    // (the other state was flagged)
    if constexpr (!STRICT) {
      if (a.strict_other && *a.strict_other == 0) return;
    }
  }
  constexpr int NT = Tri<K>::N;
  constexpr int N0 = (PART == 2) ? NT / 2 : 0, N1 = (PART == 1) ? NT / 2 : NT;  // owned triangle entries
  constexpr bool VEC = PART != 2;                                               // owns hb, sb, sd (z)
  const T* __restrict__ q = static_cast<const T*>(a.q);
  const T* __restrict__ p = static_cast<const T*>(a.p);
  const int64_t Ilen = a.i1 - a.i0;
  const int64_t nbx = (Ilen + 255) / 256;
  const Tile2 tile = xcd_tile(nbx);
  const int64_t il = tile.x * 256 + threadIdx.x;
  const bool valid = il < Ilen;
  const int64_t i = a.i0 + (valid ? il : 0);
  const int64_t jb = a.j0 + tile.y * a.chunk;
  const int64_t je = (a.j1 < jb + a.chunk) ? a.j1 : jb + a.chunk;

  T qi[K];
#pragma unroll
  for (int k = 0; k < K; ++k) qi[k] = q[k + (int64_t)K * i];
  T H[MODE == MODE_UPDATE ? N1 - N0 : 1];
  T z[K];
  T hb[(MODE == MODE_UPDATE && TRICK) ? K : 1];
  T sb = 0, sd = 0;
  double acc = 0.0, prod = 1.0;
  int esum = 0;
#pragma unroll
  for (int n = 0; n < (MODE == MODE_UPDATE ? N1 - N0 : 1); ++n) H[n] = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) z[k] = 0;
#pragma unroll
  for (int k = 0; k < ((MODE == MODE_UPDATE && TRICK) ? K : 1); ++k) hb[k] = 0;

  const T* __restrict__ mu = static_cast<const T*>(a.mu);
  // per-code constants in LDS: the product-form log-likelihood's factors, or
  // (m*g, m*(2-g)) for every other instance
  constexpr bool LLPROD = MODE == MODE_LL && sizeof(T) == 8 && !IMP && !STRICT;
  using CE = typename std::conditional<LLPROD, CodeLL, CodeGH<T>>::type;
  __shared__ CE ctab[4];
  code_table_init(ctab);
  // (m*g, m*(2-g)) of code c in SNP j (gh = ctab[c], read one genotype ahead);
  // imputed mean for a missing code under IMP.
  auto weights = [&](uint32_t c, const CodeGH<T>& gh, int64_t j, T& gm, T& hm) {
    gm = gh.g;
    hm = gh.h;
    if constexpr (IMP) {
      if (c == 1u) {
        const T mj = mu[j];
        gm = mj;
        hm = (T)2 - mj;
      }
    }
  };
  // Per-genotype accumulation for genotype code c of SNP j and P column pv.
  auto step = [&](uint32_t c, const CE& gh, int64_t j, const T (&pv)[K]) {
    T s = qi[0] * pv[0];
#pragma unroll
    for (int k = 1; k < K; ++k) s = __builtin_fma(qi[k], pv[k], s);
    const T t = (T)1 - s;
    if constexpr (MODE == MODE_LL) {
      if constexpr (LLPROD) {
        // prod_j s^g (1-s)^(2-g), renormalised every 8 SNPs; one log per lane.
        const double x = __builtin_fma(gh.ax, (double)s, gh.bx);
        const double y = __builtin_fma(gh.ay, (double)s, gh.by);
        prod *= x * y;
      } else {
        // Terms in T, accumulated in Float64 (src/loops.jl:523, 535); a missing
        // entry adds 0 log s + 0 log(1-s), NaN where s is 0 or 1, as there.
        T gm, hm;
        weights(c, gh, j, gm, hm);
        const T term = gm * tlog<T>(s) + hm * tlog<T>(t);
        acc += (double)term;
      }
    } else {
      T gm, hm;
      weights(c, gh, j, gm, hm);
      const T r = rcp(s * t);
      const T is = t * r, it = s * r;
      const T cc = gm * is, dd = hm * it;
      if constexpr (MODE == MODE_EM && sizeof(T) == 8) {
        // FP64: sum c p + d (1-p) as sum (c-d) p + (sum d) 1, K+1 FMAs per genotype
        const T u = cc - dd;
        sd += dd;
#pragma unroll
        for (int k = 0; k < K; ++k) z[k] = __builtin_fma(u, pv[k], z[k]);
      } else if constexpr (MODE == MODE_EM) {
#pragma unroll
        for (int k = 0; k < K; ++k) z[k] = __builtin_fma(cc, pv[k], __builtin_fma(dd, (T)1 - pv[k], z[k]));
      } else if constexpr (TRICK) {
        const T aa = cc * is, bb = dd * it;
        const T w = aa + bb;
        if constexpr (VEC) {
          sb += bb;
          sd += dd;
        }
        int n = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const T wp = w * pv[k];
          if constexpr (VEC) hb[k] = __builtin_fma(bb, pv[k], hb[k]);
#pragma unroll
          for (int k2 = k; k2 < K; ++k2, ++n)
            if (n >= N0 && n < N1) H[n - N0] = __builtin_fma(wp, pv[k2], H[n - N0]);
        }
      } else {
        const T aa = cc * is, bb = dd * it;
        int n = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const T mk = (T)1 - pv[k];
          const T ap = aa * pv[k], bm = bb * mk;
          if constexpr (VEC) z[k] = __builtin_fma(cc, pv[k], __builtin_fma(dd, mk, z[k]));
#pragma unroll
          for (int k2 = k; k2 < K; ++k2, ++n)
            if (n >= N0 && n < N1)
              H[n - N0] = __builtin_fma(ap, pv[k2], __builtin_fma(bm, (T)1 - pv[k2], H[n - N0]));
        }
      }
    }
  };
  auto renorm = [&]() {
    if constexpr (MODE == MODE_LL && sizeof(T) == 8 && !IMP && !STRICT) {
      int e;
      prod = __builtin_frexp(prod, &e);
      esum += e;
    }
  };

  // Genotype codes, 16 SNPs per step.  The wave's 64 individuals occupy at most
  // 17 bytes of an SNP row, inside the 32-byte window starting at the 16-byte
  // boundary `wbase`.  Lane l fetches the two 32-bit words 4(l%4) and 16+4(l%4)
  // of that window in SNP row l/4, and the 16 rows x 32 bytes are transposed
  // through 512 bytes of LDS per wave, after which each lane reads the bytes of
  // its own individual in the 16 SNPs with one 128-bit load.  The window can
  // run up to 31 bytes past a row's last genotype byte, into the next row or
  // the BED's padding rows; those bytes belong to lanes past the last
  // individual, whose sums are discarded.  The P column of SNP j+1 is loaded
  // while SNP j is computed; the words of step js+16 while step js is.
  __shared__ __align__(16) uint8_t tbuf[4][512];
  uint8_t* __restrict__ tb = tbuf[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int64_t wrow0 = tile.x * 256 + (threadIdx.x & ~63);
  const bool wave_live = wrow0 < Ilen;  // wave-uniform
  const int64_t ifirst = a.i0 + wrow0;
  const int64_t wbase = (ifirst >> 2) & ~(int64_t)15;
  const int bsel = (int)(((ifirst + lane) >> 2) - wbase);  // 0..31
  const uint32_t shw = (uint32_t)((ifirst + lane) & 3) * 2u;
  const uint8_t* __restrict__ wrow = a.bed + wbase + 4 * (lane & 3) + a.pitch * (lane >> 2);

  int64_t js = jb;
  if (wave_live && je - jb >= 16) {
    uint32_t w0 = *reinterpret_cast<const uint32_t*>(wrow + a.pitch * jb);
    uint32_t w1 = *reinterpret_cast<const uint32_t*>(wrow + a.pitch * jb + 16);
    T pn[K];
#pragma unroll
    for (int k = 0; k < K; ++k) pn[k] = p[(int64_t)K * jb + k];
    for (; js + 16 <= je; js += 16) {
      // byte e of word w0 (w1) is window byte 4(l%4)+e (+16) of SNP row l/4 -> tb[byte][row]
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        tb[(4 * (lane & 3) + e) * 16 + (lane >> 2)] = (uint8_t)(w0 >> (8 * e));
        tb[(16 + 4 * (lane & 3) + e) * 16 + (lane >> 2)] = (uint8_t)(w1 >> (8 * e));
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const u32x4 cur = *reinterpret_cast<const u32x4*>(tb + 16 * bsel);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (js + 32 <= je) {
        w0 = *reinterpret_cast<const uint32_t*>(wrow + a.pitch * (js + 16));
        w1 = *reinterpret_cast<const uint32_t*>(wrow + a.pitch * (js + 16) + 16);
      }
      auto code = [&](int u) { return (cur[u >> 2] >> (8 * (u & 3) + shw)) & 3u; };
      CE ghn = ctab[code(0)];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        T pv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) pv[k] = pn[k];
        const int64_t jn = (u < 15) ? js + u + 1 : ((js + 16 < je) ? js + 16 : js + 15);
#pragma unroll
        for (int k = 0; k < K; ++k) pn[k] = p[(int64_t)K * jn + k];
        const CE gh = ghn;
        if (u < 15) ghn = ctab[code(u + 1)];
        // the loads above (the next P column by scalar loads, the next code's
        // constants from LDS) stay ahead of this genotype's arithmetic
        __builtin_amdgcn_sched_barrier(0);
        step(code(u), gh, js + u, pv);
        if ((u & 7) == 7) renorm();
      }
    }
  }
  // The chunk's last < 16 SNPs: one byte per SNP (lanes past the last
  // individual read individual i0's byte).
  if (wave_live) {
    const uint8_t* __restrict__ gp = a.bed + (i >> 2);
    const uint32_t sh = (uint32_t)(i & 3) * 2u;
    for (; js < je; ++js) {
      T pv[K];
#pragma unroll
      for (int k = 0; k < K; ++k) pv[k] = p[(int64_t)K * js + k];
      const uint32_t c = (gp[a.pitch * js] >> sh) & 3u;
      step(c, ctab[c], js, pv);
      renorm();
    }
  }

  if constexpr (MODE == MODE_LL) {
    __shared__ double sh_red[256];
    // product form: acc is 0 and prod holds the terms; term by term: prod = 1, esum = 0
    if constexpr (sizeof(T) == 8 && !IMP) acc += log(prod) + (double)esum * 0.69314718055994530942;
    const double tot = block_sum<256>(valid ? acc : 0.0, sh_red);
    if (threadIdx.x == 0) static_cast<double*>(a.part)[tile.y * nbx + tile.x] = tot;
    return;
  } else {
    if (!valid) return;
    T* __restrict__ part = static_cast<T*>(a.part);
    const int NA = (MODE == MODE_EM) ? (sizeof(T) == 8 ? K + 1 : K) : (TRICK ? NT + K + 2 : NT + K);
    T* out = part + tile.y * NA * Ilen + il;
    int o = 0;
    if constexpr (MODE == MODE_UPDATE) {
#pragma unroll
      for (int n = N0; n < N1; ++n) out[(int64_t)n * Ilen] = H[n - N0];
      o = NT;
    }
    if constexpr (VEC) {
      if constexpr (MODE == MODE_UPDATE && TRICK) {
#pragma unroll
        for (int k = 0; k < K; ++k) out[(o++) * Ilen] = hb[k];
        out[(o++) * Ilen] = sb;
        out[(o++) * Ilen] = sd;
      } else {
#pragma unroll
        for (int k = 0; k < K; ++k) out[(o++) * Ilen] = z[k];
        if constexpr (MODE == MODE_EM && sizeof(T) == 8) out[(o++) * Ilen] = sd;
      }
    }
  }
}

// ------------------------------------------------------------------ two log-likelihoods
// The product-form log-likelihoods of two states (q, p) and (q2, p2) in one
// pass over the genotypes: each genotype's code and its factor constants are
// read once for both, and each state's sum is formed exactly as k_qside's
// product-form instance forms it (same order of products, renormalisation,
// log and block sum; partials of the second state at a.part2).  Runs only
// when both state flags are clear; a flagged state takes the single-state
// instances (see launch_qside_ll2).
template <int K>
__global__ __launch_bounds__(256) void k_qside_ll2(PassArgs a) {
  if (*a.strict != 0 || *a.strict2 != 0) return;  // uniform
  const double* __restrict__ q = static_cast<const double*>(a.q);
  const double* __restrict__ p = static_cast<const double*>(a.p);
  const double* __restrict__ q2 = static_cast<const double*>(a.q2);
  const double* __restrict__ p2 = static_cast<const double*>(a.p2);
  const int64_t Ilen = a.i1 - a.i0;
  const int64_t nbx = (Ilen + 255) / 256;
  const Tile2 tile = xcd_tile(nbx);
  const int64_t il = tile.x * 256 + threadIdx.x;
  const bool valid = il < Ilen;
  const int64_t i = a.i0 + (valid ? il : 0);
  const int64_t jb = a.j0 + tile.y * a.chunk;
  const int64_t je = (a.j1 < jb + a.chunk) ? a.j1 : jb + a.chunk;

  double qa[K], qb[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    qa[k] = q[k + (int64_t)K * i];
    qb[k] = q2[k + (int64_t)K * i];
  }
  double prod = 1.0, prod2 = 1.0;
  int esum = 0, esum2 = 0;
  __shared__ CodeLL ctab[4];
  code_table_init(ctab);
  auto step = [&](const CodeLL& e, const double (&pv)[K], const double (&pw)[K]) {
    double s = qa[0] * pv[0], s2 = qb[0] * pw[0];
#pragma unroll
    for (int k = 1; k < K; ++k) {
      s = __builtin_fma(qa[k], pv[k], s);
      s2 = __builtin_fma(qb[k], pw[k], s2);
    }
    prod *= __builtin_fma(e.ax, s, e.bx) * __builtin_fma(e.ay, s, e.by);
    prod2 *= __builtin_fma(e.ax, s2, e.bx) * __builtin_fma(e.ay, s2, e.by);
  };
  auto renorm = [&]() {
    int ex;
    prod = __builtin_frexp(prod, &ex);
    esum += ex;
    prod2 = __builtin_frexp(prod2, &ex);
    esum2 += ex;
  };

  // Genotype codes 16 SNPs per step through a per-wave LDS transpose, as in k_qside.
  __shared__ __align__(16) uint8_t tbuf[4][512];
  uint8_t* __restrict__ tb = tbuf[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  const int64_t wrow0 = tile.x * 256 + (threadIdx.x & ~63);
  const bool wave_live = wrow0 < Ilen;  // wave-uniform
  const int64_t ifirst = a.i0 + wrow0;
  const int64_t wbase = (ifirst >> 2) & ~(int64_t)15;
  const int bsel = (int)(((ifirst + lane) >> 2) - wbase);
  const uint32_t shw = (uint32_t)((ifirst + lane) & 3) * 2u;
  const uint8_t* __restrict__ wrow = a.bed + wbase + 4 * (lane & 3) + a.pitch * (lane >> 2);

  int64_t js = jb;
  if (wave_live && je - jb >= 16) {
    uint32_t w0 = *reinterpret_cast<const uint32_t*>(wrow + a.pitch * jb);
    uint32_t w1 = *reinterpret_cast<const uint32_t*>(wrow + a.pitch * jb + 16);
    double pn[K], pn2[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      pn[k] = p[(int64_t)K * jb + k];
      pn2[k] = p2[(int64_t)K * jb + k];
    }
    for (; js + 16 <= je; js += 16) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        tb[(4 * (lane & 3) + e) * 16 + (lane >> 2)] = (uint8_t)(w0 >> (8 * e));
        tb[(16 + 4 * (lane & 3) + e) * 16 + (lane >> 2)] = (uint8_t)(w1 >> (8 * e));
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const u32x4 cur = *reinterpret_cast<const u32x4*>(tb + 16 * bsel);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (js + 32 <= je) {
        w0 = *reinterpret_cast<const uint32_t*>(wrow + a.pitch * (js + 16));
        w1 = *reinterpret_cast<const uint32_t*>(wrow + a.pitch * (js + 16) + 16);
      }
      auto code = [&](int u) { return (cur[u >> 2] >> (8 * (u & 3) + shw)) & 3u; };
      CodeLL en = ctab[code(0)];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        double pv[K], pw[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          pv[k] = pn[k];
          pw[k] = pn2[k];
        }
        const int64_t jn = (u < 15) ? js + u + 1 : ((js + 16 < je) ? js + 16 : js + 15);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          pn[k] = p[(int64_t)K * jn + k];
          pn2[k] = p2[(int64_t)K * jn + k];
        }
        const CodeLL e = en;
        if (u < 15) en = ctab[code(u + 1)];
        __builtin_amdgcn_sched_barrier(0);  // the loads above stay a genotype ahead
        step(e, pv, pw);
        if ((u & 7) == 7) renorm();
      }
    }
  }
  if (wave_live) {
    const uint8_t* __restrict__ gp = a.bed + (i >> 2);
    const uint32_t sh = (uint32_t)(i & 3) * 2u;
    for (; js < je; ++js) {
      double pv[K], pw[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        pv[k] = p[(int64_t)K * js + k];
        pw[k] = p2[(int64_t)K * js + k];
      }
      step(ctab[(gp[a.pitch * js] >> sh) & 3u], pv, pw);
      renorm();
    }
  }

  __shared__ double sh_red[256];
  double acc = 0.0, acc2 = 0.0;
  acc += log(prod) + (double)esum * 0.69314718055994530942;
  acc2 += log(prod2) + (double)esum2 * 0.69314718055994530942;
  const double tot = block_sum<256>(valid ? acc : 0.0, sh_red);
  const double tot2 = block_sum<256>(valid ? acc2 : 0.0, sh_red);
  if (threadIdx.x == 0) {
    static_cast<double*>(a.part)[tile.y * nbx + tile.x] = tot;
    static_cast<double*>(a.part2)[tile.y * nbx + tile.x] = tot2;
  }
}

// ------------------------------------------------------------------ P side
//   Xtz_p[k,j]    += (c - d) q_k               (exact dLL/dp)
//   XtX_p[k2,k,j] += (a + b) q_k q_k2          (exact -d2LL/dp2)
//   EM: p_tmp += c q_k p_k ; p_next += d q_k (1 - p_k)
// PART as for k_qside: at large K the update pass runs twice, each launch owning
// half of the Hessian triangle (PART 1 also the gradient).
template <typename T, int K, int MODE, bool IMP, int PART = 0>
__global__ __launch_bounds__(256) void k_pside(PassArgs a) {
  constexpr int NT = Tri<K>::N;
  constexpr int N0 = (PART == 2) ? NT / 2 : 0, N1 = (PART == 1) ? NT / 2 : NT;
  constexpr bool VEC = PART != 2;
  const T* __restrict__ q = static_cast<const T*>(a.q);
  const T* __restrict__ p = static_cast<const T*>(a.p);
  const int64_t Jlen = a.j1 - a.j0;
  const int64_t jl = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool valid = jl < Jlen;
  const int64_t j = a.j0 + (valid ? jl : 0);
  const int64_t ia0 = (a.i0 & ~(int64_t)255) + (int64_t)blockIdx.y * a.chunk;  // chunk: multiple of 256
  const int64_t ia1 = (a.i1 < ia0 + a.chunk) ? a.i1 : ia0 + a.chunk;

  T pj[K];
#pragma unroll
  for (int k = 0; k < K; ++k) pj[k] = p[k + (int64_t)K * j];
  const uint8_t* __restrict__ col = a.bed + a.pitch * j;
  __shared__ CodeGH<T> ctab[4];
  code_table_init(ctab);
  T mj = 0;
  if constexpr (IMP) mj = static_cast<const T*>(a.mu)[j];

  T H[MODE == MODE_UPDATE ? N1 - N0 : 1];
  T z[K], e2[MODE == MODE_EM ? K : 1];
#pragma unroll
  for (int n = 0; n < (MODE == MODE_UPDATE ? N1 - N0 : 1); ++n) H[n] = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) z[k] = 0;
#pragma unroll
  for (int k = 0; k < (MODE == MODE_EM ? K : 1); ++k) e2[k] = 0;

  // Per-genotype work on the wave-uniform q column qn of the current individual
  // (scalar registers), while the next individual's column is being loaded:
  // body() takes qn, issues the scalar loads of individual `next` into qn and
  // then computes, so each wait covers one load issued a genotype earlier.
  // inr = 0 marks an individual outside [i0, i1) whose code was set to missing:
  // it adds nothing, also under IMP (only genuinely missing entries are imputed).
  // Individual indices are below 2^31 (checked at context creation): the clamp
  // compares 32-bit values, which the scalar unit does (a 64-bit signed compare
  // would take a vector instruction per genotype).
  T qn[K];
  const int32_t I32 = (int32_t)a.I;
  // q is read through the constant address space (no store of the pass can
  // alias it): the wave-uniform loads then stay scalar loads next to the EM
  // pass's scheduling barrier below.
  auto loadq = [&](int64_t i) {
    const int32_t i32 = (int32_t)i;
    const __attribute__((address_space(4))) T* src =
        (const __attribute__((address_space(4))) T*)(q + (int64_t)K * (i32 < I32 ? i32 : I32 - 1));
#pragma unroll
    for (int k = 0; k < K; ++k) qn[k] = src[k];
  };
  auto body = [&](uint32_t c, const CodeGH<T>& gh, int64_t next, uint32_t inr) {
    T qv[K];
#pragma unroll
    for (int k = 0; k < K; ++k) qv[k] = qn[k];
    loadq(next);
    // the next q column's scalar loads and the next code's LDS read stay a
    // genotype ahead of this genotype's arithmetic: EM, and the update pass
    // from K = 12 (I = J = 100,000: em_p 12.25 -> 11.0 ms at K = 8, update_p
    // 39.5 -> 37.8 ms at K = 12; the update pass no faster at K = 4, 8 and 11,
    // about 1% slower at 9 and 10)
    if constexpr (MODE == MODE_EM || K >= 12) __builtin_amdgcn_sched_barrier(0);
    T s = qv[0] * pj[0];
#pragma unroll
    for (int k = 1; k < K; ++k) s = __builtin_fma(qv[k], pj[k], s);
    const T tt = (T)1 - s;
    T gm = gh.g, hm = gh.h;
    if constexpr (IMP) {
      if (c == 1u && inr) {
        gm = mj;
        hm = (T)2 - mj;
      }
    }
    const T r = rcp(s * tt);
    const T is = tt * r, it = s * r;
    const T cc = gm * is, dd = hm * it;
    if constexpr (MODE == MODE_EM) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        z[k] = __builtin_fma(cc, qv[k], z[k]);
        e2[k] = __builtin_fma(dd, qv[k], e2[k]);
      }
    } else {
      const T w2 = cc * is + dd * it, u = cc - dd;
      int n = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const T wq = w2 * qv[k];
        if constexpr (VEC) z[k] = __builtin_fma(u, qv[k], z[k]);
#pragma unroll
        for (int k2 = k; k2 < K; ++k2, ++n)
          if (n >= N0 && n < N1) H[n - N0] = __builtin_fma(wq, qv[k2], H[n - N0]);
      }
    }
  };

  // Genotypes in tiles of 256 individuals: the 64 bytes of the lane's SNP row
  // come with four 16-byte loads issued together, one tile ahead, so every
  // 64-byte segment of a row is requested once, while it is needed (a lane that
  // streamed its row word by word kept ~64 rows per wave live in L2 at once).
  // The landed tile is parked in LDS as [word][lane] (each lane touches only
  // its own column: no barrier) and read back one 32-bit word per 16 genotypes.
  // Tiles stay inside the padded pitch: a row holds round_up(ceil(I/4), 64) bytes.
  __shared__ uint32_t tw[16][256];
  u32x4 nx[4];
  auto fetch = [&](int64_t ib) {
    const u32x4* __restrict__ src = reinterpret_cast<const u32x4*>(col + (ib >> 2));
#pragma unroll
    for (int c = 0; c < 4; ++c) nx[c] = src[c];
  };
  if (ia0 < ia1) fetch(ia0);
  for (int64_t it = ia0; it < ia1; it += 256) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int e = 0; e < 4; ++e) tw[4 * c + e][threadIdx.x] = nx[c][e];
    if (it + 256 < ia1) fetch(it + 256);
    const int s0 = (it < a.i0) ? (int)((a.i0 - it) >> 4) : 0;
    const int s1 = (int)((((ia1 - it) < 256 ? (ia1 - it) : 256) + 15) >> 4);
    loadq(it + 16 * s0);
    for (int sw = s0; sw < s1; ++sw) {
      const int64_t ib = it + 16 * sw;
      uint32_t w = tw[sw][threadIdx.x];
      uint32_t keep = 0xffffffffu;
      if (ib < a.i0 || ib + 16 > a.i1) {  // uniform: mark out-of-range individuals missing
        keep = 0;
#pragma unroll
        for (int t = 0; t < 16; ++t)
          if (ib + t >= a.i0 && ib + t < a.i1) keep |= 3u << (2 * t);
        w = (w & keep) | (0x55555555u & ~keep);
      }
      CodeGH<T> ghn = ctab[w & 3u];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const CodeGH<T> gh = ghn;
        if (t < 15) ghn = ctab[(w >> (2 * t + 2)) & 3u];
        body((w >> (2 * t)) & 3u, gh, ib + t + 1, (keep >> (2 * t)) & 1u);
      }
    }
  }
  if (!valid) return;
  T* __restrict__ part = static_cast<T*>(a.part);
  const int NA = (MODE == MODE_EM) ? 2 * K : NT + K;
  const int64_t ldp = a.ldp ? a.ldp : Jlen;
  T* out = part + (int64_t)blockIdx.y * NA * ldp + a.col0 + jl;
  int o = 0;
  if constexpr (MODE == MODE_UPDATE) {
#pragma unroll
    for (int n = N0; n < N1; ++n) out[(int64_t)n * ldp] = H[n - N0];
    o = NT;
  }
  if constexpr (VEC) {
#pragma unroll
    for (int k = 0; k < K; ++k) out[(o++) * ldp] = z[k];
  }
  if constexpr (MODE == MODE_EM) {
#pragma unroll
    for (int k = 0; k < K; ++k) out[(o++) * ldp] = e2[k];
  }
}

// ------------------------------------------------------------------ launchers
template <typename T, int K, bool IMP>
static void qside_ki(int mode, const PassArgs& a, int nchunk, hipStream_t s) {
  const int64_t n = a.i1 - a.i0;
  dim3 grid((unsigned)(((n + 255) / 256) * nchunk));  // 1-D, tiles ordered per XCD (xcd_tile)
  constexpr bool trick = sizeof(T) == 8;
  // FP64 from K = 9 on, the triangle can run in two launches (see k_qside);
  // a.split < 0 takes the default threshold K >= 13.  One launch holds
  // K(K+1)/2 + K + 2 FP64 accumulators per lane: 235 / 255 VGPRs at K = 11 / 12
  // (two waves per SIMD; I = J = 100,000: 41.6 / 45.2 ms against 47.6 / 58.1 ms
  // for two launches), 296 at K = 13 (one wave).
  if (mode == MODE_UPDATE) {
    bool done = false;
    if constexpr (sizeof(T) == 8 && K >= 9) {
      if (a.split > 0 || (a.split < 0 && K >= 13)) {
        hipLaunchKernelGGL((k_qside<T, K, MODE_UPDATE, trick, IMP, 1>), grid, dim3(256), 0, s, a);
        hipLaunchKernelGGL((k_qside<T, K, MODE_UPDATE, trick, IMP, 2>), grid, dim3(256), 0, s, a);
        done = true;
      }
    }
    if (!done) hipLaunchKernelGGL((k_qside<T, K, MODE_UPDATE, trick, IMP>), grid, dim3(256), 0, s, a);
  }
  else if (mode == MODE_EM) hipLaunchKernelGGL((k_qside<T, K, MODE_EM, false, IMP>), grid, dim3(256), 0, s, a);
  else {
    hipLaunchKernelGGL((k_qside<T, K, MODE_LL, false, IMP>), grid, dim3(256), 0, s, a);
    if constexpr (sizeof(T) == 8 && !IMP)
      if (a.strict) hipLaunchKernelGGL((k_qside<T, K, MODE_LL, false, IMP, 0, true>), grid, dim3(256), 0, s, a);
  }
}
template <typename T, int K>
static void qside_k(int mode, const PassArgs& a, int nchunk, hipStream_t s) {
  if (a.mu) qside_ki<T, K, true>(mode, a, nchunk, s);
  else qside_ki<T, K, false>(mode, a, nchunk, s);
}

template <typename T, int K, bool IMP>
static void pside_ki(int mode, const PassArgs& a, int nchunk, hipStream_t s) {
  const int64_t n = a.j1 - a.j0;
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)nchunk);
  // FP64 from K = 9 on, the triangle can run in two launches (see k_pside);
  // a.split < 0 takes the default threshold K >= 13
  if (mode == MODE_UPDATE) {
    bool done = false;
    if constexpr (sizeof(T) == 8 && K >= 9) {
      if (a.split > 0 || (a.split < 0 && K >= 13)) {
        hipLaunchKernelGGL((k_pside<T, K, MODE_UPDATE, IMP, 1>), grid, dim3(256), 0, s, a);
        hipLaunchKernelGGL((k_pside<T, K, MODE_UPDATE, IMP, 2>), grid, dim3(256), 0, s, a);
        done = true;
      }
    }
    if (!done) hipLaunchKernelGGL((k_pside<T, K, MODE_UPDATE, IMP>), grid, dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_pside<T, K, MODE_EM, IMP>), grid, dim3(256), 0, s, a);
  }
}
template <typename T, int K>
static void pside_k(int mode, const PassArgs& a, int nchunk, hipStream_t s) {
  if (a.mu) pside_ki<T, K, true>(mode, a, nchunk, s);
  else pside_ki<T, K, false>(mode, a, nchunk, s);
}

#ifdef ADMX_PASSES_F32_FORM
// The FP32 build serves only the update passes of the FP32 form (ADMX_PREC_F32
// on an FP64 context, skipmissing loops): every context's other passes, FP32
// contexts included, run the FP64 kernels.
void launch_qside_f32_update(int K, const PassArgs& a, int nchunk, hipStream_t s) {
  ADMX_DISPATCH_K(K, (qside_ki<float, KK, false>(MODE_UPDATE, a, nchunk, s)));
}
void launch_pside_f32_update(int K, const PassArgs& a, int nchunk, hipStream_t s) {
  ADMX_DISPATCH_K(K, (pside_ki<float, KK, false>(MODE_UPDATE, a, nchunk, s)));
}
#else
template <typename T> void launch_qside(int mode, int K, const PassArgs& a, int nchunk, hipStream_t s) {
  ADMX_DISPATCH_K(K, (qside_k<T, KK>(mode, a, nchunk, s)));
}
template <typename T> void launch_pside(int mode, int K, const PassArgs& a, int nchunk, hipStream_t s) {
  ADMX_DISPATCH_K(K, (pside_k<T, KK>(mode, a, nchunk, s)));
}

template void launch_qside<ADMX_T>(int, int, const PassArgs&, int, hipStream_t);

template <int K>
static void qside_ll_strict_k(const PassArgs& a, int nchunk, hipStream_t s) {
  const dim3 grid((unsigned)(((a.i1 - a.i0 + 255) / 256) * nchunk));
  hipLaunchKernelGGL((k_qside<double, K, MODE_LL, false, false, 0, true>), grid, dim3(256), 0, s, a);
}
void launch_qside_ll_strict(int K, const PassArgs& a, int nchunk, hipStream_t s) {
  ADMX_DISPATCH_K(K, (qside_ll_strict_k<KK>(a, nchunk, s)));
}
template void launch_pside<ADMX_T>(int, int, const PassArgs&, int, hipStream_t);

template <int K>
static void qside_ll2_k(const PassArgs& a, int nchunk, hipStream_t s) {
  PassArgs a1 = a, a2 = a;
  a2.q = a.q2;
  a2.p = a.p2;
  a2.part = a.part2;
  a2.strict = a.strict2;
  // Up to K = 8 the fused kernel sums both states; the single-state instances
  // then run only for a flagged state (each returns at once where the fused
  // pass ran or where its own flag selects the other form).  From K = 9 on
  // (where the fused kernel holds too few waves) they run as two passes.
  if constexpr (K <= 8) {
    const dim3 grid((unsigned)(((a.i1 - a.i0 + 255) / 256) * nchunk));
    hipLaunchKernelGGL((k_qside_ll2<K>), grid, dim3(256), 0, s, a);
    a1.strict_other = a.strict2;
    a2.strict_other = a.strict;
  }
  qside_ki<double, K, false>(MODE_LL, a1, nchunk, s);
  qside_ki<double, K, false>(MODE_LL, a2, nchunk, s);
}
void launch_qside_ll2(int K, const PassArgs& a, int nchunk, hipStream_t s) {
  ADMX_DISPATCH_K(K, (qside_ll2_k<KK>(a, nchunk, s)));
}
#endif

}  // namespace admx
