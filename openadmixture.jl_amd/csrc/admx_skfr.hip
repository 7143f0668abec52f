// admx_skfr.hip -- sparse k-means with feature ranking (SKFR) on the device
// BED: the AIM pre-selection step of run_admixture (src/driver.jl:56-66,
// 80-101), which the reference delegates to SparseKmeansFeatureRanking.jl
// (sparsekmeans1 / sparsekmeans2 on an ImputedSnpMatrix).
//
// Genotypes enter standardized and mean-imputed: x_ij = (g_ij - mu_j) / sd_j
// with mu_j the mean over the observed entries, sd_j^2 = v_j = mu_j (1 - mu_j/2)
// and x_ij = 0 where g_ij is missing (a SNP with v_j = 0 is all zeros).  With
// clusters C_k of sizes n_k, per SNP and cluster the integer counts
//   gs_kj = sum_{i in C_k observed} g_ij,  ob_kj = #observed i in C_k
// give e_kj = gs_kj - ob_kj mu_j and the cluster mean m_kj = e_kj / (n_k sd_j).
// Feature ranking uses n_k m_kj^2 = (e_kj^2 / n_k) / v_j (summed over k for the
// global ranking); the distance of individual i to the sparse center c_k
// (m_kj on the selected SNPs, 0 elsewhere) drops the term sum_j x_ij^2 common
// to all k:
//   D_ik = sum_{j sel} c_kj (c_kj - 2 x_ij) = sum_{j sel} a_kj (b_kj - 2 g~_ij)
// with a_kj = e_kj / (n_k v_j), b_kj = e_kj / n_k + 2 mu_j and g~ the imputed
// genotype, so each (k, selected j) is a 4-entry table over the PLINK code.
//
// Every floating-point value is FP64 and formed in a fixed operation order
// (this file is compiled without FMA contraction): the cluster labels and the
// AIM lists are index results and match a sequential evaluation in the same
// order bit for bit.  Integer counts are exact.
#include <rocprim/device/device_radix_sort.hpp>

#include "admx_common.hpp"
#include "admx_internal.hpp"

namespace admx {
namespace {

// ------------------------------------------------------------------ cluster masks
// cm[dw * (k-1) + c], c < k-1: 0b11 in the 2-bit field of every individual of
// BED word dw (individuals 16 dw .. 16 dw + 15, < I) that is in cluster c; the
// last cluster's counts follow from the SNP totals.  nk[c] += |C_c| for all c.
__global__ __launch_bounds__(256) void k_skfr_masks(const int32_t* __restrict__ lab, int64_t I, int k, int64_t nwords,
                                                    uint32_t* __restrict__ cm, unsigned* __restrict__ nk) {
  const int64_t dw = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int lb[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int64_t i = 16 * dw + t;
    lb[t] = (dw < nwords && i < I) ? lab[i] : -1;
  }
  for (int c = 0; c < k; ++c) {
    uint32_t m = 0;
#pragma unroll
    for (int t = 0; t < 16; ++t) m |= (uint32_t)(lb[t] == c) << (2 * t);
    if (dw < nwords && c < k - 1) cm[dw * (k - 1) + c] = m | (m << 1);
    unsigned n = __popc(m);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(&nk[c], n);
  }
}

// ------------------------------------------------------------------ centers pass
// One lane per SNP over a slab of 64-byte tiles (256 individuals) of its row,
// the next tile's four 16-byte loads issued before the current tile is used.
// Per 32-bit word (16 individuals) with code bit planes lo/hi: the observed
// genotype is g = hi + (hi & lo) and missing = lo & ~hi, so with
// comb = hi | (hi & lo) << 1 one popcount of comb & M gives the genotype sum
// of the individuals in mask M (0b11 per member) and one of missing & M their
// missing count.  The masks are the same for every lane (scalar loads).
// Clusters 0 .. KA-1 only (KA = k - 1).  Slab partials part[slab][2 KA][J]
// (genotype sums, then missing counts) are summed exactly by the score kernel.
template <int KA>
__global__ __launch_bounds__(256) void k_skfr_centers(const uint8_t* __restrict__ bed, int64_t pitch, int64_t J,
                                                      int64_t ntiles, int64_t tper, const uint32_t* __restrict__ cm,
                                                      uint32_t* __restrict__ part) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t t0 = (int64_t)blockIdx.y * tper;
  const int64_t t1 = (t0 + tper < ntiles) ? t0 + tper : ntiles;
  const int64_t jr = j < J ? j : J - 1;
  const u32x4* __restrict__ col = reinterpret_cast<const u32x4*>(bed + pitch * jr);
  uint32_t gs[KA], ms[KA];
#pragma unroll
  for (int c = 0; c < KA; ++c) gs[c] = ms[c] = 0;
  u32x4 nx[4];
  if (t0 < t1) {
#pragma unroll
    for (int e = 0; e < 4; ++e) nx[e] = col[4 * t0 + e];
  }
  for (int64_t t = t0; t < t1; ++t) {
    u32x4 cur[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) cur[e] = nx[e];
    if (t + 1 < t1) {
#pragma unroll
      for (int e = 0; e < 4; ++e) nx[e] = col[4 * (t + 1) + e];
    }
    const uint32_t* __restrict__ m = cm + 16 * t * KA;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      const uint32_t x = cur[w >> 2][w & 3];
      const uint32_t lo = x & 0x55555555u, hi = (x >> 1) & 0x55555555u;
      const uint32_t hl = hi & lo;
      const uint32_t comb = hi | (hl << 1), mi = lo ^ hl;
#pragma unroll
      for (int c = 0; c < KA; ++c) {
        const uint32_t M = m[w * KA + c];
        gs[c] += __popc(comb & M);
        ms[c] += __popc(mi & M);
      }
    }
  }
  if (j < J) {
    uint32_t* o = part + (int64_t)blockIdx.y * 2 * KA * J + j;
#pragma unroll
    for (int c = 0; c < KA; ++c) {
      o[c * J] = gs[c];
      o[(KA + c) * J] = ms[c];
    }
  }
}

// ------------------------------------------------------------------ feature scores
// mu_j = mean over the observed entries (0 if none), v_j = mu_j (1 - mu_j / 2).
__global__ void k_skfr_stats(const unsigned long long* __restrict__ nm, const unsigned long long* __restrict__ alt,
                             int64_t J, double* __restrict__ mu, double* __restrict__ var) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= J) return;
  const double m = nm[j] ? (double)alt[j] / (double)nm[j] : 0.0;
  mu[j] = m;
  var[j] = m * (1.0 - m / 2.0);
}

// E[c][j] = e_cj and the sort keys: global key[j] = (sum_c e^2/n_c) / v_j;
// local key[c][j] = (e^2/n_c) / v_j (0 for v_j = 0 or an empty cluster).  The
// last cluster's genotype sum and missing count are the SNP's totals minus
// those of the other clusters.
template <int KC>
__global__ __launch_bounds__(256) void k_skfr_score(const uint32_t* __restrict__ part, int64_t nslab,
                                                    const unsigned* __restrict__ nk,
                                                    const unsigned long long* __restrict__ nm,
                                                    const unsigned long long* __restrict__ alt, int64_t I,
                                                    const double* __restrict__ mu_j,
                                                    const double* __restrict__ var_j, int64_t J, int local,
                                                    double* __restrict__ E, double* __restrict__ key,
                                                    int32_t* __restrict__ idx) {
  constexpr int KA = KC - 1;
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= J) return;
  const double mu = mu_j[j], v = var_j[j];
  uint32_t gsum = 0, msum = 0;
  double acc = 0.0;
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    uint32_t gs = 0, ms = 0;
    if (c < KA) {
      for (int64_t sl = 0; sl < nslab; ++sl) {
        gs += part[(sl * 2 * KA + c) * J + j];
        ms += part[(sl * 2 * KA + KA + c) * J + j];
      }
      gsum += gs;
      msum += ms;
    } else {
      gs = (uint32_t)alt[j] - gsum;
      ms = (uint32_t)((uint64_t)I - nm[j]) - msum;
    }
    const unsigned n = nk[c];
    const double ob = (double)(n - ms);
    const double e = (double)gs - ob * mu;
    E[c * J + j] = e;
    double t = 0.0;
    if (n > 0) t = (e * e) / (double)n;
    if (local) key[c * J + j] = (v > 0.0) ? t / v : 0.0;
    acc = acc + t;
  }
  if (!local) key[j] = (v > 0.0) ? acc / v : 0.0;
  idx[j] = (int32_t)j;
}

// flag[c][j] = 1 for the first s entries of cluster c's ranking
__global__ void k_skfr_flag(const int32_t* __restrict__ sorted, int64_t J, int64_t s, int k,
                            uint8_t* __restrict__ flag) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= s) return;
  for (int c = 0; c < k; ++c) flag[c * J + sorted[c * J + t]] = 1;
}

// union key of the local ranking: max over the clusters that selected j, -1 if none
__global__ void k_skfr_union(const uint8_t* __restrict__ flag, const double* __restrict__ key, int64_t J, int k,
                             double* __restrict__ ukey, int32_t* __restrict__ idx,
                             unsigned long long* __restrict__ count) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  bool any = false;
  if (j < J) {
    double u = -1.0;
    for (int c = 0; c < k; ++c)
      if (flag[c * J + j] && key[c * J + j] > u) u = key[c * J + j];
    ukey[j] = u;
    idx[j] = (int32_t)j;
    any = u >= 0.0;
  }
  const unsigned long long b = __ballot(any);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, (unsigned long long)__popcll(b));
}

// Distance tables of the selected SNPs cols[0..ns): T[(t k + c) 4 + code] =
// a (b - 2 g~(code)), codes 00 (g 0), 01 (missing, g~ = mu), 10 (g 1), 11 (g 2).
__global__ void k_skfr_tables(const int32_t* __restrict__ cols, int64_t ns, int k, const uint8_t* __restrict__ flag,
                              const unsigned* __restrict__ nk, const double* __restrict__ mu,
                              const double* __restrict__ var, const double* __restrict__ E, int64_t J,
                              double* __restrict__ T) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= ns) return;
  const int64_t j = cols[t];
  const double m = mu[j], v = var[j];
  for (int c = 0; c < k; ++c) {
    const unsigned n = nk[c];
    const bool sel = flag ? flag[c * J + j] != 0 : true;
    double a = 0.0, b = 0.0;
    if (sel && n > 0 && v > 0.0) {
      const double e = E[c * J + j];
      a = e / ((double)n * v);
      b = e / (double)n + 2.0 * m;
    }
    double* o = T + (t * k + c) * 4;
    o[0] = a * b;
    o[1] = a * (b - 2.0 * m);
    o[2] = a * (b - 2.0);
    o[3] = a * (b - 4.0);
  }
}

// Squared-distance tables of all SNPs against one individual's imputed
// genotype (*center >= 0) or against the SNP means (center == nullptr, the
// per-individual sum_j x_ij^2): T[j 4 + code] = (g~(code) - g~_c)^2 / v_j.
__global__ void k_skfr_pp_table(const uint8_t* __restrict__ bed, int64_t pitch, const int32_t* __restrict__ center,
                                const double* __restrict__ mu, const double* __restrict__ var, int64_t J,
                                double* __restrict__ T) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= J) return;
  const double m = mu[j], v = var[j];
  double gc = m;
  if (center) {
    const int64_t ci = *center;
    const uint32_t cc = (bed[pitch * j + (ci >> 2)] >> (2 * (ci & 3))) & 3u;
    gc = cc == 1u ? m : (double)code_g(cc);
  }
  for (int code = 0; code < 4; ++code) {
    const double g = code == 1 ? m : (double)code_g((uint32_t)code);
    const double d = g - gc;
    T[j * 4 + code] = (v > 0.0) ? (d * d) / v : 0.0;
  }
}

// Table sums over a column list: P[sl][r][i] = sum_{t in slice sl} T[t][r][code(i, cols[t])]
// in increasing t from 0.0; slice sl = blockIdx.y covers t in
// [sl L, min((sl+1) L, ncols)); cols == nullptr means column t.  A lane owns B
// bytes (4 B individuals) of every column.  The block stages the tables of
// CT columns at a time in LDS, each lane issues the loads of its bytes of all
// CT columns at once, then every (individual, row) adds one LDS entry indexed
// by the 2-bit code.
template <int R, int B>
__global__ __launch_bounds__(256) void k_skfr_tabsum(const uint8_t* __restrict__ bed, int64_t pitch, int64_t I,
                                                     int64_t nb, const int32_t* __restrict__ cols, int64_t ncols,
                                                     int64_t L, const double* __restrict__ T,
                                                     double* __restrict__ P) {
  constexpr int CT = 32;
  constexpr int NI = 4 * B;
  __shared__ double tl[CT * R * 4];
  __shared__ int64_t off[CT];
  const int64_t b0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * B;
  const bool live = b0 < nb;
  const int64_t sl = blockIdx.y;
  const int64_t t0 = sl * L;
  const int64_t t1 = (t0 + L < ncols) ? t0 + L : ncols;
  double acc[R][NI];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int e = 0; e < NI; ++e) acc[r][e] = 0.0;
  for (int64_t tc = t0; tc < t1; tc += CT) {
    const int n = (int)((t1 - tc < CT) ? t1 - tc : CT);
    __syncthreads();
    for (int e = threadIdx.x; e < n * R * 4; e += 256) tl[e] = T[tc * R * 4 + e];
    for (int e = threadIdx.x; e < n; e += 256) off[e] = pitch * (int64_t)(cols ? cols[tc + e] : tc + e);
    __syncthreads();
    if (live) {
      uint32_t by[CT];
#pragma unroll
      for (int tt = 0; tt < CT; ++tt) {
        by[tt] = 0;
        if (tt < n) {
          const uint8_t* src = bed + off[tt] + b0;
          if constexpr (B == 1) by[tt] = *src;
          if constexpr (B == 2) by[tt] = *reinterpret_cast<const uint16_t*>(src);
          if constexpr (B == 4) by[tt] = *reinterpret_cast<const uint32_t*>(src);
        }
      }
#pragma unroll
      for (int tt = 0; tt < CT; ++tt) {
        if (tt < n) {
          const double* row = tl + tt * R * 4;
#pragma unroll
          for (int e = 0; e < NI; ++e) {
            const uint32_t code = (by[tt] >> (2 * e)) & 3u;
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r][e] += row[r * 4 + code];
          }
        }
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int e = 0; e < NI; ++e) {
      const int64_t i = 4 * b0 + e;
      if (i < I) P[(sl * R + r) * I + i] = acc[r][e];
    }
}

// Nearest nonempty sparse center: D_ir = sum over slices (in order, from 0.0);
// argmin with ties to the lowest cluster; counts label changes.
__global__ void k_skfr_assign(const double* __restrict__ P, int nsplit, int R, int64_t I,
                              const unsigned* __restrict__ nk, int32_t* __restrict__ lab, double* __restrict__ dmin,
                              unsigned long long* __restrict__ changes) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= I) return;
  int best = -1;
  double bd = 0.0;
  for (int r = 0; r < R; ++r) {
    if (nk[r] == 0) continue;
    double d = 0.0;
    for (int sl = 0; sl < nsplit; ++sl) d = d + P[((int64_t)sl * R + r) * I + i];
    if (best < 0 || d < bd) {
      best = r;
      bd = d;
    }
  }
  if (best < 0) best = lab[i];
  if (lab[i] != best) atomicAdd(changes, 1ull);
  lab[i] = best;
  dmin[i] = bd;
}

// out[i] = sum over slices of P[sl][i] (R = 1), in order from 0.0; with
// D2 != nullptr also the k-means++ bookkeeping: Dc[i] = out, D2 = min(D2, out).
__global__ void k_skfr_combine(const double* __restrict__ P, int nsplit, int64_t I, double* __restrict__ out,
                               double* __restrict__ D2, int first) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= I) return;
  double d = 0.0;
  for (int sl = 0; sl < nsplit; ++sl) d = d + P[(int64_t)sl * I + i];
  out[i] = d;
  if (D2) D2[i] = (first || d < D2[i]) ? d : D2[i];
}

// k-means++ draw of the next center with probability D2_i / sum D2 (one block
// of 256 threads): thread t sums a contiguous chunk of ceil(I/256) entries in
// order, thread 0 accumulates the chunk sums in order, finds the chunk whose
// running total first exceeds u * total and then the entry inside it.  The
// first center, and any draw with total = 0, is floor(u I).
__global__ __launch_bounds__(256) void k_skfr_pick(const double* __restrict__ D2, int64_t I, double u, int first,
                                                   int32_t* __restrict__ center) {
  __shared__ double ps[256];
  const int t = threadIdx.x;
  const int64_t chunk = (I + 255) / 256;
  double s = 0.0;
  if (!first) {
    const int64_t a = t * chunk;
    const int64_t e = (a + chunk < I) ? a + chunk : I;
    for (int64_t i = a; i < e; ++i) s = s + D2[i];
  }
  ps[t] = s;
  __syncthreads();
  if (t != 0) return;
  int64_t pick = -1;
  if (!first) {
    double total = 0.0;
    for (int tt = 0; tt < 256; ++tt) total = total + ps[tt];
    if (total > 0.0) {
      const double target = u * total;
      double cum = 0.0;
      int tsel = -1;
      for (int tt = 0; tt < 256; ++tt) {
        if (cum + ps[tt] > target) {
          tsel = tt;
          break;
        }
        cum = cum + ps[tt];
      }
      if (tsel >= 0) {
        const int64_t a = tsel * chunk;
        const int64_t e = (a + chunk < I) ? a + chunk : I;
        for (int64_t i = a; i < e; ++i) {
          cum = cum + D2[i];
          if (cum > target) {
            pick = i;
            break;
          }
        }
        if (pick < 0)
          for (int64_t i = e - 1; i >= a; --i)
            if (D2[i] > 0.0) {
              pick = i;
              break;
            }
      }
      if (pick < 0)
        for (int64_t i = I - 1; i >= 0; --i)
          if (D2[i] > 0.0) {
            pick = i;
            break;
          }
    }
  }
  if (pick < 0) {
    pick = (int64_t)(u * (double)I);
    if (pick >= I) pick = I - 1;
  }
  *center = (int32_t)pick;
}

// labels of the k-means++ start: nearest chosen center, ties to the lowest
__global__ void k_skfr_argmin(const double* __restrict__ Dc, int k, int64_t I, int32_t* __restrict__ lab) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= I) return;
  int best = 0;
  double bd = Dc[i];
  for (int c = 1; c < k; ++c) {
    const double d = Dc[(int64_t)c * I + i];
    if (d < bd) {
      best = c;
      bd = d;
    }
  }
  lab[i] = best;
}

// Within-cluster sums of squares: WSS_c = sum_{i in C_c} (sum_j x_ij^2 + D_i),
// thread t over a contiguous chunk in order, then thread 0 over the threads;
// out[k] = total objective.
__global__ __launch_bounds__(256) void k_skfr_wss(const int32_t* __restrict__ lab, const double* __restrict__ x2,
                                                  const double* __restrict__ dmin, int64_t I, int k,
                                                  double* __restrict__ out) {
  __shared__ double ws[256 * kMaxK];
  const int t = threadIdx.x;
  for (int c = 0; c < kMaxK; ++c) ws[t * kMaxK + c] = 0.0;
  const int64_t chunk = (I + 255) / 256;
  const int64_t a = t * chunk;
  const int64_t e = (a + chunk < I) ? a + chunk : I;
  for (int64_t i = a; i < e; ++i) ws[t * kMaxK + lab[i]] += x2[i] + dmin[i];
  __syncthreads();
  if (t != 0) return;
  double tot = 0.0;
  for (int c = 0; c < k; ++c) {
    double w = 0.0;
    for (int tt = 0; tt < 256; ++tt) w = w + ws[tt * kMaxK + c];
    out[c] = w;
    tot = tot + w;
  }
  out[k] = tot;
}

// centers[c + k j] = e_cj / (n_c sd_j) on the selected SNPs, 0 elsewhere
__global__ void k_skfr_center_out(const uint8_t* __restrict__ sel, const uint8_t* __restrict__ flag,
                                  const unsigned* __restrict__ nk, const double* __restrict__ var,
                                  const double* __restrict__ E, int64_t J, int k, double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= J) return;
  const double v = var[j];
  for (int c = 0; c < k; ++c) {
    const bool on = flag ? flag[c * J + j] != 0 : sel[j] != 0;
    const unsigned n = nk[c];
    out[c + (int64_t)k * j] = (on && n > 0 && v > 0.0) ? E[c * J + j] / ((double)n * sqrt(v)) : 0.0;
  }
}

__global__ void k_skfr_mark(const int32_t* __restrict__ cols, int64_t ns, uint8_t* __restrict__ sel) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < ns) sel[cols[t]] = 1;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void launch_skfr_masks(const int32_t* lab, int64_t I, int k, int64_t nwords, uint32_t* cm, unsigned* nk,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_masks, dim3(nblk(nwords)), dim3(256), 0, s, lab, I, k, nwords, cm, nk);
}

int64_t skfr_slabs(int64_t J, int64_t ntiles, int64_t* tper) {
  const int64_t blocks = (J + 255) / 256;
  int64_t ns = (4096 + blocks - 1) / blocks;  // about 16384 waves
  if (ns > ntiles) ns = ntiles;
  if (ns < 1) ns = 1;
  *tper = (ntiles + ns - 1) / ns;
  return (ntiles + *tper - 1) / *tper;
}

void launch_skfr_centers(int k, const uint8_t* bed, int64_t pitch, int64_t J, int64_t ntiles, int64_t tper,
                         int64_t nslab, const uint32_t* cm, uint32_t* part, hipStream_t s) {
  if (k < 2) return;  // one cluster: the SNP totals are its counts
  const dim3 grid(nblk(J), (unsigned)nslab);
  ADMX_DISPATCH_K(k - 1, (k_skfr_centers<KK><<<grid, dim3(256), 0, s>>>(bed, pitch, J, ntiles, tper, cm, part)));
}

void launch_skfr_stats(const unsigned long long* nm, const unsigned long long* alt, int64_t J, double* mu,
                       double* var, hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_stats, dim3(nblk(J)), dim3(256), 0, s, nm, alt, J, mu, var);
}

void launch_skfr_score(int k, int local, const uint32_t* part, int64_t nslab, const unsigned* nk,
                       const unsigned long long* nm, const unsigned long long* alt, int64_t I, const double* mu,
                       const double* var, int64_t J, double* E, double* key, int32_t* idx, hipStream_t s) {
  ADMX_DISPATCH_K(k, (k_skfr_score<KK><<<dim3(nblk(J)), dim3(256), 0, s>>>(part, nslab, nk, nm, alt, I, mu, var, J,
                                                                            local, E, key, idx)));
}

size_t skfr_sort_bytes(int64_t n) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs_desc(nullptr, bytes, (const double*)nullptr, (double*)nullptr,
                                       (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)n, 0, 64, nullptr);
  return bytes;
}

// Stable descending sort of (key, index) pairs: equal keys keep increasing index.
void launch_skfr_sort(void* tmp, size_t bytes, const double* kin, double* kout, const int32_t* vin, int32_t* vout,
                      int64_t n, hipStream_t s) {
  (void)rocprim::radix_sort_pairs_desc(tmp, bytes, kin, kout, vin, vout, (size_t)n, 0, 64, s);
}

void launch_skfr_flag(const int32_t* sorted, int64_t J, int64_t sp, int k, uint8_t* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_flag, dim3(nblk(sp)), dim3(256), 0, s, sorted, J, sp, k, flag);
}

void launch_skfr_union(const uint8_t* flag, const double* key, int64_t J, int k, double* ukey, int32_t* idx,
                       unsigned long long* count, hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_union, dim3(nblk(J)), dim3(256), 0, s, flag, key, J, k, ukey, idx, count);
}

void launch_skfr_tables(const int32_t* cols, int64_t ns, int k, const uint8_t* flag, const unsigned* nk,
                        const double* mu, const double* var, const double* E, int64_t J, double* T, hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_tables, dim3(nblk(ns)), dim3(256), 0, s, cols, ns, k, flag, nk, mu, var, E, J, T);
}

void launch_skfr_pp_table(const uint8_t* bed, int64_t pitch, const int32_t* center, const double* mu,
                          const double* var, int64_t J, double* T, hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_pp_table, dim3(nblk(J)), dim3(256), 0, s, bed, pitch, center, mu, var, J, T);
}

// Bytes per lane of the table-sum kernel: 16 individuals for one row, 8 for
// up to four rows, 4 beyond (at most 64 FP64 accumulators per lane).
static int tabsum_bytes(int R) { return R == 1 ? 4 : (R <= 4 ? 2 : 1); }

int skfr_nsplit(int64_t I, int64_t ncols, int R) {
  const int64_t nb = (I + 3) / 4;
  const int64_t lanes = (nb + tabsum_bytes(R) - 1) / tabsum_bytes(R);
  const int64_t blocks = (lanes + 255) / 256;
  int64_t n = (1024 + blocks - 1) / blocks;  // about 4096 waves
  if (n > 64) n = 64;
  if (n > ncols) n = ncols;
  if (n < 1) n = 1;
  return (int)n;
}

template <int R>
static void tabsum_r(const uint8_t* bed, int64_t pitch, int64_t I, const int32_t* cols, int64_t ncols, int nsplit,
                     const double* T, double* P, hipStream_t s) {
  const int64_t nb = (I + 3) / 4;
  const int64_t L = (ncols + nsplit - 1) / nsplit;
  constexpr int B = R == 1 ? 4 : (R <= 4 ? 2 : 1);
  const int64_t lanes = (nb + B - 1) / B;
  const dim3 grid(nblk(lanes), (unsigned)nsplit);
  hipLaunchKernelGGL((k_skfr_tabsum<R, B>), grid, dim3(256), 0, s, bed, pitch, I, nb, cols, ncols, L, T, P);
}

void launch_skfr_tabsum(int R, const uint8_t* bed, int64_t pitch, int64_t I, const int32_t* cols, int64_t ncols,
                        int nsplit, const double* T, double* P, hipStream_t s) {
  ADMX_DISPATCH_K(R, (tabsum_r<KK>(bed, pitch, I, cols, ncols, nsplit, T, P, s)));
}

void launch_skfr_assign(const double* P, int nsplit, int R, int64_t I, const unsigned* nk, int32_t* lab,
                        double* dmin, unsigned long long* changes, hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_assign, dim3(nblk(I)), dim3(256), 0, s, P, nsplit, R, I, nk, lab, dmin, changes);
}

void launch_skfr_combine(const double* P, int nsplit, int64_t I, double* out, double* D2, int first,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_combine, dim3(nblk(I)), dim3(256), 0, s, P, nsplit, I, out, D2, first);
}

void launch_skfr_pick(const double* D2, int64_t I, double u, int first, int32_t* center, hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_pick, dim3(1), dim3(256), 0, s, D2, I, u, first, center);
}

void launch_skfr_argmin(const double* Dc, int k, int64_t I, int32_t* lab, hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_argmin, dim3(nblk(I)), dim3(256), 0, s, Dc, k, I, lab);
}

void launch_skfr_wss(const int32_t* lab, const double* x2, const double* dmin, int64_t I, int k, double* out,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_wss, dim3(1), dim3(256), 0, s, lab, x2, dmin, I, k, out);
}

void launch_skfr_mark(const int32_t* cols, int64_t ns, uint8_t* sel, hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_mark, dim3(nblk(ns)), dim3(256), 0, s, cols, ns, sel);
}

void launch_skfr_center_out(const uint8_t* sel, const uint8_t* flag, const unsigned* nk, const double* var,
                            const double* E, int64_t J, int k, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_skfr_center_out, dim3(nblk(J)), dim3(256), 0, s, sel, flag, nk, var, E, J, k, out);
}

}  // namespace admx
