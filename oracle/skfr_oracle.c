/*
 * skfr_oracle.c -- CPU oracle of the SKFR AIM pre-selection (sparse k-means
 * with feature ranking).  TEST INFRASTRUCTURE ONLY: the checker that the HIP
 * library's admx_skfr_init / admx_skfr_run are compared against; no product
 * code links it or calls it.
 *
 * The reference calls SparseKmeansFeatureRanking.jl 0.1 from
 * src/driver.jl:56-66 and 80-101 (ImputedSnpMatrix{Float64}(g, K; rng),
 * sparsekmeans1 / sparsekmeans2 (ISM, sparsity; max_iter, squares=false),
 * sparsekmeans_repeat, sparsekmeans_path).  That package is a dependency
 * (Project.toml compat "0.1") that is not part of the reference tree, so this
 * is a restatement of the published method (Zhang, Lange and Xu, "A simple
 * approach to sparse clustering", CSDA 2020): standardized, mean-imputed
 * genotypes; cluster means; features ranked by the between-cluster sum of
 * squares n_k mean_kj^2 (summed over k for the global ranking, per cluster for
 * the local one); top-s features kept, other center coordinates set to the
 * global mean 0; nearest-center reassignment until no label changes.  The
 * per-SNP standardization sd_j = sqrt(mu_j (1 - mu_j / 2)) is SnpArrays'
 * convention.  Its pins are in tests/test_skfr_oracle.py: Lloyd's k-means on
 * a dense standardized matrix (sparsity = J), brute-force between-cluster
 * sums of squares for the ranking, planted ancestry-informative SNPs, allele
 * flip invariance and dense k-means++ draws.
 *
 * Operation order follows the device kernels exactly (sequential sums in the
 * same grouping, no FMA contraction: -ffp-contract=off), so labels and AIM
 * lists agree bit for bit for the same nsplit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SK_MAXK 16

static inline int sk_code(const uint8_t *bed, int64_t pitch, int64_t i, int64_t j) {
  return (bed[(i >> 2) + pitch * j] >> ((i & 3) << 1)) & 3;
}
/* g of the PLINK codes 00, 01 (missing), 10, 11 */
static const double SK_G[4] = {0.0, 0.0, 1.0, 2.0};

static inline uint64_t sk_smix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* draw u in [0,1) of k-means++ center cc */
double orc_skfr_u(uint64_t seed, int cc) {
  return (double)(sk_smix(seed ^ (0xD1B54A32D192ED03ULL * (uint64_t)(cc + 1))) >> 11) * 0x1.0p-53;
}

/* slices of the distance sums: requested nsplit clamped to [1, ncols], or the
 * automatic count for nsplit = 0 (about 4096 waves; a lane of the device
 * kernel covers 16 individuals for R = 1 row, 8 for R <= 4, else 4) */
int orc_skfr_nsplit(int64_t I, int64_t ncols, int nsplit, int R) {
  int64_t n;
  if (nsplit > 0) {
    n = nsplit;
  } else {
    const int64_t per = R == 1 ? 4 : (R <= 4 ? 2 : 1);
    const int64_t lanes = ((I + 3) / 4 + per - 1) / per;
    const int64_t blocks = (lanes + 255) / 256;
    n = (1024 + blocks - 1) / blocks;
    if (n > 64) n = 64;
  }
  if (n > ncols) n = ncols;
  if (n < 1) n = 1;
  return (int)n;
}

/* mu_j over observed entries (0 if none) and v_j = mu_j (1 - mu_j / 2) */
void orc_skfr_stats(const uint8_t *bed, int64_t pitch, int64_t I, int64_t J, double *mu, double *var) {
  for (int64_t j = 0; j < J; ++j) {
    uint64_t nm = 0, alt = 0;
    for (int64_t i = 0; i < I; ++i) {
      const int c = sk_code(bed, pitch, i, j);
      if (c != 1) {
        nm += 1;
        alt += (uint64_t)SK_G[c];
      }
    }
    const double m = nm ? (double)alt / (double)nm : 0.0;
    mu[j] = m;
    var[j] = m * (1.0 - m / 2.0);
  }
}

/* D[r][i] = sum over slices (in order, from 0.0) of the slice sums (in order,
 * from 0.0) of T[(t R + r) 4 + code(i, col t)] */
static void sk_tabsum(const uint8_t *bed, int64_t pitch, int64_t I, const int32_t *cols, int64_t ncols, int nsplit,
                      int R, const double *T, double *D) {
  const int64_t L = (ncols + nsplit - 1) / nsplit;
  for (int64_t x = 0; x < (int64_t)R * I; ++x) D[x] = 0.0;
  for (int sl = 0; sl < nsplit; ++sl) {
    const int64_t t0 = (int64_t)sl * L;
    const int64_t t1 = (t0 + L < ncols) ? t0 + L : ncols;
    for (int r = 0; r < R; ++r)
      for (int64_t i = 0; i < I; ++i) {
        double acc = 0.0;
        for (int64_t t = t0; t < t1; ++t) {
          const int64_t j = cols ? cols[t] : t;
          acc = acc + T[(t * R + r) * 4 + sk_code(bed, pitch, i, j)];
        }
        D[(int64_t)r * I + i] = D[(int64_t)r * I + i] + acc;
      }
  }
}

/* squared-distance table against individual ci's imputed genotypes (ci >= 0)
 * or against the means (ci < 0) */
static void sk_pp_table(const uint8_t *bed, int64_t pitch, int64_t ci, const double *mu, const double *var,
                        int64_t J, double *T) {
  for (int64_t j = 0; j < J; ++j) {
    const double m = mu[j], v = var[j];
    double gc = m;
    if (ci >= 0) {
      const int cc = sk_code(bed, pitch, ci, j);
      gc = cc == 1 ? m : SK_G[cc];
    }
    for (int code = 0; code < 4; ++code) {
      const double g = code == 1 ? m : SK_G[code];
      const double d = g - gc;
      T[j * 4 + code] = (v > 0.0) ? (d * d) / v : 0.0;
    }
  }
}

static int64_t sk_pick(const double *D2, int64_t I, double u, int first) {
  const int64_t chunk = (I + 255) / 256;
  double ps[256];
  for (int t = 0; t < 256; ++t) {
    double s = 0.0;
    if (!first) {
      const int64_t a = t * chunk;
      const int64_t e = (a + chunk < I) ? a + chunk : I;
      for (int64_t i = a; i < e; ++i) s = s + D2[i];
    }
    ps[t] = s;
  }
  int64_t pick = -1;
  if (!first) {
    double total = 0.0;
    for (int t = 0; t < 256; ++t) total = total + ps[t];
    if (total > 0.0) {
      const double target = u * total;
      double cum = 0.0;
      int tsel = -1;
      for (int t = 0; t < 256; ++t) {
        if (cum + ps[t] > target) {
          tsel = t;
          break;
        }
        cum = cum + ps[t];
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
  return pick;
}

/* k-means++ start; centers_out (k, nullable) receives the drawn individuals */
int orc_skfr_init(const uint8_t *bed, int64_t pitch, int64_t I, int64_t J, int k, uint64_t seed, int nsplit,
                  int32_t *clusters, int64_t *centers_out) {
  if (k < 1 || k > SK_MAXK || k > I) return 1;
  double *mu = malloc(sizeof(double) * J), *var = malloc(sizeof(double) * J);
  double *T = malloc(sizeof(double) * 4 * J), *Dc = malloc(sizeof(double) * k * I), *D2 = malloc(sizeof(double) * I);
  orc_skfr_stats(bed, pitch, I, J, mu, var);
  const int ns = orc_skfr_nsplit(I, J, nsplit, 1);
  for (int cc = 0; cc < k; ++cc) {
    const int64_t ci = sk_pick(D2, I, orc_skfr_u(seed, cc), cc == 0);
    if (centers_out) centers_out[cc] = ci;
    sk_pp_table(bed, pitch, ci, mu, var, J, T);
    double *d = Dc + (int64_t)cc * I;
    sk_tabsum(bed, pitch, I, NULL, J, ns, 1, T, d);
    for (int64_t i = 0; i < I; ++i) D2[i] = (cc == 0 || d[i] < D2[i]) ? d[i] : D2[i];
  }
  for (int64_t i = 0; i < I; ++i) {
    int best = 0;
    double bd = Dc[i];
    for (int cc = 1; cc < k; ++cc)
      if (Dc[(int64_t)cc * I + i] < bd) {
        best = cc;
        bd = Dc[(int64_t)cc * I + i];
      }
    clusters[i] = best;
  }
  free(mu);
  free(var);
  free(T);
  free(Dc);
  free(D2);
  return 0;
}

/* (key desc, index asc) */
static const double *sk_sort_key;
static int sk_cmp(const void *a, const void *b) {
  const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  const double kx = sk_sort_key[x], ky = sk_sort_key[y];
  if (kx > ky) return -1;
  if (kx < ky) return 1;
  return (x > y) - (x < y);
}
static void sk_rank(const double *key, int64_t J, int32_t *order) {
  for (int64_t j = 0; j < J; ++j) order[j] = (int32_t)j;
  sk_sort_key = key;
  qsort(order, (size_t)J, sizeof(int32_t), sk_cmp);
}

int orc_skfr_run(const uint8_t *bed, int64_t pitch, int64_t I, int64_t J, int k, int mode, int64_t s, int max_iter,
                 int nsplit, int32_t *clusters, int64_t *aims, int64_t *n_aims, double *centers, double *wss,
                 double *obj, int32_t *iters) {
  if (k < 1 || k > SK_MAXK || s < 1 || s > J || max_iter < 1 || (mode != 0 && mode != 1)) return 1;
  for (int64_t i = 0; i < I; ++i)
    if (clusters[i] < 0 || clusters[i] >= k) return 1;
  const int local = mode == 1;
  double *mu = malloc(sizeof(double) * J), *var = malloc(sizeof(double) * J);
  double *E = malloc(sizeof(double) * k * J), *key = malloc(sizeof(double) * k * J), *ukey = malloc(sizeof(double) * J);
  int32_t *order = malloc(sizeof(int32_t) * J), *cols = malloc(sizeof(int32_t) * J);
  uint8_t *flag = calloc((size_t)k * J, 1);
  const int64_t ncols_max = local ? ((int64_t)k * s < J ? (int64_t)k * s : J) : s;
  const int64_t trows = 4 * J > 4 * k * ncols_max ? 4 * J : 4 * k * ncols_max;
  double *T = malloc(sizeof(double) * trows), *D = malloc(sizeof(double) * k * I);
  double *x2 = malloc(sizeof(double) * I), *dmin = malloc(sizeof(double) * I);
  uint32_t *gs = malloc(sizeof(uint32_t) * k * J), *ms = malloc(sizeof(uint32_t) * k * J);
  unsigned nk[SK_MAXK];
  orc_skfr_stats(bed, pitch, I, J, mu, var);
  sk_pp_table(bed, pitch, -1, mu, var, J, T);
  sk_tabsum(bed, pitch, I, NULL, J, orc_skfr_nsplit(I, J, nsplit, 1), 1, T, x2);
  int it;
  int64_t ncols = 0;
  for (it = 1; it <= max_iter; ++it) {
    memset(nk, 0, sizeof(nk));
    for (int64_t i = 0; i < I; ++i) nk[clusters[i]] += 1;
    memset(gs, 0, sizeof(uint32_t) * k * J);
    memset(ms, 0, sizeof(uint32_t) * k * J);
    for (int64_t j = 0; j < J; ++j)
      for (int64_t i = 0; i < I; ++i) {
        const int c = sk_code(bed, pitch, i, j), cl = clusters[i];
        if (c == 1)
          ms[cl * J + j] += 1;
        else
          gs[cl * J + j] += (uint32_t)SK_G[c];
      }
    for (int64_t j = 0; j < J; ++j) {
      const double m = mu[j], v = var[j];
      double acc = 0.0;
      for (int c = 0; c < k; ++c) {
        const unsigned n = nk[c];
        const double ob = (double)(n - ms[c * J + j]);
        const double e = (double)gs[c * J + j] - ob * m;
        E[c * J + j] = e;
        double t = 0.0;
        if (n > 0) t = (e * e) / (double)n;
        if (local) key[c * J + j] = (v > 0.0) ? t / v : 0.0;
        acc = acc + t;
      }
      if (!local) key[j] = (v > 0.0) ? acc / v : 0.0;
    }
    if (!local) {
      sk_rank(key, J, order);
      ncols = s;
      memcpy(cols, order, sizeof(int32_t) * s);
    } else {
      memset(flag, 0, (size_t)k * J);
      for (int c = 0; c < k; ++c) {
        sk_rank(key + (int64_t)c * J, J, order);
        for (int64_t t = 0; t < s; ++t) flag[c * J + order[t]] = 1;
      }
      ncols = 0;
      for (int64_t j = 0; j < J; ++j) {
        double u = -1.0;
        for (int c = 0; c < k; ++c)
          if (flag[c * J + j] && key[c * J + j] > u) u = key[c * J + j];
        ukey[j] = u;
        if (u >= 0.0) ncols += 1;
      }
      sk_rank(ukey, J, order);
      memcpy(cols, order, sizeof(int32_t) * ncols);
    }
    for (int64_t t = 0; t < ncols; ++t) {
      const int64_t j = cols[t];
      const double m = mu[j], v = var[j];
      for (int c = 0; c < k; ++c) {
        const unsigned n = nk[c];
        const int sel = local ? flag[c * J + j] != 0 : 1;
        double a = 0.0, b = 0.0;
        if (sel && n > 0 && v > 0.0) {
          const double e = E[c * J + j];
          a = e / ((double)n * v);
          b = e / (double)n + 2.0 * m;
        }
        double *o = T + (t * k + c) * 4;
        o[0] = a * b;
        o[1] = a * (b - 2.0 * m);
        o[2] = a * (b - 2.0);
        o[3] = a * (b - 4.0);
      }
    }
    sk_tabsum(bed, pitch, I, cols, ncols, orc_skfr_nsplit(I, ncols, nsplit, k), k, T, D);
    int64_t changes = 0;
    for (int64_t i = 0; i < I; ++i) {
      int best = -1;
      double bd = 0.0;
      for (int r = 0; r < k; ++r) {
        if (nk[r] == 0) continue;
        const double d = D[(int64_t)r * I + i];
        if (best < 0 || d < bd) {
          best = r;
          bd = d;
        }
      }
      if (best < 0) best = clusters[i];
      if (clusters[i] != best) changes += 1;
      clusters[i] = best;
      dmin[i] = bd;
    }
    if (changes == 0) break;
  }
  if (it > max_iter) it = max_iter;
  /* WSS in the device's grouping: 256 contiguous chunks, then per cluster */
  {
    const int64_t chunk = (I + 255) / 256;
    double *ws = calloc((size_t)256 * SK_MAXK, sizeof(double));
    for (int t = 0; t < 256; ++t) {
      const int64_t a = t * chunk;
      const int64_t e = (a + chunk < I) ? a + chunk : I;
      for (int64_t i = a; i < e; ++i) ws[t * SK_MAXK + clusters[i]] = ws[t * SK_MAXK + clusters[i]] + (x2[i] + dmin[i]);
    }
    double tot = 0.0;
    for (int c = 0; c < k; ++c) {
      double w = 0.0;
      for (int t = 0; t < 256; ++t) w = w + ws[t * SK_MAXK + c];
      if (wss) wss[c] = w;
      tot = tot + w;
    }
    if (obj) *obj = tot;
    free(ws);
  }
  if (centers) {
    uint8_t *sel = calloc((size_t)J, 1);
    if (!local)
      for (int64_t t = 0; t < ncols; ++t) sel[cols[t]] = 1;
    for (int64_t j = 0; j < J; ++j)
      for (int c = 0; c < k; ++c) {
        const int on = local ? flag[c * J + j] != 0 : sel[j] != 0;
        const unsigned n = nk[c];
        centers[c + (int64_t)k * j] = (on && n > 0 && var[j] > 0.0) ? E[c * J + j] / ((double)n * sqrt(var[j])) : 0.0;
      }
    free(sel);
  }
  for (int64_t t = 0; t < ncols; ++t) aims[t] = cols[t];
  *n_aims = ncols;
  if (iters) *iters = it;
  free(mu);
  free(var);
  free(E);
  free(key);
  free(ukey);
  free(order);
  free(cols);
  free(flag);
  free(T);
  free(D);
  free(x2);
  free(dmin);
  free(gs);
  free(ms);
  return 0;
}
