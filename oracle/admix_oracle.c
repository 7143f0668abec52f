/*
 * admix_oracle.c -- CPU oracle for the OpenADMIXTURE.jl likelihood / EM / SQP
 * inner loops.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP
 * library is compared against (tests, smoke check and the bench's CPU
 * baseline leg).  No product code links it or calls it.
 *
 * The reference is pure Julia (no Julia toolchain exists where this is built),
 * so this is a restatement in C of the reference's own CPU loops, following
 * each function's algebraic form and operation order (file:line citations at
 * each function; relative to the OpenADMIXTURE.jl source tree).  Parity of this
 * restatement is pinned by the tests under tests/ against the PLINK-1 format,
 * closed forms, finite differences, brute-force QP enumeration and
 * textbook identities (sweep inverse, simplex projection).
 *
 * Built by oracle/Makefile into oracle/liboracle.so (gcc -O2 -fopenmp,
 * -ffp-contract=off so that a*b+c is rounded twice as Julia does outside
 * @turbo).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* PLINK-1 SNP-major 2-bit codes, low bits first: 00 hom A1 (g=0), 01 missing,
 * 10 het (g=1), 11 hom A2 (g=2); tables of src/loops.jl:2-5. */
static const double G_MAP[4] = {0.0, 0.0, 1.0, 2.0};
static const double NM_MAP[4] = {1.0, 0.0, 1.0, 1.0};

static inline int orc_code(const uint8_t *bed, int64_t pitch, int64_t i, int64_t j) {
  return (bed[(i >> 2) + pitch * j] >> ((i & 3) << 1)) & 3;
}

/* cleave(range, step)/findcleft (src/loops.jl:16-28) on a 0-based half-open
 * range; returns the start of the second half. */
static int64_t orc_cleave(int64_t a, int64_t b, int64_t step) {
  const int64_t len = b - a;
  if (len >= 2 * step) return a + step * (len / (step * 2));
  int64_t h = len / 2; /* round(Int, len/2), ties to even */
  if ((len & 1) && (h & 1)) h += 1;
  return a + h;
}

/* Gauss-Jordan inverse with partial pivoting of an m x m column-major matrix. */
static void orc_invert(const double *A, double *Ainv, int64_t m) {
  double W[16 * 32];
  const int64_t w = 2 * m;
  for (int64_t r = 0; r < m; ++r)
    for (int64_t c = 0; c < m; ++c) {
      W[r * w + c] = A[r + m * c];
      W[r * w + m + c] = (r == c) ? 1.0 : 0.0;
    }
  for (int64_t c = 0; c < m; ++c) {
    int64_t piv = c;
    for (int64_t r = c + 1; r < m; ++r)
      if (fabs(W[r * w + c]) > fabs(W[piv * w + c])) piv = r;
    if (piv != c)
      for (int64_t t = 0; t < w; ++t) {
        const double tt = W[c * w + t];
        W[c * w + t] = W[piv * w + t];
        W[piv * w + t] = tt;
      }
    const double inv = 1.0 / W[c * w + c];
    for (int64_t t = 0; t < w; ++t) W[c * w + t] *= inv;
    for (int64_t r = 0; r < m; ++r)
      if (r != c) {
        const double f = W[r * w + c];
        if (f != 0.0)
          for (int64_t t = 0; t < w; ++t) W[r * w + t] -= f * W[c * w + t];
      }
  }
  for (int64_t r = 0; r < m; ++r)
    for (int64_t c = 0; c < m; ++c) Ainv[r + m * c] = W[r * w + m + c];
}

/* ---- BED decode and counts (format: src/loops.jl:2-5, 529-534; the
 * reference has no count function, these are brute-force definitions). */
void orc_decode(const uint8_t *bed, int64_t pitch, int64_t I, int64_t J, int8_t *g) {
  for (int64_t j = 0; j < J; ++j)
    for (int64_t i = 0; i < I; ++i) {
      const int c = orc_code(bed, pitch, i, j);
      g[i + I * j] = (int8_t)(c == 1 ? -1 : (int)G_MAP[c]);
    }
}

void orc_counts(const uint8_t *bed, int64_t pitch, int64_t I, int64_t J,
                int64_t *snp_nonmissing, int64_t *snp_alt, int64_t *ind_nonmissing) {
  for (int64_t i = 0; i < I; ++i) ind_nonmissing[i] = 0;
  for (int64_t j = 0; j < J; ++j) {
    int64_t nm = 0, alt = 0;
    for (int64_t i = 0; i < I; ++i) {
      const int c = orc_code(bed, pitch, i, j);
      if (c != 1) {
        ++nm;
        alt += (int64_t)G_MAP[c];
        ++ind_nonmissing[i];
      }
    }
    snp_nonmissing[j] = nm;
    snp_alt[j] = alt;
  }
}

/* ---- Synthetic cohort genotypes (not in the reference; the generator used
 * for every workload).  g_ij ~ Binomial(2, s_ij) with s_ij = sum_k Q[k,i] P[k,j]
 * evaluated as an fma chain in double; missing with probability miss_rate.
 * Uniforms come from a counter-based splitmix64 hash of (seed, genotype index),
 * so the device generator reproduces these bytes exactly. */
static inline uint64_t orc_smix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline double orc_u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

void orc_gen_bed(const double *Qt, const double *Pt, int64_t I, int64_t J, int64_t K,
                 uint64_t seed, double miss_rate, uint8_t *bed, int64_t pitch) {
  const uint64_t base = orc_smix(seed ^ 0x5851F42D4C957F2DULL);
  for (int64_t j = 0; j < J; ++j) {
    memset(bed + pitch * j, 0, (size_t)pitch);
    for (int64_t i = 0; i < I; ++i) {
      double s = Qt[K * i] * Pt[K * j];
      for (int64_t k = 1; k < K; ++k) s = fma(Qt[k + K * i], Pt[k + K * j], s);
      const uint64_t ctr = (uint64_t)j * (uint64_t)I + (uint64_t)i;
      const double u1 = orc_u01(orc_smix(orc_smix(3 * ctr + 0) ^ base));
      const double u2 = orc_u01(orc_smix(orc_smix(3 * ctr + 1) ^ base));
      const double u3 = orc_u01(orc_smix(orc_smix(3 * ctr + 2) ^ base));
      const int g = (u1 < s) + (u2 < s);
      int code = g == 0 ? 0 : (g == 1 ? 2 : 3);
      if (u3 < miss_rate) code = 1;
      bed[(i >> 2) + pitch * j] |= (uint8_t)(code << ((i & 3) << 1));
    }
  }
}

/* ---- Cross-validation holdout (not in the reference, which has no CV; the
 * project's definition, parity unpinned against the reference).  Entry (i, j)
 * belongs to fold floor(nfolds * hi32(smix(j*I + i ^ base)) / 2^32) with
 * base = smix(seed ^ 0x2545F4914F6CDD1D). */
int orc_fold_of(int64_t i, int64_t j, int64_t I, uint64_t seed, int nfolds) {
  const uint64_t base = orc_smix(seed ^ 0x2545F4914F6CDD1DULL);
  const uint64_t ctr = (uint64_t)j * (uint64_t)I + (uint64_t)i;
  return (int)(((orc_smix(ctr ^ base) >> 32) * (uint64_t)nfolds) >> 32);
}

/* dst = src with the entries of `fold` set to missing (01). */
void orc_mask_fold(const uint8_t *src, uint8_t *dst, int64_t pitch, int64_t I, int64_t J, int fold, int nfolds,
                   uint64_t seed) {
  memcpy(dst, src, (size_t)(pitch * J));
  for (int64_t j = 0; j < J; ++j)
    for (int64_t i = 0; i < I; ++i)
      if (orc_fold_of(i, j, I, seed, nfolds) == fold) {
        uint8_t *b = dst + (i >> 2) + pitch * j;
        const int sh = (int)((i & 3) << 1);
        *b = (uint8_t)((*b & ~(3 << sh)) | (1 << sh));
      }
}

/* Held-out binomial deviance of (q, p) over the observed entries of `fold`:
 * sum 2[g log(g/2s) + (2-g) log((2-g)/2(1-s))] with 0 log 0 = 0, and the count. */
double orc_heldout_deviance(const uint8_t *bed, int64_t pitch, const double *q, const double *p, int64_t I,
                            int64_t J, int64_t K, int fold, int nfolds, uint64_t seed, int64_t *count) {
  double dev = 0.0;
  int64_t n = 0;
  for (int64_t j = 0; j < J; ++j)
    for (int64_t i = 0; i < I; ++i) {
      const int c = orc_code(bed, pitch, i, j);
      if (c == 1 || orc_fold_of(i, j, I, seed, nfolds) != fold) continue;
      double s = 0.0;
      for (int64_t k = 0; k < K; ++k) s += q[k + K * i] * p[k + K * j];
      const double g = G_MAP[c];
      double d = 0.0;
      if (g > 0.0) d += g * log(g / (2.0 * s));
      if (g < 2.0) d += (2.0 - g) * log((2.0 - g) / (2.0 * (1.0 - s)));
      dev += 2.0 * d;
      ++n;
    }
  *count = n;
  return dev;
}

/* ---- Per-precision bodies.  TILE_MAXITER = tile_maxiter(T): 512 bytes / sizeof(T)
 * (src/loops.jl:1, 7-10). */
#define T double
#define SFX _f64
#define LOGT log
#define TILE_MAXITER 64
#include "oracle_body.h"
#undef T
#undef SFX
#undef LOGT
#undef TILE_MAXITER

#define T float
#define SFX _f32
#define LOGT logf
#define TILE_MAXITER 128
#include "oracle_body.h"
#undef T
#undef SFX
#undef LOGT
#undef TILE_MAXITER
