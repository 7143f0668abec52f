/*
 * oracle_body.h -- per-precision body of the CPU oracle (test infrastructure).
 *
 * Included twice by admix_oracle.c, once with T=double (suffix _f64) and once
 * with T=float (suffix _f32).  Every function restates one function of the
 * reference (OpenADMIXTURE.jl) with the same algebraic form and operation
 * order, cited file:line.  Arrays are column-major exactly as in Julia:
 *   q[k,i] -> q[k + K*i], p[k,j] -> p[k + K*j], XtX[k2,k,i] -> XtX[k2 + K*k + K*K*i]
 * Index ranges are 0-based half-open [i0,i1) x [j0,j1) (the reference's
 * 1-based inclusive UnitRange a:b is [a-1, b)).
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* ---- qp_block!  (src/loops.jl:303-314): s[i,j] = sum_k q[k,i] p[k,j], k ascending */
static void FN(qp_block)(T *qp, const T *q, const T *p, int64_t K, int64_t i0,
                         int64_t i1, int64_t j0, int64_t j1) {
  const int64_t li = i1 - i0;
  for (int64_t j = j0; j < j1; ++j)
    for (int64_t i = i0; i < i1; ++i) {
      T s = 0;
      for (int64_t k = 0; k < K; ++k) s += q[k + K * i] * p[k + K * j];
      qp[(i - i0) + li * (j - j0)] = s;
    }
}

/* Genotype weight g and nonmissing flag m of code c in SNP j.  mu == NULL: the
 * _skipmissing loops (g_map, nonmissing_map; src/loops.jl:2-5).  Otherwise the
 * mean-imputation loops: g = g_map[c] + (c == missing) * mu[j], m = 1
 * (src/loops.jl:493-514, 589-608, 679-698, 840-865, 896-921). */
static inline void FN(gw)(int c, const T *mu, int64_t j, T *g, T *m) {
  if (mu) {
    *g = (T)G_MAP[c] + (T)(c == 1) * mu[j];
    *m = (T)1;
  } else {
    *g = (T)G_MAP[c];
    *m = (T)NM_MAP[c];
  }
}

/* ---- loglikelihood_loop_skipmissing(g::SnpLinAlg,...)  (src/loops.jl:517-539)
 * Terms are evaluated in T, accumulated in Float64 (loops.jl:523). */
static double FN(ll_tile)(const uint8_t *bed, int64_t pitch, const T *q,
                          const T *p, int64_t K, int64_t i0, int64_t i1,
                          int64_t j0, int64_t j1, const T *mu, T *qp) {
  FN(qp_block)(qp, q, p, K, i0, i1, j0, j1);
  const int64_t li = i1 - i0;
  double r = 0.0;
  for (int64_t j = j0; j < j1; ++j)
    for (int64_t i = i0; i < i1; ++i) {
      const int c = orc_code(bed, pitch, i, j);
      T g, m;
      FN(gw)(c, mu, j, &g, &m);
      const T s = qp[(i - i0) + li * (j - j0)];
      const T term = m * ((g * LOGT(s)) + (((T)2 - g) * LOGT((T)1 - s)));
      r += (double)term;
    }
  return r;
}

/* threader_scalar / tiler_scalar recursion (src/loops.jl:203-258): split the
 * longer side with cleave(range) (step 4) until both sides are <= maxL and sum
 * the two halves.  Two-child atomic sums are order independent (a+b == b+a),
 * so this tree reproduces the reference's summation tree exactly. */
static double FN(ll_tree)(const uint8_t *bed, int64_t pitch, const T *q,
                          const T *p, int64_t K, int64_t i0, int64_t i1,
                          int64_t j0, int64_t j1, const T *mu, int depth) {
  const int64_t ilen = i1 - i0, jlen = j1 - j0, maxL = TILE_MAXITER;
  if (ilen > maxL || jlen > maxL) {
    double r1 = 0, r2 = 0;
    if (ilen > jlen) {
      const int64_t m = orc_cleave(i0, i1, 4);
#pragma omp task shared(r1) if (depth < 8)
      r1 = FN(ll_tree)(bed, pitch, q, p, K, i0, m, j0, j1, mu, depth + 1);
      r2 = FN(ll_tree)(bed, pitch, q, p, K, m, i1, j0, j1, mu, depth + 1);
#pragma omp taskwait
    } else {
      const int64_t m = orc_cleave(j0, j1, 4);
#pragma omp task shared(r1) if (depth < 8)
      r1 = FN(ll_tree)(bed, pitch, q, p, K, i0, i1, j0, m, mu, depth + 1);
      r2 = FN(ll_tree)(bed, pitch, q, p, K, i0, i1, m, j1, mu, depth + 1);
#pragma omp taskwait
    }
    return r1 + r2;
  }
  T qp[TILE_MAXITER * TILE_MAXITER];
  return FN(ll_tile)(bed, pitch, q, p, K, i0, i1, j0, j1, mu, qp);
}

/* loglikelihood(g::SnpLinAlg, q, p, qp_small, K, true)  (src/algorithms_inner.jl:27-35) */
double FN(orc_loglikelihood)(const uint8_t *bed, int64_t pitch, const T *q,
                             const T *p, int64_t K, int64_t i0, int64_t i1,
                             int64_t j0, int64_t j1, const T *mu, int nthreads) {
  double r = 0;
  if (i1 <= i0 || j1 <= j0) return 0.0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
#pragma omp single
  r = FN(ll_tree)(bed, pitch, q, p, K, i0, i1, j0, j1, mu, 0);
  return r;
}

/* ---- em_q_loop_skipmissing!(q_next, g::SnpLinAlg, ...)  (src/loops.jl:610-632) */
static void FN(em_q_tile)(T *q_next, const uint8_t *bed, int64_t pitch,
                          const T *q, const T *p, int64_t K, int64_t i0,
                          int64_t i1, int64_t j0, int64_t j1, const T *mu, T *qp) {
  FN(qp_block)(qp, q, p, K, i0, i1, j0, j1);
  const int64_t li = i1 - i0;
  for (int64_t j = j0; j < j1; ++j)
    for (int64_t i = i0; i < i1; ++i) {
      const int c = orc_code(bed, pitch, i, j);
      T g, m;
      FN(gw)(c, mu, j, &g, &m);
      const T s = qp[(i - i0) + li * (j - j0)];
      for (int64_t k = 0; k < K; ++k) {
        const T qk = q[k + K * i], pk = p[k + K * j];
        const T tmp = (g * qk * pk / s + ((T)2 - g) * qk * ((T)1 - pk) / ((T)1 - s));
        q_next[k + K * i] += m * tmp;
      }
    }
}

/* ---- em_p_loop_skipmissing!(f_next, g::SnpLinAlg, q, p, p_tmp, ...)  (src/loops.jl:700-721) */
static void FN(em_p_tile)(T *p_next, T *p_tmp, const uint8_t *bed,
                          int64_t pitch, const T *q, const T *p, int64_t K,
                          int64_t i0, int64_t i1, int64_t j0, int64_t j1,
                          const T *mu, T *qp) {
  FN(qp_block)(qp, q, p, K, i0, i1, j0, j1);
  const int64_t li = i1 - i0;
  for (int64_t j = j0; j < j1; ++j)
    for (int64_t i = i0; i < i1; ++i) {
      const int c = orc_code(bed, pitch, i, j);
      T g, m;
      FN(gw)(c, mu, j, &g, &m);
      const T s = qp[(i - i0) + li * (j - j0)];
      for (int64_t k = 0; k < K; ++k) {
        const T qk = q[k + K * i], pk = p[k + K * j];
        p_tmp[k + K * j] += m * (g * qk * pk / s);
        p_next[k + K * j] += m * (((T)2 - g) * qk * ((T)1 - pk) / ((T)1 - s));
      }
    }
}

/* ---- update_q_loop_skipmissing!(XtX, Xtz, g::SnpLinAlg, ...)  (src/loops.jl:867-894) */
static void FN(update_q_tile)(T *XtX, T *Xtz, const uint8_t *bed, int64_t pitch,
                              const T *q, const T *p, int64_t K, int64_t i0,
                              int64_t i1, int64_t j0, int64_t j1, const T *mu, T *qp) {
  FN(qp_block)(qp, q, p, K, i0, i1, j0, j1);
  const int64_t li = i1 - i0;
  for (int64_t j = j0; j < j1; ++j)
    for (int64_t i = i0; i < i1; ++i) {
      const int c = orc_code(bed, pitch, i, j);
      T g, m;
      FN(gw)(c, mu, j, &g, &m);
      const T s = qp[(i - i0) + li * (j - j0)];
      for (int64_t k = 0; k < K; ++k) {
        const T pk = p[k + K * j];
        Xtz[k + K * i] += m * (g * pk / s + ((T)2 - g) * ((T)1 - pk) / ((T)1 - s));
        for (int64_t k2 = 0; k2 < K; ++k2) {
          const T pk2 = p[k2 + K * j];
          XtX[k2 + K * k + K * K * i] +=
              m * (g / (s * s) * pk * pk2 +
                   ((T)2 - g) / (((T)1 - s) * ((T)1 - s)) * ((T)1 - pk) * ((T)1 - pk2));
        }
      }
    }
}

/* ---- update_p_loop_skipmissing!(XtX, Xtz, g::SnpLinAlg, ...)  (src/loops.jl:923-950) */
static void FN(update_p_tile)(T *XtX, T *Xtz, const uint8_t *bed, int64_t pitch,
                              const T *q, const T *p, int64_t K, int64_t i0,
                              int64_t i1, int64_t j0, int64_t j1, const T *mu, T *qp) {
  FN(qp_block)(qp, q, p, K, i0, i1, j0, j1);
  const int64_t li = i1 - i0;
  for (int64_t j = j0; j < j1; ++j)
    for (int64_t i = i0; i < i1; ++i) {
      const int c = orc_code(bed, pitch, i, j);
      T g, m;
      FN(gw)(c, mu, j, &g, &m);
      const T s = qp[(i - i0) + li * (j - j0)];
      for (int64_t k = 0; k < K; ++k) {
        const T qk = q[k + K * i];
        Xtz[k + K * j] += m * (g * qk / s - ((T)2 - g) * qk / ((T)1 - s));
        for (int64_t k2 = 0; k2 < K; ++k2) {
          const T qk2 = q[k2 + K * i];
          XtX[k2 + K * k + K * K * j] +=
              m * (g / (s * s) * qk * qk2 + ((T)2 - g) / (((T)1 - s) * ((T)1 - s)) * qk * qk2);
        }
      }
    }
}

/* Tile-loop entry points with the reference's accumulate (+=) semantics over an
 * arbitrary (irange, jrange).  Work is split over i (Q side) or j (P side), as
 * threader! does (src/loops.jl:63-122); each output column is owned by one
 * thread and accumulates over the other index in ascending order, tile after
 * tile, so the result does not depend on the thread count. */
void FN(orc_em_q_loop)(T *q_next, const uint8_t *bed, int64_t pitch, const T *q,
                       const T *p, int64_t K, int64_t i0, int64_t i1, int64_t j0,
                       int64_t j1, const T *mu, int nthreads) {
  const int64_t L = TILE_MAXITER;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t ib = i0; ib < i1; ib += L) {
    T qp[TILE_MAXITER * TILE_MAXITER];
    const int64_t ie = ib + L < i1 ? ib + L : i1;
    for (int64_t jb = j0; jb < j1; jb += L)
      FN(em_q_tile)(q_next, bed, pitch, q, p, K, ib, ie, jb, jb + L < j1 ? jb + L : j1, mu, qp);
  }
}

void FN(orc_em_p_loop)(T *p_next, T *p_tmp, const uint8_t *bed, int64_t pitch,
                       const T *q, const T *p, int64_t K, int64_t i0, int64_t i1,
                       int64_t j0, int64_t j1, const T *mu, int nthreads) {
  const int64_t L = TILE_MAXITER;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t jb = j0; jb < j1; jb += L) {
    T qp[TILE_MAXITER * TILE_MAXITER];
    const int64_t je = jb + L < j1 ? jb + L : j1;
    for (int64_t ib = i0; ib < i1; ib += L)
      FN(em_p_tile)(p_next, p_tmp, bed, pitch, q, p, K, ib, ib + L < i1 ? ib + L : i1, jb, je, mu, qp);
  }
}

void FN(orc_update_q_loop)(T *XtX, T *Xtz, const uint8_t *bed, int64_t pitch,
                           const T *q, const T *p, int64_t K, int64_t i0,
                           int64_t i1, int64_t j0, int64_t j1, const T *mu, int nthreads) {
  const int64_t L = TILE_MAXITER;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t ib = i0; ib < i1; ib += L) {
    T qp[TILE_MAXITER * TILE_MAXITER];
    const int64_t ie = ib + L < i1 ? ib + L : i1;
    for (int64_t jb = j0; jb < j1; jb += L)
      FN(update_q_tile)(XtX, Xtz, bed, pitch, q, p, K, ib, ie, jb, jb + L < j1 ? jb + L : j1, mu, qp);
  }
}

void FN(orc_update_p_loop)(T *XtX, T *Xtz, const uint8_t *bed, int64_t pitch,
                           const T *q, const T *p, int64_t K, int64_t i0,
                           int64_t i1, int64_t j0, int64_t j1, const T *mu, int nthreads) {
  const int64_t L = TILE_MAXITER;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t jb = j0; jb < j1; jb += L) {
    T qp[TILE_MAXITER * TILE_MAXITER];
    const int64_t je = jb + L < j1 ? jb + L : j1;
    for (int64_t ib = i0; ib < i1; ib += L)
      FN(update_p_tile)(XtX, Xtz, bed, pitch, q, p, K, ib, ib + L < i1 ? ib + L : i1, jb, je, mu, qp);
  }
}

/* ---- sweep!  (src/qp.jl:297-317).  A is n x n column-major. */
void FN(orc_sweep)(T *A, int64_t n, int64_t k, T *tmp, int inverse) {
  const T pk = (T)1 / A[k + n * k];
  for (int64_t i = 0; i < n; ++i) {
    tmp[i] = A[i + n * k];
    A[i + n * k] = 0;
    A[k + n * i] = 0;
  }
  tmp[k] = inverse ? (T)1 : (T)-1;
  for (int64_t i = 0; i < n; ++i) {
    const T pv = pk * tmp[i];
    for (int64_t j = 0; j < n; ++j) A[j + n * i] = A[j + n * i] - pv * tmp[j];
  }
}

/* ---- create_tableau!(tableau, matrix_q, r, x, v, tmp_k, simplex)  (src/qp.jl:8-76)
 * The Gerschgorin constant mu is computed at qp.jl:54-57 through
 * mul!(tmp_k, v, tmp_k), whose output aliases its input.  LinearAlgebra hands
 * that to BLAS gemv with beta = 0, which clears y before reading x, so the
 * executed value is mu = 0; mu only shifts the objective by a constant on the
 * simplex constraint set {1'delta = 1 - sum(x)}.  `mu` is taken as an argument
 * so either reading can be checked. */
void FN(orc_create_tableau)(T *tab, const T *H, const T *r, const T *x,
                            int64_t K, int simplex, T mu) {
  const int64_t sz = simplex ? K + 2 : K + 1;
  for (int64_t a = 0; a < sz * sz; ++a) tab[a] = 0;
  for (int64_t j = 0; j < K; ++j)
    for (int64_t i = 0; i < K; ++i) tab[i + sz * j] = H[i + K * j];
  for (int64_t i = 0; i < K; ++i) {
    tab[(sz - 1) + sz * i] = -r[i];
    tab[i + sz * (sz - 1)] = -r[i];
  }
  tab[(sz - 1) + sz * (sz - 1)] = 0;
  if (!simplex) return;
  for (int64_t j = 0; j < K; ++j)
    for (int64_t i = 0; i < K; ++i) tab[i + sz * j] += mu;
  for (int64_t i = 0; i < K; ++i) {
    tab[K + sz * i] = 1;
    tab[i + sz * K] = 1;
  }
  tab[K + sz * K] = 0;
  T sx = 0;
  for (int64_t i = 0; i < K; ++i) sx += x[i];
  tab[(K + 1) + sz * K] = (T)1 - sx;
  tab[K + sz * (K + 1)] = (T)1 - sx;
  tab[(K + 1) + sz * (K + 1)] = 0;
}

/* ---- quadratic_program!(delta, tableau, par, pmin, pmax, p, c, d, tmp, swept)
 * (src/qp.jl:171-292), Jennrich-Sampson active-set method with sweeps.
 * Returns the iteration count, or 0 on a failed pivot test (delta = 0) or on
 * exhausting 1000 iterations (delta keeps its partial value). */
int FN(orc_quadratic_program)(T *delta, T *tab, const T *par, const T *pmin,
                              const T *pmax, int64_t p, int64_t c, T *d, T *tmp,
                              uint8_t *swept) {
  const int64_t sz = p + c + 1, e = sz - 1;
  const T small = (T)1e-5, tol = (T)1e-8;
  for (int64_t i = 0; i < p; ++i) delta[i] = 0;
  for (int64_t i = 0; i < sz; ++i) d[i] = tab[i + sz * i];
  for (int64_t i = 0; i < p; ++i) {
    if (d[i] <= (T)0 || tab[i + sz * i] < d[i] * tol) return 0;
    FN(orc_sweep)(tab, sz, i, tmp, 0);
  }
  for (int64_t i = 0; i < p; ++i) swept[i] = 1;
  for (int64_t i = p; i < p + c; ++i) {
    if (tab[i + sz * i] >= (T)0) return 0;
    FN(orc_sweep)(tab, sz, i, tmp, 0);
  }
  for (int iteration = 1; iteration <= 1000; ++iteration) {
    T a = 1;
    for (int64_t i = 0; i < p; ++i) {
      if (swept[i]) {
        const T ui = tab[i + sz * e];
        T ai;
        if (ui > (T)0) ai = pmax[i] - par[i] - delta[i];
        else ai = pmin[i] - par[i] - delta[i];
        if ((T)fabs((double)ui) > (T)1e-10) { const T t = ai / ui; a = t < a ? t : a; } /* abs(ui) > T(1e-10) */
      }
    }
    for (int64_t i = 0; i < p; ++i) {
      if (swept[i]) {
        const T ui = tab[i + sz * e];
        delta[i] = delta[i] + a * ui;
        tab[i + sz * e] = ((T)1 - a) * ui;
        tab[e + sz * i] = tab[i + sz * e];
      }
    }
    int cycle = 0;
    for (int64_t i = 0; i < p; ++i) {
      int critical = pmin[i] >= par[i] + delta[i] - small;
      critical = critical || pmax[i] <= par[i] + delta[i] + small;
      if (swept[i] && fabs((double)tab[i + sz * i]) > 1e-10 && critical) { /* Float64 literal */
        FN(orc_sweep)(tab, sz, i, tmp, 1);
        swept[i] = 0;
        cycle = 1;
        break;
      }
    }
    if (cycle) continue;
    for (int64_t i = 0; i < p; ++i) {
      const T ui = tab[i + sz * e];
      int violation = ui > (T)0 && pmin[i] >= par[i] + delta[i] - small;
      violation = violation || (ui < (T)0 && pmax[i] <= par[i] + delta[i] + small);
      if (!swept[i] && violation) {
        FN(orc_sweep)(tab, sz, i, tmp, 0);
        swept[i] = 1;
        cycle = 1;
        break;
      }
    }
    if (cycle) continue;
    return iteration;
  }
  return 0;
}

/* ---- project_q!(b, idx; pseudocount=1e-5) on each column of a K x n matrix
 * (src/projections.jl:1-63).  The custom quicksort orders indices by value,
 * and only the sorted values enter the result, so any descending sort gives
 * the same bits. */
void FN(orc_project_q)(T *b, int64_t K, int64_t n) {
  const T pc = (T)1e-5;
  for (int64_t col = 0; col < n; ++col) {
    T *v = b + K * col;
    T srt[64];
    const T tau = (T)1 - (T)K * pc;
    for (int64_t i = 0; i < K; ++i) v[i] -= pc;
    for (int64_t i = 0; i < K; ++i) srt[i] = v[i];
    for (int64_t i = 1; i < K; ++i) { /* insertion sort, descending */
      T x = srt[i];
      int64_t j = i - 1;
      while (j >= 0 && srt[j] < x) { srt[j + 1] = srt[j]; --j; }
      srt[j + 1] = x;
    }
    T tsum = 0, tmax = 0;
    int bget = 0;
    for (int64_t i = 0; i < K - 1; ++i) {
      tsum += srt[i];
      tmax = (tsum - tau) / (T)(i + 1);
      if (tmax >= srt[i + 1]) { bget = 1; break; }
    }
    if (!bget) tmax = (tsum + srt[K - 1] - tau) / (T)K;
    for (int64_t i = 0; i < K; ++i) {
      const T t = v[i] - tmax;
      v[i] = (t > (T)0 ? t : (T)0) + pc;
    }
  }
}

/* ---- project_p!(b; pseudocount=1e-5)  (src/projections.jl:69-74) */
void FN(orc_project_p)(T *b, int64_t n) {
  const T pc = (T)1e-5;
  for (int64_t i = 0; i < n; ++i) {
    T v = b[i] > pc ? b[i] : pc;
    const T hi = (T)1 - pc;
    b[i] = v < hi ? v : hi;
  }
}

/* ---- update_UV!  (src/quasi_newton.jl:5-9).  iter is 1-based as in the reference. */
void FN(orc_update_UV)(T *U, T *V, const T *x, const T *x_next,
                       const T *x_next2, int64_t N, int64_t iter, int64_t Q) {
  const int64_t idx = (iter - 1) % Q;
  for (int64_t n = 0; n < N; ++n) {
    V[n + N * idx] = x_next2[n] - x_next[n];
    U[n + N * idx] = x_next[n] - x[n];
  }
}

/* ---- update_QN!(x_next, x_mapped, x, U, V)  (src/quasi_newton.jl:29-31):
 *   x_next = x_mapped - V * (inv(U'(U - V)) * (U'(x - x_mapped)))
 * using the first m columns of U and V.  Inner products are accumulated in
 * double; inv is Gauss-Jordan elimination with partial pivoting. */
void FN(orc_update_QN)(T *x_out, const T *x_mapped, const T *x, const T *U,
                       const T *V, int64_t N, int64_t m) {
  double A[16 * 16] = {0}, Ainv[16 * 16] = {0}, rhs[16] = {0}, coef[16] = {0};
  for (int64_t a = 0; a < m; ++a) {
    double s = 0;
    for (int64_t n = 0; n < N; ++n) s += (double)U[n + N * a] * (double)(x[n] - x_mapped[n]);
    rhs[a] = s;
    for (int64_t b = 0; b < m; ++b) {
      double t = 0;
      for (int64_t n = 0; n < N; ++n)
        t += (double)U[n + N * a] * (double)(U[n + N * b] - V[n + N * b]);
      A[a + m * b] = t;
    }
  }
  orc_invert(A, Ainv, m);
  for (int64_t a = 0; a < m; ++a) {
    double s = 0;
    for (int64_t b = 0; b < m; ++b) s += Ainv[a + m * b] * rhs[b];
    coef[a] = s;
  }
  for (int64_t n = 0; n < N; ++n) {
    T s = 0;
    for (int64_t a = 0; a < m; ++a) s += V[n + N * a] * (T)coef[a];
    x_out[n] = x_mapped[n] - s;
  }
}

/* ---- update_UV_LBQN!  (src/quasi_newton.jl:15-23): shift the first
 * min(iter, Q) - 1 columns of U and V right by one, then
 * U[:,1] = x_next - x and V[:,1] = (x_next2 - x_next) - U[:,1]. */
void FN(orc_update_UV_LBQN)(T *U, T *V, const T *x, const T *x_next,
                            const T *x_next2, int64_t N, int64_t iter, int64_t Q) {
  const int64_t m = iter < Q ? iter : Q;
  for (int64_t i = m - 1; i >= 1; --i) {
    memcpy(U + N * i, U + N * (i - 1), sizeof(T) * N);
    memcpy(V + N * i, V + N * (i - 1), sizeof(T) * N);
  }
  for (int64_t n = 0; n < N; ++n) {
    U[n] = x_next[n] - x[n];
    V[n] = x_next2[n] - x_next[n];
    V[n] = V[n] - U[n];
  }
}

/* ---- update_QN_LBQN!(x_next, x, x_q, x_r, U, V)  (src/quasi_newton.jl:37-58)
 * with the first m columns: gamma = u'v / v'v (u, v the first columns),
 * x_q = u; for i = 1..m: alpha_i = (V_i'x_q) / (V_i'V_i), x_q -= alpha_i V_i;
 * x_r = gamma x_q + sum_i alpha_i U_i; x_next = x - x_r.  Inner products in
 * double, vector updates in T in the reference's order. */
void FN(orc_update_QN_LBQN)(T *x_out, const T *x, const T *U, const T *V, int64_t N, int64_t m) {
  T *xq = (T *)malloc(sizeof(T) * N), *xr = (T *)malloc(sizeof(T) * N);
  double uv = 0, vv = 0, alpha[16] = {0};
  for (int64_t n = 0; n < N; ++n) {
    uv += (double)U[n] * (double)V[n];
    vv += (double)V[n] * (double)V[n];
  }
  const double gamma = uv / vv;
  memcpy(xq, U, sizeof(T) * N);
  for (int64_t i = 0; i < m; ++i) {
    const T *vi = V + N * i;
    double d2 = 0, dq = 0;
    for (int64_t n = 0; n < N; ++n) {
      d2 += (double)vi[n] * (double)vi[n];
      dq += (double)vi[n] * (double)xq[n];
    }
    alpha[i] = (1.0 / d2) * dq;
    for (int64_t n = 0; n < N; ++n) xq[n] = xq[n] - (T)alpha[i] * vi[n];
  }
  for (int64_t n = 0; n < N; ++n) xr[n] = (T)gamma * xq[n];
  for (int64_t i = 0; i < m; ++i)
    for (int64_t n = 0; n < N; ++n) xr[n] = xr[n] + (T)alpha[i] * U[n + N * i];
  for (int64_t n = 0; n < N; ++n) x_out[n] = x[n] - xr[n];
  free(xq);
  free(xr);
}

/* ----------------------------------------------------------------------------
 * Whole-algorithm state (AdmixData{T}, src/structs.jl:9-91), host arrays only.
 */
typedef struct {
  int64_t I, J, K, Q;
  const uint8_t *bed;
  int64_t pitch;
  T *x, *x_next, *x_next2, *x_tmp; /* K x (I+J) each; q = cols [0,I), p = cols [I,I+J) */
  T *XtX_q, *Xtz_q, *XtX_p, *Xtz_p;
  T *U, *V; /* K(I+J) x Q */
  double ll_prev, ll_new;
  int nthreads;
  const T *mu; /* NULL: skipmissing; else per-SNP means for the mean-imputation loops */
} FN(orc_state);

static T *FN(qpart)(FN(orc_state) * st, T *x) { (void)st; return x; }
static T *FN(ppart)(FN(orc_state) * st, T *x) { return x + st->K * st->I; }

static double FN(st_ll)(FN(orc_state) * st, T *x) {
  return FN(orc_loglikelihood)(st->bed, st->pitch, FN(qpart)(st, x), FN(ppart)(st, x),
                               st->K, 0, st->I, 0, st->J, st->mu, st->nthreads);
}

/* em_q!(d, g; mode)  (src/algorithms_inner.jl:68-92), divided by 2J.
 * mode 0 (:base) and 1 (:fallback): q_next from (q, p); 2 (:fallback2):
 * q_next2 from (q_next, p_next). */
void FN(orc_em_q_mode)(FN(orc_state) * st, int mode) {
  const int64_t K = st->K, I = st->I, J = st->J;
  T *dst = mode == 2 ? st->x_next2 : st->x_next, *src = mode == 2 ? st->x_next : st->x;
  T *q_next = FN(qpart)(st, dst);
  memset(q_next, 0, sizeof(T) * K * I);
  FN(orc_em_q_loop)(q_next, st->bed, st->pitch, FN(qpart)(st, src), FN(ppart)(st, src), K, 0, I, 0, J, st->mu,
                    st->nthreads);
  const T div = (T)(2 * J);
  for (int64_t a = 0; a < K * I; ++a) q_next[a] /= div;
}

/* em_p!(d, g; mode)  (src/algorithms_inner.jl:108-141): p_next = p_tmp / (p_tmp + p_next).
 * mode 0 (:base): p_next from (q, p); 1 (:fallback): p_next from (q_next, p);
 * 2 (:fallback2): p_next2 from (q_next2, p_next). */
void FN(orc_em_p_mode)(FN(orc_state) * st, int mode) {
  const int64_t K = st->K, I = st->I, J = st->J;
  T *dst = mode == 2 ? st->x_next2 : st->x_next;
  const T *q = FN(qpart)(st, mode == 0 ? st->x : (mode == 1 ? st->x_next : st->x_next2));
  const T *p = FN(ppart)(st, mode == 2 ? st->x_next : st->x);
  T *p_next = FN(ppart)(st, dst), *p_tmp = FN(ppart)(st, st->x_tmp);
  memset(p_next, 0, sizeof(T) * K * J);
  memset(p_tmp, 0, sizeof(T) * K * J);
  FN(orc_em_p_loop)(p_next, p_tmp, st->bed, st->pitch, q, p, K, 0, I, 0, J, st->mu, st->nthreads);
  for (int64_t a = 0; a < K * J; ++a) p_next[a] = p_tmp[a] / (p_tmp[a] + p_next[a]);
}

void FN(orc_em_q)(FN(orc_state) * st) { FN(orc_em_q_mode)(st, 0); }
void FN(orc_em_p)(FN(orc_state) * st) { FN(orc_em_p_mode)(st, 0); }

/* init_em!(d, g, iter)  (src/algorithms_outer.jl:12-45); lls[it] = LL after step it */
void FN(orc_init_em)(FN(orc_state) * st, int64_t iters, double *lls) {
  const int64_t KN = st->K * (st->I + st->J);
  for (int64_t it = 0; it < iters; ++it) {
    FN(orc_em_q)(st);
    FN(orc_em_p)(st);
    memcpy(FN(ppart)(st, st->x), FN(ppart)(st, st->x_next), sizeof(T) * st->K * st->J);
    memcpy(FN(qpart)(st, st->x), FN(qpart)(st, st->x_next), sizeof(T) * st->K * st->I);
    if (lls) lls[it] = FN(st_ll)(st, st->x);
  }
  (void)KN;
  st->ll_new = FN(st_ll)(st, st->x);
}

/* The host half of update_q! (src/algorithms_inner.jl:197-230), which the
 * reference runs on the CPU after its GPU branch has copied XtX and Xtz back:
 * Xtz .*= -1, one simplex QP per individual (tableau, quadratic_program!),
 * q_next = q + qdiff, project_q!.  XtX (K x K x I) and Xtz (K x I, negated in
 * place) are the pass outputs; qdiff is K x I scratch. */
void FN(orc_qp_update_q)(T *XtX, T *Xtz, const T *q, T *q_next, T *qdiff, int64_t K, int64_t I, int32_t *status,
                         int nthreads) {
  for (int64_t a = 0; a < K * I; ++a) Xtz[a] *= (T)-1;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < I; ++i) {
    T tab[18 * 18], dd[18], tmp[18], pmin[16], pmax[16];
    uint8_t swept[16];
    for (int64_t k = 0; k < K; ++k) { pmin[k] = 0; pmax[k] = 1; }
    FN(orc_create_tableau)(tab, XtX + K * K * i, Xtz + K * i, q + K * i, K, 1, (T)0);
    const int r = FN(orc_quadratic_program)(qdiff + K * i, tab, q + K * i, pmin, pmax, K, 1, dd, tmp, swept);
    if (status) status[i] = r;
  }
  for (int64_t a = 0; a < K * I; ++a) q_next[a] = q[a] + qdiff[a];
  FN(orc_project_q)(q_next, K, I);
}

/* The host half of update_p! (src/algorithms_inner.jl:273-307): Xtz .*= -1,
 * one box QP per SNP, p_next = p + pdiff, project_p!. */
void FN(orc_qp_update_p)(T *XtX, T *Xtz, const T *p, T *p_next, T *pdiff, int64_t K, int64_t J, int32_t *status,
                         int nthreads) {
  for (int64_t a = 0; a < K * J; ++a) Xtz[a] *= (T)-1;
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < J; ++j) {
    T tab[18 * 18], dd[18], tmp[18], pmin[16], pmax[16];
    uint8_t swept[16];
    for (int64_t k = 0; k < K; ++k) { pmin[k] = 0; pmax[k] = 1; }
    FN(orc_create_tableau)(tab, XtX + K * K * j, Xtz + K * j, p + K * j, K, 0, (T)0);
    const int r = FN(orc_quadratic_program)(pdiff + K * j, tab, p + K * j, pmin, pmax, K, 0, dd, tmp, swept);
    if (status) status[j] = r;
  }
  for (int64_t a = 0; a < K * J; ++a) p_next[a] = p[a] + pdiff[a];
  FN(orc_project_p)(p_next, K * J);
}

/* update_q!(d, g, update2)  (src/algorithms_inner.jl:167-231) */
void FN(orc_update_q)(FN(orc_state) * st, int update2, int32_t *status) {
  const int64_t K = st->K, I = st->I, J = st->J;
  T *q_next = FN(qpart)(st, update2 ? st->x_next2 : st->x_next);
  const T *q = FN(qpart)(st, update2 ? st->x_next : st->x);
  const T *p = FN(ppart)(st, update2 ? st->x_next : st->x);
  T *qdiff = FN(qpart)(st, st->x_tmp);
  st->ll_prev = st->ll_new;
  memset(st->XtX_q, 0, sizeof(T) * K * K * I);
  memset(st->Xtz_q, 0, sizeof(T) * K * I);
  FN(orc_update_q_loop)(st->XtX_q, st->Xtz_q, st->bed, st->pitch, q, p, K, 0, I, 0, J, st->mu, st->nthreads);
  FN(orc_qp_update_q)(st->XtX_q, st->Xtz_q, q, q_next, qdiff, K, I, status, st->nthreads);
}

/* update_p!(d, g, update2)  (src/algorithms_inner.jl:245-310); uses q_next (Gauss-Seidel) */
void FN(orc_update_p)(FN(orc_state) * st, int update2, int32_t *status) {
  const int64_t K = st->K, I = st->I, J = st->J;
  T *p_next = FN(ppart)(st, update2 ? st->x_next2 : st->x_next);
  const T *q = FN(qpart)(st, update2 ? st->x_next2 : st->x_next);
  const T *p = FN(ppart)(st, update2 ? st->x_next : st->x);
  T *pdiff = FN(ppart)(st, st->x_tmp);
  st->ll_prev = st->ll_new;
  memset(st->XtX_p, 0, sizeof(T) * K * K * J);
  memset(st->Xtz_p, 0, sizeof(T) * K * J);
  FN(orc_update_p_loop)(st->XtX_p, st->Xtz_p, st->bed, st->pitch, q, p, K, 0, I, 0, J, st->mu, st->nthreads);
  FN(orc_qp_update_p)(st->XtX_p, st->Xtz_p, p, p_next, pdiff, K, J, status, st->nthreads);
}

/* One iteration of admixture_qn!(d, g; mode)  (src/algorithms_outer.jl:92-159),
 * mode 0 = :ZAL, 1 = :LBQN (:127-136).  iter is the 1-based outer index i.
 * out = {prev, basic, qn, new, reldiff}. */
void FN(orc_qn_iteration_mode)(FN(orc_state) * st, int64_t iter, int fix_q, int fix_p, int mode, double *out) {
  const int64_t K = st->K, I = st->I, J = st->J, N = K * (I + J);
  st->ll_prev = st->ll_new;
  const double ll_prev = st->ll_prev;
  if (!fix_q) FN(orc_update_q)(st, 0, NULL);
  else memcpy(st->x_next, st->x, sizeof(T) * K * I);
  if (!fix_p) FN(orc_update_p)(st, 0, NULL);
  else memcpy(FN(ppart)(st, st->x_next), FN(ppart)(st, st->x), sizeof(T) * K * J);
  if (!fix_q) FN(orc_update_q)(st, 1, NULL);
  else memcpy(st->x_next2, st->x, sizeof(T) * K * I);
  if (!fix_p) FN(orc_update_p)(st, 1, NULL);
  else memcpy(FN(ppart)(st, st->x_next2), FN(ppart)(st, st->x), sizeof(T) * K * J);
  /* update_q!/update_p! set d.ll_prev = d.ll_new; the outer value is the one compared */
  st->ll_prev = ll_prev;
  const double ll_basic = FN(st_ll)(st, st->x_next2);
  const int64_t m = iter < st->Q ? iter : st->Q;
  if (mode == 1) {
    FN(orc_update_UV_LBQN)(st->U, st->V, st->x, st->x_next, st->x_next2, N, iter, st->Q);
    FN(orc_update_QN_LBQN)(st->x_tmp, st->x, st->U, st->V, N, m);
  } else {
    FN(orc_update_UV)(st->U, st->V, st->x, st->x_next, st->x_next2, N, iter, st->Q);
    FN(orc_update_QN)(st->x_tmp, st->x_next, st->x, st->U, st->V, N, m);
  }
  FN(orc_project_p)(FN(ppart)(st, st->x_tmp), K * J);
  FN(orc_project_q)(FN(qpart)(st, st->x_tmp), K, I);
  const double ll_qn = FN(st_ll)(st, st->x_tmp);
  if (st->ll_prev < ll_qn) {
    memcpy(st->x, st->x_tmp, sizeof(T) * N);
    st->ll_new = ll_qn;
  } else {
    memcpy(st->x, st->x_next2, sizeof(T) * N);
    st->ll_new = ll_basic;
  }
  out[0] = st->ll_prev;
  out[1] = ll_basic;
  out[2] = ll_qn;
  out[3] = st->ll_new;
  out[4] = fabs((st->ll_new - st->ll_prev) / st->ll_prev);
}

void FN(orc_qn_iteration)(FN(orc_state) * st, int64_t iter, int fix_q, int fix_p, double *out) {
  FN(orc_qn_iteration_mode)(st, iter, fix_q, fix_p, 0, out);
}

#undef FN
#undef CAT
#undef CAT2
