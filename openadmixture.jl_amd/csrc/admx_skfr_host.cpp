// admx_skfr_host.cpp -- host orchestration of sparse k-means feature ranking
// (SKFR, the AIM pre-selection of run_admixture(...; sparsity),
// src/driver.jl:56-69, 80-127) on a context's resident BED: admx_skfr_init and
// admx_skfr_run.  The arithmetic runs in admx_skfr.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "admixhip.h"
#include "admx_host.hpp"
#include "admx_internal.hpp"

using namespace admx;
using namespace admx_host;

namespace {

// ---- sparse k-means feature ranking (SKFR)
struct SkfrView {
  int32_t* lab;
  uint32_t* cm;
  unsigned* nk;
  unsigned long long *changes, *count;
  int32_t* center;
  unsigned long long *nm, *alt;
  uint32_t* part;
  double *mu, *var, *ukey, *E, *key, *keys;
  int32_t *idx, *sidx;
  uint8_t *flag, *sel;
  double *T, *P, *x2, *dmin, *D2, *out, *Dc;
  void* tmp;
  size_t tmp_bytes;
  int64_t ntiles, nwords, tper, nslab;
};

// Sizes every SKFR buffer for k clusters, tables of up to ncols_max selected
// SNPs and p_elems table-sum partials, and computes the SNP counts and
// (mu_j, v_j) of the context's genotypes.
int skfr_prepare(admx_ctx* c, int k, int64_t ncols_max, int64_t p_elems, bool need_dc, SkfrView* v) {
  const int64_t I = c->I, J = c->J;
  const int64_t nb = (I + 3) / 4;
  v->ntiles = (nb + 63) / 64;  // 64-byte tiles; the row pitch is a multiple of 64
  v->nwords = 16 * v->ntiles;
  v->nslab = skfr_slabs(J, v->ntiles, &v->tper);
  CHK(c->sk_lab.grow(sizeof(int32_t) * I));
  CHK(c->sk_cm.grow(sizeof(uint32_t) * v->nwords * std::max(k - 1, 1)));
  CHK(c->sk_small.grow(256));
  CHK(c->sk_cnt.grow(sizeof(unsigned long long) * 2 * J));
  CHK(c->sk_part.grow(sizeof(uint32_t) * v->nslab * 2 * std::max(k - 1, 1) * J));
  CHK(c->sk_stat.grow(sizeof(double) * 3 * J));
  CHK(c->sk_E.grow(sizeof(double) * k * J));
  CHK(c->sk_key.grow(sizeof(double) * k * J));
  CHK(c->sk_keys.grow(sizeof(double) * k * J));
  CHK(c->sk_idx.grow(sizeof(int32_t) * J));
  CHK(c->sk_sidx.grow(sizeof(int32_t) * k * J));
  CHK(c->sk_flag.grow((size_t)k * J));
  CHK(c->sk_sel.grow((size_t)J));
  const int64_t trows = std::max<int64_t>(std::max<int64_t>(4 * J, (int64_t)k * J), 4 * k * ncols_max);
  CHK(c->sk_T.grow(sizeof(double) * trows));
  CHK(c->sk_P.grow(sizeof(double) * (size_t)p_elems));
  CHK(c->sk_vec.grow(sizeof(double) * (3 * I + kMaxK + 1)));
  if (need_dc) CHK(c->sk_Dc.grow(sizeof(double) * k * I));
  const size_t tb = skfr_sort_bytes(J);
  CHK(c->sk_tmp.grow(tb));
  char* sm = static_cast<char*>(c->sk_small.p);
  v->lab = static_cast<int32_t*>(c->sk_lab.p);
  v->cm = static_cast<uint32_t*>(c->sk_cm.p);
  v->nk = reinterpret_cast<unsigned*>(sm);
  v->changes = reinterpret_cast<unsigned long long*>(sm + 64);
  v->count = reinterpret_cast<unsigned long long*>(sm + 72);
  v->center = reinterpret_cast<int32_t*>(sm + 80);
  v->nm = static_cast<unsigned long long*>(c->sk_cnt.p);
  v->alt = v->nm + J;
  v->part = static_cast<uint32_t*>(c->sk_part.p);
  v->mu = static_cast<double*>(c->sk_stat.p);
  v->var = v->mu + J;
  v->ukey = v->var + J;
  v->E = static_cast<double*>(c->sk_E.p);
  v->key = static_cast<double*>(c->sk_key.p);
  v->keys = static_cast<double*>(c->sk_keys.p);
  v->idx = static_cast<int32_t*>(c->sk_idx.p);
  v->sidx = static_cast<int32_t*>(c->sk_sidx.p);
  v->flag = static_cast<uint8_t*>(c->sk_flag.p);
  v->sel = static_cast<uint8_t*>(c->sk_sel.p);
  v->T = static_cast<double*>(c->sk_T.p);
  v->P = static_cast<double*>(c->sk_P.p);
  v->x2 = static_cast<double*>(c->sk_vec.p);
  v->dmin = v->x2 + I;
  v->D2 = v->dmin + I;
  v->out = v->D2 + I;
  v->Dc = static_cast<double*>(c->sk_Dc.p);
  v->tmp = c->sk_tmp.p;
  v->tmp_bytes = c->sk_tmp.n;
  HIPCHK(hipSetDevice(c->device));
  {
    ProfScope ps(c, "skfr_stats");
    launch_counts_snp(c->bed, c->pitch, I, J, v->nm, v->alt, c->s);
    launch_skfr_stats(v->nm, v->alt, J, v->mu, v->var, c->s);
  }
  return check_launch();
}

int skfr_split(admx_ctx* c, int nsplit, int64_t ncols, int R) {
  if (nsplit <= 0) return skfr_nsplit(c->I, ncols, R);
  return (int)std::max<int64_t>(1, std::min<int64_t>(nsplit, ncols));
}

// k-means++ draw u of center cc in [0, 1)
double skfr_u(uint64_t seed, int cc) {
  return (double)(host_smix(seed ^ (0xD1B54A32D192ED03ULL * (uint64_t)(cc + 1))) >> 11) * 0x1.0p-53;
}

}  // namespace

extern "C" {

int admx_skfr_init(admx_ctx* c, int32_t k, uint64_t seed, int32_t nsplit, int32_t* clusters) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(need_resident(c, "SKFR"));
    if (!clusters) return bad("null output");
    if (k < 1 || k > kMaxK) return bad("need 1 <= k <= 16");
    if (k > c->I) return bad("more clusters than individuals");
    if (nsplit < 0 || nsplit > 64) return bad("need 0 <= nsplit <= 64");
    if (c->I >= (1LL << 31) || c->J >= (1LL << 31)) return bad("SKFR needs I, J < 2^31");
    SkfrView v;
    const int ns = skfr_split(c, nsplit, c->J, 1);
    CHK(skfr_prepare(c, k, 0, (int64_t)ns * c->I, true, &v));
    {
      ProfScope ps(c, "skfr_kmeanspp");
      for (int cc = 0; cc < k; ++cc) {
        launch_skfr_pick(v.D2, c->I, skfr_u(seed, cc), cc == 0, v.center, c->s);
        launch_skfr_pp_table(c->bed, c->pitch, v.center, v.mu, v.var, c->J, v.T, c->s);
        launch_skfr_tabsum(1, c->bed, c->pitch, c->I, nullptr, c->J, ns, v.T, v.P, c->s);
        launch_skfr_combine(v.P, ns, c->I, v.Dc + (int64_t)cc * c->I, v.D2, cc == 0, c->s);
      }
      launch_skfr_argmin(v.Dc, k, c->I, v.lab, c->s);
    }
    CHK(check_launch());
    HIPCHK(hipMemcpyAsync(clusters, v.lab, sizeof(int32_t) * c->I, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_skfr_run(admx_ctx* c, int32_t k, int32_t mode, int64_t sparsity, int32_t max_iter, int32_t nsplit,
                  int32_t* clusters, int64_t* aims, int64_t* n_aims, double* centers, double* wss, double* obj,
                  int32_t* iters) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(need_resident(c, "SKFR"));
    if (!clusters || !aims || !n_aims) return bad("null clusters or aims");
    if (k < 1 || k > kMaxK) return bad("need 1 <= k <= 16");
    if (mode != ADMX_SKFR_GLOBAL && mode != ADMX_SKFR_LOCAL) return bad("mode must be ADMX_SKFR_GLOBAL or _LOCAL");
    if (sparsity < 1 || sparsity > c->J) return bad("need 1 <= sparsity <= J");
    if (max_iter < 1) return bad("need max_iter >= 1");
    if (nsplit < 0 || nsplit > 64) return bad("need 0 <= nsplit <= 64");
    if (c->I >= (1LL << 31) || c->J >= (1LL << 31)) return bad("SKFR needs I, J < 2^31");
    for (int64_t i = 0; i < c->I; ++i)
      if (clusters[i] < 0 || clusters[i] >= k) return bad("cluster labels must lie in [0, k)");
    const bool local = mode == ADMX_SKFR_LOCAL;
    const int64_t I = c->I, J = c->J;
    const int64_t ncols_max = local ? std::min<int64_t>(J, (int64_t)k * sparsity) : sparsity;
    // slices are nondecreasing in the column count
    const int64_t p_elems = std::max<int64_t>((int64_t)skfr_split(c, nsplit, J, 1) * I,
                                              (int64_t)skfr_split(c, nsplit, ncols_max, k) * k * I);
    SkfrView v;
    CHK(skfr_prepare(c, k, ncols_max, p_elems, false, &v));
    HIPCHK(hipMemcpyAsync(v.lab, clusters, sizeof(int32_t) * I, hipMemcpyHostToDevice, c->s));
    {
      // sum_j x_ij^2 of every individual (the WSS term common to all clusters)
      ProfScope ps(c, "skfr_x2");
      const int ns = skfr_split(c, nsplit, J, 1);
      launch_skfr_pp_table(c->bed, c->pitch, nullptr, v.mu, v.var, J, v.T, c->s);
      launch_skfr_tabsum(1, c->bed, c->pitch, I, nullptr, J, ns, v.T, v.P, c->s);
      launch_skfr_combine(v.P, ns, I, v.x2, nullptr, 0, c->s);
    }
    CHK(check_launch());
    int it = 0;
    int64_t ncols = 0;
    for (it = 1; it <= max_iter; ++it) {
      HIPCHK(hipMemsetAsync(v.nk, 0, sizeof(unsigned) * kMaxK, c->s));
      {
        ProfScope ps(c, "skfr_masks");
        launch_skfr_masks(v.lab, I, k, v.nwords, v.cm, v.nk, c->s);
      }
      {
        ProfScope ps(c, "skfr_centers");
        launch_skfr_centers(k, c->bed, c->pitch, J, v.ntiles, v.tper, v.nslab, v.cm, v.part, c->s);
      }
      {
        ProfScope ps(c, "skfr_score");
        launch_skfr_score(k, local, v.part, v.nslab, v.nk, v.nm, v.alt, I, v.mu, v.var, J, v.E, v.key, v.idx, c->s);
      }
      {
        ProfScope ps(c, "skfr_select");
        if (!local) {
          launch_skfr_sort(v.tmp, v.tmp_bytes, v.key, v.keys, v.idx, v.sidx, J, c->s);
          ncols = sparsity;
        } else {
          for (int cc = 0; cc < k; ++cc)
            launch_skfr_sort(v.tmp, v.tmp_bytes, v.key + (int64_t)cc * J, v.keys + (int64_t)cc * J, v.idx,
                             v.sidx + (int64_t)cc * J, J, c->s);
          HIPCHK(hipMemsetAsync(v.flag, 0, (size_t)k * J, c->s));
          HIPCHK(hipMemsetAsync(v.count, 0, sizeof(unsigned long long), c->s));
          launch_skfr_flag(v.sidx, J, sparsity, k, v.flag, c->s);
          launch_skfr_union(v.flag, v.key, J, k, v.ukey, v.idx, v.count, c->s);
          launch_skfr_sort(v.tmp, v.tmp_bytes, v.ukey, v.keys, v.idx, v.sidx, J, c->s);
        }
      }
      CHK(check_launch());
      if (local) {
        unsigned long long h = 0;
        HIPCHK(hipMemcpyAsync(&h, v.count, sizeof(h), hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        ncols = (int64_t)h;
        if (ncols < 1 || ncols > ncols_max) return fail(ADMX_EHIP, "SKFR: inconsistent union of local selections");
      }
      const int ns = skfr_split(c, nsplit, ncols, k);
      HIPCHK(hipMemsetAsync(v.changes, 0, sizeof(unsigned long long), c->s));
      {
        ProfScope ps(c, "skfr_tables");
        launch_skfr_tables(v.sidx, ncols, k, local ? v.flag : nullptr, v.nk, v.mu, v.var, v.E, J, v.T, c->s);
      }
      {
        ProfScope ps(c, "skfr_tabsum");
        launch_skfr_tabsum(k, c->bed, c->pitch, I, v.sidx, ncols, ns, v.T, v.P, c->s);
      }
      {
        ProfScope ps(c, "skfr_assign");
        launch_skfr_assign(v.P, ns, k, I, v.nk, v.lab, v.dmin, v.changes, c->s);
      }
      CHK(check_launch());
      unsigned long long h = 0;
      HIPCHK(hipMemcpyAsync(&h, v.changes, sizeof(h), hipMemcpyDeviceToHost, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
      if (h == 0) break;
    }
    if (it > max_iter) it = max_iter;
    {
      ProfScope ps(c, "skfr_wss");
      launch_skfr_wss(v.lab, v.x2, v.dmin, I, k, v.out, c->s);
      if (centers) {
        if (!local) {
          HIPCHK(hipMemsetAsync(v.sel, 0, (size_t)J, c->s));
          launch_skfr_mark(v.sidx, ncols, v.sel, c->s);
        }
        // the distance tables are no longer needed: reuse them for the centers
        launch_skfr_center_out(v.sel, local ? v.flag : nullptr, v.nk, v.var, v.E, J, k, v.T, c->s);
      }
    }
    CHK(check_launch());
    std::vector<int32_t> h_aims((size_t)ncols);
    std::vector<double> h_out((size_t)k + 1);
    HIPCHK(hipMemcpyAsync(clusters, v.lab, sizeof(int32_t) * I, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(h_aims.data(), v.sidx, sizeof(int32_t) * ncols, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(h_out.data(), v.out, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, c->s));
    if (centers) HIPCHK(hipMemcpyAsync(centers, v.T, sizeof(double) * k * J, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    for (int64_t t = 0; t < ncols; ++t) aims[t] = h_aims[(size_t)t];
    *n_aims = ncols;
    if (wss)
      for (int cc = 0; cc < k; ++cc) wss[cc] = h_out[(size_t)cc];
    if (obj) *obj = h_out[(size_t)k];
    if (iters) *iters = it;
    return ADMX_OK;
  });
}

}  // extern "C"
