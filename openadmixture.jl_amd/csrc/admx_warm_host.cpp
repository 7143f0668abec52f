// admx_warm_host.cpp -- host side of the learned warm start: the FP32 network
// (admx_mlp_*: parameters, loss and gradient, prediction, training with Adam)
// and the genotype-projection features and network proposal of a context
// (admx_features, admx_warm_q).  The arithmetic runs in admx_mlp.hip.
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

// ---- warm-start network helpers
struct MlpActs {
  float *Xb, *Tb, *H1, *H2, *Z, *Y, *dZ, *dH2, *dH1, *rowloss, *loss;
};

int mlp_acts(admx_mlp* m, int64_t B, MlpActs* a) {
  const int64_t per = m->in + 4 * (int64_t)m->out + 4 * (int64_t)m->hid + 1;
  if (B > m->act_rows) {
    CHK(m->act.grow(sizeof(float) * (size_t)(per * B + 64)));
    m->act_rows = B;
  }
  float* p = static_cast<float*>(m->act.p);
  const int64_t R = m->act_rows;
  a->Xb = p;
  a->Tb = a->Xb + R * m->in;
  a->H1 = a->Tb + R * m->out;
  a->H2 = a->H1 + R * m->hid;
  a->Z = a->H2 + R * m->hid;
  a->Y = a->Z + R * m->out;
  a->dZ = a->Y + R * m->out;
  a->dH2 = a->dZ + R * m->out;
  a->dH1 = a->dH2 + R * m->hid;
  a->rowloss = a->dH1 + R * m->hid;
  a->loss = a->rowloss + R;
  return ADMX_OK;
}

// Forward over B rows of X into a->H1, H2, Z and Y (and dZ, rowloss when T).
void mlp_forward(admx_mlp* m, const float* X, const float* T, int B, const MlpActs& a, float* Y, hipStream_t s) {
  const float* w = static_cast<const float*>(m->w.p);
  launch_sgemm(false, true, B, m->hid, m->in, X, m->in, w + m->oW1(), m->in, a.H1, m->hid, w + m->ob1(), true,
               nullptr, false, s);
  launch_sgemm(false, true, B, m->hid, m->hid, a.H1, m->hid, w + m->oW2(), m->hid, a.H2, m->hid, w + m->ob2(), true,
               nullptr, false, s);
  launch_sgemm(false, true, B, m->out, m->hid, a.H2, m->hid, w + m->oW3(), m->hid, a.Z, m->out, w + m->ob3(), false,
               nullptr, false, s);
  launch_softmax_ce(a.Z, B, m->out, T, Y, T ? a.dZ : nullptr, T ? a.rowloss : nullptr, s);
}

// Gradients of the mean soft-target cross entropy into m->g; the loss sum of
// the batch into *loss_dev.
void mlp_backward(admx_mlp* m, const float* X, int B, const MlpActs& a, float* loss_dev, hipStream_t s) {
  const float* w = static_cast<const float*>(m->w.p);
  float* g = static_cast<float*>(m->g.p);
  launch_sgemm(true, false, m->out, m->hid, B, a.dZ, m->out, a.H2, m->hid, g + m->oW3(), m->hid, nullptr, false,
               nullptr, false, s);
  launch_colsum(a.dZ, B, m->out, g + m->ob3(), s);
  launch_sgemm(false, false, B, m->hid, m->out, a.dZ, m->out, w + m->oW3(), m->hid, a.dH2, m->hid, nullptr, false,
               a.H2, false, s);
  launch_sgemm(true, false, m->hid, m->hid, B, a.dH2, m->hid, a.H1, m->hid, g + m->oW2(), m->hid, nullptr, false,
               nullptr, false, s);
  launch_colsum(a.dH2, B, m->hid, g + m->ob2(), s);
  launch_sgemm(false, false, B, m->hid, m->hid, a.dH2, m->hid, w + m->oW2(), m->hid, a.dH1, m->hid, nullptr, false,
               a.H1, false, s);
  launch_sgemm(true, false, m->hid, m->in, B, a.dH1, m->hid, X, m->in, g + m->oW1(), m->in, nullptr, false, nullptr,
               false, s);
  launch_colsum(a.dH1, B, m->hid, g + m->ob1(), s);
  launch_colsum(a.rowloss, B, 1, loss_dev, s);
}

}  // namespace

extern "C" {

// ---- learned warm start
int admx_mlp_create(admx_mlp** out, int32_t in_dim, int32_t hidden, int32_t out_dim, uint64_t seed) {
  return guarded([&]() -> int {
    if (!out) return bad("null output handle");
    *out = nullptr;
    if (in_dim < 1 || hidden < 1 || out_dim < 2 || out_dim > 16) return bad("need in >= 1, hidden >= 1, 2 <= out <= 16");
    admx_mlp* m = new admx_mlp();
    m->in = in_dim;
    m->hid = hidden;
    m->out = out_dim;
    m->P = m->ob3() + out_dim;
    auto cleanup = [&](int rc) {
      admx_mlp_destroy(m);
      return rc;
    };
    hipError_t e = hipGetDevice(&m->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(fail(ADMX_EHIP, std::string("mlp stream: ") + hipGetErrorString(e)));
    const size_t pb = sizeof(float) * (size_t)m->P;
    int rc = m->w.grow(pb);
    if (rc == ADMX_OK) rc = m->g.grow(pb);
    if (rc == ADMX_OK) rc = m->m.grow(pb);
    if (rc == ADMX_OK) rc = m->v.grow(pb);
    if (rc != ADMX_OK) return cleanup(rc);
    // He-uniform weights U(-sqrt(6/fan_in), sqrt(6/fan_in)), zero biases
    std::vector<float> w((size_t)m->P, 0.f);
    uint64_t ctr = seed * 0x9E3779B97F4A7C15ULL;
    auto fill = [&](int64_t off, int64_t n, int fan_in) {
      const double a = std::sqrt(6.0 / fan_in);
      for (int64_t e2 = 0; e2 < n; ++e2) {
        const double u = (double)(host_smix(ctr++) >> 11) * 0x1.0p-53;
        w[(size_t)(off + e2)] = (float)((2.0 * u - 1.0) * a);
      }
    };
    fill(m->oW1(), (int64_t)hidden * in_dim, in_dim);
    fill(m->oW2(), (int64_t)hidden * hidden, hidden);
    fill(m->oW3(), (int64_t)out_dim * hidden, hidden);
    e = hipMemcpy(m->w.p, w.data(), pb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->m.p, 0, pb);
    if (e == hipSuccess) e = hipMemset(m->v.p, 0, pb);
    if (e != hipSuccess) return cleanup(fail(ADMX_EHIP, std::string("mlp init: ") + hipGetErrorString(e)));
    *out = m;
    return ADMX_OK;
  });
}

int admx_mlp_destroy(admx_mlp* m) {
  return guarded([&]() -> int {
    if (!m) return ADMX_OK;
    (void)hipSetDevice(m->device);
    if (m->s) (void)hipStreamSynchronize(m->s);
    DevBuf* bufs[] = {&m->w, &m->g, &m->m, &m->v, &m->X, &m->T, &m->idx, &m->act};
    for (DevBuf* b : bufs) b->release();
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
    return ADMX_OK;
  });
}

int admx_mlp_num_params(admx_mlp* m, int64_t* n) {
  return guarded([&]() -> int {
    if (!m || !n) return bad("null argument");
    *n = m->P;
    return ADMX_OK;
  });
}

int admx_mlp_get_params(admx_mlp* m, float* w) {
  return guarded([&]() -> int {
    if (!m || !w) return bad("null argument");
    MGuard g(m);
    HIPCHK(hipSetDevice(m->device));
    HIPCHK(hipMemcpyAsync(w, m->w.p, sizeof(float) * (size_t)m->P, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    return ADMX_OK;
  });
}

int admx_mlp_set_params(admx_mlp* m, const float* w) {
  return guarded([&]() -> int {
    if (!m || !w) return bad("null argument");
    MGuard g(m);
    HIPCHK(hipSetDevice(m->device));
    HIPCHK(hipMemcpyAsync(m->w.p, w, sizeof(float) * (size_t)m->P, hipMemcpyHostToDevice, m->s));
    HIPCHK(hipMemsetAsync(m->m.p, 0, sizeof(float) * (size_t)m->P, m->s));
    HIPCHK(hipMemsetAsync(m->v.p, 0, sizeof(float) * (size_t)m->P, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    m->step = 0;
    return ADMX_OK;
  });
}

int admx_mlp_loss_grad(admx_mlp* m, const float* X, const float* T, int64_t B, float* loss, float* grad) {
  return guarded([&]() -> int {
    if (!m || !X || !T || !loss || !grad) return bad("null argument");
    if (B < 1 || B > (1 << 30)) return bad("batch rows out of range");
    MGuard g(m);
    HIPCHK(hipSetDevice(m->device));
    MlpActs a;
    CHK(mlp_acts(m, B, &a));
    HIPCHK(hipMemcpyAsync(a.Xb, X, sizeof(float) * (size_t)(B * m->in), hipMemcpyHostToDevice, m->s));
    HIPCHK(hipMemcpyAsync(a.Tb, T, sizeof(float) * (size_t)(B * m->out), hipMemcpyHostToDevice, m->s));
    mlp_forward(m, a.Xb, a.Tb, (int)B, a, a.Y, m->s);
    mlp_backward(m, a.Xb, (int)B, a, a.loss, m->s);
    CHK(check_launch());
    float ls = 0.f;
    HIPCHK(hipMemcpyAsync(&ls, a.loss, sizeof(float), hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipMemcpyAsync(grad, m->g.p, sizeof(float) * (size_t)m->P, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    *loss = ls / (float)B;
    return ADMX_OK;
  });
}

int admx_mlp_predict(admx_mlp* m, const float* X, int64_t n, float* Y) {
  return guarded([&]() -> int {
    if (!m || !X || !Y) return bad("null argument");
    if (n < 0) return bad("negative row count");
    MGuard g(m);
    HIPCHK(hipSetDevice(m->device));
    const int64_t Bmax = 65536;
    MlpActs a;
    CHK(mlp_acts(m, std::min<int64_t>(std::max<int64_t>(n, 1), Bmax), &a));
    for (int64_t r0 = 0; r0 < n; r0 += Bmax) {
      const int64_t B = std::min(Bmax, n - r0);
      HIPCHK(hipMemcpyAsync(a.Xb, X + r0 * m->in, sizeof(float) * (size_t)(B * m->in), hipMemcpyHostToDevice, m->s));
      mlp_forward(m, a.Xb, nullptr, (int)B, a, a.Y, m->s);
      CHK(check_launch());
      HIPCHK(hipMemcpyAsync(Y + r0 * m->out, a.Y, sizeof(float) * (size_t)(B * m->out), hipMemcpyDeviceToHost, m->s));
    }
    HIPCHK(hipStreamSynchronize(m->s));
    return ADMX_OK;
  });
}

int admx_mlp_train(admx_mlp* m, const float* X, const float* T, int64_t n, int32_t epochs, int32_t batch, float lr,
                   uint64_t seed, float* epoch_loss) {
  return guarded([&]() -> int {
    if (!m || !X || !T) return bad("null argument");
    if (n < 1 || n > (1 << 30) || epochs < 0 || batch < 1) return bad("need n >= 1, epochs >= 0, batch >= 1");
    MGuard g(m);
    HIPCHK(hipSetDevice(m->device));
    CHK(m->X.grow(sizeof(float) * (size_t)(n * m->in)));
    CHK(m->T.grow(sizeof(float) * (size_t)(n * m->out)));
    CHK(m->idx.grow(sizeof(int32_t) * (size_t)n));
    HIPCHK(hipMemcpyAsync(m->X.p, X, sizeof(float) * (size_t)(n * m->in), hipMemcpyHostToDevice, m->s));
    HIPCHK(hipMemcpyAsync(m->T.p, T, sizeof(float) * (size_t)(n * m->out), hipMemcpyHostToDevice, m->s));
    m->n = n;
    const int64_t B = std::min<int64_t>(batch, n);
    const int64_t nb = (n + B - 1) / B;
    MlpActs a;
    CHK(mlp_acts(m, B, &a));
    DevBuf lossbuf;
    CHK(lossbuf.grow(sizeof(float) * (size_t)nb));
    std::vector<int32_t> perm((size_t)n);
    std::vector<float> lh((size_t)nb);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int rc = ADMX_OK;
    for (int ep = 0; ep < epochs && rc == ADMX_OK; ++ep) {
      // Fisher-Yates shuffle of the rows from the seeded counter hash
      for (int64_t r = 0; r < n; ++r) perm[(size_t)r] = (int32_t)r;
      for (int64_t r = n - 1; r > 0; --r) {
        const uint64_t h = host_smix(seed ^ host_smix((uint64_t)ep * 0x100000001B3ULL + (uint64_t)r));
        const int64_t k = (int64_t)(h % (uint64_t)(r + 1));
        std::swap(perm[(size_t)r], perm[(size_t)k]);
      }
      HIPCHK(hipMemcpyAsync(m->idx.p, perm.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, m->s));
      for (int64_t b = 0; b < nb; ++b) {
        const int64_t r0 = b * B;
        const int Bb = (int)std::min(B, n - r0);
        const int32_t* ix = static_cast<const int32_t*>(m->idx.p) + r0;
        launch_gather_rows(static_cast<const float*>(m->X.p), m->in, ix, Bb, a.Xb, m->s);
        launch_gather_rows(static_cast<const float*>(m->T.p), m->out, ix, Bb, a.Tb, m->s);
        mlp_forward(m, a.Xb, a.Tb, Bb, a, a.Y, m->s);
        mlp_backward(m, a.Xb, Bb, a, static_cast<float*>(lossbuf.p) + b, m->s);
        ++m->step;
        const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
        launch_adam(m->P, static_cast<float*>(m->w.p), static_cast<const float*>(m->g.p), static_cast<float*>(m->m.p),
                    static_cast<float*>(m->v.p), lr, b1, b2, c1, c2, eps, m->s);
      }
      rc = check_launch();
      if (rc != ADMX_OK) break;
      HIPCHK(hipMemcpyAsync(lh.data(), lossbuf.p, sizeof(float) * (size_t)nb, hipMemcpyDeviceToHost, m->s));
      HIPCHK(hipStreamSynchronize(m->s));
      double tot = 0.0;
      for (float v : lh) tot += v;
      if (epoch_loss) epoch_loss[ep] = (float)(tot / (double)n);
    }
    HIPCHK(hipStreamSynchronize(m->s));
    lossbuf.release();
    return rc;
  });
}

int admx_features(admx_ctx* c, const float* R, const float* mu, const float* isd, int64_t Jp, float* F) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(need_resident(c, "the warm-start features"));
    if (!R || !mu || !isd) return bad("null panel array");
    if (Jp < 1 || Jp > c->J) return bad("panel SNP count must be in [1, J]");
    const size_t rb = sizeof(float) * (size_t)(Jp * kFeat), vb = sizeof(float) * (size_t)Jp;
    CHK(stage_in(c, 0, R, rb, rb));
    CHK(stage_in(c, 1, mu, vb, vb));
    CHK(stage_in(c, 2, isd, vb, vb));
    const int64_t waves = (c->I + 63) / 64;
    int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(64, c->target_waves / std::max<int64_t>(waves, 1) / 8));
    nchunk = std::min(nchunk, Jp);
    const int64_t jchunk = (Jp + nchunk - 1) / nchunk;
    nchunk = (Jp + jchunk - 1) / jchunk;
    CHK(c->feat.grow(sizeof(float) * (size_t)(c->I * kFeat)));
    CHK(c->featpart.grow(sizeof(float) * (size_t)(nchunk * c->I * kFeat)));
    {
      ProfScope ps(c, "features");
      launch_features(c->bed, c->pitch, c->I, Jp, static_cast<const float*>(c->st[0].p),
                      static_cast<const float*>(c->st[1].p), static_cast<const float*>(c->st[2].p), (int)nchunk, jchunk,
                      static_cast<float*>(c->featpart.p), static_cast<float*>(c->feat.p), c->s);
    }
    CHK(check_launch());
    c->feat_rows = c->I;
    if (F) HIPCHK(hipMemcpyAsync(F, c->feat.p, sizeof(float) * (size_t)(c->I * kFeat), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_warm_q(admx_ctx* c, admx_mlp* m) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    if (!m) return bad("null network");
    Guard g(c);
    MGuard gm(m);
    if (c->dtype != ADMX_F64) return bad("the warm start writes a FP64 context");
    if (m->device != c->device) return bad("network and context must live on the same device");
    if (m->in != kFeat || m->out != c->K) return bad("network input must be 64 features and its output K");
    if (c->feat_rows != c->I) return fail(ADMX_ESTATE, "no features: call admx_features first");
    HIPCHK(hipStreamSynchronize(m->s));
    const int64_t Bmax = 65536;
    MlpActs a;
    CHK(mlp_acts(m, std::min<int64_t>(c->I, Bmax), &a));
    CHK(c->st[3].grow(sizeof(float) * (size_t)(c->I * c->K)));
    float* Y = static_cast<float*>(c->st[3].p);
    const float* Fd = static_cast<const float*>(c->feat.p);
    {
      ProfScope ps(c, "warm_q");
      for (int64_t r0 = 0; r0 < c->I; r0 += Bmax) {
        const int64_t B = std::min(Bmax, c->I - r0);
        mlp_forward(m, Fd + r0 * kFeat, nullptr, (int)B, a, Y + r0 * c->K, c->s);
      }
      double* q = static_cast<double*>(c->x[ADMX_X]);
      launch_probs_to_q(Y, c->K * c->I, q, c->s);
      launch_project_q<double>((int)c->K, c->I, q, c->s);
    }
    CHK(check_launch());
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_warm_p(admx_ctx* c, double* ll_hw) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (c->dtype != ADMX_F64) return bad("the warm start writes a FP64 context");
    CHK(need_bed(c));
    CHK(need_resident(c, "the warm start's allele frequencies"));
    const size_t jb = sizeof(unsigned long long) * (size_t)c->J;
    const int64_t nblk = (c->J + 255) / 256;
    CHK(c->st[0].grow(jb));
    CHK(c->st[1].grow(jb));
    CHK(c->st[2].grow(sizeof(double) * (size_t)(nblk + 1)));
    auto* nm = static_cast<unsigned long long*>(c->st[0].p);
    auto* alt = static_cast<unsigned long long*>(c->st[1].p);
    double* part = static_cast<double*>(c->st[2].p);
    {
      ProfScope ps(c, "warm_p");
      launch_counts_snp(c->bed, c->pitch, c->I, c->J, nm, alt, c->s);
      launch_freq_to_p(nm, alt, c->J, (int)c->K, static_cast<double*>(c->pp(c->x[ADMX_X])), ll_hw ? part : nullptr,
                       c->s);
      if (ll_hw) launch_reduce_sum(part, nblk, part + nblk, 0, c->s);
    }
    CHK(check_launch());
    if (ll_hw) HIPCHK(hipMemcpyAsync(ll_hw, part + nblk, sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

}  // extern "C"
