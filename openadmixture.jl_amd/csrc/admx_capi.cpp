// admx_capi.cpp -- host runtime of libadmixhip: the context (device mirror of
// AdmixData{T}), the C ABI declared in include/admixhip.h, and the
// device-resident orchestration of update_q!/update_p!/init_em!/admixture_qn!
// on one HIP stream.  Only plumbing lives here; all arithmetic on genotypes,
// QPs, projections and quasi-Newton steps runs in the HIP kernels.
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "admixhip.h"
#include "admx_host.hpp"
#include "admx_internal.hpp"

using namespace admx;
using namespace admx_host;

thread_local std::string admx_host::g_err;

namespace {

// ---- chunking of the genotype passes
int qside_chunks(admx_ctx* c, int64_t Ilen, int64_t Jlen, int64_t* chunk, int rows_per_wave = 64) {
  const int64_t waves = (Ilen + rows_per_wave - 1) / rows_per_wave;
  int64_t n = (c->target_waves + waves - 1) / std::max<int64_t>(waves, 1);
  n = std::max<int64_t>(1, std::min<int64_t>(n, std::max<int64_t>(1, Jlen / 64)));
  *chunk = std::max<int64_t>(1, (Jlen + n - 1) / n);
  return (int)std::max<int64_t>(1, (Jlen + *chunk - 1) / *chunk);
}

// P-side passes stream individuals in tiles of 256 from a 256-aligned start
// (the first tile may begin before i0; its leading individuals are masked).
int pside_chunks(admx_ctx* c, int64_t i0, int64_t i1, int64_t Jlen, int64_t* chunk, int rows_per_wave = 64) {
  const int64_t waves = (Jlen + rows_per_wave - 1) / rows_per_wave;
  const int64_t span = i1 - (i0 & ~(int64_t)255);
  int64_t n = (c->target_waves + waves - 1) / std::max<int64_t>(waves, 1);
  n = std::max<int64_t>(1, std::min<int64_t>(n, std::max<int64_t>(1, span / 256)));
  int64_t ch = (span + n - 1) / n;
  ch = (ch + 255) & ~(int64_t)255;
  *chunk = std::max<int64_t>(256, ch);
  return (int)std::max<int64_t>(1, (span + *chunk - 1) / *chunk);
}

const char* mode_name(bool qside, int mode) {
  if (qside) return mode == MODE_UPDATE ? "qside_update" : (mode == MODE_EM ? "qside_em" : "qside_ll");
  return mode == MODE_UPDATE ? "pside_update" : "pside_em";
}

// Rows [r0, r1) (r1 <= J) of a file-backed BED into `dst`, ibytes per row, as
// they lie in the file.  pread from the page cache or the disk, split over up
// to 8 host threads of at least 4 MiB each.
int read_file_rows(admx_ctx* c, int64_t r0, int64_t r1, uint8_t* dst) {
  const int64_t n = (r1 - r0) * c->ibytes;
  if (n <= 0) return ADMX_OK;
  const int64_t base = 3 + r0 * c->ibytes;
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(8, n >> 22));
  std::vector<int> err((size_t)nt, 0);
  auto work = [&](int t) {
    int64_t a = n * t / nt;
    const int64_t e = n * (t + 1) / nt;
    while (a < e) {
      const ssize_t got = pread(c->bed_fd, dst + a, (size_t)(e - a), (off_t)(base + a));
      if (got < 0 && errno == EINTR) continue;
      if (got <= 0) {
        err[(size_t)t] = got < 0 ? errno : EIO;
        return;
      }
      a += got;
    }
  };
  std::vector<std::thread> th;
  int started = 1;  // slice 0 runs on the calling thread
  try {
    for (; started < nt; ++started) th.emplace_back(work, started);
  } catch (...) {  // no more threads: the calling thread reads the rest
  }
  work(0);
  for (int t = started; t < nt; ++t) work(t);
  for (auto& t : th) t.join();
  for (int e : err)
    if (e) return fail(ADMX_EIO, std::string("reading the .bed file: ") + std::strerror(e));
  return ADMX_OK;
}

// Enqueue on the copy stream: rows [r0, r1) (r1 <= J + 16) of a streamed BED
// into `dst` (pitch per row), using chunk slot b's buffers.  Pinned host BED:
// one copy.  File-backed: the host reads the rows into slot b's pinned staging
// buffer (after its previous copy has finished), copies them to HBM and
// re-pitches them there (rows past J become missing codes).
int stage_rows(admx_ctx* c, int b, int64_t r0, int64_t r1, uint8_t* dst) {
  if (c->hbed) {
    HIPCHK(hipMemcpyAsync(dst, c->hbed + c->pitch * r0, (size_t)(c->pitch * (r1 - r0)), hipMemcpyHostToDevice, c->cs));
    return ADMX_OK;
  }
  if (c->ev_used[b]) HIPCHK(hipEventSynchronize(c->ev_ready[b]));
  const int64_t real = std::max<int64_t>(0, std::min(r1, c->J) - r0);
  CHK(read_file_rows(c, r0, r0 + real, c->hstage[b]));
  if (real) HIPCHK(hipMemcpyAsync(c->draw[b], c->hstage[b], (size_t)(c->ibytes * real), hipMemcpyHostToDevice, c->cs));
  launch_repitch(c->draw[b], c->ibytes, real, r1 - r0, dst, c->pitch, c->cs);
  return check_launch();
}

// SNP chunks of a pass over [j0, j1): one for a resident context; for a
// streamed one, the part below res_snps, then chunks of stream_snps SNPs.
// for_chunks calls its function once per chunk, in this order.
std::vector<std::pair<int64_t, int64_t>> snp_chunks(const admx_ctx* c, int64_t j0, int64_t j1) {
  std::vector<std::pair<int64_t, int64_t>> r;
  if (!c->stream_snps) {
    if (j0 < j1) r.push_back({j0, j1});
    return r;
  }
  if (j0 < std::min(j1, c->res_snps)) {
    r.push_back({j0, std::min(j1, c->res_snps)});
    j0 = std::min(j1, c->res_snps);
  }
  for (int64_t a = j0; a < j1; a += c->stream_snps) r.push_back({a, std::min(j1, a + c->stream_snps)});
  return r;
}

// Genotype source of a pass over SNPs [j0, j1): fn(base, cj0, cj1, k) launches
// on c->s for chunk k of snp_chunks, with bed rows reachable as base + pitch * j
// for j in [cj0, cj1 + 16).  Resident context: one call with the HBM BED.
// Streamed context: the HBM-resident chunk first, then the streamed chunks,
// alternately forwards and backwards from one pass to the next, through two
// device buffers: the copy of the next chunk (copy stream) overlaps the kernels
// of the current one, and a buffer is refilled only after the kernels that
// read it have finished (events).  A chunk whose rows a buffer already holds
// (the previous pass's last chunk, by the alternation) is not copied again.
// The base pointer of a chunk is the buffer address minus pitch * (its first row).
template <typename F>
int for_chunks(admx_ctx* c, int64_t j0, int64_t j1, F&& fn) {
  if (!c->stream_snps) return fn(static_cast<const uint8_t*>(c->bed), j0, j1, (size_t)0);
  const auto chunks = snp_chunks(c, j0, j1);
  std::vector<size_t> order;
  size_t first = 0;
  if (!chunks.empty() && chunks[0].second <= c->res_snps) order.push_back(first++);
  c->stream_rev = !c->stream_rev;
  for (size_t n = first; n < chunks.size(); ++n) order.push_back(c->stream_rev ? chunks.size() - 1 - (n - first) : n);
  for (size_t k : order) {
    const int64_t cj0 = chunks[k].first, cj1 = chunks[k].second;
    if (cj1 <= c->res_snps) {  // HBM-resident rows
      CHK(fn(static_cast<const uint8_t*>(c->dres), cj0, cj1, k));
      continue;
    }
    const int64_t r1 = std::min(cj1 + 16, c->J + 16);
    int b = -1;
    for (int bb = 0; bb < 2; ++bb)
      if (c->held[bb][0] <= cj0 && r1 <= c->held[bb][1]) b = bb;
    if (b < 0) {
      b = c->last_buf ^ 1;
      c->held[b][0] = c->held[b][1] = 0;
      HIPCHK(hipStreamWaitEvent(c->cs, c->ev_free[b], 0));
      CHK(stage_rows(c, b, cj0, r1, c->dchunk[b]));
      HIPCHK(hipEventRecord(c->ev_ready[b], c->cs));
      c->ev_used[b] = true;
      c->held[b][0] = cj0;
      c->held[b][1] = r1;
    }
    HIPCHK(hipStreamWaitEvent(c->s, c->ev_ready[b], 0));
    CHK(fn(static_cast<const uint8_t*>(c->dchunk[b]) - c->pitch * c->held[b][0], cj0, cj1, k));
    HIPCHK(hipEventRecord(c->ev_free[b], c->s));
    c->last_buf = b;
  }
  return ADMX_OK;
}

// The chunk buffers no longer hold current genotype rows.
void forget_chunks(admx_ctx* c) {
  for (auto& h : c->held) h[0] = h[1] = 0;
}

// Copy SNPs [0, res_snps + 16) of a streamed BED into the HBM-resident rows
// (after the genotypes change or the resident count is set); synchronous.
int refresh_resident(admx_ctx* c) {
  if (!c->stream_snps || !c->res_snps) return ADMX_OK;
  const int64_t R = std::min(c->res_snps + 16, c->J + 16), S = c->stream_snps;
  HIPCHK(hipStreamSynchronize(c->s));
  for (int64_t r0 = 0; r0 < R; r0 += S) {
    CHK(stage_rows(c, 0, r0, std::min(R, r0 + S), c->dres + c->pitch * r0));
    HIPCHK(hipEventRecord(c->ev_ready[0], c->cs));
    c->ev_used[0] = true;
  }
  HIPCHK(hipStreamSynchronize(c->cs));
  return ADMX_OK;
}

// Per-SNP means of the current genotypes for the mean-imputation loops
// (computed once per genotype upload).
template <typename T>
int ensure_mu(admx_ctx* c) {
  if (!c->impute || c->mu_valid) return ADMX_OK;
  CHK(need_resident(c, "the mean-imputation loops"));
  CHK(c->snp_mu.grow(sizeof(double) * (size_t)c->J));
  CHK(c->cnt.grow(sizeof(unsigned long long) * (size_t)(2 * c->J + c->I)));
  auto* nm = static_cast<unsigned long long*>(c->cnt.p);
  {
    ProfScope ps(c, "snp_mean");
    launch_counts(c->bed, c->pitch, c->I, c->J, nm, nm + c->J, nm + 2 * c->J, c->s);
    launch_snp_mean<double>(nm, nm + c->J, c->J, static_cast<double*>(c->snp_mu.p), c->s);
  }
  CHK(check_launch());
  c->mu_valid = true;
  return ADMX_OK;
}

// The genotype passes run in FP64 in every context: on gfx950 FP32 and FP64
// FMAs issue at the same rate, and the FP64 rank-1 Hessian form needs fewer
// instructions than the two-term form FP32 would need.  An FP32 context hands
// the pass FP64 copies of its q and p (scratch slot w, n values); an FP64
// context's arrays pass through.
template <typename T>
int widen(admx_ctx* c, int w, const void* x, int64_t n, const void** out) {
  if constexpr (sizeof(T) == 8) {
    *out = x;
    return ADMX_OK;
  } else {
    if (!x) {
      *out = nullptr;
      return ADMX_OK;
    }
    CHK(c->wide[w].grow(sizeof(double) * (size_t)n));
    launch_widen(static_cast<const float*>(x), n, static_cast<double*>(c->wide[w].p), c->s);
    *out = c->wide[w].p;
    return check_launch();
  }
}

// Run a Q-side pass; returns the chunk count (partials in c->part or c->llpart).
// MODE_LL with a second state (q2, p2): both log-likelihoods from one pass over
// the genotypes (the second state's partials follow the first's *nparts).
template <typename T>
int pass_q(admx_ctx* c, int mode, const void* q, const void* p, int64_t i0, int64_t i1, int64_t j0, int64_t j1,
           int* nchunk, int64_t* nparts, const void* q2 = nullptr, const void* p2 = nullptr) {
  const int64_t Ilen = i1 - i0, Jlen = j1 - j0;
  *nchunk = 0;
  *nparts = 0;
  if (Ilen <= 0 || Jlen <= 0) return ADMX_OK;
  CHK(ensure_mu<T>(c));
  CHK(widen<T>(c, 0, q, c->K * c->I, &q));
  CHK(widen<T>(c, 1, p, c->K * c->J, &p));
  CHK(widen<T>(c, 2, q2, c->K * c->I, &q2));
  CHK(widen<T>(c, 3, p2, c->K * c->J, &p2));
  // one launch per SNP chunk (one for a resident BED); their partial slots sit side by side
  const auto chunks = snp_chunks(c, j0, j1);
  std::vector<int> ns;
  std::vector<int64_t> chs;
  int total = 0;
  for (const auto& r : chunks) {
    int64_t ch = 0;
    ns.push_back(qside_chunks(c, Ilen, r.second - r.first, &ch));
    chs.push_back(ch);
    total += ns.back();
  }
  const int64_t nbx = (Ilen + 255) / 256;
  const int na = pass_na(mode, true, (int)c->K, true);
  // log-likelihood: the term-by-term form for states where some s may be 0 or 1
  int* flags = nullptr;
  if (mode == MODE_LL && !c->impute) {
    CHK(c->llflag.grow(2 * sizeof(int)));
    flags = static_cast<int*>(c->llflag.p);
    const bool f32 = c->dtype == ADMX_F32;  // widened FP32 state
    launch_ll_state_check((int)c->K, static_cast<const double*>(q), c->I, static_cast<const double*>(p), c->J,
                          flags, c->s, f32);
    if (q2)
      launch_ll_state_check((int)c->K, static_cast<const double*>(q2), c->I, static_cast<const double*>(p2), c->J,
                            flags + 1, c->s, f32);
    CHK(check_launch());
  }
  if (mode == MODE_LL) CHK(c->llpart.grow((size_t)(nbx * total) * (q2 ? 2 : 1) * sizeof(double)));
  else CHK(c->part.grow((size_t)total * na * Ilen * sizeof(double)));
  // chunk k's partial slots start after those of chunks 0..k-1, whatever order
  // the chunks run in
  std::vector<int64_t> first(chunks.size(), 0);
  for (size_t k = 1; k < chunks.size(); ++k) first[k] = first[k - 1] + ns[k - 1];
  {
    ProfScope ps(c, mode_name(true, mode));
    CHK(for_chunks(c, j0, j1, [&](const uint8_t* base, int64_t cj0, int64_t cj1, size_t k) {
      PassArgs a{base, c->pitch, c->I, q, p, i0, i1, cj0, cj1, chs[k], nullptr, c->impute ? c->snp_mu.p : nullptr};
      a.split = c->split_q;
      if (mode == MODE_LL) {
        a.part = static_cast<double*>(c->llpart.p) + nbx * first[k];
        a.strict = flags;
        if (q2 && flags) {  // both states (one fused pass over the genotypes up to K = 8)
          a.q2 = q2;
          a.p2 = p2;
          a.part2 = static_cast<double*>(c->llpart.p) + nbx * (total + first[k]);
          a.strict2 = flags + 1;
          launch_qside_ll2((int)c->K, a, ns[k], c->s);
          return check_launch();
        }
        if (q2) {
          PassArgs a2 = a;
          a2.q = q2;
          a2.p = p2;
          a2.strict = flags ? flags + 1 : nullptr;
          a2.part = static_cast<double*>(c->llpart.p) + nbx * (total + first[k]);
          launch_qside<double>(mode, (int)c->K, a2, ns[k], c->s);
        }
      } else {
        a.part = static_cast<double*>(c->part.p) + first[k] * na * Ilen;
      }
      launch_qside<double>(mode, (int)c->K, a, ns[k], c->s);
      return check_launch();
    }));
  }
  *nchunk = total;
  if (mode == MODE_LL) *nparts = nbx * total;
  return ADMX_OK;
}

template <typename T>
int pass_p(admx_ctx* c, int mode, const void* q, const void* p, int64_t i0, int64_t i1, int64_t j0, int64_t j1,
           int* nchunk) {
  const int64_t Ilen = i1 - i0, Jlen = j1 - j0;
  *nchunk = 0;
  if (Ilen <= 0 || Jlen <= 0) return ADMX_OK;
  int64_t chunk = 0;
  const int n = pside_chunks(c, i0, i1, Jlen, &chunk);
  const int na = pass_na(mode, false, (int)c->K, true);
  CHK(c->part.grow((size_t)n * na * Jlen * sizeof(double)));
  CHK(ensure_mu<T>(c));
  CHK(widen<T>(c, 0, q, c->K * c->I, &q));
  CHK(widen<T>(c, 1, p, c->K * c->J, &p));
  {
    ProfScope ps(c, mode_name(false, mode));
    CHK(for_chunks(c, j0, j1, [&](const uint8_t* base, int64_t cj0, int64_t cj1, size_t) {
      PassArgs a{base, c->pitch, c->I, q, p, i0, i1, cj0, cj1, chunk, c->part.p, c->impute ? c->snp_mu.p : nullptr,
                 Jlen, cj0 - j0};
      a.split = c->split_p;
      launch_pside<double>(mode, (int)c->K, a, n, c->s);
      return check_launch();
    }));
  }
  *nchunk = n;
  return ADMX_OK;
}

// BF16 form: the log-likelihood of the whole cohort with s and t from the
// matrix cores (split tables of q and p built per call).
int ll_mixed(admx_ctx* c, const double* q, const double* p, int slot) {
  const int64_t I = c->I, J = c->J, K = c->K;
  const int KD = split_table_cols((int)K);
  const int64_t rows_q = (I + 255) / 256 * 256, rows_p = J + 32;
  CHK(c->lqt.grow(sizeof(uint16_t) * KD * rows_q));
  CHK(c->lpt.grow(sizeof(uint16_t) * KD * rows_p));
  CHK(c->lmt.grow(sizeof(uint16_t) * KD * rows_p));
  const int64_t nbx = (I + 255) / 256;
  int64_t nchunk = std::max<int64_t>(1, (c->target_waves / 4 + nbx - 1) / nbx);
  nchunk = std::min<int64_t>(nchunk, std::max<int64_t>(1, J / 256));
  const int64_t jchunk = ((J + nchunk - 1) / nchunk + 31) / 32 * 32;
  nchunk = (J + jchunk - 1) / jchunk;
  // A state that may give s = q.p of exactly 0 or 1 (FRAPPE EM on a monomorphic
  // SNP) gets the reference's NaN / -Inf terms from the FP64 term-by-term pass
  // instead: both passes are launched, the device flag picks the one that works
  // (no host round trip, so captured graphs stay valid).
  int64_t sch = 0;
  const int sn = qside_chunks(c, I, J, &sch);
  const int64_t nmf = nbx * nchunk, nst = nbx * sn;
  CHK(c->llpart.grow(sizeof(double) * (nmf + nst)));
  CHK(c->llflag.grow(2 * sizeof(int)));
  int* flag = static_cast<int*>(c->llflag.p);
  double* parts = static_cast<double*>(c->llpart.p);
  {
    ProfScope ps(c, "qside_ll_mfma");
    launch_ll_state_check((int)K, q, I, p, J, flag, c->s);
    launch_split_tables((int)K, q, I, rows_q, 0, c->lqt.p, nullptr, c->s);
    launch_split_tables((int)K, p, J, rows_p, 1, c->lpt.p, c->lmt.p, c->s);
    launch_ll_mfma((int)K, c->bed, c->pitch, I, J, jchunk, (int)nchunk, q, c->lqt.p, c->lpt.p, c->lmt.p, parts,
                   flag, c->s);
    PassArgs a{c->bed, c->pitch, I, q, p, 0, I, 0, J, sch, parts + nmf};
    a.strict = flag;
    launch_qside_ll_strict((int)K, a, sn, c->s);
  }
  CHK(check_launch());
  ProfScope ps(c, "ll_reduce");
  launch_reduce_sum_if(parts, nmf, c->lls + slot, flag, 0, c->s);
  launch_reduce_sum_if(parts + nmf, nst, c->lls + slot, flag, 1, c->s);
  return check_launch();
}

// BF16 form: the FRAPPE EM sums of the state x on the matrix cores, into the
// FP64 partial layout of the FP64 EM passes (chunk count in *nchunk).
bool em_on_matrix_cores(const admx_ctx* c) {
  return c->dtype == ADMX_F64 && c->precision == ADMX_PREC_BF16 && !c->impute && !c->stream_snps;
}

// The matrix-core EM clamps s (1 - s) at 1e-30, so on a state where some
// s = q.p is exactly 0 or 1 it adds 0 where the reference's 0/0 gives NaN.
// Such states (flagged by the log-likelihood state check) take the FP64 EM
// pass; the check reads its flag back, which is fine outside captured graphs
// (EM steps are never captured).
int em_state_is_safe(admx_ctx* c, const double* q, const double* p, bool* safe) {
  CHK(c->llflag.grow(2 * sizeof(int)));
  int* flag = static_cast<int*>(c->llflag.p);
  launch_ll_state_check((int)c->K, q, c->I, p, c->J, flag, c->s);
  CHK(check_launch());
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  *safe = h == 0;
  return ADMX_OK;
}

int em_mixed(admx_ctx* c, bool qside, const double* q, const double* p, int* nchunk_out) {
  const int64_t I = c->I, J = c->J, K = c->K;
  const int KD = split_table_cols((int)K);
  const int64_t rows_q = (I + 255) / 256 * 256, rows_p = J + 32;
  CHK(c->lqt.grow(sizeof(uint16_t) * KD * rows_q));
  CHK(c->lpt.grow(sizeof(uint16_t) * KD * rows_p));
  CHK(c->lmt.grow(sizeof(uint16_t) * KD * rows_p));
  const int64_t ld = (qside ? (J + 31) / 32 * 32 : rows_q) + 32;  // split rows of p (em_q) or q (em_p)
  CHK(c->lps.grow(sizeof(uint16_t) * (qside ? 64 : 32) * ld));
  if (!qside) CHK(c->lqc.grow(sizeof(float) * I));
  int64_t nchunk = 0;
  {
    ProfScope ps(c, qside ? "qside_em_mfma" : "pside_em_mfma");
    launch_split_tables((int)K, q, I, rows_q, 0, c->lqt.p, nullptr, c->s);
    if (!qside || c->qside_mfma == 0) launch_split_tables((int)K, p, J, rows_p, 1, c->lpt.p, c->lmt.p, c->s);
    if (qside && c->qside_mfma != 0) {
      // the matrix-core update pass's design (block images of p and 1 - p,
      // 16x16x32 sums): 8-13% faster than k_emq_mfma at K = 4, 8 and 12
      const int64_t nblk = (J + 31) / 32, nbx = (I + 127) / 128;
      CHK(c->qimg.grow((size_t)(nblk + 1) * qimg_bytes((int)K, true)));
      nchunk = std::max<int64_t>(1, (c->target_waves / 4 + nbx - 1) / nbx);
      nchunk = std::min<int64_t>(nchunk, nblk);
      const int64_t jchunk = ((J + nchunk - 1) / nchunk + 31) / 32 * 32;
      nchunk = (J + jchunk - 1) / jchunk;
      CHK(c->part.grow(sizeof(double) * (size_t)(nchunk * (K + 1) * I)));
      launch_qimg((int)K, p, J, c->qimg.p, true, c->s);
      launch_emq_img((int)K, c->bed, c->pitch, I, J, jchunk, (int)nchunk, q, c->lqt.p, c->qimg.p,
                     static_cast<double*>(c->part.p), c->s);
    } else if (qside) {
      uint16_t* ps = static_cast<uint16_t*>(c->lps.p);  // BF16 rows: p, then 1 - p
      launch_em_split((int)K, p, J, ld, ps, ps + 32 * ld, nullptr, c->s);
      const int64_t nbx = (I + 255) / 256;
      nchunk = std::max<int64_t>(1, (c->target_waves / 4 + nbx - 1) / nbx);
      nchunk = std::min<int64_t>(nchunk, std::max<int64_t>(1, J / 256));
      const int64_t jchunk = ((J + nchunk - 1) / nchunk + 31) / 32 * 32;
      nchunk = (J + jchunk - 1) / jchunk;
      CHK(c->part.grow(sizeof(double) * (size_t)(nchunk * (K + 1) * I)));
      launch_emq_mfma((int)K, c->bed, c->pitch, I, J, jchunk, (int)nchunk, q, c->lqt.p, c->lpt.p, c->lmt.p, ps,
                      ps + 32 * ld, ld, static_cast<double*>(c->part.p), c->s);
    } else {
      launch_em_split((int)K, q, I, ld, c->lps.p, nullptr, static_cast<float*>(c->lqc.p), c->s);
      const int64_t nbj = (J + 255) / 256;
      nchunk = std::max<int64_t>(1, (c->target_waves / 4 + nbj - 1) / nbj);
      nchunk = std::min<int64_t>(nchunk, std::max<int64_t>(1, I / 256));
      // chunks of whole 128-byte BED lines (512 individuals), so no line is split between two workgroups
      const int64_t ichunk = ((I + nchunk - 1) / nchunk + 511) / 512 * 512;
      nchunk = (I + ichunk - 1) / ichunk;
      CHK(c->part.grow(sizeof(double) * (size_t)(nchunk * 2 * K * J)));
      launch_emp_mfma((int)K, c->bed, c->pitch, I, J, ichunk, (int)nchunk, static_cast<const float*>(c->lqc.p),
                      c->lqt.p, c->lpt.p, c->lmt.p, c->lps.p, ld, static_cast<double*>(c->part.p), c->s);
    }
  }
  *nchunk_out = (int)nchunk;
  return check_launch();
}

// LL of (q, p) over a range into device slot lls[slot] (accumulate optional).
// The BF16 form takes the matrix-core path for the device state's full-cohort
// LL (allow_mixed); the tile-loop entry point always runs the FP64/FP32 loop.
template <typename T>
int ll_to_slot(admx_ctx* c, const void* q, const void* p, int64_t i0, int64_t i1, int64_t j0, int64_t j1, int slot,
               bool allow_mixed = true) {
  if (allow_mixed && sizeof(T) == 8 && c->precision == ADMX_PREC_BF16 && !c->impute && !c->stream_snps && i0 == 0 &&
      i1 == c->I && j0 == 0 && j1 == c->J)
    return ll_mixed(c, static_cast<const double*>(q), static_cast<const double*>(p), slot);
  int nchunk = 0;
  int64_t np = 0;
  CHK(pass_q<T>(c, MODE_LL, q, p, i0, i1, j0, j1, &nchunk, &np));
  ProfScope ps(c, "ll_reduce");
  launch_reduce_sum(static_cast<const double*>(c->llpart.p), np, c->lls + slot, 0, c->s);
  return check_launch();
}

// Full-cohort LLs of two states into two slots from one pass over the
// genotypes (FP64/FP32 loop; one fused kernel for the product form, see
// launch_qside_ll2): each genotype's code is read once for both states, and a
// streamed context transfers the BED once instead of twice; each LL is summed
// exactly as ll_to_slot sums it.
template <typename T>
int ll2_to_slots(admx_ctx* c, void* x1, int slot1, void* x2, int slot2) {
  int nchunk = 0;
  int64_t np = 0;
  CHK(pass_q<T>(c, MODE_LL, x1, c->pp(x1), 0, c->I, 0, c->J, &nchunk, &np, x2, c->pp(x2)));
  ProfScope ps(c, "ll_reduce");
  launch_reduce_sum(static_cast<const double*>(c->llpart.p), np, c->lls + slot1, 0, c->s);
  launch_reduce_sum(static_cast<const double*>(c->llpart.p) + np, np, c->lls + slot2, 0, c->s);
  return check_launch();
}

// Measured at I = J = 100,000: the matrix-core update_q pass is faster from
// K = 7 (7.3 vs 7.5 ms; 7.9 vs 9.2 at K = 8, 13.8 vs 46.6 at K = 16) and slower
// below (6.9 vs 6.0 ms at K = 4), where s and the gradient are few FMAs.
bool qside_on_matrix_cores(const admx_ctx* c) {
  return c->qside_mfma == 1 || (c->qside_mfma == -1 && c->K >= 7);
}

// BF16 form, Q side: s, t, the Hessian and the gradient on the matrix cores
// (the split table of q, one image slot per 32-SNP block of p), then the FP64
// finalize + QP + projection of the other mixed passes.
int mixed_update_q_mfma(admx_ctx* c, const double* q, const double* p, const double* x_in, double* x_out,
                        int32_t* status_dev) {
  const int64_t I = c->I, J = c->J, K = c->K;
  const int KD = split_table_cols((int)K);
  const int64_t rows_q = (I + 255) / 256 * 256, nblk = (J + 31) / 32;
  CHK(c->lqt.grow(sizeof(uint16_t) * KD * rows_q));
  CHK(c->qimg.grow((size_t)(nblk + 1) * qimg_bytes((int)K)));
  const int64_t nbx = (I + 127) / 128;  // 128 individuals (4 waves) per workgroup
  int64_t nchunk = std::max<int64_t>(1, (c->target_waves / 4 + nbx - 1) / nbx);
  nchunk = std::min<int64_t>(nchunk, nblk);
  const int64_t jchunk = ((J + nchunk - 1) / nchunk + 31) / 32 * 32;
  nchunk = (J + jchunk - 1) / jchunk;
  const int na = (int)(K * (K + 1) / 2 + K);
  CHK(c->part.grow((size_t)nchunk * na * I * sizeof(float)));
  {
    ProfScope ps(c, "mixed_prep");
    launch_split_tables((int)K, q, I, rows_q, 0, c->lqt.p, nullptr, c->s);
    launch_qimg((int)K, p, J, c->qimg.p, false, c->s);
  }
  CHK(check_launch());
  {
    ProfScope ps(c, "qside_update_bf16");
    launch_qside_mfma((int)K, c->bed, c->pitch, I, J, jchunk, (int)nchunk, c->lqt.p, c->qimg.p,
                      static_cast<float*>(c->part.p), c->s);
  }
  CHK(check_launch());
  FinArgs f{c->part.p, I, (int)nchunk, 0, J, x_in, x_out, nullptr, c->keep_hess ? c->XtX_q : nullptr,
            c->keep_hess ? c->Xtz_q : nullptr, 0, 1, status_dev, c->XtX_q, c->Xtz_q};
  {
    ProfScope ps(c, "fin_q_update_qp");
    launch_fin_q_update_mixed((int)K, f, c->s);
  }
  return check_launch();
}

bool pside_on_matrix_cores(const admx_ctx* c) {
  return c->pside_mfma == 1 || (c->pside_mfma == -1 && c->K >= 7);
}

// BF16 form, P side: the mirror of mixed_update_q_mfma (split table of p, one
// image slot per 32-individual block of q).
int mixed_update_p_mfma(admx_ctx* c, const double* q, const double* p, const double* x_in, double* x_out,
                        int32_t* status_dev) {
  const int64_t I = c->I, J = c->J, K = c->K;
  const int KD = split_table_cols((int)K);
  // the kernel reads split rows up to the last 128-SNP workgroup's end
  const int64_t rows_p = (J + 127) / 128 * 128 + 32, nblk = (I + 31) / 32;
  CHK(c->lpt.grow(sizeof(uint16_t) * KD * rows_p));
  CHK(c->lmt.grow(sizeof(uint16_t) * KD * rows_p));
  CHK(c->pimg.grow((size_t)(nblk + 1) * pimg_bytes((int)K)));
  const int64_t nbx = (J + 127) / 128;  // 128 SNPs (4 waves) per workgroup
  int64_t nchunk = std::max<int64_t>(1, (c->target_waves / 4 + nbx - 1) / nbx);
  nchunk = std::min<int64_t>(nchunk, std::max<int64_t>(1, I / 512));
  // chunks of whole 128-byte BED lines (512 individuals) where I allows
  const int64_t ichunk = ((I + nchunk - 1) / nchunk + 511) / 512 * 512;
  nchunk = (I + ichunk - 1) / ichunk;
  const int na = (int)(K * (K + 1) / 2 + K);
  CHK(c->part.grow((size_t)nchunk * na * J * sizeof(float)));
  {
    ProfScope ps(c, "mixed_prep");
    launch_split_tables((int)K, p, J, rows_p, 1, c->lpt.p, c->lmt.p, c->s);
    launch_pimg((int)K, q, I, c->pimg.p, c->s);
  }
  CHK(check_launch());
  {
    ProfScope ps(c, "pside_update_bf16");
    launch_pside_mfma((int)K, c->bed, c->pitch, I, J, ichunk, (int)nchunk, c->lpt.p, c->pimg.p,
                      static_cast<float*>(c->part.p), c->s);
  }
  CHK(check_launch());
  FinArgs f{c->part.p, J, (int)nchunk, 0, J, x_in, x_out, nullptr, c->keep_hess ? c->XtX_p : nullptr,
            c->keep_hess ? c->Xtz_p : nullptr, 0, 1, status_dev, c->XtX_p, c->Xtz_p};
  {
    ProfScope ps(c, "fin_p_update_qp");
    launch_fin_p_update_mixed((int)K, f, c->s);
  }
  return check_launch();
}

// Mixed-precision update passes of a FP64 context: FP32 copies of (q, p), the
// BF16 pair tables, the FP32/BF16 pass, then the FP64 finalize + QP + projection.
int mixed_update(admx_ctx* c, bool qside, const double* q, const double* p, const double* x_in, double* x_out,
                 int32_t* status_dev) {
  if (c->impute) return bad("the mean-imputation loops run with the FP64/FP32 passes (ADMX_PREC_F64)");
  CHK(need_resident(c, "the mixed-precision passes"));
  if (qside && c->precision == ADMX_PREC_BF16 && qside_on_matrix_cores(c))
    return mixed_update_q_mfma(c, q, p, x_in, x_out, status_dev);
  if (!qside && c->precision == ADMX_PREC_BF16 && pside_on_matrix_cores(c))
    return mixed_update_p_mfma(c, q, p, x_in, x_out, status_dev);
  const int64_t I = c->I, J = c->J, K = c->K;
  if (c->fq.n < sizeof(float) * K * (I + 16)) {
    CHK(c->fq.grow(sizeof(float) * K * (I + 16)));
    std::vector<float> pad(K * 16, 0.5f / (float)K);  // padding individuals: s stays inside (0, 1)
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(static_cast<float*>(c->fq.p) + K * I, pad.data(), sizeof(float) * pad.size(),
                     hipMemcpyHostToDevice));
  }
  if (c->fp.n < sizeof(float) * K * (J + 17)) {
    CHK(c->fp.grow(sizeof(float) * K * (J + 17)));
    std::vector<float> half(K * 17, 0.5f);  // padding SNP columns: s stays inside (0, 1)
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(static_cast<float*>(c->fp.p) + K * J, half.data(), sizeof(float) * half.size(),
                     hipMemcpyHostToDevice));
  }
  float* fq = static_cast<float*>(c->fq.p);
  float* fp = static_cast<float*>(c->fp.p);
  const bool bf16 = c->precision == ADMX_PREC_BF16;
  const int rows = pair_table_rows((int)K);
  const int64_t ldj = (J + 15) & ~(int64_t)15, ldi = (I + 15) & ~(int64_t)15;
  {
    ProfScope ps(c, "mixed_prep");
    launch_to_float(q, K * I, fq, c->s);
    launch_to_float(p, K * J, fp, c->s);
    if (bf16) {
      if (qside) {
        CHK(c->ppt.grow((size_t)rows * ldj * 2));
        CHK(c->mmt.grow((size_t)rows * ldj * 2));
        launch_pair_tables((int)K, p, J, ldj, c->ppt.p, c->mmt.p, c->s);
      } else {
        CHK(c->qqt.grow((size_t)rows * ldi * 2));
        launch_pair_tables((int)K, q, I, ldi, c->qqt.p, nullptr, c->s);
      }
    }
  }
  CHK(check_launch());
  const int na = (int)(K * (K + 1) / 2 + K);  // direct form: Hessian triangle + gradient
  int64_t chunk = 0;
  int n = 0;
  if (qside) {
    n = qside_chunks(c, I, J, &chunk);
    chunk = (chunk + 15) & ~(int64_t)15;
    n = (int)((J + chunk - 1) / chunk);
    CHK(c->part.grow((size_t)n * na * I * sizeof(float)));
  } else {
    n = pside_chunks(c, 0, I, J, &chunk);
    CHK(c->part.grow((size_t)n * na * J * sizeof(float)));
  }
  PassArgs a{c->bed, c->pitch, I, fq, fp, 0, I, 0, J, chunk, c->part.p};
  {
    ProfScope ps(c, qside ? (bf16 ? "qside_update_bf16" : "qside_update_f32")
                          : (bf16 ? "pside_update_bf16" : "pside_update_f32"));
    if (qside) {
      if (bf16) launch_qside_bf16((int)K, a, n, fq, fp, c->ppt.p, c->mmt.p, ldj, c->s);
      else launch_qside_f32_update((int)K, a, n, c->s);
    } else {
      if (bf16) launch_pside_bf16((int)K, a, n, fq, fp, c->qqt.p, ldi, c->s);
      else launch_pside_f32_update((int)K, a, n, c->s);
    }
  }
  CHK(check_launch());
  FinArgs f{c->part.p, qside ? I : J, n, 0, J, x_in, x_out, nullptr,
            c->keep_hess ? (qside ? c->XtX_q : c->XtX_p) : nullptr, c->keep_hess ? (qside ? c->Xtz_q : c->Xtz_p) : nullptr,
            0, 1, status_dev, qside ? c->XtX_q : c->XtX_p, qside ? c->Xtz_q : c->Xtz_p};
  {
    ProfScope ps(c, qside ? "fin_q_update_qp" : "fin_p_update_qp");
    if (qside) launch_fin_q_update_mixed((int)K, f, c->s);
    else launch_fin_p_update_mixed((int)K, f, c->s);
  }
  return check_launch();
}

template <typename T>
int dev_update_q(admx_ctx* c, bool update2, int32_t* status_host) {
  void* X = c->x[ADMX_X];
  void* XN = c->x[ADMX_X_NEXT];
  void* XN2 = c->x[ADMX_X_NEXT2];
  void* q_next = update2 ? XN2 : XN;
  void* q = update2 ? XN : X;
  void* p = c->pp(update2 ? XN : X);
  if (sizeof(T) == 8 && c->precision != ADMX_PREC_F64) {
    CHK(mixed_update(c, true, (const double*)q, (const double*)p, (const double*)q, (double*)q_next,
                     status_host ? c->status : nullptr));
    if (status_host) {
      HIPCHK(hipMemcpyAsync(status_host, c->status, sizeof(int32_t) * c->I, hipMemcpyDeviceToHost, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
    }
    return ADMX_OK;
  }
  int nchunk = 0;
  int64_t np = 0;
  CHK(pass_q<T>(c, MODE_UPDATE, q, p, 0, c->I, 0, c->J, &nchunk, &np));
  FinArgs f{c->part.p, c->I, nchunk, 0, c->J, q, q_next, nullptr,
            c->keep_hess ? c->XtX_q : nullptr, c->keep_hess ? c->Xtz_q : nullptr, 0, 1,
            status_host ? c->status : nullptr, c->XtX_q, c->Xtz_q};
  {
    ProfScope ps(c, "fin_q_update_qp");
    launch_fin_q_update<T>((int)c->K, f, true, c->s);
  }
  CHK(check_launch());
  if (status_host) {
    HIPCHK(hipMemcpyAsync(status_host, c->status, sizeof(int32_t) * c->I, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
  }
  return ADMX_OK;
}

template <typename T>
int dev_update_p(admx_ctx* c, bool update2, int32_t* status_host) {
  void* X = c->x[ADMX_X];
  void* XN = c->x[ADMX_X_NEXT];
  void* XN2 = c->x[ADMX_X_NEXT2];
  void* p_next = c->pp(update2 ? XN2 : XN);
  void* q = update2 ? XN2 : XN;
  void* p = c->pp(update2 ? XN : X);
  if (sizeof(T) == 8 && c->precision != ADMX_PREC_F64) {
    CHK(mixed_update(c, false, (const double*)q, (const double*)p, (const double*)p, (double*)p_next,
                     status_host ? c->status : nullptr));
    if (status_host) {
      HIPCHK(hipMemcpyAsync(status_host, c->status, sizeof(int32_t) * c->J, hipMemcpyDeviceToHost, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
    }
    return ADMX_OK;
  }
  int nchunk = 0;
  CHK(pass_p<T>(c, MODE_UPDATE, q, p, 0, c->I, 0, c->J, &nchunk));
  FinArgs f{c->part.p, c->J, nchunk, 0, c->J, p, p_next, nullptr,
            c->keep_hess ? c->XtX_p : nullptr, c->keep_hess ? c->Xtz_p : nullptr, 0, 1,
            status_host ? c->status : nullptr, c->XtX_p, c->Xtz_p};
  {
    ProfScope ps(c, "fin_p_update_qp");
    launch_fin_p_update<T>((int)c->K, f, true, c->s);
  }
  CHK(check_launch());
  if (status_host) {
    HIPCHK(hipMemcpyAsync(status_host, c->status, sizeof(int32_t) * c->J, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
  }
  return ADMX_OK;
}

// update_q_cuda!(d_cu, g_cu) / update_p_cuda!(d_cu, g_cu) (src/cuda/runners.jl:69-111):
// XtX and Xtz of the state x's (q, p) into the context's Hessian arrays, zeroed
// then written (the reference fills them with zeros and accumulates), Xtz not
// negated; no QP.  The FP64 passes and finalize of the tile loops.
template <typename T>
int dev_grad_hessian(admx_ctx* c, bool qside) {
  void* X = c->x[ADMX_X];
  int nchunk = 0;
  int64_t np = 0;
  if (qside) CHK(pass_q<T>(c, MODE_UPDATE, X, c->pp(X), 0, c->I, 0, c->J, &nchunk, &np));
  else CHK(pass_p<T>(c, MODE_UPDATE, X, c->pp(X), 0, c->I, 0, c->J, &nchunk));
  FinArgs f{c->part.p, qside ? c->I : c->J, nchunk, 0, c->J, qside ? X : c->pp(X), nullptr, nullptr,
            qside ? c->XtX_q : c->XtX_p, qside ? c->Xtz_q : c->Xtz_p, 0, 0, nullptr};
  {
    ProfScope ps(c, qside ? "fin_q_update" : "fin_p_update");
    if (qside) launch_fin_q_update<T>((int)c->K, f, false, c->s);
    else launch_fin_p_update<T>((int)c->K, f, false, c->s);
  }
  return check_launch();
}

// em_q!(d, g; mode) (src/algorithms_inner.jl:68-92): q of x_next from (q, p) of x
// (ADMX_EM_BASE, ADMX_EM_FALLBACK), or q of x_next2 from (q, p) of x_next
// (ADMX_EM_FALLBACK2).
template <typename T>
int dev_em_q(admx_ctx* c, int mode = ADMX_EM_BASE) {
  void* src = c->x[mode == ADMX_EM_FALLBACK2 ? ADMX_X_NEXT : ADMX_X];
  void* dst = c->x[mode == ADMX_EM_FALLBACK2 ? ADMX_X_NEXT2 : ADMX_X_NEXT];
  const void* q = src;
  const void* p = c->pp(src);
  int nchunk = 0;
  int64_t np = 0;
  bool mfma = em_on_matrix_cores(c);
  if (mfma) CHK(em_state_is_safe(c, static_cast<const double*>(q), static_cast<const double*>(p), &mfma));
  if (mfma) {
    CHK(need_resident(c, "the matrix-core EM passes"));
    CHK(em_mixed(c, true, static_cast<const double*>(q), static_cast<const double*>(p), &nchunk));
  } else {
    CHK(pass_q<T>(c, MODE_EM, q, p, 0, c->I, 0, c->J, &nchunk, &np));
  }
  FinArgs f{c->part.p, c->I, nchunk, 0, c->J, q, dst, nullptr, nullptr, nullptr, 0, 0, nullptr};
  ProfScope ps(c, "fin_q_em");
  launch_fin_q_em<T>((int)c->K, f, c->s);
  return check_launch();
}

// em_p!(d, g; mode) (src/algorithms_inner.jl:108-141): p of x_next from q, p of x
// (ADMX_EM_BASE), from q of x_next and p of x (ADMX_EM_FALLBACK), or p of
// x_next2 from q of x_next2 and p of x_next (ADMX_EM_FALLBACK2).
template <typename T>
int dev_em_p(admx_ctx* c, int mode = ADMX_EM_BASE) {
  const void* q = c->x[mode == ADMX_EM_BASE ? ADMX_X : (mode == ADMX_EM_FALLBACK ? ADMX_X_NEXT : ADMX_X_NEXT2)];
  const void* p = c->pp(c->x[mode == ADMX_EM_FALLBACK2 ? ADMX_X_NEXT : ADMX_X]);
  void* dst = c->pp(c->x[mode == ADMX_EM_FALLBACK2 ? ADMX_X_NEXT2 : ADMX_X_NEXT]);
  int nchunk = 0;
  bool mfma = em_on_matrix_cores(c);
  if (mfma) CHK(em_state_is_safe(c, static_cast<const double*>(q), static_cast<const double*>(p), &mfma));
  if (mfma) {
    CHK(need_resident(c, "the matrix-core EM passes"));
    CHK(em_mixed(c, false, static_cast<const double*>(q), static_cast<const double*>(p), &nchunk));
  } else {
    CHK(pass_p<T>(c, MODE_EM, q, p, 0, c->I, 0, c->J, &nchunk));
  }
  FinArgs f{c->part.p, c->J, nchunk, 0, c->J, p, dst, nullptr, nullptr, nullptr, 0, 0, nullptr};
  ProfScope ps(c, "fin_p_em");
  launch_fin_p_em<T>((int)c->K, f, c->s);
  return check_launch();
}

template <typename T>
int dev_zal(admx_ctx* c, int64_t iter, const void* x, const void* xn, const void* xn2, void* U, void* V, void* out) {
  const int col = (int)((iter - 1) % c->Q);
  const int m = (int)std::min<int64_t>(iter, c->Q);
  constexpr int NP = kMaxQ * kMaxQ + kMaxQ;
  const int64_t len = 512 * NP + kMaxQ;
  CHK(c->zal.grow((size_t)len * sizeof(double)));
  ProfScope ps(c, "zal");
  launch_zal<T>((int)c->K, c->I, c->J, x, xn, xn2, U, V, out, col, m, (int)c->Q, static_cast<double*>(c->zal.p), len,
                c->s);
  return check_launch();
}

// LBQN step (update_UV_LBQN! + update_QN_LBQN!, src/quasi_newton.jl:15-58, and
// the projections): the history columns shift right by device copies, then
// the new first column, the Gram inner products, the recursion and the apply.
template <typename T>
int dev_lbqn(admx_ctx* c, int64_t iter, const void* x, const void* xn, const void* xn2, void* U, void* V,
             void* out) {
  const int m = (int)std::min<int64_t>(iter, c->Q);
  const size_t colb = (size_t)c->N() * sizeof(T);
  for (int col = m - 1; col >= 1; --col) {
    HIPCHK(hipMemcpyAsync(static_cast<char*>(U) + colb * col, static_cast<char*>(U) + colb * (col - 1), colb,
                          hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipMemcpyAsync(static_cast<char*>(V) + colb * col, static_cast<char*>(V) + colb * (col - 1), colb,
                          hipMemcpyDeviceToDevice, c->s));
  }
  constexpr int NP = kMaxQ * kMaxQ + kMaxQ;
  const int64_t len = 512 * NP + 2 * kMaxQ;
  CHK(c->zal.grow((size_t)len * sizeof(double)));
  ProfScope ps(c, "lbqn");
  launch_lbqn<T>((int)c->K, c->I, c->J, x, xn, xn2, U, V, out, m, static_cast<double*>(c->zal.p), len, c->s);
  return check_launch();
}

// Whether an iteration's two log-likelihoods share one pass (ll2_to_slots):
// always on a streamed context (one BED transfer instead of two), and where
// ll_to_slot would run the FP64 product form for both up to K = 8 (the fused
// kernel holds both states' q: measured at I = J = 100,000, 8.6 ms against
// 2 x 4.7 at K = 8, 6.1 against 2 x 3.7 at K = 4, even at K = 9, and
// 14.3 against 2 x 6.0 at K = 12, where it drops to four waves per SIMD).
bool ll_pairs(const admx_ctx* c) {
  const bool mixed_ll = c->dtype == ADMX_F64 && c->precision == ADMX_PREC_BF16 && !c->impute && !c->stream_snps;
  return c->stream_snps || (!c->impute && !mixed_ll && c->K <= 8);
}

// Everything of one admixture_qn! iteration up to the accept kernel, enqueued
// on the context stream without host synchronisation (capturable).
template <typename T>
int enqueue_qn_iteration(admx_ctx* c, int64_t iter, int fix_q, int fix_p) {
  const size_t qbytes = (size_t)(c->K * c->I) * c->esz, pbytes = (size_t)(c->K * c->J) * c->esz;
  void* X = c->x[ADMX_X];
  void* XN = c->x[ADMX_X_NEXT];
  void* XN2 = c->x[ADMX_X_NEXT2];
  void* XT = c->x[ADMX_X_TMP];
  HIPCHK(hipMemcpyAsync(c->lls + LL_PREV, c->lls + LL_NEW, sizeof(double), hipMemcpyDeviceToDevice, c->s));
  if (!fix_q) CHK(dev_update_q<T>(c, false, nullptr));
  else HIPCHK(hipMemcpyAsync(XN, X, qbytes, hipMemcpyDeviceToDevice, c->s));
  if (!fix_p) CHK(dev_update_p<T>(c, false, nullptr));
  else HIPCHK(hipMemcpyAsync(c->pp(XN), c->pp(X), pbytes, hipMemcpyDeviceToDevice, c->s));
  if (!fix_q) CHK(dev_update_q<T>(c, true, nullptr));
  else HIPCHK(hipMemcpyAsync(XN2, X, qbytes, hipMemcpyDeviceToDevice, c->s));
  if (!fix_p) CHK(dev_update_p<T>(c, true, nullptr));
  else HIPCHK(hipMemcpyAsync(c->pp(XN2), c->pp(X), pbytes, hipMemcpyDeviceToDevice, c->s));
  // ll_basic is evaluated after the quasi-Newton step (which does not read
  // it), in the same pass over the genotypes as ll_qn, wherever both run the
  // FP64 loop: the matrix-core log-likelihood of the BF16 form and the
  // mean-imputation loops keep one pass per state.
  if (!ll_pairs(c)) CHK(ll_to_slot<T>(c, XN2, c->pp(XN2), 0, c->I, 0, c->J, LL_BASIC));
  if (c->qn_mode == ADMX_QN_LBQN) CHK(dev_lbqn<T>(c, iter, X, XN, XN2, c->U, c->V, XT));
  else CHK(dev_zal<T>(c, iter, X, XN, XN2, c->U, c->V, XT));
  if (ll_pairs(c)) CHK(ll2_to_slots<T>(c, XN2, LL_BASIC, XT, LL_QN));
  else CHK(ll_to_slot<T>(c, XT, c->pp(XT), 0, c->I, 0, c->J, LL_QN));
  {
    ProfScope ps(c, "accept");
    launch_accept<T>(c->N(), X, XT, XN2, c->lls, c->s);
  }
  return check_launch();
}

// Signature of everything a captured iteration bakes in: scratch buffer
// addresses and sizes and the tuning that shapes the launches.
uint64_t layout_sig(const admx_ctx* c) {
  uint64_t h = 1469598103934665603ULL;
  auto mix = [&](uint64_t v) {
    h ^= v;
    h *= 1099511628211ULL;
  };
  const DevBuf* bufs[] = {&c->part,    &c->llpart,  &c->zal,     &c->fq,  &c->fp,  &c->ppt,
                          &c->mmt,     &c->qqt,     &c->wide[0], &c->wide[1], &c->wide[2], &c->wide[3],
                          &c->llflag,  &c->lqt,     &c->lpt,     &c->lmt, &c->lps, &c->lqc, &c->qimg, &c->pimg};
  for (const DevBuf* b : bufs) {
    mix((uint64_t)(uintptr_t)b->p);
    mix((uint64_t)b->n);
  }
  mix((uint64_t)c->target_waves);
  mix((uint64_t)c->keep_hess);
  mix((uint64_t)c->precision);
  mix((uint64_t)c->qn_mode);
  mix((uint64_t)c->impute);
  mix((uint64_t)(c->split_q + 2));
  mix((uint64_t)(c->split_p + 2));
  mix((uint64_t)(c->qside_mfma + 2));
  mix((uint64_t)(c->pside_mfma + 2));
  mix((uint64_t)(uintptr_t)c->snp_mu.p);
  return h;
}

void drop_graphs(admx_ctx* c) {
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
  c->graph_warm = 0;
}

template <typename T>
int dev_qn_iteration(admx_ctx* c, int64_t iter, int fix_q, int fix_p, double* out) {
  const unsigned kind = (unsigned)(fix_q ? 1 : 0) | (fix_p ? 2u : 0u);
  if (layout_sig(c) != c->graph_sig) {
    drop_graphs(c);
    c->graph_sig = layout_sig(c);
  }
  if (!c->use_graphs || c->stream_snps || c->prof.on || iter <= c->Q || !(c->graph_warm & (1u << kind))) {
    CHK(enqueue_qn_iteration<T>(c, iter, fix_q, fix_p));
    // scratch buffers may have grown in this eager run: graphs need a run
    // under an unchanged signature before capture (no allocation is captured)
    if (layout_sig(c) != c->graph_sig) {
      drop_graphs(c);
      c->graph_sig = layout_sig(c);
    } else {
      c->graph_warm |= 1u << kind;
    }
  } else {
    const int64_t key = ((iter - 1) % c->Q) | ((int64_t)kind << 8);
    hipGraphExec_t exec = nullptr;
    for (auto& g : c->graphs)
      if (g.key == key) exec = g.exec;
    if (!exec) {
      HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_qn_iteration<T>(c, iter, fix_q, fix_p);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(c->s, &graph);
      if (rc != ADMX_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (e != hipSuccess) return fail(ADMX_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e2 != hipSuccess) return fail(ADMX_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
      c->graphs.push_back({key, exec});
    }
    HIPCHK(hipGraphLaunch(exec, c->s));
  }
  HIPCHK(hipMemcpyAsync(c->h_lls, c->lls, sizeof(double) * LL_NSLOTS, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  out[0] = c->h_lls[LL_PREV];
  out[1] = c->h_lls[LL_BASIC];
  out[2] = c->h_lls[LL_QN];
  out[3] = c->h_lls[LL_NEW];
  out[4] = std::fabs((out[3] - out[0]) / out[0]);
  return ADMX_OK;
}

template <typename T>
int dev_init_em(admx_ctx* c, int64_t iters, double* lls) {
  const size_t qbytes = (size_t)(c->K * c->I) * c->esz, pbytes = (size_t)(c->K * c->J) * c->esz;
  void* X = c->x[ADMX_X];
  void* XN = c->x[ADMX_X_NEXT];
  for (int64_t it = 0; it < iters; ++it) {
    CHK(dev_em_q<T>(c));
    CHK(dev_em_p<T>(c));
    HIPCHK(hipMemcpyAsync(c->pp(X), c->pp(XN), pbytes, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipMemcpyAsync(X, XN, qbytes, hipMemcpyDeviceToDevice, c->s));
    if (lls) {
      CHK(ll_to_slot<T>(c, X, c->pp(X), 0, c->I, 0, c->J, LL_TMP));
      HIPCHK(hipMemcpyAsync(lls + it, c->lls + LL_TMP, sizeof(double), hipMemcpyDeviceToHost, c->s));
    }
  }
  // d.ll_new is the log-likelihood of the state the last EM step left, which
  // that step's pass has just summed (same kernels and order: the same bits).
  if (lls && iters > 0)
    HIPCHK(hipMemcpyAsync(c->lls + LL_NEW, c->lls + LL_TMP, sizeof(double), hipMemcpyDeviceToDevice, c->s));
  else
    CHK(ll_to_slot<T>(c, X, c->pp(X), 0, c->I, 0, c->J, LL_NEW));
  HIPCHK(hipStreamSynchronize(c->s));
  return ADMX_OK;
}

// ---- host-array tile loops (stage, run, accumulate on device, copy back)

template <typename T>
int loop_ll(admx_ctx* c, const void* q, const void* p, int64_t i0, int64_t i1, int64_t j0, int64_t j1, double* out) {
  const size_t qb = (size_t)(c->K * c->I) * sizeof(T), pb = (size_t)(c->K * c->J) * sizeof(T);
  CHK(stage_in(c, 0, q, qb, qb));
  CHK(stage_in(c, 1, p, pb, pb));
  if (i1 > i0 && j1 > j0) {
    CHK(ll_to_slot<T>(c, c->st[0].p, c->st[1].p, i0, i1, j0, j1, LL_TMP, false));
    HIPCHK(hipMemcpyAsync(out, c->lls + LL_TMP, sizeof(double), hipMemcpyDeviceToHost, c->s));
  } else {
    *out = 0.0;
  }
  HIPCHK(hipStreamSynchronize(c->s));
  return ADMX_OK;
}

template <typename T>
int loop_update(admx_ctx* c, bool qside, void* XtX, void* Xtz, const void* q, const void* p, int64_t i0, int64_t i1,
                int64_t j0, int64_t j1) {
  const int64_t K = c->K;
  const size_t qb = (size_t)(K * c->I) * sizeof(T), pb = (size_t)(K * c->J) * sizeof(T);
  const int64_t ncol = qside ? c->I : c->J;
  const int64_t c0 = qside ? i0 : j0, c1 = qside ? i1 : j1;
  const size_t hb = (size_t)(K * K) * sizeof(T), zb = (size_t)K * sizeof(T);
  CHK(stage_in(c, 0, q, qb, qb));
  CHK(stage_in(c, 1, p, pb, pb));
  CHK(c->st[2].grow(hb * ncol));
  CHK(c->st[3].grow(zb * ncol));
  const size_t nb = (size_t)(c1 - c0);
  char* dX = static_cast<char*>(c->st[2].p) + hb * c0;
  char* dZ = static_cast<char*>(c->st[3].p) + zb * c0;
  if (nb) {
    HIPCHK(hipMemcpyAsync(dX, static_cast<char*>(XtX) + hb * c0, hb * nb, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(dZ, static_cast<char*>(Xtz) + zb * c0, zb * nb, hipMemcpyHostToDevice, c->s));
  }
  int nchunk = 0;
  int64_t np = 0;
  if (qside) CHK(pass_q<T>(c, MODE_UPDATE, c->st[0].p, c->st[1].p, i0, i1, j0, j1, &nchunk, &np));
  else CHK(pass_p<T>(c, MODE_UPDATE, c->st[0].p, c->st[1].p, i0, i1, j0, j1, &nchunk));
  if (nb) {
    FinArgs f{c->part.p, (int64_t)nb, nchunk, c0, c->J, qside ? c->st[0].p : c->st[1].p, nullptr, nullptr,
              c->st[2].p, c->st[3].p, 1, 0, nullptr};
    {
      ProfScope ps(c, qside ? "fin_q_update" : "fin_p_update");
      if (qside) launch_fin_q_update<T>((int)K, f, false, c->s);
      else launch_fin_p_update<T>((int)K, f, false, c->s);
    }
    CHK(check_launch());
    HIPCHK(hipMemcpyAsync(static_cast<char*>(XtX) + hb * c0, dX, hb * nb, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(static_cast<char*>(Xtz) + zb * c0, dZ, zb * nb, hipMemcpyDeviceToHost, c->s));
  }
  HIPCHK(hipStreamSynchronize(c->s));
  return ADMX_OK;
}

template <typename T>
int loop_em_q(admx_ctx* c, void* q_next, const void* q, const void* p, int64_t i0, int64_t i1, int64_t j0,
              int64_t j1) {
  const int64_t K = c->K;
  const size_t qb = (size_t)(K * c->I) * sizeof(T), pb = (size_t)(K * c->J) * sizeof(T), cb = K * sizeof(T);
  CHK(stage_in(c, 0, q, qb, qb));
  CHK(stage_in(c, 1, p, pb, pb));
  CHK(c->st[2].grow(qb));
  const size_t nb = (size_t)(i1 - i0);
  char* dO = static_cast<char*>(c->st[2].p) + cb * i0;
  if (nb) HIPCHK(hipMemcpyAsync(dO, static_cast<char*>(q_next) + cb * i0, cb * nb, hipMemcpyHostToDevice, c->s));
  int nchunk = 0;
  int64_t np = 0;
  CHK(pass_q<T>(c, MODE_EM, c->st[0].p, c->st[1].p, i0, i1, j0, j1, &nchunk, &np));
  if (nb) {
    FinArgs f{c->part.p, (int64_t)nb, nchunk, i0, c->J, c->st[0].p, c->st[2].p, nullptr, nullptr, nullptr, 1, 0,
              nullptr};
    {
      ProfScope ps(c, "fin_q_em");
      launch_fin_q_em<T>((int)K, f, c->s);
    }
    CHK(check_launch());
    HIPCHK(hipMemcpyAsync(static_cast<char*>(q_next) + cb * i0, dO, cb * nb, hipMemcpyDeviceToHost, c->s));
  }
  HIPCHK(hipStreamSynchronize(c->s));
  return ADMX_OK;
}

template <typename T>
int loop_em_p(admx_ctx* c, void* p_next, void* p_tmp, const void* q, const void* p, int64_t i0, int64_t i1,
              int64_t j0, int64_t j1) {
  const int64_t K = c->K;
  const size_t qb = (size_t)(K * c->I) * sizeof(T), pb = (size_t)(K * c->J) * sizeof(T), cb = K * sizeof(T);
  CHK(stage_in(c, 0, q, qb, qb));
  CHK(stage_in(c, 1, p, pb, pb));
  CHK(c->st[2].grow(pb));
  CHK(c->st[3].grow(pb));
  const size_t nb = (size_t)(j1 - j0);
  char* dO = static_cast<char*>(c->st[2].p) + cb * j0;
  char* dT = static_cast<char*>(c->st[3].p) + cb * j0;
  if (nb) {
    HIPCHK(hipMemcpyAsync(dO, static_cast<char*>(p_next) + cb * j0, cb * nb, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(dT, static_cast<char*>(p_tmp) + cb * j0, cb * nb, hipMemcpyHostToDevice, c->s));
  }
  int nchunk = 0;
  CHK(pass_p<T>(c, MODE_EM, c->st[0].p, c->st[1].p, i0, i1, j0, j1, &nchunk));
  if (nb) {
    FinArgs f{c->part.p, (int64_t)nb, nchunk, j0, c->J, c->st[1].p, c->st[2].p, c->st[3].p, nullptr, nullptr, 1, 0,
              nullptr};
    {
      ProfScope ps(c, "fin_p_em");
      launch_fin_p_em<T>((int)K, f, c->s);
    }
    CHK(check_launch());
    HIPCHK(hipMemcpyAsync(static_cast<char*>(p_next) + cb * j0, dO, cb * nb, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(static_cast<char*>(p_tmp) + cb * j0, dT, cb * nb, hipMemcpyDeviceToHost, c->s));
  }
  HIPCHK(hipStreamSynchronize(c->s));
  return ADMX_OK;
}

template <typename T>
int host_qp_batch(admx_ctx* c, int simplex, int64_t n, const void* XtX, const void* Xtz, const void* x, void* delta,
                  int32_t* status) {
  const int64_t K = c->K;
  const size_t hb = (size_t)(K * K * n) * sizeof(T), zb = (size_t)(K * n) * sizeof(T);
  CHK(stage_in(c, 0, XtX, hb, hb));
  CHK(stage_in(c, 1, Xtz, zb, zb));
  CHK(stage_in(c, 2, x, zb, zb));
  CHK(c->st[3].grow(zb));
  CHK(c->st[4].grow(sizeof(int32_t) * n));
  {
    ProfScope ps(c, "qp_batch");
    launch_qp_batch<T>((int)K, simplex != 0, n, c->st[0].p, c->st[1].p, c->st[2].p, c->st[3].p,
                       static_cast<int32_t*>(c->st[4].p), c->s);
  }
  CHK(check_launch());
  HIPCHK(hipMemcpyAsync(delta, c->st[3].p, zb, hipMemcpyDeviceToHost, c->s));
  if (status) HIPCHK(hipMemcpyAsync(status, c->st[4].p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return ADMX_OK;
}

template <typename T>
int host_project(admx_ctx* c, bool simplex, void* b, int64_t n) {
  const size_t bytes = (size_t)(simplex ? c->K * n : n) * sizeof(T);
  CHK(stage_in(c, 0, b, bytes, bytes));
  {
    ProfScope ps(c, simplex ? "project_q" : "project_p");
    if (simplex) launch_project_q<T>((int)c->K, n, c->st[0].p, c->s);
    else launch_project_p<T>(n, c->st[0].p, c->s);
  }
  CHK(check_launch());
  HIPCHK(hipMemcpyAsync(b, c->st[0].p, bytes, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return ADMX_OK;
}

template <typename T>
int host_zal(admx_ctx* c, int64_t iter, const void* x, const void* xn, const void* xn2, void* U, void* V,
             void* x_tmp) {
  const size_t nb = (size_t)c->N() * sizeof(T), ub = nb * (size_t)c->Q;
  CHK(stage_in(c, 0, x, nb, nb));
  CHK(stage_in(c, 1, xn, nb, nb));
  CHK(stage_in(c, 2, xn2, nb, nb));
  CHK(stage_in(c, 3, U, ub, ub));
  CHK(stage_in(c, 4, V, ub, ub));
  CHK(c->st[5].grow(nb));
  CHK(dev_zal<T>(c, iter, c->st[0].p, c->st[1].p, c->st[2].p, c->st[3].p, c->st[4].p, c->st[5].p));
  HIPCHK(hipMemcpyAsync(U, c->st[3].p, ub, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipMemcpyAsync(V, c->st[4].p, ub, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipMemcpyAsync(x_tmp, c->st[5].p, nb, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return ADMX_OK;
}

template <typename T>
int host_lbqn(admx_ctx* c, int64_t iter, const void* x, const void* xn, const void* xn2, void* U, void* V,
              void* x_tmp) {
  const size_t nb = (size_t)c->N() * sizeof(T), ub = nb * (size_t)c->Q;
  CHK(stage_in(c, 0, x, nb, nb));
  CHK(stage_in(c, 1, xn, nb, nb));
  CHK(stage_in(c, 2, xn2, nb, nb));
  CHK(stage_in(c, 3, U, ub, ub));
  CHK(stage_in(c, 4, V, ub, ub));
  CHK(c->st[5].grow(nb));
  CHK(dev_lbqn<T>(c, iter, c->st[0].p, c->st[1].p, c->st[2].p, c->st[3].p, c->st[4].p, c->st[5].p));
  HIPCHK(hipMemcpyAsync(U, c->st[3].p, ub, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipMemcpyAsync(V, c->st[4].p, ub, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipMemcpyAsync(x_tmp, c->st[5].p, nb, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return ADMX_OK;
}

}  // namespace

// =============================================================================== C ABI
extern "C" {

const char* admx_version(void) { return "libadmixhip 0.1.0 (gfx950)"; }

const char* admx_last_error(void) { return g_err.c_str(); }

int admx_device_count(int* n) {
  return guarded([&]() -> int {
    if (!n) return bad("null pointer");
    int c = 0;
    HIPCHK(hipGetDeviceCount(&c));
    *n = c;
    return ADMX_OK;
  });
}

}  // extern "C"

namespace {
int create_impl(admx_ctx** out, int64_t I, int64_t J, int64_t K, int64_t Q, int dtype, int64_t stream_snps,
                const char* bed_path = nullptr) {
  if (!out) return bad("null output pointer");
  *out = nullptr;
  if (I < 1 || J < 1) return bad("I and J must be >= 1");
  if (I >= (1LL << 31) || J >= (1LL << 31)) return bad("I and J must be below 2^31");
  if (K < 1 || K > kMaxK) return bad("K must be in 1..16");
  if (Q < 1 || Q > kMaxQ) return bad("Q must be in 1..8");
  if (dtype != ADMX_F64 && dtype != ADMX_F32) return bad("dtype must be ADMX_F64 or ADMX_F32");
  admx_ctx* c = new admx_ctx();
  c->I = I;
  c->J = J;
  c->K = K;
  c->Q = Q;
  c->dtype = dtype;
  c->esz = dtype == ADMX_F64 ? 8 : 4;
  c->ibytes = (I + 3) / 4;
  c->pitch = (c->ibytes + 63) & ~(int64_t)63;
  c->stream_snps = stream_snps;
  auto cleanup = [&](int rc) {
    admx_destroy(c);
    return rc;
  };
  if (bed_path) {
    // .bed file: PLINK-1 magic, SNP-major mode, ceil(I/4) * J payload bytes
    c->bed_fd = open(bed_path, O_RDONLY | O_CLOEXEC);
    if (c->bed_fd < 0) return cleanup(fail(ADMX_EIO, std::string("cannot open ") + bed_path + ": " + std::strerror(errno)));
    struct stat st;
    uint8_t magic[3] = {0, 0, 0};
    if (fstat(c->bed_fd, &st) != 0 || pread(c->bed_fd, magic, 3, 0) != 3)
      return cleanup(fail(ADMX_EIO, std::string("cannot read ") + bed_path));
    if (magic[0] != 0x6C || magic[1] != 0x1B || magic[2] != 0x01)
      return cleanup(fail(ADMX_EINVAL, std::string(bed_path) + " is not a SNP-major PLINK .bed file"));
    if ((int64_t)st.st_size != 3 + c->ibytes * J)
      return cleanup(fail(ADMX_EINVAL, std::string(bed_path) + " does not hold 3 + ceil(I/4) * J bytes"));
  }
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) return cleanup(fail(ADMX_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e)));
  // set-up of the allocation-free entry point's scratch on this device
  if (int rc = noalloc_setup(c->device, nullptr); rc != ADMX_OK) return cleanup(rc);
  if ((e = hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking)) != hipSuccess)
    return cleanup(fail(ADMX_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
  const size_t xb = (size_t)c->N() * c->esz;
  auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, bytes < 256 ? 256 : bytes); };
  if (stream_snps && bed_path) {
    const size_t rb = (size_t)c->ibytes * (stream_snps + 16);
    for (int b = 0; b < 2; ++b) {
      if (hipHostMalloc((void**)&c->hstage[b], rb, hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(ADMX_EHIP, "hipHostMalloc failed for the .bed staging buffers"));
      if (alloc((void**)&c->draw[b], rb) != hipSuccess)
        return cleanup(fail(ADMX_EHIP, "hipMalloc failed for the .bed staging buffers"));
    }
    c->have_bed = true;
  } else if (stream_snps) {
    // BED in pinned host memory (zero, padding rows = missing code)
    const size_t hb = (size_t)c->pitch * (J + 16);
    if (hipHostMalloc((void**)&c->hbed, hb, hipHostMallocDefault) != hipSuccess)
      return cleanup(fail(ADMX_EHIP, "hipHostMalloc failed for the streamed BED"));
    std::memset(c->hbed, 0, (size_t)c->pitch * J);
    std::memset(c->hbed + (size_t)c->pitch * J, 0x55, (size_t)c->pitch * 16);
  }
  if (stream_snps) {  // two chunk buffers, a copy stream and its events
    const size_t cb = (size_t)c->pitch * (stream_snps + 16);
    if (alloc((void**)&c->dchunk[0], cb) != hipSuccess || alloc((void**)&c->dchunk[1], cb) != hipSuccess)
      return cleanup(fail(ADMX_EHIP, "hipMalloc failed for the BED chunk buffers"));
    if (hipStreamCreateWithFlags(&c->cs, hipStreamNonBlocking) != hipSuccess)
      return cleanup(fail(ADMX_EHIP, "hipStreamCreate failed (copy stream)"));
    for (int b = 0; b < 2; ++b)
      if (hipEventCreateWithFlags(&c->ev_ready[b], hipEventDisableTiming) != hipSuccess ||
          hipEventCreateWithFlags(&c->ev_free[b], hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(ADMX_EHIP, "hipEventCreate failed"));
  } else if (alloc((void**)&c->bed, (size_t)c->pitch * (J + 16)) != hipSuccess) {
    return cleanup(fail(ADMX_EHIP, "hipMalloc failed for the BED"));
  }
  if (alloc(&c->x[0], xb) != hipSuccess ||
      alloc(&c->x[1], xb) != hipSuccess || alloc(&c->x[2], xb) != hipSuccess || alloc(&c->x[3], xb) != hipSuccess ||
      alloc(&c->XtX_q, (size_t)(K * K * I) * c->esz) != hipSuccess ||
      alloc(&c->Xtz_q, (size_t)(K * I) * c->esz) != hipSuccess ||
      alloc(&c->XtX_p, (size_t)(K * K * J) * c->esz) != hipSuccess ||
      alloc(&c->Xtz_p, (size_t)(K * J) * c->esz) != hipSuccess || alloc(&c->U, xb * Q) != hipSuccess ||
      alloc(&c->V, xb * Q) != hipSuccess || alloc((void**)&c->lls, sizeof(double) * LL_NSLOTS) != hipSuccess ||
      alloc((void**)&c->status, sizeof(int32_t) * std::max(I, J)) != hipSuccess)
    return cleanup(fail(ADMX_EHIP, "hipMalloc failed for the context buffers"));
  if (hipHostMalloc((void**)&c->h_lls, sizeof(double) * LL_NSLOTS, hipHostMallocDefault) != hipSuccess)
    return cleanup(fail(ADMX_EHIP, "hipHostMalloc failed"));
  double nanv[LL_NSLOTS];
  for (int s = 0; s < LL_NSLOTS; ++s) nanv[s] = std::nan("");
  if (hipMemcpy(c->lls, nanv, sizeof(nanv), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(ADMX_EHIP, "hipMemcpy failed"));
  *out = c;
  return ADMX_OK;
}
}  // namespace

extern "C" {

int admx_create(admx_ctx** out, int64_t I, int64_t J, int64_t K, int64_t Q, int dtype) {
  return guarded([&]() -> int {
    return create_impl(out, I, J, K, Q, dtype, 0);
  });
}

int admx_create_streamed(admx_ctx** out, int64_t I, int64_t J, int64_t K, int64_t Q, int dtype,
                         int64_t snps_per_chunk) {
  return guarded([&]() -> int {
    if (snps_per_chunk < 16) return bad("a streamed context needs at least 16 SNPs per chunk");
    return create_impl(out, I, J, K, Q, dtype, std::min<int64_t>(snps_per_chunk, J));
  });
}

int admx_create_streamed_file(admx_ctx** out, const char* bed_path, int64_t I, int64_t J, int64_t K, int64_t Q,
                              int dtype, int64_t snps_per_chunk) {
  return guarded([&]() -> int {
    if (out) *out = nullptr;
    if (!bed_path) return bad("null .bed path");
    if (snps_per_chunk < 16) return bad("a streamed context needs at least 16 SNPs per chunk");
    return create_impl(out, I, J, K, Q, dtype, std::min<int64_t>(snps_per_chunk, J), bed_path);
  });
}

int admx_destroy(admx_ctx* c) {
  return guarded([&]() -> int {
    if (!c) return ADMX_OK;
    (void)hipSetDevice(c->device);
    if (c->s) (void)hipStreamSynchronize(c->s);
    for (auto& e : c->prof.pend) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
    for (auto e : c->prof.pool) (void)hipEventDestroy(e);
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
    void* bufs[] = {c->bed, c->x[0], c->x[1], c->x[2], c->x[3], c->XtX_q, c->Xtz_q, c->XtX_p,
                    c->Xtz_p, c->U,   c->V,    c->lls,  c->status};
    for (void* b : bufs)
      if (b) (void)hipFree(b);
    if (c->h_lls) (void)hipHostFree(c->h_lls);
    if (c->cs) (void)hipStreamSynchronize(c->cs);
    if (c->hbed) (void)hipHostFree(c->hbed);
    if (c->dres) (void)hipFree(c->dres);
    if (c->bed_fd >= 0) (void)close(c->bed_fd);
    for (int b = 0; b < 2; ++b) {
      if (c->hstage[b]) (void)hipHostFree(c->hstage[b]);
      if (c->draw[b]) (void)hipFree(c->draw[b]);
      if (c->dchunk[b]) (void)hipFree(c->dchunk[b]);
      if (c->ev_ready[b]) (void)hipEventDestroy(c->ev_ready[b]);
      if (c->ev_free[b]) (void)hipEventDestroy(c->ev_free[b]);
    }
    if (c->cs) (void)hipStreamDestroy(c->cs);
    c->part.release();
    c->llpart.release();
    c->zal.release();
    for (auto& b : c->st) b.release();
    c->fq.release();
    c->fp.release();
    c->ppt.release();
    c->mmt.release();
    c->qqt.release();
    c->lqt.release();
    c->lpt.release();
    c->lmt.release();
    c->lps.release();
    c->lqc.release();
    c->qimg.release();
    c->pimg.release();
    c->feat.release();
    c->featpart.release();
    c->snp_mu.release();
    c->cnt.release();
    for (auto& b : c->wide) b.release();
    c->llflag.release();
    for (DevBuf* b : {&c->sk_lab, &c->sk_cm, &c->sk_small, &c->sk_cnt, &c->sk_part, &c->sk_stat, &c->sk_E, &c->sk_key,
                      &c->sk_keys, &c->sk_idx, &c->sk_sidx, &c->sk_flag, &c->sk_sel, &c->sk_T, &c->sk_P, &c->sk_vec,
                      &c->sk_Dc, &c->sk_tmp})
      b->release();
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
    return ADMX_OK;
  });
}

int admx_upload_bed(admx_ctx* c, const uint8_t* bed, int64_t bytes_per_snp) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (!bed) return bad("null BED pointer");
    if (bytes_per_snp < c->ibytes) return bad("bytes_per_snp < ceil(I/4)");
    if (c->bed_fd >= 0) return fail(ADMX_ESTATE, "a file-backed context reads its genotypes from the .bed file");
    if (c->stream_snps) {  // host copy into the pinned BED; the padding rows stay missing
      HIPCHK(hipStreamSynchronize(c->s));  // no copy out of the pinned BED is still in flight
      HIPCHK(hipStreamSynchronize(c->cs));
      for (int64_t j = 0; j < c->J; ++j) {
        uint8_t* row = c->hbed + (size_t)c->pitch * j;
        std::memcpy(row, bed + (size_t)bytes_per_snp * j, (size_t)c->ibytes);
        std::memset(row + c->ibytes, 0, (size_t)(c->pitch - c->ibytes));
      }
      c->have_bed = true;
      c->mu_valid = false;
      forget_chunks(c);
      return refresh_resident(c);
    }
    HIPCHK(hipMemsetAsync(c->bed, 0, (size_t)c->pitch * c->J, c->s));
    CHK(pad_bed(c));
    HIPCHK(hipMemcpy2DAsync(c->bed, c->pitch, bed, bytes_per_snp, c->ibytes, c->J, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->have_bed = true;
    c->mu_valid = false;
    return ADMX_OK;
  });
}

int admx_download_bed(admx_ctx* c, uint8_t* bed) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    if (!bed) return bad("null BED pointer");
    if (c->bed_fd >= 0) return read_file_rows(c, 0, c->J, bed);
    if (c->stream_snps) {
      for (int64_t j = 0; j < c->J; ++j)
        std::memcpy(bed + (size_t)c->ibytes * j, c->hbed + (size_t)c->pitch * j, (size_t)c->ibytes);
      return ADMX_OK;
    }
    HIPCHK(hipMemcpy2DAsync(bed, c->ibytes, c->bed, c->pitch, c->ibytes, c->J, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_generate_bed(admx_ctx* c, const double* Qt, const double* Pt, uint64_t seed, double miss_rate) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (!Qt || !Pt) return bad("null Q or P");
    if (c->bed_fd >= 0) return fail(ADMX_ESTATE, "a file-backed context reads its genotypes from the .bed file");
    const size_t qb = (size_t)(c->K * c->I) * sizeof(double), pb = (size_t)(c->K * c->J) * sizeof(double);
    CHK(stage_in(c, 0, Qt, qb, qb));
    CHK(stage_in(c, 1, Pt, pb, pb));
    if (c->stream_snps) {  // chunk by chunk through a device buffer into the pinned BED
      const int64_t S = c->stream_snps;
      for (int64_t j0 = 0; j0 < c->J; j0 += S) {
        const int64_t j1 = std::min(c->J, j0 + S);
        launch_gen_bed_rows(static_cast<const double*>(c->st[0].p), static_cast<const double*>(c->st[1].p), c->I, j0,
                            j1, (int)c->K, seed, miss_rate, c->dchunk[0] - c->pitch * j0, c->pitch, c->s);
        CHK(check_launch());
        HIPCHK(hipMemcpyAsync(c->hbed + (size_t)c->pitch * j0, c->dchunk[0], (size_t)(c->pitch * (j1 - j0)),
                              hipMemcpyDeviceToHost, c->s));
      }
      HIPCHK(hipStreamSynchronize(c->s));
      c->have_bed = true;
      c->mu_valid = false;
      forget_chunks(c);
      return refresh_resident(c);
    }
    {
      ProfScope ps(c, "gen_bed");
      launch_gen_bed(static_cast<const double*>(c->st[0].p), static_cast<const double*>(c->st[1].p), c->I, c->J,
                     (int)c->K, seed, miss_rate, c->bed, c->pitch, c->s);
    }
    CHK(pad_bed(c));
    CHK(check_launch());
    HIPCHK(hipStreamSynchronize(c->s));
    c->have_bed = true;
    c->mu_valid = false;
    return ADMX_OK;
  });
}

int admx_bed_from(admx_ctx* dst, admx_ctx* src, int32_t fold, int32_t nfolds, uint64_t seed) {
  return guarded([&]() -> int {
    CHK(check_ctx(dst));
    CHK(check_ctx(src));
    if (dst == src) return bad("source and destination must be different contexts");
    Guard2 g(dst, src);
    CHK(need_bed(src));
    CHK(need_resident(src, "fold masking"));
    CHK(need_resident(dst, "fold masking"));
    if (dst->I != src->I || dst->J != src->J) return bad("contexts differ in I or J");
    if (fold >= 0 && (nfolds < 2 || fold >= nfolds)) return bad("need 0 <= fold < nfolds and nfolds >= 2");
    if (dst->device != src->device) return bad("contexts must live on the same device");
    HIPCHK(hipStreamSynchronize(src->s));
    HIPCHK(hipSetDevice(dst->device));
    {
      ProfScope ps(dst, "mask_fold");
      launch_mask_fold(src->bed, dst->bed, dst->pitch, dst->I, dst->J, fold, fold >= 0 ? nfolds : 1, seed, dst->s);
    }
    CHK(check_launch());
    CHK(pad_bed(dst));
    HIPCHK(hipStreamSynchronize(dst->s));
    dst->have_bed = true;
    dst->mu_valid = false;
    return ADMX_OK;
  });
}

int admx_heldout_deviance(admx_ctx* fit, admx_ctx* src, int32_t fold, int32_t nfolds, uint64_t seed,
                          double* deviance_sum, int64_t* count) {
  return guarded([&]() -> int {
    CHK(check_ctx(fit));
    CHK(check_ctx(src));
    Guard2 g(fit, src);
    CHK(need_bed(src));
    CHK(need_resident(src, "the held-out deviance"));
    if (!deviance_sum || !count) return bad("null output");
    if (fit->I != src->I || fit->J != src->J) return bad("contexts differ in I or J");
    if (fit->dtype != ADMX_F64) return bad("held-out deviance needs a FP64 fit context");
    if (nfolds < 2 || fold < 0 || fold >= nfolds) return bad("need 0 <= fold < nfolds and nfolds >= 2");
    if (fit->device != src->device) return bad("contexts must live on the same device");
    if (src != fit) HIPCHK(hipStreamSynchronize(src->s));
    HIPCHK(hipSetDevice(fit->device));
    int64_t jchunk = 0;
    const int64_t nb = heldout_dev_blocks(fit->I, fit->J, &jchunk);
    CHK(fit->st[0].grow(sizeof(double) * 2 * nb + 2 * sizeof(double)));
    double* part = static_cast<double*>(fit->st[0].p);
    double* tot = part + 2 * nb;
    const double* q = static_cast<const double*>(fit->x[ADMX_X]);
    {
      ProfScope ps(fit, "heldout_dev");
      launch_heldout_dev(src->bed, src->pitch, fit->I, fit->J, (int)fit->K, q, q + fit->K * fit->I, fold, nfolds, seed,
                         part, fit->s);
      launch_reduce_sum(part, nb, tot, 0, fit->s);
      launch_reduce_sum(part + nb, nb, tot + 1, 0, fit->s);
    }
    CHK(check_launch());
    double h[2];
    HIPCHK(hipMemcpyAsync(h, tot, sizeof(h), hipMemcpyDeviceToHost, fit->s));
    HIPCHK(hipStreamSynchronize(fit->s));
    *deviance_sum = h[0];
    *count = (int64_t)h[1];
    return ADMX_OK;
  });
}


int admx_decode(admx_ctx* c, int8_t* gout) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(need_resident(c, "decode"));
    if (!gout) return bad("null output");
    const size_t bytes = (size_t)(c->I * c->J);
    CHK(c->st[0].grow(bytes));
    {
      ProfScope ps(c, "decode");
      launch_decode(c->bed, c->pitch, c->I, c->J, static_cast<int8_t*>(c->st[0].p), c->s);
    }
    CHK(check_launch());
    HIPCHK(hipMemcpyAsync(gout, c->st[0].p, bytes, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_counts(admx_ctx* c, int64_t* snp_nm, int64_t* snp_alt, int64_t* ind_nm) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(need_resident(c, "counts"));
    if (!snp_nm || !snp_alt || !ind_nm) return bad("null output");
    const size_t jb = sizeof(unsigned long long) * c->J, ib = sizeof(unsigned long long) * c->I;
    CHK(c->st[0].grow(jb));
    CHK(c->st[1].grow(jb));
    CHK(c->st[2].grow(ib));
    HIPCHK(hipMemsetAsync(c->st[2].p, 0, ib, c->s));
    {
      ProfScope ps(c, "counts");
      launch_counts(c->bed, c->pitch, c->I, c->J, static_cast<unsigned long long*>(c->st[0].p),
                    static_cast<unsigned long long*>(c->st[1].p), static_cast<unsigned long long*>(c->st[2].p), c->s);
    }
    CHK(check_launch());
    HIPCHK(hipMemcpyAsync(snp_nm, c->st[0].p, jb, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(snp_alt, c->st[1].p, jb, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(ind_nm, c->st[2].p, ib, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_loglikelihood_loop(admx_ctx* c, const void* q, const void* p, int64_t i0, int64_t i1, int64_t j0,
                            int64_t j1, double* out) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(check_ranges(c, i0, i1, j0, j1));
    if (!q || !p || !out) return bad("null pointer");
    return BY_DTYPE(c, loop_ll, c, q, p, i0, i1, j0, j1, out);
  });
}

int admx_em_q_loop(admx_ctx* c, void* q_next, const void* q, const void* p, int64_t i0, int64_t i1, int64_t j0,
                   int64_t j1) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(check_ranges(c, i0, i1, j0, j1));
    if (!q_next || !q || !p) return bad("null pointer");
    return BY_DTYPE(c, loop_em_q, c, q_next, q, p, i0, i1, j0, j1);
  });
}

int admx_em_p_loop(admx_ctx* c, void* p_next, void* p_tmp, const void* q, const void* p, int64_t i0, int64_t i1,
                   int64_t j0, int64_t j1) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(check_ranges(c, i0, i1, j0, j1));
    if (!p_next || !p_tmp || !q || !p) return bad("null pointer");
    return BY_DTYPE(c, loop_em_p, c, p_next, p_tmp, q, p, i0, i1, j0, j1);
  });
}

int admx_update_q_loop(admx_ctx* c, void* XtX, void* Xtz, const void* q, const void* p, int64_t i0, int64_t i1,
                       int64_t j0, int64_t j1) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(check_ranges(c, i0, i1, j0, j1));
    if (!XtX || !Xtz || !q || !p) return bad("null pointer");
    return BY_DTYPE(c, loop_update, c, true, XtX, Xtz, q, p, i0, i1, j0, j1);
  });
}

int admx_update_p_loop(admx_ctx* c, void* XtX, void* Xtz, const void* q, const void* p, int64_t i0, int64_t i1,
                       int64_t j0, int64_t j1) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(check_ranges(c, i0, i1, j0, j1));
    if (!XtX || !Xtz || !q || !p) return bad("null pointer");
    return BY_DTYPE(c, loop_update, c, false, XtX, Xtz, q, p, i0, i1, j0, j1);
  });
}

int admx_qp_batch(admx_ctx* c, int simplex, int64_t n, const void* XtX, const void* Xtz_neg, const void* x,
                  void* delta, int32_t* status) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (n < 0) return bad("n < 0");
    if (n == 0) return ADMX_OK;
    if (!XtX || !Xtz_neg || !x || !delta) return bad("null pointer");
    return BY_DTYPE(c, host_qp_batch, c, simplex, n, XtX, Xtz_neg, x, delta, status);
  });
}

int admx_project_q(admx_ctx* c, void* b, int64_t n) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (n < 0 || (n > 0 && !b)) return bad("bad buffer");
    if (n == 0) return ADMX_OK;
    return BY_DTYPE(c, host_project, c, true, b, n);
  });
}

int admx_project_p(admx_ctx* c, void* b, int64_t n) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (n < 0 || (n > 0 && !b)) return bad("bad buffer");
    if (n == 0) return ADMX_OK;
    return BY_DTYPE(c, host_project, c, false, b, n);
  });
}

int admx_zal_step(admx_ctx* c, int64_t iter, const void* x, const void* xn, const void* xn2, void* U, void* V,
                  void* x_tmp) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (iter < 1) return bad("iter must be >= 1");
    if (!x || !xn || !xn2 || !U || !V || !x_tmp) return bad("null pointer");
    return BY_DTYPE(c, host_zal, c, iter, x, xn, xn2, U, V, x_tmp);
  });
}

int admx_lbqn_step(admx_ctx* c, int64_t iter, const void* x, const void* xn, const void* xn2, void* U, void* V,
                   void* x_tmp) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (iter < 1) return bad("iter must be >= 1");
    if (!x || !xn || !xn2 || !U || !V || !x_tmp) return bad("null pointer");
    return BY_DTYPE(c, host_lbqn, c, iter, x, xn, xn2, U, V, x_tmp);
  });
}

int admx_set_x(admx_ctx* c, int which, const void* host) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (which < 0 || which > 3 || !host) return bad("bad state vector");
    HIPCHK(hipMemcpyAsync(c->x[which], host, (size_t)c->N() * c->esz, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_copy_x(admx_ctx* c, int dst, int src) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (dst < 0 || dst > 3 || src < 0 || src > 3) return bad("bad state vector");
    if (dst != src)
      HIPCHK(hipMemcpyAsync(c->x[dst], c->x[src], (size_t)c->N() * c->esz, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_get_x(admx_ctx* c, int which, void* host) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (which < 0 || which > 3 || !host) return bad("bad state vector");
    HIPCHK(hipMemcpyAsync(host, c->x[which], (size_t)c->N() * c->esz, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_set_uv(admx_ctx* c, const void* U, const void* V) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (!U || !V) return bad("null pointer");
    const size_t b = (size_t)c->N() * c->Q * c->esz;
    HIPCHK(hipMemcpyAsync(c->U, U, b, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->V, V, b, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_get_uv(admx_ctx* c, void* U, void* V) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (!U || !V) return bad("null pointer");
    const size_t b = (size_t)c->N() * c->Q * c->esz;
    HIPCHK(hipMemcpyAsync(U, c->U, b, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(V, c->V, b, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_set_ll(admx_ctx* c, double ll_prev, double ll_new) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    double v[2] = {ll_prev, ll_new};
    HIPCHK(hipMemcpyAsync(c->lls + LL_PREV, v, sizeof(v), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_get_ll(admx_ctx* c, double* ll_prev, double* ll_new) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    HIPCHK(hipMemcpyAsync(c->h_lls, c->lls, sizeof(double) * LL_NSLOTS, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (ll_prev) *ll_prev = c->h_lls[LL_PREV];
    if (ll_new) *ll_new = c->h_lls[LL_NEW];
    return ADMX_OK;
  });
}

int admx_get_hessians(admx_ctx* c, int side, void* XtX, void* Xtz) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (!XtX && !Xtz) return bad("null pointer");
    if (side != ADMX_SIDE_Q && side != ADMX_SIDE_P) return bad("side must be ADMX_SIDE_Q or ADMX_SIDE_P");
    const int64_t n = side == ADMX_SIDE_Q ? c->I : c->J;
    if (XtX)
      HIPCHK(hipMemcpyAsync(XtX, side == ADMX_SIDE_Q ? c->XtX_q : c->XtX_p, (size_t)(c->K * c->K * n) * c->esz,
                            hipMemcpyDeviceToHost, c->s));
    if (Xtz)
      HIPCHK(hipMemcpyAsync(Xtz, side == ADMX_SIDE_Q ? c->Xtz_q : c->Xtz_p, (size_t)(c->K * n) * c->esz,
                            hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_set_block(admx_ctx* c, int which, int side, const void* host) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (which < 0 || which > 3 || !host) return bad("bad state vector");
    if (side != ADMX_SIDE_Q && side != ADMX_SIDE_P) return bad("side must be ADMX_SIDE_Q or ADMX_SIDE_P");
    void* X = c->x[which];
    void* dst = side == ADMX_SIDE_Q ? X : c->pp(X);
    const size_t b = (size_t)(c->K * (side == ADMX_SIDE_Q ? c->I : c->J)) * c->esz;
    HIPCHK(hipMemcpyAsync(dst, host, b, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_get_block(admx_ctx* c, int which, int side, void* host) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (which < 0 || which > 3 || !host) return bad("bad state vector");
    if (side != ADMX_SIDE_Q && side != ADMX_SIDE_P) return bad("side must be ADMX_SIDE_Q or ADMX_SIDE_P");
    void* X = c->x[which];
    const void* src = side == ADMX_SIDE_Q ? X : c->pp(X);
    const size_t b = (size_t)(c->K * (side == ADMX_SIDE_Q ? c->I : c->J)) * c->esz;
    HIPCHK(hipMemcpyAsync(host, src, b, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_copy_block(admx_ctx* c, int dst, int src, int side) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (dst < 0 || dst > 3 || src < 0 || src > 3) return bad("bad state vector");
    if (side != ADMX_SIDE_Q && side != ADMX_SIDE_P) return bad("side must be ADMX_SIDE_Q or ADMX_SIDE_P");
    if (dst != src) {
      char* d = static_cast<char*>(side == ADMX_SIDE_Q ? c->x[dst] : c->pp(c->x[dst]));
      const char* s = static_cast<const char*>(side == ADMX_SIDE_Q ? c->x[src] : c->pp(c->x[src]));
      const size_t b = (size_t)(c->K * (side == ADMX_SIDE_Q ? c->I : c->J)) * c->esz;
      HIPCHK(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, c->s));
    }
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_grad_hessian(admx_ctx* c, int side) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    if (side != ADMX_SIDE_Q && side != ADMX_SIDE_P) return bad("side must be ADMX_SIDE_Q or ADMX_SIDE_P");
    CHK(BY_DTYPE(c, dev_grad_hessian, c, side == ADMX_SIDE_Q));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_loglikelihood(admx_ctx* c, int which, double* out) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    if (which < 0 || which > 3 || !out) return bad("bad state vector");
    void* X = c->x[which];
    CHK(BY_DTYPE(c, ll_to_slot, c, X, c->pp(X), 0, c->I, 0, c->J, LL_TMP));
    HIPCHK(hipMemcpyAsync(out, c->lls + LL_TMP, sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_loglikelihood2(admx_ctx* c, int which1, int which2, double* out) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    if (which1 < 0 || which1 > 3 || which2 < 0 || which2 > 3 || !out) return bad("bad state vector");
    void* X1 = c->x[which1];
    void* X2 = c->x[which2];
    if (ll_pairs(c)) {
      CHK(BY_DTYPE(c, ll2_to_slots, c, X1, LL_TMP, X2, LL_TMP2));
    } else {
      CHK(BY_DTYPE(c, ll_to_slot, c, X1, c->pp(X1), 0, c->I, 0, c->J, LL_TMP));
      CHK(BY_DTYPE(c, ll_to_slot, c, X2, c->pp(X2), 0, c->I, 0, c->J, LL_TMP2));
    }
    HIPCHK(hipMemcpyAsync(out, c->lls + LL_TMP, 2 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_em_q(admx_ctx* c) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(BY_DTYPE(c, dev_em_q, c));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_em_p(admx_ctx* c) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(BY_DTYPE(c, dev_em_p, c));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_em_q_mode(admx_ctx* c, int mode) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    if (mode < ADMX_EM_BASE || mode > ADMX_EM_FALLBACK2) return bad("mode must be ADMX_EM_BASE, _FALLBACK or _FALLBACK2");
    CHK(BY_DTYPE(c, dev_em_q, c, mode));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_em_p_mode(admx_ctx* c, int mode) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    if (mode < ADMX_EM_BASE || mode > ADMX_EM_FALLBACK2) return bad("mode must be ADMX_EM_BASE, _FALLBACK or _FALLBACK2");
    CHK(BY_DTYPE(c, dev_em_p, c, mode));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_init_em(admx_ctx* c, int64_t iters, double* lls) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    if (iters < 0) return bad("iters < 0");
    return BY_DTYPE(c, dev_init_em, c, iters, lls);
  });
}

int admx_update_q(admx_ctx* c, int update2, int32_t* status) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(BY_DTYPE(c, dev_update_q, c, update2 != 0, status));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_update_p(admx_ctx* c, int update2, int32_t* status) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    CHK(BY_DTYPE(c, dev_update_p, c, update2 != 0, status));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_qn_iteration(admx_ctx* c, int64_t iter, int fix_q, int fix_p, double* out) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(need_bed(c));
    if (iter < 1 || !out) return bad("iter must be >= 1 and out non-null");
    return BY_DTYPE(c, dev_qn_iteration, c, iter, fix_q, fix_p, out);
  });
}

int admx_synchronize(admx_ctx* c) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    HIPCHK(hipStreamSynchronize(c->s));
    return ADMX_OK;
  });
}

int admx_profile_enable(admx_ctx* c, int on) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(prof_collect(c));
    c->prof.on = on != 0;
    return ADMX_OK;
  });
}

int admx_profile_reset(admx_ctx* c) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(prof_collect(c));
    c->prof.acc.clear();
    return ADMX_OK;
  });
}

int admx_profile_report(admx_ctx* c, char* buf, size_t len) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    CHK(prof_collect(c));
    std::string s;
    char line[256];
    for (auto& kv : c->prof.acc) {
      std::snprintf(line, sizeof(line), "%s %ld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
      s += line;
    }
    if (!buf || len == 0) return bad("null buffer");
    if (s.size() + 1 > len) return bad("buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return ADMX_OK;
  });
}

int admx_set_tuning(admx_ctx* c, int key, int64_t value) {
  return guarded([&]() -> int {
    CHK(check_ctx(c));
    Guard g(c);
    if (key == ADMX_TUNE_TARGET_WAVES) {
      if (value < 1) return bad("target waves must be >= 1");
      c->target_waves = value;
    } else if (key == ADMX_TUNE_KEEP_HESSIANS) {
      c->keep_hess = value != 0;
    } else if (key == ADMX_TUNE_PRECISION) {
      if (value < ADMX_PREC_F64 || value > ADMX_PREC_BF16) return bad("precision must be ADMX_PREC_F64/F32/BF16");
      if (value != ADMX_PREC_F64 && c->dtype != ADMX_F64) return bad("mixed precision needs a FP64 context");
      c->precision = (int)value;
    } else if (key == ADMX_TUNE_SPLIT_Q || key == ADMX_TUNE_SPLIT_P) {
      if (value < -1 || value > 1) return bad("Hessian split must be -1 (default), 0 or 1");
      (key == ADMX_TUNE_SPLIT_Q ? c->split_q : c->split_p) = (int)value;
    } else if (key == ADMX_TUNE_QSIDE_MFMA || key == ADMX_TUNE_PSIDE_MFMA) {
      if (value < -1 || value > 1) return bad("matrix-core update pass must be -1 (default), 0 or 1");
      (key == ADMX_TUNE_QSIDE_MFMA ? c->qside_mfma : c->pside_mfma) = (int)value;
    } else if (key == ADMX_TUNE_SKIPMISSING) {
      c->impute = value == 0;
    } else if (key == ADMX_TUNE_QN_MODE) {
      if (value != ADMX_QN_ZAL && value != ADMX_QN_LBQN) return bad("quasi-Newton mode must be ADMX_QN_ZAL or ADMX_QN_LBQN");
      c->qn_mode = (int)value;
    } else if (key == ADMX_TUNE_GRAPHS) {
      c->use_graphs = value != 0;
      if (!c->use_graphs) drop_graphs(c);
    } else if (key == ADMX_TUNE_RESIDENT_SNPS) {
      if (!c->stream_snps) return bad("resident SNPs apply to a streamed context");
      if (value < -1) return bad("resident SNPs must be >= 0, or -1 for as many as fit");
      HIPCHK(hipStreamSynchronize(c->s));
      if (c->dres) HIPCHK(hipFree(c->dres));
      c->dres = nullptr;
      c->res_snps = 0;
      int64_t want = value;
      if (want == -1) {  // as many rows as the free HBM holds beyond a reserve for pass scratch
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        const size_t reserve = std::max<size_t>((size_t)4 << 30, tot / 10);
        want = fr > reserve ? std::max<int64_t>((int64_t)((fr - reserve) / (size_t)c->pitch) - 16, 0) : 0;
      }
      const int64_t R = std::min(want, c->J);
      if (R > 0) {
        if (hipMalloc((void**)&c->dres, (size_t)c->pitch * (R + 16)) != hipSuccess) {
          (void)hipGetLastError();
          return fail(ADMX_EHIP, "hipMalloc failed for the resident SNPs");
        }
        c->res_snps = R;
        const int rc = refresh_resident(c);
        if (rc != ADMX_OK) {
          (void)hipFree(c->dres);
          c->dres = nullptr;
          c->res_snps = 0;
          return rc;
        }
      }
    } else {
      return bad("unknown tuning key");
    }
    return ADMX_OK;
  });
}

}  // extern "C"
