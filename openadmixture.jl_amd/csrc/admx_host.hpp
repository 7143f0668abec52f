// admx_host.hpp -- host-side state and helpers shared by the translation units
// of libadmixhip's C ABI (admx_capi.cpp, admx_skfr_host.cpp,
// admx_warm_host.cpp): the context and network structs, error reporting, the
// exception guard of every entry point, device buffers, profiling scopes and
// locks.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "admixhip.h"
#include "admx_internal.hpp"

namespace admx_host {

// Message of the calling thread's last failed call (admx_last_error).
extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Host-side exceptions (allocation failure, thread creation) never cross the C
// ABI: every entry point runs inside guarded(), which turns them into
// ADMX_EHOST with the message in admx_last_error().
inline int fail_nothrow(const char* msg) noexcept {
  try {
    g_err = msg;
  } catch (...) {
    g_err.clear();
  }
  return ADMX_EHOST;
}

template <typename F>
inline int guarded(F&& f) noexcept {
  try {
    return f();
  } catch (const std::exception& e) {
    return fail_nothrow(e.what());
  } catch (...) {
    return fail_nothrow("unknown host-side exception");
  }
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return fail(ADMX_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHK(expr)            \
  do {                       \
    int r_ = (expr);         \
    if (r_ != ADMX_OK) return r_; \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  int grow(size_t bytes) {
    if (bytes <= n) return ADMX_OK;
    if (p) HIPCHK(hipFree(p));
    p = nullptr;
    n = 0;
    const size_t want = bytes < 256 ? 256 : bytes;
    HIPCHK(hipMalloc(&p, want));
    n = want;
    return ADMX_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct ProfEv {
  const char* name;
  hipEvent_t a, b;
};

struct Prof {
  bool on = false;
  std::vector<ProfEv> pend;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<long, double>> acc;
};

}  // namespace admx_host

struct admx_ctx {
  int64_t I = 0, J = 0, K = 0, Q = 0;
  int dtype = ADMX_F64;
  size_t esz = 8;
  int64_t ibytes = 0, pitch = 0;
  int device = 0;
  hipStream_t s = nullptr;
  uint8_t* bed = nullptr;
  bool have_bed = false;
  void* x[4] = {nullptr, nullptr, nullptr, nullptr};
  void *XtX_q = nullptr, *Xtz_q = nullptr, *XtX_p = nullptr, *Xtz_p = nullptr;
  void *U = nullptr, *V = nullptr;
  double* lls = nullptr;
  double* h_lls = nullptr;
  int32_t* status = nullptr;
  admx_host::DevBuf part, llpart, zal;
  admx_host::DevBuf st[8];  // staging for host-array entry points
  int64_t target_waves = 65536;
  int split_q = -1, split_p = -1;  // ADMX_TUNE_SPLIT_Q / _P (-1: the default K thresholds)
  int qside_mfma = -1;             // ADMX_TUNE_QSIDE_MFMA (-1: on from K = 7)
  int pside_mfma = -1;             // ADMX_TUNE_PSIDE_MFMA (-1: on from K = 7)
  int keep_hess = 1;
  int precision = ADMX_PREC_F64;
  admx_host::DevBuf fq, fp, ppt, mmt, qqt;  // mixed-precision pass inputs
  admx_host::DevBuf lqt, lpt, lmt;           // BF16 split tables of the matrix-core log-likelihood
  admx_host::DevBuf lps, lqc;                // BF16 form's EM: two-term split rows of q or p, 1 - sum_k q_ik
  admx_host::DevBuf qimg;                    // BF16 form's matrix-core update_q: 32-SNP block images of p
  admx_host::DevBuf pimg;                    // ... and update_p: 32-individual block images of q
  admx_host::DevBuf feat, featpart;          // warm-start features (I x 64) and their chunk partials
  // J-chunk streaming: the BED lives in pinned host memory, or in a PLINK .bed
  // file read chunk by chunk, and every pass streams it through two device
  // chunk buffers on a copy stream.  The first res_snps SNPs can stay in HBM.
  int64_t stream_snps = 0;        // SNPs per chunk (0: the BED is resident in HBM)
  uint8_t* hbed = nullptr;        // pinned host BED, pitch * (J + 16)
  uint8_t* dchunk[2] = {nullptr, nullptr};
  hipStream_t cs = nullptr;       // copy stream (default priority)
  hipEvent_t ev_ready[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
  bool ev_used[2] = {false, false};  // ev_ready[b] has been recorded
  int64_t held[2][2] = {{0, 0}, {0, 0}};  // rows [held[b][0], held[b][1]) of the BED are in dchunk[b]
  int last_buf = 1;               // chunk buffer of the latest streamed chunk
  bool stream_rev = true;         // the latest pass ran its streamed chunks backwards
  int bed_fd = -1;                // file-backed: the .bed file (rows of ibytes at offset 3)
  uint8_t* hstage[2] = {nullptr, nullptr};  // file-backed: pinned raw rows, ibytes * (S + 16)
  uint8_t* draw[2] = {nullptr, nullptr};    // file-backed: the same rows in HBM, before re-pitching
  int64_t res_snps = 0;           // SNPs [0, res_snps) resident in `dres`
  uint8_t* dres = nullptr;        // pitch * (res_snps + 16)
  int impute = 0;                 // 1: mean-imputation loops (skipmissing = false)
  bool mu_valid = false;          // mu holds the per-SNP means of the current genotypes
  admx_host::DevBuf snp_mu, cnt;             // per-SNP means (FP64, J); count scratch
  admx_host::DevBuf wide[4];                 // FP32 context: FP64 copies of a pass's q and p (and a second state's)
  admx_host::DevBuf llflag;                  // two ints: log-likelihood state checks of a pass's two states
  int64_t feat_rows = 0;          // individuals with features in `feat` (0: none)
  // sparse k-means feature ranking scratch: labels, cluster masks, small
  // counters, SNP counts, centers-pass partials, SNP stats (mu, v, union key),
  // E, keys and sorted keys, index and sorted index, flags, distance tables,
  // table-sum partials, per-individual vectors (x2, dmin, D2, WSS), k-means++
  // distances, sort scratch
  admx_host::DevBuf sk_lab, sk_cm, sk_small, sk_cnt, sk_part, sk_stat, sk_E, sk_key, sk_keys, sk_idx, sk_sidx, sk_flag, sk_sel,
      sk_T, sk_P, sk_vec, sk_Dc, sk_tmp;
  std::mutex mu;
  admx_host::Prof prof;
  // hipGraph replay of steady-state quasi-Newton iterations (iter > Q): one
  // executable graph per (history column, fix_q, fix_p), valid while the
  // buffer layout signature graph_sig holds; graph_warm has bit (fix_q | 2 fix_p)
  // set once an eager iteration of that kind ran under the signature.
  int use_graphs = 0;
  int qn_mode = ADMX_QN_ZAL;
  struct GraphEntry {
    int64_t key;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  uint64_t graph_sig = 0;
  unsigned graph_warm = 0;

  int64_t N() const { return K * (I + J); }
  void* qp(void* v) const { return v; }
  void* pp(void* v) const { return static_cast<char*>(v) + (size_t)(K * I) * esz; }
};

// The warm-start network: FP32 parameters, gradients and Adam moments as flat
// vectors [W1 (hid x in), b1, W2 (hid x hid), b2, W3 (out x hid), b3], plus
// training data and per-batch activations.
struct admx_mlp {
  int in = 0, hid = 0, out = 0;
  int device = 0;
  hipStream_t s = nullptr;
  int64_t P = 0;
  int64_t step = 0;
  admx_host::DevBuf w, g, m, v;
  admx_host::DevBuf X, T, idx;  // training set (n x in, n x out) and the epoch's row order
  int64_t n = 0;
  admx_host::DevBuf act;        // Xb, Tb, H1, H2, Z, Y, dZ, dH2, dH1, rowloss, loss slots for a batch
  int64_t act_rows = 0;
  std::mutex mu;
  int64_t oW1() const { return 0; }
  int64_t ob1() const { return (int64_t)hid * in; }
  int64_t oW2() const { return ob1() + hid; }
  int64_t ob2() const { return oW2() + (int64_t)hid * hid; }
  int64_t oW3() const { return ob2() + hid; }
  int64_t ob3() const { return oW3() + (int64_t)out * hid; }
};

namespace admx_host {

inline hipEvent_t ev_get(admx_ctx* c) {
  if (!c->prof.pool.empty()) {
    hipEvent_t e = c->prof.pool.back();
    c->prof.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

struct ProfScope {
  admx_ctx* c;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(admx_ctx* c_, const char* n) : c(c_), name(n) {
    if (c->prof.on) {
      a = ev_get(c);
      b = ev_get(c);
      (void)hipEventRecord(a, c->s);
    }
  }
  ~ProfScope() {
    if (c->prof.on) {
      (void)hipEventRecord(b, c->s);
      c->prof.pend.push_back({name, a, b});
    }
  }
};

inline int prof_collect(admx_ctx* c) {
  if (c->prof.pend.empty()) return ADMX_OK;
  HIPCHK(hipStreamSynchronize(c->s));
  for (auto& e : c->prof.pend) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
    auto& slot = c->prof.acc[e.name];
    slot.first += 1;
    slot.second += ms;
    c->prof.pool.push_back(e.a);
    c->prof.pool.push_back(e.b);
  }
  c->prof.pend.clear();
  return ADMX_OK;
}

// The 16 SNP rows after the last one hold the missing code (01 in every field):
// passes that step 16 SNPs at a time read them in the cohort's last step.
inline int pad_bed(admx_ctx* c) {
  HIPCHK(hipMemsetAsync(c->bed + (size_t)c->pitch * c->J, 0x55, (size_t)c->pitch * 16, c->s));
  return ADMX_OK;
}

inline int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ADMX_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return ADMX_OK;
}

inline int bad(const char* msg) { return fail(ADMX_EINVAL, msg); }

inline int check_ctx(admx_ctx* c) {
  if (!c) return bad("null context");
  HIPCHK(hipSetDevice(c->device));
  return ADMX_OK;
}

inline int check_ranges(admx_ctx* c, int64_t i0, int64_t i1, int64_t j0, int64_t j1) {
  if (i0 < 0 || i1 < i0 || i1 > c->I || j0 < 0 || j1 < j0 || j1 > c->J)
    return bad("range outside [0,I) x [0,J) or reversed");
  return ADMX_OK;
}

inline int need_bed(admx_ctx* c) {
  if (!c->have_bed) return fail(ADMX_ESTATE, "no genotypes: call admx_upload_bed or admx_generate_bed first");
  return ADMX_OK;
}

inline int need_resident(admx_ctx* c, const char* what) {
  if (c->stream_snps) return fail(ADMX_ESTATE, std::string(what) + " needs a resident BED (not a streamed context)");
  return ADMX_OK;
}

inline int stage_in(admx_ctx* c, int slot, const void* host, size_t bytes, size_t cap) {
  CHK(c->st[slot].grow(cap));
  if (bytes) HIPCHK(hipMemcpyAsync(c->st[slot].p, host, bytes, hipMemcpyHostToDevice, c->s));
  return ADMX_OK;
}

inline uint64_t host_smix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

#define BY_DTYPE(c, fn, ...) ((c)->dtype == ADMX_F64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

struct Guard {
  std::unique_lock<std::mutex> lk;
  explicit Guard(admx_ctx* c) : lk(c->mu) {}
};

// Both contexts' locks, taken in address order (or one lock when a == b).
struct Guard2 {
  std::unique_lock<std::mutex> l1, l2;
  Guard2(admx_ctx* a, admx_ctx* b) {
    admx_ctx* lo = a < b ? a : b;
    admx_ctx* hi = a < b ? b : a;
    l1 = std::unique_lock<std::mutex>(lo->mu);
    if (hi != lo) l2 = std::unique_lock<std::mutex>(hi->mu);
  }
};

struct MGuard {
  std::unique_lock<std::mutex> lk;
  explicit MGuard(admx_mlp* m) : lk(m->mu) {}
};

}  // namespace admx_host
