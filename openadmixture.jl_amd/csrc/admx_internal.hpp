// admx_internal.hpp -- host-side declarations shared by the translation units
// of libadmixhip (launchers of the CDNA4 kernels).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace admx {

constexpr int kMaxK = 16;  // populations supported (templated kernels)
constexpr int kMaxQ = 8;   // quasi-Newton history columns supported on device
constexpr int kFeat = 64;  // genotype-projection feature dimension of the warm start

enum PassMode : int { MODE_UPDATE = 0, MODE_EM = 1, MODE_LL = 2 };

// Arguments of one genotype pass over [i0,i1) x [j0,j1).
struct PassArgs {
  const uint8_t* bed;
  int64_t pitch;   // bytes per SNP column (multiple of 16)
  int64_t I;       // total individuals (bound for clamped loads)
  const void* q;   // K x I
  const void* p;   // K x J
  int64_t i0, i1, j0, j1;
  int64_t chunk;   // SNPs per chunk (Q side) / individuals per chunk, multiple of 16 (P side)
  void* part;      // partial sums [nchunk][NA][n] (UPDATE/EM) or per-block doubles (LL)
  const void* mu = nullptr;  // per-SNP means (FP64, J): mean-imputation loops; nullptr: skipmissing
  int64_t ldp = 0;           // P side: column stride of the partials (0: j1 - j0)
  int64_t col0 = 0;          // P side: column of SNP j0 in the partials
  // FP64 log-likelihood: device flag set when some s = q.p may be 0 or 1 (see
  // launch_ll_state_check); the pass then sums g log s + (2-g) log(1-s) term by
  // term instead of the product form.  nullptr: product form.
  const int* strict = nullptr;
  // second state of the fused two-state log-likelihood pass (launch_qside_ll2):
  // its q, p, partials and state flag; strict_other (product-form instance)
  // names the other state's flag: the instance runs only where that flag is set
  const void* q2 = nullptr;
  const void* p2 = nullptr;
  void* part2 = nullptr;
  const int* strict2 = nullptr;
  const int* strict_other = nullptr;
  // FP64 update passes, K >= 9: 1 runs the Hessian triangle in two launches of
  // half the accumulators each, 0 in one; -1 the default (from K = 13 on both
  // sides).  Results are identical either way.
  int split = -1;
};

// Per-column reduction of pass partials, optional QP + projection.
struct FinArgs {
  const void* part;
  int64_t n;        // columns in this pass (Ilen or Jlen)
  int nchunk;
  int64_t off;      // global index of the first column (i0 or j0)
  int64_t J;        // total SNPs (EM-q divisor 2J)
  const void* x_in; // K x ncols parameters of the pass (q for Q side, p for P side)
  void* x_out;      // K x ncols output (q_next / p_next) or EM output
  void* aux;        // EM-p: p_tmp output (loop API)
  void* XtX;        // K x K x ncols (nullable)
  void* Xtz;        // K x ncols (nullable)
  int accumulate;   // loop API: outputs += partial sums
  int neg_xtz;      // store -Xtz (update_q!/update_p! negate Xtz before the QP)
  int32_t* status;  // QP return codes (nullable)
  // K x K x ncols / K x ncols buffers the split reduce-then-QP path may use
  // when XtX / Xtz are null (the context's own Hessian arrays)
  void* XtX_scratch = nullptr;
  void* Xtz_scratch = nullptr;
};

// Number of per-column accumulators of a pass.
inline int pass_na(int mode, bool qside, int K, bool f64) {
  const int tri = K * (K + 1) / 2;
  if (mode == MODE_EM) return qside ? (f64 ? K + 1 : K) : 2 * K;  // FP64 EM-q: sum (c-d) p and sum d
  if (mode == MODE_UPDATE) {
    if (qside) return f64 ? tri + K + 2 : tri + K;  // FP64: W triangle, sum b p, sum b, sum d
    return tri + K;
  }
  return 1;
}

// Genotype passes (FP64 kernels; instantiated for T = double only).
template <typename T> void launch_qside(int mode, int K, const PassArgs& a, int nchunk, hipStream_t s);
template <typename T> void launch_pside(int mode, int K, const PassArgs& a, int nchunk, hipStream_t s);
// FP32 update passes of the FP32 form (FP32 direct-form partials).
void launch_qside_f32_update(int K, const PassArgs& a, int nchunk, hipStream_t s);
void launch_pside_f32_update(int K, const PassArgs& a, int nchunk, hipStream_t s);

// Q side: update finalize (+QP simplex +project_q); EM-q finalize.
template <typename T> void launch_fin_q_update(int K, const FinArgs& f, bool do_qp, hipStream_t s);
template <typename T> void launch_fin_q_em(int K, const FinArgs& f, hipStream_t s);
// P side: update finalize (+QP box +project_p); EM-p finalize.
template <typename T> void launch_fin_p_update(int K, const FinArgs& f, bool do_qp, hipStream_t s);
template <typename T> void launch_fin_p_em(int K, const FinArgs& f, hipStream_t s);

// Mixed precision: FP64 finalize (+QP) of FP32 partials in the direct Hessian form.
void launch_fin_q_update_mixed(int K, const FinArgs& f, hipStream_t s);
void launch_fin_p_update_mixed(int K, const FinArgs& f, hipStream_t s);
// Mixed precision passes (FP32 weights/gradient, BF16 MFMA Hessian) and their inputs.
void launch_to_float(const double* x, int64_t n, float* y, hipStream_t s);
// BF16 pair tables PP (and MM) [pair][column] of x, the B operands of the MFMA passes.
void launch_pair_tables(int K, const double* x, int64_t n, int64_t ld, void* PP, void* MM, hipStream_t s);
int pair_table_rows(int K);
void launch_qside_bf16(int K, const PassArgs& a, int nchunk, const float* fq, const float* fp, const void* PPt,
                       const void* MMt, int64_t ld, hipStream_t s);
void launch_pside_bf16(int K, const PassArgs& a, int nchunk, const float* fq, const float* fp, const void* QQt,
                       int64_t ld, hipStream_t s);
// Log-likelihood with s = q.p and t = q.(1-p) + (1 - sum q) on the matrix cores
// (three-term BF16 splits, FP32 accumulation), products and reduction in FP64: tables of
// split_table_cols(K) BF16 per row (individuals: role 0; SNPs: role 1, T2 from
// 1 - p), then one partial per (256-individual block, SNP chunk).
int split_table_cols(int K);
void launch_split_tables(int K, const double* x, int64_t n, int64_t rows, int role, void* T1, void* T2,
                         hipStream_t s);
// skip (device, nullable): when *skip != 0 every workgroup returns at once (the
// state check flagged s = 0 or 1; launch_qside_ll_strict then sums the terms).
void launch_ll_mfma(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk, int nchunk,
                    const double* q, const void* QT, const void* PT, const void* MT, double* part, const int* skip,
                    hipStream_t s);
// FRAPPE EM sums of the BF16 form on the matrix cores: s and t as for the
// log-likelihood, the per-genotype weights in FP32, and the sums over SNPs
// (em_q) or individuals (em_p) as BF16 products against the two-term split
// rows of p and 1 - p, or of q ([32][ld], rows 0..K-1 the leading BF16 term,
// rows 16..16+K-1 the residual), accumulated in FP32 and written as the FP64
// partials of the FP64 passes (K+1 per individual, 2K per SNP, per chunk).
void launch_em_split(int K, const double* x, int64_t n, int64_t ld, void* S, void* S2, float* qc, hipStream_t s);
void launch_emq_mfma(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk, int nchunk,
                     const double* q, const void* QT, const void* PT, const void* MT, const void* PS, const void* MS,
                     int64_t ldps, double* part, hipStream_t s);
void launch_emp_mfma(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t ichunk, int nchunk,
                     const float* qc, const void* QT, const void* PT, const void* MT, const void* QS, int64_t ldqs,
                     double* part, hipStream_t s);

// BF16 form's update_q pass with s, t, the Hessian and the gradient on the
// matrix cores: ceil(J/32) + 1 image slots (qimg_bytes(K) bytes each, built
// from p by launch_qimg) of the SNP-side operands in LDS layout; QT is the individual
// split table of launch_split_tables (role 0).  jchunk: a multiple of 32.
// Partials [chunk][K(K+1)/2 + K][I] in FP32, as launch_qside_bf16 writes them.
int qimg_bytes(int K, bool em = false);
void launch_qimg(int K, const double* p, int64_t J, void* img, bool em, hipStream_t s);
void launch_qside_mfma(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk, int nchunk,
                       const void* QT, const void* img, float* part, hipStream_t s);

// em_q! on the same design (em = true images: split rows of p and 1 - p, no pair
// rows): sum_j c p_k + d (1 - p_k) into FP64 partials [chunk][K+1][I] (row K: 0),
// those of the FP64 EM pass; q the FP64 state (t's 1 - sum_k q_ik).
void launch_emq_img(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t jchunk, int nchunk,
                    const double* q, const void* QT, const void* img, double* part, hipStream_t s);

// The P-side mirror: ceil(I/32) + 1 individual-side image slots (pimg_bytes(K)
// each, from q by launch_pimg), PT the SNP split table of launch_split_tables
// (role 1); ichunk a multiple of 32; partials [chunk][K(K+1)/2 + K][J] in FP32.
int pimg_bytes(int K);
void launch_pimg(int K, const double* q, int64_t I, void* img, hipStream_t s);
void launch_pside_mfma(int K, const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t ichunk, int nchunk,
                       const void* PT, const void* img, float* part, hipStream_t s);

// Batched QP on host-provided tableau inputs (XtX, negated Xtz, x): delta out.
template <typename T> void launch_qp_batch(int K, bool simplex, int64_t n, const void* XtX, const void* Xtz_neg,
                                           const void* x, void* delta, int32_t* status, hipStream_t s);
template <typename T> void launch_project_q(int K, int64_t n, void* b, hipStream_t s);
template <typename T> void launch_project_p(int64_t n, void* b, hipStream_t s);

// LL partial reduction: sum nparts doubles (fixed order) into *out (device).
void launch_reduce_sum(const double* parts, int64_t nparts, double* out, int accumulate, hipStream_t s);
// The same, done only when (*flag != 0) == want (device flag).
void launch_reduce_sum_if(const double* parts, int64_t nparts, double* out, const int* flag, int want,
                          hipStream_t s);
// FP64 log-likelihood pass, term-by-term instance only, run when *a.strict != 0
// (a.strict non-null): the fallback of a product-form or matrix-core pass.
void launch_qside_ll_strict(int K, const PassArgs& a, int nchunk, hipStream_t s);

// FP64 log-likelihoods of (a.q, a.p) and (a.q2, a.p2) (partials at a.part and
// a.part2, laid out as one single-state pass lays them out), each bit-identical
// to its single-state pass: up to K = 8 one fused pass over the genotypes for
// both product forms, from K = 9 on two passes.  a.strict and a.strict2 must be
// the two states' check flags: a flagged state is summed by the single-state
// instances.
void launch_qside_ll2(int K, const PassArgs& a, int nchunk, hipStream_t s);

// ZAL quasi-Newton (src/quasi_newton.jl:5-31) on flat x of length N.
template <typename T> void launch_zal(int K, int64_t I, int64_t J, const void* x, const void* xn, const void* xn2,
                                      void* U, void* V, void* x_out, int col, int m, int Qcols,
                                      double* scratch, int64_t scratch_len, hipStream_t s);
// LBQN quasi-Newton (src/quasi_newton.jl:15-58); the caller shifts the history first.
template <typename T> void launch_lbqn(int K, int64_t I, int64_t J, const void* x, const void* xn, const void* xn2,
                                       void* U, void* V, void* x_out, int m, double* scratch, int64_t scratch_len,
                                       hipStream_t s);
// x <- (lls[prev] < lls[qn]) ? x_tmp : x_next2 ; lls[new] accordingly.
template <typename T> void launch_accept(int64_t N, void* x, const void* x_tmp, const void* x_next2, double* lls,
                                         hipStream_t s);

// BED decode / counts.
void launch_decode(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int8_t* g, hipStream_t s);
void launch_counts(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, unsigned long long* snp_nm,
                   unsigned long long* snp_alt, unsigned long long* ind_nm, hipStream_t s);
// Per-SNP counts only.
void launch_counts_snp(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, unsigned long long* snp_nm,
                       unsigned long long* snp_alt, hipStream_t s);
// Per-SNP mean genotype over the observed entries (mean-imputation loops).
template <typename T> void launch_snp_mean(const unsigned long long* snp_nm, const unsigned long long* snp_alt,
                                           int64_t J, T* mu, hipStream_t s);
// Cross-validation folds: masked copy of a BED, held-out deviance partials.
void launch_mask_fold(const uint8_t* src, uint8_t* dst, int64_t pitch, int64_t I, int64_t J, int fold, int nfolds,
                      uint64_t seed, hipStream_t s);
int64_t heldout_dev_blocks(int64_t I, int64_t J, int64_t* jchunk);
void launch_heldout_dev(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int K, const double* q,
                        const double* p, int fold, int nfolds, uint64_t seed, double* part, hipStream_t s);
// Synthetic genotypes from (Q,P) (double), seeded counter hash.
void launch_gen_bed(const double* Qt, const double* Pt, int64_t I, int64_t J, int K, uint64_t seed,
                    double miss_rate, uint8_t* bed, int64_t pitch, hipStream_t s);
// Rows [j0, j1) only, into bed + pitch * j (a streamed context's chunk buffer).
void launch_gen_bed_rows(const double* Qt, const double* Pt, int64_t I, int64_t j0, int64_t j1, int K,
                         uint64_t seed, double miss_rate, uint8_t* bed, int64_t pitch, hipStream_t s);
// *flag = 1 if the FP64 state (q: K x I, p: K x J) may give some s = q.p of
// exactly 0 or 1, or holds a non-finite value; else 0.  Safe states: every p
// in [1e-12, 1 - 1e-12], every q finite and >= 0, every column sum of q in
// [1e-200, 1 + 1e-13]; then every s and 1 - s lies strictly inside (0, 1).
void launch_ll_state_check(int K, const double* q, int64_t I, const double* p, int64_t J, int* flag, hipStream_t s,
                           bool f32_state = false);
// y = (double)x, n values.
void launch_widen(const float* x, int64_t n, double* y, hipStream_t s);
// `rows` rows of a .bed file's payload (ibytes each, back to back; `real` of
// them present, the rest become missing codes) into the pitched device layout.
void launch_repitch(const uint8_t* raw, int64_t ibytes, int64_t real, int64_t rows, uint8_t* dst, int64_t pitch,
                    hipStream_t s);

// Device address of the allocation-free log-likelihood's fixed scratch on
// `device` (admx_noalloc.hip), resolved once per device; *out may be null.
int noalloc_setup(int device, double** out);

// Slots of the device log-likelihood vector.
enum LLSlot : int { LL_PREV = 0, LL_NEW = 1, LL_BASIC = 2, LL_QN = 3, LL_TMP = 4, LL_TMP2 = 5, LL_NSLOTS = 8 };

// learned warm start (admx_mlp.hip)
void launch_features(const uint8_t* bed, int64_t pitch, int64_t I, int64_t Jp, const float* R, const float* mu,
                     const float* isd, int nchunk, int64_t jchunk, float* part, float* F, hipStream_t s);
void launch_sgemm(bool ta, bool tb, int M, int N, int Kd, const float* A, int lda, const float* B, int ldb, float* C,
                  int ldc, const float* bias, bool relu, const float* mask, bool accumulate, hipStream_t s);
void launch_colsum(const float* X, int M, int N, float* out, hipStream_t s);
void launch_softmax_ce(const float* Z, int M, int K, const float* T, float* Y, float* dZ, float* rowloss,
                       hipStream_t s);
void launch_gather_rows(const float* X, int ld, const int32_t* idx, int n, float* Y, hipStream_t s);
void launch_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float c1,
                 float c2, float eps, hipStream_t s);
void launch_probs_to_q(const float* Y, int64_t n, double* q, hipStream_t s);
// p (K x J, FP64) <- clipped observed allele frequencies from per-SNP counts;
// with ll_part (one double per 256 SNPs, or null), the per-workgroup partial
// sums of the log-likelihood at that p, which any q on the simplex shares.
void launch_freq_to_p(const unsigned long long* nm, const unsigned long long* alt, int64_t J, int K, double* p,
                      double* ll_part, hipStream_t s);

// sparse k-means feature ranking (admx_skfr.hip)
void launch_skfr_masks(const int32_t* lab, int64_t I, int k, int64_t nwords, uint32_t* cm, unsigned* nk,
                       hipStream_t s);
int64_t skfr_slabs(int64_t J, int64_t ntiles, int64_t* tper);
void launch_skfr_centers(int k, const uint8_t* bed, int64_t pitch, int64_t J, int64_t ntiles, int64_t tper,
                         int64_t nslab, const uint32_t* cm, uint32_t* part, hipStream_t s);
void launch_skfr_stats(const unsigned long long* nm, const unsigned long long* alt, int64_t J, double* mu,
                       double* var, hipStream_t s);
void launch_skfr_score(int k, int local, const uint32_t* part, int64_t nslab, const unsigned* nk,
                       const unsigned long long* nm, const unsigned long long* alt, int64_t I, const double* mu,
                       const double* var, int64_t J, double* E, double* key, int32_t* idx, hipStream_t s);
size_t skfr_sort_bytes(int64_t n);
void launch_skfr_sort(void* tmp, size_t bytes, const double* kin, double* kout, const int32_t* vin, int32_t* vout,
                      int64_t n, hipStream_t s);
void launch_skfr_flag(const int32_t* sorted, int64_t J, int64_t sp, int k, uint8_t* flag, hipStream_t s);
void launch_skfr_union(const uint8_t* flag, const double* key, int64_t J, int k, double* ukey, int32_t* idx,
                       unsigned long long* count, hipStream_t s);
void launch_skfr_tables(const int32_t* cols, int64_t ns, int k, const uint8_t* flag, const unsigned* nk,
                        const double* mu, const double* var, const double* E, int64_t J, double* T, hipStream_t s);
void launch_skfr_pp_table(const uint8_t* bed, int64_t pitch, const int32_t* center, const double* mu,
                          const double* var, int64_t J, double* T, hipStream_t s);
int skfr_nsplit(int64_t I, int64_t ncols, int R);
void launch_skfr_tabsum(int R, const uint8_t* bed, int64_t pitch, int64_t I, const int32_t* cols, int64_t ncols,
                        int nsplit, const double* T, double* P, hipStream_t s);
void launch_skfr_assign(const double* P, int nsplit, int R, int64_t I, const unsigned* nk, int32_t* lab,
                        double* dmin, unsigned long long* changes, hipStream_t s);
void launch_skfr_combine(const double* P, int nsplit, int64_t I, double* out, double* D2, int first,
                         hipStream_t s);
void launch_skfr_pick(const double* D2, int64_t I, double u, int first, int32_t* center, hipStream_t s);
void launch_skfr_argmin(const double* Dc, int k, int64_t I, int32_t* lab, hipStream_t s);
void launch_skfr_wss(const int32_t* lab, const double* x2, const double* dmin, int64_t I, int k, double* out,
                     hipStream_t s);
void launch_skfr_mark(const int32_t* cols, int64_t ns, uint8_t* sel, hipStream_t s);
void launch_skfr_center_out(const uint8_t* sel, const uint8_t* flag, const unsigned* nk, const double* var,
                            const double* E, int64_t J, int k, double* out, hipStream_t s);

}  // namespace admx
