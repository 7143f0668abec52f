/*
 * admixhip.h -- C ABI of libadmixhip, the MI355X (gfx950) engine for the
 * likelihood / EM / SQP inner loops of OpenADMIXTURE.jl.
 *
 * Conventions (identical to the reference's Julia arrays):
 *   - every matrix is column-major: q[k,i] at k + K*i (K x I), p[k,j] at
 *     k + K*j (K x J), x = [q p] (K x (I+J)), XtX[k2,k,n] at k2 + K*k + K*K*n,
 *     U/V at n + K*(I+J)*c (K(I+J) x Q);
 *   - element type is the context's dtype: double (ADMX_F64) or float (ADMX_F32);
 *     log-likelihoods are always double.  The genotype passes of an ADMX_F32
 *     context run in FP64 on widened copies of q and p (same speed on gfx950,
 *     fewer instructions); its state, Hessians and QPs stay float;
 *   - index ranges are 0-based half-open [i0,i1) x [j0,j1); the reference's
 *     1-based UnitRange a:b is [a-1, b);
 *   - host pointers are read or written only during the call (the caller keeps
 *     them alive, e.g. GC.@preserve); device memory is owned by the context;
 *   - every function returns ADMX_OK (0) or an error code, with a message in
 *     admx_last_error() (thread-local).  A QP that fails its pivot test is not
 *     an error: it yields delta = 0 and status 0, as in the reference;
 *   - a context holds one HIP stream and is not re-entrant: calls on the same
 *     context are serialised by an internal mutex.
 */
#ifndef ADMIXHIP_H
#define ADMIXHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ADMX_OK 0
#define ADMX_EINVAL 1 /* bad argument (shape, range, dtype, K > 16, Q > 8) */
#define ADMX_EHIP 2   /* HIP runtime error */
#define ADMX_ESTATE 3 /* missing state (e.g. no genotypes uploaded) */
#define ADMX_EIO 4    /* reading a .bed file failed */
#define ADMX_EHOST 5  /* host-side failure (out of host memory, no thread) */

#define ADMX_F64 0
#define ADMX_F32 1

/* which state vector (AdmixData fields x, x_next, x_next2, x_tmp; src/structs.jl:16-19) */
#define ADMX_X 0
#define ADMX_X_NEXT 1
#define ADMX_X_NEXT2 2
#define ADMX_X_TMP 3

#define ADMX_SIDE_Q 0
#define ADMX_SIDE_P 1

/* tuning keys for admx_set_tuning */
#define ADMX_TUNE_TARGET_WAVES 0 /* waves per genotype pass (sets the chunk count) */
#define ADMX_TUNE_KEEP_HESSIANS 1 /* 1: update_q/update_p also store XtX and Xtz */
#define ADMX_TUNE_PRECISION 2     /* FP64 contexts: precision of the device-resident update passes */
#define ADMX_TUNE_GRAPHS 3        /* 1: admx_qn_iteration replays steady-state iterations (iter > Q) from
                                   * captured hipGraphs (not while profiling); results are identical */
#define ADMX_TUNE_QN_MODE 4       /* quasi-Newton step of admx_qn_iteration: ADMX_QN_ZAL or ADMX_QN_LBQN */
#define ADMX_TUNE_SKIPMISSING 5   /* 1 (default): the *_skipmissing loops; 0: the mean-imputation loops
                                   * (src/loops.jl:493-514, 589-608, 679-698, 840-865, 896-921), where a
                                   * missing genotype counts as the SNP's mean over its observed entries;
                                   * every pass and entry point follows it (FP64/FP32 passes only) */
#define ADMX_TUNE_RESIDENT_SNPS 6 /* streamed contexts: keep SNPs [0, value) in HBM (copied once, and
                                   * again after admx_upload_bed / admx_generate_bed); passes stream
                                   * only the rest.  0 (default): stream everything; -1: as many as
                                   * the free HBM holds, less max(4 GiB, 10%) kept for pass scratch */
#define ADMX_TUNE_SPLIT_Q 7       /* FP64 update_q pass, K >= 9: 1 runs the Hessian triangle in two launches
                                   * of half the accumulators each (higher occupancy, the genotypes read
                                   * twice), 0 in one; -1 (default): two from K = 13.  Same results. */
#define ADMX_TUNE_SPLIT_P 8       /* the same for the FP64 update_p pass; -1: two from K = 13 */
#define ADMX_TUNE_QSIDE_MFMA 9    /* BF16 form's update_q pass: 1 forms s, the Hessian and the gradient on the
                                   * matrix cores (gradient from two-term BF16 products), 0 the Hessian only
                                   * (s and the gradient in FP32 on the vector ALU); -1 (default): 1 from
                                   * K = 7 (faster there, slower below).  Both within the BF16 form's error
                                   * bars against the FP64 oracle.  The BF16 form's em_q sums take the same
                                   * block-image kernel unless this is 0. */
#define ADMX_TUNE_PSIDE_MFMA 10   /* the same for the BF16 form's update_p pass (w = a + b Hessian weight,
                                   * gradient u = c - d from two-term BF16 products) */

/* values of ADMX_TUNE_QN_MODE: admixture_qn!'s `mode` (src/algorithms_outer.jl:127-136) */
#define ADMX_QN_ZAL 0
#define ADMX_QN_LBQN 1

/* values of ADMX_TUNE_PRECISION.  QP, projections, quasi-Newton, the
 * log-likelihood reduction and the convergence test stay FP64 in every mode. */
#define ADMX_PREC_F64 0  /* plain: everything FP64 */
#define ADMX_PREC_F32 1  /* gradient + Hessian accumulated in FP32 (vector ALU) */
#define ADMX_PREC_BF16 2 /* gradient in FP32, Hessian on BF16 matrix cores with FP32 accumulation; the
                          * device state's log-likelihood takes s = q.p and 1 - s = q.(1-p) from the matrix
                          * cores too (three-term BF16 splits, FP32 accumulation, relative error a few 1e-7),
                          * with the per-genotype products and the reduction in FP64; admx_em_q / admx_em_p
                          * and admx_init_em form their sums on the matrix cores (FP32 weights, two-term BF16
                          * products, FP32 accumulation; q_next, p_next within 2e-5 of FP64) */

typedef struct admx_ctx admx_ctx;

const char *admx_version(void);
const char *admx_last_error(void);
/* Number of visible HIP devices. */
int admx_device_count(int *n);

/* ---- context: replaces CuAdmixData{T} + _cu_admixture_base
 * (src/cuda/structs.jl:1-38, stub src/driver.jl:204-206). */
int admx_create(admx_ctx **out, int64_t I, int64_t J, int64_t K, int64_t Q, int dtype);
/* J-chunk streaming: a context whose BED stays in pinned host memory (for
 * cohorts larger than HBM).  Every genotype pass streams it in chunks of
 * snps_per_chunk SNPs through two device buffers, the copy of the next chunk
 * overlapping the kernels of the current one.  Same results as a resident
 * context.  FP64/FP32 passes with the skipmissing loops; decode, counts, fold
 * masking, the warm-start features, the mixed-precision passes and the
 * mean-imputation loops need a resident BED.  No reference counterpart (its
 * joffset/width hooks, src/cuda/kernels.jl:3,20 and src/cuda/structs.jl:13, are
 * unfinished). */
int admx_create_streamed(admx_ctx **out, int64_t I, int64_t J, int64_t K, int64_t Q, int dtype,
                         int64_t snps_per_chunk);
/* J-chunk streaming straight from a PLINK .bed file (SNP-major, the 3-byte
 * magic 6C 1B 01 followed by ceil(I/4) * J bytes), for cohorts larger than
 * host memory as well: no host copy of the BED is kept.  Each chunk is read
 * (pread, up to 8 host threads, from the page cache or the disk) into one of
 * two pinned staging buffers, copied to HBM and re-pitched there while the
 * previous chunk's kernels run.  The file stays open, read-only, until
 * admx_destroy; admx_upload_bed and admx_generate_bed are refused
 * (ADMX_ESTATE), admx_download_bed reads the file.  Same pass coverage and
 * results as admx_create_streamed.  No reference counterpart (the reference
 * reads the whole file through SnpArrays, src/driver.jl:144-145). */
int admx_create_streamed_file(admx_ctx **out, const char *bed_path, int64_t I, int64_t J, int64_t K, int64_t Q,
                              int dtype, int64_t snps_per_chunk);
int admx_destroy(admx_ctx *ctx);
/* Upload the raw SNP-major BED payload: J columns of bytes_per_snp bytes
 * (bytes_per_snp >= ceil(I/4); the 3-byte PLINK magic already stripped),
 * i.e. g.s.data of SnpArrays (src/cuda/structs.jl:35-37). */
int admx_upload_bed(admx_ctx *ctx, const uint8_t *bed, int64_t bytes_per_snp);
/* Download the BED payload as ceil(I/4) x J bytes. */
int admx_download_bed(admx_ctx *ctx, uint8_t *bed);
/* Synthetic genotypes on device from true Q (K x I) and P (K x J) in double:
 * g ~ Binomial(2, q_i.p_j), missing with probability miss_rate, seeded
 * counter hash.  No reference counterpart (workload generator). */
int admx_generate_bed(admx_ctx *ctx, const double *Qt, const double *Pt, uint64_t seed, double miss_rate);

/* ---- cross-validation holdout.  No reference counterpart (the reference has
 * no cross-validation); these serve the K sweep with nfolds-fold CV.  Entry
 * (i, j) belongs to fold floor(nfolds * hi32(smix(j*I + i ^ base(seed))) / 2^32),
 * so the folds partition the genotype entries. */
/* dst's genotypes = src's with the entries of fold `fold` set to missing;
 * fold < 0 copies.  Same I, J and device; device-to-device. */
int admx_bed_from(admx_ctx *dst, admx_ctx *src, int32_t fold, int32_t nfolds, uint64_t seed);
/* Binomial deviance 2[g log(g/2s) + (2-g) log((2-g)/2(1-s))] of fit's state x
 * (FP64 context) summed over the observed entries of fold `fold` of src's
 * (unmasked) genotypes, and the number of those entries. */
int admx_heldout_deviance(admx_ctx *fit, admx_ctx *src, int32_t fold, int32_t nfolds, uint64_t seed,
                          double *deviance_sum, int64_t *count);

/* ---- sparse k-means feature ranking (SKFR): the AIM pre-selection of
 * run_admixture(...; sparsity) (src/driver.jl:56-66, 80-101), where the
 * reference calls SparseKmeansFeatureRanking.jl 0.1 (ImputedSnpMatrix,
 * sparsekmeans1 / sparsekmeans2).  Genotypes are standardized by
 * x = (g - mu_j) / sqrt(mu_j (1 - mu_j / 2)) with mu_j the mean over observed
 * entries, missing entries x = 0.  Each iteration forms the cluster means,
 * ranks the SNPs by sum_k n_k mean_kj^2 (global) or n_k mean_kj^2 per cluster
 * (local), keeps the top `sparsity` (ties: lower SNP index first), zeroes the
 * other center coordinates and moves every individual to its nearest nonempty
 * center (ties: lower cluster); it stops when no label changes or after
 * max_iter iterations.  Labels and SNP indices are 0-based.  nsplit (0: chosen
 * from I and the column count, else 1..64) is the number of SNP slices whose
 * distance partials are added in order; labels and AIMs are reproducible for
 * a given nsplit.  Needs a resident BED; k <= 16. */
#define ADMX_SKFR_GLOBAL 0 /* sparsekmeans1: one ranking over all clusters */
#define ADMX_SKFR_LOCAL 1  /* sparsekmeans2: a ranking per cluster; AIMs = union */
/* k-means++ start (the ImputedSnpMatrix constructor's initial clustering):
 * centers drawn with probability proportional to the squared distance to the
 * nearest chosen one, over all SNPs; clusters (I) = nearest center. */
int admx_skfr_init(admx_ctx *ctx, int32_t k, uint64_t seed, int32_t nsplit, int32_t *clusters);
/* sparsekmeans1 / sparsekmeans2 (ISM, sparsity; max_iter) from the labels in
 * clusters (I, in/out).  aims receives the selected SNPs in decreasing score
 * (capacity sparsity for ADMX_SKFR_GLOBAL, min(J, k * sparsity) for
 * ADMX_SKFR_LOCAL, ranked by the largest per-cluster score), n_aims their
 * number; centers (k x J column-major, nullable) the sparse centers in
 * standardized units; wss (k, nullable) the within-cluster sums of squares
 * over all SNPs and obj (nullable) their total; iters the iterations run. */
int admx_skfr_run(admx_ctx *ctx, int32_t k, int32_t mode, int64_t sparsity, int32_t max_iter, int32_t nsplit,
                  int32_t *clusters, int64_t *aims, int64_t *n_aims, double *centers, double *wss, double *obj,
                  int32_t *iters);

/* ---- learned warm start.  No reference counterpart (the reference starts
 * from a random point plus FRAPPE EM, src/algorithms_outer.jl:12-45).
 * Features: per individual, (1/sqrt(Jp)) sum_{j < Jp observed} (g_ij - mu_j) isd_j R[j, :]
 * over the first Jp SNPs of a reference panel (R: Jp x 64 row-major, FP32).
 * Network: FP32 MLP in -> hidden -> hidden -> out, ReLU, softmax head, trained
 * by Adam on the soft-target cross entropy; parameters flat
 * [W1 (hidden x in), b1, W2, b2, W3 (out x hidden), b3], row-major arrays. */
typedef struct admx_mlp admx_mlp;
int admx_mlp_create(admx_mlp **out, int32_t in_dim, int32_t hidden, int32_t out_dim, uint64_t seed);
int admx_mlp_destroy(admx_mlp *m);
int admx_mlp_num_params(admx_mlp *m, int64_t *n);
int admx_mlp_get_params(admx_mlp *m, float *w);
int admx_mlp_set_params(admx_mlp *m, const float *w); /* also resets the Adam state */
/* mean loss and its gradient (flat, like the parameters) over B rows X (B x in), T (B x out) */
int admx_mlp_loss_grad(admx_mlp *m, const float *X, const float *T, int64_t B, float *loss, float *grad);
int admx_mlp_predict(admx_mlp *m, const float *X, int64_t n, float *Y); /* Y: n x out probabilities */
/* Adam (lr, 0.9, 0.999, 1e-8) over shuffled batches; epoch_loss[epochs] may be NULL */
int admx_mlp_train(admx_mlp *m, const float *X, const float *T, int64_t n, int32_t epochs, int32_t batch, float lr,
                   uint64_t seed, float *epoch_loss);
/* features of ctx's genotypes into ctx (kept for admx_warm_q); F (I x 64) may be NULL */
int admx_features(admx_ctx *ctx, const float *R, const float *mu, const float *isd, int64_t Jp, float *F);
/* q of the state x <- simplex projection (pseudocount 1e-5) of the network's
 * output on ctx's features; FP64 context, network out == K, in == 64 */
int admx_warm_q(admx_ctx *ctx, admx_mlp *m);
/* p of the state x <- every SNP's observed allele frequency (alt / 2 nonmissing,
 * clipped to [1e-5, 1 - 1e-5]) for all K populations, computed on the device
 * from the resident BED; FP64 context.  With ll_hw non-null, *ll_hw <- the
 * log-likelihood at that p, which every q on the simplex gives (s = q.p = f):
 * sum over SNPs of alt log f + (2 nonmissing - alt) log(1 - f), from the same
 * counts instead of a genotype pass.  No reference counterpart (the warm
 * start's first p). */
int admx_warm_p(admx_ctx *ctx, double *ll_hw);

/* ---- BED decode and counts (decode tables src/loops.jl:2-5).  No count
 * function exists in the reference; these are exact integer counts. */
int admx_decode(admx_ctx *ctx, int8_t *g); /* I x J, -1 = missing */
int admx_counts(admx_ctx *ctx, int64_t *snp_nonmissing, int64_t *snp_alt, int64_t *ind_nonmissing);

/* ---- tile loops of src/loops.jl (SnpLinAlg, skipmissing variants), host
 * arrays, accumulate (+=) semantics into the caller's outputs. */
/* loglikelihood_loop_skipmissing(g, q, p, qp_small, irange, jrange, K)::Float64  (loops.jl:517) */
int admx_loglikelihood_loop(admx_ctx *ctx, const void *q, const void *p, int64_t i0, int64_t i1, int64_t j0,
                            int64_t j1, double *out);
/* em_q_loop_skipmissing!(q_next, g, q, p, qp_small, irange, jrange, K)  (loops.jl:610) */
int admx_em_q_loop(admx_ctx *ctx, void *q_next, const void *q, const void *p, int64_t i0, int64_t i1, int64_t j0,
                   int64_t j1);
/* em_p_loop_skipmissing!(p_next, g, q, p, p_tmp, qp_small, irange, jrange, K)  (loops.jl:700) */
int admx_em_p_loop(admx_ctx *ctx, void *p_next, void *p_tmp, const void *q, const void *p, int64_t i0, int64_t i1,
                   int64_t j0, int64_t j1);
/* update_q_loop_skipmissing!(XtX, Xtz, g, q, p, qp_small, irange, jrange, K)  (loops.jl:867) */
int admx_update_q_loop(admx_ctx *ctx, void *XtX, void *Xtz, const void *q, const void *p, int64_t i0, int64_t i1,
                       int64_t j0, int64_t j1);
/* update_p_loop_skipmissing!(XtX, Xtz, g, q, p, qp_small, irange, jrange, K)  (loops.jl:923) */
int admx_update_p_loop(admx_ctx *ctx, void *XtX, void *Xtz, const void *q, const void *p, int64_t i0, int64_t i1,
                       int64_t j0, int64_t j1);

/* ---- batched QP / projections / quasi-Newton on host arrays. */
/* create_tableau! + quadratic_program! for n problems (src/qp.jl:8-292,
 * batched as in src/algorithms_inner.jl:205-221 (simplex) and :281-298 (box)).
 * Xtz_neg is the negated gradient as stored after `Xtz .*= -1`. */
int admx_qp_batch(admx_ctx *ctx, int simplex, int64_t n, const void *XtX, const void *Xtz_neg, const void *x,
                  void *delta, int32_t *status);
/* project_q!(b) on n columns of K (src/projections.jl:1-63) */
int admx_project_q(admx_ctx *ctx, void *b, int64_t n);
/* project_p!(b) on n entries (src/projections.jl:69-74) */
int admx_project_p(admx_ctx *ctx, void *b, int64_t n);
/* update_UV! + update_QN! + project_p!/project_q! of one ZAL step
 * (src/quasi_newton.jl:5-9, 29-31; src/algorithms_outer.jl:127-142): iter is
 * the 1-based outer index; U, V are updated in place, x_tmp receives the step. */
int admx_zal_step(admx_ctx *ctx, int64_t iter, const void *x, const void *x_next, const void *x_next2, void *U,
                  void *V, void *x_tmp);
/* update_UV_LBQN! + update_QN_LBQN! + project_p!/project_q!  (src/quasi_newton.jl:15-58;
 * src/algorithms_outer.jl:132-142), host arrays, same shapes as admx_zal_step */
int admx_lbqn_step(admx_ctx *ctx, int64_t iter, const void *x, const void *x_next, const void *x_next2, void *U,
                   void *V, void *x_tmp);

/* ---- device-resident state (AdmixData mirrored in HBM) */
int admx_set_x(admx_ctx *ctx, int which, const void *host);
int admx_get_x(admx_ctx *ctx, int which, void *host);
/* state vector dst <- state vector src, on the device (copyto! between the
 * device mirror's x, x_next, x_next2, x_tmp) */
int admx_copy_x(admx_ctx *ctx, int dst, int src);
int admx_set_uv(admx_ctx *ctx, const void *U, const void *V);
int admx_get_uv(admx_ctx *ctx, void *U, void *V);
int admx_set_ll(admx_ctx *ctx, double ll_prev, double ll_new);
int admx_get_ll(admx_ctx *ctx, double *ll_prev, double *ll_new);
/* XtX/Xtz of the last update on side ADMX_SIDE_Q / ADMX_SIDE_P: after
 * admx_update_q / admx_update_p Xtz is negated, as update_q!/update_p! leave it
 * (src/algorithms_inner.jl:199, 275); after admx_grad_hessian it is the raw
 * gradient sum, as update_q_cuda!/update_p_cuda! leave it.  Either pointer may
 * be NULL (that array is not copied). */
int admx_get_hessians(admx_ctx *ctx, int side, void *XtX, void *Xtz);

/* ---- per-array device mirror of CuAdmixData (src/cuda/structs.jl:1-31), for
 * a backend that plugs into the reference's own GPU branch: its device arrays
 * q, p, q_next, p_next, p_tmp are blocks of the state vectors (q of x = d_cu.q,
 * p of x = d_cu.p, q/p of x_next = d_cu.q_next/p_next, p of x_tmp = d_cu.p_tmp). */
/* copyto!(d_cu.q or d_cu.p, host) of copyto_sync! (src/cuda/transfer.jl:1-7):
 * the K x I (side Q) or K x J (side P) block of state vector `which` from a
 * column-major host array. */
int admx_set_block(admx_ctx *ctx, int which, int side, const void *host);
/* copyto!(host, d_cu.q or d_cu.p): the block to a host array. */
int admx_get_block(admx_ctx *ctx, int which, int side, void *host);
/* d_cu.p .= d_cu.p_next, d_cu.q .= d_cu.q_next (src/algorithms_outer.jl:28-29):
 * block `side` of state vector dst <- the same block of src, on the device. */
int admx_copy_block(admx_ctx *ctx, int dst, int src, int side);
/* update_q_cuda!(d_cu, g_cu) / update_p_cuda!(d_cu, g_cu) (src/cuda/runners.jl:69-111,
 * called at src/algorithms_inner.jl:193, 269): XtX and Xtz of state x's (q, p) on
 * side Q (K x K x I, K x I) or P (K x K x J, K x J) into the context's Hessian
 * arrays, zeroed then written, Xtz not negated, no QP (FP64 passes); read them
 * with admx_get_hessians. */
int admx_grad_hessian(admx_ctx *ctx, int side);

/* ---- algorithms on the device-resident state */
/* loglikelihood(g, q, p, ...) of state vector `which` (src/algorithms_inner.jl:27-35) */
int admx_loglikelihood(admx_ctx *ctx, int which, double *out);
/* The log-likelihoods of state vectors which1 and which2 into out[0], out[1],
 * each identical to admx_loglikelihood's, from one pass over the genotypes
 * where admixture_qn! pairs them too (FP64 loop up to K = 8, any streamed
 * context); otherwise one pass per state.  No single counterpart in the
 * reference: it fuses the two loglikelihood calls of admixture_qn! (ll_basic
 * and ll_qn, src/algorithms_outer.jl:109-114, 134-139). */
int admx_loglikelihood2(admx_ctx *ctx, int which1, int which2, double *out);
/* em_q!(d, g): x_next.q from x  (src/algorithms_inner.jl:68-92) */
int admx_em_q(admx_ctx *ctx);
/* em_p!(d, g): x_next.p from x  (src/algorithms_inner.jl:108-141) */
int admx_em_p(admx_ctx *ctx);
/* em_q!(d, g; mode) / em_p!(d, g; mode) with the reference's modes
 * (src/algorithms_inner.jl:54-141): ADMX_EM_BASE as admx_em_q / admx_em_p;
 * ADMX_EM_FALLBACK: em_q! as base, em_p! writes p of x_next from q of x_next
 * and p of x; ADMX_EM_FALLBACK2: em_q! writes q of x_next2 from q, p of x_next,
 * em_p! writes p of x_next2 from q of x_next2 and p of x_next. */
#define ADMX_EM_BASE 0
#define ADMX_EM_FALLBACK 1
#define ADMX_EM_FALLBACK2 2
int admx_em_q_mode(admx_ctx *ctx, int mode);
int admx_em_p_mode(admx_ctx *ctx, int mode);
/* init_em!(d, g, iters): lls[it] = LL after EM step it; sets ll_new
 * (src/algorithms_outer.jl:12-45) */
int admx_init_em(admx_ctx *ctx, int64_t iters, double *lls);
/* update_q!(d, g, update2) / update_p!(d, g, update2)
 * (src/algorithms_inner.jl:167-231, 245-310); status (nullable) receives the
 * QP return codes, I or J entries. */
int admx_update_q(admx_ctx *ctx, int update2, int32_t *status);
int admx_update_p(admx_ctx *ctx, int update2, int32_t *status);
/* One admixture_qn! iteration with mode=:ZAL (src/algorithms_outer.jl:92-159),
 * iter 1-based.  out = {ll_prev, ll_basic, ll_qn, ll_new, reldiff}. */
int admx_qn_iteration(admx_ctx *ctx, int64_t iter, int fix_q, int fix_p, double *out);
int admx_synchronize(admx_ctx *ctx);

/* ---- allocation-free log-likelihood.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The log-likelihood
 * of admx_loglikelihood / admx_loglikelihood_loop over all I x J genotypes of
 * a caller-held device BED (SNP-major, `pitch` bytes per SNP: pitch >=
 * ceil(I/4), a multiple of 16, pitch * J + 32 bytes readable) and device
 * FP64 q (K x I) and p (K x J); the double goes to the device pointer `out`.
 * Enqueued on `stream` (a hipStream_t; NULL = the null stream) without a
 * synchronisation.  It calls no hipMalloc / hipFree: the pass's partial sums
 * and state-check flag live in a fixed device array of
 * (ADMX_NOALLOC_LL_PARTS + 32) doubles that each GPU holds until the process
 * ends; cohorts with ceil(I/256) x (SNP chunks) > ADMX_NOALLOC_LL_PARTS are
 * refused (ADMX_EINVAL; below 16.7 million individuals every J fits).  Calls
 * that overlap on one GPU share that array and race: serialise them.  Same
 * kernels and chunking as a fresh context, so the same result bit for bit. */
#define ADMX_NOALLOC_LL_PARTS 65536
int admx_loglikelihood_noalloc(const uint8_t *bed, int64_t pitch, int64_t I, int64_t J, int64_t K, const double *q,
                               const double *p, double *out, void *stream);

/* ---- measurement: HIP-event timing of every kernel launch on the context's
 * stream while enabled.  Report lines: "<kernel> <launches> <total_ms>\n". */
int admx_profile_enable(admx_ctx *ctx, int on);
int admx_profile_reset(admx_ctx *ctx);
int admx_profile_report(admx_ctx *ctx, char *buf, size_t len);
int admx_set_tuning(admx_ctx *ctx, int key, int64_t value);

#ifdef __cplusplus
}
#endif
#endif /* ADMIXHIP_H */
