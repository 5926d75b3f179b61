/*
 * fastpls_hip.h - C ABI of the MI355X (gfx950) HIP backend for fastPLS.
 *
 * Drop-in counterpart of the reference's resident device engine ABI
 * (src/cuda_resident_api.h:8-80). Every fastpls_hip_* entry below mirrors the
 * fastpls_* entry named in its comment: same argument order, same meaning,
 * same conventions:
 *   - precision 32 -> float, 64 -> double (api.h:10); matrices are
 *     column-major and contiguous (leading dimension = rows);
 *   - labels are one-based and must lie in 1..q;
 *   - inputs are host pointers, copied by the call; outputs are caller-owned
 *     host buffers of exact size;
 *   - int return 0 = success, 1 = error; create returns NULL on error; the
 *     message is written to (error, error_capacity) with snprintf semantics;
 *   - every call is synchronous; calls on one handle must not overlap.
 *
 * Precision 32 stores X as FP32 on the device. SIMPLS (fits, batches, CV and
 * their predictions) widens it to FP64 as it is loaded and computes every
 * product over X in FP64: its component recursion magnifies the rounding of
 * these products by up to ~1e3. PLS-SVD, fastsvd and fastcor run their
 * products over FP32 X on the matrix cores in bf16x6 split precision
 * (FP32-accurate products, FP32 sums over one 32- to 64-element step folded
 * into FP64 accumulators). The cross-covariance, the randomized SVD, SIMPLS
 * deflation/orthogonalisation, PLS-SVD solves and LDA run in FP64 for both
 * precisions. Gaussian sketches are drawn on the host with
 * std::mt19937(seed) and std::normal_distribution<double>, the same stream
 * as the reference CPU core (simpls.hpp:244-249, rsvd.hpp:205-210).
 *
 * There is no CPU fallback: every entry fails with an error when no HIP
 * device is usable.
 */
#ifndef FASTPLS_HIP_H
#define FASTPLS_HIP_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Library and device identification (no reference counterpart): writes a
 * NUL-terminated description; returns 0 when a device is usable. */
int fastpls_hip_device_info(char* out, size_t capacity);

/* fastpls_cuda_gemm (src/cuda_resident_api.h:11-13): output(rows x columns)
 * = left(rows x inner) * right(inner x columns), host buffers. */
int fastpls_hip_gemm(const void* left, const void* right, int precision, int rows,
                     int inner, int columns, void* output, char* error,
                     size_t error_capacity);

/* fastpls_resident_simpls_create (src/cuda_resident_api.h:40-43). Exactly one
 * of y (n x q) and labels (n, one-based, q classes) is non-NULL.
 * retain_scores (0 or 1): with 0 a classification fit with components >= 8,
 * n >= 8p and n x components scores above 256 MiB forms its LDA moments and
 * drops its training scores after the fit (src/cuda_resident_simpls.cuh:
 * 213-216); every other fit keeps them. Fits with components >= 20, p <= n,
 * n >= 8p and p <= 512 take their loadings from a cached X_sᵀX_s (the CPU
 * core's rule, src/r_api.cpp:1165-1176). */
void* fastpls_hip_simpls_create(const void* x, const void* y, const int* labels,
                                int precision, int n, int p, int q, int components,
                                int scaling, int oversample, int power,
                                int retain_scores, unsigned long long seed,
                                char* error, size_t error_capacity);

/* fastpls_resident_plssvd_create (src/cuda_resident_api.h:44-47).
 * retain_scores as above, with the PLS-SVD rule (2 x components >= p instead
 * of components >= 8; src/cuda_resident_plssvd.cuh:166-169). */
void* fastpls_hip_plssvd_create(const void* x, const void* y, const int* labels,
                                int precision, int n, int p, int q, int components,
                                int scaling, int oversample, int power,
                                int retain_scores, unsigned long long seed,
                                char* error, size_t error_capacity);

/* fastpls_resident_opls_create (src/cuda_resident_api.h:48-51): resident OPLS,
 * `north` >= 0 orthogonal filter components (rank-one randomized direction of
 * the current cross-covariance, seed + component), then SIMPLS on the
 * filtered predictors with the same seed; query rows are standardised and
 * filtered on the device. Export field 0 holds R_eff, the weights that take
 * standardised rows to the scores; fields 3/4 the outer centre and scale.
 * fastpls_resident_kernelpls_create (src/cuda_resident_api.h:52-56): resident
 * nonlinear kernel PLS, kernel 2 radial basis exp(-gamma |a - b|^2) or 3
 * polynomial (gamma a.b + coef0)^degree on the standardised predictors,
 * double-centred, then SIMPLS on the n x n kernel; fields 0 and 6 have n
 * rows. The n x n kernel plus the predictors must fit in 3/5 of the free
 * device memory, as must a query's rows x n test kernel. */
void* fastpls_hip_opls_create(const void* x, const void* y, const int* labels,
                              int precision, int n, int p, int q, int components,
                              int scaling, int oversample, int power, int retain_scores,
                              unsigned long long seed, int north, char* error,
                              size_t error_capacity);
void* fastpls_hip_kernelpls_create(const void* x, const void* y, const int* labels,
                                   int precision, int n, int p, int q, int components,
                                   int scaling, int oversample, int power,
                                   int retain_scores, unsigned long long seed, int kernel,
                                   double gamma, int degree, double coef0, char* error,
                                   size_t error_capacity);

/* fastpls_resident_predict_path (src/cuda_resident_api.h:57-59): output is
 * rows x q x prefix_count; lda != 0 returns LDA discriminant scores.
 * Queries on host rows (this entry and the classify entries) go through the
 * device in row blocks when their workspace would exceed 2/5 of the free
 * device memory, as the reference blocks classify
 * (src/cuda_resident_api.cuh:259-297); the environment variable
 * FASTPLS_HIP_QUERY_FREE_BYTES, when set, stands for the free memory. */
int fastpls_hip_predict_path(void* model, const void* x, int rows, const int* prefixes,
                             int prefix_count, int lda, void* predictions, char* error,
                             size_t error_capacity);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): fastpls_hip_predict_path for one prefix, without LDA,
 * that allocates nothing. x_device (rows x p) and predictions_device (rows x
 * q, the model's precision) are DEVICE pointers. Every intermediate (split
 * partials, scores, FP64 predictions, the bf16x6 operand workspace) lives in
 * one 32 MiB device array of the library per GPU: requests needing more
 * (about rows * (prefix * (splits + 1) + q) doubles) fail with an error.
 * Calls on one GPU share that array, so calls that overlap on one GPU race;
 * calls on one handle must not overlap anyway. Synchronous, like every call. */
int fastpls_hip_predict_noalloc(void* model, const void* x_device, int rows, int prefix,
                                void* predictions_device, char* error, size_t error_capacity);

/* fastpls_resident_simpls_destroy (src/cuda_resident_api.h:60). */
void fastpls_hip_destroy(void* model);

/* fastpls_resident_export (src/cuda_resident_api.h:61-62). Fields: 0 R (p*a),
 * 1 Q (q*a), 2 T (n*a), 3 x means (p), 4 x scales (p), 5 y means / priors (q),
 * 6 X loadings P (p*a), 7 variance sums of squares (a+1). */
int fastpls_hip_export(void* model, int field, void* out, size_t elements, char* error,
                       size_t error_capacity);

/* fastpls_resident_classify_path (src/cuda_resident_api.h:63-65): labels are
 * rows x top x prefix_count, one-based. */
int fastpls_hip_classify_path(void* model, const void* x, int rows, const int* prefixes,
                              int prefix_count, int lda, int top, int* labels, char* error,
                              size_t error_capacity);

/* fastpls_resident_classify_response_path (src/cuda_resident_api.h:66-68). */
int fastpls_hip_classify_response_path(void* model, const void* x, int rows,
                                       const int* prefixes, int prefix_count, int lda,
                                       int top, int* labels, void* predictions, char* error,
                                       size_t error_capacity);

/* fastpls_resident_response_sums (src/cuda_resident_api.h:69-70): out is 3 x q
 * (SSE, SS about the training mean, SS about the observed mean). */
int fastpls_hip_response_sums(void* model, const void* x, const void* y, const int* labels,
                              int rows, int prefix, void* out, char* error,
                              size_t error_capacity);

/* fastpls_resident_project (src/cuda_resident_api.h:71-72): scores rows x prefix. */
int fastpls_hip_project(void* model, const void* x, int rows, int prefix, void* out,
                        char* error, size_t error_capacity);

/* fastpls_resident_controls (src/cuda_resident_api.h:73-75).
 * predictor_crossprod_cache is 1 when the fit took the cached X_sᵀX_s route. */
int fastpls_hip_controls(void* model, int* oversample, int* power, int* block,
                         int* block_limit, int* implicit_operator,
                         int* predictor_crossprod_cache, char* error,
                         size_t error_capacity);

/* fastpls_resident_compact (src/cuda_resident_api.h:76-77). */
int fastpls_hip_compact(void* model, int prepare_lda, char* error, size_t error_capacity);

/* Device-resident fit (no reference counterpart: the reference's
 * fit_borrowed_device_* members of ResidentSimpls/ResidentPlssvd,
 * src/cuda_resident_simpls.cuh:287-312, are not exported). x, y and labels
 * are DEVICE pointers that the handle borrows until compact or destroy;
 * method 0 = SIMPLS, 1 = PLS-SVD. Used when inputs already live in HBM. */
void* fastpls_hip_fit_device(int method, const void* x_device, const void* y_device,
                             const int* labels_device, int precision, int n, int p, int q,
                             int components, int scaling, int oversample, int power,
                             unsigned long long seed, char* error, size_t error_capacity);

/* Seconds of device time of the last fit, measured with events on the
 * handle's stream (no reference counterpart). */
int fastpls_hip_last_fit_seconds(void* model, double* seconds, char* error,
                                 size_t error_capacity);

/* Device memory helpers for callers that keep inputs resident (no reference
 * counterpart). kind: 0 host->device, 1 device->host, 2 device->device. */
int fastpls_hip_device_alloc(size_t bytes, void** ptr, char* error, size_t error_capacity);
int fastpls_hip_device_free(void* ptr, char* error, size_t error_capacity);
int fastpls_hip_memcpy(void* dst, const void* src, size_t bytes, int kind, char* error,
                       size_t error_capacity);

/* Device-time breakdown of the last fit from events on the handle's stream
 * (no reference counterpart): out[0] fit seconds, out[1] the one-pass
 * statistics + cross-product kernel over X, out[2] the PLS-SVD score pass
 * over X (SIMPLS: the whole component loop), out[3] bytes of X, out[4]
 * seconds of the projection of the last query's rows (predict / classify /
 * project), out[5] that projection's algorithmic bytes. */
int fastpls_hip_fit_phases(void* model, double* out, int count, char* error,
                           size_t error_capacity);

/* ------------------------------------------------------------------ batches
 * Many independent SIMPLS fits in one grouped execution (no reference
 * counterpart: the reference fits CV folds one after another,
 * src/cuda_resident_api.cuh:770-1023, and permutation refits on the CPU,
 * R/main.R:9419-9450). Fits are grouped by a shared device-resident X
 * (x_device[g], n[g] x p[g], precision 32/64); fit f uses group[f], either
 * y_device[f] (n x q[f], same precision) or labels_device[f] (one-based,
 * q[f] classes), an optional device permutation perm_device[f] of the
 * response rows, components[f] and seeds[f] (rSVD seed, component a uses
 * seed + a as in simpls.hpp:555-556). Every phase of every fit is one launch
 * per step; the call returns when all fits are done. Arrays of pointers are
 * host arrays; y_device / labels_device / perm_device may be NULL as a whole.
 * reorth: -1 follows the FP64 routing (one Gram-Schmidt pass of the
 * loadings; precision 32 computes in FP64 too, so the reference's
 * re-orthogonalisation for its float core, src/r_api.cpp:1416, does not
 * apply); 0 forces that routing, 1 the float core's: twice per component the
 * score's projection on the earlier scores is removed from the score and the
 * weight (simpls.hpp:161-200, 635-642) and the loadings are orthogonalised
 * twice. */
void* fastpls_hip_simpls_batch_create(int precision, int group_count, const void* const* x_device,
                                      const int* n, const int* p, int fit_count,
                                      const int* group, const void* const* y_device,
                                      const int* const* labels_device,
                                      const int* const* perm_device, const int* q,
                                      const int* components, const unsigned long long* seeds,
                                      int scaling, int oversample, int power, int reorth,
                                      char* error, size_t error_capacity);
/* Per-fit completed component counts and status flags (0 = ok) and the device
 * seconds of the whole batch. */
int fastpls_hip_batch_status(void* batch, int* completed, int* flags, double* seconds,
                             char* error, size_t error_capacity);
/* Device-time phases of the batch (no reference counterpart), measured with
 * events around every launch when the environment variable
 * FASTPLS_BENCH_PHASE_TIMING=1 was set at creation (the reference's switch,
 * src/r_api.cpp:1182-1184), zero otherwise: out[0] batch device seconds;
 * phase i (0 statistics + cross-products, 1 refresh, 2 / 3 score passes on the
 * VALU / matrix cores, 4 score slab sums, 5 / 6 loading passes on the VALU /
 * matrix cores, 7 update) fills out[1 + 3i] seconds, out[2 + 3i] algorithmic
 * bytes and out[3 + 3i] launches. count <= 25 entries are written. */
int fastpls_hip_batch_phases(void* batch, double* out, int count, char* error,
                             size_t error_capacity);
/* Fields 0 R (p*A), 1 Q (q*A), 2 T (n*A), 3 x means (p), 4 x scales (p),
 * 5 y means (q) of fit `fit`, in the batch precision. */
int fastpls_hip_batch_export(void* batch, int fit, int field, void* out, size_t elements,
                             char* error, size_t error_capacity);
/* The same fields as the FP64 values the fit holds, whatever the batch
 * precision (no reference counterpart; for accuracy audits: a precision-32
 * export rounds R, Q and T to FP32, and B = R Qᵀ of an ill-conditioned fit
 * magnifies that rounding). */
int fastpls_hip_batch_export64(void* batch, int fit, int field, double* out, size_t elements,
                               char* error, size_t error_capacity);
/* Executes every fit of the batch again on the current contents of the
 * device inputs it was created with (same pointers and shapes: new response
 * permutations, responses or X values written in place), as a permutation
 * test's rounds of refits (no reference counterpart: the reference refits
 * permutations on the CPU, R/main.R:9419-9450). The first call captures the
 * batch's whole launch sequence (statistics, cross-products, the component
 * loop) into a HIP graph and every call replays it: no planning, upload or
 * per-kernel launch on the host. Status, seconds and exports then describe
 * the new execution; the rSVD sketches are those of the creation seeds. */
int fastpls_hip_batch_rerun(void* batch, char* error, size_t error_capacity);
/* Split lengths of the batched passes (no reference counterpart): [0] the
 * longest row split of the cross-product and loading passes and [1] the
 * longest column split of the score pass of groups on the VALU, [2] / [3] the
 * same for groups whose per-step passes run on the FP64 matrix cores; each in
 * 128..4096. New batches take the process default: built-in lengths measured
 * on the 256-fit workload, or those of the file named by the environment
 * variable FASTPLS_HIP_TUNE_FILE when it exists at the first batch, or the
 * last autotune's or set's.
 * fastpls_hip_batch_autotune chooses them by measurement on this batch:
 * starting from its lengths, each length in turn is tried at 1/4, 1/2, 2 and
 * 4 times its value (wide lengths only with matrix-core groups, the others
 * only with VALU groups), the whole batch re-planned and run `trials` (1..16)
 * times per candidate, and a candidate kept when it lowers the fastest device
 * time by more than 1%. The batch then holds the fits of the chosen plan
 * (split sums differ from the first plan's in rounding only), the choice
 * becomes the process default and, with FASTPLS_HIP_TUNE_FILE set, is written
 * there. report (report_count <= 7 entries): [0..3] the chosen lengths, [4]
 * the chosen plan's device seconds, [5] the starting plan's, [6] plans run.
 * fastpls_hip_batch_splits writes a batch's lengths (the process default for
 * a NULL batch); fastpls_hip_batch_set_splits replaces the default. */
int fastpls_hip_batch_autotune(void* batch, int trials, double* report, int report_count,
                               char* error, size_t error_capacity);
int fastpls_hip_batch_splits(void* batch, long long* lengths, char* error,
                             size_t error_capacity);
int fastpls_hip_batch_set_splits(const long long* lengths, char* error, size_t error_capacity);
/* Lanes of a batch (no reference counterpart): the shared-X groups are dealt
 * to sub-batches by estimated work (X elements times components), each run on
 * its own stream of a set that shares no hardware queue, so the per-fit
 * refresh and update kernels of one lane overlap the streaming passes of the
 * others. Each fit's arithmetic, and so its results, are those of a one-lane
 * batch. 2 lanes for 32 fits or more over at least two groups, else 1; the
 * environment variable FASTPLS_HIP_BATCH_LANES (1..min(4, GPU_MAX_HW_QUEUES))
 * sets the count at creation; 1 with FASTPLS_BENCH_PHASE_TIMING=1. Writes the
 * batch's lane count. */
int fastpls_hip_batch_lanes(void* batch, int* lanes, char* error, size_t error_capacity);
void fastpls_hip_batch_destroy(void* batch);

/* fastpls_resident_simpls_cv_regression (src/cuda_resident_api.h:20-26):
 * same arguments, outputs and metric ladder (2 observed R^2, 3 Q^2, 4 RMSD;
 * src/cuda_resident_api.cuh:772-1023). The folds are fitted in batches of as
 * many folds as fit in 2/5 of the free device memory (all folds in one batch
 * when they fit, one at a time at worst; FASTPLS_HIP_CV_FREE_BYTES, when set,
 * stands for the free memory); SIMPLS folds share the seed, as in the
 * reference. */
int fastpls_hip_simpls_cv_regression(const void* predictors, const void* responses,
                                     const int* folds, int precision, int n, int p, int q,
                                     const int* prefixes, int prefix_count, int scaling,
                                     int metric, int oversample, int power,
                                     unsigned long long seed, int store_predictions,
                                     double* predictions, int* status, double* metrics,
                                     double* q2, double* rmsd, double* observed_r2, char* error,
                                     size_t error_capacity);

/* fastpls_resident_plssvd_cv_regression (src/cuda_resident_api.h:33-39):
 * the same contract as the SIMPLS entry above; folds are fitted one after
 * another and fold f (0-based) uses seed + f (src/cuda_resident_api.cuh:934-937). */
int fastpls_hip_plssvd_cv_regression(const void* predictors, const void* responses,
                                     const int* folds, int precision, int n, int p, int q,
                                     const int* prefixes, int prefix_count, int scaling,
                                     int metric, int oversample, int power,
                                     unsigned long long seed, int store_predictions,
                                     double* predictions, int* status, double* metrics,
                                     double* q2, double* rmsd, double* observed_r2, char* error,
                                     size_t error_capacity);

/* fastpls_resident_simpls_cv_classification (src/cuda_resident_api.h:14-19)
 * and fastpls_resident_plssvd_cv_classification (src/cuda_resident_api.h:27-32);
 * semantics of src/cuda_resident_api.cuh:538-768. Labels are 1..classes.
 * Per fold: status 1 fitted, 2 no held-out rows, 4 training rows hold one
 * class or none (that class, or 1, is predicted for the fold). Classes absent
 * from a fold's training rows are dropped from its fit. predictions
 * (n x prefix_count, int) and scores (n x classes x prefix_count, in the
 * call's precision, zero for dropped classes) are written when requested;
 * metrics are accuracies per prefix. SIMPLS folds share the seed, PLS-SVD
 * fold f uses seed + f. */
int fastpls_hip_simpls_cv_classification(const void* predictors, const int* labels,
                                         const int* folds, int precision, int n, int p,
                                         int classes, const int* prefixes, int prefix_count,
                                         int scaling, int lda, int oversample, int power,
                                         unsigned long long seed, int store_predictions,
                                         int store_scores, int* predictions, void* scores,
                                         int* status, double* metrics, char* error,
                                         size_t error_capacity);
int fastpls_hip_plssvd_cv_classification(const void* predictors, const int* labels,
                                         const int* folds, int precision, int n, int p,
                                         int classes, const int* prefixes, int prefix_count,
                                         int scaling, int lda, int oversample, int power,
                                         unsigned long long seed, int store_predictions,
                                         int store_scores, int* predictions, void* scores,
                                         int* status, double* metrics, char* error,
                                         size_t error_capacity);

/* Request server (no counterpart in the reference, which has no serving loop;
 * each request runs the resident handle API of src/cuda_resident_api.cuh):
 * continuous batching of fit and predict requests on one GPU under an HBM
 * budget (bytes), with `workers` worker threads, at most max_batch_rows
 * rows per batched predict and at most max_concurrent_fits fits running at
 * once (<= 0: two; at most four), so predicts always find workers. Fits run
 * on one stream per fit slot and predict batches on two query streams,
 * whatever the number of resident models; fit slots and query streams
 * together stay within the process's hardware queues (GPU_MAX_HW_QUEUES, 4
 * unless set), so a predict never queues behind a fit on a shared hardware
 * queue (the fit cap is lowered to fit them). Fits are
 * admitted when their device-byte estimate fits the budget left by running
 * fits and resident models; a fit
 * whose estimate exceeds the whole budget is rejected at submission. Fitted
 * models stay resident (compacted) until released; the model id is the fit
 * request id. Predicts queued for the same model and prefix run as one
 * product over their concatenated rows; results are bitwise the same alone
 * and in a mix. Inputs and outputs must stay valid until the request is
 * polled as finished; a finished request can be polled once. Destroying the
 * server waits for running work and drops queued requests. */
void* fastpls_hip_server_create(double hbm_budget_bytes, int workers, int max_batch_rows,
                                int max_concurrent_fits, char* error, size_t error_capacity);
/* Returns the request id (> 0), 0 on error. Arguments as fastpls_hip_fit_device
 * (device_inputs != 0) or fastpls_hip_simpls_create / _plssvd_create. */
long long fastpls_hip_server_submit_fit(void* server, int method, const void* x, const void* y,
                                        const int* labels, int precision, int device_inputs,
                                        int n, int p, int q, int components, int scaling,
                                        int oversample, int power, unsigned long long seed,
                                        char* error, size_t error_capacity);
/* rows x p predictors with leading dimension ldx (host, or device memory when
 * device_inputs != 0); predictions rows x q (host) in the model's precision.
 * Returns the request id (> 0), 0 on error. */
long long fastpls_hip_server_submit_predict(void* server, long long model, const void* x,
                                            long long ldx, int device_inputs, int rows,
                                            int prefix, void* predictions, char* error,
                                            size_t error_capacity);
/* state: 0 queued, 1 running, 2 done, 3 failed (message in error; the call
 * itself returns 0). wait != 0 blocks until the request finishes. */
int fastpls_hip_server_poll(void* server, long long request, int wait, int* state,
                            double* latency_ms, char* error, size_t error_capacity);
int fastpls_hip_server_release(void* server, long long model, char* error,
                               size_t error_capacity);
/* Fields as fastpls_hip_export. */
int fastpls_hip_server_export(void* server, long long model, int field, void* out,
                              size_t elements, char* error, size_t error_capacity);
/* out[0..17]: completed, failed, predict p50 ms, predict p99 ms, fit p50 ms,
 * fit p99 ms, completed per second, peak bytes in use, budget bytes, peak
 * concurrent batches, predict batches, mean rows per predict batch, most
 * requests in one batch, seconds from first submission to last completion,
 * bytes in use, resident models, mean fit service ms (start to completion),
 * mean predict service ms (of the batch it ran in). */
int fastpls_hip_server_stats(void* server, double* out, int count, char* error,
                             size_t error_capacity);
void fastpls_hip_server_destroy(void* server);

/* fastsvd: replaces _fastPLS_fastsvd_core_cpp / _fastPLS_fastsvd_float32_core_cpp
 * (src/r_api.cpp:2521-2600, reached from fastsvd in R/main.R:8181-8204), the
 * audited operator rSVD of an explicit matrix
 * (inst/include/fastpls/core/audited_rsvd.hpp:298-414). x is m x n,
 * column-major, precision 32 or 64; with device = 1, x, U, s and Vt are
 * device pointers, otherwise host pointers. Writes U (m x kept), s (kept),
 * Vt (kept x n, untouched when left_only), *got = kept = min(k, m, n), and
 * audit[10] = {certified, deterministic_fallback, attempts,
 * effective_oversample, effective_power, effective_seed, subspace_error,
 * singular_value_error, triplet_residual, omitted_direction_ratio}. Fails
 * (returns 1, no result) when no attempt or recovery is certified, with the
 * reference's message, and when a sketch would exceed 512 columns. */
int fastpls_hip_rsvd(const void* x, int precision, int device, int m, int n, int k,
                     int oversample, int power, unsigned int seed, int left_only, double* U,
                     double* s, double* Vt, int* got, double* audit, char* error,
                     size_t error_capacity);

/* fastcor: replaces _fastPLS_fastcor_core_cpp (src/r_api.cpp:4144-4166) and
 * fastcor_values / normalized_units (src/r_api.cpp:4054-4142). Units are the
 * columns of a (a_rows x a_cols, column-major), or its rows when by_row; each
 * is centred, scaled to unit norm and rounded to the input precision (NaN
 * for a unit with a non-finite value or zero norm). out = L Rᵀ (units_a x
 * units_b, column-major FP64) of the units of a and b (a with itself when b
 * is NULL, which ignores diagonal), or the units_a matching-unit dot products
 * when diagonal. With device = 1, a, b and out are device pointers. */
int fastpls_hip_fastcor(const void* a, int a_rows, int a_cols, const void* b, int b_rows,
                        int b_cols, int precision, int device, int by_row, int diagonal,
                        double* out, char* error, size_t error_capacity);

/* ------------------------------------------------ small operators
 * LDA from score moments (no ABI counterpart in the reference; its core's
 * train_lda_prefixes_from_moments + finalize_lda_prefixes, lda.hpp:226-346,
 * and its resident CUDA LDA, src/cuda_resident_lda.cuh:41-121, which works
 * from device moments): gram (a x a) = TᵀT of the training scores,
 * class_sums (classes x a) the per-class score sums, counts the class sizes
 * (summing to n). For the first `prefix` components writes linear (classes x
 * prefix, column-major), constants (classes) and the ridge chosen from
 * {1e-8, 1e-6, 1e-5, 1e-4, 1e-3, 1e-2} x trace(pooled) / prefix; fails when
 * every ridge leaves the pooled covariance without a Cholesky factor.
 * prefix <= 64. */
int fastpls_hip_lda_moments(const double* gram, int a, const double* class_sums,
                            const int* counts, int classes, long long n, int prefix,
                            double* linear, double* constants, double* ridge, char* error,
                            size_t error_capacity);
/* PLS-SVD prefix weights (no ABI counterpart; plssvd.hpp:138-164): for k =
 * 1..components, L_k solves G_k L = diag(sigma_1..k) for the leading k x k
 * block of the score Gram (Cholesky, pivoted elimination when it fails,
 * linalg.hpp:57-177) and W_k = L_k Q_kᵀ (k x q, leading dimension k) is
 * written at weights + (k - 1) * components * q. Q is q x components.
 * components <= 64; fails when a prefix has no solution. */
int fastpls_hip_plssvd_weights(const double* gram, const double* sigma, const double* Q,
                               int components, int q, double* weights, char* error,
                               size_t error_capacity);

/* ------------------------------------------------- FP32 core hooks
 * Counterparts of the reference's hybrid accelerator hooks for its float
 * core (src/accelerator_core_backend.h:10-36: cuda_core_gemm_f32,
 * cuda_core_workspace_create_f32 / _destroy_f32, cuda_core_gemm_into_f32,
 * cuda_core_self_gram_into_f32; implemented in src/core_cuda_backend.cu:38-293
 * and called by RoutedLinearAlgebraF32, src/r_api.cpp:421-660). The C++
 * ConstMatrixView<float> / MatrixView<float> arguments become (pointer, rows,
 * columns, leading dimension) of host column-major matrices; bool returns
 * become 0 / 1 with the message in the error buffer (C++ exceptions never
 * cross the ABI). No CPU fallback: without a device every call fails.
 *
 * The workspace owns one stream and device buffers that grow as needed and
 * persist across calls. Products are FP32-accurate (bf16x6 matrix-core
 * products or FP64 VALU sums, FP64 accumulation) and rounded once to FP32. */
void* fastpls_hip_core_workspace_create_f32(char* error, size_t error_capacity);
void fastpls_hip_core_workspace_destroy_f32(void* workspace);
/* output (output_rows x output_columns, leading dimension output_ld) =
 * op(left) op(right), op = transpose when the flag is set. A product with an
 * empty dimension succeeds without touching output, as the reference's. */
int fastpls_hip_core_gemm_into_f32(void* workspace, const float* left, size_t left_rows,
                                   size_t left_columns, size_t left_ld, const float* right,
                                   size_t right_rows, size_t right_columns, size_t right_ld,
                                   int transpose_left, int transpose_right, float* output,
                                   size_t output_rows, size_t output_columns, size_t output_ld,
                                   char* error, size_t error_capacity);
/* output (d x d) = input inputᵀ (transpose_input 0, d = rows) or inputᵀ input
 * (d = columns): the lower triangle, with the strict upper triangle zero, or
 * mirrored from the lower one when full_output != 0 (cublasSsyrk lower +
 * mirror_lower_to_upper in the reference). */
int fastpls_hip_core_self_gram_into_f32(void* workspace, const float* input, size_t rows,
                                        size_t columns, size_t ld, int transpose_input,
                                        float* output, size_t output_rows, size_t output_columns,
                                        size_t output_ld, int full_output, char* error,
                                        size_t error_capacity);
/* cuda_core_gemm_f32: op(left) op(right) into output (contiguous, leading
 * dimension = its rows), with a workspace of its own for the one call. */
int fastpls_hip_core_gemm_f32(const float* left, size_t left_rows, size_t left_columns,
                              size_t left_ld, const float* right, size_t right_rows,
                              size_t right_columns, size_t right_ld, int transpose_left,
                              int transpose_right, float* output, char* error,
                              size_t error_capacity);

/* Synthetic latent-factor data written straight into device memory
 * (Philox4x32-10 + Box-Muller): X (n x p, precision 32/64) = T P' + noise E,
 * Y (n x q, same precision) = T Q' + noise F, with factor column f scaled by
 * decay^f. x_device / y_device are caller-allocated device buffers. With
 * round32 = 1 and precision 64 the FP64 buffers hold exactly the FP32 data
 * (every value rounded to FP32), so both precisions can fit identical inputs.
 * No reference counterpart (benchmark input). */
int fastpls_hip_latent_data(void* x_device, void* y_device, int precision, long long n,
                            int p, int q, int factors, double decay, double noise,
                            unsigned long long seed, int round32, char* error,
                            size_t error_capacity);

#ifdef __cplusplus
}
#endif
#endif
