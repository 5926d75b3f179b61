// Descriptors and host launchers of the per-fit kernels (small_kernels.hip,
// prep_kernels.hip). All small-matrix arithmetic is FP64.
#pragma once

#include "common.hpp"

namespace fpls {

// One refresh of SIMPLS directions: component index, candidate block,
// sketch width and the offset of its Gaussian sketch in the sketch pool.
struct RefreshStep {
  int a;
  int b;
  int w;
  int pad;
  long long omega_off;
};

// Device-resident state of one SIMPLS fit (all pointers are device memory).
struct SimplsFitDev {
  int n, p, q, A;          // A = maximum number of components
  int power, use_right_gram, reorth, nsteps;
  int bmax, ts_splits, ls_splits, pad;
  long long ls_slab;       // distance between loading-pass split slabs
  const RefreshStep* steps;
  const double* omega;     // sketch pool
  double* S;               // p x q, deflated in place
  const double* S0;        // p x q, initial cross-covariance
  double* G;               // q x q right Gram (use_right_gram)
  const double* sx;        // p predictor scales
  double* cand;            // p x bmax candidate directions
  double* cand_scaled;     // p x bmax, cand / sx (operand of the score pass)
  double* t_raw;           // n x bmax reduced scores (input of the loading pass)
  double* ts_part;         // score-pass partials: ts_splits slabs of n x b
  double* ls_part;         // loading-pass partials: ls_splits slabs of p x b
  double* R;               // p x A
  double* Q;               // q x A
  double* V;               // p x A
  double* T;               // n x A
  double* scratch;         // refresh / update scratch
  int* state;              // [0] completed, [1] stopped, [2] error flags
  // matrix-free route (no S): Ỹᵀ t_raw (q x bmax) of the candidate block;
  // S, S0 and G are null and the update neither deflates S nor reads S0
  const double* qraw;
  // cached cross-product route (simpls.hpp:600-620): the p x p X_sᵀ X_s;
  // the update forms the loading X_sᵀ X_s r and |t|² = rᵀ X_sᵀ X_s r from it
  // (no loading pass; T, when not null, is t_raw / |t|)
  const double* xgram;
  // reorth: p x A loadings X_sᵀ T of the components so far (before their
  // orthogonalisation), so that the loading of a score with the earlier
  // scores projected out follows without another pass over X; needs T
  double* P;
};

// Error flags written to state[2].
enum FitFlags : int {
  kFlagBadLabel = 1,
  kFlagEmptyClass = 2,
  kFlagNonFinite = 4,
  kFlagDecomposition = 8,
  kFlagZeroNorm = 16,
};

// Doubles of refresh scratch needed by one fit.
long long simpls_scratch_doubles(int p, int q, int wmax);

void launch_simpls_refresh(const SimplsFitDev* fits, int count, int step, hipStream_t s);
void launch_simpls_update(const SimplsFitDev* fits, int count, int step, hipStream_t s);
// Steps of the multi-launch refresh and update of a wide single fit (the
// products between them run through the xw / xtw passes; state is the fit's
// [completed, stopped, flags]).
// Top-b eigenvectors of Bm Bmᵀ (Bm w x q), or of Bm itself (w x w) with gram.
void launch_refresh_eigen(const double* Bm, int w, int q, int b, double* vtop, double* scratch,
                          int* state, hipStream_t s, bool gram = false);
void launch_refresh_gate(int* state, hipStream_t s);
void launch_wide_component(long long n, int p, const double* t, const double* r,
                           const double* part, int splits, long long slab, const double* sx,
                           const double* ss, double* Tc, double* Rc, double* Vc, int* state,
                           hipStream_t s);
void launch_wide_subtract(double* v, const double* d, int p, const int* state, hipStream_t s);
void launch_wide_unit(double* v, int p, const double* vv, int* state, hipStream_t s);
void launch_wide_deflate(double* S, int p, int q, const double* v, const double* row,
                         const int* state, hipStream_t s);
void launch_wide_finish(double* G, int q, const double* row, int comp, int* state,
                        hipStream_t s);

// ---- single-work-group helpers for the PLS-SVD randomized SVD
// Householder QR: A (m x n) destroyed, Q (m x n) written. flags |= decomposition.
void launch_qr(double* A, int m, int n, double* Q, int* flags, hipStream_t s);
// CholQR2 pieces: Rinv = (chol(G))^{-T} for a w x w Gram (w <= 64), *fail set
// when the factorisation is unusable; the fallback recomputes Q by
// Householder from A only when *fail is set.
void launch_chol_inv(const double* G, int w, double* Rinv, int* fail, hipStream_t s);
// Whole CholQR2 (Householder when a Cholesky factor is unusable) of a small
// A (m x w, m >= w, w <= 32, m·w <= 4096) in one work-group; A is not modified.
bool small_qr_fits(int m, int w);
void launch_cholqr2_small(const double* A, int m, int w, double* Q, int* flags, hipStream_t s);
void launch_qr_fallback(double* A, int m, int n, double* Q, const int* fail, int* flags,
                        hipStream_t s);
// Finalize a sketch: P (w x q) = Qbᵀ S is given; writes Up (w x k), sigma (k),
// Vt (k x q) and *usable. Uses the eigen route when w <= q, the dense SVD
// route otherwise (rsvd.hpp:85-165). scratch: 4*w*w + 2*w*q + w doubles.
void launch_rsvd_finalize(const double* P, int w, int q, int k, double* Up,
                          double* sigma, double* Vt, int* usable, double* scratch,
                          int* flags, hipStream_t s);
// Dense economy SVD of a small square or tall matrix M (m x n, m >= n), the
// first k triplets: U (m x k), sigma (k), V (n x k, right vectors as columns).
// Eigen route of the operator rSVD's finalisation (operator_rsvd.hpp:64-168)
// on the projected Gram (w x w) of a block w wide and `cols` long: writes the
// top-r eigenvectors (left, w x r), the same divided by σ (left_scaled), σ (r)
// and status[0] = 1, or status[0] = 0 when the dense route has to run.
// scratch: 3 w² + w doubles (used when w > 64).
void launch_operator_finalize(const double* gram, int w, int cols, int r, double* left,
                              double* left_scaled, double* sigma, double* scratch, int* status,
                              hipStream_t s);
void launch_small_svd(const double* M, int m, int n, int k, double* U, double* sigma,
                      double* V, double* scratch, int* flags, hipStream_t s);

// PLS-SVD prediction weights for prefixes 1..A from G = TᵀT (A x A), sigma and
// Q (q x A): slab k-1 of W (A*q doubles) holds the k x q weights (ld k).
void launch_plssvd_weights(const double* G, const double* sigma, const double* Q,
                           int A, int q, double* W, int* flags, hipStream_t s);

// ---- preprocessing
// Column means of Y (n x q, element T) and centred copy Yc (FP64) plus the
// column sums of Yc (ysum) used by the shifted cross-product.
// `scratch` holds kCenterScratch doubles per response.
constexpr int kCenterScratch = 128;
template<class T>
void launch_center_response(const T* Y, long long n, int q, double* mean, double* Yc,
                            double* ysum, double* scratch, hipStream_t s);
// The same for many response columns in one launch (one work-group per
// column): column `y` (n values, read through `perm` when it is not null) ->
// *mean, yc (n doubles) and *ysum.
template<class T>
struct CenterColumn {
  const T* y;
  const int* perm;
  long long n;
  double* yc;
  double* mean;
  double* ysum;
  // row view (null: every row): rows with mask 0 are left out of the mean and
  // get a zero centred value; used = the number of rows with mask 1
  const int* mask;
  long long used;
};
template<class T>
void launch_center_columns(const CenterColumn<T>* device_list, int count, hipStream_t s);
// Labels (one-based) -> class counts, priors, centred one-hot Yc, plain
// one-hot H (may be null) and ysum; invalid labels / empty classes set flags.
void launch_label_response(const int* labels, long long n, int K, int* counts,
                           double* priors, double* Yc, double* H, double* ysum,
                           int* flags, hipStream_t s);
// From the stats/cross-product partials of a kShiftStats pass compute mean,
// scale and S (p x q), and copy S into S0 when S0 is non-null. Non-finite
// statistics set `flags`; a non-finite entry of S column c sets `flags`, or
// flags[cross_offsets[c]] when a device array of per-column offsets is given
// (columns of a batch belong to different fits).
template<class T>
void launch_finalize_crossprod(const T* X, long long n, int p, int q, int scaling,
                               const double* stats_part, const double* s_part, int splits,
                               const double* ysum, double* mean, double* scale,
                               double* colss, double* S, double* S0, int* flags,
                               hipStream_t s, int* cross_flags = nullptr,
                               const int* cross_offsets = nullptr,
                               const double* hold_stats = nullptr, long long used = -1);
// Row views: the shifted column sums and sums of squares (2 x p: sum d, sum
// d^2 with d = x - X[0, j], the shift of the statistics pass over the full X,
// ld n) of the held-out rows, from their gathered copy hx (hn x p); the
// statistics of the view are the full ones minus these.
template<class T>
void launch_holdout_stats(const T* hx, long long hn, int p, const T* X, long long n, double* out,
                          hipStream_t s);
// t[rows[i] + c ld] = 0 for i < count, c < cols (the held-out rows of a row
// view's scores, so that the products and norms that follow sum its rows only)
void launch_zero_rows(double* t, long long ld, int cols, const int* rows, int count,
                      hipStream_t s);
// out (n x cols, FP64) = (X[:, c0 + c] - mean[c0 + c]) / scale[c0 + c]: columns
// c0 .. c0 + cols of the standardised X (leading dimension n)
template<class T>
void launch_standardize_columns(const T* X, long long n, int c0, int cols, const double* mean,
                                const double* scale, double* out, hipStream_t s);
// out[i] = in[i] / d[i % rows]  (scale rows of a rows x cols matrix)
void launch_div_rows(const double* in, const double* d, long long rows, int cols,
                     double* out, hipStream_t s);
// The same for a device list of independent problems, one launch
// (work-groups over grid x, problems over grid y).
struct DivRows {
  const double* in;
  const double* d;
  long long rows;
  int cols;
  double* out;
};
void launch_div_rows_list(const DivRows* device_list, int count, long long max_elements,
                          hipStream_t s);

// ---- prediction / classification helpers
// out (rows x q, element T and/or FP64) = Tsc[:, :k] (ld rows) * M + bias,
// with M(l, c) = Mr[l * rs + c * cs].
template<class T>
void launch_apply_small(const double* Tsc, long long rows, int k, const double* Mr,
                        long long rs, long long cs, int q, const double* bias, T* out,
                        double* out64, hipStream_t s);
// out (cols x rows) = inᵀ
void launch_transpose(const double* in, long long rows, long long cols, double* out,
                      hipStream_t s);
// Per-row top-k (one-based indices), first maximum wins; values (rows x K)
// plus optional per-class constants.
void launch_topk(const double* values, long long rows, int K, const double* constants,
                 int top, int* out, hipStream_t s);
// Three sums per response (SSE, SS about training mean, SS about observed mean).
template<class T>
void launch_response_sums(const double* pred, const T* observed, const int* labels,
                          const double* train_mean, long long n, int q, double* out,
                          int* flags, hipStream_t s);

// LDA: from score Gram (a x a), class sums (K x a, ld K), counts and n,
// prepare prefix k: linear (K x k), constants (K), ridge.
void launch_lda_prepare(const double* gram, const double* class_sums, const int* counts,
                        long long n, int a, int K, int k, double* linear,
                        double* constants, double* ridge, double* scratch, int* flags,
                        hipStream_t s);

// Variance explained (cuda_resident_variance.cuh semantics) from cross =
// X_sᵀT (p x a), gram = TᵀT and the standardized column sums of squares.
void launch_variance(const double* cross, const double* gram, const double* colss,
                     int p, int a, double* loadings, double* ss, double* scratch,
                     hipStream_t s);

// p[0..count) = value (a kernel, so streams stay graph-capturable).
void launch_fill_int(int* p, int count, int value, hipStream_t s);
// y += alpha x over `count` doubles.
void launch_axpy(double alpha, const double* x, double* y, long long count, hipStream_t s);
// Columns of v (rows x cols) scaled to unit norm; a column that cannot be
// (norm non-finite or <= eps) is left as it is and sets *stop to 1.
void launch_unit_columns(double* v, long long rows, int cols, int* stop, hipStream_t s);

// Block copy: dst (rows x cols, leading dimension ldd) = src (leading dimension lds).
template<class T>
void launch_copy_block(const T* src, long long lds, long long rows, long long cols, T* dst,
                       long long ldd, hipStream_t s);
// Row gather: dst (count x cols) = src[idx, :] for a column-major src (n x cols).
template<class T>
void launch_gather_rows(const T* src, long long n, long long cols, const int* idx, long long count,
                        T* dst, hipStream_t s);

// Synthetic latent-factor data on the device (Philox4x32-10 + Box-Muller).
// round32 rounds every value to FP32 before it is stored (in T).
template<class T>
void launch_latent_data(T* X, double* Y, long long n, int p, int q, int factors,
                        double decay, double noise, unsigned long long seed, int round32,
                        double* scratch, hipStream_t s);

template<class T>
void launch_convert(const double* in, T* out, long long count, hipStream_t s);

}  // namespace fpls
