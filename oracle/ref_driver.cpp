// Test infrastructure only: a C-ABI driver around the reference's own
// header-only core (inst/include/fastpls/core) and its test backend
// (core/tests/reference_backend.hpp). It is compiled from the sources where
// they lie under the reference checkout by oracle/Makefile into
// oracle/_ref/libfastpls_ref.so and is used only to write golden vectors and
// to cross-check the oracle restatement (oracle/fastpls_oracle.cpp).
//
// The routing of src/r_api.cpp:1149-1190 (simpls_controls) cannot be compiled
// here (that file needs R headers), so the same default thresholds are passed
// in explicitly by this driver.

#include <fastpls/core.hpp>
#include <fastpls/core/audited_rsvd.hpp>

#include "reference_backend.hpp"

#include <algorithm>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

namespace {
using fastpls::core::ConstMatrixView;
using fastpls::core::Matrix;
using fastpls::core::PredictorScaling;
thread_local std::string g_error;

template<class F>
int guarded(F&& f) {
  try { f(); return 0; } catch (const std::exception& e) { g_error = e.what(); return 1; }
}

Matrix<double> load(const double* x, std::size_t r, std::size_t c) {
  Matrix<double> m(r, c);
  std::memcpy(m.data(), x, sizeof(double) * r * c);
  return m;
}

void store(const Matrix<double>& m, double* out) {
  if (out && m.size()) std::memcpy(out, m.data(), sizeof(double) * m.size());
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

// simpls_candidate_block_size (simpls.hpp:43-62) as compiled from the header.
long long ref_block_size(long long remaining, long long p, long long q, int classification,
                         long long n, long long maximum) {
  return (long long)fastpls::core::simpls_candidate_block_size(
      std::size_t(remaining), std::size_t(p), std::size_t(q), classification != 0,
      std::size_t(n), std::size_t(maximum));
}

// Preprocess (dense or one-based labels) and fit SIMPLS with the given
// controls. Outputs: R p*A, Q q*A, V p*A, T n*A, mx p, sx p, my q.
int ref_simpls_fit(const double* x, const double* y, const int* labels, int n,
                   int p, int q, int A, int scaling, int oversample, int power,
                   unsigned int seed, int block, int cache, int reorth,
                   int batch_geometry, double* R, double* Q, double* V,
                   double* T, double* mx, double* sx, double* my,
                   int* completed) {
  return guarded([&] {
    ReferenceBackend<double> backend;
    Matrix<double> X = load(x, n, p);
    Matrix<double> S;
    std::vector<double> vmx, vsx, vmy;
    if (labels) {
      std::vector<std::size_t> zero_based(n);
      for (int i = 0; i < n; ++i) zero_based[i] = std::size_t(labels[i] - 1);
      auto pr = fastpls::core::prepare_scaled_label_crossprod(
        X.view(), zero_based.data(), std::size_t(n), std::size_t(q),
        static_cast<PredictorScaling>(scaling), backend);
      S = std::move(pr.crossprod);
      vmx = pr.predictor_center; vsx = pr.predictor_scale; vmy = pr.response_mean;
    } else {
      Matrix<double> Y = load(y, n, q);
      auto pr = fastpls::core::prepare_scaled_dense_crossprod(
        X.view(), ConstMatrixView<double>(Y.view()),
        static_cast<PredictorScaling>(scaling), backend);
      S = std::move(pr.crossprod);
      vmx = pr.predictor_center; vsx = pr.predictor_scale; vmy = pr.response_mean;
    }
    fastpls::core::SimplsControls controls;
    controls.components = std::size_t(A);
    controls.maximum_block = std::size_t(block);
    controls.cache_predictor_crossprod = cache != 0;
    controls.batch_candidate_geometry = batch_geometry != 0;
    controls.reorthogonalize = reorth != 0;
    controls.store_scores = true;
    controls.use_right_gram = true;
    controls.rsvd.oversample = oversample;
    controls.rsvd.power = power;
    controls.rsvd.seed = seed;
    fastpls::core::SimplsWorkspace<double> workspace;
    auto model = fastpls::core::fit_simpls_preprocessed<double>(
      ConstMatrixView<double>(X.view()), ConstMatrixView<double>(S.view()),
      controls, backend, workspace);
    store(model.weights, R); store(model.response_loadings, Q);
    store(model.deflation_basis, V); store(model.scores, T);
    std::copy(vmx.begin(), vmx.end(), mx);
    std::copy(vsx.begin(), vsx.end(), sx);
    std::copy(vmy.begin(), vmy.end(), my);
    *completed = int(model.completed_components);
  });
}

// Preprocess and fit PLS-SVD for prefixes 1..A. W holds A slabs of A*q;
// slab k-1 holds the k x q prediction weights (leading dimension k).
int ref_plssvd_fit(const double* x, const double* y, const int* labels, int n,
                   int p, int q, int A, int scaling, int oversample, int power,
                   unsigned int seed, double* R, double* Q, double* T,
                   double* sigma, double* W, double* mx, double* sx,
                   double* my) {
  return guarded([&] {
    ReferenceBackend<double> backend;
    Matrix<double> X = load(x, n, p);
    Matrix<double> S;
    std::vector<double> vmx, vsx, vmy;
    if (labels) {
      std::vector<std::size_t> zero_based(n);
      for (int i = 0; i < n; ++i) zero_based[i] = std::size_t(labels[i] - 1);
      auto pr = fastpls::core::prepare_scaled_label_crossprod(
        X.view(), zero_based.data(), std::size_t(n), std::size_t(q),
        static_cast<PredictorScaling>(scaling), backend);
      S = std::move(pr.crossprod);
      vmx = pr.predictor_center; vsx = pr.predictor_scale; vmy = pr.response_mean;
    } else {
      Matrix<double> Y = load(y, n, q);
      auto pr = fastpls::core::prepare_scaled_dense_crossprod(
        X.view(), ConstMatrixView<double>(Y.view()),
        static_cast<PredictorScaling>(scaling), backend);
      S = std::move(pr.crossprod);
      vmx = pr.predictor_center; vsx = pr.predictor_scale; vmy = pr.response_mean;
    }
    std::vector<int> comps(A);
    for (int k = 0; k < A; ++k) comps[k] = k + 1;
    fastpls::core::PlssvdControls controls;
    controls.rsvd.oversample = oversample;
    controls.rsvd.power = power;
    controls.rsvd.seed = seed;
    auto model = fastpls::core::fit_plssvd_preprocessed(
      ConstMatrixView<double>(X.view()), ConstMatrixView<double>(S.view()),
      comps.data(), comps.size(), controls, backend);
    store(model.weights, R); store(model.response_loadings, Q);
    store(model.scores, T);
    std::copy(model.singular_values.begin(), model.singular_values.end(), sigma);
    for (int k = 1; k <= A; ++k)
      std::memcpy(W + std::size_t(k - 1) * A * q, model.prediction_weights[k - 1].data(),
                  sizeof(double) * k * q);
    std::copy(vmx.begin(), vmx.end(), mx);
    std::copy(vsx.begin(), vsx.end(), sx);
    std::copy(vmy.begin(), vmy.end(), my);
  });
}

// Dense preprocessing, then the matrix-free SIMPLS of simpls.hpp:768-1047
// with the reference's centred cross-covariance and projected operators
// (operators.hpp:58-301) and the given controls. Outputs as ref_simpls_fit.
int ref_simpls_operator_fit(const double* x, const double* y, int n, int p, int q, int A,
                            int scaling, int oversample, int power, unsigned int seed,
                            int block, int cache, int reorth, int batch_geometry, int rank_one,
                            double* R, double* Q, double* V, double* T, double* mx, double* sx,
                            double* my, int* completed) {
  return guarded([&] {
    ReferenceBackend<double> backend;
    Matrix<double> X = load(x, n, p);
    Matrix<double> Y = load(y, n, q);
    auto pr = fastpls::core::prepare_scaled_dense_crossprod(
      X.view(), ConstMatrixView<double>(Y.view()),
      static_cast<PredictorScaling>(scaling), backend);
    fastpls::core::SimplsControls controls;
    controls.components = std::size_t(A);
    controls.maximum_block = std::size_t(block);
    controls.cache_predictor_crossprod = cache != 0;
    controls.batch_candidate_geometry = batch_geometry != 0;
    controls.reorthogonalize = reorth != 0;
    controls.rank_one_operator_direction = rank_one != 0;
    controls.store_scores = true;
    controls.rsvd.oversample = oversample;
    controls.rsvd.power = power;
    controls.rsvd.seed = seed;
    using Crosscov = fastpls::core::CenteredCrosscovOperator<double, ReferenceBackend<double>>;
    Crosscov initial(ConstMatrixView<double>(X.view()), ConstMatrixView<double>(Y.view()),
                     pr.response_mean.data(), pr.response_mean.size(), backend);
    fastpls::core::ProjectedOperator<double, Crosscov, ReferenceBackend<double>> projected(
      initial, controls.components, backend);
    fastpls::core::SimplsWorkspace<double> workspace;
    fastpls::core::OperatorRsvdWorkspace<double> rsvd_workspace;
    auto model = fastpls::core::fit_simpls_operator<double>(
      ConstMatrixView<double>(X.view()), initial, projected, controls, backend, workspace,
      rsvd_workspace);
    store(model.weights, R); store(model.response_loadings, Q);
    store(model.deflation_basis, V); store(model.scores, T);
    std::copy(pr.predictor_center.begin(), pr.predictor_center.end(), mx);
    std::copy(pr.predictor_scale.begin(), pr.predictor_scale.end(), sx);
    std::copy(pr.response_mean.begin(), pr.response_mean.end(), my);
    *completed = int(model.completed_components);
  });
}

// Dense preprocessing, then the matrix-free PLS-SVD of plssvd.hpp:222-248;
// outputs as ref_plssvd_fit.
int ref_plssvd_operator_fit(const double* x, const double* y, int n, int p, int q, int A,
                            int scaling, int oversample, int power, unsigned int seed,
                            double* R, double* Q, double* T, double* sigma, double* W,
                            double* mx, double* sx, double* my) {
  return guarded([&] {
    ReferenceBackend<double> backend;
    Matrix<double> X = load(x, n, p);
    Matrix<double> Y = load(y, n, q);
    auto pr = fastpls::core::prepare_scaled_dense_crossprod(
      X.view(), ConstMatrixView<double>(Y.view()),
      static_cast<PredictorScaling>(scaling), backend);
    using Crosscov = fastpls::core::CenteredCrosscovOperator<double, ReferenceBackend<double>>;
    Crosscov crosscov(ConstMatrixView<double>(X.view()), ConstMatrixView<double>(Y.view()),
                      pr.response_mean.data(), pr.response_mean.size(), backend);
    std::vector<int> comps(A);
    for (int k = 0; k < A; ++k) comps[k] = k + 1;
    fastpls::core::PlssvdControls controls;
    controls.rsvd.oversample = oversample;
    controls.rsvd.power = power;
    controls.rsvd.seed = seed;
    fastpls::core::OperatorRsvdWorkspace<double> workspace;
    auto model = fastpls::core::fit_plssvd_operator<double>(
      ConstMatrixView<double>(X.view()), crosscov, comps.data(), comps.size(), controls,
      backend, workspace);
    store(model.weights, R); store(model.response_loadings, Q);
    store(model.scores, T);
    std::copy(model.singular_values.begin(), model.singular_values.end(), sigma);
    for (int k = 1; k <= A; ++k)
      std::memcpy(W + std::size_t(k - 1) * A * q, model.prediction_weights[k - 1].data(),
                  sizeof(double) * k * q);
    std::copy(pr.predictor_center.begin(), pr.predictor_center.end(), mx);
    std::copy(pr.predictor_scale.begin(), pr.predictor_scale.end(), sx);
    std::copy(pr.response_mean.begin(), pr.response_mean.end(), my);
  });
}

// OPLS as the reference core composes it. Dense responses: fit_opls and
// predict_opls (oplspls.hpp:33-122: the randomized orthogonal filter of
// opls.hpp:314-363, then SIMPLS on the filtered predictors). One-based
// labels: the label filter of opls.hpp:418-465, then SIMPLS on the label
// cross-product of the filtered predictors (scaling none), as the resident
// engine composes them (src/cuda_resident_special.cuh:302-305). Outputs: W,
// P (p x north), *filtered, inner R, Q, T (p x A, q x A, n x A), outer mx,
// sx, my, *completed; dense fits also predict `rows` test rows at `prefix`
// components into pred (rows x q).
int ref_opls_fit(const double* x, const double* y, const int* labels, int n, int p, int q,
                 int A, int north, int scaling, int oversample, int power, unsigned int seed,
                 int block, int cache, int reorth, int batch_geometry, double* W, double* P,
                 int* filtered, double* R, double* Q, double* T, double* mx, double* sx,
                 double* my, int* completed, const double* xtest, int rows, int prefix,
                 double* pred) {
  return guarded([&] {
    ReferenceBackend<double> backend;
    fastpls::core::RsvdControls rc;
    rc.oversample = oversample; rc.power = power; rc.seed = seed;
    fastpls::core::SimplsControls controls;
    controls.components = std::size_t(A);
    controls.maximum_block = std::size_t(block);
    controls.cache_predictor_crossprod = cache != 0;
    controls.batch_candidate_geometry = batch_geometry != 0;
    controls.reorthogonalize = reorth != 0;
    controls.store_scores = true;
    controls.use_right_gram = true;
    controls.rsvd = rc;
    if (!labels) {
      Matrix<double> Y = load(y, n, q);
      fastpls::core::OplsControls oc;
      oc.orthogonal_components = std::size_t(north);
      oc.scaling = static_cast<PredictorScaling>(scaling);
      oc.randomized_filter = true;
      oc.filter_rsvd = rc;
      oc.simpls = controls;
      auto model = fastpls::core::fit_opls(load(x, n, p), ConstMatrixView<double>(Y.view()), oc,
                                           backend);
      store(model.filter.weights, W); store(model.filter.loadings, P);
      *filtered = int(model.filter.completed_components);
      store(model.inner.weights, R); store(model.inner.response_loadings, Q);
      store(model.inner.scores, T);
      std::copy(model.filter.predictor_center.begin(), model.filter.predictor_center.end(), mx);
      std::copy(model.filter.predictor_scale.begin(), model.filter.predictor_scale.end(), sx);
      std::copy(model.response_mean.begin(), model.response_mean.end(), my);
      *completed = int(model.inner.completed_components);
      if (rows > 0) {
        auto yh = fastpls::core::predict_opls(model, load(xtest, rows, p), std::size_t(prefix),
                                              backend);
        store(yh, pred);
      }
      return;
    }
    std::vector<std::size_t> zero_based(n);
    for (int i = 0; i < n; ++i) zero_based[i] = std::size_t(labels[i] - 1);
    auto filter = fastpls::core::fit_opls_filter_labels_rsvd(
      load(x, n, p), zero_based.data(), std::size_t(n), std::size_t(q), std::size_t(north),
      static_cast<PredictorScaling>(scaling), rc, backend);
    auto pr = fastpls::core::prepare_scaled_label_crossprod(
      filter.predictors.view(), zero_based.data(), std::size_t(n), std::size_t(q),
      PredictorScaling::none, backend);
    fastpls::core::SimplsWorkspace<double> workspace;
    auto model = fastpls::core::fit_simpls_preprocessed<double>(
      ConstMatrixView<double>(filter.predictors.view()), ConstMatrixView<double>(pr.crossprod.view()),
      controls, backend, workspace);
    store(filter.weights, W); store(filter.loadings, P);
    *filtered = int(filter.completed_components);
    store(model.weights, R); store(model.response_loadings, Q); store(model.scores, T);
    std::copy(filter.predictor_center.begin(), filter.predictor_center.end(), mx);
    std::copy(filter.predictor_scale.begin(), filter.predictor_scale.end(), sx);
    std::copy(pr.response_mean.begin(), pr.response_mean.end(), my);
    *completed = int(model.completed_components);
  });
}

// Kernel PLS (kernelpls.hpp:87-189): fit with the given kernel controls and
// SIMPLS controls, then predict `rows` test rows at `prefix` components.
// Outputs: R (n x A), Q, T, mx, sx, my, kernel column means (n), grand mean,
// *completed, pred (rows x q).
int ref_kernelpls_fit(const double* x, const double* y, int n, int p, int q, int A, int scaling,
                      int kind, double gamma, int degree, double offset, int oversample,
                      int power, unsigned int seed, int block, int cache, int reorth,
                      int batch_geometry, double* R, double* Q, double* T, double* mx,
                      double* sx, double* my, double* colmeans, double* grand,
                      int* completed, const double* xtest, int rows, int prefix,
                      double* pred) {
  return guarded([&] {
    ReferenceBackend<double> backend;
    Matrix<double> Y = load(y, n, q);
    fastpls::core::KernelPlsControls controls;
    controls.kernel = static_cast<fastpls::core::KernelType>(kind);
    controls.gamma = gamma; controls.degree = degree; controls.offset = offset;
    controls.scaling = static_cast<PredictorScaling>(scaling);
    controls.simpls.components = std::size_t(A);
    controls.simpls.maximum_block = std::size_t(block);
    controls.simpls.cache_predictor_crossprod = cache != 0;
    controls.simpls.batch_candidate_geometry = batch_geometry != 0;
    controls.simpls.reorthogonalize = reorth != 0;
    controls.simpls.store_scores = true;
    controls.simpls.use_right_gram = true;
    controls.simpls.rsvd.oversample = oversample;
    controls.simpls.rsvd.power = power;
    controls.simpls.rsvd.seed = seed;
    auto model = fastpls::core::fit_kernelpls(load(x, n, p), ConstMatrixView<double>(Y.view()),
                                              controls, backend);
    store(model.inner.weights, R); store(model.inner.response_loadings, Q);
    store(model.inner.scores, T);
    std::copy(model.predictor_center.begin(), model.predictor_center.end(), mx);
    std::copy(model.predictor_scale.begin(), model.predictor_scale.end(), sx);
    std::copy(model.response_mean.begin(), model.response_mean.end(), my);
    std::copy(model.kernel_column_means.begin(), model.kernel_column_means.end(), colmeans);
    *grand = model.kernel_grand_mean;
    *completed = int(model.inner.completed_components);
    if (rows > 0) {
      auto yh = fastpls::core::predict_kernelpls(model, load(xtest, rows, p),
                                                 std::size_t(prefix), backend);
      store(yh, pred);
    }
  });
}

// Explicit randomized SVD of a dense matrix (rsvd.hpp:185-237).
int ref_rsvd(const double* a, int m, int n, int k, int oversample, int power,
             unsigned int seed, double* U, double* s, double* Vt, int* got) {
  return guarded([&] {
    ReferenceBackend<double> backend;
    Matrix<double> A = load(a, m, n);
    fastpls::core::RsvdControls controls;
    controls.oversample = oversample; controls.power = power; controls.seed = seed;
    auto t = fastpls::core::randomized_svd(ConstMatrixView<double>(A.view()), k, controls, backend);
    store(t.U, U);
    std::copy(t.singular_values.begin(), t.singular_values.end(), s);
    store(t.Vt, Vt);
    *got = int(t.singular_values.size());
  });
}

// Audited operator rSVD of an explicit matrix (audited_rsvd.hpp:298-414);
// outputs as orc_audited_rsvd.
int ref_audited_rsvd(const double* a, int m, int n, int k, int oversample, int power,
                     unsigned int seed, int left_only, double* U, double* s, double* Vt,
                     int* got, double* audit) {
  return guarded([&] {
    ReferenceBackend<double> backend;
    Matrix<double> A = load(a, m, n);
    fastpls::core::ExplicitOperator<double, ReferenceBackend<double>> input(
        ConstMatrixView<double>(A.view()), backend);
    fastpls::core::RsvdControls controls;
    controls.oversample = oversample; controls.power = power; controls.seed = seed;
    controls.left_only = left_only != 0;
    auto r = fastpls::core::audited_operator_rsvd<double>(input, k, controls, backend);
    store(r.decomposition.U, U);
    std::copy(r.decomposition.singular_values.begin(), r.decomposition.singular_values.end(), s);
    if (!left_only) store(r.decomposition.Vt, Vt);
    *got = int(r.decomposition.singular_values.size());
    const auto& au = r.audit;
    const double v[10] = {double(au.certified), double(au.deterministic_fallback),
                          double(au.attempts), double(au.effective_oversample),
                          double(au.effective_power), double(au.effective_seed),
                          au.subspace_error, au.singular_value_error, au.triplet_residual,
                          au.omitted_direction_ratio};
    std::copy(v, v + 10, audit);
  });
}

// LDA on training scores, then discriminants and first-max labels of a test
// score block (lda.hpp:156-346). labels one-based.
int ref_lda(const double* T, int n, int a, const int* labels, int K, int prefix,
            const double* test_scores, int rows, double* disc, int* predicted,
            double* ridge) {
  return guarded([&] {
    Matrix<double> scores = load(T, n, a);
    const int prefixes[] = {prefix};
    auto models = fastpls::core::train_lda_prefixes(
      ConstMatrixView<double>(scores.view()), labels, std::size_t(n),
      std::size_t(K), prefixes, 1);
    Matrix<double> test = load(test_scores, rows, prefix);
    auto d = fastpls::core::lda_scores(ConstMatrixView<double>(test.view()), models[0]);
    store(d, disc);
    auto lab = fastpls::core::lda_predict(ConstMatrixView<double>(test.view()), models[0]);
    std::copy(lab.begin(), lab.end(), predicted);
    *ridge = models[0].ridge;
  });
}

// LDA from score moments: the reference's train_lda_prefixes_from_moments
// (lda.hpp:226-300) with finalize_lda_prefixes. linear K x prefix.
int ref_lda_moments(const double* gram, int a, const double* sums, const double* counts, int K,
                    long long n, int prefix, double* linear, double* constants, double* ridge,
                    double* relative) {
  return guarded([&] {
    Matrix<double> g = load(gram, a, a);
    Matrix<double> cs = load(sums, K, a);
    const int prefixes[] = {prefix};
    auto models = fastpls::core::train_lda_prefixes_from_moments(
      ConstMatrixView<double>(g.view()), ConstMatrixView<double>(cs.view()), counts,
      std::size_t(K), std::size_t(n), prefixes, 1);
    store(models[0].linear, linear);
    std::copy(models[0].constants.begin(), models[0].constants.end(), constants);
    *ridge = models[0].ridge;
    *relative = models[0].relative_ridge;
  });
}

// PLS-SVD prefix weights as plssvd.hpp:138-164 composes them, with the
// reference's own solvers (linalg.hpp:57-177): route[k - 1] = 0 when its
// Cholesky solve succeeds, 1 when the pivoted solve is needed.
int ref_plssvd_weights(const double* G, const double* sigma, const double* Q, int A, int q,
                       double* W, int* route) {
  return guarded([&] {
    for (int k = 1; k <= A; ++k) {
      Matrix<double> gk(k, k), dg(k, k), lat;
      for (int j = 0; j < k; ++j) {
        dg(j, j) = sigma[j];
        for (int i = 0; i < k; ++i) gk(i, j) = G[i + j * A];
      }
      if (fastpls::core::cholesky_solve(ConstMatrixView<double>(gk.view()),
                                        ConstMatrixView<double>(dg.view()), lat)) {
        route[k - 1] = 0;
      } else if (fastpls::core::pivoted_solve(ConstMatrixView<double>(gk.view()),
                                              ConstMatrixView<double>(dg.view()), lat)) {
        route[k - 1] = 1;
      } else {
        throw std::runtime_error("fastPLS PLS-SVD latent solve failed");
      }
      Matrix<double> wk(k, q);
      ConstMatrixView<double> qk(Q, q, k, q);
      fastpls::core::reference_gemm(ConstMatrixView<double>(lat.view()), qk, false, true,
                                    wk.view());
      store(wk, W + std::size_t(k - 1) * A * q);
    }
  });
}

}  // extern "C"
