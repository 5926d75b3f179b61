// Fold-by-fold cross-validation over the single-fit engine: classification
// CV for SIMPLS and PLS-SVD and regression CV for PLS-SVD
// (src/cuda_resident_api.cuh:538-1023). The full predictors are copied to
// the device once; every fold's training and held-out rows are gathered on
// the device and fitted in place, so a fold moves no predictor bytes over
// PCIe. (SIMPLS regression CV fits all folds in one batch, in batch.hip.)
#include "../../include/fastpls_hip.h"
#include "cv_common.hpp"
#include "handle.hpp"
#include "host_util.hpp"
#include "kernels.hpp"

#include <memory>
#include <stdexcept>
#include <vector>

namespace fpls {
namespace {

// Stream borrowed by a CV call; drained and given back on every exit path.
struct CvStream {
  hipStream_t s = nullptr;
  CvStream() : s(stream_take()) {}
  ~CvStream() { stream_give(s); }
};

// Device row gathers of one fold into buffers sized for the largest fold.
template<class T>
struct FoldRows {
  Buf index, rows;
  T* data() const { return rows.as<T>(); }
  void gather(const Buf& full, int n, int cols, const std::vector<int>& which, hipStream_t s) {
    index.reserve(sizeof(int) * which.size());
    rows.reserve(sizeof(T) * which.size() * (size_t)cols);
    FPLS_HIP(hipMemcpyAsync(index.as<int>(), which.data(), sizeof(int) * which.size(),
                            hipMemcpyHostToDevice, s));
    launch_gather_rows<T>(full.as<T>(), n, cols, index.as<int>(), (long long)which.size(),
                          rows.as<T>(), s);
  }
};

// api.cuh:538-768. Per fold: classes present in the training rows are
// renumbered 1..K'; a fold whose training rows hold at most one class
// predicts that class (or class 1) for every held-out row (status 4);
// otherwise the fold is fitted with K' responses and seed (+ fold for
// PLS-SVD) and every prefix's top-1 (first maximum wins) is mapped back to
// the original label. Metrics are accuracies per prefix.
template<class T>
void pls_cv_classification(int method, bool increment_seed, const T* predictors,
                           const int* labels, const int* folds, int n, int p, int classes,
                           const int* prefixes, int prefix_count, int scaling, bool lda,
                           int oversample, int power, unsigned long long seed,
                           bool store_predictions, bool store_scores, int* prediction_output,
                           T* score_output, int* status, double* metrics) {
  if (!predictors || !labels || !folds || !prefixes || !status || !metrics || n < 2 || p < 1 ||
      classes < 2 || prefix_count < 1 || scaling < 1 || scaling > 3)
    throw std::invalid_argument("invalid resident HIP CV input");
  const int maxp = cv_path_maximum(prefixes, prefix_count);
  for (int r = 0; r < n; ++r)
    if (labels[r] < 1 || labels[r] > classes || folds[r] < 1)
      throw std::invalid_argument("invalid HIP CV label or fold");
  if (maxp > std::min(n - 1, p)) throw std::invalid_argument("HIP CV component count exceeds rank");
  if (store_predictions && !prediction_output) throw std::invalid_argument("null HIP CV prediction output");
  if (store_scores && !score_output) throw std::invalid_argument("null HIP CV score output");
  const FoldSplit split(folds, n);
  split.require_usable();
  require_device();

  CvStream cs;
  Buf full, train_labels;
  full.alloc(sizeof(T) * (size_t)n * p);
  FPLS_HIP(hipMemcpyAsync(full.as<T>(), predictors, sizeof(T) * (size_t)n * p,
                          hipMemcpyHostToDevice, cs.s));
  FoldRows<T> train, test;
  std::fill(metrics, metrics + prefix_count, 0.0);
  std::vector<double> counted(prefix_count, 0.0);
  if (store_scores) std::fill(score_output, score_output + (size_t)n * classes * prefix_count, T(0));

  for (int fold = 0; fold < split.count; ++fold) {
    const std::vector<int>& heldout = split.test[fold];
    if (heldout.empty()) { status[fold] = 2; continue; }
    const std::vector<int> training = split.training(fold, n);
    std::vector<int> counts(classes, 0);
    for (int r : training) ++counts[labels[r] - 1];
    std::vector<int> active, active_map(classes, -1);
    for (int c = 0; c < classes; ++c)
      if (counts[c] > 0) {
        active_map[c] = int(active.size());
        active.push_back(c + 1);
      }
    if (active.size() <= 1) {
      const int fallback = active.empty() ? 1 : active.front();
      for (int j = 0; j < prefix_count; ++j)
        for (int r : heldout) {
          if (store_predictions) prediction_output[(size_t)j * n + r] = fallback;
          metrics[j] += labels[r] == fallback ? 1.0 : 0.0;
          counted[j] += 1.0;
        }
      status[fold] = 4;
      continue;
    }
    const int ntr = int(training.size()), nte = int(heldout.size()), K = int(active.size());
    if (maxp > std::min(ntr - 1, p)) throw std::invalid_argument("HIP CV component count exceeds a fold rank");
    std::vector<int> compact(ntr);
    for (int i = 0; i < ntr; ++i) compact[i] = active_map[labels[training[i]] - 1] + 1;
    train.gather(full, n, p, training, cs.s);
    test.gather(full, n, p, heldout, cs.s);
    train_labels.reserve(sizeof(int) * (size_t)ntr);
    FPLS_HIP(hipMemcpyAsync(train_labels.as<int>(), compact.data(), sizeof(int) * (size_t)ntr,
                            hipMemcpyHostToDevice, cs.s));
    // the model runs on its own stream
    FPLS_HIP(hipStreamSynchronize(cs.s));
    const unsigned long long fold_seed = seed + (increment_seed ? (unsigned long long)fold : 0ULL);
    std::unique_ptr<Handle> model(create(method, train.data(), nullptr,
                                         train_labels.as<int>(), sizeof(T) == 4 ? 32 : 64, true,
                                         ntr, p, K, maxp, scaling, oversample, power, fold_seed));
    std::vector<int> top((size_t)nte * prefix_count);
    std::vector<T> scores(store_scores ? (size_t)nte * K * prefix_count : 0);
    model->classify_path(test.data(), nte, prefixes, prefix_count, lda, 1, top.data(),
                         store_scores ? scores.data() : nullptr, true);
    model.reset();
    for (int j = 0; j < prefix_count; ++j) {
      for (int i = 0; i < nte; ++i) {
        const int predicted = active[top[(size_t)j * nte + i] - 1];
        const int r = heldout[i];
        if (store_predictions) prediction_output[(size_t)j * n + r] = predicted;
        metrics[j] += predicted == labels[r] ? 1.0 : 0.0;
        counted[j] += 1.0;
      }
      if (store_scores)
        for (int a = 0; a < K; ++a) {
          const int destination = active[a] - 1;
          for (int i = 0; i < nte; ++i)
            score_output[(size_t)j * n * classes + (size_t)destination * n + heldout[i]] =
                scores[(size_t)j * nte * K + (size_t)a * nte + i];
        }
    }
    status[fold] = 1;
  }
  for (int j = 0; j < prefix_count; ++j)
    metrics[j] = counted[j] > 0.0 ? metrics[j] / counted[j] : std::numeric_limits<double>::quiet_NaN();
}

// api.cuh:770-1023 with one fit per fold on the gathered rows, seed + fold
// for PLS-SVD.
template<class T>
void pls_cv_regression(int method, bool increment_seed, const T* predictors, const T* responses,
                       const int* folds, int n, int p, int q, const int* prefixes,
                       int prefix_count, int scaling, int metric, int oversample, int power,
                       unsigned long long seed, bool store_predictions,
                       double* prediction_output, int* status, double* metrics, double* q2,
                       double* rmsd, double* observed_r2) {
  if (!predictors || !responses || !folds || !prefixes || !status || !metrics || !q2 || !rmsd ||
      !observed_r2 || n < 2 || p < 1 || q < 1 || prefix_count < 1 || scaling < 1 ||
      scaling > 3 || metric < 2 || metric > 4)
    throw std::invalid_argument("invalid resident HIP regression CV input");
  const int maxp = cv_path_maximum(prefixes, prefix_count);
  const FoldSplit split(folds, n);
  if (maxp > std::min(n - 1, p) || (store_predictions && !prediction_output))
    throw std::invalid_argument("invalid HIP regression CV output or rank");
  split.require_usable();
  require_device();
  const long double total_ss = cv_total_ss(responses, n, q);

  CvStream cs;
  Buf fx, fy;
  fx.alloc(sizeof(T) * (size_t)n * p);
  fy.alloc(sizeof(T) * (size_t)n * q);
  FPLS_HIP(hipMemcpyAsync(fx.as<T>(), predictors, sizeof(T) * (size_t)n * p, hipMemcpyHostToDevice, cs.s));
  FPLS_HIP(hipMemcpyAsync(fy.as<T>(), responses, sizeof(T) * (size_t)n * q, hipMemcpyHostToDevice, cs.s));
  FoldRows<T> trx, try_, tex, tey;
  RegressionTally tally(prefix_count);
  for (int fold = 0; fold < split.count; ++fold) {
    const std::vector<int>& heldout = split.test[fold];
    if (heldout.empty()) { status[fold] = 2; continue; }
    const std::vector<int> training = split.training(fold, n);
    const int ntr = int(training.size()), nte = int(heldout.size());
    if (maxp > std::min(ntr - 1, p)) throw std::invalid_argument("HIP CV component count exceeds a fold rank");
    trx.gather(fx, n, p, training, cs.s);
    try_.gather(fy, n, q, training, cs.s);
    tex.gather(fx, n, p, heldout, cs.s);
    tey.gather(fy, n, q, heldout, cs.s);
    FPLS_HIP(hipStreamSynchronize(cs.s));
    const unsigned long long fold_seed = seed + (increment_seed ? (unsigned long long)fold : 0ULL);
    std::unique_ptr<Handle> model(create(method, trx.data(), try_.data(), nullptr,
                                         sizeof(T) == 4 ? 32 : 64, true, ntr, p, q, maxp,
                                         scaling, oversample, power, fold_seed));
    std::vector<double> sums(3 * (size_t)q * prefix_count);
    std::vector<double> pred(store_predictions ? (size_t)nte * q * prefix_count : 0);
    model->heldout_regression(tex.data(), tey.data(), nte, prefixes, prefix_count,
                              store_predictions ? pred.data() : nullptr, sums.data());
    model.reset();
    tally.add_fold(sums.data(), q, nte);
    if (store_predictions) scatter_predictions(pred.data(), heldout, n, q, prefix_count, prediction_output);
    status[fold] = 1;
  }
  tally.finish(total_ss, metric, metrics, q2, rmsd, observed_r2);
}

template<class F>
void by_precision(int precision, F&& f) {
  if (precision == 32) f(float());
  else if (precision == 64) f(double());
  else throw std::invalid_argument("precision must be 32 or 64");
}

int classification_entry(int method, const void* predictors, const int* labels, const int* folds,
                         int precision, int n, int p, int classes, const int* prefixes,
                         int prefix_count, int scaling, int lda, int oversample, int power,
                         unsigned long long seed, int store_predictions, int store_scores,
                         int* predictions, void* scores, int* status, double* metrics,
                         char* error, size_t error_capacity) {
  return guarded(error, error_capacity, [&] {
    by_precision(precision, [&](auto zero) {
      using T = decltype(zero);
      // PLS-SVD folds advance the seed; SIMPLS folds share it (api.cuh:1039-1045, 1110-1115)
      pls_cv_classification<T>(method, method == 1, static_cast<const T*>(predictors), labels,
                               folds, n, p, classes, prefixes, prefix_count, scaling, lda != 0,
                               oversample, power, seed, store_predictions != 0,
                               store_scores != 0, predictions, static_cast<T*>(scores), status,
                               metrics);
    });
  });
}

}  // namespace
}  // namespace fpls

extern "C" {

int fastpls_hip_simpls_cv_classification(const void* predictors, const int* labels,
                                         const int* folds, int precision, int n, int p,
                                         int classes, const int* prefixes, int prefix_count,
                                         int scaling, int lda, int oversample, int power,
                                         unsigned long long seed, int store_predictions,
                                         int store_scores, int* predictions, void* scores,
                                         int* status, double* metrics, char* error,
                                         size_t error_capacity) {
  return fpls::classification_entry(0, predictors, labels, folds, precision, n, p, classes,
                                    prefixes, prefix_count, scaling, lda, oversample, power, seed,
                                    store_predictions, store_scores, predictions, scores, status,
                                    metrics, error, error_capacity);
}

int fastpls_hip_plssvd_cv_classification(const void* predictors, const int* labels,
                                         const int* folds, int precision, int n, int p,
                                         int classes, const int* prefixes, int prefix_count,
                                         int scaling, int lda, int oversample, int power,
                                         unsigned long long seed, int store_predictions,
                                         int store_scores, int* predictions, void* scores,
                                         int* status, double* metrics, char* error,
                                         size_t error_capacity) {
  return fpls::classification_entry(1, predictors, labels, folds, precision, n, p, classes,
                                    prefixes, prefix_count, scaling, lda, oversample, power, seed,
                                    store_predictions, store_scores, predictions, scores, status,
                                    metrics, error, error_capacity);
}

int fastpls_hip_plssvd_cv_regression(const void* predictors, const void* responses,
                                     const int* folds, int precision, int n, int p, int q,
                                     const int* prefixes, int prefix_count, int scaling,
                                     int metric, int oversample, int power,
                                     unsigned long long seed, int store_predictions,
                                     double* predictions, int* status, double* metrics,
                                     double* q2, double* rmsd, double* observed_r2, char* error,
                                     size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    fpls::by_precision(precision, [&](auto zero) {
      using T = decltype(zero);
      fpls::pls_cv_regression<T>(1, true, static_cast<const T*>(predictors),
                                 static_cast<const T*>(responses), folds, n, p, q, prefixes,
                                 prefix_count, scaling, metric, oversample, power, seed,
                                 store_predictions != 0, predictions, status, metrics, q2, rmsd,
                                 observed_r2);
    });
  });
}

}  // extern "C"
