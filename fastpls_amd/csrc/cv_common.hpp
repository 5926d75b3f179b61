// Host-side pieces shared by the cross-validation drivers: fold bookkeeping,
// input validation and the regression metric ladder of
// src/cuda_resident_api.cuh:538-1023.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

namespace fpls {

// Rows of every fold: test rows are the rows carrying the fold id, training
// rows the rest, both in increasing row order (api.cuh:566-575, 617-627).
struct FoldSplit {
  int count = 0;
  int maximum_test = 0;
  int maximum_train = 0;
  std::vector<std::vector<int>> test;

  FoldSplit(const int* folds, int n) {
    for (int r = 0; r < n; ++r) {
      if (folds[r] < 1) throw std::invalid_argument("invalid HIP CV fold");
      count = std::max(count, folds[r]);
    }
    test.resize(count);
    for (int r = 0; r < n; ++r) test[folds[r] - 1].push_back(r);
    for (const auto& t : test) {
      maximum_test = std::max(maximum_test, int(t.size()));
      maximum_train = std::max(maximum_train, n - int(t.size()));
    }
  }

  std::vector<int> training(int fold, int n) const {
    std::vector<char> held(n, 0);
    for (int r : test[fold]) held[r] = 1;
    std::vector<int> rows;
    rows.reserve(n - test[fold].size());
    for (int r = 0; r < n; ++r)
      if (!held[r]) rows.push_back(r);
    return rows;
  }

  // api.cuh:571-575: at least one fold with test rows and one training set
  // of two rows
  void require_usable() const {
    if (maximum_test < 1 || maximum_train < 2)
      throw std::invalid_argument("HIP CV contains an empty train or test fold");
  }
};

// Strictly increasing prefixes (api.cuh:548-553); returns the largest.
inline int cv_path_maximum(const int* prefixes, int count) {
  int maximum = 0;
  for (int j = 0; j < count; ++j) {
    if (prefixes[j] <= maximum)
      throw std::invalid_argument("HIP CV component path must be strictly increasing");
    maximum = prefixes[j];
  }
  return maximum;
}

// Σ over responses of the sum of squares about the full-data mean, in long
// double (api.cuh:805-815).
template<class T>
long double cv_total_ss(const T* responses, int n, int q) {
  long double total = 0.0L;
  for (int c = 0; c < q; ++c) {
    long double s = 0.0L, s2 = 0.0L;
    for (int r = 0; r < n; ++r) {
      const long double v = responses[(size_t)c * n + r];
      s += v;
      s2 += v * v;
    }
    total += s2 - s * s / n;
  }
  return total;
}

// Regression CV accumulators: SSE, SS about the training mean and value
// counts per prefix (api.cuh:963-994), then the metric ladder 2 = observed
// R², 3 = Q², 4 = RMSD (api.cuh:998-1013).
struct RegressionTally {
  std::vector<double> sse, training_ss, counts;
  explicit RegressionTally(int prefixes)
      : sse(prefixes, 0.0), training_ss(prefixes, 0.0), counts(prefixes, 0.0) {}

  // sums: 3 x q per prefix (SSE, SS about the training mean, SS about the
  // held-out mean), for `rows` held-out rows
  void add_fold(const double* sums, int q, int rows) {
    for (size_t j = 0; j < sse.size(); ++j) {
      long double e = 0.0L, t = 0.0L;
      for (int c = 0; c < q; ++c) {
        e += sums[j * 3 * q + (size_t)c * 3];
        t += sums[j * 3 * q + (size_t)c * 3 + 1];
      }
      sse[j] += double(e);
      training_ss[j] += double(t);
      counts[j] += double(rows) * q;
    }
  }

  void finish(long double total_ss, int metric, double* metrics, double* q2, double* rmsd,
              double* observed_r2) const {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (size_t j = 0; j < sse.size(); ++j) {
      rmsd[j] = counts[j] > 0.0 ? std::sqrt(sse[j] / counts[j]) : nan;
      q2[j] = training_ss[j] > 0.0 ? 1.0 - sse[j] / training_ss[j] : nan;
      observed_r2[j] = total_ss > 0.0L ? 1.0 - sse[j] / double(total_ss) : nan;
      metrics[j] = metric == 4 ? rmsd[j] : metric == 3 ? q2[j] : observed_r2[j];
    }
  }
};

// Held-out predictions (rows x q x prefixes, fold order) into the full
// output (n x q x prefixes) at the fold's rows (api.cuh:984-993).
inline void scatter_predictions(const double* fold, const std::vector<int>& rows, int n, int q,
                                int prefixes, double* out) {
  const size_t m = rows.size();
  for (int j = 0; j < prefixes; ++j)
    for (int c = 0; c < q; ++c)
      for (size_t r = 0; r < m; ++r)
        out[(size_t)j * n * q + (size_t)c * n + rows[r]] = fold[(size_t)j * m * q + (size_t)c * m + r];
}

}  // namespace fpls
