// Standalone small operators of the fitted-model path on caller data:
//   * LDA from score moments (lda.hpp:226-346, train_lda_prefixes_from_moments
//     + finalize_lda_prefixes): the pooled covariance, the ridge ladder
//     {1e-8 .. 1e-2} x trace / k with Cholesky solves, linear weights and
//     constants; the reference's resident CUDA LDA does the same from device
//     moments (src/cuda_resident_lda.cuh:41-121);
//   * PLS-SVD prefix weights (plssvd.hpp:138-164): per prefix k, the
//     symmetric solve of the leading k x k score Gram against diag(sigma)
//     (Cholesky, pivoted elimination when it fails; linalg.hpp:104-177) and
//     W_k = L_k Q_kᵀ.
// Both run the kernels the resident fits use (k_lda_prepare,
// k_plssvd_weights) on moments the caller gives, so the rules that only
// rounding reaches inside a fit (a later ridge, the pivoted fallback) can be
// driven exactly.
#include "../../include/fastpls_hip.h"
#include "host_util.hpp"
#include "kernels.hpp"

#include <cmath>
#include <stdexcept>
#include <vector>

namespace fpls {
namespace {

class Scoped {
 public:
  Scoped() : s_(stream_take()) {}
  ~Scoped() { stream_give(s_); }
  hipStream_t get() const { return s_; }

 private:
  hipStream_t s_;
};

void require_finite(const double* v, size_t count, const char* what) {
  for (size_t i = 0; i < count; ++i)
    if (!std::isfinite(v[i])) throw std::invalid_argument(what);
}

void lda_moments(const double* gram, int a, const double* class_sums, const int* counts, int K,
                 long long n, int prefix, double* linear, double* constants, double* ridge) {
  if (!gram || !class_sums || !counts || !linear || !constants || !ridge || a < 1 || K < 2 ||
      n < 1 || prefix < 1 || prefix > a)
    throw std::invalid_argument("fastPLS LDA moment dimensions are invalid");
  if (prefix > 64) throw std::invalid_argument("LDA prefixes above 64 components are not supported");
  long long total = 0;
  for (int c = 0; c < K; ++c) {
    if (counts[c] <= 0) throw std::invalid_argument("fastPLS LDA received an empty class");
    total += counts[c];
  }
  if (total != n) throw std::invalid_argument("fastPLS LDA class counts do not sum to n");
  require_finite(gram, size_t(a) * a, "fastPLS LDA moments must be finite");
  require_finite(class_sums, size_t(K) * a, "fastPLS LDA moments must be finite");
  Scoped s;
  Buf g, sums, cnt, lin, con, rdg, scratch, flags;
  for (Buf* b : {&g, &sums, &cnt, &lin, &con, &rdg, &scratch, &flags}) b->bind(s.get());
  g.alloc(sizeof(double) * a * a);
  sums.alloc(sizeof(double) * K * a);
  cnt.alloc(sizeof(int) * K);
  lin.alloc(sizeof(double) * K * prefix);
  con.alloc(sizeof(double) * K);
  rdg.alloc(sizeof(double));
  scratch.alloc(sizeof(double) * K * prefix);
  flags.alloc(sizeof(int));
  FPLS_HIP(hipMemcpyAsync(g.d(), gram, sizeof(double) * a * a, hipMemcpyHostToDevice, s.get()));
  FPLS_HIP(hipMemcpyAsync(sums.d(), class_sums, sizeof(double) * K * a, hipMemcpyHostToDevice,
                          s.get()));
  FPLS_HIP(hipMemcpyAsync(cnt.as<int>(), counts, sizeof(int) * K, hipMemcpyHostToDevice, s.get()));
  launch_fill_int(flags.as<int>(), 1, 0, s.get());
  launch_lda_prepare(g.d(), sums.d(), cnt.as<int>(), n, a, K, prefix, lin.d(), con.d(), rdg.d(),
                     scratch.d(), flags.as<int>(), s.get());
  int bad = 0;
  FPLS_HIP(hipMemcpyAsync(&bad, flags.as<int>(), sizeof(int), hipMemcpyDeviceToHost, s.get()));
  FPLS_HIP(hipMemcpyAsync(linear, lin.d(), sizeof(double) * K * prefix, hipMemcpyDeviceToHost,
                          s.get()));
  FPLS_HIP(hipMemcpyAsync(constants, con.d(), sizeof(double) * K, hipMemcpyDeviceToHost, s.get()));
  FPLS_HIP(hipMemcpyAsync(ridge, rdg.d(), sizeof(double), hipMemcpyDeviceToHost, s.get()));
  FPLS_HIP(hipStreamSynchronize(s.get()));
  if (bad)
    throw std::runtime_error(
        "fastPLS LDA Cholesky factorization failed for every regularization level");
}

void plssvd_weights(const double* gram, const double* sigma, const double* Q, int A, int q,
                    double* W) {
  if (!gram || !sigma || !Q || !W || A < 1 || q < 1)
    throw std::invalid_argument("fastPLS PLS-SVD weight dimensions are invalid");
  if (A > 64) throw std::invalid_argument("PLS-SVD components above 64 are not supported");
  require_finite(gram, size_t(A) * A, "fastPLS PLS-SVD score Gram must be finite");
  Scoped s;
  Buf g, sg, qq, w, flags;
  for (Buf* b : {&g, &sg, &qq, &w, &flags}) b->bind(s.get());
  g.alloc(sizeof(double) * A * A);
  sg.alloc(sizeof(double) * A);
  qq.alloc(sizeof(double) * q * A);
  w.alloc(sizeof(double) * A * A * q);
  flags.alloc(sizeof(int));
  FPLS_HIP(hipMemcpyAsync(g.d(), gram, sizeof(double) * A * A, hipMemcpyHostToDevice, s.get()));
  FPLS_HIP(hipMemcpyAsync(sg.d(), sigma, sizeof(double) * A, hipMemcpyHostToDevice, s.get()));
  FPLS_HIP(hipMemcpyAsync(qq.d(), Q, sizeof(double) * q * A, hipMemcpyHostToDevice, s.get()));
  launch_fill_int(flags.as<int>(), 1, 0, s.get());
  launch_plssvd_weights(g.d(), sg.d(), qq.d(), A, q, w.d(), flags.as<int>(), s.get());
  int bad = 0;
  FPLS_HIP(hipMemcpyAsync(&bad, flags.as<int>(), sizeof(int), hipMemcpyDeviceToHost, s.get()));
  FPLS_HIP(hipMemcpyAsync(W, w.d(), sizeof(double) * A * A * q, hipMemcpyDeviceToHost, s.get()));
  FPLS_HIP(hipStreamSynchronize(s.get()));
  if (bad) throw std::runtime_error("fastPLS PLS-SVD latent solve failed");
}

}  // namespace
}  // namespace fpls

extern "C" {

int fastpls_hip_lda_moments(const double* gram, int a, const double* class_sums,
                            const int* counts, int classes, long long n, int prefix,
                            double* linear, double* constants, double* ridge, char* error,
                            size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    fpls::lda_moments(gram, a, class_sums, counts, classes, n, prefix, linear, constants, ridge);
  });
}

int fastpls_hip_plssvd_weights(const double* gram, const double* sigma, const double* Q,
                               int components, int q, double* weights, char* error,
                               size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    fpls::plssvd_weights(gram, sigma, Q, components, q, weights);
  });
}

}  // extern "C"
