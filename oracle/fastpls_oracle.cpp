// fastpls oracle: TEST INFRASTRUCTURE ONLY.
//
// An FP64 CPU restatement of the fastPLS reference core, written for this
// project, used solely as the checker for the HIP library (tests, the smoke
// check and the CPU-baseline leg of the benchmark). Nothing in the product
// path links or calls this file.
//
// Each function cites the reference file:line whose semantics it restates.
// Pinning (how each function is checked against something other than itself)
// is listed in the test files: golden vectors produced by compiling the
// reference's own header core (oracle/_ref), the reference's known-answer
// tests (core/tests/*.cpp), and closed-form / brute-force checks.
//
// Storage convention: column-major doubles, leading dimension = rows.
// Labels at this boundary are one-based, as in src/cuda_resident_api.h:10.

#include <algorithm>
#include <dlfcn.h>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using std::size_t;

// ---------------------------------------------------------------- matrices
struct Mat {
  size_t r = 0, c = 0;
  std::vector<double> d;
  Mat() = default;
  Mat(size_t rows, size_t cols) : r(rows), c(cols), d(rows * cols, 0.0) {}
  void resize(size_t rows, size_t cols) { r = rows; c = cols; d.assign(rows * cols, 0.0); }
  double& operator()(size_t i, size_t j) { return d[i + j * r]; }
  double operator()(size_t i, size_t j) const { return d[i + j * r]; }
  double* data() { return d.data(); }
  const double* data() const { return d.data(); }
};

// Optional BLAS dgemm (Fortran interface with 64-bit integers), bound at run
// time by orc_use_blas for CPU-baseline timing only: the reference's CPU
// adapter runs every product through BLAS dgemm (src/core_cpu_backend.cpp:441-619).
// Unset (the default, and always in the pinning tests) the plain loops below run.
typedef void (*DgemmFn)(const char*, const char*, const int64_t*, const int64_t*,
                        const int64_t*, const double*, const double*, const int64_t*,
                        const double*, const int64_t*, const double*, double*,
                        const int64_t*, size_t, size_t);
DgemmFn g_dgemm = nullptr;

// C(m x n) = op(A) * op(B); op(A) is m x k. Plain loops in contiguous order.
void gemm(const double* A, size_t lda, bool ta, const double* B, size_t ldb,
          bool tb, size_t m, size_t n, size_t k, double* C, size_t ldc) {
  if (g_dgemm && m && n) {
    const int64_t M = int64_t(m), N = int64_t(n), K = int64_t(k);
    const int64_t LA = int64_t(std::max<size_t>(lda, 1)), LB = int64_t(std::max<size_t>(ldb, 1));
    const int64_t LC = int64_t(std::max<size_t>(ldc, 1));
    const double one = 1.0, zero = 0.0;
    g_dgemm(ta ? "T" : "N", tb ? "T" : "N", &M, &N, &K, &one, A, &LA, B, &LB, &zero, C, &LC,
            1, 1);
    return;
  }
  for (size_t j = 0; j < n; ++j) {
    double* cj = C + j * ldc;
    if (ta) {
      // C[i,j] = sum_l A[l,i] * op(B)[l,j]
      for (size_t i = 0; i < m; ++i) {
        const double* ai = A + i * lda;
        double acc = 0.0;
        if (!tb) {
          const double* bj = B + j * ldb;
          for (size_t l = 0; l < k; ++l) acc += ai[l] * bj[l];
        } else {
          for (size_t l = 0; l < k; ++l) acc += ai[l] * B[j + l * ldb];
        }
        cj[i] = acc;
      }
    } else {
      for (size_t i = 0; i < m; ++i) cj[i] = 0.0;
      for (size_t l = 0; l < k; ++l) {
        const double b = tb ? B[j + l * ldb] : B[l + j * ldb];
        if (b == 0.0) continue;
        const double* al = A + l * lda;
        for (size_t i = 0; i < m; ++i) cj[i] += al[i] * b;
      }
    }
  }
}

Mat mul(const Mat& A, bool ta, const Mat& B, bool tb) {
  const size_t m = ta ? A.c : A.r, k = ta ? A.r : A.c;
  const size_t kb = tb ? B.c : B.r, n = tb ? B.r : B.c;
  if (k != kb) throw std::invalid_argument("oracle gemm: inner dimensions differ");
  Mat C(m, n);
  if (m && n) gemm(A.data(), A.r, ta, B.data(), B.r, tb, m, n, k, C.data(), m);
  return C;
}

// Householder economy QR returning Q (rows x min(rows, cols)); the same
// factorisation LAPACK dgeqrf + dorgqr performs (src/core_cpu_lapack.cpp:49-96).
bool qr_economy(const Mat& A, Mat& Q) {
  const size_t m = A.r, n = A.c, k = std::min(m, n);
  Mat W = A;
  std::vector<double> tau(k, 0.0);
  for (size_t j = 0; j < k; ++j) {
    double norm2 = 0.0;
    for (size_t i = j; i < m; ++i) norm2 += W(i, j) * W(i, j);
    const double alpha = W(j, j);
    double xnorm2 = norm2 - alpha * alpha;
    if (!std::isfinite(norm2)) return false;
    if (xnorm2 <= 0.0) { tau[j] = 0.0; continue; }
    const double norm = std::sqrt(norm2);
    const double beta = alpha >= 0 ? -norm : norm;
    tau[j] = (beta - alpha) / beta;
    const double scale = 1.0 / (alpha - beta);
    for (size_t i = j + 1; i < m; ++i) W(i, j) *= scale;
    W(j, j) = beta;
    for (size_t c = j + 1; c < n; ++c) {
      double s = W(j, c);
      for (size_t i = j + 1; i < m; ++i) s += W(i, j) * W(i, c);
      s *= tau[j];
      W(j, c) -= s;
      for (size_t i = j + 1; i < m; ++i) W(i, c) -= s * W(i, j);
    }
  }
  Q.resize(m, k);
  for (size_t j = 0; j < k; ++j) Q(j, j) = 1.0;
  for (size_t jj = k; jj-- > 0;) {
    if (tau[jj] == 0.0) continue;
    for (size_t c = jj; c < k; ++c) {
      double s = Q(jj, c);
      for (size_t i = jj + 1; i < m; ++i) s += W(i, jj) * Q(i, c);
      s *= tau[jj];
      Q(jj, c) -= s;
      for (size_t i = jj + 1; i < m; ++i) Q(i, c) -= s * W(i, jj);
    }
  }
  for (double v : Q.d) if (!std::isfinite(v)) return false;
  return true;
}

// Cyclic Jacobi symmetric eigensolver. On return the matrix holds the
// eigenvectors as columns and the values ascend, the contract of LAPACK
// dsyevd('V','L') used by the reference (src/core_cpu_lapack.cpp:98-135).
bool symmetric_eigen(Mat& A, std::vector<double>& values) {
  const size_t n = A.r;
  if (n != A.c) return false;
  for (double v : A.d) if (!std::isfinite(v)) return false;
  Mat V(n, n);
  for (size_t i = 0; i < n; ++i) V(i, i) = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (size_t j = 0; j < n; ++j) {
      diag += A(j, j) * A(j, j);
      for (size_t i = 0; i < j; ++i) off += A(i, j) * A(i, j);
    }
    if (off <= 1e-34 * diag || off == 0.0) break;
    for (size_t p = 0; p + 1 < n; ++p) {
      for (size_t q = p + 1; q < n; ++q) {
        const double apq = A(p, q);
        if (apq == 0.0) continue;
        const double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) /
          (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
        for (size_t i = 0; i < n; ++i) {
          const double aip = A(i, p), aiq = A(i, q);
          A(i, p) = cs * aip - sn * aiq;
          A(i, q) = sn * aip + cs * aiq;
        }
        for (size_t i = 0; i < n; ++i) {
          const double api = A(p, i), aqi = A(q, i);
          A(p, i) = cs * api - sn * aqi;
          A(q, i) = sn * api + cs * aqi;
        }
        for (size_t i = 0; i < n; ++i) {
          const double vip = V(i, p), viq = V(i, q);
          V(i, p) = cs * vip - sn * viq;
          V(i, q) = sn * vip + cs * viq;
        }
      }
    }
  }
  std::vector<size_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::vector<double> diag(n);
  for (size_t i = 0; i < n; ++i) diag[i] = A(i, i);
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return diag[a] < diag[b]; });
  values.resize(n);
  Mat sorted(n, n);
  for (size_t j = 0; j < n; ++j) {
    values[j] = diag[order[j]];
    for (size_t i = 0; i < n; ++i) sorted(i, j) = V(i, order[j]);
  }
  A = std::move(sorted);
  return true;
}

// One-sided Jacobi economy SVD (singular values descending), the contract of
// dgesvd/dgesdd as the reference calls them (src/core_cpu_lapack.cpp:137-205).
bool svd_economy(const Mat& A, bool left_only, Mat& U,
                 std::vector<double>& s, Mat& Vt) {
  const bool tall = A.r >= A.c;
  Mat W = tall ? A : Mat(A.c, A.r);
  if (!tall) for (size_t i = 0; i < A.r; ++i) for (size_t j = 0; j < A.c; ++j) W(j, i) = A(i, j);
  const size_t m = W.r, n = W.c;
  Mat V(n, n);
  for (size_t i = 0; i < n; ++i) V(i, i) = 1.0;
  for (int sweep = 0; sweep < 80; ++sweep) {
    bool rotated = false;
    for (size_t p = 0; p + 1 < n; ++p) {
      for (size_t q = p + 1; q < n; ++q) {
        double alpha = 0, beta = 0, gamma = 0;
        for (size_t i = 0; i < m; ++i) {
          alpha += W(i, p) * W(i, p); beta += W(i, q) * W(i, q);
          gamma += W(i, p) * W(i, q);
        }
        if (std::fabs(gamma) <= 1e-15 * std::sqrt(alpha * beta) || gamma == 0.0) continue;
        rotated = true;
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0 ? 1.0 : -1.0) /
          (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (size_t i = 0; i < m; ++i) {
          const double wp = W(i, p), wq = W(i, q);
          W(i, p) = cs * wp - sn * wq; W(i, q) = sn * wp + cs * wq;
        }
        for (size_t i = 0; i < n; ++i) {
          const double vp = V(i, p), vq = V(i, q);
          V(i, p) = cs * vp - sn * vq; V(i, q) = sn * vp + cs * vq;
        }
      }
    }
    if (!rotated) break;
  }
  std::vector<double> norms(n);
  for (size_t j = 0; j < n; ++j) {
    double acc = 0;
    for (size_t i = 0; i < m; ++i) acc += W(i, j) * W(i, j);
    norms[j] = std::sqrt(acc);
  }
  std::vector<size_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return norms[a] > norms[b]; });
  const size_t k = n;
  s.resize(k);
  Mat left(m, k), right(n, k);
  for (size_t jj = 0; jj < k; ++jj) {
    const size_t j = order[jj];
    s[jj] = norms[j];
    const double inv = norms[j] > 0 ? 1.0 / norms[j] : 0.0;
    for (size_t i = 0; i < m; ++i) left(i, jj) = W(i, j) * inv;
    for (size_t i = 0; i < n; ++i) right(i, jj) = V(i, j);
  }
  // For a wide input the roles of the factors swap.
  const Mat& u = tall ? left : right;
  const Mat& v = tall ? right : left;
  U = u;
  if (!left_only) {
    Vt.resize(v.c, v.r);
    for (size_t i = 0; i < v.r; ++i) for (size_t j = 0; j < v.c; ++j) Vt(j, i) = v(i, j);
  }
  for (double x : s) if (!std::isfinite(x)) return false;
  return true;
}

// ------------------------------------------------------------ sketches
// Gaussian sketch entries exactly as the reference draws them:
// std::mt19937(seed) + std::normal_distribution<double>(0, 1), filled in
// storage order (simpls.hpp:244-249, rsvd.hpp:205-210).
void gaussian(unsigned int seed, size_t count, double* out) {
  std::mt19937 generator(seed);
  std::normal_distribution<double> normal(0.0, 1.0);
  for (size_t i = 0; i < count; ++i) out[i] = normal(generator);
}

// ------------------------------------------------------------ rSVD
struct Triplets { Mat U; std::vector<double> s; Mat Vt; };

// rsvd.hpp:59-83
Triplets dense_triplets(const Mat& A, size_t retained, bool left_only) {
  Triplets full;
  if (!svd_economy(A, left_only, full.U, full.s, full.Vt))
    throw std::runtime_error("oracle rSVD dense decomposition failed");
  retained = std::min(retained, std::min(full.s.size(), full.U.c));
  Triplets out;
  out.U.resize(full.U.r, retained);
  for (size_t j = 0; j < retained; ++j) for (size_t i = 0; i < full.U.r; ++i) out.U(i, j) = full.U(i, j);
  out.s.assign(full.s.begin(), full.s.begin() + retained);
  if (!left_only) {
    out.Vt.resize(retained, full.Vt.c);
    for (size_t j = 0; j < full.Vt.c; ++j) for (size_t i = 0; i < retained; ++i) out.Vt(i, j) = full.Vt(i, j);
  }
  return out;
}

// rsvd.hpp:85-165
Triplets finalize_sample(const Mat& A, const Mat& sample, size_t retained, bool left_only) {
  Mat basis;
  if (!qr_economy(sample, basis)) throw std::runtime_error("oracle rSVD sketch QR failed");
  Mat projected = mul(basis, true, A, false);
  if (projected.r <= projected.c) {
    Mat gram = mul(projected, false, projected, true);
    std::vector<double> ev;
    if (symmetric_eigen(gram, ev)) {
      const double largest = ev.empty() ? 1.0 : std::max(ev.back(), 1.0);
      const double tol = std::numeric_limits<double>::epsilon() *
        double(std::max(projected.r, projected.c)) * largest;
      size_t usable = 0;
      for (size_t idx = ev.size(); idx > 0; --idx) {
        if (ev[idx - 1] <= tol || usable == retained) break;
        ++usable;
      }
      if (usable > 0) {
        Mat reduced_left(gram.r, usable);
        std::vector<double> s(usable);
        for (size_t c = 0; c < usable; ++c) {
          const size_t src = ev.size() - 1 - c;
          s[c] = std::sqrt(std::max(ev[src], 0.0));
          for (size_t i = 0; i < gram.r; ++i) reduced_left(i, c) = gram(i, src);
        }
        Triplets out;
        out.U = mul(basis, false, reduced_left, false);
        out.s = s;
        if (!left_only) {
          out.Vt = mul(reduced_left, true, projected, false);
          for (size_t i = 0; i < usable; ++i) {
            const double inv = 1.0 / s[i];
            for (size_t j = 0; j < out.Vt.c; ++j) out.Vt(i, j) *= inv;
          }
        }
        return out;
      }
    }
  }
  Triplets reduced = dense_triplets(projected, retained, left_only);
  Triplets out;
  out.U = mul(basis, false, reduced.U, false);
  out.s = reduced.s;
  out.Vt = reduced.Vt;
  return out;
}

// rsvd.hpp:185-237
Triplets randomized_svd(const Mat& A, int retained, int oversample, int power,
                        unsigned int seed, bool left_only) {
  const size_t maximum = std::min(A.r, A.c);
  const size_t target = std::min(maximum, size_t(std::max(retained, 1)));
  if (target == 0) return {};
  const size_t width = std::min(maximum, target + size_t(std::max(oversample, 0)));
  if (width >= maximum) return dense_triplets(A, target, left_only);
  Mat omega(A.c, width);
  gaussian(seed, omega.d.size(), omega.data());
  Mat sample = mul(A, false, omega, false);
  for (int it = 0; it < std::max(power, 0); ++it) {
    Mat fb;
    if (!qr_economy(sample, fb)) throw std::runtime_error("oracle rSVD forward QR failed");
    Mat rev = mul(A, true, fb, false);
    Mat rb;
    if (!qr_economy(rev, rb)) throw std::runtime_error("oracle rSVD reverse QR failed");
    sample = mul(A, false, rb, false);
  }
  return finalize_sample(A, sample, target, left_only);
}

// ------------------------------------------------------- operator rSVD
// operator_rsvd.hpp:64-168 (finalize_operator_sample): eigen route when the
// projected block is at least 4 rows and not taller than wide, with the
// squared rank tolerance and no unit floor, accepted only when it yields
// all `retained` directions; otherwise the dense SVD of the block.
Triplets finalize_operator(const Mat& basis, const Mat& projected, size_t retained,
                           bool left_only) {
  Mat small_left;
  std::vector<double> sv;
  Mat small_vt;
  bool completed = false;
  if (projected.r >= 4 && projected.r <= projected.c) {
    Mat gram = mul(projected, false, projected, true);
    std::vector<double> ev;
    if (symmetric_eigen(gram, ev)) {
      const double largest = ev.empty() ? 0.0 : ev.back();
      const double rel = std::numeric_limits<double>::epsilon() *
        double(std::max(projected.r, projected.c));
      const double tol = largest > 0.0 ? rel * rel * largest : 0.0;
      size_t usable = 0;
      for (size_t idx = ev.size(); idx > 0; --idx) {
        if (ev[idx - 1] <= tol || usable == retained) break;
        ++usable;
      }
      if (usable == retained) {
        small_left.resize(projected.r, usable);
        sv.resize(usable);
        for (size_t c = 0; c < usable; ++c) {
          const size_t src = ev.size() - 1 - c;
          sv[c] = std::sqrt(std::max(ev[src], 0.0));
          for (size_t i = 0; i < projected.r; ++i) small_left(i, c) = gram(i, src);
        }
        if (!left_only) {
          small_vt = mul(small_left, true, projected, false);
          for (size_t i = 0; i < usable; ++i) {
            const double inv = 1.0 / sv[i];
            for (size_t j = 0; j < small_vt.c; ++j) small_vt(i, j) *= inv;
          }
        }
        completed = true;
      }
    }
  }
  if (!completed && !svd_economy(projected, left_only, small_left, sv, small_vt))
    throw std::runtime_error("oracle operator rSVD reduced decomposition failed");
  retained = std::min({retained, small_left.c, sv.size()});
  if (retained == 0) throw std::runtime_error("oracle operator rSVD returned no usable directions");
  Mat kept(small_left.r, retained);
  for (size_t j = 0; j < retained; ++j) for (size_t i = 0; i < small_left.r; ++i) kept(i, j) = small_left(i, j);
  Triplets out;
  out.U = mul(basis, false, kept, false);
  out.s.assign(sv.begin(), sv.begin() + retained);
  if (!left_only) {
    out.Vt.resize(retained, projected.c);
    for (size_t j = 0; j < projected.c; ++j) for (size_t i = 0; i < retained; ++i) out.Vt(i, j) = small_vt(i, j);
  }
  return out;
}

// A linear operator (rows x cols) known only through its products A·Z and
// Aᵀ·B (the operator interface of operators.hpp).
struct LinOp {
  size_t rows = 0, cols = 0;
  std::function<Mat(const Mat&)> fwd, adj;
};

LinOp explicit_op(const Mat& A) {
  LinOp op;
  op.rows = A.r; op.cols = A.c;
  op.fwd = [&A](const Mat& Z) { return mul(A, false, Z, false); };
  op.adj = [&A](const Mat& B) { return mul(A, true, B, false); };
  return op;
}

// operator_rsvd.hpp:204-272 (randomized_operator_svd) on any operator.
Triplets operator_svd_op(const LinOp& A, int retained, int oversample, int power,
                         unsigned int seed, bool left_only) {
  const size_t maximum = std::min(A.rows, A.cols);
  const size_t target = std::min(maximum, size_t(std::max(retained, 1)));
  if (target == 0) return {};
  const size_t width = std::min(maximum, target + size_t(std::max(oversample, 0)));
  Mat omega(A.cols, width);
  gaussian(seed, omega.d.size(), omega.data());
  Mat sample = A.fwd(omega);
  Mat fb;
  for (int it = 0; it < std::max(power, 0); ++it) {
    if (!qr_economy(sample, fb)) throw std::runtime_error("oracle operator rSVD forward QR failed");
    Mat rev = A.adj(fb);
    Mat rb;
    if (!qr_economy(rev, rb)) throw std::runtime_error("oracle operator rSVD reverse QR failed");
    sample = A.fwd(rb);
  }
  if (!qr_economy(sample, fb)) throw std::runtime_error("oracle operator rSVD sketch QR failed");
  Mat rev = A.adj(fb);                         // cols x width
  Mat projected(rev.c, rev.r);                 // width x cols
  for (size_t j = 0; j < rev.c; ++j) for (size_t i = 0; i < rev.r; ++i) projected(j, i) = rev(i, j);
  return finalize_operator(fb, projected, target, left_only);
}

// The explicit operator A (operators.hpp:16-56).
Triplets operator_svd(const Mat& A, int retained, int oversample, int power, unsigned int seed,
                      bool left_only) {
  return operator_svd_op(explicit_op(A), retained, oversample, power, seed, left_only);
}

// ----------------------------------------------------- audited rSVD
struct Audit {
  bool certified = false, fallback = false;
  int attempts = 0, oversample = 0, power = 0;
  unsigned int seed = 0;
  double subspace = 0.0, singular = 0.0, residual = 0.0, omitted = 0.0;
};

bool finite_triplets(const Triplets& t) {
  if (t.U.c == 0 || t.s.empty() || t.Vt.r == 0) return false;
  for (double v : t.U.d) if (!std::isfinite(v)) return false;
  for (double v : t.s) if (!std::isfinite(v)) return false;
  for (double v : t.Vt.d) if (!std::isfinite(v)) return false;
  return true;
}

// audited_rsvd.hpp:97-148
double triplet_residual(const Mat& A, const Triplets& t) {
  const size_t rank = std::min({t.U.c, t.Vt.r, t.s.size()});
  if (rank == 0 || t.Vt.c != A.c || t.U.r != A.r) return std::numeric_limits<double>::infinity();
  Mat right(A.c, rank);
  for (size_t c = 0; c < rank; ++c) for (size_t i = 0; i < A.c; ++i) right(i, c) = t.Vt(c, i);
  Mat forward = mul(A, false, right, false);
  Mat reverse = mul(A, true, t.U, false);
  const long double scale = std::max(std::fabs((long double)t.s.front()),
                                     (long double)std::numeric_limits<double>::epsilon());
  long double worst = 0.0L;
  for (size_t c = 0; c < rank; ++c) {
    long double f2 = 0.0L, r2 = 0.0L;
    const long double sv = t.s[c];
    for (size_t i = 0; i < A.r; ++i) {
      const long double d = (long double)forward(i, c) - sv * (long double)t.U(i, c);
      f2 += d * d;
    }
    for (size_t i = 0; i < A.c; ++i) {
      const long double d = (long double)reverse(i, c) - sv * (long double)right(i, c);
      r2 += d * d;
    }
    worst = std::max(worst, std::max(std::sqrt(f2), std::sqrt(r2)) / scale);
  }
  return double(worst);
}

double col_norm(const Mat& v) {
  long double s = 0.0L;
  for (size_t i = 0; i < v.r; ++i) s += (long double)v(i, 0) * (long double)v(i, 0);
  return double(std::sqrt(s));
}

bool normalize_col(Mat& v) {
  const double nrm = col_norm(v);
  if (!std::isfinite(nrm) || nrm <= std::numeric_limits<double>::epsilon()) return false;
  for (size_t i = 0; i < v.r; ++i) v(i, 0) /= nrm;
  return true;
}

void project_away(Mat& v, const Mat& basis, size_t cols) {
  cols = std::min(cols, basis.c);
  for (size_t c = 0; c < cols; ++c) {
    long double coef = 0.0L;
    for (size_t i = 0; i < v.r; ++i) coef += (long double)basis(i, c) * (long double)v(i, 0);
    const double cv = double(coef);
    for (size_t i = 0; i < v.r; ++i) v(i, 0) -= cv * basis(i, c);
  }
}

// audited_rsvd.hpp:150-183
double omitted_ratio(const Mat& A, const Triplets& t, size_t retained, unsigned int seed) {
  retained = std::min(retained, t.U.c);
  if (retained == 0 || t.s.size() < retained) return std::numeric_limits<double>::infinity();
  const double boundary = std::max(std::fabs(t.s[retained - 1]),
                                   std::numeric_limits<double>::epsilon());
  std::mt19937 generator(seed);
  std::normal_distribution<double> normal(0.0, 1.0);
  double largest = 0.0;
  for (int probe = 0; probe < 3; ++probe) {
    Mat right(A.c, 1);
    for (size_t i = 0; i < right.r; ++i) right(i, 0) = normal(generator);
    if (!normalize_col(right)) continue;
    Mat left;
    for (int it = 0; it < 2; ++it) {
      left = mul(A, false, right, false);
      project_away(left, t.U, retained);
      if (!normalize_col(left)) break;
      right = mul(A, true, left, false);
      if (!normalize_col(right)) break;
    }
    left = mul(A, false, right, false);
    project_away(left, t.U, retained);
    largest = std::max(largest, col_norm(left));
  }
  return largest / boundary;
}

// audited_rsvd.hpp:185-205
bool audit_triplets(const Mat& A, const Triplets& t, size_t retained, unsigned int seed,
                    double& residual, double& omitted) {
  if (!finite_triplets(t) || t.U.c < retained || t.s.size() < retained || t.Vt.r < retained)
    return false;
  residual = triplet_residual(A, t);
  omitted = omitted_ratio(A, t, retained, seed + 32452843U);
  if (t.s.size() > retained) {
    const double boundary = std::max(std::fabs(t.s[retained - 1]),
                                     std::numeric_limits<double>::epsilon());
    omitted = std::max(omitted, std::fabs(t.s[retained] / boundary));
  }
  return std::isfinite(residual) && residual <= 1e-2 && std::isfinite(omitted) && omitted <= 1.01;
}

// audited_rsvd.hpp:207-252
bool consensus(const Mat& A, const Triplets& lhs, const Triplets& rhs, double& subspace,
               double& singular, double& residual) {
  if (!finite_triplets(lhs) || !finite_triplets(rhs) || lhs.U.c != rhs.U.c ||
      lhs.s.size() != rhs.s.size())
    return false;
  const size_t rank = lhs.U.c;
  if (rank == 0) return false;
  long double overlap = 0.0L;
  for (size_t a = 0; a < rank; ++a)
    for (size_t b = 0; b < rank; ++b) {
      long double inner = 0.0L;
      for (size_t i = 0; i < lhs.U.r; ++i) inner += (long double)lhs.U(i, a) * (long double)rhs.U(i, b);
      overlap += inner * inner;
    }
  overlap = std::min((long double)rank, overlap);
  subspace = double(std::sqrt(std::max(0.0L, ((long double)rank - overlap) / (long double)rank)));
  const long double scale = std::max(std::fabs((long double)rhs.s.front()),
                                     (long double)std::numeric_limits<double>::epsilon());
  long double diff = 0.0L;
  for (size_t i = 0; i < lhs.s.size(); ++i)
    diff = std::max(diff, std::fabs((long double)lhs.s[i] - (long double)rhs.s[i]));
  singular = double(diff / scale);
  residual = triplet_residual(A, rhs);
  return subspace <= 1e-3 && singular <= 1e-5 && residual <= 1e-6;
}

// audited_rsvd.hpp:254-278
void trim_triplets(Triplets& t, size_t retained, bool left_only) {
  retained = std::min({retained, t.U.c, t.s.size()});
  Mat left(t.U.r, retained);
  for (size_t j = 0; j < retained; ++j) for (size_t i = 0; i < t.U.r; ++i) left(i, j) = t.U(i, j);
  t.U = left;
  t.s.resize(retained);
  if (left_only) { t.Vt = Mat(); return; }
  Mat right(retained, t.Vt.c);
  for (size_t j = 0; j < t.Vt.c; ++j) for (size_t i = 0; i < retained; ++i) right(i, j) = t.Vt(i, j);
  t.Vt = right;
}

// audited_rsvd.hpp:298-414
Triplets audited_rsvd(const Mat& A, int retained_components, int oversample, int power,
                      unsigned int seed, bool left_only, Audit& audit) {
  const size_t maximum = std::min(A.r, A.c);
  const size_t retained = std::min(maximum, size_t(std::max(retained_components, 1)));
  if (retained == 0) throw std::invalid_argument("oracle rSVD requires a nonempty matrix");
  if (maximum < 6) {
    Triplets out;
    if (!svd_economy(A, false, out.U, out.s, out.Vt)) throw std::runtime_error("oracle dense SVD recovery failed");
    trim_triplets(out, retained, false);
    audit = Audit();
    audit.certified = true;
    audit.fallback = true;
    audit.oversample = std::max(oversample, 0);
    audit.power = std::max(power, 0);
    audit.seed = seed;
    return out;
  }
  const int audit_rank = int(std::min(maximum, retained + 1));
  int cur_os = oversample, cur_pw = power;
  unsigned int cur_seed = seed;
  Triplets previous, accepted;
  double residual = std::numeric_limits<double>::infinity();
  double omitted = std::numeric_limits<double>::infinity();
  double subspace = 0.0, singular = 0.0;
  bool certified = false;
  int attempts = 0;
  const int oversamples[] = {std::max(oversample, 0), std::max(oversample, 32), std::max(oversample, 48)};
  const int powers[] = {std::max(power, 0), std::max(power, 3), std::max(power, 4)};
  const unsigned int seeds[] = {seed, seed + 104729U, seed + 209759U};
  for (int attempt = 0; attempt < 3; ++attempt) {
    cur_os = oversamples[attempt];
    cur_pw = powers[attempt];
    cur_seed = seeds[attempt];
    accepted = operator_svd(A, audit_rank, cur_os, cur_pw, cur_seed, false);
    ++attempts;
    const bool checked = audit_triplets(A, accepted, retained, cur_seed, residual, omitted);
    bool agreed = false;
    if (attempt > 0) agreed = consensus(A, previous, accepted, subspace, singular, residual);
    certified = checked || agreed;
    if (certified) break;
    previous = accepted;
  }
  for (int recovery = 0; !certified && recovery < 4; ++recovery) {
    const int width = std::min<int>(int(maximum), audit_rank + std::max(cur_os, 0));
    const int remaining = int(maximum) - width;
    const int enlarged = std::min<int>(int(maximum), std::max(width + std::min(32, remaining),
                                                              width + std::min(width, remaining)));
    cur_os = enlarged - audit_rank;
    cur_pw = std::max(cur_pw, 6) + 2 * recovery;
    cur_seed += 104729U * unsigned(recovery + 1);
    accepted = operator_svd(A, audit_rank, cur_os, cur_pw, cur_seed, false);
    ++attempts;
    certified = audit_triplets(A, accepted, retained, cur_seed, residual, omitted);
  }
  if (!certified)
    throw std::runtime_error("rSVD recovery did not meet numerical tolerances within its "
                             "workspace/iteration budget; no unchecked result was returned");
  trim_triplets(accepted, retained, left_only);
  audit.certified = true;
  audit.fallback = false;
  audit.attempts = attempts;
  audit.oversample = cur_os;
  audit.power = cur_pw;
  audit.seed = cur_seed;
  audit.subspace = subspace;
  audit.singular = singular;
  audit.residual = residual;
  audit.omitted = omitted;
  return accepted;
}

// ------------------------------------------------------------ fastcor
// src/r_api.cpp:4054-4080 (normalized_units): per unit (a row when by_row,
// else a column), centre by the mean over its features and scale to unit
// norm, in long double; a unit with a non-finite value or zero norm is NaN.
// Returns units x features.
Mat normalized_units(const Mat& X, bool by_row) {
  const size_t units = by_row ? X.r : X.c, features = by_row ? X.c : X.r;
  Mat out(units, features);
  for (size_t u = 0; u < units; ++u) {
    long double mean = 0.0L;
    bool finite = true;
    for (size_t f = 0; f < features; ++f) {
      const double v = by_row ? X(u, f) : X(f, u);
      if (!std::isfinite(v)) finite = false;
      mean += v;
    }
    mean /= (long double)features;
    long double ss = 0.0L;
    for (size_t f = 0; f < features; ++f) {
      const long double v = (by_row ? X(u, f) : X(f, u)) - mean;
      ss += v * v;
    }
    const long double nrm = std::sqrt(ss);
    for (size_t f = 0; f < features; ++f) {
      const long double v = (by_row ? X(u, f) : X(f, u)) - mean;
      out(u, f) = finite && nrm > 0.0L ? double(v / nrm) : std::numeric_limits<double>::quiet_NaN();
    }
  }
  return out;
}

// ------------------------------------------------------------ preprocessing
struct Prepared {
  Mat S;                      // p x q cross-covariance
  std::vector<double> mx, sx, my;
  Mat Xs;                     // standardized predictors
};

// Column statistics by sums of squares (supervised.hpp:63-90,
// classification.hpp:128-153).
void column_stats(const Mat& X, int scaling, std::vector<double>& mx,
                  std::vector<double>& sx) {
  const size_t n = X.r, p = X.c;
  mx.assign(p, 0.0); sx.assign(p, 1.0);
  for (size_t j = 0; j < p; ++j) {
    double sum = 0, ss = 0;
    for (size_t i = 0; i < n; ++i) { const double v = X(i, j); sum += v; ss += v * v; }
    if (scaling != 3) {
      mx[j] = sum / double(n);
      if (scaling == 2) {
        const double css = std::max(0.0, ss - double(n) * mx[j] * mx[j]);
        double sd = std::sqrt(css / double(std::max<size_t>(n - 1, 1)));
        if (!std::isfinite(sd) || sd <= 0) sd = 1.0;
        sx[j] = sd;
      }
    }
  }
}

// supervised.hpp:27-124
// crossprod = false skips S (the matrix-free route needs only the statistics
// and the standardised predictors).
Prepared prepare_dense(const Mat& X, const Mat& Y, int scaling, bool crossprod = true) {
  if (X.r != Y.r || X.r == 0 || X.c == 0 || Y.c == 0)
    throw std::invalid_argument("oracle dense cross-product dimensions are invalid");
  Prepared out;
  const size_t n = X.r, p = X.c, q = Y.c;
  out.my.assign(q, 0.0);
  for (size_t c = 0; c < q; ++c) {
    double sum = 0;
    for (size_t i = 0; i < n; ++i) sum += Y(i, c);
    out.my[c] = sum / double(n);
  }
  column_stats(X, scaling, out.mx, out.sx);
  out.Xs.resize(n, p);
  std::vector<double> colsum(p, 0.0);
  for (size_t j = 0; j < p; ++j) {
    double s = 0;
    for (size_t i = 0; i < n; ++i) {
      const double v = (X(i, j) - out.mx[j]) / out.sx[j];
      out.Xs(i, j) = v; s += v;
    }
    colsum[j] = s;
  }
  if (!crossprod) return out;
  out.S = mul(out.Xs, true, Y, false);
  for (size_t c = 0; c < q; ++c)
    for (size_t j = 0; j < p; ++j) out.S(j, c) -= colsum[j] * out.my[c];
  return out;
}

// classification.hpp:86-184 (labels are one-based here)
Prepared prepare_labels(const Mat& X, const int* labels1, size_t K, int scaling) {
  const size_t n = X.r, p = X.c;
  if (K < 2 || n == 0 || p == 0) throw std::invalid_argument("oracle label cross-product dimensions are invalid");
  Prepared out;
  std::vector<double> counts(K, 0.0);
  for (size_t i = 0; i < n; ++i) {
    const int c = labels1[i] - 1;
    if (c < 0 || size_t(c) >= K) throw std::invalid_argument("oracle label outside 1..K");
    counts[size_t(c)] += 1.0;
  }
  out.my.assign(K, 0.0);
  for (size_t c = 0; c < K; ++c) {
    if (counts[c] == 0) throw std::invalid_argument("oracle empty class");
    out.my[c] = counts[c] / double(n);
  }
  column_stats(X, scaling, out.mx, out.sx);
  out.Xs.resize(n, p);
  out.S.resize(p, K);
  std::vector<double> sums(K);
  for (size_t j = 0; j < p; ++j) {
    std::fill(sums.begin(), sums.end(), 0.0);
    double total = 0;
    for (size_t i = 0; i < n; ++i) {
      const double v = (X(i, j) - out.mx[j]) / out.sx[j];
      out.Xs(i, j) = v;
      sums[size_t(labels1[i] - 1)] += v; total += v;
    }
    for (size_t c = 0; c < K; ++c) out.S(j, c) = sums[c] - total * out.my[c];
  }
  return out;
}

// ------------------------------------------------------------ SIMPLS
struct SimplsControls {
  size_t components = 2, block = 1;
  bool cache = false, reorth = false, store_scores = true, right_gram = true,
       batch_geometry = false;
  int oversample = 32, power = 5;
  unsigned int seed = 1;
};

// simpls.hpp:43-62
size_t candidate_block_size(size_t remaining, size_t p, size_t q, bool cls,
                            size_t n, size_t maximum) {
  const double work = double(n) * double(p) * double(q);
  if (q <= 1 || remaining < 4 || work < 5.0e8) return 1;
  if (cls && q <= 2048) return std::max<size_t>(1, std::min({maximum, remaining, p, q}));
  if (!cls && double(p) * double(q) > 64.0 * 1024.0 * 1024.0)
    return std::max<size_t>(1, std::min({size_t(8), maximum, remaining, p, q}));
  return 1;
}

// src/r_api.cpp:1149-1190 with the default environment thresholds.
SimplsControls simpls_controls(size_t n, size_t p, size_t q, size_t A, bool cls,
                               int oversample, int power, unsigned int seed) {
  SimplsControls c;
  c.components = A;
  c.block = candidate_block_size(A, p, q, cls, n, 64);
  c.cache = A >= 20 && p <= n && n >= p * 8 && p <= 512;
  c.batch_geometry = c.block > 1 && !c.cache;
  c.reorth = false;
  c.store_scores = true;
  c.right_gram = true;
  c.oversample = oversample; c.power = power; c.seed = seed;
  return c;
}

// simpls.hpp:136-159 (classical Gram-Schmidt, one or two passes)
void remove_basis(const Mat& basis, size_t used, std::vector<double>& v, bool twice) {
  if (used == 0) return;
  const size_t p = basis.r;
  std::vector<double> proj(used);
  for (int pass = 0; pass < (twice ? 2 : 1); ++pass) {
    for (size_t j = 0; j < used; ++j) {
      double s = 0;
      for (size_t i = 0; i < p; ++i) s += basis(i, j) * v[i];
      proj[j] = s;
    }
    std::vector<double> corr(p, 0.0);
    for (size_t j = 0; j < used; ++j)
      for (size_t i = 0; i < p; ++i) corr[i] += basis(i, j) * proj[j];
    for (size_t i = 0; i < p; ++i) v[i] -= corr[i];
  }
}

// simpls.hpp:161-200
void remove_score_basis(const Mat& T, const Mat& R, size_t used,
                        std::vector<double>& t, std::vector<double>& r) {
  if (used == 0) return;
  std::vector<double> proj(used);
  for (int pass = 0; pass < 2; ++pass) {
    for (size_t j = 0; j < used; ++j) {
      double s = 0;
      for (size_t i = 0; i < T.r; ++i) s += T(i, j) * t[i];
      proj[j] = s;
    }
    for (size_t i = 0; i < T.r; ++i) {
      double s = 0;
      for (size_t j = 0; j < used; ++j) s += T(i, j) * proj[j];
      t[i] -= s;
    }
    for (size_t i = 0; i < R.r; ++i) {
      double s = 0;
      for (size_t j = 0; j < used; ++j) s += R(i, j) * proj[j];
      r[i] -= s;
    }
  }
}

double norm2(const std::vector<double>& v) {
  double s = 0;
  for (double x : v) s += x * x;
  return std::sqrt(std::max(s, 0.0));
}

// simpls.hpp:226-368
bool refresh_directions(const Mat& S, const Mat& G, bool use_right_gram,
                        size_t retained, int oversample, int power,
                        unsigned int seed, Mat& directions) {
  const size_t maximum = std::min(S.r, S.c);
  const size_t width = std::min(maximum, retained + size_t(std::max(oversample, 0)));
  if (retained == 0 || width == 0) return false;
  Mat random(S.c, width);
  gaussian(seed, random.d.size(), random.data());
  Mat sample;
  if (use_right_gram) {
    Mat reverse_basis = random;
    for (int it = 0; it < std::max(power, 0); ++it) {
      Mat rs = mul(G, false, reverse_basis, false);
      if (!qr_economy(rs, reverse_basis)) return false;
    }
    if (power == 0) {
      Mat tmp;
      if (!qr_economy(reverse_basis, tmp)) return false;
      reverse_basis = tmp;
    }
    sample = mul(S, false, reverse_basis, false);
  } else {
    sample = mul(S, false, random, false);
    for (int it = 0; it < std::max(power, 0); ++it) {
      Mat db, rb;
      if (!qr_economy(sample, db)) return false;
      Mat rs = mul(S, true, db, false);
      if (!qr_economy(rs, rb)) return false;
      sample = mul(S, false, rb, false);
    }
  }
  Mat basis;
  if (!qr_economy(sample, basis)) return false;
  Mat reduced = mul(basis, true, S, false);
  Mat gram = mul(reduced, false, reduced, true);
  std::vector<double> ev;
  Mat left;
  if (!symmetric_eigen(gram, ev)) {
    std::vector<double> s; Mat unused;
    if (!svd_economy(reduced, true, left, s, unused)) return false;
  } else {
    const size_t available = std::min(retained, ev.size());
    left.resize(gram.r, available);
    for (size_t c = 0; c < available; ++c) {
      const size_t src = ev.size() - 1 - c;
      for (size_t i = 0; i < gram.r; ++i) left(i, c) = gram(i, src);
    }
  }
  const size_t available = std::min(retained, left.c);
  if (available == 0) return false;
  Mat l2(left.r, available);
  for (size_t j = 0; j < available; ++j) for (size_t i = 0; i < left.r; ++i) l2(i, j) = left(i, j);
  directions = mul(basis, false, l2, false);
  return true;
}

struct SimplsModel {
  Mat R, Q, V, T;
  size_t completed = 0;
};

// simpls.hpp:480-763 (explicit cross-covariance path)
SimplsModel fit_simpls(const Mat& X, const Mat& S0, const SimplsControls& ctl) {
  const size_t n = X.r, p = S0.r, q = S0.c;
  if (ctl.components == 0 || X.c != p) throw std::invalid_argument("oracle SIMPLS needs conformable inputs");
  const size_t maximum = std::min({ctl.components, p, std::max<size_t>(n - 1, 1)});
  SimplsModel model;
  model.R.resize(p, maximum); model.Q.resize(q, maximum); model.V.resize(p, maximum);
  const bool retain_scores = ctl.store_scores || ctl.reorth;
  if (retain_scores) model.T.resize(n, maximum);
  Mat S = S0;
  Mat XtX;
  if (ctl.cache) XtX = mul(X, true, X, false);
  const bool use_right_gram = ctl.right_gram && q <= p && q <= 512;
  Mat G;
  if (use_right_gram) G = mul(S, true, S, false);

  size_t component = 0;
  while (component < maximum) {
    const size_t block = std::min({std::max<size_t>(ctl.block, 1), maximum - component, std::min(p, q)});
    Mat candidates;
    if (!refresh_directions(S, G, use_right_gram, block, ctl.oversample, ctl.power,
                            ctl.seed + unsigned(component), candidates)) break;
    const size_t available = std::min(block, candidates.c);
    Mat cand_scores, cand_loadings;
    if (ctl.batch_geometry && !ctl.cache) {
      cand_scores = mul(X, false, candidates, false);
      cand_loadings = mul(X, true, cand_scores, false);
    }
    bool stopped = false;
    for (size_t cand = 0; cand < available && component < maximum; ++cand, ++component) {
      std::vector<double> r(p), t(n, 0.0), w(p);
      for (size_t i = 0; i < p; ++i) r[i] = candidates(i, cand);
      if (ctl.cache) {
        for (size_t i = 0; i < p; ++i) {
          double s = 0;
          for (size_t l = 0; l < p; ++l) s += XtX(i, l) * r[l];
          w[i] = s;
        }
        double ns = 0;
        for (size_t i = 0; i < p; ++i) ns += r[i] * w[i];
        if (!std::isfinite(ns) || ns <= 0) { stopped = true; break; }
        const double sn = std::sqrt(ns);
        for (double& x : r) x /= sn;
        for (double& x : w) x /= sn;
        if (retain_scores) gemm(X.data(), n, false, r.data(), p, false, n, 1, p, t.data(), n);
      } else {
        if (ctl.batch_geometry) {
          for (size_t i = 0; i < n; ++i) t[i] = cand_scores(i, cand);
        } else {
          gemm(X.data(), n, false, r.data(), p, false, n, 1, p, t.data(), n);
        }
        const bool reorthed = ctl.reorth && component > 0;
        if (reorthed) remove_score_basis(model.T, model.R, component, t, r);
        const double tn = norm2(t);
        if (!std::isfinite(tn) || tn <= 0) { stopped = true; break; }
        for (double& x : t) x /= tn;
        for (double& x : r) x /= tn;
        if (ctl.batch_geometry && !reorthed) {
          for (size_t i = 0; i < p; ++i) w[i] = cand_loadings(i, cand) / tn;
        } else {
          gemm(X.data(), n, true, t.data(), n, false, p, 1, n, w.data(), p);
        }
      }
      std::vector<double> qa(q);
      for (size_t c = 0; c < q; ++c) {
        double s = 0;
        for (size_t i = 0; i < p; ++i) s += S0(i, c) * r[i];
        qa[c] = s;
      }
      std::vector<double> v = w;
      remove_basis(model.V, component, v, ctl.reorth);
      const double vn = norm2(v);
      if (!std::isfinite(vn) || vn <= 0) { stopped = true; break; }
      for (double& x : v) x /= vn;
      std::vector<double> row(q);
      for (size_t c = 0; c < q; ++c) {
        double s = 0;
        for (size_t i = 0; i < p; ++i) s += v[i] * S(i, c);
        row[c] = s;
      }
      for (size_t c = 0; c < q; ++c)
        for (size_t i = 0; i < p; ++i) S(i, c) -= v[i] * row[c];
      if (use_right_gram) {
        for (size_t c = 0; c < q; ++c) for (size_t i = 0; i < q; ++i) G(i, c) -= row[i] * row[c];
        for (size_t c = 0; c < q; ++c) for (size_t i = 0; i < c; ++i) {
          const double val = 0.5 * (G(i, c) + G(c, i));
          G(i, c) = val; G(c, i) = val;
        }
      }
      for (size_t i = 0; i < p; ++i) { model.R(i, component) = r[i]; model.V(i, component) = v[i]; }
      for (size_t c = 0; c < q; ++c) model.Q(c, component) = qa[c];
      if (retain_scores) for (size_t i = 0; i < n; ++i) model.T(i, component) = t[i];
      model.completed = component + 1;
    }
    if (stopped) break;
  }
  return model;
}

// ------------------------------------------------------------ PLS-SVD
// linalg.hpp:57-177
bool cholesky_solve(const Mat& A, const Mat& B, Mat& X) {
  const size_t n = A.r;
  Mat L(n, n);
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j <= i; ++j) {
      double v = A(i, j);
      for (size_t k = 0; k < j; ++k) v -= L(i, k) * L(j, k);
      if (i == j) {
        if (!std::isfinite(v) || v <= 0) return false;
        L(i, j) = std::sqrt(v);
      } else {
        L(i, j) = v / L(j, j);
      }
    }
  }
  X.resize(n, B.c);
  for (size_t c = 0; c < B.c; ++c) {
    for (size_t i = 0; i < n; ++i) {
      double v = B(i, c);
      for (size_t k = 0; k < i; ++k) v -= L(i, k) * X(k, c);
      X(i, c) = v / L(i, i);
    }
    for (size_t ii = n; ii-- > 0;) {
      double v = X(ii, c);
      for (size_t k = ii + 1; k < n; ++k) v -= L(k, ii) * X(k, c);
      X(ii, c) = v / L(ii, ii);
      if (!std::isfinite(X(ii, c))) return false;
    }
  }
  return true;
}

bool pivoted_solve(const Mat& A, const Mat& B, Mat& X) {
  const size_t n = A.r;
  Mat F = A;
  X = B;
  double scale = 0;
  for (double v : F.d) scale = std::max(scale, std::fabs(v));
  const double tol = std::numeric_limits<double>::epsilon() * double(std::max<size_t>(n, 1)) * std::max(scale, 1.0);
  for (size_t piv = 0; piv < n; ++piv) {
    size_t sel = piv;
    for (size_t i = piv + 1; i < n; ++i) if (std::fabs(F(i, piv)) > std::fabs(F(sel, piv))) sel = i;
    if (!std::isfinite(F(sel, piv)) || std::fabs(F(sel, piv)) <= tol) return false;
    if (sel != piv) {
      for (size_t c = 0; c < n; ++c) std::swap(F(piv, c), F(sel, c));
      for (size_t c = 0; c < X.c; ++c) std::swap(X(piv, c), X(sel, c));
    }
    for (size_t i = piv + 1; i < n; ++i) {
      const double mlt = F(i, piv) / F(piv, piv);
      F(i, piv) = 0;
      for (size_t c = piv + 1; c < n; ++c) F(i, c) -= mlt * F(piv, c);
      for (size_t c = 0; c < X.c; ++c) X(i, c) -= mlt * X(piv, c);
    }
  }
  for (size_t c = 0; c < X.c; ++c) {
    for (size_t ii = n; ii-- > 0;) {
      double v = X(ii, c);
      for (size_t k = ii + 1; k < n; ++k) v -= F(ii, k) * X(k, c);
      X(ii, c) = v / F(ii, ii);
      if (!std::isfinite(X(ii, c))) return false;
    }
  }
  return true;
}

bool solve_symmetric(const Mat& A, const Mat& B, Mat& X) {
  return cholesky_solve(A, B, X) || pivoted_solve(A, B, X);
}

struct PlssvdModel {
  Mat R, Q, T;
  std::vector<double> s;
  std::vector<Mat> W;     // prediction weights per prefix (k x q)
  std::vector<Mat> latent;
};

PlssvdModel assemble_plssvd(const Mat& X, size_t p, size_t q, const std::vector<int>& comps,
                            size_t retained, const Triplets& dec);

// plssvd.hpp:57-166, 197-220
PlssvdModel fit_plssvd(const Mat& X, const Mat& S, const std::vector<int>& comps,
                       int oversample, int power, unsigned int seed) {
  const size_t p = S.r, q = S.c;
  const size_t bound = std::min(p, q);
  size_t retained = 0;
  for (int c : comps) {
    if (c < 1 || size_t(c) > bound) throw std::invalid_argument("oracle PLS-SVD component count exceeds cross-covariance rank");
    retained = std::max(retained, size_t(c));
  }
  return assemble_plssvd(X, p, q, comps, retained,
                         randomized_svd(S, int(retained), oversample, power, seed, false));
}

// plssvd.hpp:166-220 (assemble_plssvd_model): R, Q, T = X R and the prefix
// weights from the decomposition of the cross-covariance.
PlssvdModel assemble_plssvd(const Mat& X, size_t p, size_t q, const std::vector<int>& comps,
                            size_t retained, const Triplets& dec) {
  const size_t done = std::min({dec.U.c, dec.Vt.r, dec.s.size()});
  if (done < retained) throw std::runtime_error("oracle PLS-SVD returned fewer components than requested");
  PlssvdModel m;
  m.R.resize(p, retained); m.Q.resize(q, retained);
  m.s.assign(dec.s.begin(), dec.s.begin() + retained);
  for (size_t k = 0; k < retained; ++k) {
    for (size_t i = 0; i < p; ++i) m.R(i, k) = dec.U(i, k);
    for (size_t c = 0; c < q; ++c) m.Q(c, k) = dec.Vt(k, c);
  }
  m.T = mul(X, false, m.R, false);
  Mat G = mul(m.T, true, m.T, false);
  for (int req : comps) {
    const size_t k = size_t(req);
    Mat g(k, k), diag(k, k);
    for (size_t j = 0; j < k; ++j) {
      diag(j, j) = m.s[j];
      for (size_t i = 0; i < k; ++i) g(i, j) = G(i, j);
    }
    Mat lat;
    if (!solve_symmetric(g, diag, lat)) throw std::runtime_error("oracle PLS-SVD latent solve failed");
    Mat qk(q, k);
    for (size_t j = 0; j < k; ++j) for (size_t c = 0; c < q; ++c) qk(c, j) = m.Q(c, j);
    m.W.push_back(mul(lat, false, qk, true));
    m.latent.push_back(lat);
  }
  return m;
}

// ------------------------------------------------- matrix-free path
// operators.hpp:58-156 (CenteredCrosscovOperator): S = Xsᵀ (Y - 1 μᵀ) applied
// without forming S or the centred responses.
LinOp centered_crosscov(const Mat& Xs, const Mat& Y, const std::vector<double>& mu) {
  LinOp op;
  op.rows = Xs.c; op.cols = Y.c;
  op.fwd = [&Xs, &Y, &mu](const Mat& Z) {
    Mat inter = mul(Y, false, Z, false);
    for (size_t c = 0; c < Z.c; ++c) {
      double offset = 0.0;
      for (size_t r = 0; r < Z.r; ++r) offset += mu[r] * Z(r, c);
      for (size_t i = 0; i < inter.r; ++i) inter(i, c) -= offset;
    }
    return mul(Xs, true, inter, false);
  };
  op.adj = [&Xs, &Y, &mu](const Mat& B) {
    Mat inter = mul(Xs, false, B, false);
    Mat out = mul(Y, true, inter, false);
    for (size_t c = 0; c < B.c; ++c) {
      double sum = 0.0;
      for (size_t i = 0; i < inter.r; ++i) sum += inter(i, c);
      for (size_t r = 0; r < out.r; ++r) out(r, c) -= mu[r] * sum;
    }
    return out;
  };
  return op;
}

// operators.hpp:158-301 (ProjectedOperator): (I - V Vᵀ) S with the first
// `used` columns of V; one classical Gram-Schmidt pass per product.
void project_out(const Mat& V, size_t used, Mat& M) {
  if (used == 0) return;
  for (size_t c = 0; c < M.c; ++c) {
    std::vector<double> coef(used, 0.0);
    for (size_t j = 0; j < used; ++j)
      for (size_t i = 0; i < M.r; ++i) coef[j] += V(i, j) * M(i, c);
    std::vector<double> corr(M.r, 0.0);
    for (size_t j = 0; j < used; ++j)
      for (size_t i = 0; i < M.r; ++i) corr[i] += V(i, j) * coef[j];
    for (size_t i = 0; i < M.r; ++i) M(i, c) -= corr[i];
  }
}

LinOp projected_op(const LinOp& base, const Mat& V, const size_t& used) {
  LinOp op;
  op.rows = base.rows; op.cols = base.cols;
  op.fwd = [&base, &V, &used](const Mat& Z) {
    Mat out = base.fwd(Z);
    project_out(V, used, out);
    return out;
  };
  op.adj = [&base, &V, &used](const Mat& B) {
    Mat pb = B;
    project_out(V, used, pb);
    return base.adj(pb);
  };
  return op;
}

// operator_rsvd.hpp:37-61 (normalize_column) with a long double norm.
bool normalize_unit(Mat& v) {
  long double s = 0.0L;
  for (size_t i = 0; i < v.r; ++i) s += (long double)v(i, 0) * (long double)v(i, 0);
  const double nrm = double(std::sqrt(s));
  if (!std::isfinite(nrm) || nrm <= std::numeric_limits<double>::epsilon()) return false;
  for (size_t i = 0; i < v.r; ++i) v(i, 0) /= nrm;
  return true;
}

// operator_rsvd.hpp:175-202 (randomized_dominant_operator_direction) on the
// projected operator; stabilising is two projection passes
// (operators.hpp:262-285).
bool dominant_direction(const LinOp& projected, const Mat& V, size_t used, int power,
                        unsigned int seed, Mat& direction) {
  direction.resize(projected.rows, 1);
  std::mt19937 generator(seed);
  std::normal_distribution<double> normal(0.0, 1.0);
  for (size_t i = 0; i < direction.r; ++i) direction(i, 0) = normal(generator);
  project_out(V, used, direction);
  project_out(V, used, direction);
  if (!normalize_unit(direction)) return false;
  for (int it = 0; it < std::max(power, 1); ++it) {
    Mat reverse = projected.adj(direction);
    if (!normalize_unit(reverse)) return false;
    Mat sample = projected.fwd(reverse);
    project_out(V, used, sample);
    project_out(V, used, sample);
    if (!normalize_unit(sample)) return false;
    direction = sample;
  }
  return true;
}

// simpls.hpp:768-1047 (fit_simpls_operator) with the centred cross-covariance
// operator and its projection; rank_one selects the dominant-direction route
// taken for a cross-covariance above 512 MiB (src/r_api.cpp:1480-1493).
SimplsModel fit_simpls_operator(const Mat& Xs, const Mat& Y, const std::vector<double>& mu,
                                const SimplsControls& ctl, bool rank_one) {
  const size_t n = Xs.r, p = Xs.c, q = Y.c;
  if (ctl.components == 0) throw std::invalid_argument("oracle implicit SIMPLS needs components");
  const size_t maximum = std::min({ctl.components, p, std::max<size_t>(n - 1, 1)});
  SimplsModel model;
  model.R.resize(p, maximum); model.Q.resize(q, maximum); model.V.resize(p, maximum);
  const bool retain_scores = ctl.store_scores || ctl.reorth;
  if (retain_scores) model.T.resize(n, maximum);
  Mat XtX;
  if (ctl.cache) XtX = mul(Xs, true, Xs, false);
  const LinOp initial = centered_crosscov(Xs, Y, mu);
  size_t component = 0;
  const LinOp projected = projected_op(initial, model.V, component);
  while (component < maximum) {
    const size_t block = std::min({std::max<size_t>(ctl.block, 1), maximum - component, std::min(p, q)});
    const unsigned int seed = ctl.seed + unsigned(component);
    Mat candidates;
    if (block == 1 && rank_one) {
      if (!dominant_direction(projected, model.V, component, ctl.power, seed, candidates)) break;
    } else {
      candidates = operator_svd_op(projected, int(block), ctl.oversample, ctl.power, seed, true).U;
    }
    const size_t available = std::min(block, candidates.c);
    if (available == 0) break;
    Mat cand_scores, cand_loadings;
    if (ctl.batch_geometry && !ctl.cache) {
      cand_scores = mul(Xs, false, candidates, false);
      cand_loadings = mul(Xs, true, cand_scores, false);
    }
    bool stopped = false;
    for (size_t cand = 0; cand < available && component < maximum; ++cand, ++component) {
      std::vector<double> r(p), t(n, 0.0), w(p);
      for (size_t i = 0; i < p; ++i) r[i] = candidates(i, cand);
      if (ctl.cache) {
        for (size_t i = 0; i < p; ++i) {
          double s2 = 0;
          for (size_t l = 0; l < p; ++l) s2 += XtX(i, l) * r[l];
          w[i] = s2;
        }
        double ns = 0;
        for (size_t i = 0; i < p; ++i) ns += r[i] * w[i];
        if (!std::isfinite(ns) || ns <= 0) { stopped = true; break; }
        const double sn = std::sqrt(ns);
        for (double& x : r) x /= sn;
        for (double& x : w) x /= sn;
        if (retain_scores)
          for (size_t l = 0; l < p; ++l) for (size_t i = 0; i < n; ++i) t[i] += Xs(i, l) * r[l];
      } else {
        if (ctl.batch_geometry) {
          for (size_t i = 0; i < n; ++i) t[i] = cand_scores(i, cand);
        } else {
          for (size_t l = 0; l < p; ++l) for (size_t i = 0; i < n; ++i) t[i] += Xs(i, l) * r[l];
        }
        const bool reorthed = ctl.reorth && component > 0;
        if (reorthed) remove_score_basis(model.T, model.R, component, t, r);
        const double tn = norm2(t);
        if (!std::isfinite(tn) || tn <= 0) { stopped = true; break; }
        for (double& x : t) x /= tn;
        for (double& x : r) x /= tn;
        if (ctl.batch_geometry && !reorthed) {
          for (size_t i = 0; i < p; ++i) w[i] = cand_loadings(i, cand) / tn;
        } else {
          for (size_t j = 0; j < p; ++j) {
            double s2 = 0;
            for (size_t i = 0; i < n; ++i) s2 += Xs(i, j) * t[i];
            w[j] = s2;
          }
        }
      }
      Mat rm(p, 1);
      for (size_t i = 0; i < p; ++i) rm(i, 0) = r[i];
      const Mat qa = initial.adj(rm);          // response loadings Sᵀ r
      std::vector<double> v = w;
      remove_basis(model.V, component, v, ctl.reorth);
      const double vn = norm2(v);
      if (!std::isfinite(vn) || vn <= 0) { stopped = true; break; }
      for (double& x : v) x /= vn;
      for (size_t i = 0; i < p; ++i) { model.R(i, component) = r[i]; model.V(i, component) = v[i]; }
      for (size_t c = 0; c < q; ++c) model.Q(c, component) = qa(c, 0);
      if (retain_scores) for (size_t i = 0; i < n; ++i) model.T(i, component) = t[i];
      model.completed = component + 1;
    }
    if (stopped) break;
  }
  return model;
}

// src/r_api.cpp:1466-1493: the implicit route's controls, with the rank-one
// override for a cross-covariance above 512 MiB of `elem` bytes per entry.
SimplsControls operator_controls(size_t n, size_t p, size_t q, size_t A, int oversample,
                                 int power, unsigned int seed, size_t elem, bool& rank_one) {
  SimplsControls c = simpls_controls(n, p, q, A, false, oversample, power, seed);
  rank_one = (long double)p * (long double)q * (long double)elem > 512.0L * 1024.0L * 1024.0L;
  if (rank_one) {
    c.block = 1;
    c.batch_geometry = false;
    c.reorth = true;
  }
  return c;
}

// R/main.R:3454-3463 (.should_use_xprod_default): the CPU path forms no
// cross-covariance once p·q·8 bytes exceed 32 MiB.
bool implicit_route(size_t p, size_t q) {
  return (double)p * (double)q * 8.0 / (1024.0 * 1024.0) > 32.0;
}

// plssvd.hpp:222-248 (fit_plssvd_operator).
PlssvdModel fit_plssvd_operator(const Mat& Xs, const Mat& Y, const std::vector<double>& mu,
                                const std::vector<int>& comps, int oversample, int power,
                                unsigned int seed) {
  const size_t p = Xs.c, q = Y.c;
  const size_t bound = std::min(p, q);
  size_t retained = 0;
  for (int c : comps) {
    if (c < 1 || size_t(c) > bound) throw std::invalid_argument("oracle implicit PLS-SVD component count exceeds cross-covariance rank");
    retained = std::max(retained, size_t(c));
  }
  const LinOp crosscov = centered_crosscov(Xs, Y, mu);
  return assemble_plssvd(Xs, p, q, comps, retained,
                         operator_svd_op(crosscov, int(retained), oversample, power, seed, false));
}

// ------------------------------------------------------------ OPLS
// The orthogonal filter of opls.hpp:123-221 (fit_preprocessed_filter_inplace)
// with the randomized leading direction of fit_opls_filter_rsvd
// (opls.hpp:314-363) and fit_opls_filter_labels_rsvd (opls.hpp:418-465): a
// rank-one rSVD of the current cross-covariance with seed + component, left
// vector only, normalised. The resident device engine takes this randomized
// route at both precisions (src/cuda_resident_special.cuh:270-301).
struct OplsFilter {
  Prepared pr;        // outer statistics; pr.Xs is filtered in place
  Mat W, P;           // orthogonal weights and loadings, p x completed
  size_t completed = 0;
};

// opls.hpp:34-61 (vector_norm / normalize): plain double sums, a zero or
// non-finite norm fails
bool opls_normalize(Mat& v) {
  double sum = 0.0;
  for (size_t i = 0; i < v.r; ++i) sum += v(i, 0) * v(i, 0);
  const double nrm = std::sqrt(std::max(sum, 0.0));
  if (!std::isfinite(nrm) || nrm <= 0.0) return false;
  for (size_t i = 0; i < v.r; ++i) v(i, 0) /= nrm;
  return true;
}

double opls_dot(const Mat& a, const Mat& b) {
  double sum = 0.0;
  for (size_t i = 0; i < a.r; ++i) sum += a(i, 0) * b(i, 0);
  return sum;
}

// classification.hpp:341-385 (centered_label_crossprod) with offsets -prior
Mat label_crossprod(const Mat& X, const int* labels1, const std::vector<double>& prior) {
  const size_t K = prior.size();
  Mat S(X.c, K);
  std::vector<double> sums(K);
  for (size_t j = 0; j < X.c; ++j) {
    std::fill(sums.begin(), sums.end(), 0.0);
    double total = 0.0;
    for (size_t i = 0; i < X.r; ++i) {
      const double v = X(i, j);
      total += v;
      sums[size_t(labels1[i] - 1)] += v;
    }
    for (size_t c = 0; c < K; ++c) S(j, c) = sums[c] + total * (-prior[c]);
  }
  return S;
}

OplsFilter opls_filter(const Mat& X, const Mat* Y, const int* labels1, size_t q, size_t north,
                       int scaling, int oversample, int power, unsigned int seed) {
  OplsFilter f;
  f.pr = labels1 ? prepare_labels(X, labels1, q, scaling) : prepare_dense(X, *Y, scaling);
  Mat& Xf = f.pr.Xs;
  const size_t n = Xf.r, p = Xf.c;
  Mat Yc;
  if (!labels1) {  // opls.hpp:283-289: the centred response of the refreshes
    Yc.resize(n, q);
    for (size_t c = 0; c < q; ++c)
      for (size_t i = 0; i < n; ++i) Yc(i, c) = (*Y)(i, c) - f.pr.my[c];
  }
  Mat S = f.pr.S;
  f.W.resize(p, north);
  f.P.resize(p, north);
  for (size_t comp = 0; comp < north; ++comp) {
    if (comp > 0) S = labels1 ? label_crossprod(Xf, labels1, f.pr.my) : mul(Xf, true, Yc, false);
    Triplets tr = randomized_svd(S, 1, oversample, power, seed + unsigned(comp), true);
    if (tr.U.c == 0) break;
    Mat w(p, 1);
    for (size_t j = 0; j < p; ++j) w(j, 0) = tr.U(j, 0);
    if (!opls_normalize(w)) break;
    Mat t = mul(Xf, false, w, false);
    const double tt = opls_dot(t, t);
    if (!std::isfinite(tt) || tt <= 0.0) break;
    Mat loading = mul(Xf, true, t, false);
    for (size_t j = 0; j < p; ++j) loading(j, 0) /= tt;
    const double ww = opls_dot(w, w);
    if (!std::isfinite(ww) || ww <= 0.0) break;
    const double proj = opls_dot(w, loading) / ww;
    Mat wo(p, 1);
    for (size_t j = 0; j < p; ++j) wo(j, 0) = loading(j, 0) - w(j, 0) * proj;
    if (!opls_normalize(wo)) break;
    Mat to = mul(Xf, false, wo, false);
    const double too = opls_dot(to, to);
    if (!std::isfinite(too) || too <= 0.0) break;
    Mat po = mul(Xf, true, to, false);
    for (size_t j = 0; j < p; ++j) {
      po(j, 0) /= too;
      f.W(j, f.completed) = wo(j, 0);
      f.P(j, f.completed) = po(j, 0);
    }
    Mat corr = mul(to, false, po, true);
    for (size_t e = 0; e < Xf.d.size(); ++e) Xf.d[e] -= corr.d[e];
    ++f.completed;
  }
  return f;
}

// Weights that take standardised rows straight to the scores of the inner
// model: R_eff = (I - w_1 p_1ᵀ) ... (I - w_m p_mᵀ) R (the filtered rows are
// x_s (I - w_1 p_1ᵀ) ... (I - w_m p_mᵀ), opls.hpp:467-509), applied from the
// last filter component to the first (src/cuda_resident_special.cuh:306-312).
Mat opls_effective_weights(const Mat& R, const Mat& W, const Mat& P, size_t completed) {
  Mat out = R;
  for (size_t c = completed; c > 0; --c) {
    const size_t k = c - 1;
    for (size_t a = 0; a < out.c; ++a) {
      double coef = 0.0;
      for (size_t j = 0; j < out.r; ++j) coef += P(j, k) * out(j, a);
      for (size_t j = 0; j < out.r; ++j) out(j, a) -= W(j, k) * coef;
    }
  }
  return out;
}

// ------------------------------------------------------------ kernel PLS
// kernels.hpp:20-62 (kernel_from_dots): 2 = radial basis exp(-γ d) with d =
// |a|² + |b|² - 2 a·b (small negative d, above -1e-10, clamped to zero),
// 3 = polynomial (γ a·b + offset)^degree.
Mat kernel_dots(const Mat& A, const Mat& B, int kind, double gamma, int degree, double offset) {
  Mat K = mul(A, false, B, true);
  if (kind == 3) {
    for (double& v : K.d) v = std::pow(gamma * v + offset, degree);
    return K;
  }
  std::vector<double> ln(A.r, 0.0), rn(B.r, 0.0);
  for (size_t f = 0; f < A.c; ++f) {
    for (size_t i = 0; i < A.r; ++i) ln[i] += A(i, f) * A(i, f);
    for (size_t i = 0; i < B.r; ++i) rn[i] += B(i, f) * B(i, f);
  }
  for (size_t j = 0; j < K.c; ++j)
    for (size_t i = 0; i < K.r; ++i) {
      double d = ln[i] + rn[j] - 2.0 * K(i, j);
      if (d < 0.0 && d > -1e-10) d = 0.0;
      K(i, j) = std::exp(-gamma * d);
    }
  return K;
}

struct KernelCentre {
  std::vector<double> colmeans;
  double grand = 0.0;
};

// kernels.hpp:128-154 (center_kernel_train)
KernelCentre centre_train_kernel(Mat& K) {
  KernelCentre kc;
  kc.colmeans.assign(K.c, 0.0);
  std::vector<double> rowmeans(K.r, 0.0);
  for (size_t j = 0; j < K.c; ++j) {
    for (size_t i = 0; i < K.r; ++i) {
      kc.colmeans[j] += K(i, j);
      rowmeans[i] += K(i, j);
    }
    kc.colmeans[j] /= double(K.r);
    kc.grand += kc.colmeans[j];
  }
  kc.grand /= double(K.c);
  for (double& v : rowmeans) v /= double(K.c);
  for (size_t j = 0; j < K.c; ++j)
    for (size_t i = 0; i < K.r; ++i) K(i, j) += kc.grand - rowmeans[i] - kc.colmeans[j];
  return kc;
}

// kernels.hpp:156-178 (center_kernel_test)
void centre_test_kernel(Mat& K, const KernelCentre& kc) {
  for (size_t i = 0; i < K.r; ++i) {
    double rm = 0.0;
    for (size_t j = 0; j < K.c; ++j) rm += K(i, j);
    rm /= double(K.c);
    for (size_t j = 0; j < K.c; ++j) K(i, j) += kc.grand - rm - kc.colmeans[j];
  }
}

// ------------------------------------------------------------ LDA
struct Lda { Mat means, linear; std::vector<double> constants, priors; double ridge = 0, relative = 0; };

// lda.hpp:32-84
bool lda_chol_solve(const Mat& pooled, const Mat& means, double ridge, Mat& linear) {
  const size_t a = pooled.r, K = means.r;
  Mat L(a, a);
  for (size_t i = 0; i < a; ++i) {
    for (size_t j = 0; j <= i; ++j) {
      double v = pooled(i, j);
      if (i == j) v += ridge;
      for (size_t k = 0; k < j; ++k) v -= L(i, k) * L(j, k);
      if (i == j) {
        if (!std::isfinite(v) || v <= 0) return false;
        L(i, j) = std::sqrt(v);
      } else {
        const double dg = L(j, j);
        if (!std::isfinite(dg) || dg <= 0) return false;
        L(i, j) = v / dg;
      }
    }
  }
  Mat sol(a, K);
  for (size_t c = 0; c < K; ++c) {
    for (size_t i = 0; i < a; ++i) {
      double v = means(c, i);
      for (size_t k = 0; k < i; ++k) v -= L(i, k) * sol(k, c);
      sol(i, c) = v / L(i, i);
    }
    for (size_t ii = a; ii-- > 0;) {
      double v = sol(ii, c);
      for (size_t k = ii + 1; k < a; ++k) v -= L(k, ii) * sol(k, c);
      v /= L(ii, ii);
      if (!std::isfinite(v)) return false;
      sol(ii, c) = v;
    }
  }
  linear.resize(K, a);
  for (size_t c = 0; c < K; ++c) for (size_t j = 0; j < a; ++j) linear(c, j) = sol(j, c);
  return true;
}

std::vector<Lda> finalize_lda(const Mat& pooled, const Mat& means, const std::vector<double>& counts,
                              size_t n, const std::vector<int>& prefixes);

// lda.hpp:156-223 (train_lda_prefixes)
std::vector<Lda> train_lda(const Mat& T, const int* labels1, size_t K,
                           const std::vector<int>& prefixes) {
  const size_t n = T.r;
  size_t maximum = 0;
  for (int pfx : prefixes) {
    if (pfx < 1 || size_t(pfx) > T.c) throw std::invalid_argument("oracle LDA prefix outside score dimension");
    maximum = std::max(maximum, size_t(pfx));
  }
  std::vector<double> counts(K, 0.0);
  Mat means(K, maximum);
  for (size_t i = 0; i < n; ++i) {
    const int c = labels1[i] - 1;
    if (c < 0 || size_t(c) >= K) throw std::invalid_argument("oracle LDA label outside 1..K");
    counts[size_t(c)] += 1;
    for (size_t j = 0; j < maximum; ++j) means(size_t(c), j) += T(i, j);
  }
  for (size_t c = 0; c < K; ++c) {
    if (counts[c] <= 0) throw std::invalid_argument("oracle LDA empty class");
    for (size_t j = 0; j < maximum; ++j) means(c, j) /= counts[c];
  }
  Mat pooled(maximum, maximum);
  const double den = double(std::max<size_t>(1, n > K ? n - K : 1));
  for (size_t c2 = 0; c2 < maximum; ++c2) {
    for (size_t r2 = 0; r2 <= c2; ++r2) {
      double v = 0;
      for (size_t i = 0; i < n; ++i) v += T(i, r2) * T(i, c2);
      for (size_t c = 0; c < K; ++c) v -= counts[c] * means(c, r2) * means(c, c2);
      v /= den;
      pooled(r2, c2) = v; pooled(c2, r2) = v;
    }
  }
  return finalize_lda(pooled, means, counts, n, prefixes);
}

// lda.hpp:226-300 (train_lda_prefixes_from_moments): class means and the
// pooled covariance from the score Gram and the per-class score sums.
std::vector<Lda> train_lda_moments(const Mat& gram, const Mat& sums, const std::vector<double>& counts,
                                   size_t n, const std::vector<int>& prefixes) {
  const size_t K = sums.r;
  size_t maximum = 0;
  for (int pfx : prefixes) {
    if (pfx < 1 || size_t(pfx) > gram.c) throw std::invalid_argument("oracle LDA prefix outside moment dimension");
    maximum = std::max(maximum, size_t(pfx));
  }
  Mat means(K, maximum);
  double total = 0;
  for (size_t c = 0; c < K; ++c) {
    if (!std::isfinite(counts[c]) || counts[c] <= 0) throw std::invalid_argument("oracle LDA empty class");
    total += counts[c];
    for (size_t j = 0; j < maximum; ++j) {
      if (!std::isfinite(sums(c, j))) throw std::invalid_argument("oracle LDA moments must be finite");
      means(c, j) = sums(c, j) / counts[c];
    }
  }
  if (std::abs(total - double(n)) > 1e-8 * std::max(1.0, double(n)))
    throw std::invalid_argument("oracle LDA class counts do not sum to n");
  Mat pooled(maximum, maximum);
  const double den = double(std::max<size_t>(1, n > K ? n - K : 1));
  for (size_t c2 = 0; c2 < maximum; ++c2) {
    for (size_t r2 = 0; r2 <= c2; ++r2) {
      double v = gram(r2, c2);
      if (!std::isfinite(v)) throw std::invalid_argument("oracle LDA moments must be finite");
      for (size_t c = 0; c < K; ++c) v -= counts[c] * means(c, r2) * means(c, c2);
      v /= den;
      pooled(r2, c2) = v; pooled(c2, r2) = v;
    }
  }
  return finalize_lda(pooled, means, counts, n, prefixes);
}

// lda.hpp:86-153 (finalize_lda_prefixes): per prefix, the ridge ladder over
// the pooled covariance, Cholesky solves for the linear weights, constants.
std::vector<Lda> finalize_lda(const Mat& pooled, const Mat& means, const std::vector<double>& counts,
                              size_t n, const std::vector<int>& prefixes) {
  const size_t K = means.r;
  static const double grid[] = {1e-8, 1e-6, 1e-5, 1e-4, 1e-3, 1e-2};
  std::vector<Lda> models;
  for (int pfx : prefixes) {
    const size_t k = size_t(pfx);
    Lda m;
    m.means.resize(K, k);
    Mat cov(k, k);
    double scale = 0;
    for (size_t j = 0; j < k; ++j) {
      scale += pooled(j, j);
      for (size_t i = 0; i < k; ++i) cov(i, j) = pooled(i, j);
      for (size_t c = 0; c < K; ++c) m.means(c, j) = means(c, j);
    }
    scale /= double(k);
    if (!std::isfinite(scale) || scale <= 0) scale = 1.0;
    bool solved = false;
    for (double rel : grid) {
      if (lda_chol_solve(cov, m.means, rel * scale, m.linear)) {
        m.ridge = rel * scale; m.relative = rel; solved = true; break;
      }
    }
    if (!solved) throw std::runtime_error("oracle LDA Cholesky failed for every ridge");
    m.constants.resize(K); m.priors.resize(K);
    for (size_t c = 0; c < K; ++c) {
      m.priors[c] = counts[c] / double(n);
      double prod = 0;
      for (size_t j = 0; j < k; ++j) prod += m.means(c, j) * m.linear(c, j);
      m.constants[c] = -0.5 * prod + std::log(std::max(m.priors[c], std::numeric_limits<double>::min()));
    }
    models.push_back(std::move(m));
  }
  return models;
}

// Top-k class indices (one-based) with the reference tie rule: the earlier
// class wins a tie (src/cuda_resident_api.cuh:29-53, classification.hpp:237-243).
void topk_rows(const double* values, size_t rows, size_t classes, size_t top, int* out) {
  std::vector<size_t> idx(classes);
  for (size_t r = 0; r < rows; ++r) {
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
      const double va = values[r + a * rows], vb = values[r + b * rows];
      if (std::isnan(va)) return false;
      if (std::isnan(vb)) return true;
      return va > vb;
    });
    for (size_t k = 0; k < top; ++k) out[r + k * rows] = int(idx[k]) + 1;
  }
}

thread_local std::string g_error;

template<class F>
int guarded(F&& f) {
  try { f(); return 0; } catch (const std::exception& e) { g_error = e.what(); return 1; }
}

Mat wrap(const double* x, size_t r, size_t c) {
  Mat m(r, c);
  if (r != 0 && c != 0) std::memcpy(m.data(), x, sizeof(double) * r * c);
  return m;
}

void copy_out(const Mat& m, double* out) {
  if (out && !m.d.empty()) std::memcpy(out, m.data(), sizeof(double) * m.d.size());
}

}  // namespace

extern "C" {

// Binds the products to the BLAS dgemm exported as `symbol` by the shared
// library at `path` (CPU-baseline timing), with `threads` BLAS threads set
// through `threads_symbol` when both are given; path NULL unbinds.
int orc_use_blas(const char* path, const char* symbol, const char* threads_symbol,
                 int threads) {
  if (!path) { g_dgemm = nullptr; return 0; }
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) { g_error = dlerror(); return 1; }
  DgemmFn f = reinterpret_cast<DgemmFn>(dlsym(h, symbol));
  if (!f) { g_error = std::string("missing BLAS symbol ") + symbol; return 1; }
  if (threads_symbol && threads > 0) {
    typedef void (*SetThreads)(int);
    SetThreads st = reinterpret_cast<SetThreads>(dlsym(h, threads_symbol));
    if (!st) { g_error = std::string("missing BLAS symbol ") + threads_symbol; return 1; }
    st(threads);
  }
  g_dgemm = f;
  return 0;
}

const char* orc_last_error() { return g_error.c_str(); }

void orc_gaussian(unsigned int seed, size_t count, double* out) { gaussian(seed, count, out); }

// Dense preprocessing (supervised.hpp:27-124). xs may be null.
int orc_prepare_dense(const double* x, const double* y, int n, int p, int q,
                      int scaling, double* S, double* mx, double* sx, double* my,
                      double* xs) {
  return guarded([&] {
    Prepared pr = prepare_dense(wrap(x, n, p), wrap(y, n, q), scaling);
    copy_out(pr.S, S);
    std::copy(pr.mx.begin(), pr.mx.end(), mx);
    std::copy(pr.sx.begin(), pr.sx.end(), sx);
    std::copy(pr.my.begin(), pr.my.end(), my);
    if (xs) copy_out(pr.Xs, xs);
  });
}

// Label preprocessing (classification.hpp:86-184); labels one-based.
int orc_prepare_labels(const double* x, const int* labels, int n, int p, int K,
                       int scaling, double* S, double* mx, double* sx,
                       double* priors, double* xs) {
  return guarded([&] {
    Prepared pr = prepare_labels(wrap(x, n, p), labels, size_t(K), scaling);
    copy_out(pr.S, S);
    std::copy(pr.mx.begin(), pr.mx.end(), mx);
    std::copy(pr.sx.begin(), pr.sx.end(), sx);
    std::copy(pr.my.begin(), pr.my.end(), priors);
    if (xs) copy_out(pr.Xs, xs);
  });
}

// Routing decisions (src/r_api.cpp:1149-1190); out = {block, cache, batch_geometry}.
void orc_simpls_controls(int n, int p, int q, int A, int classification, int* out) {
  SimplsControls c = simpls_controls(n, p, q, A, classification != 0, 32, 5, 1);
  out[0] = int(c.block); out[1] = c.cache ? 1 : 0; out[2] = c.batch_geometry ? 1 : 0;
}

// Full SIMPLS fit: preprocessing + routing + component loop.
// y (n x q) or labels (one-based, q classes). Outputs sized p*A, q*A, p*A, n*A.
// reorth: -1 keeps the routing default, 0/1 forces it.
int orc_simpls_fit(const double* x, const double* y, const int* labels, int n,
                   int p, int q, int ncomp, int scaling, int oversample,
                   int power, unsigned int seed, int reorth, double* R,
                   double* Q, double* V, double* T, double* mx, double* sx,
                   double* my, int* completed) {
  return guarded([&] {
    const Mat X = wrap(x, n, p);
    Prepared pr = labels ? prepare_labels(X, labels, size_t(q), scaling)
                         : prepare_dense(X, wrap(y, n, q), scaling);
    const size_t cap = std::min<size_t>(size_t(p), std::max<size_t>(size_t(n) - 1, 1));
    const size_t A = std::max<size_t>(1, std::min<size_t>(size_t(ncomp), cap));
    SimplsControls ctl = simpls_controls(n, p, q, A, labels != nullptr, oversample, power, seed);
    if (reorth >= 0) ctl.reorth = reorth != 0;
    SimplsModel m = fit_simpls(pr.Xs, pr.S, ctl);
    copy_out(m.R, R); copy_out(m.Q, Q); copy_out(m.V, V);
    if (T) copy_out(m.T, T);
    std::copy(pr.mx.begin(), pr.mx.end(), mx);
    std::copy(pr.sx.begin(), pr.sx.end(), sx);
    std::copy(pr.my.begin(), pr.my.end(), my);
    *completed = int(m.completed);
  });
}

// SIMPLS core on prepared inputs with explicit controls (used by the
// known-answer tests restated from core/tests/test_core_simpls.cpp).
int orc_simpls_core(const double* xs, const double* S0, int n, int p, int q,
                    int components, int block, int cache, int reorth,
                    int batch_geometry, int oversample, int power,
                    unsigned int seed, double* R, double* Q, double* V,
                    double* T, int* completed) {
  return guarded([&] {
    SimplsControls ctl;
    ctl.components = size_t(components); ctl.block = size_t(block);
    ctl.cache = cache != 0; ctl.reorth = reorth != 0;
    ctl.batch_geometry = batch_geometry != 0;
    ctl.oversample = oversample; ctl.power = power; ctl.seed = seed;
    SimplsModel m = fit_simpls(wrap(xs, n, p), wrap(S0, p, q), ctl);
    copy_out(m.R, R); copy_out(m.Q, Q); copy_out(m.V, V);
    if (T) copy_out(m.T, T);
    *completed = int(m.completed);
  });
}

// Explicit randomized SVD (rsvd.hpp:185-237). U m*k, s k, Vt k*n.
int orc_rsvd(const double* a, int m, int n, int k, int oversample, int power,
             unsigned int seed, double* U, double* s, double* Vt, int* got) {
  return guarded([&] {
    Triplets t = randomized_svd(wrap(a, m, n), k, oversample, power, seed, false);
    copy_out(t.U, U);
    std::copy(t.s.begin(), t.s.end(), s);
    copy_out(t.Vt, Vt);
    *got = int(t.s.size());
  });
}

// Audited operator rSVD of an explicit matrix (audited_rsvd.hpp:298-414).
// U m*k, s k, Vt k*n (untouched when left_only). audit[10]: certified,
// deterministic fallback, attempts, effective oversample, power, seed,
// subspace error, singular value error, triplet residual, omitted ratio.
int orc_audited_rsvd(const double* a, int m, int n, int k, int oversample, int power,
                     unsigned int seed, int left_only, double* U, double* s, double* Vt,
                     int* got, double* audit) {
  return guarded([&] {
    Audit au;
    Triplets t = audited_rsvd(wrap(a, m, n), k, oversample, power, seed, left_only != 0, au);
    copy_out(t.U, U);
    std::copy(t.s.begin(), t.s.end(), s);
    if (!left_only) copy_out(t.Vt, Vt);
    *got = int(t.s.size());
    const double v[10] = {double(au.certified), double(au.fallback), double(au.attempts),
                          double(au.oversample), double(au.power), double(au.seed), au.subspace,
                          au.singular, au.residual, au.omitted};
    std::copy(v, v + 10, audit);
  });
}

// fastcor_values (src/r_api.cpp:4082-4142): out = L Rᵀ of the normalised
// units (units_a x units_b), or the per-unit dot products when diagonal.
// b null correlates a with itself (and ignores diagonal).
int orc_fastcor(const double* a, int ar, int ac, const double* b, int br, int bc, int by_row,
                int diagonal, double* out) {
  return guarded([&] {
    const Mat L = normalized_units(wrap(a, ar, ac), by_row != 0);
    Mat R;
    const Mat* right = &L;
    bool diag = diagonal != 0;
    if (b) {
      R = normalized_units(wrap(b, br, bc), by_row != 0);
      if (L.c != R.c) throw std::invalid_argument("a and b must contain the same number of correlated features");
      right = &R;
    } else {
      diag = false;
    }
    if (diag) {
      if (L.r != right->r) throw std::invalid_argument("diag = TRUE requires matching numbers of rows or columns");
      for (size_t u = 0; u < L.r; ++u) {
        long double sum = 0.0L;
        for (size_t f = 0; f < L.c; ++f) sum += (long double)L(u, f) * (*right)(u, f);
        out[u] = double(sum);
      }
      return;
    }
    copy_out(mul(L, false, *right, true), out);
  });
}

// PLS-SVD fit for prefixes 1..A (plssvd.hpp:57-220). W holds A slabs of A*q
// doubles; slab k-1 stores the k x q weights with leading dimension k.
int orc_plssvd_fit(const double* x, const double* y, const int* labels, int n,
                   int p, int q, int A, int scaling, int oversample, int power,
                   unsigned int seed, double* R, double* Q, double* T,
                   double* sigma, double* W, double* mx, double* sx,
                   double* my) {
  return guarded([&] {
    const Mat X = wrap(x, n, p);
    Prepared pr = labels ? prepare_labels(X, labels, size_t(q), scaling)
                         : prepare_dense(X, wrap(y, n, q), scaling);
    std::vector<int> comps(A);
    std::iota(comps.begin(), comps.end(), 1);
    PlssvdModel m = fit_plssvd(pr.Xs, pr.S, comps, oversample, power, seed);
    copy_out(m.R, R); copy_out(m.Q, Q);
    if (T) copy_out(m.T, T);
    std::copy(m.s.begin(), m.s.end(), sigma);
    if (W) {
      for (int k = 1; k <= A; ++k)
        std::memcpy(W + size_t(k - 1) * A * q, m.W[k - 1].data(), sizeof(double) * k * q);
    }
    std::copy(pr.mx.begin(), pr.mx.end(), mx);
    std::copy(pr.sx.begin(), pr.sx.end(), sx);
    std::copy(pr.my.begin(), pr.my.end(), my);
  });
}

// Matrix-free SIMPLS (simpls.hpp:768-1047) after dense preprocessing, with
// the implicit route's controls for entries of `elem` bytes (8 or 4).
// Outputs as orc_simpls_fit; route[0] = 1 when the rank-one route ran.
int orc_simpls_operator_fit(const double* x, const double* y, int n, int p, int q, int ncomp,
                            int scaling, int oversample, int power, unsigned int seed, int elem,
                            double* R, double* Q, double* V, double* T, double* mx, double* sx,
                            double* my, int* completed, int* route) {
  return guarded([&] {
    const Mat X = wrap(x, n, p);
    const Mat Y = wrap(y, n, q);
    Prepared pr = prepare_dense(X, Y, scaling, false);
    const size_t cap = std::min<size_t>(size_t(p), std::max<size_t>(size_t(n) - 1, 1));
    const size_t A = std::max<size_t>(1, std::min<size_t>(size_t(ncomp), cap));
    bool rank_one = false;
    SimplsControls ctl = operator_controls(n, p, q, A, oversample, power, seed, size_t(elem),
                                           rank_one);
    SimplsModel m = fit_simpls_operator(pr.Xs, Y, pr.my, ctl, rank_one);
    if (m.completed < A) throw std::runtime_error("oracle implicit SIMPLS returned fewer components than requested");
    copy_out(m.R, R); copy_out(m.Q, Q); copy_out(m.V, V);
    if (T) copy_out(m.T, T);
    std::copy(pr.mx.begin(), pr.mx.end(), mx);
    std::copy(pr.sx.begin(), pr.sx.end(), sx);
    std::copy(pr.my.begin(), pr.my.end(), my);
    *completed = int(m.completed);
    if (route) route[0] = rank_one ? 1 : 0;
  });
}

// Matrix-free PLS-SVD (plssvd.hpp:222-248) after dense preprocessing, for
// prefixes 1..A; outputs as orc_plssvd_fit.
int orc_plssvd_operator_fit(const double* x, const double* y, int n, int p, int q, int A,
                            int scaling, int oversample, int power, unsigned int seed,
                            double* R, double* Q, double* T, double* sigma, double* W,
                            double* mx, double* sx, double* my) {
  return guarded([&] {
    const Mat X = wrap(x, n, p);
    const Mat Y = wrap(y, n, q);
    Prepared pr = prepare_dense(X, Y, scaling, false);
    std::vector<int> comps(A);
    std::iota(comps.begin(), comps.end(), 1);
    PlssvdModel m = fit_plssvd_operator(pr.Xs, Y, pr.my, comps, oversample, power, seed);
    copy_out(m.R, R); copy_out(m.Q, Q);
    if (T) copy_out(m.T, T);
    std::copy(m.s.begin(), m.s.end(), sigma);
    if (W) {
      for (int k = 1; k <= A; ++k)
        std::memcpy(W + size_t(k - 1) * A * q, m.W[k - 1].data(), sizeof(double) * k * q);
    }
    std::copy(pr.mx.begin(), pr.mx.end(), mx);
    std::copy(pr.sx.begin(), pr.sx.end(), sx);
    std::copy(pr.my.begin(), pr.my.end(), my);
  });
}

int orc_implicit_route(int p, int q) { return implicit_route(size_t(p), size_t(q)) ? 1 : 0; }

// Matrix-free SIMPLS after dense preprocessing with explicit controls (the
// counterpart of ref_simpls_operator_fit, for the golden cases).
int orc_simpls_operator_core(const double* x, const double* y, int n, int p, int q, int A,
                             int scaling, int oversample, int power, unsigned int seed,
                             int block, int cache, int reorth, int batch_geometry, int rank_one,
                             double* R, double* Q, double* V, double* T, int* completed) {
  return guarded([&] {
    const Mat X = wrap(x, n, p);
    const Mat Y = wrap(y, n, q);
    Prepared pr = prepare_dense(X, Y, scaling, false);
    SimplsControls ctl;
    ctl.components = size_t(A); ctl.block = size_t(block);
    ctl.cache = cache != 0; ctl.reorth = reorth != 0; ctl.batch_geometry = batch_geometry != 0;
    ctl.oversample = oversample; ctl.power = power; ctl.seed = seed;
    SimplsModel m = fit_simpls_operator(pr.Xs, Y, pr.my, ctl, rank_one != 0);
    copy_out(m.R, R); copy_out(m.Q, Q); copy_out(m.V, V);
    if (T) copy_out(m.T, T);
    *completed = int(m.completed);
  });
}

// PLS-SVD core on a prepared predictor block and cross-product (for the
// known-answer test of core/tests/test_core_plssvd.cpp).
int orc_plssvd_core(const double* xs, const double* S, int n, int p, int q,
                    const int* comps, int count, int oversample, int power,
                    unsigned int seed, double* R, double* Q, double* T,
                    double* sigma, double* W_last) {
  return guarded([&] {
    std::vector<int> c(comps, comps + count);
    PlssvdModel m = fit_plssvd(wrap(xs, n, p), wrap(S, p, q), c, oversample, power, seed);
    copy_out(m.R, R); copy_out(m.Q, Q); copy_out(m.T, T);
    std::copy(m.s.begin(), m.s.end(), sigma);
    copy_out(m.W.back(), W_last);
  });
}

// Ŷ for SIMPLS at one prefix: ((X - mx)/sx) R_k Q_kᵀ + my
// (cuda_resident_simpls.cuh:332-370, simpls.hpp:1049-1076).
int orc_simpls_predict(const double* x, int rows, int p, int q, int A,
                       const double* R, const double* Q, const double* mx,
                       const double* sx, const double* my, int prefix,
                       double* out) {
  return guarded([&] {
    Mat Xs(rows, p);
    for (int j = 0; j < p; ++j) for (int i = 0; i < rows; ++i)
      Xs(i, j) = (x[i + size_t(j) * rows] - mx[j]) / sx[j];
    Mat Rk(p, prefix), Qk(q, prefix);
    for (int k = 0; k < prefix; ++k) {
      for (int i = 0; i < p; ++i) Rk(i, k) = R[i + size_t(k) * p];
      for (int c = 0; c < q; ++c) Qk(c, k) = Q[c + size_t(k) * q];
    }
    (void)A;
    Mat Tk = mul(Xs, false, Rk, false);
    Mat Yh = mul(Tk, false, Qk, true);
    for (int c = 0; c < q; ++c) for (int i = 0; i < rows; ++i) Yh(i, c) += my[c];
    copy_out(Yh, out);
  });
}

// LDA from score moments (lda.hpp:226-346): linear K x prefix, constants K,
// the ridge and its relative level.
int orc_lda_moments(const double* gram, int a, const double* sums, const int* counts, int K,
                    long long n, int prefix, double* linear, double* constants, double* ridge,
                    double* relative) {
  return guarded([&] {
    std::vector<double> cnt(counts, counts + K);
    auto models = train_lda_moments(wrap(gram, a, a), wrap(sums, K, a), cnt, size_t(n), {prefix});
    const Lda& m = models[0];
    std::copy(m.linear.d.begin(), m.linear.d.end(), linear);
    std::copy(m.constants.begin(), m.constants.end(), constants);
    *ridge = m.ridge;
    *relative = m.relative;
  });
}

// PLS-SVD prefix weights (plssvd.hpp:138-164): W_k = solve(G_k, diag(sigma))
// Q_kᵀ at W + (k - 1) A q (k x q); route[k - 1] = 0 Cholesky, 1 pivoted.
int orc_plssvd_weights(const double* G, const double* sigma, const double* Q, int A, int q,
                       double* W, int* route) {
  return guarded([&] {
    const Mat g = wrap(G, A, A), qm = wrap(Q, q, A);
    for (int k = 1; k <= A; ++k) {
      Mat gk(k, k), dg(k, k), lat;
      for (int j = 0; j < k; ++j) {
        dg(j, j) = sigma[j];
        for (int i = 0; i < k; ++i) gk(i, j) = g(i, j);
      }
      if (cholesky_solve(gk, dg, lat)) {
        route[k - 1] = 0;
      } else if (pivoted_solve(gk, dg, lat)) {
        route[k - 1] = 1;
      } else {
        throw std::runtime_error("oracle PLS-SVD latent solve failed");
      }
      Mat qk(q, k);
      for (int j = 0; j < k; ++j) for (int i = 0; i < q; ++i) qk(i, j) = qm(i, j);
      const Mat wk = mul(lat, false, qk, true);
      std::copy(wk.d.begin(), wk.d.end(), W + size_t(k - 1) * A * q);
    }
  });
}

// LDA training on scores (lda.hpp:156-223) followed by discriminant scores
// for a test score block (lda.hpp:308-327). disc is rows x K; labels out
// one-based by first maximum.
int orc_lda(const double* T, int n, int a, const int* labels, int K, int prefix,
            const double* test_scores, int rows, double* disc, int* predicted,
            double* ridge) {
  return guarded([&] {
    std::vector<int> pf{prefix};
    auto models = train_lda(wrap(T, n, a), labels, size_t(K), pf);
    const Lda& m = models[0];
    for (int c = 0; c < K; ++c) {
      for (int i = 0; i < rows; ++i) {
        double v = m.constants[size_t(c)];
        for (int j = 0; j < prefix; ++j) v += test_scores[i + size_t(j) * rows] * m.linear(size_t(c), size_t(j));
        disc[i + size_t(c) * rows] = v;
      }
    }
    for (int i = 0; i < rows; ++i) {
      int best = 0;
      double bv = disc[i];
      for (int c = 1; c < K; ++c) if (disc[i + size_t(c) * rows] > bv) { bv = disc[i + size_t(c) * rows]; best = c; }
      predicted[i] = best + 1;
    }
    *ridge = m.ridge;
  });
}

void orc_topk(const double* values, int rows, int classes, int top, int* out) {
  topk_rows(values, size_t(rows), size_t(classes), size_t(top), out);
}

// Per-response SSE, SS about the training mean, SS about the observed mean
// (cuda_resident_metrics.cuh:9-43). out is 3 x q.
int orc_response_sums(const double* pred, const double* observed,
                      const int* labels, const double* train_mean, int n,
                      int q, double* out) {
  return guarded([&] {
    for (int c = 0; c < q; ++c) {
      double sum = 0;
      for (int i = 0; i < n; ++i) {
        if (labels && (labels[i] < 1 || labels[i] > q)) throw std::invalid_argument("oracle label outside 1..q");
        const double y = labels ? double(labels[i] == c + 1) : observed[i + size_t(c) * n];
        if (!std::isfinite(y)) throw std::invalid_argument("oracle non-finite response");
        sum += y;
      }
      const double mean = sum / double(n);
      double sse = 0, tss = 0, oss = 0;
      for (int i = 0; i < n; ++i) {
        const double y = labels ? double(labels[i] == c + 1) : observed[i + size_t(c) * n];
        const double pr = pred[i + size_t(c) * n];
        if (!std::isfinite(pr)) throw std::invalid_argument("oracle non-finite prediction");
        const double e = y - pr, d = y - train_mean[c], v = y - mean;
        sse += e * e; tss += d * d; oss += v * v;
      }
      out[3 * c] = sse; out[3 * c + 1] = tss; out[3 * c + 2] = oss;
    }
  });
}

// X-variance explained without an n x p residual
// (cuda_resident_variance.cuh:10-98). P is p x a; ss has a + 1 entries.
int orc_variance(const double* xs, const double* T, int n, int p, int a,
                 double* P, double* ss) {
  return guarded([&] {
    Mat X = wrap(xs, n, p), Tm = wrap(T, n, a);
    Mat cross = mul(X, true, Tm, false);
    Mat gram = mul(Tm, true, Tm, false);
    Mat seq(p, a);
    for (int f = 0; f < p; ++f) {
      for (int j = 0; j < a; ++j) {
        const double den = gram(j, j);
        double v = cross(f, j);
        P[f + size_t(j) * p] = (std::isfinite(den) && den > 0) ? v / den : 0.0;
        for (int i = 0; i < j; ++i) v -= seq(f, i) * gram(i, j);
        seq(f, j) = (std::isfinite(den) && den > 0) ? v / den : 0.0;
      }
    }
    double total = 0;
    for (double v : X.d) total += v * v;
    for (int j = 0; j < a; ++j) {
      double g = 0;
      for (int f = 0; f < p; ++f) g += seq(f, j) * seq(f, j);
      const double gain = g * gram(j, j);
      ss[j] = (std::isfinite(gain) && gain > 0) ? gain : 0.0;
    }
    ss[a] = total;
  });
}

// General product with the reference's fastpls_cuda_gemm contract
// (src/cuda_resident_api.h:11-13): out (rows x cols) = left (rows x inner) * right.
// OPLS: orthogonal filter (north components) + SIMPLS on the filtered
// predictors (scaling none, the same controls and seed; the resident
// engine's inner fit, src/cuda_resident_special.cuh:302-305). y (n x q) or
// one-based labels. Outputs: W, P (p x north; *filtered components), the
// inner R (p x A), Q (q x A), T (n x A), R_eff (p x A), outer mx, sx (p),
// my (q), *completed inner components.
int orc_opls_fit(const double* x, const double* y, const int* labels, int n, int p, int q,
                 int ncomp, int north, int scaling, int oversample, int power,
                 unsigned int seed, double* W, double* P, int* filtered, double* R,
                 double* Q, double* T, double* Reff, double* mx, double* sx, double* my,
                 int* completed) {
  return guarded([&] {
    const Mat X = wrap(x, n, p);
    Mat Y;
    if (!labels) Y = wrap(y, n, q);
    OplsFilter f = opls_filter(X, labels ? nullptr : &Y, labels, size_t(q), size_t(north),
                               scaling, oversample, power, seed);
    Prepared inner = labels ? prepare_labels(f.pr.Xs, labels, size_t(q), 3)
                            : prepare_dense(f.pr.Xs, Y, 3);
    const size_t cap = std::min<size_t>(size_t(p), std::max<size_t>(size_t(n) - 1, 1));
    const size_t A = std::max<size_t>(1, std::min<size_t>(size_t(ncomp), cap));
    SimplsControls ctl = simpls_controls(n, p, q, A, labels != nullptr, oversample, power, seed);
    SimplsModel m = fit_simpls(inner.Xs, inner.S, ctl);
    for (size_t c = 0; c < f.completed; ++c)
      for (int j = 0; j < p; ++j) {
        W[j + c * size_t(p)] = f.W(size_t(j), c);
        P[j + c * size_t(p)] = f.P(size_t(j), c);
      }
    *filtered = int(f.completed);
    copy_out(m.R, R); copy_out(m.Q, Q);
    if (T) copy_out(m.T, T);
    copy_out(opls_effective_weights(m.R, f.W, f.P, f.completed), Reff);
    std::copy(f.pr.mx.begin(), f.pr.mx.end(), mx);
    std::copy(f.pr.sx.begin(), f.pr.sx.end(), sx);
    std::copy(inner.my.begin(), inner.my.end(), my);
    *completed = int(m.completed);
  });
}

// Kernel PLS (kernelpls.hpp:87-153; with one-based labels the inner SIMPLS
// takes the label cross-product of the centred kernel, the resident
// engine's classification route, src/cuda_resident_special.cuh:491-503):
// standardised X (scaling), the training kernel (kind 2 radial basis, 3
// polynomial), its double centring, SIMPLS on it (scaling none). Outputs:
// R (n x A), Q (q x A), T (n x A), mx, sx (p), my (q), kernel column means
// (n) and grand mean, *completed. Predictions of `rows` test rows at
// `prefix` components go to pred (rows x q) when rows > 0.
int orc_kernelpls_fit(const double* x, const double* y, const int* labels, int n, int p, int q,
                      int ncomp, int scaling, int kind, double gamma, int degree,
                      double offset, int oversample, int power, unsigned int seed, double* R,
                      double* Q, double* T, double* mx, double* sx, double* my,
                      double* colmeans, double* grand, int* completed, const double* xtest,
                      int rows, int prefix, double* pred) {
  return guarded([&] {
    if (kind != 2 && kind != 3) throw std::invalid_argument("oracle kernel must be 2 or 3");
    const Mat X = wrap(x, n, p);
    Mat Y;
    if (!labels) Y = wrap(y, n, q);
    Prepared outer = labels ? prepare_labels(X, labels, size_t(q), scaling)
                            : prepare_dense(X, Y, scaling, false);
    Mat K = kernel_dots(outer.Xs, outer.Xs, kind, gamma, degree, offset);
    KernelCentre kc = centre_train_kernel(K);
    Prepared inner = labels ? prepare_labels(K, labels, size_t(q), 3) : prepare_dense(K, Y, 3);
    const size_t cap = std::max<size_t>(size_t(n) - 1, 1);
    const size_t A = std::max<size_t>(1, std::min<size_t>(size_t(ncomp), cap));
    SimplsControls ctl = simpls_controls(n, n, q, A, labels != nullptr, oversample, power, seed);
    SimplsModel m = fit_simpls(inner.Xs, inner.S, ctl);
    copy_out(m.R, R); copy_out(m.Q, Q);
    if (T) copy_out(m.T, T);
    std::copy(outer.mx.begin(), outer.mx.end(), mx);
    std::copy(outer.sx.begin(), outer.sx.end(), sx);
    std::copy(inner.my.begin(), inner.my.end(), my);
    std::copy(kc.colmeans.begin(), kc.colmeans.end(), colmeans);
    *grand = kc.grand;
    *completed = int(m.completed);
    if (rows > 0) {
      // kernelpls.hpp:155-189 (predict_kernelpls)
      Mat Xt{size_t(rows), size_t(p)};
      for (int j = 0; j < p; ++j)
        for (int i = 0; i < rows; ++i)
          Xt(size_t(i), size_t(j)) = (xtest[i + size_t(j) * rows] - outer.mx[j]) / outer.sx[j];
      Mat Kt = kernel_dots(Xt, outer.Xs, kind, gamma, degree, offset);
      centre_test_kernel(Kt, kc);
      Mat Rk{size_t(n), size_t(prefix)}, Qk{size_t(q), size_t(prefix)};
      for (int k = 0; k < prefix; ++k) {
        for (int i = 0; i < n; ++i) Rk(size_t(i), size_t(k)) = m.R(size_t(i), size_t(k));
        for (int c = 0; c < q; ++c) Qk(size_t(c), size_t(k)) = m.Q(size_t(c), size_t(k));
      }
      Mat Yh = mul(mul(Kt, false, Rk, false), false, Qk, true);
      for (int c = 0; c < q; ++c)
        for (int i = 0; i < rows; ++i) Yh(size_t(i), size_t(c)) += inner.my[c];
      copy_out(Yh, pred);
    }
  });
}

void orc_gemm(const double* left, const double* right, int rows, int inner,
              int cols, double* out) {
  gemm(left, size_t(rows), false, right, size_t(inner), false, size_t(rows),
       size_t(cols), size_t(inner), out, size_t(rows));
}

// Symmetric eigen and economy SVD exposed for their closed-form checks.
int orc_symmetric_eigen(double* a, int n, double* values) {
  return guarded([&] {
    Mat m = wrap(a, n, n);
    std::vector<double> ev;
    if (!symmetric_eigen(m, ev)) throw std::runtime_error("oracle eigen failed");
    copy_out(m, a);
    std::copy(ev.begin(), ev.end(), values);
  });
}

int orc_qr(const double* a, int m, int n, double* q) {
  return guarded([&] {
    Mat Q;
    if (!qr_economy(wrap(a, m, n), Q)) throw std::runtime_error("oracle QR failed");
    copy_out(Q, q);
  });
}

}  // extern "C"
