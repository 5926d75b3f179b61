// Resident OPLS and nonlinear kernel PLS handles (the reference's
// ResidentOpls and ResidentKernelPls, src/cuda_resident_special.cuh:155-546,
// behind fastpls_resident_{opls,kernelpls}_create, src/cuda_resident_api.h:48-56).
//
// Both run their own preprocessing on the device and then fit an inner
// SIMPLS handle on a device-resident feature matrix with scaling "none":
//   * OPLS: the standardised predictors filtered by `north` orthogonal
//     components (opls.hpp:123-221 with the randomized leading direction of
//     opls.hpp:314-363 / 418-465). Query rows are standardised and filtered
//     the same way (opls.hpp:467-509) before the inner model sees them; the
//     exported weights are R_eff = (I - w_1 p_1ᵀ)...(I - w_m p_mᵀ) R, which
//     take standardised rows straight to the scores.
//   * Kernel PLS: the double-centred radial-basis or polynomial kernel of
//     the standardised predictors (kernelpls.hpp:87-189, kernels.hpp:20-178);
//     query rows become centred test-kernel rows.
// The filter direction is the rank-one randomized SVD of the current cross-
// covariance, taken from a one-component PLS-SVD handle on the filtered
// matrix (its weights are exactly that direction: plssvd.hpp keeps U[:, :k]).
// All filter and kernel arithmetic is FP64; the inner model stores its
// feature matrix in the handle's precision, as the reference does.
#include "handle.hpp"
#include "host_util.hpp"
#include "kernels.hpp"
#include "ops.hpp"
#include "stream.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace fpls {
namespace {

constexpr int kRed = 1024;  // threads of the one-work-group reductions

// Deterministic sum over the work-group (fixed tree order).
__device__ double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = kRed / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// supervised.hpp:63-90 (column statistics by sums of squares) in FP64: one
// work-group per column. scaling 1 centres, 2 autoscales (a zero or
// non-finite deviation becomes 1), 3 leaves the column as it is.
template<class T>
__global__ __launch_bounds__(kRed) void k_column_stats(const T* X, long long n, int scaling,
                                                      double* mean, double* scale, int* flags) {
  __shared__ double red[kRed];
  const long long j = blockIdx.x;
  const T* col = X + j * n;
  double s = 0.0, ss = 0.0;
  for (long long i = threadIdx.x; i < n; i += kRed) {
    const double v = double(col[i]);
    s += v;
    ss += v * v;
  }
  s = block_sum(s, red);
  ss = block_sum(ss, red);
  if (threadIdx.x == 0) {
    double m = 0.0, sd = 1.0;
    if (scaling != 3) {
      m = s / double(n);
      if (scaling == 2) {
        const double css = fmax(0.0, ss - double(n) * m * m);
        sd = sqrt(css / double(n > 1 ? n - 1 : 1));
        if (!isfinite(sd) || sd <= 0.0) sd = 1.0;
      }
    }
    if (!isfinite(s) || !isfinite(ss)) atomicOr(flags, kFlagNonFinite);
    mean[j] = m;
    scale[j] = sd;
  }
}

// v /= |v| (opls.hpp:52-61); a zero or non-finite norm sets kFlagZeroNorm.
__global__ __launch_bounds__(kRed) void k_normalize(double* v, long long len, int* flags) {
  __shared__ double red[kRed];
  double s = 0.0;
  for (long long i = threadIdx.x; i < len; i += kRed) s += v[i] * v[i];
  s = block_sum(s, red);
  const double nrm = sqrt(fmax(s, 0.0));
  if (!isfinite(nrm) || nrm <= 0.0) {
    if (threadIdx.x == 0) atomicOr(flags, kFlagZeroNorm);
    return;
  }
  for (long long i = threadIdx.x; i < len; i += kRed) v[i] /= nrm;
}

// One filter component after the direction w (opls.hpp:155-183): with
// t = X_f w and the raw loading X_fᵀ t (`loading`, p), form
// loading /= tᵀt, then w_o = loading - w (wᵀ loading / wᵀ w), normalised.
__global__ __launch_bounds__(kRed) void k_opls_orthogonal(const double* t, long long n,
                                                         double* loading, const double* w,
                                                         int p, double* wo, int* flags) {
  __shared__ double red[kRed];
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += kRed) s += t[i] * t[i];
  const double tt = block_sum(s, red);
  if (!isfinite(tt) || tt <= 0.0) {
    if (threadIdx.x == 0) atomicOr(flags, kFlagZeroNorm);
    return;
  }
  double a = 0.0, b = 0.0;
  for (int j = threadIdx.x; j < p; j += kRed) {
    const double l = loading[j] / tt;
    loading[j] = l;
    a += w[j] * l;
    b += w[j] * w[j];
  }
  a = block_sum(a, red);
  const double ww = block_sum(b, red);
  if (!isfinite(ww) || ww <= 0.0) {
    if (threadIdx.x == 0) atomicOr(flags, kFlagZeroNorm);
    return;
  }
  const double proj = a / ww;
  double n2 = 0.0;
  for (int j = threadIdx.x; j < p; j += kRed) {
    const double v = loading[j] - w[j] * proj;
    wo[j] = v;
    n2 += v * v;
  }
  const double nrm = sqrt(fmax(block_sum(n2, red), 0.0));
  if (!isfinite(nrm) || nrm <= 0.0) {
    if (threadIdx.x == 0) atomicOr(flags, kFlagZeroNorm);
    return;
  }
  for (int j = threadIdx.x; j < p; j += kRed) wo[j] /= nrm;
}

// The orthogonal loading p_o = X_fᵀ t_o / t_oᵀ t_o (opls.hpp:185-198), in place.
__global__ __launch_bounds__(kRed) void k_opls_loading(const double* to, long long n, double* po,
                                                      int p, int* flags) {
  __shared__ double red[kRed];
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += kRed) s += to[i] * to[i];
  const double too = block_sum(s, red);
  if (!isfinite(too) || too <= 0.0) {
    if (threadIdx.x == 0) atomicOr(flags, kFlagZeroNorm);
    return;
  }
  for (int j = threadIdx.x; j < p; j += kRed) po[j] /= too;
}

// X (rows x cols, column-major) -= t vᵀ  (opls.hpp:200-209, :494-507)
__global__ __launch_bounds__(kWG) void k_rank_one_sub(double* X, long long rows, long long cols,
                                                      const double* t, const double* v) {
  const long long total = rows * cols;
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < total;
       e += (long long)gridDim.x * kWG) {
    const long long i = e % rows, j = e / rows;
    X[e] -= t[i] * v[j];
  }
}

// R_eff (p x A) = (I - W_0 P_0ᵀ) ... (I - W_{m-1} P_{m-1}ᵀ) R, applied from
// the last filter component to the first (src/cuda_resident_special.cuh:306-312);
// one work-group per column of R.
__global__ __launch_bounds__(kRed) void k_opls_effective(double* R, int p, const double* W,
                                                        const double* P, int north) {
  __shared__ double red[kRed];
  double* r = R + (long long)blockIdx.x * p;
  for (int c = north - 1; c >= 0; --c) {
    const double* w = W + (long long)c * p;
    const double* pc = P + (long long)c * p;
    double s = 0.0;
    for (int j = threadIdx.x; j < p; j += kRed) s += pc[j] * r[j];
    const double coef = block_sum(s, red);
    for (int j = threadIdx.x; j < p; j += kRed) r[j] -= w[j] * coef;
    __syncthreads();
  }
}

// out[i] = sum_j X[i, j]^2 for X rows x cols (column-major): thread per row.
__global__ __launch_bounds__(kWG) void k_row_sumsq(const double* X, long long rows,
                                                   long long cols, double* out) {
  const long long i = (long long)blockIdx.x * kWG + threadIdx.x;
  if (i >= rows) return;
  double s = 0.0;
  for (long long j = 0; j < cols; ++j) {
    const double v = X[i + j * rows];
    s += v * v;
  }
  out[i] = s;
}

// kernels.hpp:20-62 (kernel_from_dots) on dot products K (rows x cols):
// kind 2 radial basis exp(-γ d), d = |a_i|² + |b_j|² - 2 K_ij with small
// negative d (above -tol) clamped to zero; kind 3 polynomial (γ K + c)^degree.
__global__ __launch_bounds__(kWG) void k_kernel_transform(double* K, long long rows,
                                                          long long cols, int kind, double gamma,
                                                          int degree, double offset,
                                                          const double* lnorm,
                                                          const double* rnorm, double tol) {
  const long long total = rows * cols;
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < total;
       e += (long long)gridDim.x * kWG) {
    const double dot = K[e];
    if (kind == 3) {
      const double base = gamma * dot + offset;
      double v = 1.0;
      for (int k = 0; k < degree; ++k) v *= base;
      K[e] = v;
    } else {
      const long long i = e % rows, j = e / rows;
      double d = lnorm[i] + rnorm[j] - 2.0 * dot;
      if (d < 0.0 && d > -tol) d = 0.0;
      K[e] = exp(-gamma * d);
    }
  }
}

// Column means of K (rows x cols): one work-group per column.
__global__ __launch_bounds__(kRed) void k_col_means(const double* K, long long rows, double* out) {
  __shared__ double red[kRed];
  const double* col = K + (long long)blockIdx.x * rows;
  double s = 0.0;
  for (long long i = threadIdx.x; i < rows; i += kRed) s += col[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s / double(rows);
}

// Row means of K (rows x cols): thread per row.
__global__ __launch_bounds__(kWG) void k_row_means(const double* K, long long rows,
                                                   long long cols, double* out) {
  const long long i = (long long)blockIdx.x * kWG + threadIdx.x;
  if (i >= rows) return;
  double s = 0.0;
  for (long long j = 0; j < cols; ++j) s += K[i + j * rows];
  out[i] = s / double(cols);
}

// Grand mean: the mean of the column means (kernels.hpp:143-145).
__global__ __launch_bounds__(kRed) void k_mean(const double* v, long long len, double* out) {
  __shared__ double red[kRed];
  double s = 0.0;
  for (long long i = threadIdx.x; i < len; i += kRed) s += v[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[0] = s / double(len);
}

// K (rows x cols) += grand - rowmean_i - colmean_j  (kernels.hpp:147-151, :167-177)
__global__ __launch_bounds__(kWG) void k_kernel_centre(double* K, long long rows, long long cols,
                                                       const double* rowmean,
                                                       const double* colmean,
                                                       const double* grand) {
  const long long total = rows * cols;
  const double g = grand[0];
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < total;
       e += (long long)gridDim.x * kWG) {
    const long long i = e % rows, j = e / rows;
    K[e] += g - rowmean[i] - colmean[j];
  }
}

template<class T>
__global__ __launch_bounds__(kWG) void k_widen_copy(const T* in, long long count, double* out) {
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < count;
       e += (long long)gridDim.x * kWG)
    out[e] = double(in[e]);
}

unsigned stream_blocks(long long count) {
  return unsigned(std::max<long long>(1, std::min<long long>(65536, (count + kWG - 1) / kWG)));
}

// Free device memory (FASTPLS_HIP_QUERY_FREE_BYTES stands in for it when set).
size_t free_device_bytes() {
  size_t free_b = 0, total_b = 0;
  if (const char* v = std::getenv("FASTPLS_HIP_QUERY_FREE_BYTES"))
    return size_t(std::strtoull(v, nullptr, 10));
  FPLS_HIP(hipMemGetInfo(&free_b, &total_b));
  return free_b;
}

// --------------------------------------------------------------- common
// Shared plumbing of the two handles: the stream, the inner SIMPLS handle on
// a device feature matrix, the outer statistics and the query forwarding.
template<class T>
class SpecialBase : public Handle, protected StreamOps {
 public:
  ~SpecialBase() override { close(); }

  void predict_path(const void* x, int rows, const int* prefixes, int count, bool lda, void* out,
                    bool device_rows) override {
    if (!x || !out || rows < 1 || count < 1) throw std::invalid_argument("invalid resident prediction-path request");
    inner_->predict_path(features(x, rows, device_rows), rows, prefixes, count, lda, out, true);
  }
  void classify_path(const void* x, int rows, const int* prefixes, int count, bool lda, int top,
                     int* labels, void* responses, bool device_rows) override {
    if (!x || !labels || rows < 1 || count < 1)
      throw std::invalid_argument("invalid resident classification-path request");
    inner_->classify_path(features(x, rows, device_rows), rows, prefixes, count, lda, top, labels,
                          responses, true);
  }
  void response_sums(const void* x, const void* y, const int* labels, int rows, int prefix,
                     void* out) override {
    if (!x || rows < 1) throw std::invalid_argument("invalid resident prediction input");
    inner_->response_sums_device(features(x, rows, false), y, labels, rows, prefix, out);
  }
  void heldout_regression(const void*, const void*, int, const int*, int, double*,
                          double*) override {
    throw std::invalid_argument("held-out evaluation runs on SIMPLS and PLS-SVD handles only");
  }
  void project(const void* x, int rows, int prefix, void* out) override {
    if (!x || !out || rows < 1) throw std::invalid_argument("invalid resident score request");
    inner_->project_device(features(x, rows, false), rows, prefix, out);
  }
  void predict_scratch(const void*, int, int, void*, double*, size_t) override {
    throw std::invalid_argument("the allocation-free prediction serves SIMPLS and PLS-SVD handles");
  }
  void controls(int* oversample, int* power, int* block, int* block_limit, int* implicit,
                int* cache) override {
    inner_->controls(oversample, power, block, block_limit, implicit, cache);
  }
  void compact(bool prepare_lda) override {
    inner_->compact(prepare_lda);
    FPLS_HIP(hipStreamSynchronize(stream_));
    feat_.release();  // the inner model no longer reads its training features
  }
  double last_fit_seconds() const override { return fit_seconds_; }
  void set_batch_invariant(bool on) override {
    invariant_ = on;
    inner_->set_batch_invariant(on);
  }
  size_t device_bytes() const override {
    size_t total = inner_ ? inner_->device_bytes() : 0;
    for (const Buf* b : const_cast<SpecialBase*>(this)->bufs()) total += b->bytes();
    return total;
  }
  void phases(double* out, int count) const override {
    for (int i = 0; i < count; ++i) out[i] = i == 0 ? fit_seconds_ : 0.0;
  }
  void attach_stream(hipStream_t stream) override {
    if (!stream || stream == stream_) return;
    FPLS_HIP(hipStreamSynchronize(stream_));
    inner_->attach_stream(stream);
    for (Buf* b : bufs()) b->rebind(stream);
    stream_give(stream_);
    stream_ = stream;
    owns_ = false;
  }

 protected:
  SpecialBase(int n, int p, int q, int scaling) : n_(n), p_(p), q_(q), scaling_(scaling) {
    f64_products_ = true;
    stream_ = stream_take();
    for (Buf* b : bufs()) b->bind(stream_);
  }

  // Device rows (rows x feature columns, element T) that the inner model
  // takes for query rows x (host unless device_rows).
  virtual const T* features(const void* x, int rows, bool device_rows) = 0;

  std::vector<Buf*> bufs() {
    std::vector<Buf*> v = {&xin_, &yin_, &lab_, &mx_, &sx_, &flags_, &feat_, &qraw_, &q64_,
                           &qfeat_};
    for (Buf* b : extra_bufs()) v.push_back(b);
    for (Buf* b : ops_bufs()) v.push_back(b);
    return v;
  }
  virtual std::vector<Buf*> extra_bufs() { return {}; }

  void close() noexcept {
    inner_.reset();
    if (!stream_) return;
    (void)hipStreamSynchronize(stream_);
    for (Buf* b : bufs()) b->release();
    if (owns_) stream_give(stream_);
    stream_ = nullptr;
  }

  // Uploads X (n x p, element T), Y (n x q) or one-based labels, computes
  // the outer column statistics and writes the standardised predictors
  // (FP64) into `xs`.
  void load_standardized(const T* x, const T* y, const int* labels, Buf& xs) {
    const size_t np = size_t(n_) * p_;
    xin_.alloc(sizeof(T) * np);
    FPLS_HIP(hipMemcpyAsync(xin_.as<T>(), x, sizeof(T) * np, hipMemcpyHostToDevice, stream_));
    if (labels) {
      lab_.alloc(sizeof(int) * size_t(n_));
      FPLS_HIP(hipMemcpyAsync(lab_.as<int>(), labels, sizeof(int) * size_t(n_),
                              hipMemcpyHostToDevice, stream_));
    } else {
      yin_.alloc(sizeof(T) * size_t(n_) * q_);
      FPLS_HIP(hipMemcpyAsync(yin_.as<T>(), y, sizeof(T) * size_t(n_) * q_,
                              hipMemcpyHostToDevice, stream_));
    }
    flags_.alloc(sizeof(int));
    fill_zero_int(flags_.as<int>(), 1);
    mx_.alloc(sizeof(double) * p_);
    sx_.alloc(sizeof(double) * p_);
    hipLaunchKernelGGL((k_column_stats<T>), dim3(unsigned(p_)), dim3(kRed), 0, stream_,
                       xin_.as<T>(), (long long)n_, scaling_, mx_.d(), sx_.d(), flags_.as<int>());
    FPLS_LAUNCHED("k_column_stats");
    xs.alloc(sizeof(double) * np);
    launch_standardize_columns<T>(xin_.as<T>(), n_, 0, p_, mx_.d(), sx_.d(), xs.d(), stream_);
    xin_.release();
  }

  // Standardised FP64 query rows (rows x p) in q64_.
  void standardize_query(const void* x, int rows, bool device_rows) {
    const T* dx = static_cast<const T*>(x);
    if (!device_rows) {
      qraw_.reserve(sizeof(T) * size_t(rows) * p_);
      FPLS_HIP(hipMemcpyAsync(qraw_.as<T>(), x, sizeof(T) * size_t(rows) * p_,
                              hipMemcpyHostToDevice, stream_));
      dx = qraw_.as<T>();
    }
    q64_.reserve(sizeof(double) * size_t(rows) * p_);
    launch_standardize_columns<T>(dx, rows, 0, p_, mx_.d(), sx_.d(), q64_.d(), stream_);
  }

  // FP64 feature matrix (rows x cols) -> the handle's precision, for the inner model.
  const T* to_inner(const double* f, long long count, Buf& out) {
    if constexpr (std::is_same<T, double>::value) {
      (void)out;
      (void)count;
      return f;
    } else {
      out.reserve(sizeof(T) * size_t(count));
      launch_convert<T>(f, out.as<T>(), count, stream_);
      return out.as<T>();
    }
  }

  // The inner SIMPLS handle on the device feature matrix (n x cols).
  void fit_inner(const T* features, int cols, int components, int oversample, int power,
                 unsigned long long seed, bool retain_scores) {
    inner_.reset(create(0, features, yin_.empty() ? nullptr : yin_.as<T>(),
                        lab_.empty() ? nullptr : lab_.as<int>(), sizeof(T) == 4 ? 32 : 64, true,
                        n_, cols, q_, components, 3, oversample, power, seed, stream_,
                        retain_scores));
  }

  void check_flags(const char* what) {
    int bad = 0;
    FPLS_HIP(hipMemcpyAsync(&bad, flags_.as<int>(), sizeof(int), hipMemcpyDeviceToHost, stream_));
    FPLS_HIP(hipStreamSynchronize(stream_));
    if (bad) throw std::runtime_error(what);
  }

  int n_, p_, q_, scaling_;
  bool owns_ = true;
  double fit_seconds_ = 0.0;
  std::unique_ptr<Handle> inner_;
  Buf xin_, yin_, lab_, mx_, sx_, flags_, feat_, qraw_, q64_, qfeat_;
};

// ----------------------------------------------------------------- OPLS
template<class T>
class OplsHandle final : public SpecialBase<T> {
  using B = SpecialBase<T>;
  using B::stream_;

 public:
  OplsHandle(const T* x, const T* y, const int* labels, int n, int p, int q, int components,
             int scaling, int oversample, int power, unsigned long long seed, int north,
             bool retain_scores)
      : B(n, p, q, scaling), A_(components), north_(north) {
    if ((y == nullptr) == (labels == nullptr))
      throw std::invalid_argument("provide responses or labels, not both");
    if (!x) throw std::invalid_argument("null predictors");
    if (n < 2 || p < 1 || q < 1 || components < 1 || components > std::min(n - 1, p) || north < 0)
      throw std::invalid_argument("invalid resident HIP OPLS dimensions");
    if (scaling < 1 || scaling > 3) throw std::invalid_argument("scaling must be 1, 2 or 3");
    if (labels && q < 2) throw std::invalid_argument("classification needs at least two classes");
    for (Buf* b : extra_bufs()) b->bind(stream_);
    try {
      Event start, stop;
      FPLS_HIP(hipEventRecord(start.e, stream_));
      // the inner model reads the filtered predictors in the handle's
      // precision: FP64 handles filter their feature buffer in place
      Buf xf;
      xf.bind(stream_);
      constexpr bool kF64 = std::is_same<T, double>::value;
      Buf& work = kF64 ? this->feat_ : xf;
      this->load_standardized(x, y, labels, work);
      filter(work, oversample, power, seed);
      const size_t np = size_t(n) * p;
      if (!kF64) {
        this->feat_.alloc(sizeof(T) * np);
        launch_convert<T>(xf.d(), this->feat_.template as<T>(), (long long)np, stream_);
        xf.release();
      }
      this->fit_inner(this->feat_.template as<T>(), p, components, oversample, power, seed,
                      retain_scores);
      reff_.alloc(sizeof(double) * size_t(p) * A_);
      FPLS_HIP(hipMemcpyAsync(reff_.d(), this->inner_->device_weights(),
                              sizeof(double) * size_t(p) * A_, hipMemcpyDeviceToDevice, stream_));
      if (north_ > 0)
        hipLaunchKernelGGL(k_opls_effective, dim3(unsigned(A_)), dim3(kRed), 0, stream_, reff_.d(),
                           p, W_.d(), P_.d(), north_);
      FPLS_HIP(hipEventRecord(stop.e, stream_));
      this->check_flags("resident HIP OPLS filtering failed");
      float ms = 0.f;
      FPLS_HIP(hipEventElapsedTime(&ms, start.e, stop.e));
      this->fit_seconds_ = double(ms) * 1e-3;
    } catch (...) {
      this->close();
      throw;
    }
  }

  ~OplsHandle() override { this->close(); }

  const double* device_weights() const override { return reff_.d(); }

  void export_field(int field, void* out, size_t elements) override {
    const int p = this->p_;
    if (field == 0 || field == 3 || field == 4) {
      const size_t want = field == 0 ? size_t(p) * A_ : size_t(p);
      if (!out || elements != want) throw std::invalid_argument("invalid resident field output size");
      const double* src = field == 0 ? reff_.d() : field == 3 ? this->mx_.d() : this->sx_.d();
      this->qfeat_.reserve(sizeof(T) * want);
      launch_convert<T>(src, this->qfeat_.template as<T>(), (long long)want, stream_);
      FPLS_HIP(hipMemcpyAsync(out, this->qfeat_.template as<T>(), sizeof(T) * want,
                              hipMemcpyDeviceToHost, stream_));
      FPLS_HIP(hipStreamSynchronize(stream_));
      return;
    }
    this->inner_->export_field(field, out, elements);
  }

 private:
  std::vector<Buf*> extra_bufs() override { return {&W_, &P_, &reff_, &w_, &t_, &ld_}; }

  // opls.hpp:123-221 on the standardised predictors xf (n x p, FP64), in place.
  void filter(Buf& xf, int oversample, int power, unsigned long long seed) {
    const int n = this->n_, p = this->p_, q = this->q_;
    W_.alloc(sizeof(double) * size_t(p) * std::max(1, north_));
    P_.alloc(sizeof(double) * size_t(p) * std::max(1, north_));
    w_.alloc(sizeof(double) * p);
    t_.alloc(sizeof(double) * n);
    ld_.alloc(sizeof(double) * p);
    // FP64 responses for the direction fits (the one-component PLS-SVD
    // handles read device inputs of their own precision)
    Buf y64;
    y64.bind(stream_);
    if (!this->yin_.empty()) {
      y64.alloc(sizeof(double) * size_t(n) * q);
      hipLaunchKernelGGL((k_widen_copy<T>), dim3(stream_blocks((long long)n * q)), dim3(kWG), 0,
                         stream_, this->yin_.template as<T>(), (long long)n * q, y64.d());
    }
    int* flags = this->flags_.template as<int>();
    for (int c = 0; c < north_; ++c) {
      double* wo = W_.d() + (size_t)c * p;
      double* po = P_.d() + (size_t)c * p;
      {
        // the leading left singular vector of the current cross-covariance
        // (rank-one randomized SVD, seed + component)
        std::unique_ptr<Handle> dir(create(1, xf.d(), y64.empty() ? nullptr : y64.d(),
                                           this->lab_.empty() ? nullptr : this->lab_.template as<int>(),
                                           64, true, n, p, q, 1, 3, oversample, power,
                                           seed + (unsigned long long)unsigned(c), stream_, true));
        FPLS_HIP(hipMemcpyAsync(w_.d(), dir->device_weights(), sizeof(double) * p,
                                hipMemcpyDeviceToDevice, stream_));
        FPLS_HIP(hipStreamSynchronize(stream_));
      }
      hipLaunchKernelGGL(k_normalize, dim3(1), dim3(kRed), 0, stream_, w_.d(), (long long)p, flags);
      this->template product<double>(xf.d(), n, p, nullptr, w_.d(), p, 1, t_.d());
      this->template tproduct<double>(xf.d(), n, p, nullptr, t_.d(), n, 1, ld_.d());
      hipLaunchKernelGGL(k_opls_orthogonal, dim3(1), dim3(kRed), 0, stream_, t_.d(), (long long)n,
                         ld_.d(), w_.d(), p, wo, flags);
      this->template product<double>(xf.d(), n, p, nullptr, wo, p, 1, t_.d());
      this->template tproduct<double>(xf.d(), n, p, nullptr, t_.d(), n, 1, po);
      hipLaunchKernelGGL(k_opls_loading, dim3(1), dim3(kRed), 0, stream_, t_.d(), (long long)n, po,
                         p, flags);
      hipLaunchKernelGGL(k_rank_one_sub, dim3(stream_blocks((long long)n * p)), dim3(kWG), 0,
                         stream_, xf.d(), (long long)n, (long long)p, t_.d(), po);
      FPLS_LAUNCHED("OPLS filter component");
      // a failed component stops the fit (the resident engine's invalid flag)
      this->check_flags("resident HIP OPLS filtering failed");
    }
  }

  // standardised, filtered query rows in the handle's precision
  const T* features(const void* x, int rows, bool device_rows) override {
    const int p = this->p_;
    this->standardize_query(x, rows, device_rows);
    double* xq = this->q64_.d();
    for (int c = 0; c < north_; ++c) {
      t_.reserve(sizeof(double) * size_t(std::max(rows, this->n_)));
      this->template product<double>(xq, rows, p, nullptr, W_.d() + (size_t)c * p, p, 1, t_.d());
      hipLaunchKernelGGL(k_rank_one_sub, dim3(stream_blocks((long long)rows * p)), dim3(kWG), 0,
                         stream_, xq, (long long)rows, (long long)p, t_.d(),
                         P_.d() + (size_t)c * p);
      FPLS_LAUNCHED("k_rank_one_sub");
    }
    return this->to_inner(xq, (long long)rows * p, this->qfeat_);
  }

  int A_, north_;
  Buf W_, P_, reff_, w_, t_, ld_;
};

// ----------------------------------------------------------- kernel PLS
template<class T>
class KernelHandle final : public SpecialBase<T> {
  using B = SpecialBase<T>;
  using B::stream_;

 public:
  KernelHandle(const T* x, const T* y, const int* labels, int n, int p, int q, int components,
               int scaling, int oversample, int power, unsigned long long seed, int kind,
               double gamma, int degree, double coef0, bool retain_scores)
      : B(n, p, q, scaling), A_(components), kind_(kind), degree_(degree), gamma_(gamma),
        coef0_(coef0) {
    if ((y == nullptr) == (labels == nullptr))
      throw std::invalid_argument("provide responses or labels, not both");
    if (!x) throw std::invalid_argument("null predictors");
    if (n < 2 || p < 1 || q < 1 || components < 1 || components > n - 1 ||
        (kind != 2 && kind != 3) || degree < 1 || !(gamma > 0.0) || !std::isfinite(gamma) ||
        !std::isfinite(coef0))
      throw std::invalid_argument("invalid resident HIP nonlinear kernel PLS dimensions or controls");
    if (scaling < 1 || scaling > 3) throw std::invalid_argument("scaling must be 1, 2 or 3");
    if (labels && q < 2) throw std::invalid_argument("classification needs at least two classes");
    // src/cuda_resident_special.cuh:427-436: the n x n Gram and the training
    // predictors must fit in 3/5 of the free device memory
    const double gram_bytes = double(n) * double(n) * sizeof(double);
    const double need = gram_bytes + double(n) * double(p) * sizeof(double);
    if (need > double(free_device_bytes()) * 0.6)
      throw std::runtime_error(
          "nonlinear kernel PLS requires an n-by-n Gram matrix that exceeds the guarded HIP memory budget");
    for (Buf* b : extra_bufs()) b->bind(stream_);
    try {
      Event start, stop;
      FPLS_HIP(hipEventRecord(start.e, stream_));
      Buf xs;
      xs.bind(stream_);
      this->load_standardized(x, y, labels, xs);
      // the training rows, transposed (p x n): the right operand of every
      // kernel product
      xst_.alloc(sizeof(double) * size_t(n) * p);
      launch_transpose(xs.d(), n, p, xst_.d(), stream_);
      rnorm_.alloc(sizeof(double) * n);
      hipLaunchKernelGGL(k_row_sumsq, dim3(unsigned((n + kWG - 1) / kWG)), dim3(kWG), 0, stream_,
                         xs.d(), (long long)n, (long long)p, rnorm_.d());
      // the centred kernel is the inner model's feature matrix: FP64 handles
      // build it in their feature buffer
      Buf Kw;
      Kw.bind(stream_);
      constexpr bool kF64 = std::is_same<T, double>::value;
      Buf& K = kF64 ? this->feat_ : Kw;
      K.alloc(sizeof(double) * size_t(n) * n);
      this->template product<double>(xs.d(), n, p, nullptr, xst_.d(), p, n, K.d());
      xs.release();
      transform(K.d(), n, rnorm_.d());
      colmean_.alloc(sizeof(double) * n);
      grand_.alloc(sizeof(double));
      Buf rowmean;
      rowmean.bind(stream_);
      rowmean.alloc(sizeof(double) * n);
      hipLaunchKernelGGL(k_col_means, dim3(unsigned(n)), dim3(kRed), 0, stream_, K.d(),
                         (long long)n, colmean_.d());
      hipLaunchKernelGGL(k_row_means, dim3(unsigned((n + kWG - 1) / kWG)), dim3(kWG), 0, stream_,
                         K.d(), (long long)n, (long long)n, rowmean.d());
      hipLaunchKernelGGL(k_mean, dim3(1), dim3(kRed), 0, stream_, colmean_.d(), (long long)n,
                         grand_.d());
      hipLaunchKernelGGL(k_kernel_centre, dim3(stream_blocks((long long)n * n)), dim3(kWG), 0,
                         stream_, K.d(), (long long)n, (long long)n, rowmean.d(), colmean_.d(),
                         grand_.d());
      FPLS_LAUNCHED("kernel centring");
      if (!kF64) {
        this->feat_.alloc(sizeof(T) * size_t(n) * n);
        launch_convert<T>(Kw.d(), this->feat_.template as<T>(), (long long)n * n, stream_);
        Kw.release();
      }
      this->fit_inner(this->feat_.template as<T>(), n, components, oversample, power, seed,
                      retain_scores);
      FPLS_HIP(hipEventRecord(stop.e, stream_));
      this->check_flags("resident HIP kernel PLS preprocessing failed");
      float ms = 0.f;
      FPLS_HIP(hipEventElapsedTime(&ms, start.e, stop.e));
      this->fit_seconds_ = double(ms) * 1e-3;
    } catch (...) {
      this->close();
      throw;
    }
  }

  ~KernelHandle() override { this->close(); }

  const double* device_weights() const override { return this->inner_->device_weights(); }

  void export_field(int field, void* out, size_t elements) override {
    if (field == 3 || field == 4) {
      const size_t want = size_t(this->p_);
      if (!out || elements != want) throw std::invalid_argument("invalid resident field output size");
      this->qfeat_.reserve(sizeof(T) * want);
      launch_convert<T>(field == 3 ? this->mx_.d() : this->sx_.d(), this->qfeat_.template as<T>(),
                        (long long)want, stream_);
      FPLS_HIP(hipMemcpyAsync(out, this->qfeat_.template as<T>(), sizeof(T) * want,
                              hipMemcpyDeviceToHost, stream_));
      FPLS_HIP(hipStreamSynchronize(stream_));
      return;
    }
    this->inner_->export_field(field, out, elements);
  }

 private:
  std::vector<Buf*> extra_bufs() override {
    return {&xst_, &rnorm_, &colmean_, &grand_, &qnorm_, &qmean_, &kt_};
  }

  void transform(double* K, long long rows, const double* lnorm) {
    const double tol = std::is_same<T, float>::value ? 1e-5 : 1e-10;
    hipLaunchKernelGGL(k_kernel_transform, dim3(stream_blocks(rows * this->n_)), dim3(kWG), 0,
                       stream_, K, rows, (long long)this->n_, kind_, gamma_, degree_, coef0_,
                       lnorm, rnorm_.d(), tol);
    FPLS_LAUNCHED("k_kernel_transform");
  }

  // centred test-kernel rows (kernelpls.hpp:155-189) in the handle's precision
  const T* features(const void* x, int rows, bool device_rows) override {
    const int n = this->n_, p = this->p_;
    // src/cuda_resident_special.cuh:388-398: the test kernel must fit in 3/5
    // of the free device memory
    const double need = double(rows) * double(n) * sizeof(double) + 2.0 * rows * sizeof(double);
    if (need > double(free_device_bytes()) * 0.6)
      throw std::runtime_error(
          "nonlinear kernel prediction exceeds the guarded HIP workspace; reduce the prediction block size or training sample count");
    this->standardize_query(x, rows, device_rows);
    qnorm_.reserve(sizeof(double) * rows);
    qmean_.reserve(sizeof(double) * rows);
    kt_.reserve(sizeof(double) * size_t(rows) * n);
    hipLaunchKernelGGL(k_row_sumsq, dim3(unsigned((rows + kWG - 1) / kWG)), dim3(kWG), 0, stream_,
                       this->q64_.d(), (long long)rows, (long long)p, qnorm_.d());
    this->template product<double>(this->q64_.d(), rows, p, nullptr, xst_.d(), p, n, kt_.d());
    transform(kt_.d(), rows, qnorm_.d());
    hipLaunchKernelGGL(k_row_means, dim3(unsigned((rows + kWG - 1) / kWG)), dim3(kWG), 0, stream_,
                       kt_.d(), (long long)rows, (long long)n, qmean_.d());
    hipLaunchKernelGGL(k_kernel_centre, dim3(stream_blocks((long long)rows * n)), dim3(kWG), 0,
                       stream_, kt_.d(), (long long)rows, (long long)n, qmean_.d(), colmean_.d(),
                       grand_.d());
    FPLS_LAUNCHED("test kernel");
    return this->to_inner(kt_.d(), (long long)rows * n, this->qfeat_);
  }

  int A_, kind_, degree_;
  double gamma_, coef0_;
  Buf xst_, rnorm_, colmean_, grand_, qnorm_, qmean_, kt_;
};

}  // namespace

Handle* create_opls(const void* x, const void* y, const int* labels, int precision, int n, int p,
                    int q, int components, int scaling, int oversample, int power,
                    unsigned long long seed, int north, bool retain_scores) {
  require_device();
  if (precision == 32)
    return new OplsHandle<float>(static_cast<const float*>(x), static_cast<const float*>(y),
                                 labels, n, p, q, components, scaling, oversample, power, seed,
                                 north, retain_scores);
  if (precision == 64)
    return new OplsHandle<double>(static_cast<const double*>(x), static_cast<const double*>(y),
                                  labels, n, p, q, components, scaling, oversample, power, seed,
                                  north, retain_scores);
  throw std::invalid_argument("invalid resident HIP OPLS input");
}

Handle* create_kernelpls(const void* x, const void* y, const int* labels, int precision, int n,
                         int p, int q, int components, int scaling, int oversample, int power,
                         unsigned long long seed, int kind, double gamma, int degree,
                         double coef0, bool retain_scores) {
  require_device();
  if (precision == 32)
    return new KernelHandle<float>(static_cast<const float*>(x), static_cast<const float*>(y),
                                   labels, n, p, q, components, scaling, oversample, power, seed,
                                   kind, gamma, degree, coef0, retain_scores);
  if (precision == 64)
    return new KernelHandle<double>(static_cast<const double*>(x), static_cast<const double*>(y),
                                    labels, n, p, q, components, scaling, oversample, power,
                                    seed, kind, gamma, degree, coef0, retain_scores);
  throw std::invalid_argument("invalid resident HIP nonlinear kernel PLS input");
}

}  // namespace fpls
