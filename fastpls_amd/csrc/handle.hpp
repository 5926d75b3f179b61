// The resident model interface shared by the C ABI thunks and the CV drivers.
#pragma once

#include <cstddef>
#include <stdexcept>

#include <hip/hip_runtime.h>

namespace fpls {

// One fitted model with its device buffers and stream
// (src/cuda_resident_api.cuh:69-479). Row inputs are host arrays unless a
// call says they are on the device.
class Handle {
 public:
  virtual ~Handle() = default;
  // Responses (or LDA discriminants) per prefix, rows x q x count, in the
  // model's precision; `device_rows` reads x from device memory.
  virtual void predict_path(const void* x, int rows, const int* prefixes, int count, bool lda,
                            void* out, bool device_rows) = 0;
  // Top-`top` class indices (1-based) per prefix; `responses` (rows x q x
  // count, optional) receives the response scores. `device_rows` reads x
  // from device memory.
  virtual void classify_path(const void* x, int rows, const int* prefixes, int count, bool lda,
                             int top, int* labels, void* responses, bool device_rows) = 0;
  virtual void response_sums(const void* x, const void* y, const int* labels, int rows,
                             int prefix, void* out) = 0;
  // Held-out regression evaluation of device rows x (rows x p) against
  // device responses y (rows x q): per prefix, FP64 predictions (rows x q,
  // optional) and the 3 x q response sums of src/cuda_resident_metrics.cuh:9-43.
  virtual void heldout_regression(const void* x, const void* y, int rows, const int* prefixes,
                                  int count, double* predictions, double* sums) = 0;
  virtual void project(const void* x, int rows, int prefix, void* out) = 0;
  // Responses of one prefix for device rows x into device `out` (rows x q, the
  // model's precision) with every intermediate in the caller's `scratch`
  // (device, `capacity` doubles): no device memory is allocated.
  virtual void predict_scratch(const void* x, int rows, int prefix, void* out, double* scratch,
                               size_t capacity) = 0;
  virtual void export_field(int field, void* out, size_t elements) = 0;
  virtual void controls(int* oversample, int* power, int* block, int* block_limit,
                        int* implicit, int* cache) = 0;
  virtual void compact(bool prepare_lda) = 0;
  virtual double last_fit_seconds() const = 0;
  // Query products split the reduction over p by a rule that depends on p
  // only, so a row's result does not depend on the rows queried with it.
  virtual void set_batch_invariant(bool on) = 0;
  // Device bytes the handle holds.
  virtual size_t device_bytes() const = 0;
  // [0] fit, [1] cross-product pass over X, [2] score pass over X (PLS-SVD)
  // or the whole component loop (SIMPLS), [3] bytes of X, seconds from events.
  virtual void phases(double* out, int count) const = 0;
  // Runs later calls on `stream`, owned by the caller (it outlives the handle
  // or another attach): the current stream is drained first, and a stream
  // the handle took from the stream cache goes back to it.
  virtual void attach_stream(hipStream_t stream) = 0;
  // The FP64 weights R (features x components, column-major) in device
  // memory, valid while the handle lives; null when the handle keeps none.
  virtual const double* device_weights() const { return nullptr; }
  // `project` and `response_sums` for query rows x already in device memory
  // (y and labels stay host arrays).
  virtual void project_device(const void* x, int rows, int prefix, void* out) {
    (void)x; (void)rows; (void)prefix; (void)out;
    throw std::invalid_argument("this handle takes host query rows only");
  }
  virtual void response_sums_device(const void* x, const void* y, const int* labels, int rows,
                                    int prefix, void* out) {
    (void)x; (void)y; (void)labels; (void)rows; (void)prefix; (void)out;
    throw std::invalid_argument("this handle takes host query rows only");
  }
};

// method 0 = SIMPLS, 1 = PLS-SVD; precision 32 or 64. With `device_inputs`
// x, y and labels are device arrays that must outlive the handle. A non-null
// `stream` (owned by the caller) carries the fit instead of a stream from the
// stream cache. `retain_scores` false lets a large classification fit drop
// its training scores after the fit (the reference's rule,
// src/cuda_resident_simpls.cuh:213-216, cuda_resident_plssvd.cuh:166-169).
Handle* create(int method, const void* x, const void* y, const int* labels, int precision,
               bool device_inputs, int n, int p, int q, int components, int scaling,
               int oversample, int power, unsigned long long seed,
               hipStream_t stream = nullptr, bool retain_scores = true);

// Resident OPLS (src/cuda_resident_special.cuh:155-358): `north` orthogonal
// filter components, then SIMPLS on the filtered predictors. Host inputs.
Handle* create_opls(const void* x, const void* y, const int* labels, int precision, int n, int p,
                    int q, int components, int scaling, int oversample, int power,
                    unsigned long long seed, int north, bool retain_scores);
// Resident nonlinear kernel PLS (src/cuda_resident_special.cuh:360-544): kind
// 2 radial basis, 3 polynomial. Host inputs.
Handle* create_kernelpls(const void* x, const void* y, const int* labels, int precision, int n,
                         int p, int q, int components, int scaling, int oversample, int power,
                         unsigned long long seed, int kind, double gamma, int degree,
                         double coef0, bool retain_scores);

}  // namespace fpls
