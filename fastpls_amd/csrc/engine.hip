// Resident model handles behind the C ABI of include/fastpls_hip.h.
//
// Host orchestration of the HIP kernels: one non-blocking stream per handle,
// all matrices resident in HBM, one host synchronisation at the end of every
// ABI call. The handle semantics follow the reference's TypedResidentHandle
// (src/cuda_resident_api.cuh:69-479) and the fitting cores follow the
// reference CPU algorithms (inst/include/fastpls/core/simpls.hpp,
// rsvd.hpp, plssvd.hpp, lda.hpp) so that results agree with the FP64 CPU
// path for the same seed.
#include "../../include/fastpls_hip.h"
#include "handle.hpp"
#include "host_util.hpp"
#include "kernels.hpp"
#include "ops.hpp"
#include "stream.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace fpls {
namespace {

// Scratch of the no-allocation prediction entry: one array per GPU (each
// device holds its own copy until the process ends). Host code must never
// pass `scratch` itself to a kernel: it is a host-side stand-in; the device
// address comes from hipGetSymbolAddress (noalloc_scratch below).
constexpr size_t kNoAllocScratchDoubles = size_t(1) << 22;  // 32 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

template<class T>
class Model final : public Handle, private StreamOps {
 public:
  enum Method { kSimpls = 0, kPlssvd = 1 };

  Model(int method, const T* x, const T* y, const int* labels, bool device_inputs, int n,
        int p, int q, int components, int scaling, int oversample, int power,
        unsigned long long seed, hipStream_t stream = nullptr, bool retain_scores = true)
      : method_(method), n_(n), p_(p), q_(q), scaling_(scaling), oversample_(oversample),
        power_(power), seed_(seed), classification_(labels != nullptr) {
    if ((y == nullptr) == (labels == nullptr))
      throw std::invalid_argument("provide responses or labels, not both");
    if (x == nullptr) throw std::invalid_argument("null predictors");
    if (n < 2 || p < 1 || q < 1 || components < 1)
      throw std::invalid_argument("invalid resident dimensions");
    if (scaling < 1 || scaling > 3) throw std::invalid_argument("scaling must be 1, 2 or 3");
    if (oversample < 0 || power < 0) throw std::invalid_argument("oversample and power must be non-negative");
    if (classification_ && q < 2) throw std::invalid_argument("classification needs at least two classes");
    if (method_ == kSimpls) {
      if (components > std::min(n - 1, p)) throw std::invalid_argument("invalid resident SIMPLS dimensions");
    } else {
      if (components > std::min(n - 1, std::min(p, q)))
        throw std::invalid_argument("invalid resident PLS-SVD dimensions");
      if (classification_ && components >= q)
        throw std::invalid_argument("classification PLS-SVD components must be below class count");
    }
    A_ = components;
    // training scores are dropped after the fit only for large
    // classification fits (src/cuda_resident_simpls.cuh:213-216,
    // src/cuda_resident_plssvd.cuh:166-169): they are needed by nothing else
    // once the LDA moments are formed
    {
      const double score_bytes = double(n) * double(components) * double(sizeof(T));
      const bool small = score_bytes <= 256.0 * 1024.0 * 1024.0 || n < 8 * p;
      keep_scores_ = retain_scores || !classification_ || small ||
                     (method_ == kSimpls ? components < 8 : 2 * components < p);
    }
    // src/r_api.cpp:1165-1176: SIMPLS takes its loadings from a cached X_sᵀX_s
    // when A >= 20, p <= n, n >= 8p and p <= 512 (the CPU core's rule)
    cache_gram_ = method_ == kSimpls && components >= 20 && p <= n && n >= 8 * p && p <= 512;
    // R/main.R:3454-3463: no cross-covariance is formed once p·q·8 bytes
    // exceed 32 MiB (dense responses); above 512 MiB of entries SIMPLS takes
    // the rank-one direction route (src/r_api.cpp:1480-1493).
    implicit_ = !classification_ && double(p) * double(q) * 8.0 / (1024.0 * 1024.0) > 32.0;
    rank_one_ = implicit_ && method_ == kSimpls &&
                double(p) * double(q) * double(sizeof(T)) > 512.0 * 1024.0 * 1024.0;
    cache_gram_ = cache_gram_ && !implicit_;
    // SIMPLS amplifies the rounding of the products over X (the later
    // directions of its component recursion move by up to ~1e3 times a
    // relative perturbation of S), so FP32-stored X is widened to FP64 for
    // every SIMPLS product; PLS-SVD keeps bf16x6 (FP32-accurate) products.
    f64_products_ = method_ == kSimpls;
    owns_stream_ = stream == nullptr;
    stream_ = owns_stream_ ? stream_take() : stream;
    for (Buf* b : all_bufs()) b->bind(stream_);
    try {
      Event start, stop;
      load_inputs(x, y, labels, device_inputs);
      if (method_ == kPlssvd && !implicit_) stage_plssvd_sketch();
      FPLS_HIP(hipEventRecord(start.e, stream_));
      prepare_crossprod();
      if (method_ == kSimpls) fit_simpls(); else fit_plssvd();
      finish_fit();
      if (!P_.empty()) {  // the re-orthogonalisation's loadings serve the fit only
        FPLS_HIP(hipStreamSynchronize(stream_));
        P_.release();
      }
      if (!keep_scores_) {
        ensure_lda_moments();  // classification only: its moments come from the scores
        FPLS_HIP(hipStreamSynchronize(stream_));
        Tsc_.release();
      }
      FPLS_HIP(hipEventRecord(stop.e, stream_));
      check_state();
      float ms = 0.f;
      FPLS_HIP(hipEventElapsedTime(&ms, start.e, stop.e));
      fit_seconds_ = double(ms) * 1e-3;
      FPLS_HIP(hipEventElapsedTime(&ms, cp0_.e, cp1_.e));
      crossprod_seconds_ = double(ms) * 1e-3;
      FPLS_HIP(hipEventElapsedTime(&ms, sc0_.e, sc1_.e));
      score_seconds_ = double(ms) * 1e-3;
    } catch (...) {
      close_stream();
      throw;
    }
  }

  ~Model() override { close_stream(); }

  double last_fit_seconds() const override { return fit_seconds_; }

  void set_batch_invariant(bool on) override { invariant_ = on; }

  size_t device_bytes() const override {
    size_t total = 0;
    for (const Buf* b : const_cast<Model*>(this)->all_bufs()) total += b->bytes();
    return total;
  }

  void phases(double* out, int count) const override {
    // the last query's projection: its events were passed before the query
    // returned (every query synchronises its stream)
    float ms = 0.f;
    if (proj_timed_) FPLS_HIP(hipEventElapsedTime(&ms, pr0_.e, pr1_.e));
    const double v[6] = {fit_seconds_, crossprod_seconds_, score_seconds_,
                         double(n_) * double(p_) * double(sizeof(T)), double(ms) * 1e-3,
                         proj_bytes_};
    for (int i = 0; i < count && i < 6; ++i) out[i] = v[i];
  }

  // ------------------------------------------------------------ queries
  void predict_path(const void* x, int rows, const int* prefixes, int count, bool lda,
                    void* out, bool device_rows) override {
    if (!x || !out || rows < 1 || count < 1) throw std::invalid_argument("invalid resident prediction-path request");
    check_path(prefixes, count);
    if (lda && !classification_) throw std::invalid_argument("LDA requires class labels");
    const int maxp = prefixes[count - 1];
    const T* xs = static_cast<const T*>(x);
    T* o = static_cast<T*>(out);
    const int block = device_rows ? rows : query_block_rows(rows, query_row_bytes(maxp, 0));
    for (int start = 0; start < rows; start += block) {
      const int cnt = std::min(block, rows - start);
      project_rows(device_rows ? xs : load_row_block(xs, rows, start, cnt), cnt, maxp);
      for (int j = 0; j < count; ++j)
        response_block(cnt, prefixes[j], lda, o + (size_t)j * rows * q_out() + start, rows);
    }
    FPLS_HIP(hipStreamSynchronize(stream_));
  }

  void classify_path(const void* x, int rows, const int* prefixes, int count, bool lda,
                     int top, int* labels, void* responses, bool device_rows) override {
    if (!classification_) throw std::invalid_argument("invalid resident classification request");
    if (!x || !labels || rows < 1 || count < 1 || top < 1 || top > q_)
      throw std::invalid_argument("invalid resident classification-path request");
    check_path(prefixes, count);
    const int maxp = prefixes[count - 1];
    const T* xs = static_cast<const T*>(x);
    const int block = device_rows ? rows : query_block_rows(rows, query_row_bytes(maxp, top));
    for (int start = 0; start < rows; start += block) {
      const int cnt = std::min(block, rows - start);
      project_rows(device_rows ? xs : load_row_block(xs, rows, start, cnt), cnt, maxp);
      top_.reserve(sizeof(int) * (size_t)cnt * top);
      for (int j = 0; j < count; ++j) {
        T* resp = responses ? static_cast<T*>(responses) + (size_t)j * rows * q_ + start : nullptr;
        if (resp) response_block(cnt, prefixes[j], false, resp, rows);
        if (lda) {
          // discriminants already include the class constants
          response_block(cnt, prefixes[j], true, nullptr, rows);
        } else if (!resp) {
          response_block(cnt, prefixes[j], false, nullptr, rows);
        }
        launch_topk(pred64_.d(), cnt, q_, nullptr, top, top_.as<int>(), stream_);
        // labels of the block: cnt x top into the rows x top output
        FPLS_HIP(hipMemcpy2DAsync(labels + (size_t)j * rows * top + start, sizeof(int) * (size_t)rows,
                                  top_.as<int>(), sizeof(int) * (size_t)cnt,
                                  sizeof(int) * (size_t)cnt, top, hipMemcpyDeviceToHost, stream_));
        // top_ is reused by the next prefix and block
        FPLS_HIP(hipStreamSynchronize(stream_));
      }
    }
    FPLS_HIP(hipStreamSynchronize(stream_));
  }

  void response_sums(const void* x, const void* y, const int* labels, int rows, int prefix,
                     void* out) override {
    response_sums_rows(x, y, labels, rows, prefix, out, false);
  }

  void response_sums_device(const void* x, const void* y, const int* labels, int rows,
                            int prefix, void* out) override {
    response_sums_rows(x, y, labels, rows, prefix, out, true);
  }

  void response_sums_rows(const void* x, const void* y, const int* labels, int rows, int prefix,
                          void* out, bool device_rows) {
    if (!out || (!y && !labels) || (y && labels) || (classification_ != (labels != nullptr)))
      throw std::invalid_argument("invalid resident response representation");
    if (!x || rows < 1 || prefix < 1 || prefix > A_) throw std::invalid_argument("invalid resident prediction input");
    project_rows(load_rows(static_cast<const T*>(x), rows, device_rows), rows, prefix);
    response_block(rows, prefix, false, nullptr, 0);
    metric_.reserve(sizeof(double) * 3 * q_);
    flags_.reserve(sizeof(int));
    fill_zero_int(flags_.as<int>(), 1);
    const T* obs = nullptr;
    const int* lab = nullptr;
    if (labels) {
      obs_labels_.reserve(sizeof(int) * (size_t)rows);
      FPLS_HIP(hipMemcpyAsync(obs_labels_.as<int>(), labels, sizeof(int) * (size_t)rows,
                              hipMemcpyHostToDevice, stream_));
      lab = obs_labels_.as<int>();
    } else {
      obs_.reserve(sizeof(T) * (size_t)rows * q_);
      FPLS_HIP(hipMemcpyAsync(obs_.as<T>(), y, sizeof(T) * (size_t)rows * q_,
                              hipMemcpyHostToDevice, stream_));
      obs = obs_.as<T>();
    }
    launch_response_sums<T>(pred64_.d(), obs, lab, my_.d(), rows, q_, metric_.d(),
                            flags_.as<int>(), stream_);
    std::vector<double> host(3 * (size_t)q_);
    int bad = 0;
    FPLS_HIP(hipMemcpyAsync(host.data(), metric_.d(), sizeof(double) * 3 * q_, hipMemcpyDeviceToHost, stream_));
    FPLS_HIP(hipMemcpyAsync(&bad, flags_.as<int>(), sizeof(int), hipMemcpyDeviceToHost, stream_));
    FPLS_HIP(hipStreamSynchronize(stream_));
    if (bad) throw std::runtime_error("nonfinite response/prediction or invalid class label in resident metrics");
    T* o = static_cast<T*>(out);
    for (size_t i = 0; i < host.size(); ++i) o[i] = T(host[i]);
  }

  void heldout_regression(const void* x, const void* y, int rows, const int* prefixes, int count,
                          double* predictions, double* sums) override {
    if (classification_) throw std::invalid_argument("held-out regression needs a response model");
    if (!x || !y || !sums || rows < 1) throw std::invalid_argument("invalid held-out regression request");
    check_path(prefixes, count);
    project_rows(static_cast<const T*>(x), rows, prefixes[count - 1]);
    metric_.reserve(sizeof(double) * 3 * q_);
    flags_.reserve(sizeof(int));
    for (int j = 0; j < count; ++j) {
      response_block(rows, prefixes[j], false, nullptr, 0);
      if (predictions)
        FPLS_HIP(hipMemcpyAsync(predictions + (size_t)j * rows * q_, pred64_.d(),
                                sizeof(double) * (size_t)rows * q_, hipMemcpyDeviceToHost, stream_));
      launch_fill_int(flags_.as<int>(), 1, 0, stream_);
      launch_response_sums<T>(pred64_.d(), static_cast<const T*>(y), nullptr, my_.d(), rows, q_,
                              metric_.d(), flags_.as<int>(), stream_);
      FPLS_HIP(hipMemcpyAsync(sums + (size_t)j * 3 * q_, metric_.d(), sizeof(double) * 3 * q_,
                              hipMemcpyDeviceToHost, stream_));
    }
    FPLS_HIP(hipStreamSynchronize(stream_));
  }

  void predict_scratch(const void* x, int rows, int prefix, void* out, double* scratch,
                       size_t capacity) override {
    if (!x || !out || !scratch || rows < 1 || prefix < 1 || prefix > A_)
      throw std::invalid_argument("invalid resident prediction request");
    const bool f64 = f64_products_;
    const int nb = std::min(prefix, 64);
    const int splits = invariant_ ? xw_splits_by_depth(p_)
                                  : xw_splits(rows, p_, nb, sizeof(T) == 4 && !f64);
    // layout (doubles): scores rows x prefix | FP64 predictions rows x q |
    // split slabs of one 64-column chunk | operand workspace
    const size_t r = size_t(rows);
    const size_t n_scores = r * prefix, n_pred = r * q_;
    const size_t n_slabs = splits > 1 ? size_t(splits) * r * nb : 0;
    const size_t n_frag = (xw_frag_bytes<T>(p_, nb, f64) + 255) / 256 * 32;
    const size_t need = n_scores + n_pred + n_slabs + n_frag;
    if (need > capacity)
      throw std::invalid_argument("prediction needs " + std::to_string(need) +
                                  " scratch doubles; the no-allocation entry holds " +
                                  std::to_string(capacity));
    double* scores = scratch;
    double* pred = scores + n_scores;
    double* slabs = pred + n_pred;
    void* fragp = n_frag ? static_cast<void*>(slabs + n_slabs) : nullptr;
    const T* xr = static_cast<const T*>(x);
    for (int c0 = 0; c0 < prefix; c0 += 64) {
      const int w = std::min(64, prefix - c0);
      XwTask<T> t{};
      t.A = xr; t.m = rows; t.k = p_; t.center = mx_.d(); t.W = Rs_.d() + (size_t)c0 * p_;
      t.ldw = p_; t.N = w; t.splits = splits; t.frag = fragp; t.ldo = rows;
      if (splits == 1) {
        t.out = scores + (size_t)c0 * r; t.slab = 0;
        launch_xw_single<T>(t, stream_, f64);
      } else {
        t.out = slabs; t.slab = (long long)r * w;
        launch_xw_single<T>(t, stream_, f64);
        launch_sum_slabs(slabs, splits, (long long)r * w, (long long)r * w, scores + (size_t)c0 * r,
                         stream_);
      }
    }
    if (method_ == kSimpls) {
      launch_apply_small<T>(scores, rows, prefix, Q_.d(), q_, 1, q_, my_.d(), nullptr, pred, stream_);
    } else {
      const double* Wk = W_.d() + (size_t)(prefix - 1) * A_ * q_;
      launch_apply_small<T>(scores, rows, prefix, Wk, 1, prefix, q_, my_.d(), nullptr, pred,
                            stream_);
    }
    launch_convert<T>(pred, static_cast<T*>(out), (long long)r * q_, stream_);
    FPLS_HIP(hipStreamSynchronize(stream_));
  }

  void project(const void* x, int rows, int prefix, void* out) override {
    project_any(x, rows, prefix, out, false);
  }

  void project_device(const void* x, int rows, int prefix, void* out) override {
    project_any(x, rows, prefix, out, true);
  }

  const double* device_weights() const override { return R_.d(); }

  void project_any(const void* x, int rows, int prefix, void* out, bool device_rows) {
    if (!x || !out || rows < 1 || prefix < 1 || prefix > A_) throw std::invalid_argument("invalid resident score request");
    project_rows(load_rows(static_cast<const T*>(x), rows, device_rows), rows, prefix);
    outT_.reserve(sizeof(T) * (size_t)rows * prefix);
    launch_convert<T>(scores_.d(), outT_.as<T>(), (long long)rows * prefix, stream_);
    FPLS_HIP(hipMemcpyAsync(out, outT_.as<T>(), sizeof(T) * (size_t)rows * prefix,
                            hipMemcpyDeviceToHost, stream_));
    FPLS_HIP(hipStreamSynchronize(stream_));
  }

  void export_field(int field, void* out, size_t elements) override {
    const size_t counts[] = {(size_t)p_ * A_, (size_t)q_ * A_, (size_t)n_ * A_, (size_t)p_,
                             (size_t)p_, (size_t)q_, (size_t)p_ * A_, (size_t)A_ + 1};
    if (field < 0 || field > 7 || !out || elements != counts[field])
      throw std::invalid_argument("invalid resident field output size");
    const double* src = nullptr;
    switch (field) {
      case 0: src = R_.d(); break;
      case 1: src = Q_.d(); break;
      case 2:
        if (Tsc_.empty())
          throw std::runtime_error(compacted_ ? "training scores were released by compact"
                                              : "training scores were not retained");
        src = Tsc_.d();
        break;
      case 3: src = mx_.d(); break;
      case 4: src = sx_.d(); break;
      case 5: src = my_.d(); break;
      default:
        ensure_variance();
        src = field == 6 ? var_loadings_.d() : var_ss_.d();
    }
    outT_.reserve(sizeof(T) * elements);
    launch_convert<T>(src, outT_.as<T>(), (long long)elements, stream_);
    FPLS_HIP(hipMemcpyAsync(out, outT_.as<T>(), sizeof(T) * elements, hipMemcpyDeviceToHost, stream_));
    FPLS_HIP(hipStreamSynchronize(stream_));
  }

  void controls(int* oversample, int* power, int* block, int* block_limit, int* implicit,
                int* cache) override {
    if (!oversample || !power || !block || !block_limit || !implicit || !cache)
      throw std::invalid_argument("null resident control output");
    *oversample = oversample_;
    *power = power_;
    *block = method_ == kSimpls ? block_ : A_;
    *block_limit = 0;
    *implicit = implicit_ ? 1 : 0;
    *cache = cache_gram_ ? 1 : 0;
  }

  void compact(bool prepare_lda) override {
    if (prepare_lda) {
      if (!classification_) throw std::invalid_argument("LDA compaction requires classification labels");
      ensure_lda_moments();
    }
    FPLS_HIP(hipStreamSynchronize(stream_));
    for (Buf* b : {&xown_, &yown_, &Yc_, &S_, &S0_, &G_, &V_, &Tsc_, &P_, &xgram_, &simpls_scratch_, &cand_,
                   &cand_scaled_, &t_raw_, &ts_part_, &ls_part_, &omega_, &steps_, &fitdev_,
                   &part_, &tmp1_, &tmp2_, &tmp3_, &H_, &labels_dev_, &frag_, &cp_part_,
                   &cp_stats_, &qr_gram_, &qr_rinv_, &qr_tmp_, &qr_fail_, &qraw_, &op_sig_,
                   &op_tmp_, &op_scaled_, &op_coef_, &op_corr_, &op_pb_, &op_rev_, &op_sample_})
      b->release();
    for (Buf& b : svd_ws_) b.release();
    for (Buf& b : wide_ws_) b.release();
    for (Buf& b : opw_) b.release();
    X_ = nullptr;
    compacted_ = true;
  }

 private:
  int q_out() const { return q_; }

  std::vector<Buf*> all_bufs() {
    std::vector<Buf*> v = {&xown_, &yown_, &labels_dev_, &counts_, &Yc_, &H_, &ysum_, &my_, &mx_,
                           &sx_, &colss_, &S_, &S0_, &G_, &R_, &Q_, &V_, &Tsc_, &P_, &xgram_, &Rs_, &sigma_,
                           &gram_, &W_, &state_, &omega_, &steps_, &fitdev_, &simpls_scratch_,
                           &cand_, &cand_scaled_, &t_raw_, &ts_part_, &ls_part_, &tmp1_,
                           &tmp2_, &tmp3_, &cp_part_, &cp_stats_, &usable_, &rows_, &scores_,
                           &pred64_, &outT_, &top_, &metric_, &flags_, &obs_, &obs_labels_,
                           &lda_gram_, &lda_sums_, &lda_linear_, &lda_constants_, &lda_ridge_,
                           &lda_scratch_, &var_loadings_, &var_ss_, &qraw_, &op_sig_,
                           &op_tmp_, &op_scaled_, &op_coef_, &op_corr_, &op_pb_, &op_rev_,
                           &op_sample_};
    for (Buf* b : ops_bufs()) v.push_back(b);
    for (Buf& b : svd_ws_) v.push_back(&b);
    for (Buf& b : wide_ws_) v.push_back(&b);
    return v;
  }

  // Gives every device block back in stream order, then the drained stream
  // back to the stream cache.
  void close_stream() noexcept {
    if (!stream_) return;
    for (Buf* b : all_bufs()) b->release();
    if (owns_stream_) stream_give(stream_);
    else (void)hipStreamSynchronize(stream_);
    stream_ = nullptr;
  }

 public:
  void attach_stream(hipStream_t stream) override {
    if (!stream || stream == stream_) return;
    FPLS_HIP(hipStreamSynchronize(stream_));
    for (Buf* b : all_bufs()) b->rebind(stream);
    if (owns_stream_) stream_give(stream_);
    stream_ = stream;
    owns_stream_ = false;
  }

 private:

  // --------------------------------------------------------- data load
  void load_inputs(const T* x, const T* y, const int* labels, bool device_inputs) {
    const size_t np = (size_t)n_ * p_;
    if (device_inputs) {
      X_ = x;
    } else {
      xown_.alloc(sizeof(T) * np);
      FPLS_HIP(hipMemcpyAsync(xown_.as<T>(), x, sizeof(T) * np, hipMemcpyHostToDevice, stream_));
      X_ = xown_.as<T>();
    }
    Yc_.alloc(sizeof(double) * (size_t)n_ * q_);
    ysum_.alloc(sizeof(double) * q_);
    my_.alloc(sizeof(double) * q_);
    state_.alloc(sizeof(int) * 4);
    fill_zero_int(state_.as<int>(), 4);
    if (labels) {
      const int* dl = labels;
      if (!device_inputs) {
        labels_dev_.alloc(sizeof(int) * (size_t)n_);
        FPLS_HIP(hipMemcpyAsync(labels_dev_.as<int>(), labels, sizeof(int) * (size_t)n_,
                                hipMemcpyHostToDevice, stream_));
        dl = labels_dev_.as<int>();
      }
      counts_.alloc(sizeof(int) * q_);
      H_.alloc(sizeof(double) * (size_t)n_ * q_);
      launch_label_response(dl, n_, q_, counts_.as<int>(), my_.d(), Yc_.d(), H_.d(), ysum_.d(),
                            state_.as<int>() + 2, stream_);
    } else {
      const T* dy = y;
      if (!device_inputs) {
        yown_.alloc(sizeof(T) * (size_t)n_ * q_);
        FPLS_HIP(hipMemcpyAsync(yown_.as<T>(), y, sizeof(T) * (size_t)n_ * q_,
                                hipMemcpyHostToDevice, stream_));
        dy = yown_.as<T>();
      }
      part_.reserve(sizeof(double) * (size_t)kCenterScratch * q_);
      launch_center_response<T>(dy, n_, q_, my_.d(), Yc_.d(), ysum_.d(), part_.d(), stream_);
    }
  }

  // ----------------------------------------------------- preprocessing
  // supervised.hpp:27-124 / classification.hpp:86-184 in one pass over X.
  // On the matrix-free route only the column statistics are kept: the pass
  // runs against the first response column and its one-column product is
  // dropped.
  void prepare_crossprod() {
    const int qc = implicit_ ? 1 : q_;
    mx_.alloc(sizeof(double) * p_);
    sx_.alloc(sizeof(double) * p_);
    colss_.alloc(sizeof(double) * p_);
    S_.alloc(sizeof(double) * (size_t)p_ * qc);
    S0_.alloc(sizeof(double) * (size_t)p_ * qc);
    const bool f64 = f64_products_;
    const int splits = xtw_splits(n_, p_, std::min(qc, 64), sizeof(T) == 4 && !f64);
    Buf& s_part = cp_part_;
    Buf& stats = cp_stats_;
    s_part.alloc(sizeof(double) * (size_t)splits * p_ * qc);
    stats.alloc(sizeof(double) * (size_t)splits * 2 * p_);
    void* wf = frag(xtw_frag_bytes<T>(n_, std::min(qc, 64), f64));
    FPLS_HIP(hipEventRecord(cp0_.e, stream_));
    for (int c0 = 0; c0 < qc; c0 += 64) {
      const int nb = std::min(64, qc - c0);
      XtwTask<T> t{};
      t.A = X_; t.m = n_; t.k = p_; t.center = nullptr;
      t.W = Yc_.d() + (long long)c0 * n_; t.ldw = n_; t.N = nb; t.mode = kShiftStats;
      t.frag = wf;
      t.splits = splits;
      t.out = s_part.d() + (long long)c0 * p_; t.ldo = p_; t.slab = (long long)p_ * qc;
      t.stats = c0 == 0 ? stats.d() : nullptr;
      launch_xtw_single<T>(t, stream_, f64);
    }
    FPLS_HIP(hipEventRecord(cp1_.e, stream_));
    launch_finalize_crossprod<T>(X_, n_, p_, qc, scaling_, stats.d(), s_part.d(), splits,
                                 ysum_.d(), mx_.d(), sx_.d(), colss_.d(), S_.d(), S0_.d(),
                                 state_.as<int>() + 2, stream_);
    if (implicit_) {
      S_.release();
      S0_.release();
    }
  }

  // ------------------------------------------------- matrix-free route
  // operators.hpp:58-156 (CenteredCrosscovOperator): S Z = X_sᵀ (Ỹ Z) for Z
  // (q x w) and Sᵀ B = Ỹᵀ (X_s B) for B (p x w), with X_s = (X - 1 mᵀ)/s
  // applied through the centred products and a row scaling.
  void crosscov_fwd(const double* Z, int w, double* out) {
    op_tmp_.reserve(sizeof(double) * (size_t)n_ * w);
    product<double>(Yc_.d(), n_, q_, nullptr, Z, q_, w, op_tmp_.d());
    tproduct<T>(X_, n_, p_, mx_.d(), op_tmp_.d(), n_, w, out);
    launch_div_rows(out, sx_.d(), p_, w, out, stream_);
  }

  void crosscov_adj(const double* B, int w, double* out) {
    op_scaled_.reserve(sizeof(double) * (size_t)p_ * w);
    launch_div_rows(B, sx_.d(), p_, w, op_scaled_.d(), stream_);
    op_tmp_.reserve(sizeof(double) * (size_t)n_ * w);
    product<T>(X_, n_, p_, mx_.d(), op_scaled_.d(), p_, w, op_tmp_.d());
    tproduct<double>(Yc_.d(), n_, q_, nullptr, op_tmp_.d(), n_, w, out);
  }

  // M (p x w) -= V (Vᵀ M) over the first `used` deflation directions (one
  // classical Gram-Schmidt pass, operators.hpp:158-301).
  void project_out(double* M, int w, int used) {
    if (used <= 0) return;
    op_coef_.reserve(sizeof(double) * (size_t)used * w);
    op_corr_.reserve(sizeof(double) * (size_t)p_ * w);
    tproduct<double>(V_.d(), p_, used, nullptr, M, p_, w, op_coef_.d());
    product<double>(V_.d(), p_, used, nullptr, op_coef_.d(), used, w, op_corr_.d());
    launch_axpy(-1.0, op_corr_.d(), M, (long long)p_ * w, stream_);
  }

  // (I - V Vᵀ) S Z and Sᵀ (I - V Vᵀ) B (ProjectedOperator::multiply).
  void projected_fwd(const double* Z, int w, double* out, int used) {
    crosscov_fwd(Z, w, out);
    project_out(out, w, used);
  }

  void projected_adj(const double* B, int w, double* out, int used) {
    op_pb_.reserve(sizeof(double) * (size_t)p_ * w);
    FPLS_HIP(hipMemcpyAsync(op_pb_.d(), B, sizeof(double) * (size_t)p_ * w,
                            hipMemcpyDeviceToDevice, stream_));
    project_out(op_pb_.d(), w, used);
    crosscov_adj(op_pb_.d(), w, out);
  }

  // operator_rsvd.hpp:175-202 (randomized_dominant_operator_direction): a
  // p-vector from the host stream of `seed`, stabilised (two projections),
  // normalised, then max(power, 1) rounds of Sᵀ, normalise, S, stabilise,
  // normalise. `normals` holds the draw; a direction that cannot be
  // normalised stops the fit (state[1]).
  void dominant_direction(const double* normals, int used, double* dir) {
    int* stop = state_.as<int>() + 1;
    op_rev_.reserve(sizeof(double) * (size_t)q_);
    op_sample_.reserve(sizeof(double) * (size_t)p_);
    FPLS_HIP(hipMemcpyAsync(dir, normals, sizeof(double) * (size_t)p_, hipMemcpyDeviceToDevice,
                            stream_));
    project_out(dir, 1, used);
    project_out(dir, 1, used);
    launch_unit_columns(dir, p_, 1, stop, stream_);
    for (int it = 0; it < std::max(power_, 1); ++it) {
      projected_adj(dir, 1, op_rev_.d(), used);
      launch_unit_columns(op_rev_.d(), q_, 1, stop, stream_);
      projected_fwd(op_rev_.d(), 1, op_sample_.d(), used);
      project_out(op_sample_.d(), 1, used);
      project_out(op_sample_.d(), 1, used);
      launch_unit_columns(op_sample_.d(), p_, 1, stop, stream_);
      FPLS_HIP(hipMemcpyAsync(dir, op_sample_.d(), sizeof(double) * (size_t)p_,
                              hipMemcpyDeviceToDevice, stream_));
    }
  }

  // ------------------------------------------------------------ SIMPLS
  // simpls.hpp:480-763 with the routing of src/r_api.cpp:1149-1190.
  void fit_simpls() {
    const int maximum = std::min({A_, p_, std::max(n_ - 1, 1)});
    A_ = maximum;
    // the rank-one route refreshes one direction at a time
    block_ = rank_one_ ? 1 : candidate_block_size(maximum, p_, q_, classification_, n_, 64);
    const bool right_gram = !implicit_ && q_ <= p_ && q_ <= 512;
    const int maxdim = std::min(p_, q_);
    std::vector<RefreshStep> steps;
    std::vector<double> pool;
    int wmax = 1, bmax = 1;
    for (int a = 0; a < maximum;) {
      const int b = std::min({std::max(block_, 1), maximum - a, maxdim});
      const int w = std::min(maxdim, b + oversample_);
      RefreshStep st{};
      st.a = a; st.b = b; st.w = w; st.omega_off = (long long)pool.size();
      // sketches: q x w for the explicit refresh; the rank-one route draws
      // a p-vector per step; the operator rSVD draws its own (same stream)
      const size_t count = !implicit_ ? (size_t)q_ * w : rank_one_ ? (size_t)p_ : 0;
      const size_t base = pool.size();
      pool.resize(base + count);
      if (count) gaussian(unsigned(seed_) + unsigned(a), count, pool.data() + base);
      steps.push_back(st);
      wmax = std::max(wmax, w);
      bmax = std::max(bmax, b);
      a += b;
    }
    if (wmax > 512) throw std::invalid_argument("sketch width above 512 is not supported");
    R_.alloc(sizeof(double) * (size_t)p_ * A_);
    Q_.alloc(sizeof(double) * (size_t)q_ * A_);
    V_.alloc(sizeof(double) * (size_t)p_ * A_);
    Tsc_.alloc(sizeof(double) * (size_t)n_ * A_);
    if (right_gram) {
      G_.alloc(sizeof(double) * (size_t)q_ * q_);
      tproduct<double>(S_.d(), p_, q_, nullptr, S_.d(), p_, q_, G_.d());
    }
    // the cached route's X_sᵀX_s (simpls.hpp:521-532): columns of X_s in FP64,
    // 64 at a time, against the centred X, rows then divided by the scales
    if (cache_gram_) {
      xgram_.alloc(sizeof(double) * (size_t)p_ * p_);
      Buf& chunk = tmp3_;
      chunk.reserve(sizeof(double) * (size_t)n_ * std::min(64, p_));
      for (int c0 = 0; c0 < p_; c0 += 64) {
        const int w = std::min(64, p_ - c0);
        launch_standardize_columns<T>(X_, n_, c0, w, mx_.d(), sx_.d(), chunk.d(), stream_);
        double* out = xgram_.d() + (size_t)c0 * p_;
        tproduct<T>(X_, n_, p_, mx_.d(), chunk.d(), n_, w, out);
        launch_div_rows(out, sx_.d(), p_, w, out, stream_);
      }
    }
    // host copies are members: they outlive the asynchronous uploads
    omega_host_ = std::move(pool);
    steps_host_ = std::move(steps);
    omega_.alloc(sizeof(double) * omega_host_.size());
    if (!omega_host_.empty())
      FPLS_HIP(hipMemcpyAsync(omega_.d(), omega_host_.data(),
                              sizeof(double) * omega_host_.size(), hipMemcpyHostToDevice,
                              stream_));
    steps_.alloc(sizeof(RefreshStep) * steps_host_.size());
    FPLS_HIP(hipMemcpyAsync(steps_.as<RefreshStep>(), steps_host_.data(),
                            sizeof(RefreshStep) * steps_host_.size(), hipMemcpyHostToDevice, stream_));
    const std::vector<RefreshStep>& steps_ref = steps_host_;
    simpls_scratch_.alloc(sizeof(double) * (size_t)simpls_scratch_doubles(p_, q_, wmax));
    cand_.alloc(sizeof(double) * (size_t)p_ * bmax);
    cand_scaled_.alloc(sizeof(double) * (size_t)p_ * bmax);
    t_raw_.alloc(sizeof(double) * (size_t)n_ * bmax);
    const bool f64 = f64_products_;
    const int xs = xw_splits(n_, p_, bmax, sizeof(T) == 4 && !f64);
    const int ls = xtw_splits(n_, p_, bmax, sizeof(T) == 4 && !f64);
    // Fits with many predictors run their refresh and update as several
    // launches whose products use the whole device: one work-group walking
    // p-long vectors and p x q matrices streams at tens of GB/s (both took
    // 1.4 ms per component at p = 20000, as long as a pass over an 8 GB X;
    // the whole 1e5 x 2e4 x 25 fit 145 -> 85 ms).
    // FASTPLS_HIP_WIDE_P, when set, replaces the p threshold.
    {
      long long threshold = kWideP;
      if (const char* v = std::getenv("FASTPLS_HIP_WIDE_P")) threshold = std::atoll(v);
      wide_ = !implicit_ && !rank_one_ && !cache_gram_ && p_ >= threshold;
    }
    ts_part_.alloc(sizeof(double) * (size_t)xs * n_ * bmax);
    ls_part_.alloc(sizeof(double) * (size_t)ls * p_ * bmax);

    SimplsFitDev f{};
    f.n = n_; f.p = p_; f.q = q_; f.A = A_;
    f.power = power_; f.use_right_gram = right_gram ? 1 : 0;
    // the rank-one route re-orthogonalises (src/r_api.cpp:1491): scores and
    // weights against the earlier scores, loadings twice; the reference's
    // float core does too (src/r_api.cpp:1416) to offset float arithmetic,
    // which this engine's FP32-stored fits do not have: their products and
    // small math are FP64, so both precisions take the FP64 routing (one
    // pass) and give the same fit on the same values
    f.reorth = rank_one_ ? 1 : 0;
    f.nsteps = int(steps_ref.size()); f.bmax = bmax; f.ts_splits = xs; f.ls_splits = ls;
    f.ls_slab = (long long)p_ * bmax;
    f.steps = steps_.as<RefreshStep>(); f.omega = omega_.d();
    f.S = implicit_ ? nullptr : S_.d(); f.S0 = implicit_ ? nullptr : S0_.d();
    f.G = right_gram ? G_.d() : nullptr; f.sx = sx_.d();
    if (implicit_) {
      qraw_.alloc(sizeof(double) * (size_t)q_ * bmax);
      op_sig_.alloc(sizeof(double) * bmax);
      f.qraw = qraw_.d();
    }
    f.cand = cand_.d(); f.cand_scaled = cand_scaled_.d(); f.t_raw = t_raw_.d();
    f.ts_part = ts_part_.d(); f.ls_part = ls_part_.d();
    f.R = R_.d(); f.Q = Q_.d(); f.V = V_.d(); f.T = Tsc_.d();
    f.xgram = cache_gram_ ? xgram_.d() : nullptr;
    if (f.reorth) P_.alloc(sizeof(double) * (size_t)p_ * A_);
    f.P = f.reorth ? P_.d() : nullptr;
    f.scratch = simpls_scratch_.d(); f.state = state_.as<int>();
    fitdev_host_ = f;
    fitdev_.alloc(sizeof(SimplsFitDev));
    FPLS_HIP(hipMemcpyAsync(fitdev_.as<SimplsFitDev>(), &fitdev_host_, sizeof(f), hipMemcpyHostToDevice,
                            stream_));

    FPLS_HIP(hipEventRecord(sc0_.e, stream_));
    for (size_t s = 0; s < steps_ref.size(); ++s) {
      const int b = steps_ref[s].b;
      if (wide_) {
        wide_refresh(steps_ref[s], right_gram);
      } else if (!implicit_) {
        launch_simpls_refresh(fitdev_.as<SimplsFitDev>(), 1, int(s), stream_);
      } else {
        // simpls.hpp:824-858: candidates from the projected operator
        const int used = steps_ref[s].a;
        if (rank_one_) {
          dominant_direction(omega_.d() + steps_ref[s].omega_off, used, cand_.d());
        } else {
          operator_rsvd(
              [&](const double* Z, int w, double* out) { projected_fwd(Z, w, out, used); },
              [&](const double* B, int w, double* out) { projected_adj(B, w, out, used); },
              p_, q_, b, oversample_, power_, unsigned(seed_) + unsigned(used), true, cand_.d(),
              op_sig_.d(), nullptr, state_.as<int>() + 2);
        }
        launch_div_rows(cand_.d(), sx_.d(), p_, b, cand_scaled_.d(), stream_);
      }
      // scores of the candidate block: X_s C = (X - 1 mᵀ)(C / s)
      XwTask<T> xw{};
      xw.A = X_; xw.m = n_; xw.k = p_; xw.center = mx_.d(); xw.W = cand_scaled_.d(); xw.ldw = p_;
      xw.N = b; xw.splits = xs; xw.out = ts_part_.d(); xw.ldo = n_; xw.slab = (long long)n_ * b;
      xw.frag = frag(xw_frag_bytes<T>(p_, b, f64));
      launch_xw_single<T>(xw, stream_, f64);
      launch_sum_slabs(ts_part_.d(), xs, (long long)n_ * b, (long long)n_ * b, t_raw_.d(), stream_);
      if (cache_gram_) {
        // cached route: the loadings come from X_sᵀX_s in the update
        launch_simpls_update(fitdev_.as<SimplsFitDev>(), 1, int(s), stream_);
        continue;
      }
      // loadings: (X - 1 mᵀ)ᵀ t  (divided by s in the update)
      XtwTask<T> xt{};
      xt.A = X_; xt.m = n_; xt.k = p_; xt.center = mx_.d(); xt.W = t_raw_.d(); xt.ldw = n_;
      xt.N = b; xt.mode = kCentered; xt.splits = ls; xt.out = ls_part_.d(); xt.ldo = p_;
      xt.slab = (long long)p_ * bmax;
      xt.frag = frag(xtw_frag_bytes<T>(n_, b, f64));
      launch_xtw_single<T>(xt, stream_, f64);
      // response loadings without S0: Ỹᵀ t (scaled by 1/|t| in the update)
      if (implicit_) tproduct<double>(Yc_.d(), n_, q_, nullptr, t_raw_.d(), n_, b, qraw_.d());
      if (wide_)
        wide_update(steps_ref[s], ls, (long long)p_ * bmax, right_gram);
      else
        launch_simpls_update(fitdev_.as<SimplsFitDev>(), 1, int(s), stream_);
    }
    FPLS_HIP(hipEventRecord(sc1_.e, stream_));
  }

  // refresh_directions (simpls.hpp:226-368) of a wide fit through the
  // multi-work-group products: the q-side power iterations on the right Gram
  // (or on S when the right Gram is not used), the sample S Z, its CholQR2
  // basis Qb, Bm = Qbᵀ S, the top-b eigenvectors of Bm Bmᵀ (one work-group)
  // and the candidates Qb v, divided by the scales for the score pass.
  void wide_refresh(const RefreshStep& st, bool right_gram) {
    const int w = st.w, b = st.b;
    int* state = state_.as<int>();
    int* flags = state + 2;
    Buf &Z = wide_ws_[0], &Y2 = wide_ws_[1], &sample = wide_ws_[2], &basis = wide_ws_[3],
        &Bm = wide_ws_[4], &vtop = wide_ws_[5], &scratch = wide_ws_[6];
    Z.reserve(sizeof(double) * (size_t)q_ * w);
    Y2.reserve(sizeof(double) * (size_t)q_ * w);
    sample.reserve(sizeof(double) * (size_t)p_ * w);
    basis.reserve(sizeof(double) * (size_t)p_ * w);
    Bm.reserve(sizeof(double) * (size_t)w * q_);
    vtop.reserve(sizeof(double) * (size_t)w * b);
    scratch.reserve(sizeof(double) * ((size_t)3 * w * w + w));
    if (right_gram && w == q_) {
      // a sketch as wide as the response count spans ℝ^q: the directions are
      // the top-b left singular vectors of S, S v_c for the eigenvectors v_c
      // of SᵀS formed from the current S (k_simpls_refresh's square-sketch
      // step; the update divides each candidate by its score norm, so their
      // scale does not matter)
      tproduct<double>(S_.d(), p_, q_, nullptr, S_.d(), p_, q_, Bm.d());
      launch_refresh_eigen(Bm.d(), w, q_, b, vtop.d(), scratch.d(), state, stream_, true);
      product<double>(S_.d(), p_, q_, nullptr, vtop.d(), q_, b, cand_.d());
      launch_div_rows(cand_.d(), sx_.d(), p_, b, cand_scaled_.d(), stream_);
      return;
    }
    const double* omega = omega_.d() + st.omega_off;
    if (right_gram) {
      FPLS_HIP(hipMemcpyAsync(Z.d(), omega, sizeof(double) * (size_t)q_ * w,
                              hipMemcpyDeviceToDevice, stream_));
      for (int it = 0; it < power_; ++it) {
        product<double>(G_.d(), q_, q_, nullptr, Z.d(), q_, w, Y2.d());
        qr_tall(Y2.d(), q_, w, Z.d(), flags);
      }
      if (power_ == 0) {
        FPLS_HIP(hipMemcpyAsync(Y2.d(), Z.d(), sizeof(double) * (size_t)q_ * w,
                                hipMemcpyDeviceToDevice, stream_));
        qr_tall(Y2.d(), q_, w, Z.d(), flags);
      }
      product<double>(S_.d(), p_, q_, nullptr, Z.d(), q_, w, sample.d());
    } else {
      product<double>(S_.d(), p_, q_, nullptr, omega, q_, w, sample.d());
      for (int it = 0; it < power_; ++it) {
        qr_tall(sample.d(), p_, w, basis.d(), flags);
        tproduct<double>(S_.d(), p_, q_, nullptr, basis.d(), p_, w, Y2.d());
        qr_tall(Y2.d(), q_, w, Z.d(), flags);
        product<double>(S_.d(), p_, q_, nullptr, Z.d(), q_, w, sample.d());
      }
    }
    qr_tall(sample.d(), p_, w, basis.d(), flags);
    tproduct<double>(basis.d(), p_, w, nullptr, S_.d(), p_, q_, Bm.d());
    launch_refresh_gate(state, stream_);
    launch_refresh_eigen(Bm.d(), w, q_, b, vtop.d(), scratch.d(), state, stream_);
    product<double>(basis.d(), p_, w, nullptr, vtop.d(), w, b, cand_.d());
    launch_div_rows(cand_.d(), sx_.d(), p_, b, cand_scaled_.d(), stream_);
  }

  // The component update (simpls.hpp:600-756, k_simpls_update's steps) of a
  // wide fit, one candidate of the block at a time: |t|² and the rescaled
  // T, R and loading; q = S0ᵀ r; the loading's projection on the earlier
  // loadings removed (once, twice when re-orthogonalising) and normalised;
  // the deflation row vᵀ S, S -= v row and the right-Gram downdate.
  void wide_update(const RefreshStep& st, int splits, long long slab, bool right_gram) {
    int* state = state_.as<int>();
    Buf &scal = wide_ws_[7], &proj = wide_ws_[8], &row = wide_ws_[9], &dv = wide_ws_[10];
    scal.reserve(sizeof(double) * 2);
    proj.reserve(sizeof(double) * (size_t)std::max(A_, 1));
    row.reserve(sizeof(double) * (size_t)q_);
    dv.reserve(sizeof(double) * (size_t)p_);
    for (int j = 0; j < st.b; ++j) {
      const int comp = st.a + j;
      if (comp >= A_) break;
      const double* t = t_raw_.d() + (size_t)j * n_;
      double* Rc = R_.d() + (size_t)comp * p_;
      double* Vc = V_.d() + (size_t)comp * p_;
      tproduct<double>(t, n_, 1, nullptr, t, n_, 1, scal.d());
      launch_wide_component(n_, p_, t, cand_.d() + (size_t)j * p_, ls_part_.d() + (size_t)j * p_,
                            splits, slab, sx_.d(), scal.d(),
                            Tsc_.empty() ? nullptr : Tsc_.d() + (size_t)comp * n_, Rc, Vc, state,
                            stream_);
      tproduct<double>(S0_.d(), p_, q_, nullptr, Rc, p_, 1, Q_.d() + (size_t)comp * q_);
      if (comp > 0) {
        for (int pass = 0; pass < (fitdev_host_.reorth ? 2 : 1); ++pass) {
          tproduct<double>(V_.d(), p_, comp, nullptr, Vc, p_, 1, proj.d());
          product<double>(V_.d(), p_, comp, nullptr, proj.d(), comp, 1, dv.d());
          launch_wide_subtract(Vc, dv.d(), p_, state, stream_);
        }
      }
      tproduct<double>(Vc, p_, 1, nullptr, Vc, p_, 1, scal.d() + 1);
      launch_wide_unit(Vc, p_, scal.d() + 1, state, stream_);
      tproduct<double>(S_.d(), p_, q_, nullptr, Vc, p_, 1, row.d());
      launch_wide_deflate(S_.d(), p_, q_, Vc, row.d(), state, stream_);
      launch_wide_finish(right_gram ? G_.d() : nullptr, q_, row.d(), comp, state, stream_);
    }
  }

  // ----------------------------------------------------------- PLS-SVD
  // plssvd.hpp:197-220 -> rsvd.hpp:185-237 -> plssvd.hpp:57-166.
  void fit_plssvd() {
    const int k = A_;
    const int maxdim = std::min(p_, q_);
    const int width = std::min(maxdim, k + oversample_);
    R_.alloc(sizeof(double) * (size_t)p_ * k);
    Q_.alloc(sizeof(double) * (size_t)q_ * k);
    sigma_.alloc(sizeof(double) * k);
    usable_.alloc(sizeof(int));
    int* flags = state_.as<int>() + 2;
    if (implicit_) {
      // plssvd.hpp:222-248: the operator rSVD of S = X_sᵀ Ỹ without S
      operator_rsvd([&](const double* Z, int w, double* out) { crosscov_fwd(Z, w, out); },
                    [&](const double* B, int w, double* out) { crosscov_adj(B, w, out); }, p_,
                    q_, k, oversample_, power_, unsigned(seed_), false, R_.d(), sigma_.d(),
                    Q_.d(), flags);
      launch_fill_int(usable_.as<int>(), 1, k, stream_);
    } else {
      reserve_plssvd_parts(k, width);
      if (width >= maxdim) {
        dense_svd(k, flags);
        launch_fill_int(usable_.as<int>(), 1, k, stream_);
      } else {
        randomized_svd(k, width, usable_.as<int>(), flags);
      }
    }
    // (the usable component count is checked with the other fit flags)
    // scores T = X_s R, Gram TᵀT, prefix weights
    Rs_.alloc(sizeof(double) * (size_t)p_ * k);
    launch_div_rows(R_.d(), sx_.d(), p_, k, Rs_.d(), stream_);
    Tsc_.alloc(sizeof(double) * (size_t)n_ * k);
    {
      // scores T = (X - 1 mᵀ)(R / s): the second and last pass over X
      const int splits = xw_splits(n_, p_, k, sizeof(T) == 4);
      part_.reserve(sizeof(double) * (size_t)splits * n_ * k);
      XwTask<T> t{};
      t.A = X_; t.m = n_; t.k = p_; t.center = mx_.d(); t.W = Rs_.d(); t.ldw = p_; t.N = k;
      t.frag = frag(xw_frag_bytes<T>(p_, k));
      t.splits = splits; t.out = splits == 1 ? Tsc_.d() : part_.d(); t.ldo = n_;
      t.slab = (long long)n_ * k;
      FPLS_HIP(hipEventRecord(sc0_.e, stream_));
      launch_xw_single<T>(t, stream_);
      FPLS_HIP(hipEventRecord(sc1_.e, stream_));
      if (splits > 1)
        launch_sum_slabs(part_.d(), splits, (long long)n_ * k, (long long)n_ * k, Tsc_.d(), stream_);
    }
    gram_.alloc(sizeof(double) * (size_t)k * k);
    tproduct<double>(Tsc_.d(), n_, k, nullptr, Tsc_.d(), n_, k, gram_.d());
    W_.alloc(sizeof(double) * (size_t)k * k * q_);
    launch_plssvd_weights(gram_.d(), sigma_.d(), Q_.d(), k, q_, W_.d(), flags, stream_);
  }

  // Split partials of every product of a PLS-SVD fit fit in part_ at once,
  // so no product has to grow (and free) it while queued kernels read it.
  void reserve_plssvd_parts(int k, int w) {
    auto xtw_part = [&](long long m, long long kk, int N) {
      return (long long)xtw_splits(m, kk, std::min(N, 64)) * kk * std::min(N, 64);
    };
    auto xw_part = [&](long long m, long long kk, int N) {
      return (long long)xw_splits(m, kk, std::min(N, 64)) * m * std::min(N, 64);
    };
    const int mx = std::max({w, q_, k, 1});
    long long need = 0;
    for (long long v : {xtw_part(p_, q_, mx), xtw_part(p_, mx, mx), xtw_part(q_, mx, mx),
                        xtw_part(p_, mx, q_), xtw_part(q_, p_, mx), xw_part(p_, q_, mx),
                        xw_part(p_, mx, mx), xw_part(q_, p_, mx), xw_part(q_, mx, mx),
                        xw_part(p_, mx, k), xtw_part(n_, k, k), xw_part(n_, p_, k)})
      need = std::max(need, v);
    part_.reserve(sizeof(double) * (size_t)need);
  }

  // The Gaussian sketch of the randomized PLS-SVD (rsvd.hpp:205-210) goes to
  // tmp1_ before the first pass over X is queued: a copy from pageable host
  // memory returns only once the stream has reached it, so issued later it
  // would hold the host for the length of that pass.
  void stage_plssvd_sketch() {
    const int maxdim = std::min(p_, q_);
    const int w = std::min(maxdim, A_ + oversample_);
    if (w >= maxdim) return;
    std::vector<double>& om = omega_host_;
    om.resize((size_t)q_ * w);
    gaussian(unsigned(seed_), om.size(), om.data());
    tmp1_.alloc(sizeof(double) * om.size());
    // om is a member: it outlives the asynchronous copy
    FPLS_HIP(hipMemcpyAsync(tmp1_.d(), om.data(), sizeof(double) * om.size(), hipMemcpyHostToDevice,
                            stream_));
  }

  void randomized_svd(int k, int w, int* usable, int* flags) {
    Buf &sample = svd_ws_[0], &basis = svd_ws_[1], &rev = svd_ws_[2], &rbasis = svd_ws_[3],
        &P = svd_ws_[4], &Up = svd_ws_[5], &Vt = svd_ws_[6], &scratch = svd_ws_[7];
    sample.alloc(sizeof(double) * (size_t)p_ * w);
    basis.alloc(sizeof(double) * (size_t)p_ * w);
    rev.alloc(sizeof(double) * (size_t)q_ * w);
    rbasis.alloc(sizeof(double) * (size_t)q_ * w);
    product<double>(S_.d(), p_, q_, nullptr, tmp1_.d(), q_, w, sample.d());
    for (int it = 0; it < power_; ++it) {
      qr_tall(sample.d(), p_, w, basis.d(), flags);
      tproduct<double>(S_.d(), p_, q_, nullptr, basis.d(), p_, w, rev.d());
      qr_tall(rev.d(), q_, w, rbasis.d(), flags);
      product<double>(S_.d(), p_, q_, nullptr, rbasis.d(), q_, w, sample.d());
    }
    qr_tall(sample.d(), p_, w, basis.d(), flags);
    P.alloc(sizeof(double) * (size_t)w * q_);
    tproduct<double>(basis.d(), p_, w, nullptr, S_.d(), p_, q_, P.d());
    Up.alloc(sizeof(double) * (size_t)w * k);
    Vt.alloc(sizeof(double) * (size_t)k * q_);
    const int mx = std::max(w, q_);
    scratch.alloc(sizeof(double) * ((size_t)3 * w * w + w + q_ + 2 * (size_t)w * q_ +
                                    (size_t)mx * mx + mx + 64));
    launch_rsvd_finalize(P.d(), w, q_, k, Up.d(), sigma_.d(), Vt.d(), usable, scratch.d(), flags,
                         stream_);
    product<double>(basis.d(), p_, w, nullptr, Up.d(), w, k, R_.d());
    launch_transpose(Vt.d(), k, q_, Q_.d(), stream_);
  }

  void dense_svd(int k, int* flags) {
    Buf &Sc = svd_ws_[0], &Qs = svd_ws_[1], &Rm = svd_ws_[2], &Ur = svd_ws_[3],
        &scratch = svd_ws_[4], &St = svd_ws_[5];
    if (p_ >= q_) {
      Sc.alloc(sizeof(double) * (size_t)p_ * q_);
      FPLS_HIP(hipMemcpyAsync(Sc.d(), S_.d(), sizeof(double) * (size_t)p_ * q_, hipMemcpyDeviceToDevice, stream_));
      Qs.alloc(sizeof(double) * (size_t)p_ * q_);
      qr_tall(Sc.d(), p_, q_, Qs.d(), flags);
      Rm.alloc(sizeof(double) * (size_t)q_ * q_);
      tproduct<double>(Qs.d(), p_, q_, nullptr, S_.d(), p_, q_, Rm.d());
      Ur.alloc(sizeof(double) * (size_t)q_ * k);
      scratch.alloc(sizeof(double) * (3 * (size_t)q_ * q_ + q_ + 64));
      launch_small_svd(Rm.d(), q_, q_, k, Ur.d(), sigma_.d(), Q_.d(), scratch.d(), flags, stream_);
      product<double>(Qs.d(), p_, q_, nullptr, Ur.d(), q_, k, R_.d());
    } else {
      Sc.alloc(sizeof(double) * (size_t)p_ * q_);
      launch_transpose(S_.d(), p_, q_, Sc.d(), stream_);  // q x p
      Qs.alloc(sizeof(double) * (size_t)q_ * p_);
      St.alloc(sizeof(double) * (size_t)q_ * p_);
      FPLS_HIP(hipMemcpyAsync(St.d(), Sc.d(), sizeof(double) * (size_t)p_ * q_, hipMemcpyDeviceToDevice, stream_));
      launch_qr(Sc.d(), q_, p_, Qs.d(), flags, stream_);
      Rm.alloc(sizeof(double) * (size_t)p_ * p_);
      tproduct<double>(Qs.d(), q_, p_, nullptr, St.d(), q_, p_, Rm.d());
      Ur.alloc(sizeof(double) * (size_t)p_ * k);
      scratch.alloc(sizeof(double) * (3 * (size_t)p_ * p_ + p_ + 64));
      // Sᵀ = Qs Ur s Vrᵀ  =>  S = Vr s (Qs Ur)ᵀ
      launch_small_svd(Rm.d(), p_, p_, k, Ur.d(), sigma_.d(), R_.d(), scratch.d(), flags, stream_);
      product<double>(Qs.d(), q_, p_, nullptr, Ur.d(), p_, k, Q_.d());
    }
  }

  // ------------------------------------------------------------ common
  void finish_fit() {
    // prediction operands: R / s (p x A) and Qᵀ is read with strides
    if (method_ == kSimpls) {
      Rs_.alloc(sizeof(double) * (size_t)p_ * A_);
      launch_div_rows(R_.d(), sx_.d(), p_, A_, Rs_.d(), stream_);
    }
  }

  void check_state() {
    int st[4] = {0, 0, 0, 0};
    FPLS_HIP(hipMemcpyAsync(st, state_.as<int>(), sizeof(st), hipMemcpyDeviceToHost, stream_));
    FPLS_HIP(hipStreamSynchronize(stream_));
    if (st[2] & kFlagBadLabel) throw std::runtime_error("resident fit failed: labels must lie in 1..q");
    if (st[2] & kFlagEmptyClass) throw std::runtime_error("resident fit failed: a class has no samples");
    if (st[2] & kFlagNonFinite) throw std::runtime_error("resident fit failed: non-finite values in the inputs or components");
    if (st[2] & kFlagDecomposition) throw std::runtime_error("resident fit failed: decomposition or score-Gram solve failed");
    if (method_ == kSimpls && st[0] < A_)
      throw std::runtime_error("fastPLS SIMPLS returned fewer components than requested");
    if (method_ == kPlssvd && !usable_.empty()) {
      int got = 0;
      FPLS_HIP(hipMemcpyAsync(&got, usable_.as<int>(), sizeof(int), hipMemcpyDeviceToHost, stream_));
      FPLS_HIP(hipStreamSynchronize(stream_));
      if (got < A_) throw std::runtime_error("fastPLS PLS-SVD returned fewer components than requested");
    }
  }

  void check_path(const int* prefixes, int count) {
    if (!prefixes || count < 1) throw std::invalid_argument("empty component path");
    int previous = 0;
    for (int j = 0; j < count; ++j) {
      if (prefixes[j] <= previous || prefixes[j] > A_)
        throw std::invalid_argument("component path must be strictly increasing");
      previous = prefixes[j];
    }
  }

  // Device address of `rows` x p query rows: host rows are staged in rows_,
  // device rows are read in place.
  const T* load_rows(const T* x, int rows, bool on_device = false) {
    if (on_device) return x;
    rows_.reserve(sizeof(T) * (size_t)rows * p_);
    FPLS_HIP(hipMemcpyAsync(rows_.as<T>(), x, sizeof(T) * (size_t)rows * p_, hipMemcpyHostToDevice, stream_));
    return rows_.as<T>();
  }

  // Rows start .. start + count of the host rows x p matrix x (leading
  // dimension rows) into the device row buffer (count x p).
  const T* load_row_block(const T* x, int rows, int start, int count) {
    if (start == 0 && count == rows) return load_rows(x, rows);
    rows_.reserve(sizeof(T) * (size_t)count * p_);
    FPLS_HIP(hipMemcpy2DAsync(rows_.as<T>(), sizeof(T) * (size_t)count, x + start,
                              sizeof(T) * (size_t)rows, sizeof(T) * (size_t)count, p_,
                              hipMemcpyHostToDevice, stream_));
    return rows_.as<T>();
  }

  // Device bytes a query holds per host row: the row, its scores, FP64 and
  // output responses and top-k labels.
  size_t query_row_bytes(int prefix, int top) const {
    return sizeof(T) * (size_t)p_ + sizeof(double) * size_t(prefix + q_) + sizeof(T) * (size_t)q_ +
           sizeof(int) * (size_t)top;
  }

  // Rows per block of a query on host rows (src/cuda_resident_api.cuh:259-297,
  // where the reference blocks classify; here predictions as well): when the
  // query's workspace (rows x bytes_per_row) exceeds 2/5 of the free device
  // memory, rows go through in blocks of max(1024, min(rows, 65536,
  // max(1024 rows' bytes, 3/20 of free) / bytes_per_row)) rows, else all at
  // once. FASTPLS_HIP_QUERY_FREE_BYTES, when set, stands for the free memory
  // (so blocking can be exercised at small sizes).
  int query_block_rows(int rows, size_t bytes_per_row) const {
    size_t free_b = 0, total_b = 0;
    if (const char* v = std::getenv("FASTPLS_HIP_QUERY_FREE_BYTES"))
      free_b = size_t(std::strtoull(v, nullptr, 10));
    else
      FPLS_HIP(hipMemGetInfo(&free_b, &total_b));
    if (size_t(rows) * bytes_per_row <= free_b / 5 * 2) return rows;
    const size_t target = std::max(bytes_per_row * 1024, free_b / 20 * 3);
    const long long fit = (long long)(target / bytes_per_row);
    return int(std::max<long long>(1024, std::min<long long>({(long long)rows, 65536LL, fit})));
  }

  // scores (rows x prefix) = ((x - m) / s) R[:, :prefix]
  void project_rows(const T* x, int rows, int prefix) {
    scores_.reserve(sizeof(double) * (size_t)rows * prefix);
    FPLS_HIP(hipEventRecord(pr0_.e, stream_));
    product<T>(x, rows, p_, mx_.d(), Rs_.d(), p_, prefix, scores_.d());
    FPLS_HIP(hipEventRecord(pr1_.e, stream_));
    proj_timed_ = true;
    // algorithmic bytes: the query rows, R / s and the scores once each
    proj_bytes_ = double(rows) * p_ * sizeof(T) + 8.0 * p_ * prefix + 8.0 * rows * prefix;
    scores_rows_ = rows;
  }

  // Responses (or LDA discriminants) of one prefix from scores_, FP64 into
  // pred64_, converted copy to host `out` (rows x q, leading dimension ldo;
  // 0 = rows) when non-null.
  void response_block(int rows, int prefix, bool lda, T* out, int ldo) {
    pred64_.reserve(sizeof(double) * (size_t)rows * q_);
    if (lda) {
      ensure_lda(prefix);
      // disc = T_k linearᵀ + constants ; linear is K x k (ld K)
      launch_apply_small<T>(scores_.d(), rows, prefix, lda_linear_.d(), q_, 1, q_,
                            lda_constants_.d(), nullptr, pred64_.d(), stream_);
    } else if (method_ == kSimpls) {
      // Ŷ = T_k Q_kᵀ + mean_y ; Q is q x A (ld q)
      launch_apply_small<T>(scores_.d(), rows, prefix, Q_.d(), q_, 1, q_, my_.d(), nullptr,
                            pred64_.d(), stream_);
    } else {
      // Ŷ = T_k W_k + mean_y ; W_k is k x q (ld k)
      const double* Wk = W_.d() + (size_t)(prefix - 1) * A_ * q_;
      launch_apply_small<T>(scores_.d(), rows, prefix, Wk, 1, prefix, q_, my_.d(), nullptr,
                            pred64_.d(), stream_);
    }
    if (out) {
      outT_.reserve(sizeof(T) * (size_t)rows * q_);
      launch_convert<T>(pred64_.d(), outT_.as<T>(), (long long)rows * q_, stream_);
      if (ldo <= 0 || ldo == rows)
        FPLS_HIP(hipMemcpyAsync(out, outT_.as<T>(), sizeof(T) * (size_t)rows * q_,
                                hipMemcpyDeviceToHost, stream_));
      else
        FPLS_HIP(hipMemcpy2DAsync(out, sizeof(T) * (size_t)ldo, outT_.as<T>(),
                                  sizeof(T) * (size_t)rows, sizeof(T) * (size_t)rows, q_,
                                  hipMemcpyDeviceToHost, stream_));
      // outT_ is reused by the next block: keep the copy ordered before it
      FPLS_HIP(hipStreamSynchronize(stream_));
    }
  }

  // --------------------------------------------------------------- LDA
  // lda.hpp:156-223 through the moments TᵀT and class sums of T.
  void ensure_lda_moments() {
    if (!lda_gram_.empty()) return;
    if (Tsc_.empty() || H_.empty()) throw std::runtime_error("LDA moments need the training scores and labels");
    lda_gram_.alloc(sizeof(double) * (size_t)A_ * A_);
    tproduct<double>(Tsc_.d(), n_, A_, nullptr, Tsc_.d(), n_, A_, lda_gram_.d());
    lda_sums_.alloc(sizeof(double) * (size_t)q_ * A_);
    tproduct<double>(H_.d(), n_, q_, nullptr, Tsc_.d(), n_, A_, lda_sums_.d());
    FPLS_HIP(hipStreamSynchronize(stream_));
  }

  void ensure_lda(int k) {
    if (!classification_) throw std::invalid_argument("LDA requires class labels");
    ensure_lda_moments();
    if (lda_prefix_ == k) return;
    lda_linear_.reserve(sizeof(double) * (size_t)q_ * A_);
    lda_constants_.reserve(sizeof(double) * q_);
    lda_ridge_.reserve(sizeof(double));
    lda_scratch_.reserve(sizeof(double) * (size_t)q_ * A_);
    Buf fl;
    fl.bind(stream_);
    fl.alloc(sizeof(int));
    fill_zero_int(fl.as<int>(), 1);
    launch_lda_prepare(lda_gram_.d(), lda_sums_.d(), counts_.as<int>(), n_, A_, q_, k,
                       lda_linear_.d(), lda_constants_.d(), lda_ridge_.d(), lda_scratch_.d(),
                       fl.as<int>(), stream_);
    int bad = 0;
    FPLS_HIP(hipMemcpyAsync(&bad, fl.as<int>(), sizeof(int), hipMemcpyDeviceToHost, stream_));
    FPLS_HIP(hipStreamSynchronize(stream_));
    if (bad) throw std::runtime_error("resident LDA Cholesky failed for every regularization value");
    lda_prefix_ = k;
  }

  // ---------------------------------------------------------- variance
  void ensure_variance() {
    if (!var_ss_.empty()) return;
    if (X_ == nullptr || Tsc_.empty()) throw std::runtime_error("variance needs the training data; export it before compact");
    Buf cross, gram, seq;
    for (Buf* b : {&cross, &gram, &seq}) b->bind(stream_);
    cross.alloc(sizeof(double) * (size_t)p_ * A_);
    tproduct<T>(X_, n_, p_, mx_.d(), Tsc_.d(), n_, A_, cross.d());
    launch_div_rows(cross.d(), sx_.d(), p_, A_, cross.d(), stream_);
    gram.alloc(sizeof(double) * (size_t)A_ * A_);
    tproduct<double>(Tsc_.d(), n_, A_, nullptr, Tsc_.d(), n_, A_, gram.d());
    seq.alloc(sizeof(double) * (size_t)p_ * A_);
    var_loadings_.alloc(sizeof(double) * (size_t)p_ * A_);
    var_ss_.alloc(sizeof(double) * (A_ + 1));
    launch_variance(cross.d(), gram.d(), colss_.d(), p_, A_, var_loadings_.d(), var_ss_.d(),
                    seq.d(), stream_);
    FPLS_HIP(hipStreamSynchronize(stream_));
  }

  // ------------------------------------------------------------ members
  int method_, n_, p_, q_, A_ = 0, scaling_, oversample_, power_;
  unsigned long long seed_;
  bool classification_;
  int block_ = 1;
  bool compacted_ = false;
  double fit_seconds_ = 0.0, crossprod_seconds_ = 0.0, score_seconds_ = 0.0;
  Event cp0_, cp1_, sc0_, sc1_, pr0_, pr1_;
  bool owns_stream_ = true;
  bool proj_timed_ = false;
  double proj_bytes_ = 0.0;
  const T* X_ = nullptr;
  Buf xown_, yown_, labels_dev_, counts_, Yc_, H_, ysum_, my_, mx_, sx_, colss_, S_, S0_, G_;
  Buf R_, Q_, V_, Tsc_, Rs_, sigma_, gram_, W_, state_;
  Buf P_;  // loadings X_sᵀ T of a re-orthogonalising fit (fit time only)
  Buf omega_, steps_, fitdev_, simpls_scratch_, cand_, cand_scaled_, t_raw_, ts_part_, ls_part_;
  Buf tmp1_, tmp2_, tmp3_;
  // fit workspaces that outlive the kernels reading them (no mid-fit frees)
  Buf cp_part_, cp_stats_, usable_, svd_ws_[8];
  std::vector<double> omega_host_;
  std::vector<RefreshStep> steps_host_;
  SimplsFitDev fitdev_host_{};
  Buf rows_, scores_, pred64_, outT_, top_, metric_, flags_, obs_, obs_labels_;
  Buf lda_gram_, lda_sums_, lda_linear_, lda_constants_, lda_ridge_, lda_scratch_;
  Buf var_loadings_, var_ss_;
  int lda_prefix_ = 0;
  int scores_rows_ = 0;
  bool implicit_ = false, rank_one_ = false;
  bool cache_gram_ = false;   // SIMPLS loadings from the cached X_sᵀX_s (xgram_)
  bool wide_ = false;         // multi-launch refresh and update (many predictors)
  // measured crossover (FP32 fits, q 3..17): the one-work-group route wins
  // at p = 1000 (4.1 against 5.0 ms), the multi-launch route at p = 2186
  // (11.8 against 16.5 ms) and beyond
  static constexpr long long kWideP = 1536;
  Buf wide_ws_[11];
  bool keep_scores_ = true;   // training scores kept after the fit
  Buf xgram_;
  // matrix-free route workspaces
  Buf qraw_, op_sig_, op_tmp_, op_scaled_, op_coef_, op_corr_, op_pb_, op_rev_, op_sample_;
};

}  // namespace

Handle* create(int method, const void* x, const void* y, const int* labels, int precision,
               bool device, int n, int p, int q, int components, int scaling, int oversample,
               int power, unsigned long long seed, hipStream_t stream, bool retain_scores) {
  require_device();
  if (precision == 32)
    return new Model<float>(method, static_cast<const float*>(x), static_cast<const float*>(y),
                            labels, device, n, p, q, components, scaling, oversample, power,
                            seed, stream, retain_scores);
  if (precision == 64)
    return new Model<double>(method, static_cast<const double*>(x),
                             static_cast<const double*>(y), labels, device, n, p, q, components,
                             scaling, oversample, power, seed, stream, retain_scores);
  throw std::invalid_argument("precision must be 32 or 64");
}

namespace {

template<class T>
void host_gemm(const T* left, const T* right, int rows, int inner, int cols, T* out) {
  hipStream_t s = stream_take();
  Buf a, b, c, c64, part, wf;
  try {
    a.alloc(sizeof(T) * (size_t)rows * inner);
    b.alloc(sizeof(T) * (size_t)inner * cols);
    Buf b64;
    b64.alloc(sizeof(double) * (size_t)inner * cols);
    c64.alloc(sizeof(double) * (size_t)rows * cols);
    c.alloc(sizeof(T) * (size_t)rows * cols);
    FPLS_HIP(hipMemcpyAsync(a.as<T>(), left, sizeof(T) * (size_t)rows * inner, hipMemcpyHostToDevice, s));
    std::vector<double> r64((size_t)inner * cols);
    for (size_t i = 0; i < r64.size(); ++i) r64[i] = double(right[i]);
    FPLS_HIP(hipMemcpyAsync(b64.d(), r64.data(), sizeof(double) * r64.size(), hipMemcpyHostToDevice, s));
    for (int c0 = 0; c0 < cols; c0 += 64) {
      const int nb = std::min(64, cols - c0);
      const int splits = xw_splits(rows, inner, nb);
      part.reserve(sizeof(double) * (size_t)splits * rows * nb);
      wf.reserve(xw_frag_bytes<T>(inner, nb));
      XwTask<T> t{};
      t.A = a.as<T>(); t.m = rows; t.k = inner; t.center = nullptr;
      t.W = b64.d() + (size_t)c0 * inner; t.ldw = inner; t.N = nb; t.splits = splits;
      t.frag = wf.as<void>();
      t.out = part.d(); t.ldo = rows; t.slab = (long long)rows * nb;
      launch_xw_single<T>(t, s);
      launch_sum_slabs(part.d(), splits, (long long)rows * nb, (long long)rows * nb,
                       c64.d() + (size_t)c0 * rows, s);
      FPLS_HIP(hipStreamSynchronize(s));
    }
    launch_convert<T>(c64.d(), c.as<T>(), (long long)rows * cols, s);
    FPLS_HIP(hipMemcpyAsync(out, c.as<T>(), sizeof(T) * (size_t)rows * cols, hipMemcpyDeviceToHost, s));
    FPLS_HIP(hipStreamSynchronize(s));
  } catch (...) {
    stream_give(s);
    throw;
  }
  stream_give(s);
}

}  // namespace

// Device address of `scratch` on the current device, looked up once per device.
static double* noalloc_scratch() {
  static std::mutex mu;
  static double* addr[64] = {};
  int dev = 0;
  FPLS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw std::runtime_error("device index outside 0..63");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") +
                               hipGetErrorString(e));
    addr[dev] = static_cast<double*>(p);
  }
  return addr[dev];
}

}  // namespace fpls

using fpls::Handle;

extern "C" {

int fastpls_hip_device_info(char* out, size_t capacity) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    fpls::write_error(out, capacity, "no HIP device");
    return 1;
  }
  hipDeviceProp_t prop{};
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
    fpls::write_error(out, capacity, "hipGetDeviceProperties failed");
    return 1;
  }
  if (out && capacity)
    std::snprintf(out, capacity, "%s (%s), %d CUs, %.1f GiB", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount, double(prop.totalGlobalMem) / (1024.0 * 1024.0 * 1024.0));
  return 0;
}

int fastpls_hip_gemm(const void* left, const void* right, int precision, int rows, int inner,
                     int columns, void* output, char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    fpls::require_device();
    if (!left || !right || !output || rows < 1 || inner < 1 || columns < 1)
      throw std::invalid_argument("invalid GEMM arguments");
    if (precision == 32)
      fpls::host_gemm<float>(static_cast<const float*>(left), static_cast<const float*>(right),
                             rows, inner, columns, static_cast<float*>(output));
    else if (precision == 64)
      fpls::host_gemm<double>(static_cast<const double*>(left), static_cast<const double*>(right),
                              rows, inner, columns, static_cast<double*>(output));
    else
      throw std::invalid_argument("precision must be 32 or 64");
  });
}

void* fastpls_hip_simpls_create(const void* x, const void* y, const int* labels, int precision,
                                int n, int p, int q, int components, int scaling, int oversample,
                                int power, int retain_scores, unsigned long long seed,
                                char* error, size_t error_capacity) {
  Handle* h = nullptr;
  fpls::guarded(error, error_capacity, [&] {
    if (retain_scores != 0 && retain_scores != 1)
      throw std::invalid_argument("invalid resident score-retention request");
    h = fpls::create(0, x, y, labels, precision, false, n, p, q, components, scaling,
                     oversample, power, seed, nullptr, retain_scores == 1);
  });
  return h;
}

void* fastpls_hip_plssvd_create(const void* x, const void* y, const int* labels, int precision,
                                int n, int p, int q, int components, int scaling, int oversample,
                                int power, int retain_scores, unsigned long long seed,
                                char* error, size_t error_capacity) {
  Handle* h = nullptr;
  fpls::guarded(error, error_capacity, [&] {
    if (retain_scores != 0 && retain_scores != 1)
      throw std::invalid_argument("invalid resident score-retention request");
    h = fpls::create(1, x, y, labels, precision, false, n, p, q, components, scaling,
                     oversample, power, seed, nullptr, retain_scores == 1);
  });
  return h;
}

void* fastpls_hip_opls_create(const void* x, const void* y, const int* labels, int precision,
                              int n, int p, int q, int components, int scaling, int oversample,
                              int power, int retain_scores, unsigned long long seed, int north,
                              char* error, size_t error_capacity) {
  Handle* h = nullptr;
  fpls::guarded(error, error_capacity, [&] {
    if (!x || (precision != 32 && precision != 64) || north < 0)
      throw std::invalid_argument("invalid resident HIP OPLS input");
    if (retain_scores != 0 && retain_scores != 1)
      throw std::invalid_argument("invalid resident score-retention request");
    h = fpls::create_opls(x, y, labels, precision, n, p, q, components, scaling, oversample,
                          power, seed, north, retain_scores == 1);
  });
  return h;
}

void* fastpls_hip_kernelpls_create(const void* x, const void* y, const int* labels, int precision,
                                   int n, int p, int q, int components, int scaling,
                                   int oversample, int power, int retain_scores,
                                   unsigned long long seed, int kernel, double gamma, int degree,
                                   double coef0, char* error, size_t error_capacity) {
  Handle* h = nullptr;
  fpls::guarded(error, error_capacity, [&] {
    if (!x || (precision != 32 && precision != 64) || (kernel != 2 && kernel != 3))
      throw std::invalid_argument("invalid resident HIP nonlinear kernel PLS input");
    if (retain_scores != 0 && retain_scores != 1)
      throw std::invalid_argument("invalid resident score-retention request");
    h = fpls::create_kernelpls(x, y, labels, precision, n, p, q, components, scaling, oversample,
                               power, seed, kernel, gamma, degree, coef0, retain_scores == 1);
  });
  return h;
}

void* fastpls_hip_fit_device(int method, const void* x_device, const void* y_device,
                             const int* labels_device, int precision, int n, int p, int q,
                             int components, int scaling, int oversample, int power,
                             unsigned long long seed, char* error, size_t error_capacity) {
  Handle* h = nullptr;
  fpls::guarded(error, error_capacity, [&] {
    if (method != 0 && method != 1) throw std::invalid_argument("method must be 0 (SIMPLS) or 1 (PLS-SVD)");
    h = fpls::create(method, x_device, y_device, labels_device, precision, true, n, p, q,
                     components, scaling, oversample, power, seed);
  });
  return h;
}


int fastpls_hip_predict_noalloc(void* model, const void* x_device, int rows, int prefix,
                                void* predictions_device, char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!model) throw std::invalid_argument("null resident model");
    static_cast<Handle*>(model)->predict_scratch(x_device, rows, prefix, predictions_device,
                                                 fpls::noalloc_scratch(),
                                                 fpls::kNoAllocScratchDoubles);
  });
}

int fastpls_hip_predict_path(void* model, const void* x, int rows, const int* prefixes,
                             int prefix_count, int lda, void* predictions, char* error,
                             size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!model) throw std::invalid_argument("null resident model");
    static_cast<Handle*>(model)->predict_path(x, rows, prefixes, prefix_count, lda != 0, predictions,
                                              false);
  });
}

void fastpls_hip_destroy(void* model) { delete static_cast<Handle*>(model); }

int fastpls_hip_export(void* model, int field, void* out, size_t elements, char* error,
                       size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!model) throw std::invalid_argument("null resident model");
    static_cast<Handle*>(model)->export_field(field, out, elements);
  });
}

int fastpls_hip_classify_path(void* model, const void* x, int rows, const int* prefixes,
                              int prefix_count, int lda, int top, int* labels, char* error,
                              size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!model) throw std::invalid_argument("null resident model");
    static_cast<Handle*>(model)->classify_path(x, rows, prefixes, prefix_count, lda != 0, top,
                                               labels, nullptr, false);
  });
}

int fastpls_hip_classify_response_path(void* model, const void* x, int rows,
                                       const int* prefixes, int prefix_count, int lda, int top,
                                       int* labels, void* predictions, char* error,
                                       size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!model) throw std::invalid_argument("null resident model");
    if (!predictions) throw std::invalid_argument("invalid resident classification-response path request");
    static_cast<Handle*>(model)->classify_path(x, rows, prefixes, prefix_count, lda != 0, top,
                                               labels, predictions, false);
  });
}

int fastpls_hip_response_sums(void* model, const void* x, const void* y, const int* labels,
                              int rows, int prefix, void* out, char* error,
                              size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!model) throw std::invalid_argument("null resident model");
    static_cast<Handle*>(model)->response_sums(x, y, labels, rows, prefix, out);
  });
}

int fastpls_hip_project(void* model, const void* x, int rows, int prefix, void* out, char* error,
                        size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!model) throw std::invalid_argument("null resident model");
    static_cast<Handle*>(model)->project(x, rows, prefix, out);
  });
}

int fastpls_hip_controls(void* model, int* oversample, int* power, int* block, int* block_limit,
                         int* implicit_operator, int* predictor_crossprod_cache, char* error,
                         size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!model) throw std::invalid_argument("null resident model");
    static_cast<Handle*>(model)->controls(oversample, power, block, block_limit,
                                          implicit_operator, predictor_crossprod_cache);
  });
}

int fastpls_hip_compact(void* model, int prepare_lda, char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!model) throw std::invalid_argument("null resident model");
    static_cast<Handle*>(model)->compact(prepare_lda != 0);
  });
}

int fastpls_hip_last_fit_seconds(void* model, double* seconds, char* error,
                                 size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!model || !seconds) throw std::invalid_argument("null argument");
    *seconds = static_cast<Handle*>(model)->last_fit_seconds();
  });
}

int fastpls_hip_device_alloc(size_t bytes, void** ptr, char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    fpls::require_device();
    if (!ptr) throw std::invalid_argument("null output pointer");
    *ptr = nullptr;
    FPLS_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  });
}

int fastpls_hip_device_free(void* ptr, char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (ptr) FPLS_HIP(hipFree(ptr));
  });
}

int fastpls_hip_memcpy(void* dst, const void* src, size_t bytes, int kind, char* error,
                       size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!dst || !src) throw std::invalid_argument("null copy pointer");
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                    : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    FPLS_HIP(hipMemcpy(dst, src, bytes, k));
  });
}

int fastpls_hip_fit_phases(void* model, double* out, int count, char* error,
                           size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!model || !out || count < 1) throw std::invalid_argument("null argument");
    static_cast<Handle*>(model)->phases(out, count);
  });
}

int fastpls_hip_latent_data(void* x_device, void* y_device, int precision, long long n, int p,
                            int q, int factors, double decay, double noise,
                            unsigned long long seed, int round32, char* error,
                            size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    fpls::require_device();
    if (!x_device || !y_device || n < 1 || p < 1 || q < 1 || factors < 1)
      throw std::invalid_argument("invalid synthetic data request");
    hipStream_t s = fpls::stream_take();
    struct Give {
      hipStream_t s;
      ~Give() { fpls::stream_give(s); }
    } give{s};
    fpls::Buf scratch, y64;
    scratch.alloc(sizeof(double) * (size_t)(n + p + q) * factors);
    y64.alloc(sizeof(double) * (size_t)n * q);
    if (precision == 32) {
      fpls::launch_latent_data<float>(static_cast<float*>(x_device), y64.d(), n, p, q, factors,
                                      decay, noise, seed, 1, scratch.d(), s);
      fpls::launch_convert<float>(y64.d(), static_cast<float*>(y_device), n * q, s);
    } else if (precision == 64) {
      fpls::launch_latent_data<double>(static_cast<double*>(x_device), y64.d(), n, p, q,
                                       factors, decay, noise, seed, round32 ? 1 : 0, scratch.d(),
                                       s);
      fpls::launch_convert<double>(y64.d(), static_cast<double*>(y_device), n * q, s);
    } else {
      throw std::invalid_argument("precision must be 32 or 64");
    }
    FPLS_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
