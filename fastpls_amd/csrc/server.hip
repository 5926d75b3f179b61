// Request server on one GPU: continuous batching of fit and predict requests
// under an HBM budget.
//
// Worker threads take requests from one FIFO queue as device memory and
// workers free up: a fit is admitted when its device-byte estimate fits the
// budget left by running fits and resident models and fewer than the fit
// cap are running, so predicts always find workers (a fit that has to wait
// may be passed by later requests that can run); a predict runs as soon as its
// model is idle, together with every predict queued behind it for the same
// model and prefix (up to a row cap), as one product over the concatenated
// rows. Fitted models stay resident (compacted) until released. Query products
// use a depth-only split rule and every reduction has a fixed order, so a
// request gives bitwise the same result alone and in a mix. Latency is
// measured from submission to completion.
//
// Streams: the server owns one stream per fit slot and a few query streams,
// whatever the number of models, and never more than the device's hardware
// queues can hold. Work on one hardware queue runs in order, so a stream that
// shares a queue with a fit's stream waits behind the fit's kernels (hundreds
// of milliseconds for a large fit), and queues beyond the hardware slots are
// time-sliced; fits and queries therefore run on disjoint streams: a fit on
// its slot's stream, every predict batch on a query stream its model is
// attached to for the batch, with its rows staged on that same stream.
//
// The reference has no serving loop; its resident handle API
// (src/cuda_resident_api.cuh:69-479) is what each request runs.
#include "../../include/fastpls_hip.h"
#include "handle.hpp"
#include "host_util.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace fpls {
namespace {

using Clock = std::chrono::steady_clock;

enum ReqKind { kFitRequest = 0, kPredictRequest = 1 };
enum ReqState { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Request {
  long long id = 0;
  int kind = kFitRequest;
  // fit
  int method = 0, precision = 64, n = 0, p = 0, q = 0, components = 0, scaling = 2;
  int oversample = 32, power = 5;
  unsigned long long seed = 1;
  const void* x = nullptr;
  const void* y = nullptr;
  const int* labels = nullptr;
  bool device_inputs = false;
  size_t bytes = 0;  // admission estimate of a fit
  int slot = -1;      // fit stream slot while running
  // predict
  long long model = 0;
  int rows = 0, prefix = 0;
  long long ldx = 0;  // leading dimension of the predict rows
  void* out = nullptr;
  // outcome
  int state = kQueued;
  std::string error;
  Clock::time_point submitted, started, done;
};

struct Resident {
  std::unique_ptr<Handle> handle;
  int precision = 64, p = 0, q = 0, components = 0;
  size_t bytes = 0;
  bool busy = false;
};

// Device bytes a fit may hold at its peak (inputs, centred responses, the
// cross-product partials, the bf16 operand workspace, scores and the score
// pass partials), rounded up generously; admission only.
size_t fit_estimate(int method, int precision, bool device_inputs, long long n, long long p,
                    long long q, long long A, bool labels) {
  const double s = precision == 32 ? 4.0 : 8.0;
  double b = device_inputs ? 0.0 : double(n) * double(p) * s + double(n) * double(q) * s;
  b += double(n) * double(q) * 8.0 * (labels ? 2.0 : 1.0);   // centred responses (+ one-hot)
  const double splits = std::min<double>(4096.0, std::max<double>(1.0, double(n) / 4096.0) + 1.0);
  b += double(p) * double(q) * 8.0 * (3.0 + splits);          // S, S0, partials
  if (precision == 32) b += double(n + 64) * 64.0 * 6.0;      // xtw operand planes
  b += double(n) * double(A) * 8.0 * 8.0;                     // scores and score partials
  b += double(p) * double(std::max<long long>(A, 64)) * 8.0 * 8.0;  // R, V, sketches, parts
  if (method == 1) b += double(A) * double(A) * double(q) * 8.0;
  return size_t(b * 1.25) + (size_t(64) << 20);
}

class Server {
 public:
  Server(double budget, int workers, int max_rows, int max_fits)
      : budget_(size_t(budget)), max_rows_(max_rows),
        max_fits_(max_fits > 0 ? std::min(max_fits, kMaxFitStreams) : 2) {
    if (!(budget > 0.0)) throw std::invalid_argument("server HBM budget must be positive");
    if (workers < 1 || workers > 64) throw std::invalid_argument("server workers must lie in 1..64");
    if (max_rows < 1) throw std::invalid_argument("server batch row cap must be positive");
    require_device();
    int device = 0;
    FPLS_HIP(hipGetDevice(&device));
    device_ = device;
    // fit and query streams together stay within the hardware queues HIP
    // gives the process (GPU_MAX_HW_QUEUES, 4 unless set)
    const int hwq = hardware_queues();
    const int queries = std::max(1, std::min({kMaxQueryStreams, workers - 1, hwq - 1}));
    max_fits_ = std::max(1, std::min(max_fits_, hwq - queries));
    // Each stream is bound to a hardware queue when it is created: once the
    // GPU_MAX_HW_QUEUES queues exist, to the one the fewest live streams use
    // (the first such queue on ties). Two streams on one queue run in order,
    // so a query stream that shared a queue with a fit stream would wait
    // behind the fit's kernels. Other live streams of the process (idle
    // cached ones, those of open handles) leave the queues unevenly used, so
    // the server creates a run of streams: the first ones even the usage
    // out, after which consecutive streams land on consecutive queues, and
    // the server keeps the last ones (which then share no queue) and
    // destroys the rest, which never ran work. Idle cached streams are
    // destroyed first, so the run is short.
    stream_cache_drop();
    const int keep = queries + max_fits_;
    const int run = std::max(keep, 8 * hwq);
    std::vector<hipStream_t> made;
    try {
      for (int i = 0; i < run; ++i) made.push_back(stream_create());
    } catch (...) {
      for (hipStream_t st : made) (void)hipStreamDestroy(st);
      throw;
    }
    for (int i = 0; i < run - keep; ++i) (void)hipStreamDestroy(made[size_t(i)]);
    made.erase(made.begin(), made.begin() + (run - keep));
    for (int i = 0; i < queries; ++i) query_streams_.push_back(made[size_t(i)]);
    for (int i = 0; i < max_fits_; ++i) {
      fit_streams_.push_back(made[size_t(queries + i)]);
      free_slots_.push_back(i);
    }
    for (int w = 0; w < workers; ++w) threads_.emplace_back([this, w] { worker(w); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
    models_.clear();   // resident handles drain their (server) streams
    give_streams();
  }

  long long submit_fit(int method, const void* x, const void* y, const int* labels, int precision,
                       bool device_inputs, int n, int p, int q, int components, int scaling,
                       int oversample, int power, unsigned long long seed) {
    if (method != 0 && method != 1) throw std::invalid_argument("method must be 0 (SIMPLS) or 1 (PLS-SVD)");
    if (precision != 32 && precision != 64) throw std::invalid_argument("precision must be 32 or 64");
    if (!x || (y == nullptr) == (labels == nullptr)) throw std::invalid_argument("provide predictors and responses or labels");
    if (n < 2 || p < 1 || q < 1 || components < 1) throw std::invalid_argument("invalid fit request dimensions");
    auto r = std::make_unique<Request>();
    r->kind = kFitRequest;
    r->method = method; r->precision = precision; r->n = n; r->p = p; r->q = q;
    r->components = components; r->scaling = scaling; r->oversample = oversample; r->power = power;
    r->seed = seed; r->x = x; r->y = y; r->labels = labels; r->device_inputs = device_inputs;
    r->bytes = fit_estimate(method, precision, device_inputs, n, p, q, components, labels != nullptr);
    if (r->bytes > budget_) throw std::invalid_argument("fit request exceeds the server HBM budget");
    return enqueue(std::move(r));
  }

  long long submit_predict(long long model, const void* x, long long ldx, bool device_inputs,
                           int rows, int prefix, void* out) {
    if (!x || !out || rows < 1 || prefix < 1 || ldx < rows) throw std::invalid_argument("invalid predict request");
    auto r = std::make_unique<Request>();
    r->kind = kPredictRequest;
    r->model = model; r->x = x; r->ldx = ldx; r->device_inputs = device_inputs; r->rows = rows;
    r->prefix = prefix; r->out = out;
    {
      std::lock_guard<std::mutex> lock(mu_);
      auto it = models_.find(model);
      if (it == models_.end()) throw std::invalid_argument("unknown or released server model");
      if (prefix > it->second.components) throw std::invalid_argument("prefix exceeds the model's components");
      if (rows > max_rows_) throw std::invalid_argument("predict request exceeds the server batch row cap");
    }
    return enqueue(std::move(r));
  }

  int poll(long long id, bool wait, double* latency_ms, std::string* error) {
    std::unique_lock<std::mutex> lock(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end()) throw std::invalid_argument("unknown server request");
    Request* r = it->second.get();
    if (wait) done_cv_.wait(lock, [&] { return r->state == kDone || r->state == kFailed; });
    if (latency_ms && (r->state == kDone || r->state == kFailed))
      *latency_ms = std::chrono::duration<double, std::milli>(r->done - r->submitted).count();
    if (error) *error = r->error;
    const int state = r->state;
    if (state == kDone || state == kFailed) requests_.erase(id);   // polled once when finished
    return state;
  }

  void release(long long model) {
    std::unique_ptr<Handle> drop;
    {
      std::unique_lock<std::mutex> lock(mu_);
      done_cv_.wait(lock, [&] {
        auto f = models_.find(model);
        return f == models_.end() || !f->second.busy;
      });
      auto it = models_.find(model);
      if (it == models_.end()) throw std::invalid_argument("unknown or released server model");
      drop = std::move(it->second.handle);
      in_use_ -= it->second.bytes;
      models_.erase(it);
    }
    drop.reset();
    cv_.notify_all();
  }

  void export_field(long long model, int field, void* out, size_t elements) {
    Handle* h = nullptr;
    {
      std::unique_lock<std::mutex> lock(mu_);
      done_cv_.wait(lock, [&] {
        auto f = models_.find(model);
        return f == models_.end() || !f->second.busy;
      });
      auto it = models_.find(model);
      if (it == models_.end()) throw std::invalid_argument("unknown or released server model");
      it->second.busy = true;
      h = it->second.handle.get();
    }
    std::string err;
    try {
      h->export_field(field, out, elements);
    } catch (const std::exception& e) {
      err = e.what();
    }
    {
      std::lock_guard<std::mutex> lock(mu_);
      models_[model].busy = false;
    }
    done_cv_.notify_all();
    cv_.notify_all();
    if (!err.empty()) throw std::runtime_error(err);
  }

  void stats(double* out, int count) {
    std::lock_guard<std::mutex> lock(mu_);
    auto pct = [](std::vector<double> v, double f) {
      if (v.empty()) return 0.0;
      std::sort(v.begin(), v.end());
      const size_t i = std::min(v.size() - 1, size_t(f * double(v.size() - 1) + 0.5));
      return v[i];
    };
    const double span = completed_ > 0
        ? std::chrono::duration<double>(last_done_ - first_submit_).count() : 0.0;
    const double v[18] = {double(completed_), double(failed_), pct(lat_predict_, 0.5),
                          pct(lat_predict_, 0.99), pct(lat_fit_, 0.5), pct(lat_fit_, 0.99),
                          span > 0.0 ? double(completed_) / span : 0.0, double(peak_in_use_),
                          double(budget_), double(peak_running_), double(batches_),
                          batches_ ? double(batch_rows_) / double(batches_) : 0.0,
                          double(max_batch_requests_), span, double(in_use_),
                          double(models_.size()),
                          lat_fit_.empty() ? 0.0 : service_fit_ms_ / double(lat_fit_.size()),
                          lat_predict_.empty() ? 0.0 : service_predict_ms_ / double(lat_predict_.size())};
    for (int i = 0; i < count && i < 18; ++i) out[i] = v[i];
  }

 private:
  long long enqueue(std::unique_ptr<Request> r) {
    long long id;
    {
      std::lock_guard<std::mutex> lock(mu_);
      id = next_id_++;
      r->id = id;
      r->submitted = Clock::now();
      if (!any_submitted_) { first_submit_ = r->submitted; any_submitted_ = true; }
      queue_.push_back(r.get());
      requests_[id] = std::move(r);
    }
    cv_.notify_one();
    return id;
  }

  static constexpr int kMaxFitStreams = 4;
  static constexpr int kMaxQueryStreams = 2;

  void give_streams() noexcept {
    for (hipStream_t st : fit_streams_) stream_give(st);
    for (hipStream_t st : query_streams_) stream_give(st);
    fit_streams_.clear();
    query_streams_.clear();
  }

  // Under mu_: the next runnable work, FIFO with bounded bypass. Predicts on
  // an unknown model fail here.
  bool pick(std::vector<Request*>& work) {
    constexpr size_t kScan = 256;
    for (size_t i = 0; i < queue_.size() && i < kScan; ++i) {
      Request* r = queue_[i];
      if (r->kind == kFitRequest) {
        if (running_fits_ >= max_fits_ || in_use_ + r->bytes > budget_) continue;
        in_use_ += r->bytes;
        ++running_fits_;
        r->slot = free_slots_.back();
        free_slots_.pop_back();
        work.push_back(r);
        queue_.erase(queue_.begin() + long(i));
        return true;
      }
      auto it = models_.find(r->model);
      if (it == models_.end()) {
        finish(r, "unknown or released server model");
        queue_.erase(queue_.begin() + long(i));
        --i;
        continue;
      }
      if (it->second.busy) continue;
      int rows = 0;
      for (size_t j = i; j < queue_.size();) {
        Request* o = queue_[j];
        if (o->kind == kPredictRequest && o->model == r->model && o->prefix == r->prefix &&
            rows + o->rows <= max_rows_) {
          rows += o->rows;
          work.push_back(o);
          queue_.erase(queue_.begin() + long(j));
        } else {
          ++j;
        }
      }
      it->second.busy = true;
      return true;
    }
    return false;
  }

  // Under mu_.
  void finish(Request* r, const std::string& error) {
    r->done = Clock::now();
    r->error = error;
    r->state = error.empty() ? kDone : kFailed;
    if (error.empty()) {
      ++completed_;
      (r->kind == kFitRequest ? lat_fit_ : lat_predict_)
          .push_back(std::chrono::duration<double, std::milli>(r->done - r->submitted).count());
      (r->kind == kFitRequest ? service_fit_ms_ : service_predict_ms_) +=
          std::chrono::duration<double, std::milli>(r->done - r->started).count();
    } else {
      ++failed_;
    }
    last_done_ = r->done;
  }

  void worker(int w) {
    (void)hipSetDevice(device_);
    // this worker's predict batches run on one query stream, where its
    // staging buffers live too
    hipStream_t qs = query_streams_[size_t(w) % query_streams_.size()];
    Buf staging;                 // concatenated predict rows on the device
    Buf upload;                  // one request's host rows
    staging.bind(qs);
    upload.bind(qs);
    std::vector<char> host_out;  // concatenated predictions
    for (;;) {
      std::vector<Request*> work;
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [&] { return stop_ || pick(work); });
        if (work.empty()) {  // stopping
          lock.unlock();
          staging.release();
          upload.release();
          return;
        }
        const Clock::time_point now = Clock::now();
        for (Request* r : work) { r->state = kRunning; r->started = now; }
        ++running_;
        peak_running_ = std::max(peak_running_, running_);
        peak_in_use_ = std::max(peak_in_use_, in_use_);
      }
      done_cv_.notify_all();
      if (work[0]->kind == kFitRequest) run_fit(work[0], qs);
      else run_predict(work, qs, staging, upload, host_out);
      {
        std::lock_guard<std::mutex> lock(mu_);
        --running_;
      }
      done_cv_.notify_all();
      cv_.notify_all();
    }
  }

  void run_fit(Request* r, hipStream_t qs) {
    std::unique_ptr<Handle> h;
    std::string error;
    size_t resident = 0;
    try {
      h.reset(create(r->method, r->x, r->y, r->labels, r->precision, r->device_inputs, r->n,
                     r->p, r->q, r->components, r->scaling, r->oversample, r->power, r->seed,
                     fit_streams_[size_t(r->slot)]));
      h->compact(false);
      h->set_batch_invariant(true);
      // off the fit slot's stream before the slot is handed on
      h->attach_stream(qs);
      resident = h->device_bytes();
    } catch (const std::exception& e) {
      error = e.what();
      h.reset();
    }
    std::lock_guard<std::mutex> lock(mu_);
    in_use_ -= r->bytes;
    --running_fits_;
    free_slots_.push_back(r->slot);
    if (h) {
      Resident m;
      m.handle = std::move(h);
      m.precision = r->precision; m.p = r->p; m.q = r->q; m.components = r->components;
      m.bytes = resident;
      in_use_ += resident;
      peak_in_use_ = std::max(peak_in_use_, in_use_);
      models_[r->id] = std::move(m);
    }
    finish(r, error);
  }

  void run_predict(std::vector<Request*>& work, hipStream_t qs, Buf& staging, Buf& upload,
                   std::vector<char>& host_out) {
    Handle* h = nullptr;
    int p = 0, q = 0, precision = 64;
    {
      std::lock_guard<std::mutex> lock(mu_);
      Resident& m = models_.at(work[0]->model);
      h = m.handle.get();
      p = m.p; q = m.q; precision = m.precision;
    }
    const size_t s = precision == 32 ? 4 : 8;
    int total = 0;
    for (Request* r : work) total += r->rows;
    std::string error;
    try {
      // the model is busy for this batch: no other work runs on it
      h->attach_stream(qs);
      staging.reserve(s * (size_t)total * p);
      // column-major concatenation: request i fills rows [off_i, off_i + rows_i);
      // host rows go up in one contiguous copy, then a device block copy
      // places them
      int off = 0;
      for (Request* r : work) {
        const void* src = r->x;
        long long lds = r->ldx;
        if (!r->device_inputs) {
          upload.reserve(s * (size_t)r->ldx * p);
          FPLS_HIP(hipMemcpyAsync(upload.as<void>(), r->x, s * (size_t)r->ldx * p,
                                  hipMemcpyHostToDevice, qs));
          src = upload.as<void>();
        }
        if (precision == 32)
          launch_copy_block<float>(static_cast<const float*>(src), lds, r->rows, p,
                                   staging.as<float>() + off, total, qs);
        else
          launch_copy_block<double>(static_cast<const double*>(src), lds, r->rows, p,
                                    staging.as<double>() + off, total, qs);
        if (!r->device_inputs) FPLS_HIP(hipStreamSynchronize(qs));  // upload is reused
        off += r->rows;
      }
      host_out.resize(s * (size_t)total * q);
      const int prefix = work[0]->prefix;
      h->predict_path(staging.as<void>(), total, &prefix, 1, false, host_out.data(), true);
      off = 0;
      for (Request* r : work) {
        char* dst = static_cast<char*>(r->out);
        for (int c = 0; c < q; ++c)
          std::memcpy(dst + s * (size_t)c * r->rows, host_out.data() + s * ((size_t)c * total + off),
                      s * (size_t)r->rows);
        off += r->rows;
      }
    } catch (const std::exception& e) {
      error = e.what();
    }
    std::lock_guard<std::mutex> lock(mu_);
    models_.at(work[0]->model).busy = false;
    ++batches_;
    batch_rows_ += total;
    max_batch_requests_ = std::max(max_batch_requests_, (long long)work.size());
    for (Request* r : work) finish(r, error);
  }

  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<Request*> queue_;
  std::unordered_map<long long, std::unique_ptr<Request>> requests_;
  std::unordered_map<long long, Resident> models_;
  size_t budget_ = 0, in_use_ = 0, peak_in_use_ = 0;
  int max_rows_ = 0;
  int max_fits_ = 1;
  int running_fits_ = 0;
  std::vector<hipStream_t> fit_streams_, query_streams_;
  std::vector<int> free_slots_;   // fit stream slots not in use
  int device_ = 0;
  long long next_id_ = 1;
  bool stop_ = false;
  int running_ = 0, peak_running_ = 0;
  long long completed_ = 0, failed_ = 0, batches_ = 0, batch_rows_ = 0, max_batch_requests_ = 0;
  bool any_submitted_ = false;
  Clock::time_point first_submit_, last_done_;
  std::vector<double> lat_fit_, lat_predict_;
  double service_fit_ms_ = 0.0, service_predict_ms_ = 0.0;
  std::vector<std::thread> threads_;
};

}  // namespace
}  // namespace fpls

using fpls::Server;

extern "C" {

void* fastpls_hip_server_create(double hbm_budget_bytes, int workers, int max_batch_rows,
                                int max_concurrent_fits, char* error, size_t error_capacity) {
  Server* s = nullptr;
  fpls::guarded(error, error_capacity, [&] {
    s = new Server(hbm_budget_bytes, workers, max_batch_rows, max_concurrent_fits);
  });
  return s;
}

long long fastpls_hip_server_submit_fit(void* server, int method, const void* x, const void* y,
                                        const int* labels, int precision, int device_inputs,
                                        int n, int p, int q, int components, int scaling,
                                        int oversample, int power, unsigned long long seed,
                                        char* error, size_t error_capacity) {
  long long id = 0;
  fpls::guarded(error, error_capacity, [&] {
    if (!server) throw std::invalid_argument("null server");
    id = static_cast<Server*>(server)->submit_fit(method, x, y, labels, precision,
                                                  device_inputs != 0, n, p, q, components,
                                                  scaling, oversample, power, seed);
  });
  return id;
}

long long fastpls_hip_server_submit_predict(void* server, long long model, const void* x,
                                            long long ldx, int device_inputs, int rows,
                                            int prefix, void* predictions, char* error,
                                            size_t error_capacity) {
  long long id = 0;
  fpls::guarded(error, error_capacity, [&] {
    if (!server) throw std::invalid_argument("null server");
    id = static_cast<Server*>(server)->submit_predict(model, x, ldx, device_inputs != 0, rows,
                                                      prefix, predictions);
  });
  return id;
}

int fastpls_hip_server_poll(void* server, long long request, int wait, int* state,
                            double* latency_ms, char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!server || !state) throw std::invalid_argument("null server or state output");
    std::string message;
    *state = static_cast<Server*>(server)->poll(request, wait != 0, latency_ms, &message);
    if (*state == fpls::kFailed) fpls::write_error(error, error_capacity, message.c_str());
  });
}

int fastpls_hip_server_release(void* server, long long model, char* error,
                               size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!server) throw std::invalid_argument("null server");
    static_cast<Server*>(server)->release(model);
  });
}

int fastpls_hip_server_export(void* server, long long model, int field, void* out,
                              size_t elements, char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!server) throw std::invalid_argument("null server");
    static_cast<Server*>(server)->export_field(model, field, out, elements);
  });
}

int fastpls_hip_server_stats(void* server, double* out, int count, char* error,
                             size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!server || !out) throw std::invalid_argument("null server or stats output");
    static_cast<Server*>(server)->stats(out, count);
  });
}

void fastpls_hip_server_destroy(void* server) { delete static_cast<Server*>(server); }

}  // extern "C"
