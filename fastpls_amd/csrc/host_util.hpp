// Host-side helpers shared by the single-fit engine and the batch engine.
#pragma once

#include "common.hpp"
#include "stream.hpp"

#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <random>
#include <stdexcept>
#include <vector>

namespace fpls {

// ------------------------------------------------------------- buffers
// Device block cache shared by every handle of the process (pool.hip). A
// block given back with the stream that last used it is handed out again at
// once to that stream (stream order covers the earlier work) and to any
// other stream once an event recorded at the give-back has completed, so
// freeing never synchronises the device.
void* pool_take(size_t bytes, hipStream_t stream, size_t* capacity);
void pool_give(void* ptr, size_t capacity, hipStream_t stream) noexcept;
// Called once `stream` has been drained and before it is destroyed: its
// blocks become idle and no event recorded on it is queried again.
void pool_forget_stream(hipStream_t stream) noexcept;

// A non-blocking stream of default priority from the process-wide stream
// cache (created when none is idle), and its return: stream_give drains the
// stream and keeps it for the next owner.
hipStream_t stream_take();
void stream_give(hipStream_t stream) noexcept;
// A new stream, never one from the cache (the runtime binds a stream to one
// of its hardware queues when the stream is created, and a cached stream
// keeps the queue it got then), and the destruction of every idle cached
// stream.
hipStream_t stream_create();
void stream_cache_drop() noexcept;
// Hardware queues HIP gives the process (GPU_MAX_HW_QUEUES, 4 unless set).
inline int hardware_queues() {
  const char* v = std::getenv("GPU_MAX_HW_QUEUES");
  const int n = v ? std::atoi(v) : 0;
  return n > 1 ? n : 4;
}
// `count` streams (at most hardware_queues()) that share no hardware queue,
// for work split over streams that should run at the same time, owned by the
// caller until stream_set_give drains them and keeps the set for the next
// caller asking for as many. A new set is the last `count` of a run of
// streams created back to back (the runtime binds a new stream to the
// hardware queue the fewest live streams use, so the first ones of the run
// even the usage out and consecutive ones then land on consecutive queues);
// the rest of the run, which never ran work, is destroyed.
std::vector<hipStream_t> stream_set_take(int count);
void stream_set_give(std::vector<hipStream_t>& set) noexcept;

// Device buffer. Unbound buffers own their memory directly (hipMalloc /
// hipFree); buffers bound to a stream take it from the block cache and give
// it back in that stream's order, so every use of a bound buffer must be
// ordered on its stream.
class Buf {
 public:
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() { release(); }
  void bind(hipStream_t stream) {
    if (ptr_) throw std::logic_error("bind a device buffer before its first allocation");
    stream_ = stream;
    bound_ = true;
  }
  // Moves a bound buffer to another stream. The caller has synchronised the
  // old stream, so no queued work uses the memory any more.
  void rebind(hipStream_t stream) {
    if (bound_) stream_ = stream;
  }
  void alloc(size_t bytes) {
    release();
    if (bytes) {
      if (bound_) ptr_ = pool_take(bytes, stream_, &cap_);
      else FPLS_HIP(hipMalloc(&ptr_, bytes));
    }
    bytes_ = bytes;
  }
  void reserve(size_t bytes) {
    if (bytes > bytes_) alloc(bytes);
  }
  void release() noexcept {
    if (ptr_) {
      if (bound_) pool_give(ptr_, cap_, stream_);
      else (void)hipFree(ptr_);
    }
    ptr_ = nullptr;
    bytes_ = 0;
    cap_ = 0;
  }
  template<class U> U* as() const { return static_cast<U*>(ptr_); }
  double* d() const { return static_cast<double*>(ptr_); }
  bool empty() const { return ptr_ == nullptr; }
  size_t bytes() const { return bytes_; }

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
  size_t cap_ = 0;
  hipStream_t stream_ = nullptr;
  bool bound_ = false;
};

// Gaussian sketch exactly as the reference draws it (simpls.hpp:244-249).
inline void gaussian(unsigned int seed, size_t count, double* out) {
  std::mt19937 generator(seed);
  std::normal_distribution<double> normal(0.0, 1.0);
  for (size_t i = 0; i < count; ++i) out[i] = normal(generator);
}

// Split counts: aim for >= 2048 work-groups, keep every split >= 32 wide.
inline int splits_for(long long tiles_per_split, long long extent, long long min_chunk) {
  long long s = (2048 + tiles_per_split - 1) / std::max<long long>(tiles_per_split, 1);
  const long long cap = std::max<long long>(1, extent / min_chunk);
  s = std::max<long long>(1, std::min(s, cap));
  return int(std::min<long long>(s, 256));
}

// Compute units of the current device, queried once per process.
inline int device_cus() {
  static const int cus = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      return v;
    return 256;
  }();
  return cus;
}

// Split count of a bf16x6 matrix-core pass (one work-group per CU at a time).
// The pass runs in ceil(tiles·s / CUs) rounds of work-groups, each covering
// 1/s of the streamed extent, and every split adds one FP64 partial slab that
// is written and summed again. Picks the s in [1, min(256, extent / quantum)]
// minimising
//   stream_bytes · (CUs / tiles) · ceil(tiles·s / CUs) / s + 2 s · slab_bytes,
// the smallest on ties (measured on MI355X: the PLS-SVD preprocessing pass
// over a 1e5 x 2e4 FP32 X runs in 1.94 ms at 13 splits and 2.23 ms at 25).
inline int balanced_splits(long long tiles, long long extent, long long quantum,
                           double stream_bytes, double slab_bytes) {
  const long long cus = device_cus();
  const long long cap = std::max<long long>(1, std::min<long long>(256, extent / quantum));
  const long long t = std::max<long long>(tiles, 1);
  int best = 1;
  double best_cost = 0.0;
  for (long long s = 1; s <= cap; ++s) {
    const double rounds = double((t * s + cus - 1) / cus);
    const double cost = stream_bytes * double(cus) / double(t) * rounds / double(s) +
                        2.0 * double(s) * slab_bytes;
    if (s == 1 || cost < best_cost * (1.0 - 1e-9)) {
      best_cost = cost;
      best = int(s);
    }
  }
  return best;
}

// Row splits of an xtw pass over an m x k matrix with N right-hand columns.
// FP32 A on the matrix cores (bf6) balances the rounds of work-groups against
// the partial slabs; otherwise enough work-groups to fill the device, and for
// the FP64 matrix-core kernel no split longer than kXtwSplitRows.
inline int xtw_splits(long long m, long long k, int N, bool bf6 = false) {
  if (bf6 && N >= kMfmaMinN)
    return balanced_splits(xtw_tiles(k, N), m, 128, 4.0 * double(m) * double(k),
                           8.0 * double(k) * double(std::min(N, 64)));
  int s = splits_for(xtw_tiles(k, N), m, 128);
  if (N >= kMfmaMinN) {
    const long long need = (m + kXtwSplitRows - 1) / kXtwSplitRows;
    if (need > s) s = int(std::min<long long>(need, 4096));
  }
  return s;
}

// Column splits of an xw pass (m x k times k x N), chosen as for xtw.
inline int xw_splits(long long m, long long k, int N, bool bf6 = false) {
  if (bf6 && N >= kMfmaMinN)
    return balanced_splits(xw_row_tiles(m), k, 64, 4.0 * double(m) * double(k),
                           8.0 * double(m) * double(std::min(N, 64)));
  int s = splits_for(xw_row_tiles(m), k, 32);
  if (N >= kMfmaMinN) {
    const long long need = (k + kXwSplitCols - 1) / kXwSplitCols;
    if (need > s) s = int(std::min<long long>(need, 4096));
  }
  return s;
}

// Column splits of an xw pass that depend on the depth k only (batch-
// invariant queries): no split deeper than 2048 columns.
inline int xw_splits_by_depth(long long k) {
  return int(std::min<long long>(4096, std::max<long long>(1, (k + 2047) / 2048)));
}

// simpls.hpp:43-62
inline int candidate_block_size(long long remaining, long long p, long long q, bool cls, long long n,
                         long long maximum) {
  const double work = double(n) * double(p) * double(q);
  if (q <= 1 || remaining < 4 || work < 5.0e8) return 1;
  if (cls && q <= 2048) return int(std::max<long long>(1, std::min({maximum, remaining, p, q})));
  if (!cls && double(p) * double(q) > 64.0 * 1024.0 * 1024.0)
    return int(std::max<long long>(1, std::min({8LL, maximum, remaining, p, q})));
  return 1;
}

struct Event {
  hipEvent_t e = nullptr;
  Event() { FPLS_HIP(hipEventCreate(&e)); }
  ~Event() { if (e) (void)hipEventDestroy(e); }
};

// ------------------------------------------------------------ C ABI glue
// Exceptions never cross the C ABI: every entry point runs its body through
// `guarded`, which returns 0 on success and 1 with the message copied into
// the caller's buffer (src/cuda_resident_api.cuh:480-482).
inline void write_error(char* out, size_t cap, const char* message) noexcept {
  if (out && cap) std::snprintf(out, cap, "%s", message);
}

template<class F>
int guarded(char* err, size_t cap, F&& f) noexcept {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    write_error(err, cap, e.what());
  } catch (...) {
    write_error(err, cap, "unknown fastPLS HIP failure");
  }
  return 1;
}

inline void require_device() {
  int count = 0;
  const hipError_t st = hipGetDeviceCount(&count);
  if (st != hipSuccess || count < 1)
    throw std::runtime_error("no HIP device is available (the HIP backend has no CPU fallback)");
}

}  // namespace fpls
