// Process-wide caches of device blocks for stream-bound buffers and of
// streams.
//
// Sizes are rounded up to one of four classes per power of two (at most 25%
// slack). A given-back block keeps the stream that last used it and an event
// recorded there; a later request of the same class takes it at once on that
// stream, or on another stream once the event has completed. Blocks are
// never returned to the device while the process runs, except that the
// cache drops completed blocks (largest first) when what it holds idle
// exceeds kIdleLimit.
#include "host_util.hpp"

#include <map>
#include <mutex>
#include <vector>

namespace fpls {
namespace {

constexpr size_t kIdleLimit = size_t(48) << 30;

struct Block {
  void* ptr;
  hipStream_t stream;   // null once the stream is drained and forgotten
  hipEvent_t ready;     // null once the block is known to be idle
};

struct Pool {
  std::mutex mu;
  std::map<size_t, std::vector<Block>> idle;   // by size class
  std::vector<hipEvent_t> spare_events;
  size_t idle_bytes = 0;
};

// Whether the work that last used a block has finished; a block seen idle
// once drops its event, so it is never queried again. hipEventQuery reports
// "not ready" through the thread's last-error slot, which the launch checks
// read, so the slot is cleared. Called with the lock held.
bool idle_now(Pool& P, Block& b) {
  if (b.ready == nullptr) return true;
  const hipError_t st = hipEventQuery(b.ready);
  (void)hipGetLastError();
  if (st != hipSuccess) return false;
  P.spare_events.push_back(b.ready);
  b.ready = nullptr;
  return true;
}

Pool& pool() {
  static Pool* p = new Pool();   // lives until the process ends
  return *p;
}

size_t size_class(size_t bytes) {
  if (bytes <= 4096) return 4096;
  size_t top = 1;
  while ((top << 1) <= bytes) top <<= 1;   // largest power of two <= bytes
  const size_t step = top / 4;
  return (bytes + step - 1) / step * step;
}

// Drop completed idle blocks (largest first) until the idle total is under
// the limit; called with the lock held.
void trim(Pool& P) {
  for (auto it = P.idle.rbegin(); it != P.idle.rend() && P.idle_bytes > kIdleLimit; ++it) {
    auto& v = it->second;
    for (size_t i = 0; i < v.size() && P.idle_bytes > kIdleLimit;) {
      if (idle_now(P, v[i])) {
        (void)hipFree(v[i].ptr);
        P.idle_bytes -= it->first;
        v.erase(v.begin() + long(i));
      } else {
        ++i;
      }
    }
  }
}

}  // namespace

void* pool_take(size_t bytes, hipStream_t stream, size_t* capacity) {
  const size_t cls = size_class(bytes);
  *capacity = cls;
  Pool& P = pool();
  {
    std::lock_guard<std::mutex> lock(P.mu);
    auto it = P.idle.find(cls);
    if (it != P.idle.end()) {
      auto& v = it->second;
      for (size_t i = 0; i < v.size(); ++i) {
        if ((v[i].stream != nullptr && v[i].stream == stream) || idle_now(P, v[i])) {
          void* ptr = v[i].ptr;
          if (v[i].ready) P.spare_events.push_back(v[i].ready);
          v.erase(v.begin() + long(i));
          P.idle_bytes -= cls;
          return ptr;
        }
      }
    }
  }
  void* ptr = nullptr;
  hipError_t st = hipMalloc(&ptr, cls);
  if (st != hipSuccess) {
    // give idle memory back to the device and try once more
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> lock(P.mu);
      for (auto& kv : P.idle)
        for (Block& b : kv.second) {
          if (b.ready) {
            (void)hipEventSynchronize(b.ready);
            P.spare_events.push_back(b.ready);
          }
          (void)hipFree(b.ptr);
        }
      P.idle.clear();
      P.idle_bytes = 0;
    }
    FPLS_HIP(hipMalloc(&ptr, cls));
  }
  return ptr;
}

void pool_give(void* ptr, size_t capacity, hipStream_t stream) noexcept {
  Pool& P = pool();
  std::lock_guard<std::mutex> lock(P.mu);
  hipEvent_t ev = nullptr;
  if (!P.spare_events.empty()) {
    ev = P.spare_events.back();
    P.spare_events.pop_back();
  } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
    (void)hipStreamSynchronize(stream);   // no event: the block is free once the stream drains
    (void)hipFree(ptr);
    return;
  }
  if (hipEventRecord(ev, stream) != hipSuccess) {
    (void)hipStreamSynchronize(stream);
    (void)hipFree(ptr);
    P.spare_events.push_back(ev);
    return;
  }
  P.idle[capacity].push_back(Block{ptr, stream, ev});
  P.idle_bytes += capacity;
  if (P.idle_bytes > kIdleLimit) trim(P);
}

// ----------------------------------------------------------- streams
// Creating a stream sets up a hardware queue on this runtime (milliseconds),
// so handles borrow non-blocking streams from this cache and give them back
// drained. Streams keep their identity across owners, so cached blocks
// tagged with a given-back stream stay valid.
namespace {

constexpr size_t kIdleStreams = 64;

struct StreamCache {
  std::mutex mu;
  std::vector<hipStream_t> idle;
};

StreamCache& streams() {
  static StreamCache* c = new StreamCache();   // lives until the process ends
  return *c;
}

}  // namespace

hipStream_t stream_take() {
  StreamCache& C = streams();
  {
    std::lock_guard<std::mutex> lock(C.mu);
    if (!C.idle.empty()) {
      hipStream_t s = C.idle.back();
      C.idle.pop_back();
      return s;
    }
  }
  return stream_create();
}

hipStream_t stream_create() {
  hipStream_t s = nullptr;
  FPLS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}

void stream_cache_drop() noexcept {
  StreamCache& C = streams();
  std::vector<hipStream_t> idle;
  {
    std::lock_guard<std::mutex> lock(C.mu);
    idle.swap(C.idle);
  }
  // drained and forgotten by the pool when they were given back
  for (hipStream_t s : idle) (void)hipStreamDestroy(s);
}

namespace {

struct StreamSets {
  std::mutex mu;
  std::vector<std::vector<hipStream_t>> idle;
};

StreamSets& stream_sets() {
  static StreamSets* c = new StreamSets();   // lives until the process ends
  return *c;
}

}  // namespace

std::vector<hipStream_t> stream_set_take(int count) {
  if (count < 1 || count > hardware_queues())
    throw std::invalid_argument("a stream set holds 1 to GPU_MAX_HW_QUEUES streams");
  StreamSets& C = stream_sets();
  {
    std::lock_guard<std::mutex> lock(C.mu);
    for (size_t i = 0; i < C.idle.size(); ++i)
      if (int(C.idle[i].size()) == count) {
        std::vector<hipStream_t> set = std::move(C.idle[i]);
        C.idle.erase(C.idle.begin() + long(i));
        return set;
      }
  }
  const int run = std::max(count, 8 * hardware_queues());
  std::vector<hipStream_t> made;
  try {
    for (int i = 0; i < run; ++i) made.push_back(stream_create());
  } catch (...) {
    for (hipStream_t st : made) (void)hipStreamDestroy(st);
    throw;
  }
  for (int i = 0; i < run - count; ++i) (void)hipStreamDestroy(made[size_t(i)]);
  made.erase(made.begin(), made.begin() + (run - count));
  return made;
}

void stream_set_give(std::vector<hipStream_t>& set) noexcept {
  if (set.empty()) return;
  for (hipStream_t s : set) {
    (void)hipStreamSynchronize(s);
    pool_forget_stream(s);
  }
  StreamSets& C = stream_sets();
  std::lock_guard<std::mutex> lock(C.mu);
  if (C.idle.size() < 8) {
    C.idle.push_back(std::move(set));
  } else {
    for (hipStream_t s : set) (void)hipStreamDestroy(s);
  }
  set.clear();
}

void stream_give(hipStream_t stream) noexcept {
  if (!stream) return;
  (void)hipStreamSynchronize(stream);
  pool_forget_stream(stream);   // its blocks are idle now: no event to query
  StreamCache& C = streams();
  {
    std::lock_guard<std::mutex> lock(C.mu);
    if (C.idle.size() < kIdleStreams) {
      C.idle.push_back(stream);
      return;
    }
  }
  (void)hipStreamDestroy(stream);
}

void pool_forget_stream(hipStream_t stream) noexcept {
  Pool& P = pool();
  std::lock_guard<std::mutex> lock(P.mu);
  for (auto& kv : P.idle)
    for (Block& b : kv.second)
      if (b.stream == stream) {
        b.stream = nullptr;
        if (b.ready) P.spare_events.push_back(b.ready);
        b.ready = nullptr;
      }
}

}  // namespace fpls
