// Grouped execution of many independent SIMPLS fits.
//
// The reference fits cross-validation folds one after another on one stream
// (src/cuda_resident_api.cuh:770-1023) and runs permutation refits on the CPU
// (R/main.R:9419-9450). Here every phase of every fit in a batch is one
// launch: the column statistics and cross-products of all fits, the right
// Grams, then per component step one refresh launch (one work-group per fit),
// one score pass and one loading pass over every distinct X (fits that share
// X, e.g. the permutation refits of one fold, form one wider product whose
// right-hand side holds all their candidates), and one update launch. A fit
// whose component count is reached simply idles in later steps.
#include "../../include/fastpls_hip.h"
#include "cv_common.hpp"
#include "host_util.hpp"
#include "kernels.hpp"
#include "stream.hpp"

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <vector>

namespace fpls {

namespace {

// A bump allocator over one device allocation (256-byte aligned pieces).
class Arena {
 public:
  size_t add(size_t bytes) {
    const size_t off = total_;
    total_ += (bytes + 255) / 256 * 256;
    return off;
  }
  void allocate() { buf_.alloc(total_ ? total_ : 256); }
  template<class U> U* at(size_t off) const { return reinterpret_cast<U*>(buf_.as<char>() + off); }
  size_t bytes() const { return total_; }

 private:
  size_t total_ = 0;
  Buf buf_;
};

struct GroupIn {
  const void* x;
  int n, p;
  // row view (all null / 0: every row): the group's fits use the rows whose
  // mask (device, n) is 1; hold (device, ascending) lists the other hold_n
  // rows and hold_x (device, hold_n x p) holds their copy
  const int* mask = nullptr;
  const int* hold = nullptr;
  const void* hold_x = nullptr;
  int hold_n = 0;
};

struct FitIn {
  int group, q, A;
  const void* y;        // device responses (n x q) or null
  const int* labels;    // device one-based labels or null
  const int* perm;      // device permutation of response rows or null
  unsigned long long seed;
};

// Split lengths of the batched passes: the longest row split of the
// cross-product and loading passes and the longest column split of the score
// pass, for groups on the VALU and for groups whose per-step passes run on
// the FP64 matrix cores (shared-X refits: those work-groups are compute-
// bound, so shorter splits keep more of them in flight). The built-in values
// were measured on the 256-fit workload (shared-X score pass 8.0 -> 4.5 ms,
// loading pass 5.5 -> 4.0 ms per batch at 256 / 512 against 1024 / 2048);
// fastpls_hip_batch_autotune replaces them with the fastest of a candidate
// set on the caller's own batch. Every length lies in [kMinSplit, kMaxSplit]
// (the upper bound is the matrix-core kernels' limit, kXtwSplitRows and
// kXwSplitCols).
struct BatchSplits {
  long long step_rows = 2048, step_cols = 1024, wide_rows = 512, wide_cols = 256;
  long long& at(int i) {
    return i == 0 ? step_rows : i == 1 ? step_cols : i == 2 ? wide_rows : wide_cols;
  }
  long long get(int i) const { return const_cast<BatchSplits*>(this)->at(i); }
};
constexpr long long kMinSplit = 128;
constexpr long long kMaxSplit = 4096;
static_assert(kMaxSplit <= kXtwSplitRows && kMaxSplit <= kXwSplitCols, "split limit");
constexpr const char* kSplitsTag = "fastpls-hip-batch-splits-v1";

bool valid_splits(const long long* v) {
  for (int i = 0; i < 4; ++i)
    if (v[i] < kMinSplit || v[i] > kMaxSplit) return false;
  return true;
}

std::mutex& splits_mutex() {
  static std::mutex m;
  return m;
}

// The process-wide default for new batches: the built-in lengths, or those of
// the file named by FASTPLS_HIP_TUNE_FILE (as an autotune wrote it) when it
// exists and holds valid lengths, until an autotune or a set replaces them.
// Called with splits_mutex held.
BatchSplits& splits_state() {
  static BatchSplits state = [] {
    BatchSplits d;
    const char* path = std::getenv("FASTPLS_HIP_TUNE_FILE");
    if (path == nullptr) return d;
    FILE* f = std::fopen(path, "r");
    if (f == nullptr) return d;
    char tag[64] = {};
    long long v[4] = {};
    if (std::fscanf(f, "%63s %lld %lld %lld %lld", tag, &v[0], &v[1], &v[2], &v[3]) == 5 &&
        std::strcmp(tag, kSplitsTag) == 0 && valid_splits(v))
      for (int i = 0; i < 4; ++i) d.at(i) = v[i];
    std::fclose(f);
    return d;
  }();
  return state;
}

BatchSplits default_splits() {
  std::lock_guard<std::mutex> lock(splits_mutex());
  return splits_state();
}

// Replaces the default; with FASTPLS_HIP_TUNE_FILE set, also writes it there.
void set_default_splits(const BatchSplits& b) {
  std::lock_guard<std::mutex> lock(splits_mutex());
  splits_state() = b;
  const char* path = std::getenv("FASTPLS_HIP_TUNE_FILE");
  if (path == nullptr) return;
  FILE* f = std::fopen(path, "w");
  if (f == nullptr) throw std::runtime_error("cannot write the split file named by FASTPLS_HIP_TUNE_FILE");
  std::fprintf(f, "%s %lld %lld %lld %lld\n", kSplitsTag, b.step_rows, b.step_cols, b.wide_rows,
               b.wide_cols);
  std::fclose(f);
}

template<class T>
class SimplsBatch {
  // Every product over X runs in FP64 (FP32-stored X is widened as it is
  // loaded): SIMPLS amplifies the rounding of these products by up to ~1e3 in
  // its later directions, so FP32-accurate (bf16x6) sums would not hold B to
  // 1e-5 of the FP64 fit.
  static constexpr bool kF64 = true;
  // right-hand chunk of the per-step score pass
  static constexpr int kXwChunk = 64;
  // Launch classes by right-hand width: 0 = VALU (below kMfmaMinN columns),
  // 1..4 = matrix-core launches compiled for 16, 32, 48, 64 columns.
  static constexpr int kWidthClasses = 5;
  static int width_class(int N) { return N < kMfmaMinN ? 0 : mfma_width(N) / 16; }

 public:
  // Plans the batch on `stream` (a stream of its own from the stream cache
  // when null) and, unless `deferred`, executes it and waits for it.
  SimplsBatch(const std::vector<GroupIn>& groups, const std::vector<FitIn>& fits, int scaling,
              int oversample, int power, int reorth, const BatchSplits& splits = default_splits(),
              hipStream_t stream = nullptr, bool deferred = false)
      : groups_(groups), fits_(fits), scaling_(scaling), oversample_(oversample), power_(power),
        reorth_(reorth), splits_(splits) {
    if (reorth < -1 || reorth > 1) throw std::invalid_argument("reorth must be -1, 0 or 1");
    if (groups_.empty() || fits_.empty()) throw std::invalid_argument("empty batch");
    if (scaling < 1 || scaling > 3) throw std::invalid_argument("scaling must be 1, 2 or 3");
    if (oversample < 0 || power < 0) throw std::invalid_argument("oversample and power must be non-negative");
    // per-phase device times with events around every launch, as the
    // reference's FASTPLS_BENCH_PHASE_TIMING switch (src/r_api.cpp:1182-1184)
    const char* pt = std::getenv("FASTPLS_BENCH_PHASE_TIMING");
    timing_ = pt != nullptr && std::strcmp(pt, "1") == 0;
    own_stream_ = stream == nullptr;
    stream_ = own_stream_ ? stream_take() : stream;
    try {
      plan();
      if (!deferred) {
        launch(nullptr, false);
        finish();
      }
    } catch (...) {
      for (TimedLaunch& t : timed_) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
      }
      timed_.clear();
      if (own_stream_) stream_give(stream_);
      else (void)hipStreamSynchronize(stream_);
      stream_ = nullptr;
      throw;
    }
  }

  ~SimplsBatch() {
    if (graph_) (void)hipGraphExecDestroy(graph_);
    if (stream_) {
      if (own_stream_) stream_give(stream_);
      else (void)hipStreamSynchronize(stream_);
    }
  }

  // Enqueues the whole execution on the batch's stream after `after` (an
  // event recorded on another stream, or null), without waiting: run()'s
  // launches, or with `graph` the graph run() was captured into on the
  // batch's stream at the first such call (so a repeated execution costs one
  // graph launch and no planning, upload or per-kernel launch on the host).
  // With phase timing on, run() is launched eagerly (events are not
  // captured).
  void launch(hipEvent_t after, bool graph) {
    const bool use_graph = graph && !timing_;
    if (use_graph && !graph_) capture();
    if (after) FPLS_HIP(hipStreamWaitEvent(stream_, after, 0));
    FPLS_HIP(hipEventRecord(start_.e, stream_));
    if (use_graph) FPLS_HIP(hipGraphLaunch(graph_, stream_));
    else run();
    FPLS_HIP(hipEventRecord(stop_.e, stream_));
  }

  // Waits for the execution launch() enqueued and reads its statuses and
  // device time (and phase times with phase timing on).
  void finish() {
    FPLS_HIP(hipEventSynchronize(stop_.e));
    collect();
    float ms = 0.f;
    FPLS_HIP(hipEventElapsedTime(&ms, start_.e, stop_.e));
    seconds_ = double(ms) * 1e-3;
    if (timing_) {
      for (auto& ph : phase_) ph[0] = ph[1] = ph[2] = 0.0;
      collect_phases();
    }
  }

  // Device seconds from `e` (recorded before launch()) to the end of the
  // finished execution.
  double seconds_since(hipEvent_t e) const {
    float ms = 0.f;
    FPLS_HIP(hipEventElapsedTime(&ms, e, stop_.e));
    return double(ms) * 1e-3;
  }

  // Executes every fit again on the current contents of the caller's device
  // inputs (X, responses or labels, permutations: same pointers and shapes),
  // as a permutation test refits with new permutations, from the captured
  // graph.
  void rerun() {
    launch(nullptr, true);
    finish();
  }

  int count() const { return int(fits_.size()); }
  double seconds() const { return seconds_; }
  const BatchSplits& splits() const { return splits_; }
  // whether some group's per-step passes run on the FP64 matrix cores
  // (wide_rows / wide_cols apply) and whether some run on the VALU
  bool has_wide() const { return has_wide_; }
  bool has_narrow() const { return has_narrow_; }

  // Phases of the batch: 0 column statistics + cross-products, 1 refresh,
  // 2 / 3 score passes (VALU / matrix-core launches), 4 score slab sums,
  // 5 / 6 loading passes (VALU / matrix-core), 7 update. out[0] = device
  // seconds of the batch; phase i fills out[1 + 3i] seconds, out[2 + 3i]
  // algorithmic bytes (X, the right-hand side and the output once each) and
  // out[3 + 3i] launches. Zero unless FASTPLS_BENCH_PHASE_TIMING=1 was set
  // when the batch was created.
  static constexpr int kPhases = 8;
  void phases(double* out, int count) const {
    for (int i = 0; i < count; ++i) out[i] = 0.0;
    if (count > 0) out[0] = seconds_;
    for (int ph = 0; ph < kPhases; ++ph)
      for (int j = 0; j < 3; ++j)
        if (1 + 3 * ph + j < count) out[1 + 3 * ph + j] = phase_[ph][j];
  }
  int completed(int f) const { return completed_[f]; }
  int flags(int f) const { return flags_[f]; }

  // Fields as fastpls_hip_export: 0 R, 1 Q, 2 T, 3 x means, 4 x scales, 5 y
  // means; in the batch precision, or as the FP64 values the fit holds (f64).
  void export_field(int f, int field, void* out, size_t elements, bool f64 = false) {
    check_fit(f);
    const Fit& F = plan_[f];
    const Group& G = gplan_[fits_[f].group];
    const size_t counts[] = {(size_t)G.p * F.A, (size_t)F.q * F.A, (size_t)G.n * F.A,
                             (size_t)G.p, (size_t)G.p, (size_t)F.q};
    if (field < 0 || field > 5 || !out || elements != counts[field])
      throw std::invalid_argument("invalid batch field output size");
    const double* src = field == 0 ? F.R : field == 1 ? F.Q : field == 2 ? F.Tsc
                      : field == 3 ? G.mx : field == 4 ? G.sx : F.my;
    if (f64) {
      FPLS_HIP(hipMemcpy(out, src, sizeof(double) * elements, hipMemcpyDeviceToHost));
      return;
    }
    std::vector<double> host(elements);
    FPLS_HIP(hipMemcpy(host.data(), src, sizeof(double) * elements, hipMemcpyDeviceToHost));
    T* o = static_cast<T*>(out);
    for (size_t i = 0; i < elements; ++i) o[i] = T(host[i]);
  }

  // Scores of device rows (rows x p, element T) for fit f: rows x prefix, FP64 (device).
  void project_device(int f, const T* x, int rows, int prefix, double* scores, Buf& part) {
    const Fit& F = plan_[f];
    const Group& G = gplan_[fits_[f].group];
    XwTask<T> t{};
    t.A = x; t.m = rows; t.k = G.p; t.center = G.mx; t.W = F.Rs; t.ldw = G.p; t.N = prefix;
    const int splits = xw_splits(rows, G.p, prefix);
    t.splits = splits; t.ldo = rows; t.slab = (long long)rows * prefix;
    const size_t fb = xw_frag_bytes<T>(G.p, prefix, kF64);
    if (fb > proj_frag_.bytes()) {
      FPLS_HIP(hipStreamSynchronize(stream_));
      proj_frag_.alloc(fb);
    }
    t.frag = fb ? proj_frag_.as<void>() : nullptr;
    if (splits == 1) {
      t.out = scores;
      launch_xw_single<T>(t, stream_, kF64);
    } else {
      const size_t pb = sizeof(double) * (size_t)splits * rows * prefix;
      if (pb > part.bytes()) {
        FPLS_HIP(hipStreamSynchronize(stream_));
        part.alloc(pb);
      }
      t.out = part.d();
      launch_xw_single<T>(t, stream_, kF64);
      launch_sum_slabs(part.d(), splits, (long long)rows * prefix, (long long)rows * prefix,
                       scores, stream_);
    }
  }

  // Ŷ (rows x q, FP64 device) = scores[:, :k] Q_kᵀ + mean_y of fit f.
  void predict_from_scores(int f, const double* scores, int rows, int k, double* out) {
    const Fit& F = plan_[f];
    launch_apply_small<T>(scores, rows, k, F.Q, F.q, 1, F.q, F.my, nullptr, out, stream_);
  }

  const double* y_mean(int f) const { return plan_[f].my; }
  hipStream_t stream() const { return stream_; }

 private:
  struct Group {
    const T* X;
    int n, p, qcat = 0, bcat = 0, max_steps = 0;
    // row view: see GroupIn; hstats (2 x p) holds the held-out rows' shifted sums
    const int* mask = nullptr;
    const int* hold = nullptr;
    const T* hold_x = nullptr;
    int hold_n = 0;
    double* hstats = nullptr;
    int used() const { return n - hold_n; }  // rows the fits see
    double *Yc, *ysum, *S, *S0, *mx, *sx, *colss, *stats, *spart;
    double *cand, *cand_scaled, *t_raw, *tpart, *loads;
    int splits = 1, ls_splits = 1, xs_splits = 1;
    int* flags;
    int* col_flags;  // per response column: offset of its fit's flag word in the status block
  };
  struct Fit {
    int q, A, bmax, wmax, nsteps, off, boff;
    bool right_gram;
    double *my, *G, *R, *Q, *V, *Tsc, *Rs, *scratch;
    int* state;
    int* counts;
    int* lperm;
    RefreshStep* steps;
    double* omega;
    std::vector<RefreshStep> host_steps;
    std::vector<double> host_omega;
  };

  void check_fit(int f) const {
    if (f < 0 || f >= count()) throw std::invalid_argument("fit index outside the batch");
  }

  // The Gaussian sketch of every refresh step (simpls.hpp:244-249, one seed
  // per fit and step), drawn on up to 8 host threads: each draw is
  // independent, so the values do not depend on the thread count.
  void draw_sketches() {
    const size_t nf = plan_.size();
    auto work = [&](size_t first, size_t stride) {
      for (size_t f = first; f < nf; f += stride)
        for (const RefreshStep& st : plan_[f].host_steps)
          gaussian(unsigned(fits_[f].seed) + unsigned(st.a), (size_t)plan_[f].q * st.w,
                   plan_[f].host_omega.data() + st.omega_off);
    };
    const size_t hw = std::max<size_t>(1, std::min<size_t>(8, std::thread::hardware_concurrency()));
    const size_t nt = std::min(hw, nf);
    std::vector<std::thread> pool;
    size_t started = 1;
    try {
      for (; started < nt; ++started) pool.emplace_back(work, started, nt);
    } catch (...) {
      // fewer threads than asked for: this thread draws the rest
    }
    work(0, nt);
    for (std::thread& t : pool) t.join();
    for (size_t t = started; t < nt; ++t) work(t, nt);
  }

  // Sizes, schedules and one arena for everything.
  void plan() {
    gplan_.resize(groups_.size());
    for (size_t g = 0; g < groups_.size(); ++g) {
      const GroupIn& in = groups_[g];
      if (!in.x || in.n < 2 || in.p < 1) throw std::invalid_argument("invalid batch group");
      gplan_[g].X = static_cast<const T*>(in.x);
      gplan_[g].n = in.n;
      gplan_[g].p = in.p;
      if (in.hold_n != 0) {
        if (in.hold_n < 0 || in.n - in.hold_n < 2 || !in.mask || !in.hold || !in.hold_x)
          throw std::invalid_argument("invalid batch row view");
        gplan_[g].mask = in.mask;
        gplan_[g].hold = in.hold;
        gplan_[g].hold_x = static_cast<const T*>(in.hold_x);
        gplan_[g].hold_n = in.hold_n;
      }
    }
    plan_.resize(fits_.size());
    for (size_t f = 0; f < fits_.size(); ++f) {
      const FitIn& in = fits_[f];
      if (in.group < 0 || in.group >= int(groups_.size())) throw std::invalid_argument("fit refers to a missing group");
      if ((in.y == nullptr) == (in.labels == nullptr)) throw std::invalid_argument("provide responses or labels, not both");
      Group& G = gplan_[in.group];
      if (G.hold_n > 0 && (in.labels || in.perm))
        throw std::invalid_argument("row views take unpermuted dense responses");
      Fit& F = plan_[f];
      F.q = in.q;
      if (F.q < 1 || (in.labels && F.q < 2)) throw std::invalid_argument("invalid response count");
      if (F.q > 512) throw std::invalid_argument("more than 512 responses is not supported");
      if (in.A < 1 || in.A > std::min(G.used() - 1, G.p)) throw std::invalid_argument("invalid batch SIMPLS dimensions");
      F.A = in.A;
      const int block = candidate_block_size(F.A, G.p, F.q, in.labels != nullptr, G.used(), 64);
      const int maxdim = std::min(G.p, F.q);
      F.right_gram = F.q <= G.p && F.q <= 512;
      F.bmax = 1; F.wmax = 1;
      for (int a = 0; a < F.A;) {
        const int b = std::min({std::max(block, 1), F.A - a, maxdim});
        const int w = std::min(maxdim, b + oversample_);
        RefreshStep st{};
        st.a = a; st.b = b; st.w = w; st.omega_off = (long long)F.host_omega.size();
        const size_t base = F.host_omega.size();
        F.host_omega.resize(base + (size_t)F.q * w);
        F.host_steps.push_back(st);
        F.bmax = std::max(F.bmax, b);
        F.wmax = std::max(F.wmax, w);
        a += b;
      }
      if (F.wmax > 512) throw std::invalid_argument("sketch width above 512 is not supported");
      F.nsteps = int(F.host_steps.size());
      F.off = G.qcat;
      G.qcat += F.q;
      F.boff = G.bcat;
      G.bcat += F.bmax;
      G.max_steps = std::max(G.max_steps, F.nsteps);
      max_steps_ = std::max(max_steps_, F.nsteps);
    }
    draw_sketches();
    // groups in decreasing order of component steps: the groups still active
    // at step s are a prefix of this order
    order_.resize(gplan_.size());
    for (size_t g = 0; g < gplan_.size(); ++g) order_[g] = int(g);
    std::stable_sort(order_.begin(), order_.end(),
                     [&](int a, int b) { return gplan_[a].max_steps > gplan_[b].max_steps; });
    // preprocessing splits (one launch over every group) and the per-step splits
    long long tiles1 = 0;
    int maxn_pre = 0, maxn_step = 0, maxn_xw = 0;
    for (const Group& G : gplan_) {
      tiles1 += xtw_col_tiles(G.p) * ((G.qcat + 63) / 64);
      maxn_pre = std::max(maxn_pre, std::min(G.qcat, 64));
      maxn_step = std::max(maxn_step, std::min(G.bcat, 64));
      maxn_xw = std::max(maxn_xw, std::min(G.bcat, kXwChunk));
    }
    for (Group& G : gplan_) {
      long long s = (2048 + tiles1 - 1) / std::max<long long>(tiles1, 1);
      s = std::max<long long>(1, std::min<long long>(s, G.n / 256));
      // every work-group of a pass streams at most splits_.step_rows rows
      // (the cross-product and loading passes) or splits_.step_cols columns
      // (the score pass) of its group, so work-groups of all groups are alike
      // in size and a launch over groups of very different shapes has no
      // long tail
      s = std::max<long long>(s, (G.n + splits_.step_rows - 1) / splits_.step_rows);
      G.splits = int(std::min<long long>(s, 4096));
      const bool wide = G.bcat >= kMfmaMinN;
      (wide ? has_wide_ : has_narrow_) = true;
      const long long rows = wide ? splits_.wide_rows : splits_.step_rows;
      const long long cols = wide ? splits_.wide_cols : splits_.step_cols;
      G.ls_splits = int(std::min<long long>(4096, (G.n + rows - 1) / rows));
      G.xs_splits = int(std::min<long long>(4096, (G.p + cols - 1) / cols));
    }
    // arena layout; the status words (zero at upload) lead the uploaded arena
    status_off_ = up_.add(sizeof(int) * 4 * (gplan_.size() + plan_.size()));
    constexpr int kGo = 17, kFo = 15;
    std::vector<size_t> go(gplan_.size() * kGo), fo(plan_.size() * kFo);
    for (size_t g = 0; g < gplan_.size(); ++g) {
      const Group& G = gplan_[g];
      const size_t n = G.n, p = G.p, Q = G.qcat, B = G.bcat;
      size_t* o = &go[g * kGo];
      o[0] = arena_.add(8 * n * Q);                    // Yc
      o[1] = arena_.add(8 * Q);                        // ysum
      o[2] = arena_.add(8 * p * Q);                    // S
      o[3] = arena_.add(8 * p * Q);                    // S0
      o[4] = arena_.add(8 * p);                        // mx
      o[5] = arena_.add(8 * p);                        // sx
      o[6] = arena_.add(8 * p);                        // colss
      o[7] = arena_.add(8 * (size_t)G.splits * 2 * p); // stats
      o[8] = arena_.add(8 * (size_t)G.splits * p * Q); // spart
      o[9] = arena_.add(8 * p * B);                    // cand
      o[10] = arena_.add(8 * p * B);                   // cand_scaled
      o[11] = arena_.add(8 * n * B);                   // t_raw
      o[12] = arena_.add(8 * (size_t)G.ls_splits * p * B);  // loads (split slabs)
      o[13] = 0;                                       // flags: in the status block
      o[14] = G.xs_splits > 1 ? arena_.add(8 * (size_t)G.xs_splits * n * B) : 0;  // score slabs
      // operand workspaces: one per right-hand chunk of each pass
      const size_t xw_chunks = (B + kXwChunk - 1) / kXwChunk, xt_chunks = (B + 63) / 64;
      const size_t pre_chunks = (Q + 63) / 64;
      o[15] = arena_.add(xw_chunks * xw_frag_bytes<T>(G.p, maxn_xw, kF64) +
                         xt_chunks * xtw_frag_bytes<T>(G.n, maxn_step, kF64) +
                         pre_chunks * xtw_frag_bytes<T>(G.n, maxn_pre, kF64) +
                         256 * (xw_chunks + xt_chunks + pre_chunks));
      o[16] = G.hold_n > 0 ? arena_.add(8 * 2 * p) : 0;  // held-out sums
    }
    for (size_t f = 0; f < plan_.size(); ++f) {
      const Fit& F = plan_[f];
      const Group& G = gplan_[fits_[f].group];
      const size_t n = G.n, p = G.p, q = F.q, A = F.A;
      size_t* o = &fo[f * kFo];
      o[0] = arena_.add(8 * q);                                   // my
      o[1] = arena_.add(8 * q * q);                               // G
      o[2] = arena_.add(8 * p * A);                               // R
      o[3] = arena_.add(8 * q * A);                               // Q
      o[4] = arena_.add(8 * p * A);                               // V
      o[5] = arena_.add(8 * n * A);                               // T
      o[6] = arena_.add(8 * p * A);                               // Rs
      o[7] = arena_.add(8 * (size_t)simpls_scratch_doubles(G.p, F.q, F.wmax));
      o[8] = 0;                                                   // state: in the status block
      o[9] = arena_.add(sizeof(int) * q);                         // counts
      o[10] = reorth_ == 1 ? arena_.add(8 * p * A) : 0;           // P (re-orthogonalising)
      o[11] = fits_[f].perm && fits_[f].labels ? arena_.add(sizeof(int) * n) : 0;  // permuted labels
      o[12] = up_.add(sizeof(RefreshStep) * F.host_steps.size());  // uploaded
      o[13] = up_.add(8 * F.host_omega.size());                    // uploaded
    }
    devfits_off_ = up_.add(sizeof(SimplsFitDev) * plan_.size());
    center_count_ = 0;
    for (size_t f = 0; f < plan_.size(); ++f)
      if (!fits_[f].labels) center_count_ += plan_[f].q;
    const size_t center_off = up_.add(sizeof(CenterColumn<T>) * (size_t)center_count_);
    const size_t div_off = up_.add(sizeof(DivRows) * plan_.size());
    std::vector<size_t> colflag_off(gplan_.size());
    for (size_t g = 0; g < gplan_.size(); ++g)
      colflag_off[g] = up_.add(sizeof(int) * std::max(gplan_[g].qcat, 1));
    arena_.allocate();
    up_.allocate();
    // host image of the uploaded pieces (states and flags start at zero):
    // one copy for the whole batch
    std::vector<char> image(up_.bytes(), 0);
    std::vector<char*> frag_cursor(gplan_.size());
    for (size_t g = 0; g < gplan_.size(); ++g) {
      Group& G = gplan_[g];
      const size_t* o = &go[g * kGo];
      G.Yc = arena_.at<double>(o[0]); G.ysum = arena_.at<double>(o[1]);
      G.S = arena_.at<double>(o[2]); G.S0 = arena_.at<double>(o[3]);
      G.mx = arena_.at<double>(o[4]); G.sx = arena_.at<double>(o[5]);
      G.colss = arena_.at<double>(o[6]); G.stats = arena_.at<double>(o[7]);
      G.spart = arena_.at<double>(o[8]); G.cand = arena_.at<double>(o[9]);
      G.cand_scaled = arena_.at<double>(o[10]); G.t_raw = arena_.at<double>(o[11]);
      G.loads = arena_.at<double>(o[12]); G.flags = up_.at<int>(status_off_) + 4 * g;
      G.col_flags = up_.at<int>(colflag_off[g]);
      G.tpart = G.xs_splits > 1 ? arena_.at<double>(o[14]) : nullptr;
      G.hstats = G.hold_n > 0 ? arena_.at<double>(o[16]) : nullptr;
      frag_cursor[g] = arena_.at<char>(o[15]);
    }
    auto take_frag = [&](int g, size_t bytes) -> void* {
      if (bytes == 0) return nullptr;
      void* ptr = frag_cursor[g];
      frag_cursor[g] += (bytes + 255) / 256 * 256;
      return ptr;
    };
    std::vector<SimplsFitDev> dev(plan_.size());
    for (size_t f = 0; f < plan_.size(); ++f) {
      Fit& F = plan_[f];
      Group& G = gplan_[fits_[f].group];
      const size_t* o = &fo[f * kFo];
      F.my = arena_.at<double>(o[0]); F.G = arena_.at<double>(o[1]);
      F.R = arena_.at<double>(o[2]); F.Q = arena_.at<double>(o[3]);
      F.V = arena_.at<double>(o[4]); F.Tsc = arena_.at<double>(o[5]);
      F.Rs = arena_.at<double>(o[6]); F.scratch = arena_.at<double>(o[7]);
      F.state = up_.at<int>(status_off_) + 4 * (gplan_.size() + f); F.counts = arena_.at<int>(o[9]);
      F.lperm = fits_[f].perm && fits_[f].labels ? arena_.at<int>(o[11]) : nullptr;
      F.steps = up_.at<RefreshStep>(o[12]);
      F.omega = up_.at<double>(o[13]);
      std::memcpy(image.data() + o[12], F.host_steps.data(), sizeof(RefreshStep) * F.host_steps.size());
      std::memcpy(image.data() + o[13], F.host_omega.data(), sizeof(double) * F.host_omega.size());
      SimplsFitDev& d = dev[f];
      d.n = G.n; d.p = G.p; d.q = F.q; d.A = F.A;
      d.power = power_; d.use_right_gram = F.right_gram ? 1 : 0;
      // src/r_api.cpp:1416: two passes for float, one for double, unless forced
      // one Gram-Schmidt pass unless forced: FP32-stored fits compute in FP64
      // and take the FP64 routing (src/r_api.cpp:1416 doubles the passes
      // only to offset float arithmetic)
      d.reorth = reorth_ >= 0 ? reorth_ : 0;
      d.nsteps = F.nsteps; d.bmax = F.bmax; d.ts_splits = 1; d.ls_splits = G.ls_splits;
      d.ls_slab = (long long)G.p * G.bcat;
      d.steps = F.steps; d.omega = F.omega;
      d.S = G.S + (size_t)F.off * G.p; d.S0 = G.S0 + (size_t)F.off * G.p;
      d.G = F.right_gram ? F.G : nullptr; d.sx = G.sx;
      d.cand = G.cand + (size_t)F.boff * G.p;
      d.cand_scaled = G.cand_scaled + (size_t)F.boff * G.p;
      d.t_raw = G.t_raw + (size_t)F.boff * G.n;
      d.ts_part = nullptr;
      d.ls_part = G.loads + (size_t)F.boff * G.p;
      d.R = F.R; d.Q = F.Q; d.V = F.V; d.T = F.Tsc;
      d.P = reorth_ == 1 ? arena_.at<double>(o[10]) : nullptr;
      d.scratch = F.scratch; d.state = F.state;
    }
    // a non-finite cross-product column flags the fit that owns it
    status_base_ = up_.at<int>(status_off_);
    for (size_t f = 0; f < plan_.size(); ++f) {
      const Fit& F = plan_[f];
      int* cf = reinterpret_cast<int*>(image.data() + colflag_off[fits_[f].group]);
      for (int c = 0; c < F.q; ++c) cf[F.off + c] = int(4 * (gplan_.size() + f) + 2);
    }
    devfits_ = up_.at<SimplsFitDev>(devfits_off_);
    std::memcpy(image.data() + devfits_off_, dev.data(), sizeof(SimplsFitDev) * dev.size());
    // R / sx (the coefficients on the original predictor scale) of every fit
    div_list_ = up_.at<DivRows>(div_off);
    div_max_ = 0;
    DivRows* dl = reinterpret_cast<DivRows*>(image.data() + div_off);
    for (size_t f = 0; f < plan_.size(); ++f) {
      const Fit& F = plan_[f];
      const Group& G = gplan_[fits_[f].group];
      dl[f] = DivRows{F.R, G.sx, (long long)G.p, F.A, F.Rs};
      div_max_ = std::max(div_max_, (long long)G.p * F.A);
    }
    // response columns to centre: column c of fit f goes to column off + c of
    // its group's centred block
    center_list_ = up_.at<CenterColumn<T>>(center_off);
    CenterColumn<T>* cl = reinterpret_cast<CenterColumn<T>*>(image.data() + center_off);
    for (size_t f = 0; f < plan_.size(); ++f) {
      if (fits_[f].labels) continue;
      const Fit& F = plan_[f];
      const Group& G = gplan_[fits_[f].group];
      for (int c = 0; c < F.q; ++c, ++cl) {
        cl->y = static_cast<const T*>(fits_[f].y) + (size_t)c * G.n;
        cl->perm = fits_[f].perm;
        cl->n = G.n;
        cl->yc = G.Yc + (size_t)(F.off + c) * G.n;
        cl->mean = F.my + c;
        cl->ysum = G.ysum + F.off + c;
        cl->mask = G.hold_n > 0 ? G.mask : nullptr;
        cl->used = G.used();
      }
    }
    upload(up_.at<char>(0), image.data(), image.size());
    // task lists of the per-step passes (identical for every step), in group
    // order; prefix bookkeeping for the active groups of each step. Tasks are
    // kept in width classes, each launched on its own: narrow right-hand sides
    // (below kMfmaMinN columns, e.g. a singleton fit's one candidate) on the
    // VALU kernels, wider ones (groups of refits sharing X) on the matrix
    // cores at the width they need, so no task pays for another's columns.
    std::vector<XwTask<T>> xw[kWidthClasses];
    std::vector<XtwTask<T>> xt[kWidthClasses];
    std::vector<SlabSum> sums;
    const size_t ng = order_.size();
    std::vector<int> xw_end[kWidthClasses], xt_end[kWidthClasses], xw_mn[kWidthClasses],
        xt_mn[kWidthClasses];
    for (int c = 0; c < kWidthClasses; ++c) {
      xw_end[c].assign(ng + 1, 0); xt_end[c].assign(ng + 1, 0);
      xw_mn[c].assign(ng + 1, 0); xt_mn[c].assign(ng + 1, 0);
    }
    std::vector<int> sum_end(ng + 1, 0);
    std::vector<long long> sum_max(ng + 1, 0);
    for (size_t gi = 0; gi < ng; ++gi) {
      const int g = order_[gi];
      Group& G = gplan_[g];
      for (int c = 0; c < kWidthClasses; ++c) {
        xw_mn[c][gi + 1] = xw_mn[c][gi];
        xt_mn[c][gi + 1] = xt_mn[c][gi];
      }
      for (int c0 = 0; c0 < G.bcat; c0 += kXwChunk) {
        const int nb = std::min(kXwChunk, G.bcat - c0);
        const int cls = width_class(nb);
        XwTask<T> a{};
        a.A = G.X; a.m = G.n; a.k = G.p; a.center = G.mx;
        a.W = G.cand_scaled + (size_t)c0 * G.p; a.ldw = G.p; a.N = nb; a.splits = G.xs_splits;
        if (G.xs_splits > 1) {
          a.out = G.tpart + (size_t)c0 * G.n; a.slab = (long long)G.n * G.bcat;
        } else {
          a.out = G.t_raw + (size_t)c0 * G.n; a.slab = 0;
        }
        a.ldo = G.n;
        a.frag = take_frag(g, xw_frag_bytes<T>(G.p, maxn_xw, kF64));
        xw[cls].push_back(a);
        xw_mn[cls][gi + 1] = std::max(xw_mn[cls][gi + 1], nb);
      }
      for (int c0 = 0; c0 < G.bcat; c0 += 64) {
        const int nb = std::min(64, G.bcat - c0);
        const int cls = width_class(nb);
        XtwTask<T> b{};
        b.A = G.X; b.m = G.n; b.k = G.p; b.center = G.mx; b.mode = kCentered;
        b.W = G.t_raw + (size_t)c0 * G.n; b.ldw = G.n; b.N = nb; b.splits = G.ls_splits;
        b.out = G.loads + (size_t)c0 * G.p; b.ldo = G.p; b.slab = (long long)G.p * G.bcat;
        b.frag = take_frag(g, xtw_frag_bytes<T>(G.n, maxn_step, kF64));
        xt[cls].push_back(b);
        xt_mn[cls][gi + 1] = std::max(xt_mn[cls][gi + 1], nb);
      }
      if (G.xs_splits > 1) {
        const long long cnt = (long long)G.n * G.bcat;
        sums.push_back(SlabSum{G.tpart, G.t_raw, cnt, cnt, G.xs_splits});
        sum_max[gi + 1] = std::max(sum_max[gi], cnt);
      } else {
        sum_max[gi + 1] = sum_max[gi];
      }
      for (int c = 0; c < kWidthClasses; ++c) {
        xw_end[c][gi + 1] = int(xw[c].size());
        xt_end[c][gi + 1] = int(xt[c].size());
      }
      sum_end[gi + 1] = int(sums.size());
    }
    // tiles of a prefix: tile0 of the first task after it (within a class
    // every task has the same kernel, so its tile count is prefix-independent)
    steps_.assign(ng + 1, StepLaunch{});
    for (int c = 0; c < kWidthClasses; ++c) {
      int mn = 0;
      const long long xw_total = pack_xw(xw[c].data(), int(xw[c].size()), &mn, kF64);
      const long long xt_total = pack_xtw(xt[c].data(), int(xt[c].size()), &mn, kF64);
      // algorithmic bytes of each prefix: A, W and the output once each
      std::vector<double> xwb(xw[c].size() + 1, 0.0), xtb(xt[c].size() + 1, 0.0);
      for (size_t i = 0; i < xw[c].size(); ++i) {
        const XwTask<T>& a = xw[c][i];
        xwb[i + 1] = xwb[i] + double(a.m) * a.k * sizeof(T) + 8.0 * a.k * a.N + 8.0 * a.m * a.N;
      }
      for (size_t i = 0; i < xt[c].size(); ++i) {
        const XtwTask<T>& b = xt[c][i];
        xtb[i + 1] = xtb[i] + double(b.m) * b.k * sizeof(T) + 8.0 * b.m * b.N + 8.0 * b.k * b.N;
      }
      for (size_t na = 1; na <= ng; ++na) {
        auto& P = steps_[na].pass[c];
        P.xw_bytes = xwb[xw_end[c][na]];
        P.xt_bytes = xtb[xt_end[c][na]];
        P.xw_count = xw_end[c][na];
        P.xt_count = xt_end[c][na];
        P.xw_maxn = xw_mn[c][na];
        P.xt_maxn = xt_mn[c][na];
        P.xw_tiles = P.xw_count < int(xw[c].size()) ? xw[c][P.xw_count].tile0 : xw_total;
        P.xt_tiles = P.xt_count < int(xt[c].size()) ? xt[c][P.xt_count].tile0 : xt_total;
      }
      if (!xw[c].empty()) {
        xw_tasks_[c].alloc(sizeof(XwTask<T>) * xw[c].size());
        upload(xw_tasks_[c].as<XwTask<T>>(), xw[c].data(), sizeof(XwTask<T>) * xw[c].size());
      }
      if (!xt[c].empty()) {
        xt_tasks_[c].alloc(sizeof(XtwTask<T>) * xt[c].size());
        upload(xt_tasks_[c].as<XtwTask<T>>(), xt[c].data(), sizeof(XtwTask<T>) * xt[c].size());
      }
    }
    // The one-column loading pass of a step launches its tasks lightest
    // work-groups first (fewest rows per row split), so the heavy ones
    // finish together at the end: one launch over 56 fits of 500..2e4 rows
    // took 0.836 against 0.886 ms in random order (0.2-1% of a 256-fit
    // batch's device time). Each step's prefix of active groups gets its own
    // sorted copy; the tasks write separate slabs, so the order changes no
    // result.
    {
      std::vector<char> used(ng + 1, 0);
      for (int st = 0; st < max_steps_; ++st) {
        size_t na = 0;
        while (na < ng && gplan_[order_[na]].max_steps > st) ++na;
        used[na] = 1;
      }
      std::vector<XtwTask<T>> sorted;
      for (size_t na = 1; na <= ng; ++na) {
        const int cnt = xt_end[0][na];
        if (!used[na] || cnt < 2) continue;
        std::vector<XtwTask<T>> v(xt[0].begin(), xt[0].begin() + cnt);
        std::stable_sort(v.begin(), v.end(), [](const XtwTask<T>& a, const XtwTask<T>& b) {
          return (a.m + a.splits - 1) / a.splits < (b.m + b.splits - 1) / b.splits;
        });
        int mn = 0;
        const long long tiles = pack_xtw(v.data(), cnt, &mn, kF64);
        steps_[na].pass[0].xt_first = (long long)sorted.size();
        steps_[na].pass[0].xt_tiles = tiles;
        sorted.insert(sorted.end(), v.begin(), v.end());
      }
      if (!sorted.empty()) {
        xt_sorted_.alloc(sizeof(XtwTask<T>) * sorted.size());
        upload(xt_sorted_.as<XtwTask<T>>(), sorted.data(), sizeof(XtwTask<T>) * sorted.size());
      }
    }
    for (size_t na = 1; na <= ng; ++na) {
      steps_[na].sum_count = sum_end[na];
      steps_[na].sum_max = sum_max[na];
    }
    if (!sums.empty()) {
      sums_.alloc(sizeof(SlabSum) * sums.size());
      upload(sums_.as<SlabSum>(), sums.data(), sizeof(SlabSum) * sums.size());
    }
    // preprocessing task lists: column statistics + cross-products of every
    // group, one launch per width class
    for (const int g : order_) {
      Group& G = gplan_[g];
      for (int c0 = 0; c0 < G.qcat; c0 += 64) {
        const int nb = std::min(64, G.qcat - c0);
        XtwTask<T> t{};
        t.A = G.X; t.m = G.n; t.k = G.p; t.W = G.Yc + (size_t)c0 * G.n; t.ldw = G.n; t.N = nb;
        t.mode = kShiftStats; t.splits = G.splits;
        t.out = G.spart + (size_t)c0 * G.p; t.ldo = G.p; t.slab = (long long)G.p * G.qcat;
        t.stats = c0 == 0 ? G.stats : nullptr;
        t.frag = take_frag(g, xtw_frag_bytes<T>(G.n, maxn_pre, kF64));
        // several response columns ride the FP64 matrix cores even below
        // kMfmaMinN (the VALU kernel for 2..7 columns is LDS-bound); one
        // column has its own VALU kernel
        pre_[nb <= 1 ? 0 : std::max(1, width_class(nb))].push_back(t);
      }
    }
    for (int c = 0; c < kWidthClasses; ++c) {
      if (pre_[c].empty()) continue;
      for (const XtwTask<T>& t : pre_[c])
        pre_bytes_[c] += double(t.m) * t.k * sizeof(T) + 8.0 * t.m * t.N + 8.0 * t.k * t.N;
      pre_tiles_[c] = pack_xtw(pre_[c].data(), int(pre_[c].size()), &pre_maxn_[c], kF64,
                               c == 0 ? 0 : kMfmaMinN);
      pre_dev_[c].alloc(sizeof(XtwTask<T>) * pre_[c].size());
      upload(pre_dev_[c].as<XtwTask<T>>(), pre_[c].data(), sizeof(XtwTask<T>) * pre_[c].size());
    }
    // right Grams SᵀS of every fit that works in the response space: one
    // task list, launched after the cross-products
    std::vector<XtwTask<double>> gram;
    for (size_t f = 0; f < plan_.size(); ++f) {
      Fit& F = plan_[f];
      if (!F.right_gram) continue;
      Group& G = gplan_[fits_[f].group];
      const double* S = G.S + (size_t)F.off * G.p;
      for (int c0 = 0; c0 < F.q; c0 += 64) {
        const int nb = std::min(64, F.q - c0);
        XtwTask<double> t{};
        t.A = S; t.m = G.p; t.k = F.q; t.W = S + (size_t)c0 * G.p; t.ldw = G.p; t.N = nb;
        t.mode = kPlain; t.splits = 1; t.out = F.G + (size_t)c0 * F.q; t.ldo = F.q; t.slab = 0;
        gram.push_back(t);
      }
    }
    if (!gram.empty()) {
      gram_count_ = int(gram.size());
      gram_tiles_ = pack_xtw(gram.data(), gram_count_, &gram_maxn_);
      gram_dev_.alloc(sizeof(XtwTask<double>) * gram.size());
      upload(gram_dev_.as<XtwTask<double>>(), gram.data(), sizeof(XtwTask<double>) * gram.size());
    }
  }

  // Host-to-device upload of planning data on the batch's own stream, ahead
  // of its launches. (A synchronous copy runs on the null stream, whose
  // hardware queue may be the one another lane's running kernels hold, and
  // would wait behind them.) The host bytes are kept until the batch goes.
  void upload(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    const char* b = static_cast<const char*>(src);
    staged_.emplace_back(b, b + bytes);
    FPLS_HIP(hipMemcpyAsync(dst, staged_.back().data(), bytes, hipMemcpyHostToDevice, stream_));
  }

  // The whole execution as launches on stream_ (no host synchronisation, no
  // memory set or copy), so it can be captured into a graph and replayed.
  void run() {
    // group flags and fit states start at zero (a replay clears what the
    // previous execution left; a fill kernel, since memset nodes are kept
    // out of graphs)
    launch_fill_int(status_base_, int(4 * (gplan_.size() + plan_.size())), 0, stream_);
    // responses of every fit into its columns of the group's centred block
    for (size_t f = 0; f < plan_.size(); ++f) {
      Fit& F = plan_[f];
      Group& G = gplan_[fits_[f].group];
      const FitIn& in = fits_[f];
      double* yc = G.Yc + (size_t)F.off * G.n;
      if (in.labels) {
        const int* lab = in.labels;
        if (in.perm) {
          launch_gather_rows<int>(in.labels, G.n, 1, in.perm, G.n, F.lperm, stream_);
          lab = F.lperm;
        }
        launch_label_response(lab, G.n, F.q, F.counts, F.my, yc, nullptr, G.ysum + F.off,
                              F.state + 2, stream_);
      }
    }
    // every response column of every fit: one launch
    launch_center_columns<T>(center_list_, center_count_, stream_);
    // one launch: column statistics + cross-products of every group
    for (int c = 0; c < kWidthClasses; ++c)
      if (!pre_[c].empty())
        timed(0, pre_bytes_[c], [&] {
          launch_xtw<T>(pre_dev_[c].as<XtwTask<T>>(), int(pre_[c].size()), pre_tiles_[c],
                        pre_maxn_[c], stream_, kF64, /*stats=*/true);
        });
    for (Group& G : gplan_) {
      // a row view: the pass ran over every row (zero responses on the
      // held-out ones); its column sums lose the held-out rows'
      if (G.hold_n > 0)
        launch_holdout_stats<T>(G.hold_x, G.hold_n, G.p, G.X, G.n, G.hstats, stream_);
      launch_finalize_crossprod<T>(G.X, G.n, G.p, G.qcat, scaling_, G.stats, G.spart, G.splits,
                                   G.ysum, G.mx, G.sx, G.colss, G.S, G.S0, G.flags, stream_,
                                   status_base_, G.col_flags, G.hstats,
                                   G.hold_n > 0 ? G.used() : -1);
    }
    // right Grams SᵀS of every fit: one launch
    if (gram_count_ > 0)
      launch_xtw<double>(gram_dev_.as<XtwTask<double>>(), gram_count_, gram_tiles_, gram_maxn_,
                         stream_);
    // component steps: refresh, score pass, loading pass, update over the
    // groups still active at the step (a prefix of the step order)
    const int nf = int(plan_.size());
    for (int s = 0; s < max_steps_; ++s) {
      int na = 0;
      while (na < int(order_.size()) && gplan_[order_[na]].max_steps > s) ++na;
      const StepLaunch& L = steps_[na];
      timed(1, 0.0, [&] { launch_simpls_refresh(devfits_, nf, s, stream_); });
      for (int c = 0; c < kWidthClasses; ++c) {
        const auto& P = L.pass[c];
        if (P.xw_count > 0)
          timed(c == 0 ? 2 : 3, P.xw_bytes, [&] {
            launch_xw<T>(xw_tasks_[c].as<XwTask<T>>(), P.xw_count, P.xw_tiles, P.xw_maxn,
                         stream_, kF64);
          });
      }
      if (L.sum_count > 0)
        timed(4, 0.0, [&] {
          launch_sum_slabs_list(sums_.as<SlabSum>(), L.sum_count, L.sum_max, stream_);
        });
      // row views: the held-out rows' scores are zeroed, so that the loading
      // pass and the score norms sum the view's rows only
      for (const Group& G : gplan_)
        if (G.hold_n > 0 && G.max_steps > s)
          launch_zero_rows(G.t_raw, G.n, G.bcat, G.hold, G.hold_n, stream_);
      for (int c = 0; c < kWidthClasses; ++c) {
        const auto& P = L.pass[c];
        if (P.xt_count > 0)
          timed(c == 0 ? 5 : 6, P.xt_bytes, [&] {
            launch_xtw<T>(P.xt_first >= 0 ? xt_sorted_.as<XtwTask<T>>() + P.xt_first
                                          : xt_tasks_[c].as<XtwTask<T>>(),
                          P.xt_count, P.xt_tiles, P.xt_maxn,
                          stream_, kF64);
          });
      }
      timed(7, 0.0, [&] { launch_simpls_update(devfits_, nf, s, stream_); });
    }
    // R / sx of every fit: one launch
    launch_div_rows_list(div_list_, int(plan_.size()), div_max_, stream_);
  }

  // Runs `launch`; with phase timing on, between two events on the stream.
  template<class F>
  void timed(int phase, double bytes, F&& launch) {
    if (!timing_) {
      launch();
      return;
    }
    TimedLaunch t{phase, bytes, nullptr, nullptr};
    FPLS_HIP(hipEventCreate(&t.a));
    timed_.push_back(t);
    FPLS_HIP(hipEventCreate(&timed_.back().b));
    FPLS_HIP(hipEventRecord(timed_.back().a, stream_));
    launch();
    FPLS_HIP(hipEventRecord(timed_.back().b, stream_));
  }

  void collect_phases() {
    for (TimedLaunch& t : timed_) {
      float ms = 0.f;
      FPLS_HIP(hipEventElapsedTime(&ms, t.a, t.b));
      phase_[t.phase][0] += double(ms) * 1e-3;
      phase_[t.phase][1] += t.bytes;
      phase_[t.phase][2] += 1.0;
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.b);
    }
    timed_.clear();
  }

  void collect() {
    // group flags then fit states, one copy
    const size_t ng = gplan_.size(), nf = plan_.size();
    std::vector<int> status(4 * (ng + nf));
    // on the batch's stream (after its work), not behind another lane's
    FPLS_HIP(hipMemcpyAsync(status.data(), up_.at<int>(status_off_), sizeof(int) * status.size(),
                            hipMemcpyDeviceToHost, stream_));
    FPLS_HIP(hipStreamSynchronize(stream_));
    completed_.assign(nf, 0);
    flags_.assign(nf, 0);
    for (size_t f = 0; f < nf; ++f) {
      const int* st = &status[4 * (ng + f)];
      completed_[f] = st[0];
      flags_[f] = st[2] | status[4 * fits_[f].group];
    }
  }

  struct StepLaunch {
    struct Pass {  // one width class of the score and loading passes
      int xw_count = 0, xt_count = 0, xw_maxn = 0, xt_maxn = 0;
      long long xw_tiles = 0, xt_tiles = 0;
      long long xt_first = -1;  // the loading pass's own task order (xt_sorted_), or -1
      double xw_bytes = 0.0, xt_bytes = 0.0;  // algorithmic bytes of the launch
    } pass[kWidthClasses];
    int sum_count = 0;
    long long sum_max = 0;
  };

  std::vector<GroupIn> groups_;
  std::vector<FitIn> fits_;
  std::vector<Group> gplan_;
  std::vector<Fit> plan_;
  std::vector<int> order_;
  std::vector<StepLaunch> steps_;
  int scaling_, oversample_, power_, reorth_;
  int max_steps_ = 0;
  Arena arena_;
  Arena up_;  // pieces filled from the host: schedules, sketches, fit descriptors
  CenterColumn<T>* center_list_ = nullptr;
  DivRows* div_list_ = nullptr;
  long long div_max_ = 0;
  int center_count_ = 0;
  int* status_base_ = nullptr;
  size_t devfits_off_ = 0;
  size_t status_off_ = 0;  // int[4] flags per group, then int[4] state per fit
  SimplsFitDev* devfits_ = nullptr;
  Buf xw_tasks_[kWidthClasses], xt_tasks_[kWidthClasses], xt_sorted_, sums_, proj_frag_;
  Buf pre_dev_[kWidthClasses];
  std::vector<XtwTask<T>> pre_[kWidthClasses];
  Buf gram_dev_;
  int gram_count_ = 0, gram_maxn_ = 0;
  long long gram_tiles_ = 0;
  hipGraphExec_t graph_ = nullptr;  // run() captured by the first rerun
  std::vector<std::vector<char>> staged_;  // host sources of the uploads
  BatchSplits splits_;
  bool has_wide_ = false, has_narrow_ = false;
  long long pre_tiles_[kWidthClasses] = {};
  int pre_maxn_[kWidthClasses] = {};
  double pre_bytes_[kWidthClasses] = {};
  struct TimedLaunch {
    int phase;
    double bytes;
    hipEvent_t a, b;
  };
  bool timing_ = false;
  std::vector<TimedLaunch> timed_;
  double phase_[kPhases][3] = {};
  // run() captured into graph_ on the batch's stream
  void capture() {
    hipGraph_t g = nullptr;
    FPLS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      run();
    } catch (...) {
      hipGraph_t partial = nullptr;
      (void)hipStreamEndCapture(stream_, &partial);
      if (partial) (void)hipGraphDestroy(partial);
      throw;
    }
    FPLS_HIP(hipStreamEndCapture(stream_, &g));
    const hipError_t st = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    FPLS_HIP(st);
  }

  hipStream_t stream_ = nullptr;
  bool own_stream_ = true;
  Event start_, stop_;   // around the last launch()
  double seconds_ = 0.0;
  std::vector<int> completed_, flags_;
};

class BatchHandle {
 public:
  virtual ~BatchHandle() = default;
  virtual int count() const = 0;
  virtual double seconds() const = 0;
  virtual int completed(int f) const = 0;
  virtual int flags(int f) const = 0;
  virtual void export_field(int f, int field, void* out, size_t elements, bool f64) = 0;
  virtual void phases(double* out, int count) const = 0;
  virtual void rerun() = 0;
  virtual void autotune(int trials, double* report, int count) = 0;
  virtual BatchSplits splits() const = 0;
  virtual int lanes() const = 0;
};

// A batch split into lanes: the shared-X groups are dealt to 1..4 sub-batches
// (largest estimated work first, each to the lane with the least so far),
// and each lane runs on its own stream of a stream set, so the lanes share
// no hardware queue and run at the same time. The per-fit kernels of a lane
// (refresh, update: one work-group per fit, latency-bound) then overlap the
// bandwidth-bound passes of the others instead of leaving the device idle
// between passes. Every fit's arithmetic is the same as in one lane. The lane
// count is 2 for batches of 32 fits or more over at least two groups, 1
// otherwise, FASTPLS_HIP_BATCH_LANES when set, and 1 with phase timing on
// (per-launch events of concurrent lanes would overlap).
template<class T>
class TypedBatch final : public BatchHandle {
 public:
  TypedBatch(const std::vector<GroupIn>& groups, const std::vector<FitIn>& fits, int scaling,
             int oversample, int power, int reorth)
      : groups_(groups), fits_(fits), scaling_(scaling), oversample_(oversample), power_(power),
        reorth_(reorth) {
    if (groups_.empty() || fits_.empty()) throw std::invalid_argument("empty batch");
    for (const FitIn& f : fits_)
      if (f.group < 0 || f.group >= int(groups_.size()))
        throw std::invalid_argument("fit refers to a missing group");
    partition(lane_count());
    plan_ = build_launch(default_splits());
  }
  int count() const override { return int(fits_.size()); }
  double seconds() const override { return plan_->seconds; }
  int completed(int f) const override { return lane(f).completed(local_[size_t(f)]); }
  int flags(int f) const override { return lane(f).flags(local_[size_t(f)]); }
  void export_field(int f, int field, void* out, size_t elements, bool f64) override {
    if (f < 0 || f >= count()) throw std::invalid_argument("fit index outside the batch");
    lane(f).export_field(local_[size_t(f)], field, out, elements, f64);
  }
  // Phase times of a one-lane batch (phase timing runs one lane).
  void phases(double* out, int count) const override {
    if (plan_->b.size() == 1) {
      plan_->b[0]->phases(out, count);
      return;
    }
    for (int i = 0; i < count; ++i) out[i] = 0.0;
    if (count > 0) out[0] = plan_->seconds;
  }
  void rerun() override { execute(*plan_, true); }
  BatchSplits splits() const override { return plan_->b[0]->splits(); }
  int lanes() const override { return int(plan_->b.size()); }

  // Chooses the split lengths for this batch by measurement: starting from
  // the batch's own lengths, each length in turn is tried at 1/4, 1/2, 2 and
  // 4 times its current value (within [kMinSplit, kMaxSplit]; the wide
  // lengths only when some group runs on the matrix cores, the others only
  // when some runs on the VALU), re-planning and running the whole batch
  // (all its lanes) `trials` times per candidate and keeping a candidate that
  // lowers the fastest device time by more than 1%. The batch then holds the
  // fits of the chosen plan (the same fits; split sums differ in rounding
  // only) and the choice becomes the process default for new batches.
  // report: [0..3] the chosen lengths (step_rows, step_cols, wide_rows,
  // wide_cols), [4] the chosen plan's device seconds, [5] the starting
  // plan's, [6] plans run.
  void autotune(int trials, double* report, int count) override {
    if (trials < 1 || trials > 16) throw std::invalid_argument("autotune trials must lie in 1..16");
    int runs = 0;
    auto measure = [&](const BatchSplits& cfg, std::unique_ptr<Plan>& keep) {
      double best = std::numeric_limits<double>::infinity();
      for (int t = 0; t < trials; ++t) {
        std::unique_ptr<Plan> P = build(cfg);
        execute(*P, false);
        ++runs;
        if (P->seconds < best) {
          best = P->seconds;
          keep = std::move(P);
        }
      }
      return best;
    };
    BatchSplits best = plan_->b[0]->splits();
    plan_.reset();
    std::unique_ptr<Plan> best_plan;
    double best_s = measure(best, best_plan);
    const double start_s = best_s;
    bool narrow = false, wide = false;
    for (const auto& b : best_plan->b) {
      narrow = narrow || b->has_narrow();
      wide = wide || b->has_wide();
    }
    for (int i = 0; i < 4; ++i) {
      if ((i < 2 && !narrow) || (i >= 2 && !wide)) continue;
      const long long base = best.get(i);
      for (const double f : {0.25, 0.5, 2.0, 4.0}) {
        const long long v = (long long)(double(base) * f);
        if (v < kMinSplit || v > kMaxSplit) continue;
        BatchSplits cfg = best;
        cfg.at(i) = v;
        std::unique_ptr<Plan> cand;
        const double t = measure(cfg, cand);
        if (t < best_s * 0.99) {
          best = cfg;
          best_s = t;
          best_plan = std::move(cand);
        }
      }
    }
    plan_ = std::move(best_plan);
    set_default_splits(best);
    const double out[7] = {double(best.step_rows), double(best.step_cols), double(best.wide_rows),
                           double(best.wide_cols), best_s, start_s, double(runs)};
    for (int i = 0; i < count && i < 7; ++i) report[i] = out[i];
  }

 private:
  // The lanes of one plan (split lengths) and their streams.
  struct Plan {
    std::vector<hipStream_t> streams;   // a stream set when there are several lanes
    std::vector<std::unique_ptr<SimplsBatch<T>>> b;
    double seconds = 0.0;
    ~Plan() {
      b.clear();   // each drains its stream before its buffers go
      stream_set_give(streams);
    }
  };

  int lane_count() const {
    const char* pt = std::getenv("FASTPLS_BENCH_PHASE_TIMING");
    if (pt != nullptr && std::strcmp(pt, "1") == 0) return 1;
    int used = 0;
    std::vector<char> has(groups_.size(), 0);
    for (const FitIn& f : fits_)
      if (!has[size_t(f.group)]) {
        has[size_t(f.group)] = 1;
        ++used;
      }
    int L = fits_.size() >= 32 ? 2 : 1;
    if (const char* v = std::getenv("FASTPLS_HIP_BATCH_LANES")) {
      L = std::atoi(v);
      if (L < 1 || L > std::min(4, hardware_queues()))
        throw std::invalid_argument("FASTPLS_HIP_BATCH_LANES must lie in 1..min(4, GPU_MAX_HW_QUEUES)");
    }
    return std::max(1, std::min(L, used));
  }

  // Groups to lanes by estimated work (X elements times components over the
  // group's fits), largest first, each to the least-loaded lane; groups no
  // fit uses stay with lane 0. Fits keep their order within a lane.
  void partition(int L) {
    const size_t ng = groups_.size();
    std::vector<double> work(ng, 0.0);
    std::vector<char> used(ng, 0);
    for (const FitIn& f : fits_) {
      const GroupIn& g = groups_[size_t(f.group)];
      work[size_t(f.group)] += double(g.n) * double(g.p) * double(std::max(f.A, 1));
      used[size_t(f.group)] = 1;
    }
    std::vector<int> order(ng);
    for (size_t g = 0; g < ng; ++g) order[g] = int(g);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return work[size_t(a)] > work[size_t(b)]; });
    std::vector<int> lane_of_group(ng, 0), local_group(ng, 0);
    std::vector<double> load(size_t(L), 0.0);
    lane_groups_.assign(size_t(L), {});
    for (int g : order) {
      int l = 0;
      if (used[size_t(g)])
        for (int k = 1; k < L; ++k)
          if (load[size_t(k)] < load[size_t(l)]) l = k;
      load[size_t(l)] += work[size_t(g)];
      lane_of_group[size_t(g)] = l;
      local_group[size_t(g)] = int(lane_groups_[size_t(l)].size());
      lane_groups_[size_t(l)].push_back(groups_[size_t(g)]);
    }
    lane_fits_.assign(size_t(L), {});
    lane_of_.resize(fits_.size());
    local_.resize(fits_.size());
    for (size_t f = 0; f < fits_.size(); ++f) {
      const int l = lane_of_group[size_t(fits_[f].group)];
      FitIn in = fits_[f];
      in.group = local_group[size_t(fits_[f].group)];
      lane_of_[f] = l;
      local_[f] = int(lane_fits_[size_t(l)].size());
      lane_fits_[size_t(l)].push_back(in);
    }
  }

  std::unique_ptr<Plan> build(const BatchSplits& cfg) {
    auto P = std::make_unique<Plan>();
    const int L = int(lane_groups_.size());
    if (L > 1) P->streams = stream_set_take(L);
    for (int l = 0; l < L; ++l)
      P->b.push_back(std::make_unique<SimplsBatch<T>>(
          lane_groups_[size_t(l)], lane_fits_[size_t(l)], scaling_, oversample_, power_, reorth_,
          cfg, L > 1 ? P->streams[size_t(l)] : nullptr, true));
    return P;
  }

  // The first execution of a plan: each lane is planned and at once launched,
  // so the host planning of the later lanes (sizes, sketches, uploads)
  // overlaps the device work of the earlier ones instead of delaying the
  // first launch. The device seconds run from one event recorded before the
  // first launch to the last lane's end.
  std::unique_ptr<Plan> build_launch(const BatchSplits& cfg) {
    const int L = int(lane_groups_.size());
    if (L == 1) {
      auto P = build(cfg);
      execute(*P, false);
      return P;
    }
    Event fork;  // declared first: destroyed after the plan has drained its streams
    auto P = std::make_unique<Plan>();
    P->streams = stream_set_take(L);
    for (int l = 0; l < L; ++l) {
      P->b.push_back(std::make_unique<SimplsBatch<T>>(
          lane_groups_[size_t(l)], lane_fits_[size_t(l)], scaling_, oversample_, power_, reorth_,
          cfg, P->streams[size_t(l)], true));
      if (l == 0) FPLS_HIP(hipEventRecord(fork.e, P->streams[0]));
      P->b.back()->launch(fork.e, false);
    }
    double s = 0.0;
    for (auto& b : P->b) {
      b->finish();
      s = std::max(s, b->seconds_since(fork.e));
    }
    P->seconds = s;
    return P;
  }

  // Every lane launched (after one event, so the device time counts from a
  // common start), then waited for; the plan's device seconds run from that
  // event to the last lane's end.
  static void execute(Plan& P, bool graph) {
    if (P.b.size() == 1) {
      P.b[0]->launch(nullptr, graph);
      P.b[0]->finish();
      P.seconds = P.b[0]->seconds();
      return;
    }
    Event fork;
    FPLS_HIP(hipEventRecord(fork.e, P.streams[0]));
    for (auto& b : P.b) b->launch(fork.e, graph);
    double s = 0.0;
    for (auto& b : P.b) {
      b->finish();
      s = std::max(s, b->seconds_since(fork.e));
    }
    P.seconds = s;
  }

  const SimplsBatch<T>& lane(int f) const {
    if (f < 0 || f >= int(fits_.size())) throw std::invalid_argument("fit index outside the batch");
    return *plan_->b[size_t(lane_of_[size_t(f)])];
  }
  SimplsBatch<T>& lane(int f) {
    if (f < 0 || f >= int(fits_.size())) throw std::invalid_argument("fit index outside the batch");
    return *plan_->b[size_t(lane_of_[size_t(f)])];
  }

  std::vector<GroupIn> groups_;
  std::vector<FitIn> fits_;
  int scaling_, oversample_, power_, reorth_;
  std::vector<std::vector<GroupIn>> lane_groups_;
  std::vector<std::vector<FitIn>> lane_fits_;
  std::vector<int> lane_of_, local_;   // per fit: its lane and index there
  std::unique_ptr<Plan> plan_;
};

// ------------------------------------------------------------------- CV
// src/cuda_resident_api.cuh:770-1023: regression CV with per-fold fits, test
// predictions for every prefix, SSE / training-mean SS per fold and the
// metric ladder (2 = observed R², 3 = Q², 4 = RMSD). All folds are fitted
// in one batch instead of one after another.
template<class T>
void simpls_cv_regression(const T* predictors, const T* responses, const int* folds, int n, int p,
                          int q, const int* prefixes, int prefix_count, int scaling, int metric,
                          int oversample, int power, unsigned long long seed,
                          bool store_predictions, double* prediction_output, int* status,
                          double* metrics, double* q2_values, double* rmsd_values,
                          double* observed_r2_values) {
  if (!predictors || !responses || !folds || !prefixes || !status || !metrics || !q2_values ||
      !rmsd_values || !observed_r2_values || n < 2 || p < 1 || q < 1 || prefix_count < 1 ||
      scaling < 1 || scaling > 3 || metric < 2 || metric > 4)
    throw std::invalid_argument("invalid resident HIP regression CV input");
  const int maxp = cv_path_maximum(prefixes, prefix_count);
  const FoldSplit split(folds, n);
  const int fold_count = split.count;
  if (maxp > std::min(n - 1, p) || (store_predictions && !prediction_output))
    throw std::invalid_argument("invalid HIP regression CV output or rank");
  split.require_usable();
  const std::vector<std::vector<int>>& test_rows = split.test;
  const long double total_ss = cv_total_ss(responses, n, q);
  hipStream_t stream = stream_take();
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { stream_give(s); }
  } guard_stream{stream};
  Buf xfull, yfull, idx;
  xfull.alloc(sizeof(T) * (size_t)n * p);
  yfull.alloc(sizeof(T) * (size_t)n * q);
  FPLS_HIP(hipMemcpy(xfull.as<T>(), predictors, sizeof(T) * (size_t)n * p, hipMemcpyHostToDevice));
  FPLS_HIP(hipMemcpy(yfull.as<T>(), responses, sizeof(T) * (size_t)n * q, hipMemcpyHostToDevice));
  std::vector<int> all_idx;
  std::vector<size_t> te_off(fold_count);
  for (int f = 0; f < fold_count; ++f) {
    te_off[f] = all_idx.size();
    all_idx.insert(all_idx.end(), test_rows[f].begin(), test_rows[f].end());
  }
  idx.alloc(sizeof(int) * all_idx.size());
  FPLS_HIP(hipMemcpy(idx.as<int>(), all_idx.data(), sizeof(int) * all_idx.size(), hipMemcpyHostToDevice));
  // Folds are fitted together in batches of as many folds as fit. A fold is a
  // row view of the one resident X: its fit runs over every row of X with
  // zero responses and zero scores on its held-out rows, and its column
  // statistics are the full ones minus the held-out rows'. A fold holds a
  // gathered copy of its held-out rows only (for those statistics and its
  // test predictions), a row mask and its fit's device state. The reference
  // fits one fold at a time on a gathered training copy
  // (src/cuda_resident_api.cuh:770-1023); here the folds of a sub-batch
  // share every launch, and a sub-batch holds at most 2/5 of the free device
  // memory (as the reference's query workspace rule, :259-297), down to one
  // fold per sub-batch.
  // FASTPLS_HIP_CV_FREE_BYTES, when set, stands for the free memory.
  std::vector<int> usable;
  for (int f = 0; f < fold_count; ++f) {
    if (test_rows[f].empty()) {
      status[f] = 2;
      continue;
    }
    if (maxp > std::min(n - int(test_rows[f].size()) - 1, p))
      throw std::invalid_argument("HIP CV component count exceeds a fold rank");
    usable.push_back(f);
  }
  if (usable.empty()) throw std::invalid_argument("HIP CV has no non-empty fold");
  size_t max_test = 0;
  for (int f : usable) max_test = std::max(max_test, test_rows[f].size());
  const double fold_bytes =
      1.25 * (double(sizeof(T)) * double(max_test) * double(p + q) + 4.0 * double(n) +
              8.0 * double(n) * double(2 * maxp + 2 * q + 4) +
              8.0 * double(p) * double(4 * maxp + 3 * q + 96) + 8.0 * double(q) * double(q));
  size_t free_b = 0, total_b = 0;
  if (const char* v = std::getenv("FASTPLS_HIP_CV_FREE_BYTES"))
    free_b = size_t(std::strtoull(v, nullptr, 10));
  else
    FPLS_HIP(hipMemGetInfo(&free_b, &total_b));
  const int per_batch = int(std::max<double>(
      1.0, std::min<double>(double(usable.size()), std::floor(double(free_b) * 0.4 / fold_bytes))));
  RegressionTally tally(prefix_count);
  for (size_t b0 = 0; b0 < usable.size(); b0 += size_t(per_batch)) {
    const size_t b1 = std::min(usable.size(), b0 + size_t(per_batch));
    std::vector<std::unique_ptr<Buf>> masks, xte, yte;
    std::vector<GroupIn> groups;
    std::vector<FitIn> fits;
    std::vector<int> fold_of;
    std::vector<int> host_mask(n);
    for (size_t u = b0; u < b1; ++u) {
      const int f = usable[u];
      const int nte = int(test_rows[f].size());
      for (auto* v : {&masks, &xte, &yte}) v->emplace_back(new Buf);
      masks.back()->alloc(sizeof(int) * (size_t)n);
      xte.back()->alloc(sizeof(T) * (size_t)nte * p);
      yte.back()->alloc(sizeof(T) * (size_t)nte * q);
      std::fill(host_mask.begin(), host_mask.end(), 1);
      for (int r : test_rows[f]) host_mask[size_t(r)] = 0;
      FPLS_HIP(hipMemcpyAsync(masks.back()->as<int>(), host_mask.data(), sizeof(int) * (size_t)n,
                              hipMemcpyHostToDevice, stream));
      FPLS_HIP(hipStreamSynchronize(stream));  // host_mask is refilled for the next fold
      launch_gather_rows<T>(xfull.as<T>(), n, p, idx.as<int>() + te_off[f], nte, xte.back()->as<T>(), stream);
      launch_gather_rows<T>(yfull.as<T>(), n, q, idx.as<int>() + te_off[f], nte, yte.back()->as<T>(), stream);
      GroupIn gi{};
      gi.x = xfull.as<T>(); gi.n = n; gi.p = p;
      gi.mask = masks.back()->as<int>();
      gi.hold = idx.as<int>() + te_off[f];
      gi.hold_x = xte.back()->as<T>();
      gi.hold_n = nte;
      groups.push_back(gi);
      FitIn fi{};
      fi.group = int(groups.size()) - 1; fi.q = q; fi.A = maxp; fi.y = yfull.as<T>();
      fi.labels = nullptr; fi.perm = nullptr; fi.seed = seed;  // SIMPLS folds share the seed
      fits.push_back(fi);
      fold_of.push_back(f);
    }
    FPLS_HIP(hipStreamSynchronize(stream));
    SimplsBatch<T> batch(groups, fits, scaling, oversample, power, -1);
    for (int i = 0; i < batch.count(); ++i)
      if (batch.completed(i) < maxp || batch.flags(i))
        throw std::runtime_error("resident SIMPLS failed: invalid labels, decomposition or component norm");
    Buf scores, part, pred, sums, flag;
    sums.alloc(sizeof(double) * 3 * q);
    flag.alloc(sizeof(int));
    hipStream_t bs = batch.stream();
    for (int i = 0; i < batch.count(); ++i) {
      const int f = fold_of[i];
      const int nte = int(test_rows[f].size());
      scores.reserve(sizeof(double) * (size_t)nte * maxp);
      pred.reserve(sizeof(double) * (size_t)nte * q);
      batch.project_device(i, xte[i]->as<T>(), nte, maxp, scores.d(), part);
      std::vector<double> host_metric(3 * (size_t)q * prefix_count);
      std::vector<double> host_pred(store_predictions ? (size_t)nte * q * prefix_count : 0);
      for (int j = 0; j < prefix_count; ++j) {
        batch.predict_from_scores(i, scores.d(), nte, prefixes[j], pred.d());
        if (store_predictions)
          FPLS_HIP(hipMemcpyAsync(host_pred.data() + (size_t)j * nte * q, pred.d(),
                                  sizeof(double) * (size_t)nte * q, hipMemcpyDeviceToHost, bs));
        launch_fill_int(flag.as<int>(), 1, 0, bs);
        launch_response_sums<T>(pred.d(), yte[i]->as<T>(), nullptr, batch.y_mean(i), nte, q,
                                sums.d(), flag.as<int>(), bs);
        FPLS_HIP(hipMemcpyAsync(host_metric.data() + (size_t)j * 3 * q, sums.d(),
                                sizeof(double) * 3 * q, hipMemcpyDeviceToHost, bs));
        FPLS_HIP(hipStreamSynchronize(bs));
      }
      tally.add_fold(host_metric.data(), q, nte);
      if (store_predictions)
        scatter_predictions(host_pred.data(), test_rows[f], n, q, prefix_count, prediction_output);
      status[f] = 1;
    }
  }
  tally.finish(total_ss, metric, metrics, q2_values, rmsd_values, observed_r2_values);
}

}  // namespace
}  // namespace fpls

using fpls::BatchHandle;

extern "C" {

void* fastpls_hip_simpls_batch_create(int precision, int group_count, const void* const* x_device,
                                      const int* n, const int* p, int fit_count,
                                      const int* group, const void* const* y_device,
                                      const int* const* labels_device,
                                      const int* const* perm_device, const int* q,
                                      const int* components, const unsigned long long* seeds,
                                      int scaling, int oversample, int power, int reorth,
                                      char* error, size_t error_capacity) {
  BatchHandle* h = nullptr;
  fpls::guarded(error, error_capacity, [&] {
    fpls::require_device();
    if (group_count < 1 || fit_count < 1 || !x_device || !n || !p || !group || !q ||
        !components || !seeds)
      throw std::invalid_argument("invalid batch description");
    std::vector<fpls::GroupIn> groups(group_count);
    for (int g = 0; g < group_count; ++g) groups[g] = {x_device[g], n[g], p[g]};
    std::vector<fpls::FitIn> fits(fit_count);
    for (int f = 0; f < fit_count; ++f) {
      fits[f].group = group[f];
      fits[f].q = q[f];
      fits[f].A = components[f];
      fits[f].y = y_device ? y_device[f] : nullptr;
      fits[f].labels = labels_device ? labels_device[f] : nullptr;
      fits[f].perm = perm_device ? perm_device[f] : nullptr;
      fits[f].seed = seeds[f];
    }
    if (precision == 32)
      h = new fpls::TypedBatch<float>(groups, fits, scaling, oversample, power, reorth);
    else if (precision == 64)
      h = new fpls::TypedBatch<double>(groups, fits, scaling, oversample, power, reorth);
    else
      throw std::invalid_argument("precision must be 32 or 64");
  });
  return h;
}

int fastpls_hip_batch_status(void* batch, int* completed, int* flags, double* seconds,
                             char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!batch) throw std::invalid_argument("null batch");
    BatchHandle* b = static_cast<BatchHandle*>(batch);
    for (int f = 0; f < b->count(); ++f) {
      if (completed) completed[f] = b->completed(f);
      if (flags) flags[f] = b->flags(f);
    }
    if (seconds) *seconds = b->seconds();
  });
}

int fastpls_hip_batch_phases(void* batch, double* out, int count, char* error,
                             size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!batch || !out || count < 1) throw std::invalid_argument("invalid batch phase request");
    static_cast<BatchHandle*>(batch)->phases(out, count);
  });
}

int fastpls_hip_batch_export(void* batch, int fit, int field, void* out, size_t elements,
                             char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!batch) throw std::invalid_argument("null batch");
    static_cast<BatchHandle*>(batch)->export_field(fit, field, out, elements, false);
  });
}

int fastpls_hip_batch_export64(void* batch, int fit, int field, double* out, size_t elements,
                             char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!batch) throw std::invalid_argument("null batch");
    static_cast<BatchHandle*>(batch)->export_field(fit, field, out, elements, true);
  });
}

int fastpls_hip_batch_autotune(void* batch, int trials, double* report, int report_count,
                               char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!batch || (report_count > 0 && !report)) throw std::invalid_argument("invalid autotune request");
    static_cast<BatchHandle*>(batch)->autotune(trials, report, report_count);
  });
}

int fastpls_hip_batch_splits(void* batch, long long* lengths, char* error,
                             size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!lengths) throw std::invalid_argument("null split output");
    const fpls::BatchSplits b =
        batch ? static_cast<BatchHandle*>(batch)->splits() : fpls::default_splits();
    for (int i = 0; i < 4; ++i) lengths[i] = b.get(i);
  });
}

int fastpls_hip_batch_lanes(void* batch, int* lanes, char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!batch || !lanes) throw std::invalid_argument("null batch or lane output");
    *lanes = static_cast<BatchHandle*>(batch)->lanes();
  });
}

int fastpls_hip_batch_set_splits(const long long* lengths, char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!lengths || !fpls::valid_splits(lengths))
      throw std::invalid_argument("split lengths must lie in 128..4096");
    fpls::BatchSplits b;
    for (int i = 0; i < 4; ++i) b.at(i) = lengths[i];
    fpls::set_default_splits(b);
  });
}

int fastpls_hip_batch_rerun(void* batch, char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!batch) throw std::invalid_argument("null batch");
    static_cast<BatchHandle*>(batch)->rerun();
  });
}

void fastpls_hip_batch_destroy(void* batch) { delete static_cast<BatchHandle*>(batch); }

int fastpls_hip_simpls_cv_regression(const void* predictors, const void* responses,
                                     const int* folds, int precision, int n, int p, int q,
                                     const int* prefixes, int prefix_count, int scaling,
                                     int metric, int oversample, int power,
                                     unsigned long long seed, int store_predictions,
                                     double* predictions, int* status, double* metrics,
                                     double* q2, double* rmsd, double* observed_r2, char* error,
                                     size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    fpls::require_device();
    if (precision == 32)
      fpls::simpls_cv_regression<float>(static_cast<const float*>(predictors),
          static_cast<const float*>(responses), folds, n, p, q, prefixes, prefix_count, scaling,
          metric, oversample, power, seed, store_predictions != 0, predictions, status, metrics,
          q2, rmsd, observed_r2);
    else if (precision == 64)
      fpls::simpls_cv_regression<double>(static_cast<const double*>(predictors),
          static_cast<const double*>(responses), folds, n, p, q, prefixes, prefix_count, scaling,
          metric, oversample, power, seed, store_predictions != 0, predictions, status, metrics,
          q2, rmsd, observed_r2);
    else
      throw std::invalid_argument("precision must be 32 or 64");
  });
}

}  // extern "C"
