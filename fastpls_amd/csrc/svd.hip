// fastsvd and fastcor behind the C ABI of include/fastpls_hip.h.
//
// fastsvd is the audited operator rSVD of an explicit matrix X (m x n,
// column-major, resident in HBM). Every attempt of the reference schedule
// (inst/include/fastpls/core/audited_rsvd.hpp:298-414) runs as operator
// products over X on the streaming matrix-core kernels, CholQR2 / Householder
// QR and single-work-group FP64 kernels for the small decompositions. The
// audit of an attempt (triplet residuals, omitted-direction probes, consensus
// with the previous attempt) is reduced on the device; five scalars come back
// to the host, which walks the attempt and recovery schedule.
//
// fastcor (src/r_api.cpp:4054-4142) normalises the units on the device and
// forms L Rᵀ with the same products, or the per-unit dot products.
#include "../../include/fastpls_hip.h"
#include "host_util.hpp"
#include "kernels.hpp"
#include "ops.hpp"
#include "wg_linalg.hpp"

#include <cfloat>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

namespace fpls {
namespace {

constexpr int kSvdMaxW = 512;     // widest sketch (single-work-group QR / SVD bound)
constexpr int kAuditChunks = 64;  // row chunks of the residual reduction
constexpr int kProbes = 3;        // omitted-direction probes (audited_rsvd.hpp:159)
constexpr int kRecord = 5;        // residual, finite, omitted, subspace, singular

// std::max(a, b) as the reference applies it: b only when a < b, so a NaN
// candidate never replaces the running value.
__device__ inline double keep_max(double a, double b) { return a < b ? b : a; }

// ------------------------------------------------------------ kernels
// Sums of squares of F - U diag(σ) (side 0: F = A V, m rows) and of
// R - V diag(σ) (side 1: R = Aᵀ U, n rows) per (row chunk, column), and the
// count of non-finite entries of U and V (audited_rsvd.hpp:97-148).
// Grid: (kAuditChunks, r, 2).
__global__ __launch_bounds__(kWG) void k_triplet_partials(const double* F, const double* U,
                                                          long long m, const double* R,
                                                          const double* V, long long n,
                                                          const double* s, int r, double* part,
                                                          double* bad) {
  __shared__ double red[kWaves];
  const int c = blockIdx.y, side = blockIdx.z;
  const long long rows = side ? n : m;
  const double* P = (side ? R : F) + (long long)c * rows;
  const double* B = (side ? V : U) + (long long)c * rows;
  const long long chunk = (rows + gridDim.x - 1) / gridDim.x;
  const long long r0 = (long long)blockIdx.x * chunk;
  const long long r1 = r0 + chunk < rows ? r0 + chunk : rows;
  const double sv = s[c];
  double acc = 0.0, nb = 0.0;
  for (long long i = r0 + threadIdx.x; i < r1; i += kWG) {
    const double b = B[i];
    const double d = P[i] - sv * b;
    acc = fma(d, d, acc);
    nb += isfinite(b) ? 0.0 : 1.0;
  }
  acc = wg_sum(acc, red);
  nb = wg_sum(nb, red);
  if (threadIdx.x == 0) {
    const long long slot = ((long long)side * r + c) * gridDim.x + blockIdx.x;
    part[slot] = acc;
    bad[slot] = nb;
  }
}

// One step of the omitted-direction probes (audited_rsvd.hpp:150-183), one
// work-group per probe column j of the rows x kProbes blocks; state[j] is 1
// while probe j iterates, 0 once it stopped and -1 when its draw could not be
// normalised (the probe is skipped).
//   mode 0: normalise dst;
//   mode 1: dst = src - sub (the left vector with the retained directions
//           projected away), then normalise it while the probe iterates;
//   mode 2: while the probe iterates, dst = src normalised (src as it is, and
//           the probe stops, when it cannot be normalised);
//   mode 3: norms[j] = |src - sub| (-1 for a skipped probe).
__global__ __launch_bounds__(kWG) void k_probe(double* dst, const double* src, const double* sub,
                                               long long rows, int mode, int* state,
                                               double* norms) {
  __shared__ double red[kWaves];
  const int j = blockIdx.x, tid = threadIdx.x;
  const int st = state[j];
  double* d = dst ? dst + (long long)j * rows : nullptr;
  const double* a = src ? src + (long long)j * rows : nullptr;
  const double* b = sub ? sub + (long long)j * rows : nullptr;
  auto usable = [](double nrm) { return isfinite(nrm) && nrm > DBL_EPSILON; };
  if (mode == 0) {
    double ss = 0.0;
    for (long long i = tid; i < rows; i += kWG) ss = fma(d[i], d[i], ss);
    const double nrm = sqrt(wg_sum(ss, red));
    if (!usable(nrm)) {
      if (tid == 0) state[j] = -1;
      return;
    }
    for (long long i = tid; i < rows; i += kWG) d[i] /= nrm;
  } else if (mode == 1) {
    if (st == -1) return;
    double ss = 0.0;
    for (long long i = tid; i < rows; i += kWG) {
      const double v = a[i] - b[i];
      d[i] = v;
      ss = fma(v, v, ss);
    }
    if (st != 1) return;
    const double nrm = sqrt(wg_sum(ss, red));
    if (!usable(nrm)) {
      if (tid == 0) state[j] = 0;
      return;
    }
    for (long long i = tid; i < rows; i += kWG) d[i] /= nrm;
  } else if (mode == 2) {
    if (st != 1) return;
    double ss = 0.0;
    for (long long i = tid; i < rows; i += kWG) ss = fma(a[i], a[i], ss);
    const double nrm = sqrt(wg_sum(ss, red));
    const bool ok = usable(nrm);
    for (long long i = tid; i < rows; i += kWG) d[i] = ok ? a[i] / nrm : a[i];
    if (!ok && tid == 0) state[j] = 0;
  } else {
    if (st == -1) {
      if (tid == 0) norms[j] = -1.0;
      return;
    }
    double ss = 0.0;
    for (long long i = tid; i < rows; i += kWG) {
      const double v = a[i] - b[i];
      ss = fma(v, v, ss);
    }
    const double nrm = sqrt(wg_sum(ss, red));
    if (tid == 0) norms[j] = nrm;
  }
}

// Audit scalars of one attempt (audited_rsvd.hpp:97-205): rec[0] the worst
// triplet residual relative to max(|σ0|, eps), rec[1] 1 when U, V and σ are
// finite, rec[2] the omitted-direction ratio (largest probe norm, and the
// next singular value when r > retained, over max(|σ_retained-1|, eps)).
__global__ __launch_bounds__(kWG) void k_audit_finish(const double* part, const double* bad,
                                                      int chunks, int r, const double* s,
                                                      int retained, const double* norms,
                                                      double* rec) {
  __shared__ double red[kWaves];
  __shared__ double col[2 * kSvdMaxW];
  const int tid = threadIdx.x;
  double nb = 0.0;
  for (int e = tid; e < 2 * r; e += kWG) {
    col[e] = sum_strided(part + (long long)e * chunks, 1, chunks);
    nb += sum_strided(bad + (long long)e * chunks, 1, chunks);
  }
  for (int c = tid; c < r; c += kWG) nb += isfinite(s[c]) ? 0.0 : 1.0;
  nb = wg_sum(nb, red);
  if (tid == 0) {
    const double scale = fmax(fabs(s[0]), DBL_EPSILON);
    double worst = 0.0;
    for (int c = 0; c < r; ++c) worst = keep_max(worst, keep_max(sqrt(col[c]), sqrt(col[r + c])) / scale);
    const double boundary = fmax(fabs(s[retained - 1]), DBL_EPSILON);
    double largest = 0.0;
    for (int j = 0; j < kProbes; ++j)
      if (!(norms[j] < 0.0)) largest = keep_max(largest, norms[j]);  // -1: skipped probe
    double omitted = largest / boundary;
    if (r > retained) omitted = keep_max(omitted, fabs(s[retained] / boundary));
    rec[0] = worst;
    rec[1] = nb == 0.0 ? 1.0 : 0.0;
    rec[2] = omitted;
  }
}

// audited_rsvd.hpp:207-252 on cross = U_prevᵀ U_cur (r x r): rec[3] the
// subspace distance sqrt((r - min(r, Σ cross²)) / r), rec[4] the largest
// singular value change over max(|σ_cur[0]|, eps).
__global__ __launch_bounds__(kWG) void k_consensus(const double* cross, int r,
                                                   const double* s_prev, const double* s_cur,
                                                   double* rec) {
  __shared__ double red[kWaves];
  double acc = 0.0;
  for (int e = threadIdx.x; e < r * r; e += kWG) acc = fma(cross[e], cross[e], acc);
  const double overlap = fmin(double(r), wg_sum(acc, red));
  if (threadIdx.x == 0) {
    rec[3] = sqrt(fmax(0.0, (double(r) - overlap) / double(r)));
    const double scale = fmax(fabs(s_cur[0]), DBL_EPSILON);
    double diff = 0.0;
    for (int i = 0; i < r; ++i) diff = keep_max(diff, fabs(s_prev[i] - s_cur[i]));
    rec[4] = diff / scale;
  }
}

template<class T>
__global__ __launch_bounds__(kWG) void k_widen(const T* in, long long count, double* out) {
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < count;
       e += (long long)gridDim.x * kWG)
    out[e] = double(in[e]);
}

// src/r_api.cpp:4054-4080 (normalized_units) for units that are the columns
// of X (features x units): one work-group per unit, centred by its mean and
// scaled to unit norm, rounded to the input precision (the reference's
// Scalar); a unit with a non-finite value or zero norm is NaN. The unit is
// written to column u of outT (features x units, element T) and/or outD
// (FP64 copy of the same rounded values); either may be null.
template<class T>
__global__ __launch_bounds__(kWG) void k_units_cols(const T* X, long long features, T* outT,
                                                    double* outD) {
  __shared__ double red[kWaves];
  const long long u = blockIdx.x;
  const T* x = X + u * features;
  double sum = 0.0, bad = 0.0;
  for (long long f = threadIdx.x; f < features; f += kWG) {
    const double v = double(x[f]);
    bad += isfinite(v) ? 0.0 : 1.0;
    sum += v;
  }
  sum = wg_sum(sum, red);
  bad = wg_sum(bad, red);
  const double mean = sum / double(features);
  double ss = 0.0;
  for (long long f = threadIdx.x; f < features; f += kWG) {
    const double v = double(x[f]) - mean;
    ss = fma(v, v, ss);
  }
  const double nrm = sqrt(wg_sum(ss, red));
  const bool ok = bad == 0.0 && nrm > 0.0;
  for (long long f = threadIdx.x; f < features; f += kWG) {
    const T v = ok ? T((double(x[f]) - mean) / nrm) : T(NAN);
    if (outT) outT[u * features + f] = v;
    if (outD) outD[u * features + f] = double(v);
  }
}

// The same for units that are the rows of X (units x features): 64 units per
// work-group, four feature lanes per unit for coalesced reads, and the
// normalised tile transposed through LDS so the unit-major writes are
// coalesced too.
template<class T>
__global__ __launch_bounds__(kWG) void k_units_rows(const T* X, long long units,
                                                    long long features, T* outT, double* outD) {
  __shared__ double part[4][64];
  __shared__ double stat[2][64];
  __shared__ double tile[64][65];
  const int lane = threadIdx.x & 63, quarter = threadIdx.x >> 6;
  const long long u0 = (long long)blockIdx.x * 64;
  const long long u = u0 + lane;
  const bool live = u < units;
  double sum = 0.0, bad = 0.0;
  if (live)
    for (long long f = quarter; f < features; f += 4) {
      const double v = double(X[u + f * units]);
      bad += isfinite(v) ? 0.0 : 1.0;
      sum += v;
    }
  part[quarter][lane] = sum;
  __syncthreads();
  if (quarter == 0) stat[0][lane] = (part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]) / double(features);
  __syncthreads();
  part[quarter][lane] = bad;
  __syncthreads();
  if (quarter == 0) stat[1][lane] = part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane];
  __syncthreads();
  const double mean = stat[0][lane];
  double ss = 0.0;
  if (live)
    for (long long f = quarter; f < features; f += 4) {
      const double v = double(X[u + f * units]) - mean;
      ss = fma(v, v, ss);
    }
  part[quarter][lane] = ss;
  __syncthreads();
  double nrm = sqrt(part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]);
  const bool ok = stat[1][lane] == 0.0 && nrm > 0.0;
  __syncthreads();
  for (long long f0 = 0; f0 < features; f0 += 64) {
    for (int fi = quarter; fi < 64; fi += 4) {
      const long long f = f0 + fi;
      double v = 0.0;
      if (live && f < features) v = ok ? (double(X[u + f * units]) - mean) / nrm : double(NAN);
      tile[fi][lane] = double(T(v));
    }
    __syncthreads();
    // thread (fi = lane, unit row = quarter + 4 j) writes feature f0 + lane of that unit
    for (int j = quarter; j < 64; j += 4) {
      const long long uu = u0 + j, f = f0 + lane;
      if (uu < units && f < features) {
        const double v = tile[lane][j];
        if (outT) outT[uu * features + f] = T(v);
        if (outD) outD[uu * features + f] = v;
      }
    }
    __syncthreads();
  }
}

// Per-unit dot products of the normalised units (the diag = TRUE route of
// fastcor_values, src/r_api.cpp:4099-4115): one work-group per unit.
template<class T>
__global__ __launch_bounds__(kWG) void k_unit_dots(const T* L, const double* R, long long features,
                                                   double* out) {
  __shared__ double red[kWaves];
  const long long u = blockIdx.x;
  double acc = 0.0;
  for (long long f = threadIdx.x; f < features; f += kWG)
    acc = fma(double(L[u * features + f]), R[u * features + f], acc);
  acc = wg_sum(acc, red);
  if (threadIdx.x == 0) out[u] = acc;
}

unsigned grid_1d(long long count) {
  long long g = (count + kWG - 1) / kWG;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return unsigned(g);
}

// ------------------------------------------------------------ fastsvd
struct AuditRecord {
  bool certified = false, fallback = false;
  int attempts = 0, oversample = 0, power = 0;
  unsigned int seed = 0;
  double subspace = 0.0, singular = 0.0, residual = 0.0, omitted = 0.0;
};

template<class T>
class AuditedSvd : private StreamOps {
 public:
  AuditedSvd(const T* x, bool device, int m, int n) : m_(m), n_(n) {
    stream_ = stream_take();
    for (Buf* b : bufs()) b->bind(stream_);
    try {
      if (device) {
        X_ = x;
      } else {
        xown_.alloc(sizeof(T) * (size_t)m * n);
        FPLS_HIP(hipMemcpyAsync(xown_.as<T>(), x, sizeof(T) * (size_t)m * n,
                                hipMemcpyHostToDevice, stream_));
        X_ = xown_.as<T>();
      }
      flags_.alloc(sizeof(int) * 4);
      fill_zero_int(flags_.as<int>(), 4);
      rec_.alloc(sizeof(double) * kRecord);
    } catch (...) {
      close();
      throw;
    }
  }
  ~AuditedSvd() { close(); }

  // audited_rsvd.hpp:298-414. Writes U (m x kept), σ (kept), Vt (kept x n,
  // unless left_only) and the audit record; returns kept.
  int run(int components, int oversample, int power, unsigned int seed, bool left_only,
          double* U, double* s, double* Vt, bool device_out, AuditRecord& audit) {
    const int maximum = std::min(m_, n_);
    const int retained = std::min(maximum, std::max(components, 1));
    if (maximum < 6) {
      exact();
      audit = AuditRecord();
      audit.certified = true;
      audit.fallback = true;
      audit.oversample = std::max(oversample, 0);
      audit.power = std::max(power, 0);
      audit.seed = seed;
      copy_out(0, retained, left_only, U, s, Vt, device_out);
      return retained;
    }
    const int rank = std::min(maximum, retained + 1);
    const int oversamples[] = {std::max(oversample, 0), std::max(oversample, 32),
                               std::max(oversample, 48)};
    const int powers[] = {std::max(power, 0), std::max(power, 3), std::max(power, 4)};
    const unsigned int seeds[] = {seed, seed + 104729U, seed + 209759U};
    int cur_os = oversample, cur_pw = power;
    unsigned int cur_seed = seed;
    double residual = std::numeric_limits<double>::infinity();
    double omitted = std::numeric_limits<double>::infinity();
    double subspace = 0.0, singular = 0.0;
    bool certified = false, prev_finite = false;
    int attempts = 0, cur = 0;
    for (int attempt = 0; attempt < 3; ++attempt) {
      cur_os = oversamples[attempt];
      cur_pw = powers[attempt];
      cur_seed = seeds[attempt];
      operator_svd(cur, rank, cur_os, cur_pw, cur_seed);
      ++attempts;
      launch_audit(cur, retained, cur_seed);
      if (attempt > 0) launch_consensus(1 - cur, cur, rank);
      const Scalars a = scalars();
      bool checked = false;
      if (a.finite) {
        residual = a.residual;
        omitted = a.omitted;
        checked = std::isfinite(residual) && residual <= 1e-2 && std::isfinite(omitted) &&
                  omitted <= 1.01;
      }
      bool agreed = false;
      if (attempt > 0 && a.finite && prev_finite) {
        subspace = a.subspace;
        singular = a.singular;
        agreed = subspace <= 1e-3 && singular <= 1e-5 && residual <= 1e-6;
      }
      certified = checked || agreed;
      if (certified) break;
      prev_finite = a.finite;
      cur = 1 - cur;
    }
    for (int recovery = 0; !certified && recovery < 4; ++recovery) {
      const int width = std::min(maximum, rank + std::max(cur_os, 0));
      const int remaining = maximum - width;
      const int enlarged = std::min(maximum, std::max(width + std::min(32, remaining),
                                                      width + std::min(width, remaining)));
      cur_os = enlarged - rank;
      cur_pw = std::max(cur_pw, 6) + 2 * recovery;
      cur_seed += 104729U * unsigned(recovery + 1);
      operator_svd(cur, rank, cur_os, cur_pw, cur_seed);
      ++attempts;
      launch_audit(cur, retained, cur_seed);
      const Scalars a = scalars();
      if (a.finite) {
        residual = a.residual;
        omitted = a.omitted;
        certified = std::isfinite(residual) && residual <= 1e-2 && std::isfinite(omitted) &&
                    omitted <= 1.01;
      }
    }
    if (!certified)
      throw std::runtime_error("rSVD recovery did not meet numerical tolerances within its "
                               "workspace/iteration budget; no unchecked result was returned");
    audit = AuditRecord();
    audit.certified = true;
    audit.attempts = attempts;
    audit.oversample = cur_os;
    audit.power = cur_pw;
    audit.seed = cur_seed;
    audit.subspace = subspace;
    audit.singular = singular;
    audit.residual = residual;
    audit.omitted = omitted;
    copy_out(cur, retained, left_only, U, s, Vt, device_out);
    return retained;
  }

 private:
  struct Scalars {
    double residual, omitted, subspace, singular;
    bool finite;
  };

  std::vector<Buf*> bufs() {
    std::vector<Buf*> v = ops_bufs();
    for (Buf* b : {&xown_, &flags_, &rec_, &rb_, &fin_scratch_, &rcopy_, &qr_r_, &small_u_,
                   &small_s_, &small_v_, &fwd_, &rev2_, &part_sums_, &bad_sums_, &probe_right_, &probe_raw_,
                   &probe_left_, &probe_coef_, &probe_proj_, &probe_state_, &probe_norms_,
                   &cross_, &vt_out_, &wide_, &wide_t_})
      v.push_back(b);
    for (int i = 0; i < 2; ++i)
      for (Buf& b : tri_[i]) v.push_back(&b);
    return v;
  }

  void close() noexcept {
    if (!stream_) return;
    for (Buf* b : bufs()) b->release();
    stream_give(stream_);
    stream_ = nullptr;
  }

  double* U_of(int slot) { return tri_[slot][0].d(); }
  double* s_of(int slot) { return tri_[slot][1].d(); }
  double* V_of(int slot) { return tri_[slot][2].d(); }

  void reserve_triplets(int slot, int r) {
    tri_[slot][0].reserve(sizeof(double) * (size_t)m_ * r);
    tri_[slot][1].reserve(sizeof(double) * (size_t)r);
    tri_[slot][2].reserve(sizeof(double) * (size_t)n_ * r);
  }

  void check_flags() {
    int f = 0;
    FPLS_HIP(hipMemcpyAsync(&f, flags_.as<int>(), sizeof(int), hipMemcpyDeviceToHost, stream_));
    FPLS_HIP(hipStreamSynchronize(stream_));
    if (f & kFlagDecomposition)
      throw std::runtime_error("fastsvd: a QR or SVD of the sketch failed (non-finite input?)");
  }

  // operator_rsvd.hpp:204-272 with the explicit operator X (operators.hpp:16-56):
  // triplets of rank r into slot `slot`.
  void operator_svd(int slot, int r, int oversample, int power, unsigned int seed) {
    reserve_triplets(slot, r);
    auto fwd = [&](const double* Z, int w, double* out) {
      product<T>(X_, m_, n_, nullptr, Z, n_, w, out);
    };
    auto adj = [&](const double* B, int w, double* out) {
      tproduct<T>(X_, m_, n_, nullptr, B, m_, w, out);
    };
    operator_rsvd(fwd, adj, m_, n_, r, oversample, power, seed, false, U_of(slot), s_of(slot),
                  V_of(slot), flags_.as<int>());
  }

  // audited_rsvd.hpp:185-205: residual partials and the three probes.
  void launch_audit(int slot, int retained, unsigned int seed) {
    const int r = std::min(std::min(m_, n_), retained + 1);
    const double* U = U_of(slot);
    const double* V = V_of(slot);
    fwd_.reserve(sizeof(double) * (size_t)m_ * r);
    rev2_.reserve(sizeof(double) * (size_t)n_ * r);
    part_sums_.reserve(sizeof(double) * 2 * kAuditChunks * (size_t)r);
    bad_sums_.reserve(sizeof(double) * 2 * kAuditChunks * (size_t)r);
    product<T>(X_, m_, n_, nullptr, V, n_, r, fwd_.d());
    tproduct<T>(X_, m_, n_, nullptr, U, m_, r, rev2_.d());
    hipLaunchKernelGGL(k_triplet_partials, dim3(kAuditChunks, unsigned(r), 2), dim3(kWG), 0,
                       stream_, fwd_.d(), U, (long long)m_, rev2_.d(), V, (long long)n_,
                       s_of(slot), r, part_sums_.d(), bad_sums_.d());
    FPLS_LAUNCHED("k_triplet_partials");
    // omitted-direction probes (audited_rsvd.hpp:150-183)
    probe_host_.resize((size_t)n_ * kProbes);
    gaussian(seed + 32452843U, probe_host_.size(), probe_host_.data());
    probe_right_.reserve(sizeof(double) * (size_t)n_ * kProbes);
    probe_raw_.reserve(sizeof(double) * (size_t)std::max(m_, n_) * kProbes);
    probe_left_.reserve(sizeof(double) * (size_t)m_ * kProbes);
    probe_coef_.reserve(sizeof(double) * (size_t)retained * kProbes);
    probe_proj_.reserve(sizeof(double) * (size_t)m_ * kProbes);
    probe_state_.reserve(sizeof(int) * kProbes);
    probe_norms_.reserve(sizeof(double) * kProbes);
    FPLS_HIP(hipMemcpyAsync(probe_right_.d(), probe_host_.data(),
                            sizeof(double) * probe_host_.size(), hipMemcpyHostToDevice, stream_));
    launch_fill_int(probe_state_.as<int>(), kProbes, 1, stream_);
    probe(probe_right_.d(), nullptr, nullptr, n_, 0);
    auto left_projected = [&]() {
      product<T>(X_, m_, n_, nullptr, probe_right_.d(), n_, kProbes, probe_raw_.d());
      tproduct<double>(U, m_, retained, nullptr, probe_raw_.d(), m_, kProbes, probe_coef_.d());
      product<double>(U, m_, retained, nullptr, probe_coef_.d(), retained, kProbes,
                      probe_proj_.d());
    };
    for (int it = 0; it < 2; ++it) {
      left_projected();
      probe(probe_left_.d(), probe_raw_.d(), probe_proj_.d(), m_, 1);
      tproduct<T>(X_, m_, n_, nullptr, probe_left_.d(), m_, kProbes, probe_raw_.d());
      probe(probe_right_.d(), probe_raw_.d(), nullptr, n_, 2);
    }
    left_projected();
    probe(nullptr, probe_raw_.d(), probe_proj_.d(), m_, 3);
    hipLaunchKernelGGL(k_audit_finish, dim3(1), dim3(kWG), 0, stream_, part_sums_.d(),
                       bad_sums_.d(), kAuditChunks, r, s_of(slot), retained, probe_norms_.d(),
                       rec_.d());
    FPLS_LAUNCHED("k_audit_finish");
  }

  void probe(double* dst, const double* src, const double* sub, long long rows, int mode) {
    hipLaunchKernelGGL(k_probe, dim3(kProbes), dim3(kWG), 0, stream_, dst, src, sub, rows, mode,
                       probe_state_.as<int>(), probe_norms_.d());
    FPLS_LAUNCHED("k_probe");
  }

  void launch_consensus(int prev, int cur, int r) {
    cross_.reserve(sizeof(double) * (size_t)r * r);
    tproduct<double>(U_of(prev), m_, r, nullptr, U_of(cur), m_, r, cross_.d());
    hipLaunchKernelGGL(k_consensus, dim3(1), dim3(kWG), 0, stream_, cross_.d(), r, s_of(prev),
                       s_of(cur), rec_.d());
    FPLS_LAUNCHED("k_consensus");
  }

  Scalars scalars() {
    double h[kRecord] = {0, 0, 0, 0, 0};
    FPLS_HIP(hipMemcpyAsync(h, rec_.d(), sizeof(h), hipMemcpyDeviceToHost, stream_));
    check_flags();  // synchronises the stream
    return Scalars{h[0], h[2], h[3], h[4], h[1] == 1.0};
  }

  // audited_rsvd.hpp:312-322 (exact_operator_svd): min(m, n) < 6. The thin
  // side is at most five wide: QR of the tall orientation, then the small SVD
  // of its triangle.
  void exact() {
    const bool tall = m_ >= n_;
    const int mt = tall ? m_ : n_, nt = tall ? n_ : m_;
    int* flags = flags_.as<int>();
    wide_.reserve(sizeof(double) * (size_t)m_ * n_);
    hipLaunchKernelGGL((k_widen<T>), dim3(grid_1d((long long)m_ * n_)), dim3(kWG), 0, stream_,
                       X_, (long long)m_ * n_, wide_.d());
    FPLS_LAUNCHED("k_widen");
    double* A = wide_.d();
    if (!tall) {
      wide_t_.reserve(sizeof(double) * (size_t)m_ * n_);
      launch_transpose(wide_.d(), m_, n_, wide_t_.d(), stream_);
      A = wide_t_.d();
    }
    rcopy_.reserve(sizeof(double) * (size_t)mt * nt);
    rb_.reserve(sizeof(double) * (size_t)mt * nt);
    qr_r_.reserve(sizeof(double) * (size_t)nt * nt);
    small_u_.reserve(sizeof(double) * (size_t)nt * nt);
    small_s_.reserve(sizeof(double) * (size_t)nt);
    small_v_.reserve(sizeof(double) * (size_t)nt * nt);
    fin_scratch_.reserve(sizeof(double) * (3 * (size_t)nt * nt + nt + 64));
    FPLS_HIP(hipMemcpyAsync(rcopy_.d(), A, sizeof(double) * (size_t)mt * nt,
                            hipMemcpyDeviceToDevice, stream_));
    qr_tall(rcopy_.d(), mt, nt, rb_.d(), flags);
    tproduct<double>(rb_.d(), mt, nt, nullptr, A, mt, nt, qr_r_.d());
    launch_small_svd(qr_r_.d(), nt, nt, nt, small_u_.d(), small_s_.d(), small_v_.d(),
                     fin_scratch_.d(), flags, stream_);
    reserve_triplets(0, nt);
    // tall: A = (Q Ur) σ Vrᵀ; wide: Aᵀ = (Q Ur) σ Vrᵀ, so A = Vr σ (Q Ur)ᵀ
    double* QU = tall ? U_of(0) : V_of(0);
    double* Vr = tall ? V_of(0) : U_of(0);
    product<double>(rb_.d(), mt, nt, nullptr, small_u_.d(), nt, nt, QU);
    FPLS_HIP(hipMemcpyAsync(Vr, small_v_.d(), sizeof(double) * (size_t)nt * nt,
                            hipMemcpyDeviceToDevice, stream_));
    FPLS_HIP(hipMemcpyAsync(s_of(0), small_s_.d(), sizeof(double) * nt, hipMemcpyDeviceToDevice,
                            stream_));
    check_flags();
  }

  // audited_rsvd.hpp:254-278 (trim_triplets) and the copy out.
  void copy_out(int slot, int kept, bool left_only, double* U, double* s, double* Vt,
                bool device_out) {
    const hipMemcpyKind kind = device_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    FPLS_HIP(hipMemcpyAsync(U, U_of(slot), sizeof(double) * (size_t)m_ * kept, kind, stream_));
    FPLS_HIP(hipMemcpyAsync(s, s_of(slot), sizeof(double) * kept, kind, stream_));
    if (!left_only) {
      vt_out_.reserve(sizeof(double) * (size_t)kept * n_);
      launch_transpose(V_of(slot), n_, kept, vt_out_.d(), stream_);
      FPLS_HIP(hipMemcpyAsync(Vt, vt_out_.d(), sizeof(double) * (size_t)kept * n_, kind, stream_));
    }
    FPLS_HIP(hipStreamSynchronize(stream_));
  }

  int m_, n_;
  const T* X_ = nullptr;
  std::vector<double> probe_host_;
  Buf xown_, flags_, rec_, rb_;
  Buf fin_scratch_, rcopy_, qr_r_, small_u_, small_s_, small_v_;  // exact route
  Buf fwd_, rev2_, part_sums_, bad_sums_;
  Buf probe_right_, probe_raw_, probe_left_, probe_coef_, probe_proj_, probe_state_, probe_norms_;
  Buf cross_, vt_out_, wide_, wide_t_;
  Buf tri_[2][3];  // U (m x r), σ (r), V (n x r) of the current and previous attempt
};

// ------------------------------------------------------------ fastcor
template<class T>
class Correlation : private StreamOps {
 public:
  Correlation() {
    stream_ = stream_take();
    for (Buf* b : bufs()) b->bind(stream_);
  }
  ~Correlation() { close(); }

  // fastcor_values (src/r_api.cpp:4082-4142).
  void run(const T* a, int ar, int ac, const T* b, int br, int bc, bool device, bool by_row,
           bool diagonal, double* out) {
    const long long ua = by_row ? ar : ac, fa = by_row ? ac : ar;
    long long ub = ua, fb = fa;
    if (b) {
      ub = by_row ? br : bc;
      fb = by_row ? bc : br;
      if (fa != fb)
        throw std::invalid_argument("a and b must contain the same number of correlated features");
    } else {
      diagonal = false;
    }
    if (diagonal && ua != ub)
      throw std::invalid_argument("diag = TRUE requires matching numbers of rows or columns");
    const T* da = upload(a, (size_t)ar * ac, ain_, device);
    const T* db = b ? upload(b, (size_t)br * bc, bin_, device) : nullptr;
    // left units in the input precision, right units in FP64 (the product's
    // right operand); without b both come from a
    lt_.reserve(sizeof(T) * (size_t)ua * fa);
    rd_.reserve(sizeof(double) * (size_t)ub * fb);
    normalise(da, ar, ac, by_row, lt_.as<T>(), db ? nullptr : rd_.d());
    if (db) normalise(db, br, bc, by_row, nullptr, rd_.d());
    double* dst = out;
    if (!device) {
      res_.reserve(sizeof(double) * (size_t)(diagonal ? ua : ua * ub));
      dst = res_.d();
    }
    if (diagonal) {
      hipLaunchKernelGGL((k_unit_dots<T>), dim3(unsigned(ua)), dim3(kWG), 0, stream_, lt_.as<T>(),
                         rd_.d(), fa, dst);
      FPLS_LAUNCHED("k_unit_dots");
    } else {
      // C (ua x ub) = L Rᵀ with L, R stored features x units
      tproduct<T>(lt_.as<T>(), fa, ua, nullptr, rd_.d(), fa, int(ub), dst);
    }
    if (!device)
      FPLS_HIP(hipMemcpyAsync(out, dst, sizeof(double) * (size_t)(diagonal ? ua : ua * ub),
                              hipMemcpyDeviceToHost, stream_));
    FPLS_HIP(hipStreamSynchronize(stream_));
  }

 private:
  std::vector<Buf*> bufs() {
    std::vector<Buf*> v = ops_bufs();
    for (Buf* b : {&ain_, &bin_, &lt_, &rd_, &res_}) v.push_back(b);
    return v;
  }

  void close() noexcept {
    if (!stream_) return;
    for (Buf* b : bufs()) b->release();
    stream_give(stream_);
    stream_ = nullptr;
  }

  const T* upload(const T* x, size_t count, Buf& own, bool device) {
    if (device) return x;
    own.reserve(sizeof(T) * count);
    FPLS_HIP(hipMemcpyAsync(own.as<T>(), x, sizeof(T) * count, hipMemcpyHostToDevice, stream_));
    return own.as<T>();
  }

  void normalise(const T* x, int rows, int cols, bool by_row, T* outT, double* outD) {
    if (by_row) {
      hipLaunchKernelGGL((k_units_rows<T>), dim3(unsigned((rows + 63) / 64)), dim3(kWG), 0,
                         stream_, x, (long long)rows, (long long)cols, outT, outD);
      FPLS_LAUNCHED("k_units_rows");
    } else {
      hipLaunchKernelGGL((k_units_cols<T>), dim3(unsigned(cols)), dim3(kWG), 0, stream_, x,
                         (long long)rows, outT, outD);
      FPLS_LAUNCHED("k_units_cols");
    }
  }

  Buf ain_, bin_, lt_, rd_, res_;
};

}  // namespace
}  // namespace fpls

using namespace fpls;

extern "C" {

int fastpls_hip_rsvd(const void* x, int precision, int device, int m, int n, int k,
                     int oversample, int power, unsigned int seed, int left_only, double* U,
                     double* s, double* Vt, int* got, double* audit, char* error,
                     size_t error_capacity) {
  return guarded(error, error_capacity, [&] {
    if (!x || !U || !s || (!left_only && !Vt) || !got || !audit)
      throw std::invalid_argument("null fastsvd argument");
    if (m < 1 || n < 1 || k < 1 || oversample < 0 || power < 0)
      throw std::invalid_argument(precision == 32 ? "float32 core rSVD controls are invalid"
                                                  : "double core rSVD controls are invalid");
    require_device();
    AuditRecord rec;
    int kept = 0;
    if (precision == 32) {
      AuditedSvd<float> svd(static_cast<const float*>(x), device != 0, m, n);
      kept = svd.run(k, oversample, power, seed, left_only != 0, U, s, Vt, device != 0, rec);
    } else if (precision == 64) {
      AuditedSvd<double> svd(static_cast<const double*>(x), device != 0, m, n);
      kept = svd.run(k, oversample, power, seed, left_only != 0, U, s, Vt, device != 0, rec);
    } else {
      throw std::invalid_argument("precision must be 32 or 64");
    }
    *got = kept;
    const double v[10] = {double(rec.certified), double(rec.fallback), double(rec.attempts),
                          double(rec.oversample), double(rec.power), double(rec.seed),
                          rec.subspace, rec.singular, rec.residual, rec.omitted};
    for (int i = 0; i < 10; ++i) audit[i] = v[i];
  });
}

int fastpls_hip_fastcor(const void* a, int a_rows, int a_cols, const void* b, int b_rows,
                        int b_cols, int precision, int device, int by_row, int diagonal,
                        double* out, char* error, size_t error_capacity) {
  return guarded(error, error_capacity, [&] {
    if (!a || !out) throw std::invalid_argument("null fastcor argument");
    if (a_rows < 1 || a_cols < 1 || (b && (b_rows < 1 || b_cols < 1)))
      throw std::invalid_argument("fastcor inputs must be non-empty matrices");
    require_device();
    if (precision == 32) {
      Correlation<float> c;
      c.run(static_cast<const float*>(a), a_rows, a_cols, static_cast<const float*>(b), b_rows,
            b_cols, device != 0, by_row != 0, diagonal != 0, out);
    } else if (precision == 64) {
      Correlation<double> c;
      c.run(static_cast<const double*>(a), a_rows, a_cols, static_cast<const double*>(b), b_rows,
            b_cols, device != 0, by_row != 0, diagonal != 0, out);
    } else {
      throw std::invalid_argument("precision must be 32 or 64");
    }
  });
}

}  // extern "C"
