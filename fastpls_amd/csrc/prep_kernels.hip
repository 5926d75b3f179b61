// Preprocessing, prediction-side and data-generation kernels.
#include "kernels.hpp"

#include <algorithm>

#include <cfloat>

namespace fpls {

namespace {

// ------------------------------------------------------------ responses
// supervised.hpp:54-61: response means; the centred copy is what the
// shifted cross-product pass multiplies by. One work-group per column.
// Response centring in three fixed-order phases over (kCenterBlocks x q)
// work-groups: chunk sums, then the mean and the centred copy with chunk sums
// of the centred values, then one thread per response folds the chunks.
constexpr int kCenterBlocks = 64;

template<class T>
__global__ __launch_bounds__(kWG) void k_center_sums(const T* Y, long long n, double* part) {
  __shared__ double red[kWaves];
  const int c = blockIdx.y, b = blockIdx.x;
  const long long per = (n + kCenterBlocks - 1) / kCenterBlocks;
  const long long i0 = b * per, i1 = i0 + per < n ? i0 + per : n;
  const T* y = Y + (long long)c * n;
  double s = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += kWG) s += double(y[i]);
  s = wg_sum(s, red);
  if (threadIdx.x == 0) part[(long long)c * kCenterBlocks + b] = s;
}

template<class T>
__global__ __launch_bounds__(kWG) void k_center_apply(const T* Y, long long n, const double* part,
                                                      double* Yc, double* part2) {
  __shared__ double red[kWaves];
  const int c = blockIdx.y, b = blockIdx.x;
  double tot = 0.0;
  for (int k = 0; k < kCenterBlocks; ++k) tot += part[(long long)c * kCenterBlocks + k];
  const double mu = tot / double(n);
  const long long per = (n + kCenterBlocks - 1) / kCenterBlocks;
  const long long i0 = b * per, i1 = i0 + per < n ? i0 + per : n;
  const T* y = Y + (long long)c * n;
  double* yc = Yc + (long long)c * n;
  double cs = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += kWG) {
    const double v = double(y[i]) - mu;
    yc[i] = v;
    cs += v;
  }
  cs = wg_sum(cs, red);
  if (threadIdx.x == 0) part2[(long long)c * kCenterBlocks + b] = cs;
}

__global__ __launch_bounds__(kWG) void k_center_fold(const double* part, const double* part2,
                                                     long long n, int q, double* mean,
                                                     double* ysum) {
  const int c = blockIdx.x * kWG + threadIdx.x;
  if (c >= q) return;
  double tot = 0.0, cs = 0.0;
  for (int k = 0; k < kCenterBlocks; ++k) {
    tot += part[(long long)c * kCenterBlocks + k];
    cs += part2[(long long)c * kCenterBlocks + k];
  }
  mean[c] = tot / double(n);
  ysum[c] = cs;
}

// One response column per work-group: the mean, then the centred copy and
// its sum (a second read of the column, from cache).
template<class T>
__global__ __launch_bounds__(kWG) void k_center_columns(const CenterColumn<T>* list) {
  __shared__ double red[kWaves];
  const CenterColumn<T> c = list[blockIdx.x];
  double s = 0.0;
  if (c.mask) {
    for (long long i = threadIdx.x; i < c.n; i += kWG)
      if (c.mask[i]) s += double(c.y[c.perm ? c.perm[i] : i]);
  } else {
#pragma unroll 4
    for (long long i = threadIdx.x; i < c.n; i += kWG) s += double(c.y[c.perm ? c.perm[i] : i]);
  }
  const double mu = wg_sum(s, red) / double(c.mask ? c.used : c.n);
  double cs = 0.0;
  if (c.mask) {
    for (long long i = threadIdx.x; i < c.n; i += kWG) {
      const double v = c.mask[i] ? double(c.y[c.perm ? c.perm[i] : i]) - mu : 0.0;
      c.yc[i] = v;
      cs += v;
    }
  } else {
#pragma unroll 4
    for (long long i = threadIdx.x; i < c.n; i += kWG) {
      const double v = double(c.y[c.perm ? c.perm[i] : i]) - mu;
      c.yc[i] = v;
      cs += v;
    }
  }
  cs = wg_sum(cs, red);
  if (threadIdx.x == 0) {
    *c.mean = mu;
    *c.ysum = cs;
  }
}

__global__ __launch_bounds__(kWG) void k_fill_int(int* p, int count, int value) {
  for (int i = blockIdx.x * kWG + threadIdx.x; i < count; i += gridDim.x * kWG) p[i] = value;
}

__global__ __launch_bounds__(kWG) void k_axpy(double alpha, const double* x, double* y,
                                              long long count) {
  for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < count;
       i += (long long)gridDim.x * kWG)
    y[i] = fma(alpha, x[i], y[i]);
}

// Each column (one work-group) scaled to unit norm; a column whose norm is
// non-finite or at most eps stays as it is and sets *stop
// (operator_rsvd.hpp:37-61, normalize_column).
__global__ __launch_bounds__(kWG) void k_unit_columns(double* v, long long rows, int* stop) {
  __shared__ double red[kWaves];
  double* c = v + (long long)blockIdx.x * rows;
  double ss = 0.0;
  for (long long i = threadIdx.x; i < rows; i += kWG) ss = fma(c[i], c[i], ss);
  const double nrm = sqrt(wg_sum(ss, red));
  if (!isfinite(nrm) || !(nrm > DBL_EPSILON)) {
    if (threadIdx.x == 0) atomicOr(stop, 1);
    return;
  }
  for (long long i = threadIdx.x; i < rows; i += kWG) c[i] /= nrm;
}

// classification.hpp:106-123: class counts (integer atomics: order-free).
__global__ __launch_bounds__(kWG) void k_label_counts(const int* labels, long long n, int K,
                                                      int* counts, int* flags) {
  for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < n;
       i += (long long)gridDim.x * kWG) {
    const int l = labels[i];
    if (l < 1 || l > K) atomicOr(flags, kFlagBadLabel);
    else atomicAdd(&counts[l - 1], 1);
  }
}

// Centred one-hot (column c: [label == c+1] - prior_c), plain one-hot and
// priors; column sums of the centred one-hot go to ysum.
__global__ __launch_bounds__(kWG) void k_label_fill(const int* labels, long long n, int K,
                                                    const int* counts, double* priors,
                                                    double* Yc, double* H, double* ysum,
                                                    int* flags) {
  __shared__ double red[kWaves];
  const int c = blockIdx.x;
  const double prior = double(counts[c]) / double(n);
  if (threadIdx.x == 0) {
    priors[c] = prior;
    if (counts[c] == 0) atomicOr(flags, kFlagEmptyClass);
  }
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += kWG) {
    const double h = labels[i] == c + 1 ? 1.0 : 0.0;
    const double v = h - prior;
    Yc[i + (long long)c * n] = v;
    if (H) H[i + (long long)c * n] = h;
    s += v;
  }
  s = wg_sum(s, red);
  if (threadIdx.x == 0) ysum[c] = s;
}

// supervised.hpp:63-124 by shifted sums: d = x - x[0,j]; mean, the sum-of-
// squares standard deviation with the reference's clamp and fallback, and
// S[j,c] = (sum d*yc - (mean - shift) * sum yc) / scale.
template<class T>
__global__ __launch_bounds__(kWG) void k_finalize_stats(
    const T* X, long long n, int p, int scaling, const double* stats_part, int splits,
    double* mean, double* scale, double* colss, int* flags, const double* hold_stats,
    long long used) {
  const int j = blockIdx.x * kWG + threadIdx.x;
  if (j >= p) return;
  double s1 = sum_strided(stats_part + j, 2LL * p, splits);
  double s2 = sum_strided(stats_part + p + j, 2LL * p, splits);
  if (hold_stats) {  // a row view: the full sums minus the held-out rows'
    s1 -= hold_stats[j];
    s2 -= hold_stats[p + j];
  }
  const double shift = double(X[(long long)j * n]);
  const long long rows = used >= 0 ? used : n;
  const double dn = double(rows);
  double mu = 0.0, sd = 1.0;
  if (scaling != 3) {
    mu = shift + s1 / dn;
    if (scaling == 2) {
      const double css = fmax(0.0, s2 - s1 * s1 / dn);
      sd = sqrt(css / double(rows > 1 ? rows - 1 : 1));
      if (!isfinite(sd) || sd <= 0.0) sd = 1.0;
    }
  }
  mean[j] = mu;
  scale[j] = sd;
  const double delta = shift - mu;  // x - mu = d + delta
  const double ss = s2 + 2.0 * delta * s1 + dn * delta * delta;
  colss[j] = ss / (sd * sd);
  if (!(isfinite(mu) && isfinite(ss))) atomicOr(flags, kFlagNonFinite);
}

// S[j, c] = (sum_i d_ij y~_ic + (shift_j - mu_j) sum_i y~_ic) / sd_j over the
// row-split partials; grid (column blocks, responses).
template<class T>
__global__ __launch_bounds__(kWG) void k_finalize_cross(
    const T* X, long long n, int p, int q, const double* s_part, int splits,
    const double* ysum, const double* mean, const double* scale, double* S, double* S0,
    int* flags, const int* offsets) {
  const int j = blockIdx.x * kWG + threadIdx.x;
  const int c = blockIdx.y;
  if (j >= p) return;
  if (offsets) flags += offsets[c];
  const double delta = double(X[(long long)j * n]) - mean[j];
  const double acc = sum_strided(s_part + j + (long long)c * p, (long long)p * q, splits);
  const double v = (acc + delta * ysum[c]) / scale[j];
  S[j + (long long)c * p] = v;
  if (S0) S0[j + (long long)c * p] = v;
  if (!isfinite(v)) atomicOr(flags, kFlagNonFinite);
}

// Shifted sums of the held-out rows of a row view: one wave per column j,
// over the gathered copy hx (hn x p), with the shift X[0, j] of the full X.
template<class T>
__global__ __launch_bounds__(kWG) void k_holdout_stats(const T* hx, long long hn, int p,
                                                       const T* X, long long n, double* out) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (j >= p) return;
  const double shift = double(X[(long long)j * n]);
  const T* c = hx + (long long)j * hn;
  double s1 = 0.0, s2 = 0.0;
  for (long long i = lane; i < hn; i += 64) {
    const double d = double(c[i]) - shift;
    s1 += d;
    s2 = fma(d, d, s2);
  }
  for (int off = 32; off > 0; off >>= 1) {
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  if (lane == 0) {
    out[j] = s1;
    out[p + j] = s2;
  }
}

__global__ __launch_bounds__(kWG) void k_zero_rows(double* t, long long ld, int cols,
                                                   const int* rows, int count) {
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < (long long)count * cols;
       e += (long long)gridDim.x * kWG) {
    const long long c = e / count, i = e % count;
    t[rows[i] + c * ld] = 0.0;
  }
}

// dst (count x cols) = src[idx[i], j] for a column-major src with n rows.
template<class T>
__global__ __launch_bounds__(kWG) void k_gather_rows(const T* src, long long n, long long cols,
                                                     const int* idx, long long count, T* dst) {
  // columns over grid y, rows over threads (no index division)
  for (long long j = blockIdx.y; j < cols; j += gridDim.y)
    for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < count;
         i += (long long)gridDim.x * kWG)
      dst[i + j * count] = src[(long long)idx[i] + j * n];
}

// dst (rows x cols, leading dimension ldd) = src (leading dimension lds).
template<class T>
__global__ __launch_bounds__(kWG) void k_copy_block(const T* src, long long lds, long long rows,
                                                    long long cols, T* dst, long long ldd) {
  for (long long j = blockIdx.y; j < cols; j += gridDim.y)
    for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < rows;
         i += (long long)gridDim.x * kWG)
      dst[i + j * ldd] = src[i + j * lds];
}

__global__ __launch_bounds__(kWG) void k_div_rows(const double* in, const double* d,
                                                  long long rows, int cols, double* out) {
  const long long total = rows * cols;
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < total;
       e += (long long)gridDim.x * kWG) {
    out[e] = in[e] / d[e % rows];
  }
}

// out (n x cols, FP64) = (X[:, c0 + c] - mean[c0 + c]) / scale[c0 + c]
template<class T>
__global__ __launch_bounds__(kWG) void k_standardize_columns(const T* X, long long n, int c0,
                                                             int cols, const double* mean,
                                                             const double* scale, double* out) {
  const long long total = n * cols;
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < total;
       e += (long long)gridDim.x * kWG) {
    const long long c = e / n, i = e - c * n;
    const long long j = c0 + c;
    out[e] = (double(X[i + j * n]) - mean[j]) / scale[j];
  }
}

__global__ __launch_bounds__(kWG) void k_div_rows_list(const DivRows* list) {
  const DivRows t = list[blockIdx.y];
  const long long total = t.rows * t.cols;
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < total;
       e += (long long)gridDim.x * kWG) {
    t.out[e] = t.in[e] / t.d[e % t.rows];
  }
}

// ----------------------------------------------------------- prediction
// out = Tsc[:, :k] * M + bias with M(l, c) = Mr[l * rs + c * cs], one thread
// per output element.
template<class T>
__global__ __launch_bounds__(kWG) void k_apply_small(const double* Tsc, long long rows, int k,
                                                     const double* Mr, long long rs,
                                                     long long cs, int q,
                                                     const double* bias, T* out,
                                                     double* out64) {
  const long long total = rows * q;
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < total;
       e += (long long)gridDim.x * kWG) {
    const long long i = e % rows;
    const int c = int(e / rows);
    double acc = 0.0;
    for (int l = 0; l < k; ++l) acc = fma(Tsc[i + (long long)l * rows], Mr[l * rs + c * cs], acc);
    if (bias) acc += bias[c];
    if (out) out[e] = T(acc);
    if (out64) out64[e] = acc;
  }
}

// Per-row top-k with the insertion rule of cuda_resident_api.cuh:29-53: a
// value enters only when strictly greater than the current last entry and
// moves past strictly smaller ones, so the earlier class wins a tie.
constexpr int kMaxTop = 32;
__global__ __launch_bounds__(kWG) void k_topk(const double* values, long long rows, int K,
                                              const double* constants, int top, int* out) {
  const long long r = (long long)blockIdx.x * kWG + threadIdx.x;
  if (r >= rows) return;
  double bv[kMaxTop];
  int bc[kMaxTop];
  for (int t = 0; t < top; ++t) { bv[t] = -INFINITY; bc[t] = -1; }
  for (int c = 0; c < K; ++c) {
    double v = values[r + (long long)c * rows];
    if (constants) v += constants[c];
    if (v <= bv[top - 1]) continue;
    int pos = top - 1;
    while (pos > 0 && v > bv[pos - 1]) {
      bv[pos] = bv[pos - 1];
      bc[pos] = bc[pos - 1];
      --pos;
    }
    bv[pos] = v;
    bc[pos] = c + 1;
  }
  for (int t = 0; t < top; ++t) out[r + (long long)t * rows] = bc[t];
}

// cuda_resident_metrics.cuh:9-43, one work-group per response column.
template<class T>
__global__ __launch_bounds__(kWG) void k_response_sums(const double* pred, const T* observed,
                                                       const int* labels,
                                                       const double* train_mean, long long n,
                                                       int q, double* out, int* flags) {
  __shared__ double red[kWaves];
  const int c = blockIdx.x;
  int bad = 0;
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += kWG) {
    double y;
    if (labels) {
      const int l = labels[i];
      if (l < 1 || l > q) bad = 1;
      y = l == c + 1 ? 1.0 : 0.0;
    } else {
      y = double(observed[i + (long long)c * n]);
    }
    if (!isfinite(y)) bad = 1;
    s += y;
  }
  const double mean = wg_sum(s, red) / double(n);
  double sse = 0.0, tss = 0.0, oss = 0.0;
  for (long long i = threadIdx.x; i < n; i += kWG) {
    const double y = labels ? (labels[i] == c + 1 ? 1.0 : 0.0) : double(observed[i + (long long)c * n]);
    const double pr = pred[i + (long long)c * n];
    if (!isfinite(pr)) bad = 1;
    const double e = y - pr, d = y - train_mean[c], v = y - mean;
    sse = fma(e, e, sse); tss = fma(d, d, tss); oss = fma(v, v, oss);
  }
  sse = wg_sum(sse, red);
  tss = wg_sum(tss, red);
  oss = wg_sum(oss, red);
  if (threadIdx.x == 0) {
    out[3 * c] = sse; out[3 * c + 1] = tss; out[3 * c + 2] = oss;
  }
  if (bad) atomicOr(flags, kFlagNonFinite);
}

// ------------------------------------------------------ synthetic data
// Philox4x32-10 (Salmon et al., SC'11) keyed by the seed; one counter per
// element, Box-Muller on the first two words.
__device__ inline void philox(unsigned c[4], unsigned k0, unsigned k1) {
  const unsigned M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    const unsigned long long p0 = (unsigned long long)M0 * c[0];
    const unsigned long long p1 = (unsigned long long)M1 * c[2];
    const unsigned n0 = unsigned(p1 >> 32) ^ c[1] ^ k0;
    const unsigned n2 = unsigned(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0; c[1] = unsigned(p1); c[2] = n2; c[3] = unsigned(p0);
    k0 += W0; k1 += W1;
  }
}

__device__ inline double normal_at(unsigned long long seed, unsigned stream,
                                   unsigned long long index) {
  unsigned c[4] = {unsigned(index), unsigned(index >> 32), stream, 0x5eedu};
  philox(c, unsigned(seed), unsigned(seed >> 32));
  const double u1 = (double(c[0]) + 1.0) * (1.0 / 4294967297.0);
  const double u2 = double(c[1]) * (1.0 / 4294967296.0);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// factors: T (n x F, column f scaled by decay^f), P (p x F), Q (q x F)
__global__ __launch_bounds__(kWG) void k_latent_factors(double* Tl, double* Pl, double* Ql,
                                                        long long n, int p, int q, int F,
                                                        double decay,
                                                        unsigned long long seed) {
  const long long total = (n + p + q) * F;
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < total;
       e += (long long)gridDim.x * kWG) {
    if (e < n * F) {
      const int f = int(e / n);
      Tl[e] = normal_at(seed, 0, e) * pow(decay, double(f));
    } else if (e < (n + p) * F) {
      const long long o = e - n * F;
      Pl[o] = normal_at(seed, 1, o);
    } else {
      const long long o = e - (n + p) * F;
      Ql[o] = normal_at(seed, 2, o);
    }
  }
}

template<class T>
__global__ __launch_bounds__(kWG) void k_latent_matrix(T* X, const double* Tl,
                                                       const double* Pl, long long n,
                                                       long long cols, int F, double noise,
                                                       unsigned long long seed,
                                                       unsigned stream, int round32) {
  const long long total = n * cols;
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < total;
       e += (long long)gridDim.x * kWG) {
    const long long i = e % n, j = e / n;
    double acc = 0.0;
    for (int f = 0; f < F; ++f) acc = fma(Tl[i + (long long)f * n], Pl[j + (long long)f * cols], acc);
    const double v = acc + noise * normal_at(seed, stream, (unsigned long long)e);
    X[e] = T(round32 ? double(float(v)) : v);
  }
}

// out (cols x rows) = inᵀ for a column-major rows x cols input.
__global__ __launch_bounds__(kWG) void k_transpose(const double* in, long long rows,
                                                   long long cols, double* out) {
  const long long total = rows * cols;
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < total;
       e += (long long)gridDim.x * kWG) {
    const long long i = e % rows, j = e / rows;
    out[j + i * cols] = in[e];
  }
}

template<class T>
__global__ __launch_bounds__(kWG) void k_convert(const double* in, T* out, long long count) {
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < count;
       e += (long long)gridDim.x * kWG) out[e] = T(in[e]);
}

unsigned grid_for(long long count) {
  long long b = (count + kWG - 1) / kWG;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return unsigned(b);
}

}  // namespace

template<class T>
void launch_center_response(const T* Y, long long n, int q, double* mean, double* Yc,
                            double* ysum, double* scratch, hipStream_t s) {
  static_assert(kCenterBlocks * 2 == kCenterScratch, "centring scratch layout");
  if (q <= 0) return;
  double* part = scratch;
  hipLaunchKernelGGL((k_center_sums<T>), dim3(kCenterBlocks, q), dim3(kWG), 0, s, Y, n, part);
  hipLaunchKernelGGL((k_center_apply<T>), dim3(kCenterBlocks, q), dim3(kWG), 0, s, Y, n, part, Yc,
                     part + (long long)q * kCenterBlocks);
  hipLaunchKernelGGL(k_center_fold, dim3((q + kWG - 1) / kWG), dim3(kWG), 0, s, part,
                     part + (long long)q * kCenterBlocks, n, q, mean, ysum);
  FPLS_LAUNCHED("k_center_fold");
}

template<class T>
void launch_center_columns(const CenterColumn<T>* device_list, int count, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL((k_center_columns<T>), dim3(unsigned(count)), dim3(kWG), 0, s, device_list);
  FPLS_LAUNCHED("k_center_columns");
}

void launch_axpy(double alpha, const double* x, double* y, long long count, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(count)), dim3(kWG), 0, s, alpha, x, y, count);
  FPLS_LAUNCHED("k_axpy");
}

void launch_unit_columns(double* v, long long rows, int cols, int* stop, hipStream_t s) {
  if (cols <= 0 || rows <= 0) return;
  hipLaunchKernelGGL(k_unit_columns, dim3(unsigned(cols)), dim3(kWG), 0, s, v, rows, stop);
  FPLS_LAUNCHED("k_unit_columns");
}

void launch_fill_int(int* p, int count, int value, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(count)), dim3(kWG), 0, s, p, count, value);
  FPLS_LAUNCHED("k_fill_int");
}

void launch_label_response(const int* labels, long long n, int K, int* counts,
                           double* priors, double* Yc, double* H, double* ysum, int* flags,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(K)), dim3(kWG), 0, s, counts, K, 0);
  hipLaunchKernelGGL(k_label_counts, dim3(grid_for(n)), dim3(kWG), 0, s, labels, n, K, counts,
                     flags);
  hipLaunchKernelGGL(k_label_fill, dim3(K), dim3(kWG), 0, s, labels, n, K, counts, priors, Yc,
                     H, ysum, flags);
  FPLS_LAUNCHED("k_label_fill");
}

template<class T>
void launch_finalize_crossprod(const T* X, long long n, int p, int q, int scaling,
                               const double* stats_part, const double* s_part, int splits,
                               const double* ysum, double* mean, double* scale,
                               double* colss, double* S, double* S0, int* flags,
                               hipStream_t s, int* cross_flags, const int* cross_offsets,
                               const double* hold_stats, long long used) {
  const unsigned blocks = unsigned((p + kWG - 1) / kWG);
  hipLaunchKernelGGL((k_finalize_stats<T>), dim3(blocks), dim3(kWG), 0, s, X, n, p, scaling,
                     stats_part, splits, mean, scale, colss, flags, hold_stats, used);
  if (q > 0)
    hipLaunchKernelGGL((k_finalize_cross<T>), dim3(blocks, q), dim3(kWG), 0, s, X, n, p, q, s_part,
                       splits, ysum, mean, scale, S, S0,
                       cross_offsets ? cross_flags : flags, cross_offsets);
  FPLS_LAUNCHED("k_finalize_cross");
}

template<class T>
void launch_holdout_stats(const T* hx, long long hn, int p, const T* X, long long n, double* out,
                          hipStream_t s) {
  if (p <= 0) return;
  hipLaunchKernelGGL((k_holdout_stats<T>), dim3(unsigned((p + kWaves - 1) / kWaves)), dim3(kWG), 0,
                     s, hx, hn, p, X, n, out);
  FPLS_LAUNCHED("k_holdout_stats");
}

void launch_zero_rows(double* t, long long ld, int cols, const int* rows, int count,
                      hipStream_t s) {
  if (count <= 0 || cols <= 0) return;
  const long long total = (long long)count * cols;
  const unsigned blocks = unsigned(std::min<long long>((total + kWG - 1) / kWG, 1024));
  hipLaunchKernelGGL(k_zero_rows, dim3(blocks), dim3(kWG), 0, s, t, ld, cols, rows, count);
  FPLS_LAUNCHED("k_zero_rows");
}

template<class T>
void launch_gather_rows(const T* src, long long n, long long cols, const int* idx,
                        long long count, T* dst, hipStream_t s) {
  if (count <= 0 || cols <= 0) return;
  if (count <= 0 || cols <= 0) return;
  const long long bx = std::min<long long>((count + kWG - 1) / kWG, 64);
  const long long by = std::min<long long>(cols, 65535);
  hipLaunchKernelGGL((k_gather_rows<T>), dim3(unsigned(bx), unsigned(by)), dim3(kWG), 0, s, src,
                     n, cols, idx, count, dst);
  FPLS_LAUNCHED("k_gather_rows");
}

template<class T>
void launch_copy_block(const T* src, long long lds, long long rows, long long cols, T* dst,
                       long long ldd, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  const long long bx = std::min<long long>((rows + kWG - 1) / kWG, 64);
  const long long by = std::min<long long>(cols, 65535);
  hipLaunchKernelGGL((k_copy_block<T>), dim3(unsigned(bx), unsigned(by)), dim3(kWG), 0, s, src,
                     lds, rows, cols, dst, ldd);
  FPLS_LAUNCHED("k_copy_block");
}
template void launch_copy_block<float>(const float*, long long, long long, long long, float*,
                                       long long, hipStream_t);
template void launch_copy_block<double>(const double*, long long, long long, long long, double*,
                                        long long, hipStream_t);
template void launch_gather_rows<float>(const float*, long long, long long, const int*, long long,
                                        float*, hipStream_t);
template void launch_gather_rows<double>(const double*, long long, long long, const int*,
                                         long long, double*, hipStream_t);
template void launch_gather_rows<int>(const int*, long long, long long, const int*, long long,
                                      int*, hipStream_t);

void launch_div_rows(const double* in, const double* d, long long rows, int cols,
                     double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_div_rows, dim3(grid_for(rows * cols)), dim3(kWG), 0, s, in, d, rows,
                     cols, out);
  FPLS_LAUNCHED("k_div_rows");
}

void launch_div_rows_list(const DivRows* device_list, int count, long long max_elements,
                          hipStream_t s) {
  if (count <= 0 || max_elements <= 0) return;
  const unsigned bx = std::min<unsigned>(grid_for(max_elements), 64);
  for (int c0 = 0; c0 < count; c0 += 65535) {
    const int c = std::min(65535, count - c0);
    hipLaunchKernelGGL(k_div_rows_list, dim3(bx, unsigned(c)), dim3(kWG), 0, s, device_list + c0);
    FPLS_LAUNCHED("k_div_rows_list");
  }
}

template<class T>
void launch_apply_small(const double* Tsc, long long rows, int k, const double* Mr,
                        long long rs, long long cs, int q, const double* bias, T* out,
                        double* out64, hipStream_t s) {
  hipLaunchKernelGGL((k_apply_small<T>), dim3(grid_for(rows * q)), dim3(kWG), 0, s, Tsc, rows,
                     k, Mr, rs, cs, q, bias, out, out64);
  FPLS_LAUNCHED("k_apply_small");
}

void launch_topk(const double* values, long long rows, int K, const double* constants,
                 int top, int* out, hipStream_t s) {
  if (top < 1 || top > kMaxTop || top > K) throw std::invalid_argument("top must lie in 1..min(classes, 32)");
  hipLaunchKernelGGL(k_topk, dim3(unsigned((rows + kWG - 1) / kWG)), dim3(kWG), 0, s, values,
                     rows, K, constants, top, out);
  FPLS_LAUNCHED("k_topk");
}

template<class T>
void launch_response_sums(const double* pred, const T* observed, const int* labels,
                          const double* train_mean, long long n, int q, double* out,
                          int* flags, hipStream_t s) {
  hipLaunchKernelGGL((k_response_sums<T>), dim3(q), dim3(kWG), 0, s, pred, observed, labels,
                     train_mean, n, q, out, flags);
  FPLS_LAUNCHED("k_response_sums");
}

template<class T>
void launch_latent_data(T* X, double* Y, long long n, int p, int q, int factors, double decay,
                        double noise, unsigned long long seed, int round32, double* scratch,
                        hipStream_t s) {
  double* Tl = scratch;
  double* Pl = Tl + n * factors;
  double* Ql = Pl + (long long)p * factors;
  hipLaunchKernelGGL(k_latent_factors, dim3(grid_for((n + p + q) * factors)), dim3(kWG), 0, s,
                     Tl, Pl, Ql, n, p, q, factors, decay, seed);
  hipLaunchKernelGGL((k_latent_matrix<T>), dim3(grid_for(n * p)), dim3(kWG), 0, s, X, Tl, Pl,
                     n, (long long)p, factors, noise, seed, 3u, round32);
  hipLaunchKernelGGL((k_latent_matrix<double>), dim3(grid_for(n * q)), dim3(kWG), 0, s, Y, Tl,
                     Ql, n, (long long)q, factors, noise, seed, 4u, round32);
  FPLS_LAUNCHED("k_latent_matrix");
}

void launch_transpose(const double* in, long long rows, long long cols, double* out,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_transpose, dim3(grid_for(rows * cols)), dim3(kWG), 0, s, in, rows, cols,
                     out);
  FPLS_LAUNCHED("k_transpose");
}

template<class T>
void launch_standardize_columns(const T* X, long long n, int c0, int cols, const double* mean,
                                const double* scale, double* out, hipStream_t s) {
  if (n <= 0 || cols <= 0) return;
  hipLaunchKernelGGL((k_standardize_columns<T>), dim3(grid_for(n * cols)), dim3(kWG), 0, s, X, n,
                     c0, cols, mean, scale, out);
  FPLS_LAUNCHED("k_standardize_columns");
}

template<class T>
void launch_convert(const double* in, T* out, long long count, hipStream_t s) {
  hipLaunchKernelGGL((k_convert<T>), dim3(grid_for(count)), dim3(kWG), 0, s, in, out, count);
  FPLS_LAUNCHED("k_convert");
}

#define FPLS_PREP_INSTANCES(T)                                                              \
  template void launch_center_response<T>(const T*, long long, int, double*, double*,        \
                                          double*, double*, hipStream_t);                    \
  template void launch_center_columns<T>(const CenterColumn<T>*, int, hipStream_t);          \
  template void launch_finalize_crossprod<T>(const T*, long long, int, int, int,             \
                                             const double*, const double*, int,              \
                                             const double*, double*, double*, double*,       \
                                             double*, double*, int*, hipStream_t, int*,      \
                                             const int*, const double*, long long);          \
  template void launch_holdout_stats<T>(const T*, long long, int, const T*, long long,       \
                                        double*, hipStream_t);                               \
  template void launch_apply_small<T>(const double*, long long, int, const double*,          \
                                      long long, long long, int, const double*, T*, double*, \
                                      hipStream_t);                                          \
  template void launch_response_sums<T>(const double*, const T*, const int*, const double*,  \
                                        long long, int, double*, int*, hipStream_t);         \
  template void launch_latent_data<T>(T*, double*, long long, int, int, int, double, double, \
                                      unsigned long long, int, double*, hipStream_t);             \
  template void launch_convert<T>(const double*, T*, long long, hipStream_t);                \
  template void launch_standardize_columns<T>(const T*, long long, int, int, const double*,  \
                                              const double*, double*, hipStream_t);
FPLS_PREP_INSTANCES(float)
FPLS_PREP_INSTANCES(double)
#undef FPLS_PREP_INSTANCES

}  // namespace fpls
