// Multi-work-group streaming products over a tall column-major matrix.
//
// Two shapes cover every pass over the predictor matrix X (and the same
// shapes over the cross-covariance S in the randomized SVD):
//   xw : out(m x N)  = (A - 1 cᵀ) · W       W is k x N      (scores, sketches)
//   xtw: out(k x N)  = (A - 1 cᵀ)ᵀ · W      W is m x N      (XᵀY, loadings)
// A is m x k, column-major, leading dimension m, element type T (float or
// double). The optional centre c (length k) is subtracted element-wise as A
// is loaded, so standardisation never needs its own pass over A.
// Both kernels accept a task list so that many independent problems (fits of
// a batch) run in one launch; blockIdx.x indexes a flattened tile space.
#pragma once

#include "common.hpp"

namespace fpls {

enum XtwMode : int {
  kPlain = 0,      // plain product
  kCentered = 1,   // subtract centre[j] from column j
  kShiftStats = 2  // subtract A[0, j]; also return sum d and sum d^2 per column
};

template<class T>
struct XwTask {
  const T* A;            // m x k
  long long m, k;
  const double* center;  // k or null
  const double* W;       // k x N, leading dimension ldw
  void* frag;            // matrix-core operand workspace, xw_frag_bytes<T>(k, maxn) bytes
  long long ldw;
  int N;
  int splits;            // column splits (k dimension)
  double* out;           // splits slabs of m x N
  long long ldo;         // leading dimension of out (>= m)
  long long slab;        // distance between split slabs of out
  long long tile0;       // first flattened tile of this task
};

template<class T>
struct XtwTask {
  const T* A;            // m x k
  long long m, k;
  const double* center;  // k or null (kCentered)
  const double* W;       // m x N, leading dimension ldw
  void* frag;            // matrix-core operand workspace, xtw_frag_bytes<T>(m, maxn) bytes
  long long ldw;
  int N;
  int mode;
  int splits;            // row splits (m dimension)
  double* out;           // splits slabs of k x N
  long long ldo;         // leading dimension of out (>= k)
  long long slab;        // distance between split slabs of out
  double* stats;         // kShiftStats: splits slabs of 2 x k (sum, sum of squares)
  long long tile0;
};

// Host side (implemented in stream_kernels.hip). pack_* assigns tile0 to a
// host task array, checks it and returns the grid size; the caller uploads
// the packed array once and launches with the device copy, so launches carry
// no host-to-device traffic and can be captured in a graph.
//
// `f64` selects the products for FP32 A on the matrix cores: false runs them
// in bf16x6 split precision (FP32 sums per 32- to 64-long step), true widens
// A to FP64 as it is loaded and runs them on the FP64 matrix cores (FP64
// products and sums throughout; no operand workspace). FP64 A always takes
// the FP64 route. `min_width` makes pack_xtw plan the list as if its widest
// task had at least that many right-hand columns (a list of narrow tasks can
// so be sent to the matrix cores). `stats` tells launch_xtw that the list's
// tasks are kShiftStats tasks (a list is all of them or none).
template<class T>
long long pack_xw(XwTask<T>* tasks, int count, int* maxn, bool f64 = false);
template<class T>
long long pack_xtw(XtwTask<T>* tasks, int count, int* maxn, bool f64 = false, int min_width = 0);
template<class T>
void launch_xw(const XwTask<T>* device_tasks, int count, long long tiles, int maxn,
               hipStream_t stream, bool f64 = false);
template<class T>
void launch_xtw(const XtwTask<T>* device_tasks, int count, long long tiles, int maxn,
                hipStream_t stream, bool f64 = false, bool stats = false);
// One task passed by value as a kernel argument (no device task array).
template<class T>
void launch_xw_single(XwTask<T> task, hipStream_t stream, bool f64 = false);
template<class T>
void launch_xtw_single(XtwTask<T> task, hipStream_t stream, bool f64 = false);

// Number of 256-row tiles of an xw task and of 64-column tiles of xtw.
__host__ __device__ inline long long xw_row_tiles(long long m) { return (m + 255) / 256; }
__host__ __device__ inline long long xtw_col_tiles(long long k) { return (k + 63) / 64; }
// Right-hand sides of this many columns or more run on the matrix cores.
constexpr int kMfmaMinN = 8;
// Column tiles of an xtw launch whose widest task has maxn columns: the
// matrix-core kernel owns 128 columns per work-group, the VALU kernel 64.
__host__ __device__ inline long long xtw_tiles(long long k, int maxn) {
  return maxn >= kMfmaMinN ? (k + 127) / 128 : (k + 63) / 64;
}
// Longest row split of an FP64 matrix-core xtw and longest column split of an
// FP64 matrix-core xw, and of every split of the batched launches.
constexpr long long kXtwSplitRows = 4096;
constexpr long long kXwSplitCols = 4096;

// Right-hand width a matrix-core launch is compiled for.
inline int mfma_width(int maxn) { return maxn <= 16 ? 16 : maxn <= 32 ? 32 : maxn <= 48 ? 48 : 64; }

// Bytes of the operand workspace (XwTask::frag / XtwTask::frag) a task needs
// when the widest task of its launch has maxn right-hand columns; 0 means the
// launch does not read it. For FP32 A the matrix-core kernels run in bf16x6
// split precision (three bf16 terms per value, six products) and W is split
// once per launch into its high, middle and low bf16 parts, laid out in
// fragment order:
//   xw : [hi, mid, lo][NB][k rounded up to 64] (W transposed) + FP32 centre [k64]
//   xtw: [hi, mid, lo][NB][m rounded up to 64], rows permuted within 64-row tiles
constexpr int kSplitTerms = 3;
template<class T>
inline size_t xw_frag_bytes(long long k, int maxn, bool f64 = false) {
  if (sizeof(T) != 4 || f64 || maxn < kMfmaMinN) return 0;
  const size_t k64 = size_t((k + 63) / 64 * 64);
  return k64 * (size_t)mfma_width(maxn) * 2 * kSplitTerms + k64 * 4;
}
template<class T>
inline size_t xtw_frag_bytes(long long m, int maxn, bool f64 = false) {
  if (sizeof(T) != 4 || f64 || maxn < kMfmaMinN) return 0;
  const size_t m64 = size_t((m + 63) / 64 * 64);
  return m64 * (size_t)mfma_width(maxn) * 2 * kSplitTerms;
}


// Deterministic sum of `splits` slabs: out[i] = sum_s part[s * stride + i].
void launch_sum_slabs(const double* part, int splits, long long stride,
                      long long count, double* out, hipStream_t stream);

// The same over a device list of independent slab sums (one launch).
struct SlabSum {
  const double* part;
  double* out;
  long long stride, count;
  int splits;
};
void launch_sum_slabs_list(const SlabSum* device_list, int count, long long max_count,
                           hipStream_t stream);

}  // namespace fpls
