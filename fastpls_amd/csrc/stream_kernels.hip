// Streaming products over tall column-major matrices (see stream.hpp).
//
// Roofline: both kernels read A once per launch (m*k*sizeof(T) bytes) and are
// meant to run at HBM bandwidth; W tiles come from L2 through LDS. The
// arithmetic is FP64 FMA on values widened from T, so the FP32 storage path
// loses nothing to accumulation.
#include "stream.hpp"

#include <algorithm>

namespace fpls {

namespace {

template<class Task>
__device__ inline int find_task(const Task* tasks, int count, long long block) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].tile0 <= block) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------------- xw
// One thread per row of A; the k range of the split is walked in chunks of
// 32 columns whose W rows (and centres) are staged in LDS and broadcast.
template<class T, int NB>
__global__ __launch_bounds__(kWG) void k_xw(const XwTask<T>* tasks, int count, XwTask<T> single) {
  const XwTask<T>& t = tasks ? tasks[find_task(tasks, count, blockIdx.x)] : single;
  const long long local = (long long)blockIdx.x - t.tile0;
  const long long rtiles = xw_row_tiles(t.m);
  if (local >= rtiles * t.splits) return;
  const int split = int(local / rtiles);
  const long long row = (local % rtiles) * kWG + threadIdx.x;
  const long long per = (t.k + t.splits - 1) / t.splits;
  const long long j0 = split * per;
  const long long j1 = j0 + per < t.k ? j0 + per : t.k;

  __shared__ double ws[32][NB];
  __shared__ double cs[32];
  double acc[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) acc[c] = 0.0;
  const bool valid = row < t.m;
  const T* a = t.A + (valid ? row : 0);

  for (long long jc = j0; jc < j1; jc += 32) {
    const int cnt = int(j1 - jc < 32 ? j1 - jc : 32);
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * NB; e += kWG) {
      const int jj = e & 31, c = e >> 5;
      ws[jj][c] = (jj < cnt && c < t.N) ? t.W[(jc + jj) + (long long)c * t.ldw] : 0.0;
    }
    if (threadIdx.x < 32) {
      cs[threadIdx.x] = (t.center != nullptr && int(threadIdx.x) < cnt) ? t.center[jc + threadIdx.x] : 0.0;
    }
    __syncthreads();
    if (valid) {
      if (cnt == 32) {
#pragma unroll 8
        for (int jj = 0; jj < 32; ++jj) {
          const double x = double(a[(jc + jj) * t.m]) - cs[jj];
#pragma unroll
          for (int c = 0; c < NB; ++c) acc[c] = fma(x, ws[jj][c], acc[c]);
        }
      } else {
        for (int jj = 0; jj < cnt; ++jj) {
          const double x = double(a[(jc + jj) * t.m]) - cs[jj];
#pragma unroll
          for (int c = 0; c < NB; ++c) acc[c] = fma(x, ws[jj][c], acc[c]);
        }
      }
    }
  }
  if (valid) {
    double* o = t.out + (long long)split * t.slab + row;
#pragma unroll
    for (int c = 0; c < NB; ++c) if (c < t.N) o[(long long)c * t.ldo] = acc[c];
  }
}

// ------------------------------------------------------------------ xtw
// A block owns 64 columns of A and one row split. Row tiles of 32 are staged
// in LDS (A widened to double and centred on the way in, W as is); each
// thread then owns an MC x MN patch of the 64 x NB output.
template<class T, int NB>
__global__ __launch_bounds__(kWG) void k_xtw(const XtwTask<T>* tasks, int count, XtwTask<T> single) {
  constexpr int MN = NB >= 16 ? NB / 16 : 1;
  constexpr int MC = (NB / 4) / MN;
  constexpr int GROUPS = 64 / MC;  // column groups
  const XtwTask<T>& t = tasks ? tasks[find_task(tasks, count, blockIdx.x)] : single;
  const long long local = (long long)blockIdx.x - t.tile0;
  const long long ctiles = xtw_col_tiles(t.k);
  if (local >= ctiles * t.splits) return;
  const int split = int(local / ctiles);
  const long long col0 = (local % ctiles) * 64;
  const long long per = (t.m + t.splits - 1) / t.splits;
  const long long r0 = split * per;
  const long long r1 = r0 + per < t.m ? r0 + per : t.m;

  __shared__ double xs[32][65];
  __shared__ double wsh[32][NB + 1];

  const int tid = threadIdx.x;
  const int lrow = tid & 31;          // staging row
  const int lcol = tid >> 5;          // staging column base (0..7)
  // Per staged column: centre (or shift) and running stats.
  double centre[8], s1[8], s2[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const long long col = col0 + lcol + 8 * e;
    double c = 0.0;
    if (col < t.k) {
      if (t.mode == kCentered && t.center != nullptr) c = t.center[col];
      else if (t.mode == kShiftStats) c = double(t.A[col * t.m]);
    }
    centre[e] = c; s1[e] = 0.0; s2[e] = 0.0;
  }
  const int g = tid % GROUPS, h = tid / GROUPS;
  double acc[MC][MN];
#pragma unroll
  for (int a = 0; a < MC; ++a)
#pragma unroll
    for (int b = 0; b < MN; ++b) acc[a][b] = 0.0;

  for (long long i0 = r0; i0 < r1; i0 += 32) {
    const long long row = i0 + lrow;
    const bool rv = row < r1;
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long long col = col0 + lcol + 8 * e;
      double v = 0.0;
      if (rv && col < t.k) {
        v = double(t.A[row + col * t.m]) - centre[e];
        s1[e] += v;
        s2[e] = fma(v, v, s2[e]);
      }
      xs[lrow][lcol + 8 * e] = v;
    }
    for (int e = tid; e < 32 * NB; e += kWG) {
      const int r = e & 31, c = e >> 5;
      const long long wr = i0 + r;
      wsh[r][c] = (wr < r1 && c < t.N) ? t.W[wr + (long long)c * t.ldw] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int r = 0; r < 32; ++r) {
      double xa[MC], wb[MN];
#pragma unroll
      for (int a = 0; a < MC; ++a) xa[a] = xs[r][g * MC + a];
#pragma unroll
      for (int b = 0; b < MN; ++b) wb[b] = wsh[r][h * MN + b];
#pragma unroll
      for (int a = 0; a < MC; ++a)
#pragma unroll
        for (int b = 0; b < MN; ++b) acc[a][b] = fma(xa[a], wb[b], acc[a][b]);
    }
  }
  double* o = t.out + (long long)split * t.slab;
#pragma unroll
  for (int a = 0; a < MC; ++a) {
    const long long col = col0 + g * MC + a;
#pragma unroll
    for (int b = 0; b < MN; ++b) {
      const int n = h * MN + b;
      if (col < t.k && n < t.N) o[col + (long long)n * t.ldo] = acc[a][b];
    }
  }
  if (t.mode == kShiftStats && t.stats != nullptr) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      double a1 = s1[e], a2 = s2[e];
      for (int off = 1; off < 32; off <<= 1) {
        a1 += __shfl_xor(a1, off, 64);
        a2 += __shfl_xor(a2, off, 64);
      }
      const long long col = col0 + lcol + 8 * e;
      if (lrow == 0 && col < t.k) {
        double* st = t.stats + (long long)split * 2 * t.k;
        st[col] = a1;
        st[t.k + col] = a2;
      }
    }
  }
}

// ------------------------------------------------------------ MFMA paths
// Matrix-core versions of the two products for right-hand sides of 8..64
// columns. FP32 inputs use v_mfma_f32_16x16x4f32 and FP64 inputs
// v_mfma_f64_16x16x4f64; FP32 accumulators are folded into FP64 ones every
// 256 rows (xtw) or 64 columns (xw), so long reductions keep FP64 accuracy.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Output layout of the 16x16 tiles (measured on gfx950): item r of lane l
// holds D[row(l / 16, r)][l % 16], with row(g, r) = 4g + r for FP32 and
// g + 4r for FP64.
template<class T> struct Mma;
template<> struct Mma<float> {
  using acc_t = f32x4;
  __device__ static int row(int g, int r) { return 4 * g + r; }
  __device__ static acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template<> struct Mma<double> {
  using acc_t = f64x4;
  __device__ static int row(int g, int r) { return g + 4 * r; }
  __device__ static acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// xtw on matrix cores. A block owns 128 columns of A (32 per wave, two
// 16-column M-blocks) and one row split of at most kXtwSplitRows rows, over
// which the FP32 accumulators run before being written as FP64 partials.
// Each iteration covers 64 rows: lane (m, g) = (lane % 16, lane / 16) holds
// rows 16g..16g+15 of column m of each M-block (all loads issued before the
// first MFMA), and MFMA call u pairs row 16g+u of A with the same row of W.
// W tiles (64 rows x NB) are double-buffered in LDS: one barrier per
// iteration.
constexpr int kXtwMfmaCols = 128;
constexpr int kXtwRows = 64;

template<class T, int NB>
__global__ __launch_bounds__(kWG) void k_xtw_mfma(const XtwTask<T>* tasks, int count,
                                                  XtwTask<T> single) {
  using M = Mma<T>;
  constexpr int NBLK = NB / 16;
  constexpr int LD = kXtwRows + 4;
  const XtwTask<T>& t = tasks ? tasks[find_task(tasks, count, blockIdx.x)] : single;
  const long long local = (long long)blockIdx.x - t.tile0;
  const long long ctiles = (t.k + kXtwMfmaCols - 1) / kXtwMfmaCols;
  if (local >= ctiles * t.splits) return;
  const int split = int(local / ctiles);
  const long long col0 = (local % ctiles) * kXtwMfmaCols;
  long long per = (t.m + t.splits - 1) / t.splits;
  per = (per + kXtwRows - 1) / kXtwRows * kXtwRows;
  const long long r0 = split * per;
  const long long r1 = r0 + per < t.m ? r0 + per : t.m;

  __shared__ T bs[2][NB][LD];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m16 = lane & 15, g = lane >> 4;
  long long col[2];
  bool cvalid[2];
  T centre[2];
  const T* xcol[2];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
    col[mb] = col0 + wave * 32 + mb * 16 + m16;
    cvalid[mb] = col[mb] < t.k;
    double c = 0.0;
    if (cvalid[mb]) {
      if (t.mode == kCentered && t.center != nullptr) c = t.center[col[mb]];
      else if (t.mode == kShiftStats) c = double(t.A[col[mb] * t.m]);
    }
    centre[mb] = T(c);
    xcol[mb] = t.A + (cvalid[mb] ? col[mb] : 0) * t.m;
  }
  const bool vec_ok = (t.m % (16 / sizeof(T))) == 0;
  const T* Wt = t.Wt;

  // stage W rows [i0, i0 + 64) into buffer b (zero outside [r0, r1) or N)
  auto stage = [&](int b, long long i0) {
    for (int e = threadIdx.x; e < kXtwRows * NB; e += kWG) {
      const int r = e % kXtwRows, c = e / kXtwRows;
      const long long wr = i0 + r;
      T v = T(0);
      if (wr < r1 && c < t.N) v = Wt ? Wt[wr + (long long)c * t.ldw] : T(t.W[wr + (long long)c * t.ldw]);
      bs[b][c][r] = v;
    }
  };

  typename M::acc_t acc[2][NBLK];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) acc[mb][nb] = M::zero();
  double s1[2] = {0.0, 0.0}, s2[2] = {0.0, 0.0};

  if (r0 < r1) stage(0, r0);
  __syncthreads();
  int buf = 0;
  for (long long i0 = r0; i0 < r1; i0 += kXtwRows) {
    T a[2][16];
    const long long rb = i0 + 16 * g;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      if (cvalid[mb] && vec_ok && rb + 16 <= r1) {
        if constexpr (sizeof(T) == 4) {
#pragma unroll
          for (int u = 0; u < 16; u += 4) {
            const float4 v = *reinterpret_cast<const float4*>(xcol[mb] + rb + u);
            a[mb][u] = v.x; a[mb][u + 1] = v.y; a[mb][u + 2] = v.z; a[mb][u + 3] = v.w;
          }
        } else {
#pragma unroll
          for (int u = 0; u < 16; u += 2) {
            const double2 v = *reinterpret_cast<const double2*>(xcol[mb] + rb + u);
            a[mb][u] = v.x; a[mb][u + 1] = v.y;
          }
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) a[mb][u] -= centre[mb];
      } else {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const long long row = rb + u;
          a[mb][u] = (cvalid[mb] && row < r1) ? T(xcol[mb][row] - centre[mb]) : T(0);
        }
      }
    }
    // prefetch the next W tile into the other buffer while this one is used
    if (i0 + kXtwRows < r1) stage(buf ^ 1, i0 + kXtwRows);
    if (t.mode == kShiftStats) {
#pragma unroll
      for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const double v = double(a[mb][u]);
          s1[mb] += v;
          s2[mb] = fma(v, v, s2[mb]);
        }
    }
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) {
      T b[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) b[u] = bs[buf][nb * 16 + m16][16 * g + u];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        acc[0][nb] = M::mma(a[0][u], b[u], acc[0][nb]);
        acc[1][nb] = M::mma(a[1][u], b[u], acc[1][nb]);
      }
    }
    __syncthreads();
    buf ^= 1;
  }
  double* o = t.out + (long long)split * t.slab;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) {
      const int n = nb * 16 + m16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long oc = col0 + wave * 32 + mb * 16 + M::row(g, r);
        if (oc < t.k && n < t.N) o[oc + (long long)n * t.ldo] = double(acc[mb][nb][r]);
      }
    }
  if (t.mode == kShiftStats && t.stats != nullptr) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      double a1 = s1[mb], a2 = s2[mb];
      a1 += __shfl_xor(a1, 16, 64);
      a1 += __shfl_xor(a1, 32, 64);
      a2 += __shfl_xor(a2, 16, 64);
      a2 += __shfl_xor(a2, 32, 64);
      if (g == 0 && cvalid[mb]) {
        double* st = t.stats + (long long)split * 2 * t.k;
        st[col[mb]] = a1;
        st[t.k + col[mb]] = a2;
      }
    }
  }
}

// xw on matrix cores. A block owns 256 rows (64 per wave) and one column
// split; the split is walked in chunks of 64 columns whose W rows are staged
// in LDS. Within a chunk lane (m, kk) = (lane % 16, lane / 16) loads rows
// 4m..4m+3 of columns 4s+kk for all 16 K-steps s before the first MFMA, and
// MFMA call e of step s produces output rows {4m + e}. FP32 accumulators are
// folded into FP64 ones after every chunk.
template<class T, int NB>
__global__ __launch_bounds__(kWG) void k_xw_mfma(const XwTask<T>* tasks, int count,
                                                 XwTask<T> single) {
  using M = Mma<T>;
  constexpr int NBLK = NB / 16;
  constexpr int PAD = 1;
  const XwTask<T>& t = tasks ? tasks[find_task(tasks, count, blockIdx.x)] : single;
  const long long local = (long long)blockIdx.x - t.tile0;
  const long long rtiles = xw_row_tiles(t.m);
  if (local >= rtiles * t.splits) return;
  const int split = int(local / rtiles);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m16 = lane & 15, kk = lane >> 4;
  const long long i0 = (local % rtiles) * kWG + wave * 64;
  const long long rowb = i0 + 4 * m16;
  long long per = (t.k + t.splits - 1) / t.splits;
  per = (per + 63) / 64 * 64;
  const long long j0 = split * per;
  const long long j1 = j0 + per < t.k ? j0 + per : t.k;
  const bool vec_ok = (t.m % (16 / sizeof(T))) == 0 && rowb + 4 <= t.m;
  const bool rows_ok = rowb < t.m;

  __shared__ T ws[64][NB + PAD];
  __shared__ T cs[64];
  typename M::acc_t acc[4][NBLK];
  double acc64[4][NBLK][4];
#pragma unroll
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) {
      acc[e][nb] = M::zero();
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[e][nb][r] = 0.0;
    }

  for (long long jc = j0; jc < j1; jc += 64) {
    const int cnt = int(j1 - jc < 64 ? j1 - jc : 64);
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * NB; e += kWG) {
      const int jj = e & 63, c = e >> 6;
      T v = T(0);
      if (jj < cnt && c < t.N)
        v = t.Wt ? t.Wt[(jc + jj) + (long long)c * t.ldw] : T(t.W[(jc + jj) + (long long)c * t.ldw]);
      ws[jj][c] = v;
    }
    if (threadIdx.x < 64)
      cs[threadIdx.x] = (t.center != nullptr && int(threadIdx.x) < cnt) ? T(t.center[jc + threadIdx.x]) : T(0);
    // all 16 K-steps of this chunk in flight at once
    T x[16][4];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int jj = 4 * s + kk;
      const T* src = t.A + (jc + jj) * t.m + rowb;
      if (jj < cnt && vec_ok) {
        if constexpr (sizeof(T) == 4) {
          const float4 v = *reinterpret_cast<const float4*>(src);
          x[s][0] = v.x; x[s][1] = v.y; x[s][2] = v.z; x[s][3] = v.w;
        } else {
          const double2 v0 = *reinterpret_cast<const double2*>(src);
          const double2 v1 = *reinterpret_cast<const double2*>(src + 2);
          x[s][0] = v0.x; x[s][1] = v0.y; x[s][2] = v1.x; x[s][3] = v1.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          x[s][e] = (jj < cnt && rows_ok && rowb + e < t.m) ? src[e] : T(0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int jj = 4 * s + kk;
      const T c = cs[jj];
#pragma unroll
      for (int e = 0; e < 4; ++e) x[s][e] = (jj < cnt && rowb + e < t.m) ? T(x[s][e] - c) : T(0);
#pragma unroll
      for (int nb = 0; nb < NBLK; ++nb) {
        const T b = ws[jj][nb * 16 + m16];
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[e][nb] = M::mma(x[s][e], b, acc[e][nb]);
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int nb = 0; nb < NBLK; ++nb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[e][nb][r] += double(acc[e][nb][r]);
        acc[e][nb] = M::zero();
      }
  }
  double* o = t.out + (long long)split * t.slab;
#pragma unroll
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) {
      const int n = nb * 16 + m16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long row = i0 + 4 * M::row(kk, r) + e;
        if (row < t.m && n < t.N) o[row + (long long)n * t.ldo] = acc64[e][nb][r];
      }
    }
}

__global__ __launch_bounds__(kWG) void k_sum_slabs(const double* part, int splits,
                                                   long long stride, long long count,
                                                   double* out) {
  for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < count;
       i += (long long)gridDim.x * kWG) {
    double s = 0.0;
    for (int k = 0; k < splits; ++k) s += part[k * stride + i];
    out[i] = s;
  }
}

template<class T, int NB>
void run_xw(const XwTask<T>* d, int count, const XwTask<T>& single, long long tiles,
            hipStream_t stream) {
  hipLaunchKernelGGL((k_xw<T, NB>), dim3(unsigned(tiles)), dim3(kWG), 0, stream, d, count,
                     single);
}

template<class T, int NB>
void run_xtw(const XtwTask<T>* d, int count, const XtwTask<T>& single, long long tiles,
             hipStream_t stream) {
  hipLaunchKernelGGL((k_xtw<T, NB>), dim3(unsigned(tiles)), dim3(kWG), 0, stream, d, count,
                     single);
}

template<class T, int NB>
void run_xw_mfma(const XwTask<T>* d, int count, const XwTask<T>& single, long long tiles,
                 hipStream_t stream) {
  hipLaunchKernelGGL((k_xw_mfma<T, NB>), dim3(unsigned(tiles)), dim3(kWG), 0, stream, d, count,
                     single);
}

template<class T, int NB>
void run_xtw_mfma(const XtwTask<T>* d, int count, const XtwTask<T>& single, long long tiles,
                  hipStream_t stream) {
  hipLaunchKernelGGL((k_xtw_mfma<T, NB>), dim3(unsigned(tiles)), dim3(kWG), 0, stream, d,
                     count, single);
}

template<class T>
void dispatch_xw(const XwTask<T>* d, int count, const XwTask<T>& single, long long tiles,
                 int maxn, hipStream_t stream) {
  if (tiles <= 0) return;
  if (tiles > 0x7fffffffLL) throw std::invalid_argument("xw grid too large");
  if (maxn >= kMfmaMinN) {
    if (maxn <= 16) run_xw_mfma<T, 16>(d, count, single, tiles, stream);
    else if (maxn <= 32) run_xw_mfma<T, 32>(d, count, single, tiles, stream);
    else run_xw_mfma<T, 64>(d, count, single, tiles, stream);
    FPLS_LAUNCHED("k_xw_mfma");
    return;
  }
  if (maxn <= 1) run_xw<T, 1>(d, count, single, tiles, stream);
  else if (maxn <= 4) run_xw<T, 4>(d, count, single, tiles, stream);
  else if (maxn <= 16) run_xw<T, 16>(d, count, single, tiles, stream);
  else if (maxn <= 32) run_xw<T, 32>(d, count, single, tiles, stream);
  else run_xw<T, 64>(d, count, single, tiles, stream);
  FPLS_LAUNCHED("k_xw");
}

template<class T>
void dispatch_xtw(const XtwTask<T>* d, int count, const XtwTask<T>& single, long long tiles,
                  int maxn, hipStream_t stream) {
  if (tiles <= 0) return;
  if (tiles > 0x7fffffffLL) throw std::invalid_argument("xtw grid too large");
  if (maxn >= kMfmaMinN) {
    if (maxn <= 16) run_xtw_mfma<T, 16>(d, count, single, tiles, stream);
    else if (maxn <= 32) run_xtw_mfma<T, 32>(d, count, single, tiles, stream);
    else run_xtw_mfma<T, 64>(d, count, single, tiles, stream);
    FPLS_LAUNCHED("k_xtw_mfma");
    return;
  }
  if (maxn <= 4) run_xtw<T, 4>(d, count, single, tiles, stream);
  else if (maxn <= 16) run_xtw<T, 16>(d, count, single, tiles, stream);
  else run_xtw<T, 64>(d, count, single, tiles, stream);
  FPLS_LAUNCHED("k_xtw");
}

}  // namespace

template<class T>
long long pack_xw(XwTask<T>* tasks, int count, int* maxn) {
  long long tiles = 0;
  int mn = 0;
  for (int i = 0; i < count; ++i) {
    if (tasks[i].N < 1 || tasks[i].N > 64) throw std::invalid_argument("xw supports 1..64 right-hand columns");
    if (tasks[i].splits < 1) throw std::invalid_argument("xw needs at least one split");
    tasks[i].tile0 = tiles;
    tiles += xw_row_tiles(tasks[i].m) * tasks[i].splits;
    mn = std::max(mn, tasks[i].N);
  }
  if (maxn) *maxn = mn;
  return tiles;
}

template<class T>
long long pack_xtw(XtwTask<T>* tasks, int count, int* maxn) {
  int mn = 0;
  for (int i = 0; i < count; ++i) {
    if (tasks[i].N < 1 || tasks[i].N > 64) throw std::invalid_argument("xtw supports 1..64 right-hand columns");
    if (tasks[i].splits < 1) throw std::invalid_argument("xtw needs at least one split");
    mn = std::max(mn, tasks[i].N);
  }
  long long tiles = 0;
  for (int i = 0; i < count; ++i) {
    tasks[i].tile0 = tiles;
    tiles += xtw_tiles(tasks[i].k, mn) * tasks[i].splits;
  }
  if (maxn) *maxn = mn;
  return tiles;
}

template<class T>
void launch_xw(const XwTask<T>* device_tasks, int count, long long tiles, int maxn,
               hipStream_t stream) {
  if (count <= 0) return;
  dispatch_xw<T>(device_tasks, count, XwTask<T>{}, tiles, maxn, stream);
}

template<class T>
void launch_xtw(const XtwTask<T>* device_tasks, int count, long long tiles, int maxn,
                hipStream_t stream) {
  if (count <= 0) return;
  dispatch_xtw<T>(device_tasks, count, XtwTask<T>{}, tiles, maxn, stream);
}

template<class T>
void launch_xw_single(XwTask<T> task, hipStream_t stream) {
  int maxn = 0;
  const long long tiles = pack_xw(&task, 1, &maxn);
  dispatch_xw<T>(nullptr, 1, task, tiles, maxn, stream);
}

template<class T>
void launch_xtw_single(XtwTask<T> task, hipStream_t stream) {
  int maxn = 0;
  const long long tiles = pack_xtw(&task, 1, &maxn);
  dispatch_xtw<T>(nullptr, 1, task, tiles, maxn, stream);
}

void launch_sum_slabs(const double* part, int splits, long long stride,
                      long long count, double* out, hipStream_t stream) {
  if (count <= 0) return;
  long long blocks = (count + kWG - 1) / kWG;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_sum_slabs, dim3(unsigned(blocks)), dim3(kWG), 0, stream,
                     part, splits, stride, count, out);
  FPLS_LAUNCHED("k_sum_slabs");
}

#define FPLS_STREAM_INSTANCES(T)                                                        \
  template long long pack_xw<T>(XwTask<T>*, int, int*);                                  \
  template long long pack_xtw<T>(XtwTask<T>*, int, int*);                                \
  template void launch_xw<T>(const XwTask<T>*, int, long long, int, hipStream_t);       \
  template void launch_xtw<T>(const XtwTask<T>*, int, long long, int, hipStream_t);      \
  template void launch_xw_single<T>(XwTask<T>, hipStream_t);                             \
  template void launch_xtw_single<T>(XtwTask<T>, hipStream_t);
FPLS_STREAM_INSTANCES(float)
FPLS_STREAM_INSTANCES(double)
#undef FPLS_STREAM_INSTANCES

}  // namespace fpls
