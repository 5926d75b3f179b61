// Streaming products over tall column-major matrices (see stream.hpp).
//
// Roofline: every kernel reads A once per launch (m*k*sizeof(T) bytes) and is
// meant to run at HBM bandwidth; W operands come from L2 (directly or through
// LDS). Three arithmetic paths:
//   * VALU, FP64 FMA on values widened from T (right-hand sides below 8);
//   * FP32 A on the matrix cores in bf16x6 split precision: x = xh + xm + xl
//     and w = wh + wm + wl (three bf16 terms each, residual 2^-24), products
//     hh + hm + mh + hl + mm + lh (dropped terms of order 2^-24), each
//     32-long reduction step in one FP32 matrix-core chain folded into FP64
//     accumulators, FP64 partial slabs;
//   * FP64 A on the FP64 matrix cores.
// Kernel arguments are read once into registers (task lists through the
// constant address space) and A is read through global-address-space
// pointers, so no generic (flat) access stalls the load pipeline.
#include "stream.hpp"

#include <algorithm>
#include <type_traits>

namespace fpls {

namespace {


template<class T>
__device__ __forceinline__ const T FPLS_GLOBAL* gp(const T* p) { return (const T FPLS_GLOBAL*)p; }
template<class T>
__device__ __forceinline__ T FPLS_GLOBAL* gpw(T* p) { return (T FPLS_GLOBAL*)p; }

// Four consecutive elements widened to FP64, read with 16-byte loads that need
// only element alignment (gfx950 global loads accept dword-aligned 16-byte
// accesses), so a column of any length is read with vector loads.
typedef float float4_ea __attribute__((ext_vector_type(4), aligned(4)));
typedef double double2_ea __attribute__((ext_vector_type(2), aligned(8)));
__device__ __forceinline__ void load4_wide(const float FPLS_GLOBAL* p, double* out) {
  const float4_ea a = *(const float4_ea FPLS_GLOBAL*)p;
  out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = a.w;
}
__device__ __forceinline__ void load4_wide(const double FPLS_GLOBAL* p, double* out) {
  const double2_ea a = *(const double2_ea FPLS_GLOBAL*)p;
  const double2_ea b = *(const double2_ea FPLS_GLOBAL*)(p + 2);
  out[0] = a.x; out[1] = a.y; out[2] = b.x; out[3] = b.y;
}
// The same four elements kept in their own type (widened where they are used).
__device__ __forceinline__ void load4_raw(const float FPLS_GLOBAL* p, float* out) {
  const float4_ea a = *(const float4_ea FPLS_GLOBAL*)p;
  out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = a.w;
}
__device__ __forceinline__ void load4_raw(const double FPLS_GLOBAL* p, double* out) {
  load4_wide(p, out);
}

// The task of this work-group: the single by-value task, or a binary search
// of the flattened tile space of a device task list.
template<class Task>
__device__ __forceinline__ Task load_task(const Task* tasks, int count, const Task& single) {
  if (tasks == nullptr) return single;
  const Task FPLS_CONSTANT* ct = (const Task FPLS_CONSTANT*)tasks;
  int lo = 0, hi = count - 1;
  const long long b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ct[mid].tile0 <= b) lo = mid; else hi = mid - 1;
  }
  return ct[lo];
}

// Task `index` of a device task list, or the single by-value task.
template<class Task>
__device__ __forceinline__ Task task_at(const Task* tasks, unsigned index, const Task& single) {
  if (tasks == nullptr) return single;
  return ((const Task FPLS_CONSTANT*)tasks)[index];
}

// Rows of one split of an xtw task (a multiple of 64) and columns of one split
// of an xw task (a multiple of 32).
__device__ __forceinline__ long long split_extent(long long extent, int splits, int quantum) {
  const long long per = (extent + splits - 1) / splits;
  return (per + quantum - 1) / quantum * quantum;
}

// Tiles are handed to the 8 XCDs round-robin (tile b runs on XCD b % 8).
// Renumber them so that each XCD walks a contiguous range of the task's tile
// space: work-groups sharing an L2 then share row splits (and W tiles).
__device__ __forceinline__ long long xcd_local(long long local, long long total) {
  constexpr long long kXcds = 8;
  const long long q = total / kXcds, r = total % kXcds;
  const long long xcd = local % kXcds, pos = local / kXcds;
  return xcd < r ? xcd * (q + 1) + pos : r * (q + 1) + (xcd - r) * q + pos;
}

// ------------------------------------------------------------------- xw
// One thread per row of A; the k range of the split is walked in chunks of
// 32 columns whose W rows (and centres) are staged in LDS and broadcast.
template<class T, int NB>
__global__ __launch_bounds__(kWG) void k_xw(const XwTask<T>* tasks, int count, XwTask<T> single) {
  const XwTask<T> t = load_task(tasks, count, single);
  long long local = (long long)blockIdx.x - t.tile0;
  const long long rtiles = xw_row_tiles(t.m);
  if (local >= rtiles * t.splits) return;
  // neighbouring row tiles share the 128-byte line at their boundary
  // (columns of any length start unaligned): on one XCD, one L2 fetches it
  local = xcd_local(local, rtiles * t.splits);
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
  const T FPLS_GLOBAL* a = gp(t.A) + (valid ? row : 0);
  const double FPLS_GLOBAL* W = gp(t.W);

  for (long long jc = j0; jc < j1; jc += 32) {
    const int cnt = int(j1 - jc < 32 ? j1 - jc : 32);
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * NB; e += kWG) {
      const int jj = e & 31, c = e >> 5;
      ws[jj][c] = (jj < cnt && c < t.N) ? W[(jc + jj) + (long long)c * t.ldw] : 0.0;
    }
    if (threadIdx.x < 32) {
      cs[threadIdx.x] = (t.center != nullptr && int(threadIdx.x) < cnt) ? gp(t.center)[jc + threadIdx.x] : 0.0;
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
    double FPLS_GLOBAL* o = gpw(t.out) + (long long)split * t.slab + row;
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
  const XtwTask<T> t = load_task(tasks, count, single);
  const long long local = (long long)blockIdx.x - t.tile0;
  const long long ctiles = xtw_col_tiles(t.k);
  if (local >= ctiles * t.splits) return;
  const int split = int(local / ctiles);
  const long long col0 = (local % ctiles) * 64;
  const long long per = (t.m + t.splits - 1) / t.splits;
  const long long r0 = split * per;
  const long long r1 = r0 + per < t.m ? r0 + per : t.m;
  const T FPLS_GLOBAL* A = gp(t.A);
  const double FPLS_GLOBAL* W = gp(t.W);

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
      if (t.mode == kCentered && t.center != nullptr) c = gp(t.center)[col];
      else if (t.mode == kShiftStats) c = double(A[col * t.m]);
    }
    centre[e] = c; s1[e] = 0.0; s2[e] = 0.0;
  }
  const int g = tid % GROUPS, h = tid / GROUPS;
  double acc[MC][MN];
#pragma unroll
  for (int a = 0; a < MC; ++a)
#pragma unroll
    for (int b = 0; b < MN; ++b) acc[a][b] = 0.0;

  // the A and W values of the next 32-row tile are loaded into registers
  // while the current tile is multiplied out of LDS
  constexpr int WPT = (32 * NB + kWG - 1) / kWG;
  T xr[8];
  double wr[WPT];
  auto fetch = [&](long long i0) __attribute__((always_inline)) {
    const long long row = i0 + lrow;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long long col = col0 + lcol + 8 * e;
      xr[e] = (row < r1 && col < t.k) ? A[row + col * t.m] : T(0);
    }
#pragma unroll
    for (int u = 0; u < WPT; ++u) {
      const int e = tid + u * kWG;
      const int r = e & 31, c = e >> 5;
      const long long w = i0 + r;
      wr[u] = (e < 32 * NB && w < r1 && c < t.N) ? W[w + (long long)c * t.ldw] : 0.0;
    }
  };
  if (r0 < r1) fetch(r0);
  for (long long i0 = r0; i0 < r1; i0 += 32) {
    const long long row = i0 + lrow;
    const bool rv = row < r1;
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long long col = col0 + lcol + 8 * e;
      double v = 0.0;
      if (rv && col < t.k) {
        v = double(xr[e]) - centre[e];
        s1[e] += v;
        s2[e] = fma(v, v, s2[e]);
      }
      xs[lrow][lcol + 8 * e] = v;
    }
#pragma unroll
    for (int u = 0; u < WPT; ++u) {
      const int e = tid + u * kWG;
      if (e < 32 * NB) wsh[e & 31][e >> 5] = wr[u];
    }
    if (i0 + 32 < r1) fetch(i0 + 32);
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
  double FPLS_GLOBAL* o = gpw(t.out) + (long long)split * t.slab;
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
        double FPLS_GLOBAL* st = gpw(t.stats) + (long long)split * 2 * t.k;
        st[col] = a1;
        st[t.k + col] = a2;
      }
    }
  }
}

// ------------------------------------------------- one right-hand column
// xtw with N = 1 (a singleton fit's loading pass, or the statistics and
// cross-product of a one-response fit). The LDS-staged k_xtw above widens
// every element into LDS and reads it back once per output; with a single
// right-hand column that LDS traffic, not HBM, bounds it. Here a work-group
// of eight waves owns 64 columns (eight per wave) and one row split: a lane
// owns four consecutive rows, reads them from each of its wave's eight
// columns with one 16-byte load per column (all eight in flight together)
// and the four right-hand values once, and keeps eight running dot products
// (and, with kStats, the shifted sums) that the wave reduces at the end.
// Columns past k read column k - 1 and are not written, so the loads stay
// unconditional.
template<class T, bool kStats>
__global__ __launch_bounds__(512) void k_xtw1(const XtwTask<T>* tasks, int count,
                                              XtwTask<T> single) {
  constexpr int kCols = 8;  // columns per wave
  const XtwTask<T> t = load_task(tasks, count, single);
  long long local = (long long)blockIdx.x - t.tile0;
  const long long ctiles = xtw_col_tiles(t.k);
  if (local >= ctiles * t.splits) return;
  local = xcd_local(local, ctiles * t.splits);
  const int split = int(local / ctiles);
  const long long col0 = (local % ctiles) * 64;
  const long long per = (t.m + t.splits - 1) / t.splits;
  const long long r0 = split * per;
  const long long r1 = r0 + per < t.m ? r0 + per : t.m;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long jw = col0 + wave * kCols;  // this wave's first column
  if (jw >= t.k) return;                     // no barrier below
  const T FPLS_GLOBAL* A = gp(t.A);
  const double FPLS_GLOBAL* W = gp(t.W);

  const T FPLS_GLOBAL* col[kCols];
  double cv[kCols], acc[kCols], s1[kCols], s2[kCols];
#pragma unroll
  for (int u = 0; u < kCols; ++u) {
    const long long j = jw + u < t.k ? jw + u : t.k - 1;
    col[u] = A + j * t.m;
    double c = 0.0;
    if (!kStats && t.mode == kCentered && t.center != nullptr) c = gp(t.center)[j];
    if (kStats) c = double(col[u][0]);
    cv[u] = c; acc[u] = 0.0; s1[u] = 0.0; s2[u] = 0.0;
  }
  for (long long i0 = r0; i0 < r1; i0 += 256) {
    const long long row = i0 + 4 * lane;
    if (row + 3 < r1) {
      double w[4];
      load4_wide(W + row, w);
      T x[kCols][4];
#pragma unroll
      for (int u = 0; u < kCols; ++u) load4_raw(col[u] + row, x[u]);
#pragma unroll
      for (int u = 0; u < kCols; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double d = double(x[u][e]) - cv[u];
          acc[u] = fma(d, w[e], acc[u]);
          if (kStats) {
            s1[u] += d;
            s2[u] = fma(d, d, s2[u]);
          }
        }
    } else if (row < r1) {
      // the split's last rows: rows past r1 contribute nothing
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (row + e >= r1) break;
        const double w = W[row + e];
#pragma unroll
        for (int u = 0; u < kCols; ++u) {
          const double d = double(col[u][row + e]) - cv[u];
          acc[u] = fma(d, w, acc[u]);
          if (kStats) {
            s1[u] += d;
            s2[u] = fma(d, d, s2[u]);
          }
        }
      }
    }
  }
  double FPLS_GLOBAL* o = gpw(t.out) + (long long)split * t.slab;
#pragma unroll
  for (int u = 0; u < kCols; ++u) {
    double a = acc[u], b1 = s1[u], b2 = s2[u];
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      if (kStats) {
        b1 += __shfl_xor(b1, off, 64);
        b2 += __shfl_xor(b2, off, 64);
      }
    }
    const long long j = jw + u;
    if (lane == 0 && j < t.k) {
      o[j] = a;
      if (kStats && t.stats != nullptr) {
        double FPLS_GLOBAL* st = gpw(t.stats) + (long long)split * 2 * t.k;
        st[j] = b1;
        st[t.k + j] = b2;
      }
    }
  }
}

// ------------------------------------------------------------ MFMA paths
// Fragment conventions of the 16x16 matrix-core instructions (gfx950): for
// D = A·B with A 16 x K and B K x 16, lane l supplies row l % 16 of A and
// column l % 16 of B over the K slice owned by lane group l / 16 (one value
// per lane for the x4 FP64 shape, eight bf16 for the x32 shape); the products
// only pair A and B values held in the same lane slot, so any assignment of
// rows of the reduction to slots is valid as long as A and B agree. Results
// (measured on gfx950): item r of lane l holds D[4(l/16) + r][l % 16] for FP32
// accumulators and D[l/16 + 4r][l % 16] for FP64 accumulators.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  const bf16x2 v = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(unsigned, v);
}

// Three-term bf16 split, each term rounded to nearest even: x = hi + mid + lo
// + r with |r| <= 2^-24 |x| (hi carries the leading 8 significant bits, mid
// the next 8, lo the next 8; both residuals are exact in FP32). Two values per
// 32-bit word, the first in the low half.
__device__ __forceinline__ void split_pair(float a, float b, unsigned& hi, unsigned& mid,
                                           unsigned& lo) {
  hi = pack_bf16(a, b);
  const float ra = a - __builtin_bit_cast(float, hi << 16);
  const float rb = b - __builtin_bit_cast(float, hi & 0xffff0000u);
  mid = pack_bf16(ra, rb);
  const float sa = ra - __builtin_bit_cast(float, mid << 16);
  const float sb = rb - __builtin_bit_cast(float, mid & 0xffff0000u);
  lo = pack_bf16(sa, sb);
}

struct Frag3 {
  u32x4 h, m, l;
};

__device__ __forceinline__ Frag3 split8(const float (&v)[8]) {
  unsigned h[4], m[4], l[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) split_pair(v[2 * i], v[2 * i + 1], h[i], m[i], l[i]);
  return Frag3{u32x4{h[0], h[1], h[2], h[3]}, u32x4{m[0], m[1], m[2], m[3]},
               u32x4{l[0], l[1], l[2], l[3]}};
}

__device__ __forceinline__ f32x4 mma_bf16(const u32x4& a, const u32x4& b, f32x4 acc) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
}

// acc += a·b to FP32 accuracy from six bf16 products (bf16x6): the terms of
// order 2^-16 (hi·lo, mid·mid, lo·hi) first, then those of order 2^-8
// (hi·mid, mid·hi), then hi·hi; the dropped terms are of order 2^-24.
__device__ __forceinline__ f32x4 mma6(const Frag3& a, const u32x4& bh, const u32x4& bm,
                                      const u32x4& bl, f32x4 acc) {
  acc = mma_bf16(a.h, bl, acc);
  acc = mma_bf16(a.m, bm, acc);
  acc = mma_bf16(a.l, bh, acc);
  acc = mma_bf16(a.h, bm, acc);
  acc = mma_bf16(a.m, bh, acc);
  return mma_bf16(a.h, bh, acc);
}

// FP64 accumulation of one short FP32 matrix-core chain.
__device__ __forceinline__ void fold(f64x4& acc, const f32x4& c) {
  acc[0] += double(c[0]); acc[1] += double(c[1]); acc[2] += double(c[2]); acc[3] += double(c[3]);
}

__device__ __forceinline__ f64x4 mma64(double a, double b, f64x4 acc) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}

// ---------------------------------------------- bf16x6 operand splitting
// xw operand: W (k x N) -> hi/mid/lo [3][NB][k64] in W's own column order,
// plus the centre in FP32 (zero past k and past N). Grid: (blocks, tasks).
template<int NB>
__global__ __launch_bounds__(kWG) void k_xw_split(const XwTask<float>* tasks, XwTask<float> single) {
  const XwTask<float> t = task_at(tasks, blockIdx.y, single);
  const long long k32 = (t.k + 63) / 64 * 64;  // padded to 64 columns
  unsigned short FPLS_GLOBAL* hi = (unsigned short FPLS_GLOBAL*)t.frag;
  unsigned short FPLS_GLOBAL* mid = hi + (long long)NB * k32;
  unsigned short FPLS_GLOBAL* lo = mid + (long long)NB * k32;
  float FPLS_GLOBAL* cf = (float FPLS_GLOBAL*)(hi + 3LL * NB * k32);
  const double FPLS_GLOBAL* W = gp(t.W);
  const long long stride = (long long)gridDim.x * kWG;
  // two adjacent elements per thread
  for (long long e = ((long long)blockIdx.x * kWG + threadIdx.x) * 2; e < (long long)NB * k32; e += 2 * stride) {
    const long long n = e / k32, j = e % k32;  // k32 is even: j, j + 1 share n
    const float a = (n < t.N && j < t.k) ? float(W[j + n * t.ldw]) : 0.f;
    const float b = (n < t.N && j + 1 < t.k) ? float(W[j + 1 + n * t.ldw]) : 0.f;
    unsigned h, md, l;
    split_pair(a, b, h, md, l);
    *(unsigned FPLS_GLOBAL*)(hi + e) = h;
    *(unsigned FPLS_GLOBAL*)(mid + e) = md;
    *(unsigned FPLS_GLOBAL*)(lo + e) = l;
  }
  for (long long j = (long long)blockIdx.x * kWG + threadIdx.x; j < k32; j += stride)
    cf[j] = (t.center != nullptr && j < t.k) ? float(gp(t.center)[j]) : 0.f;
}

// xtw operand: W (m x N) -> hi/mid/lo [3][NB][m64]; within each 64-row tile,
// position 32h + 8g + j holds row 16(2h + j/4) + 4g + j%4, so that lane group
// g of reduction step h reads its eight values as one 16-byte chunk.
template<int NB>
__global__ __launch_bounds__(kWG) void k_xtw_split(const XtwTask<float>* tasks, XtwTask<float> single) {
  const XtwTask<float> t = task_at(tasks, blockIdx.y, single);
  const long long m64 = (t.m + 63) / 64 * 64;
  unsigned short FPLS_GLOBAL* hi = (unsigned short FPLS_GLOBAL*)t.frag;
  unsigned short FPLS_GLOBAL* mid = hi + (long long)NB * m64;
  unsigned short FPLS_GLOBAL* lo = mid + (long long)NB * m64;
  const double FPLS_GLOBAL* W = gp(t.W);
  const long long stride = (long long)gridDim.x * kWG;
  for (long long e = ((long long)blockIdx.x * kWG + threadIdx.x) * 2; e < (long long)NB * m64; e += 2 * stride) {
    const long long n = e / m64, p = e % m64;
    const long long tile = p >> 6;
    float v[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int pp = int((p + u) & 63);
      const int h = pp >> 5, g = (pp >> 3) & 3, j = pp & 7;
      const long long row = tile * 64 + 16 * (2 * h + (j >> 2)) + 4 * g + (j & 3);
      v[u] = (n < t.N && row < t.m) ? float(W[row + n * t.ldw]) : 0.f;
    }
    unsigned h, md, l;
    split_pair(v[0], v[1], h, md, l);
    *(unsigned FPLS_GLOBAL*)(hi + e) = h;
    *(unsigned FPLS_GLOBAL*)(mid + e) = md;
    *(unsigned FPLS_GLOBAL*)(lo + e) = l;
  }
}

// --------------------------------------------------- xtw, bf16x6 (FP32 A)
// A work-group of 512 threads owns 128 columns of A and one row split,
// walked in 64-row steps, with the work split by role. Waves 0-3 produce:
// wave p loads columns 32p..32p+31 (two 16-column M-blocks; lane (c, g) =
// (lane % 16, lane / 16) holds rows 16v + 4g + e, v, e < 4, of column c of
// each block, as float4), centres them, gathers the column statistics,
// splits every 32-row sub-step h (rows {16(2h + j/4) + 4g + j%4}) into bf16
// hi/mid/lo fragments and stores them in LDS next to the step's W tile.
// Waves 4-7 consume: wave 4 + c runs the matrix-core chains of producer c's
// M-blocks against every 16-column block of W. Waves w and w + 4 share a
// SIMD, so one step's vector work overlaps the previous step's matrix
// products. LDS holds two steps; producers prefetch the A rows and W tile of
// the step after the one they prepare. Each (M-block, W block) runs one FP32
// chain of twelve products per step (order-2^-16 terms of both sub-steps,
// then order-2^-8, then hi·hi) that is folded into FP64 accumulators.
constexpr int kXtwCols = 128;
constexpr int kBf6Threads = 512;

template<int NB>
__global__ __launch_bounds__(kBf6Threads) void k_xtw_bf6(const XtwTask<float>* tasks, int count,
                                                         XtwTask<float> single) {
  constexpr int NBLK = NB / 16;
  constexpr int ROWS = 64;                        // rows per step
  constexpr int CPR = ROWS / 8;                   // 16-byte W chunks per fragment row
  constexpr int WCH = 3 * NB * CPR;               // W chunks of one step
  constexpr int PT = kBf6Threads / 2;             // producer threads
  constexpr int WPT = (WCH + PT - 1) / PT;        // W chunks per producer thread
  static_assert(CPR == 8, "one 64-row tile of W per step");
  const XtwTask<float> t = load_task(tasks, count, single);
  const long long ctiles = (t.k + kXtwCols - 1) / kXtwCols;
  const long long total = ctiles * t.splits;
  long long local = (long long)blockIdx.x - t.tile0;
  if (local >= total) return;
  local = xcd_local(local, total);
  const int split = int(local / ctiles);
  const long long col0 = (local % ctiles) * kXtwCols;
  const long long per = split_extent(t.m, t.splits, 64);
  const long long r0 = split * per;
  const long long r1 = r0 + per < t.m ? r0 + per : t.m;
  const long long m64 = (t.m + 63) / 64 * 64;

  __shared__ u32x4 af[2][8][2][3][64];            // [buffer][M-block][sub-step][plane][lane]
  __shared__ u32x4 wl[2][3 * NB][CPR + 1];        // [buffer][plane NB + W column][chunk + pad]
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int role = wave & 3;
  const int m16 = lane & 15, g = lane >> 4;
  const bool vec = (t.m & 3) == 0 && (((unsigned long long)t.A) & 15) == 0;
  const bool stats = t.mode == kShiftStats;
  const long long nfull = vec ? (r1 - r0) / ROWS : 0;
  const long long cw = col0 + 32 * role;          // first column of the wave's M-blocks

  if (wave < 4) {
    // ------------------------------------------------------------ producer
    // Fast loads: a wave-uniform base (column bc) plus 32-bit lane offsets;
    // columns past k re-read column bc (their outputs are never written).
    const long long bc = cw < t.k ? cw : t.k - 1;
    const char FPLS_GLOBAL* xbase = (const char FPLS_GLOBAL*)(gp(t.A) + bc * t.m);
    unsigned xoff[2];
    bool cvalid[2];
    float centre[2];
    const float FPLS_GLOBAL* xcol[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      const long long col = cw + 16 * mb + m16;
      cvalid[mb] = col < t.k;
      xcol[mb] = gp(t.A) + (cvalid[mb] ? col : 0) * t.m;
      xoff[mb] = unsigned(((cvalid[mb] ? col - bc : 0) * t.m + 4 * g) * 4);
      centre[mb] = 0.f;
      if (cvalid[mb]) {
        if (t.mode == kCentered && t.center != nullptr) centre[mb] = float(gp(t.center)[col]);
        else if (stats) centre[mb] = xcol[mb][0];
      }
    }
    auto load_x = [&](float (&x)[2][16], long long i0) {
#pragma unroll
      for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float4 f = *(const float4 FPLS_GLOBAL*)(xbase + i0 * 4 + 64 * v + xoff[mb]);
          x[mb][4 * v] = f.x; x[mb][4 * v + 1] = f.y; x[mb][4 * v + 2] = f.z; x[mb][4 * v + 3] = f.w;
        }
    };
    // rows past r1 and columns past k load the centre: they contribute zero
    auto load_x_edge = [&](float (&x)[2][16], long long i0) {
#pragma unroll
      for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const long long row = i0 + 16 * v + 4 * g + e;
            x[mb][4 * v + e] = (cvalid[mb] && row < r1) ? xcol[mb][row] : centre[mb];
          }
    };
    double s1[2] = {0.0, 0.0}, s2[2] = {0.0, 0.0};
    auto prepare = [&](const float (&x)[2][16], int b) {
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        float p1 = 0.f, p2 = 0.f;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          float v[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) v[j] = x[mb][8 * h + j] - centre[mb];
          if (stats) {
#pragma unroll
            for (int j = 0; j < 8; ++j) { p1 += v[j]; p2 = fmaf(v[j], v[j], p2); }
          }
          const Frag3 a = split8(v);
          af[b][2 * role + mb][h][0][lane] = a.h;
          af[b][2 * role + mb][h][1][lane] = a.m;
          af[b][2 * role + mb][h][2][lane] = a.l;
        }
        if (stats) { s1[mb] += double(p1); s2[mb] += double(p2); }
      }
    };

    float xa[2][16], xb[2][16];
    if (nfull > 0) {
      auto at = [&](long long step) { return r0 + ROWS * (step < nfull ? step : nfull - 1); };
      load_x(xa, at(0));
      load_x(xb, at(1));
      prepare(xa, 0);
      __syncthreads();
      for (long long s = 0; s < nfull; s += 2) {
        load_x(xa, at(s + 2));
        if (s + 1 < nfull) prepare(xb, 1);
        __syncthreads();
        if (s + 1 >= nfull) break;
        load_x(xb, at(s + 3));
        if (s + 2 < nfull) prepare(xa, 0);
        __syncthreads();
      }
    }
    // the ragged last step (or every step of unaligned columns)
    for (long long i0 = r0 + ROWS * nfull; i0 < r1; i0 += ROWS) {
      load_x_edge(xa, i0);
      prepare(xa, 0);
      __syncthreads();
      __syncthreads();
    }
    if (stats && t.stats != nullptr) {
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        double a1 = s1[mb], a2 = s2[mb];
        a1 += __shfl_xor(a1, 16, 64);
        a1 += __shfl_xor(a1, 32, 64);
        a2 += __shfl_xor(a2, 16, 64);
        a2 += __shfl_xor(a2, 32, 64);
        const long long col = cw + 16 * mb + m16;
        if (g == 0 && cvalid[mb]) {
          double FPLS_GLOBAL* st = gpw(t.stats) + (long long)split * 2 * t.k;
          st[col] = a1;
          st[t.k + col] = a2;
        }
      }
    }
    return;
  }

  // -------------------------------------------------------------- consumer
  // The consumers also stage W: chunk c = (threadIdx.x - PT) + s * PT of a
  // step is fragment row c / 8, 16 bytes at 8 (c % 8), i.e. row 32 s +
  // (threadIdx.x - PT) / 8 of the [3 NB][m64] operand.
  const int ct = threadIdx.x - PT;
  const char FPLS_GLOBAL* frag = (const char FPLS_GLOBAL*)t.frag;
  const unsigned woff = unsigned(((ct >> 3) * m64 + 8 * (ct & 7)) * 2);
  u32x4 wr[WPT];
  auto load_w = [&](long long i0) {
#pragma unroll
    for (int s = 0; s < WPT; ++s) {
      if (WCH % PT == 0 || ct + s * PT < WCH) {
        const char FPLS_GLOBAL* b = frag + ((long long)s * (PT / CPR) * m64 + i0) * 2;
        wr[s] = *(const u32x4 FPLS_GLOBAL*)(b + woff);
      }
    }
  };
  auto store_w = [&](int b) {
#pragma unroll
    for (int s = 0; s < WPT; ++s) {
      const int c = ct + s * PT;
      if (WCH % PT == 0 || c < WCH) wl[b][c / CPR][c % CPR] = wr[s];
    }
  };
  f64x4 acc[2][NBLK];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) acc[mb][nb] = f64x4{0.0, 0.0, 0.0, 0.0};
  auto compute = [&](int b) {
    u32x4 a[2][2][3];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) a[mb][h][pl] = af[b][2 * role + mb][h][pl][lane];
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) {
      u32x4 w[2][3];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) w[h][pl] = wl[b][pl * NB + nb * 16 + m16][4 * h + g];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          c = mma_bf16(a[mb][h][0], w[h][2], c);
          c = mma_bf16(a[mb][h][1], w[h][1], c);
          c = mma_bf16(a[mb][h][2], w[h][0], c);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          c = mma_bf16(a[mb][h][0], w[h][1], c);
          c = mma_bf16(a[mb][h][1], w[h][0], c);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) c = mma_bf16(a[mb][h][0], w[h][0], c);
        fold(acc[mb][nb], c);
      }
    }
  };
  // phase s: W of step s + 1 goes from registers into LDS buffer (s + 1) % 2,
  // the W loads of step s + 2 are issued, then step s is computed
  if (nfull > 0) {
    auto at = [&](long long step) { return r0 + ROWS * (step < nfull ? step : nfull - 1); };
    load_w(at(0));
    store_w(0);
    load_w(at(1));
    __syncthreads();
    for (long long s = 0; s < nfull; s += 2) {
      store_w(1);
      load_w(at(s + 2));
      compute(0);
      __syncthreads();
      if (s + 1 >= nfull) break;
      store_w(0);
      load_w(at(s + 3));
      compute(1);
      __syncthreads();
    }
  }
  // the W operand is padded to m64 rows, so its tile is always whole
  for (long long i0 = r0 + ROWS * nfull; i0 < r1; i0 += ROWS) {
    load_w(i0);
    store_w(0);
    __syncthreads();
    compute(0);
    __syncthreads();
  }
  double FPLS_GLOBAL* o = gpw(t.out) + (long long)split * t.slab;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) {
      const int n = nb * 16 + m16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long oc = cw + 16 * mb + 4 * g + r;
        if (oc < t.k && n < t.N) o[oc + (long long)n * t.ldo] = acc[mb][nb][r];
      }
    }
}

// ---------------------------------------------------- xw, bf16x6 (FP32 A)
// A work-group of 512 threads owns 256 rows (32 per wave) and one column
// split, walked in 64-column steps of two 32-column reduction sub-steps.
// Lane (r, g) = (lane % 16, lane / 16) loads rows 2r, 2r + 1 of columns
// 32u + 8g..32u + 8g + 7 of sub-step u (eight float2 per sub-step, a full
// 128-byte line per column and instruction); MFMA call e uses row 2r + e of
// those columns as the A fragment of output rows {2r' + e}. The W fragments
// (already split, W's own column order) and the FP32 centre of a step are
// staged in LDS (double-buffered, shared by the eight waves); the next
// step's loads are issued before the current step's arithmetic. The products
// of two consecutive steps (four 32-column sub-steps) run in one FP32 chain
// per output block, folded into FP64 accumulators.
template<int NB>
__global__ __launch_bounds__(kBf6Threads) void k_xw_bf6(const XwTask<float>* tasks, int count,
                                                        XwTask<float> single) {
  constexpr int NBLK = NB / 16;
  constexpr int COLS = 64;                                     // columns per step
  constexpr int CPR = COLS / 8;                                // 16-byte chunks per W row
  constexpr int WCH = 3 * NB * CPR;                            // W chunks of one step
  constexpr int WPT = (WCH + kBf6Threads - 1) / kBf6Threads;
  const XwTask<float> t = load_task(tasks, count, single);
  // (no XCD renumbering here: measured slower on the 1e5 x 2e4 pass, whose
  // 400000-byte columns start aligned)
  const long long local = (long long)blockIdx.x - t.tile0;
  const long long rtiles = xw_row_tiles(t.m);
  if (local >= rtiles * t.splits) return;
  const int split = int(local / rtiles);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int m16 = lane & 15, g = lane >> 4;
  const long long i0w = (local % rtiles) * 256 + wave * 32;
  const long long rowb = i0w + 2 * m16;
  const long long per = split_extent(t.k, t.splits, COLS);
  const long long j0 = split * per;
  const long long j1 = j0 + per < t.k ? j0 + per : t.k;
  const long long k64 = (t.k + 63) / 64 * 64;
  const bool vec = (t.m & 1) == 0 && (((unsigned long long)t.A) & 7) == 0;
  const bool rows_full = i0w + 32 <= t.m;

  __shared__ u32x4 wl[2][3 * NB][CPR + 1];
  __shared__ float cs[2][COLS];

  const float FPLS_GLOBAL* xr = gp(t.A) + (rowb < t.m ? rowb : 0);
  // Fast loads: wave-uniform bases plus one 32-bit lane offset.
  const long long colbytes = t.m * 4;
  const char FPLS_GLOBAL* xbase = (const char FPLS_GLOBAL*)(gp(t.A) + i0w);
  const unsigned xoff = unsigned(8 * g * colbytes + 8 * m16);
  const char FPLS_GLOBAL* fbase = (const char FPLS_GLOBAL*)t.frag;
  const float FPLS_GLOBAL* cf = (const float FPLS_GLOBAL*)(fbase + 6LL * NB * k64);
  // W chunk c = threadIdx.x + s * 512: fragment row c / 8, 16 bytes at 8 (c % 8)
  const unsigned woff = unsigned(((threadIdx.x >> 3) * k64 + 8 * (threadIdx.x & 7)) * 2);

  auto load_w = [&](u32x4 (&wr)[WPT], float& cr, long long jc) {
#pragma unroll
    for (int s = 0; s < WPT; ++s) {
      if (WCH % kBf6Threads == 0 || int(threadIdx.x) + s * kBf6Threads < WCH) {
        const char FPLS_GLOBAL* b =
            fbase + ((long long)s * (kBf6Threads / CPR) * k64 + jc) * 2;
        wr[s] = *(const u32x4 FPLS_GLOBAL*)(b + woff);
      }
    }
    if (threadIdx.x < COLS) cr = cf[jc + threadIdx.x];
  };
  auto store_w = [&](const u32x4 (&wr)[WPT], float cr, int b) {
#pragma unroll
    for (int s = 0; s < WPT; ++s) {
      const int c = threadIdx.x + s * kBf6Threads;
      if (WCH % kBf6Threads == 0 || c < WCH) wl[b][c / CPR][c % CPR] = wr[s];
    }
    if (threadIdx.x < COLS) cs[b][threadIdx.x] = cr;
  };
  // a full 64-column step of 8-byte-aligned row pairs
  auto load_x = [&](float (&x)[2][8][2], long long jc) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const char FPLS_GLOBAL* b = xbase + (jc + 32 * u) * colbytes;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float2 f = *(const float2 FPLS_GLOBAL*)(b + j * colbytes + xoff);
        x[u][j][0] = f.x; x[u][j][1] = f.y;
      }
    }
  };
  // any step: columns past j1 and rows past m read as zero
  auto load_x_edge = [&](float (&x)[2][8][2], long long jc) {
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const long long cc = jc + 32 * u + 8 * g + j;
          x[u][j][e] = (cc < j1 && rowb + e < t.m) ? xr[cc * t.m + e] : 0.f;
        }
  };

  f64x4 acc[2][NBLK];
#pragma unroll
  for (int e = 0; e < 2; ++e)
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) acc[e][nb] = f64x4{0.0, 0.0, 0.0, 0.0};

  // FP32 chains run over both sub-steps of two consecutive steps (buffers 0
  // then 1) and are folded into the FP64 accumulators after the second
  f32x4 cc[2][NBLK];
  auto compute = [&](const float (&x)[2][8][2], int b, bool start, bool finish) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      float c[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) c[j] = cs[b][32 * u + 8 * g + j];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = x[u][j][e] - c[j];
        const Frag3 a = split8(v);
#pragma unroll
        for (int nb = 0; nb < NBLK; ++nb) {
          const u32x4 bh = wl[b][nb * 16 + m16][4 * u + g];
          const u32x4 bm = wl[b][NB + nb * 16 + m16][4 * u + g];
          const u32x4 bl = wl[b][2 * NB + nb * 16 + m16][4 * u + g];
          const f32x4 c0 = (start && u == 0) ? f32x4{0.f, 0.f, 0.f, 0.f} : cc[e][nb];
          const f32x4 c = mma6(a, bh, bm, bl, c0);
          if (finish && u == 1) fold(acc[e][nb], c);
          else cc[e][nb] = c;
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  };

  // Branch-free two-stage pipeline over the steps of the split. Every wave
  // of the work-group walks the same steps (barriers); a wave whose rows are
  // ragged or unaligned uses masked loads. Steps past j1 read zero W (the
  // operand is padded to k64 columns) and masked A.
  float xa[2][8][2], xb[2][8][2];
  u32x4 wr[WPT];
  float cr = 0.f;
  const bool fast = vec && rows_full;
  const long long nsteps = (j1 - j0 + COLS - 1) / COLS;
  auto load_any = [&](float (&x)[2][8][2], long long jc) {
    if (fast && jc + COLS <= j1) load_x(x, jc); else load_x_edge(x, jc);
  };
  if (nsteps > 0) {
    load_w(wr, cr, j0);
    load_any(xa, j0);
    store_w(wr, cr, 0);
    __syncthreads();
    long long s = 0;
    for (; s + 1 < nsteps; s += 2) {
      load_w(wr, cr, j0 + COLS * (s + 1));
      load_any(xb, j0 + COLS * (s + 1));
      compute(xa, 0, true, false);
      store_w(wr, cr, 1);
      __syncthreads();
      const long long nx = j0 + COLS * (s + 2 < nsteps ? s + 2 : s + 1);
      load_w(wr, cr, nx);
      load_any(xa, nx);
      compute(xb, 1, false, true);
      store_w(wr, cr, 0);
      __syncthreads();
    }
    if (s < nsteps) compute(xa, 0, true, true);
  }

  double FPLS_GLOBAL* o = gpw(t.out) + (long long)split * t.slab;
#pragma unroll
  for (int nb = 0; nb < NBLK; ++nb) {
    const int n = nb * 16 + m16;
    if (n >= t.N) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const long long row = i0w + 8 * g + 2 * r + e;
        if (row < t.m) o[row + (long long)n * t.ldo] = acc[e][nb][r];
      }
  }
}

// ------------------------------------------------------ xtw, FP64 MFMA
// Same tiling as the bf16x6 kernel (128 columns, 64-row steps, lane rows
// 16v + 4g + e); MFMA call (v, e) pairs row 16v + 4g + e of A with the same
// row of W. The loop is software-pipelined one step deep: the A values and
// the W tile of step s + 1 are loaded into registers while step s runs its
// matrix products (A from registers, W from a double-buffered LDS tile), so
// a wave never waits for global memory between steps. The W tile is laid
// out row by row with one padding column: the staging stores (consecutive
// rows per lane) and the operand reads (16 consecutive W columns per lane
// group) are both free of bank conflicts. A is FP64, or FP32 widened to
// FP64 as it is used (every product and sum is then FP64: the accurate route
// for FP32-stored X). Each accumulator sums its products in the same order
// (steps ascending, then v, then e), whatever the pipelining.
template<int NB, class TX>
__global__ __launch_bounds__(kWG) void k_xtw_f64(const XtwTask<TX>* tasks, int count,
                                                 XtwTask<TX> single) {
  constexpr int NBLK = NB / 16;
  constexpr int WLD = NB + 1;                 // LDS row length of the W tile
  constexpr int WPT = 64 * NB / kWG;          // W values per thread and step
  const XtwTask<TX> t = load_task(tasks, count, single);
  const long long ctiles = (t.k + kXtwCols - 1) / kXtwCols;
  const long long total = ctiles * t.splits;
  long long local = (long long)blockIdx.x - t.tile0;
  if (local >= total) return;
  local = xcd_local(local, total);
  const int split = int(local / ctiles);
  const long long col0 = (local % ctiles) * kXtwCols;
  const long long per = split_extent(t.m, t.splits, 64);
  const long long r0 = split * per;
  const long long r1 = r0 + per < t.m ? r0 + per : t.m;

  __shared__ double ws[2][64][WLD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m16 = lane & 15, g = lane >> 4;
  const bool stats = t.mode == kShiftStats;
  const double FPLS_GLOBAL* W = gp(t.W);

  const TX FPLS_GLOBAL* xcol[2];
  double centre[2];
  bool cvalid[2];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
    const long long col = col0 + wave * 32 + mb * 16 + m16;
    cvalid[mb] = col < t.k;
    xcol[mb] = gp(t.A) + (cvalid[mb] ? col : 0) * t.m;
    double c = 0.0;
    if (cvalid[mb]) {
      if (t.mode == kCentered && t.center != nullptr) c = gp(t.center)[col];
      else if (stats) c = double(xcol[mb][0]);
    }
    centre[mb] = c;
  }

  f64x4 acc[2][NBLK];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) acc[mb][nb] = f64x4{0.0, 0.0, 0.0, 0.0};
  double s1[2] = {0.0, 0.0}, s2[2] = {0.0, 0.0};

  // A values of a full 64-row step (raw; widened where they are used)
  auto load_x = [&](TX (&x)[2][16], long long i0) __attribute__((always_inline)) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v) load4_raw(xcol[mb] + i0 + 16 * v + 4 * g, &x[mb][4 * v]);
  };
  // W rows of a step: value u of thread tid is row (tid + 256 u) % 64 of
  // column (tid + 256 u) / 64; rows past r1 and columns past N read zero
  auto load_w = [&](double (&w)[WPT], long long i0) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < WPT; ++u) {
      const int e = tid + u * kWG;
      const int r = e & 63, n = e >> 6;
      const long long wr = i0 + r;
      w[u] = (wr < r1 && n < t.N) ? W[wr + (long long)n * t.ldw] : 0.0;
    }
  };
  auto store_w = [&](const double (&w)[WPT], int b) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < WPT; ++u) {
      const int e = tid + u * kWG;
      ws[b][e & 63][e >> 6] = w[u];
    }
  };
  // one step: the raw A values are widened and centred one reduction row
  // (v, e) at a time, right before the products that use them
  // (`edge` steps: rows at or past `lim` are zero after centring)
  auto compute = [&](const TX (&xr)[2][16], int b, auto edge, long long lim)
      __attribute__((always_inline)) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double x[2];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
          x[mb] = double(xr[mb][4 * v + e]) - centre[mb];
          if constexpr (decltype(edge)::value) {
            if (16 * v + 4 * g + e >= lim) x[mb] = 0.0;
          }
          if (stats) { s1[mb] += x[mb]; s2[mb] = fma(x[mb], x[mb], s2[mb]); }
        }
#pragma unroll
        for (int nb = 0; nb < NBLK; ++nb) {
          const double bv = ws[b][16 * v + 4 * g + e][nb * 16 + m16];
#pragma unroll
          for (int mb = 0; mb < 2; ++mb) acc[mb][nb] = mma64(x[mb], bv, acc[mb][nb]);
        }
      }
  };

  const long long nfull = (r1 - r0) / 64;
  if (nfull > 0) {
    // xn / w hold the loads of the next step while the current one (xc)
    // multiplies; LDS buffer s & 1 holds the W tile of step s
    TX xn[2][16];
    double w[WPT];
    load_x(xn, r0);
    load_w(w, r0);
    store_w(w, 0);
    __syncthreads();
    for (long long s = 0; s < nfull; ++s) {
      TX xc[2][16];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int u = 0; u < 16; ++u) xc[mb][u] = xn[mb][u];
      const bool more = s + 1 < nfull;
      if (more) { load_x(xn, r0 + 64 * (s + 1)); load_w(w, r0 + 64 * (s + 1)); }
      compute(xc, int(s & 1), std::false_type{}, 0);
      if (more) store_w(w, int((s + 1) & 1));
      __syncthreads();
    }
  }
  // the split's ragged last step: rows past r1 read the centre (zero after
  // centring) and their W rows are zero
  const long long it = r0 + 64 * nfull;
  if (it < r1) {
    TX x[2][16];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const long long row = it + 16 * v + 4 * g + e;
          x[mb][4 * v + e] = (cvalid[mb] && row < r1) ? xcol[mb][row] : TX(0);
        }
    double w[WPT];
    load_w(w, it);
    store_w(w, 0);
    __syncthreads();
    compute(x, 0, std::true_type{}, r1 - it);
  }

  double FPLS_GLOBAL* o = gpw(t.out) + (long long)split * t.slab;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) {
      const int n = nb * 16 + m16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long oc = col0 + wave * 32 + mb * 16 + g + 4 * r;
        if (oc < t.k && n < t.N) o[oc + (long long)n * t.ldo] = acc[mb][nb][r];
      }
    }
  if (stats && t.stats != nullptr) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      double a1 = s1[mb], a2 = s2[mb];
      a1 += __shfl_xor(a1, 16, 64);
      a1 += __shfl_xor(a1, 32, 64);
      a2 += __shfl_xor(a2, 16, 64);
      a2 += __shfl_xor(a2, 32, 64);
      if (g == 0 && cvalid[mb]) {
        const long long col = col0 + wave * 32 + mb * 16 + m16;
        double FPLS_GLOBAL* st = gpw(t.stats) + (long long)split * 2 * t.k;
        st[col] = a1;
        st[t.k + col] = a2;
      }
    }
  }
}

// ------------------------------------------------------- xw, FP64 MFMA
// Same tiling as the bf16x6 kernel (256 rows, 32-column steps, lane rows
// 4r..4r+3 of columns 8g..8g+7); MFMA call (e, j) pairs column 8g + j of row
// 4r + e with row 8g + j of W. Pipelined like the xtw kernel above: the A
// values of step s + 1 and its W rows (and centres) are loaded while step s
// multiplies; W is staged in a double-buffered LDS tile shared by the four
// waves (row length NB + 1: conflict-free stores and operand reads). Each
// accumulator sums its products in the order steps, then j.
template<int NB, class TX>
__global__ __launch_bounds__(kWG) void k_xw_f64(const XwTask<TX>* tasks, int count,
                                                XwTask<TX> single) {
  constexpr int NBLK = NB / 16;
  constexpr int WLD = NB + 1;
  constexpr int WPT = 32 * NB / kWG;          // W values per thread and step
  const XwTask<TX> t = load_task(tasks, count, single);
  long long local = (long long)blockIdx.x - t.tile0;
  const long long rtiles = xw_row_tiles(t.m);
  if (local >= rtiles * t.splits) return;
  // neighbouring row tiles share the 128-byte line at their boundary
  // (columns of any length start unaligned): on one XCD, one L2 fetches it
  local = xcd_local(local, rtiles * t.splits);
  const int split = int(local / rtiles);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m16 = lane & 15, g = lane >> 4;
  const long long i0w = (local % rtiles) * kWG + wave * 64;
  const long long rowb = i0w + 4 * m16;
  const long long per = split_extent(t.k, t.splits, 32);
  const long long j0 = split * per;
  const long long j1 = j0 + per < t.k ? j0 + per : t.k;
  const bool rows_full = i0w + 64 <= t.m;
  const TX FPLS_GLOBAL* xr = gp(t.A) + (rowb < t.m ? rowb : 0);
  const double FPLS_GLOBAL* W = gp(t.W);
  const double FPLS_GLOBAL* C = t.center ? gp(t.center) : nullptr;

  __shared__ double ws[2][32][WLD];
  __shared__ double cs[2][32];

  f64x4 acc[4][NBLK];
#pragma unroll
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int nb = 0; nb < NBLK; ++nb) acc[e][nb] = f64x4{0.0, 0.0, 0.0, 0.0};

  // A values of a step (columns past j1 and rows past m read zero)
  auto load_x = [&](TX (&x)[8][4], long long jc, bool full) __attribute__((always_inline)) {
    const long long cb = jc + 8 * g;
    if (rows_full && full) {
#pragma unroll
      for (int j = 0; j < 8; ++j) load4_raw(xr + (cb + j) * t.m, x[j]);
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          x[j][e] = (cb + j < j1 && rowb + e < t.m) ? xr[(cb + j) * t.m + e] : TX(0);
    }
  };
  // W rows of a step: value u of thread tid is row (tid + 256 u) % 32 of
  // column (tid + 256 u) / 32; threads below 32 also load the centres
  auto load_w = [&](double (&w)[WPT], double& c, long long jc) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < WPT; ++u) {
      const int e = tid + u * kWG;
      const int r = e & 31, n = e >> 5;
      w[u] = (jc + r < j1 && n < t.N) ? W[(jc + r) + (long long)n * t.ldw] : 0.0;
    }
    if (tid < 32) c = (C != nullptr && jc + tid < j1) ? C[jc + tid] : 0.0;
  };
  auto store_w = [&](const double (&w)[WPT], double c, int b) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < WPT; ++u) {
      const int e = tid + u * kWG;
      ws[b][e & 31][e >> 5] = w[u];
    }
    if (tid < 32) cs[b][tid] = c;
  };
  // one step: the raw A values are widened and centred one column j at a
  // time, right before the products that use them
  auto compute = [&](const TX (&xr8)[8][4], int b) __attribute__((always_inline)) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const double c = cs[b][8 * g + j];
      double x[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) x[e] = double(xr8[j][e]) - c;
#pragma unroll
      for (int nb = 0; nb < NBLK; ++nb) {
        const double bv = ws[b][8 * g + j][nb * 16 + m16];
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[e][nb] = mma64(x[e], bv, acc[e][nb]);
      }
    }
  };

  const long long nfull = (j1 - j0) / 32;
  if (nfull > 0) {
    // xn / w / c hold the loads of the next step while the current one (xc)
    // multiplies; LDS buffer s & 1 holds the W rows of step s
    TX xn[8][4];
    double w[WPT], c = 0.0;
    load_x(xn, j0, true);
    load_w(w, c, j0);
    store_w(w, c, 0);
    __syncthreads();
    for (long long s = 0; s < nfull; ++s) {
      TX xc[8][4];
#pragma unroll
      for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) xc[j][e] = xn[j][e];
      const bool more = s + 1 < nfull;
      if (more) { load_x(xn, j0 + 32 * (s + 1), true); load_w(w, c, j0 + 32 * (s + 1)); }
      compute(xc, int(s & 1));
      if (more) store_w(w, c, int((s + 1) & 1));
      __syncthreads();
    }
  }
  // the split's ragged last step
  const long long jt = j0 + 32 * nfull;
  if (jt < j1) {
    TX x[8][4];
    double w[WPT], c = 0.0;
    load_x(x, jt, false);
    load_w(w, c, jt);
    store_w(w, c, 0);
    __syncthreads();
    compute(x, 0);
  }

  double FPLS_GLOBAL* o = gpw(t.out) + (long long)split * t.slab;
#pragma unroll
  for (int nb = 0; nb < NBLK; ++nb) {
    const int n = nb * 16 + m16;
    if (n >= t.N) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const long long row = i0w + 4 * (g + 4 * r) + e;
        if (row < t.m) o[row + (long long)n * t.ldo] = acc[e][nb][r];
      }
  }
}

__global__ __launch_bounds__(kWG) void k_sum_slabs(const double* part, int splits,
                                                   long long stride, long long count,
                                                   double* out) {
  const double FPLS_GLOBAL* P = gp(part);
  double FPLS_GLOBAL* O = gpw(out);
  for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < count;
       i += (long long)gridDim.x * kWG) {
    O[i] = sum_strided(P + i, stride, splits);
  }
}

// Grid (blocks, list entries): entry blockIdx.y of a slab-sum list.
__global__ __launch_bounds__(kWG) void k_sum_slabs_list(const SlabSum* list) {
  const SlabSum e = load_const(list, blockIdx.y);
  const double FPLS_GLOBAL* P = gp(e.part);
  double FPLS_GLOBAL* O = gpw(e.out);
  for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < e.count;
       i += (long long)gridDim.x * kWG) {
    O[i] = sum_strided(P + i, e.stride, e.splits);
  }
}

// ------------------------------------------------------------- dispatch
// Widest right-hand side of one FP32 matrix-core xw launch (its FP64
// accumulators and double-buffered operands fill the register file at 32).
constexpr int kXwBf6MaxN = 32;
// Lane offsets of the bf16x6 kernels are 32-bit byte offsets spanning at most
// 32 columns of A.
constexpr long long kMfmaMaxRows = 1LL << 24;
constexpr unsigned kSplitBlocks = 512;

template<class Fn>
void with_width(int maxn, Fn&& fn) {
  switch (mfma_width(maxn)) {
    case 16: fn(std::integral_constant<int, 16>{}); break;
    case 32: fn(std::integral_constant<int, 32>{}); break;
    case 48: fn(std::integral_constant<int, 48>{}); break;
    default: fn(std::integral_constant<int, 64>{}); break;
  }
}

template<class T>
void dispatch_xw(const XwTask<T>* d, int count, const XwTask<T>& single, long long tiles,
                 int maxn, hipStream_t stream, bool f64) {
  if (tiles <= 0) return;
  if (tiles > 0x7fffffffLL) throw std::invalid_argument("xw grid too large");
  const dim3 grid((unsigned)tiles), block((unsigned)kWG);
  if (maxn >= kMfmaMinN) {
    const bool bf6 = sizeof(T) == 4 && !f64;
    if (bf6 && maxn > kXwBf6MaxN)
      throw std::invalid_argument("FP32 matrix-core xw launches take at most 32 right-hand columns");
    with_width(maxn, [&](auto w) {
      constexpr int NB = decltype(w)::value;
      if constexpr (sizeof(T) == 4 && NB <= kXwBf6MaxN) {
        if (bf6) {
          hipLaunchKernelGGL((k_xw_split<NB>), dim3(kSplitBlocks, unsigned(d ? count : 1)), block,
                             0, stream, d, single);
          FPLS_LAUNCHED("k_xw_split");
          hipLaunchKernelGGL((k_xw_bf6<NB>), grid, dim3(kBf6Threads), 0, stream, d, count, single);
          FPLS_LAUNCHED("k_xw_bf6");
          return;
        }
      }
      hipLaunchKernelGGL((k_xw_f64<NB, T>), grid, block, 0, stream, d, count, single);
      FPLS_LAUNCHED("k_xw_f64");
    });
    return;
  }
  if (maxn <= 1) hipLaunchKernelGGL((k_xw<T, 1>), grid, block, 0, stream, d, count, single);
  else if (maxn <= 4) hipLaunchKernelGGL((k_xw<T, 4>), grid, block, 0, stream, d, count, single);
  else hipLaunchKernelGGL((k_xw<T, 8>), grid, block, 0, stream, d, count, single);
  FPLS_LAUNCHED("k_xw");
}

template<class T>
void dispatch_xtw(const XtwTask<T>* d, int count, const XtwTask<T>& single, long long tiles,
                  int maxn, hipStream_t stream, bool f64, bool stats) {
  if (tiles <= 0) return;
  if (tiles > 0x7fffffffLL) throw std::invalid_argument("xtw grid too large");
  const dim3 grid((unsigned)tiles), block((unsigned)kWG);
  if (maxn >= kMfmaMinN) {
    with_width(maxn, [&](auto w) {
      constexpr int NB = decltype(w)::value;
      if constexpr (sizeof(T) == 4) {
        if (!f64) {
          hipLaunchKernelGGL((k_xtw_split<NB>), dim3(kSplitBlocks, unsigned(d ? count : 1)), block,
                             0, stream, d, single);
          FPLS_LAUNCHED("k_xtw_split");
          hipLaunchKernelGGL((k_xtw_bf6<NB>), grid, dim3(kBf6Threads), 0, stream, d, count, single);
          FPLS_LAUNCHED("k_xtw_bf6");
          return;
        }
      }
      hipLaunchKernelGGL((k_xtw_f64<NB, T>), grid, block, 0, stream, d, count, single);
      FPLS_LAUNCHED("k_xtw_f64");
    });
    return;
  }
  if (maxn <= 1) {
    // the statistics variant carries three accumulators per column; a task
    // list is all statistics tasks or none (pack_xtw checks)
    if (stats) hipLaunchKernelGGL((k_xtw1<T, true>), grid, dim3(512), 0, stream, d, count, single);
    else hipLaunchKernelGGL((k_xtw1<T, false>), grid, dim3(512), 0, stream, d, count, single);
    FPLS_LAUNCHED("k_xtw1");
    return;
  }
  if (maxn <= 4) hipLaunchKernelGGL((k_xtw<T, 4>), grid, block, 0, stream, d, count, single);
  else hipLaunchKernelGGL((k_xtw<T, 16>), grid, block, 0, stream, d, count, single);
  FPLS_LAUNCHED("k_xtw");
}

}  // namespace

template<class T>
long long pack_xw(XwTask<T>* tasks, int count, int* maxn, bool f64) {
  long long tiles = 0;
  int mn = 0;
  for (int i = 0; i < count; ++i) mn = std::max(mn, tasks[i].N);
  for (int i = 0; i < count; ++i) {
    if (tasks[i].N < 1 || tasks[i].N > 64) throw std::invalid_argument("xw supports 1..64 right-hand columns");
    if (tasks[i].splits < 1) throw std::invalid_argument("xw needs at least one split");
    if (xw_frag_bytes<T>(tasks[i].k, mn, f64) > 0 && tasks[i].frag == nullptr)
      throw std::invalid_argument("xw task without its operand workspace");
    if (xw_frag_bytes<T>(tasks[i].k, mn, f64) > 0 && tasks[i].m > kMfmaMaxRows)
      throw std::invalid_argument("matrix-core products support at most 2^24 rows");
    tasks[i].tile0 = tiles;
    tiles += xw_row_tiles(tasks[i].m) * tasks[i].splits;
  }
  if (maxn) *maxn = mn;
  return tiles;
}

template<class T>
long long pack_xtw(XtwTask<T>* tasks, int count, int* maxn, bool f64, int min_width) {
  int mn = min_width;
  for (int i = 0; i < count; ++i) {
    if (tasks[i].N < 1 || tasks[i].N > 64) throw std::invalid_argument("xtw supports 1..64 right-hand columns");
    if ((tasks[i].mode == kShiftStats) != (tasks[0].mode == kShiftStats))
      throw std::invalid_argument("an xtw task list mixes statistics and plain tasks");
    if (tasks[i].splits < 1) throw std::invalid_argument("xtw needs at least one split");
    mn = std::max(mn, tasks[i].N);
  }
  long long tiles = 0;
  for (int i = 0; i < count; ++i) {
    if (xtw_frag_bytes<T>(tasks[i].m, mn, f64) > 0 && tasks[i].frag == nullptr)
      throw std::invalid_argument("xtw task without its operand workspace");
    if (xtw_frag_bytes<T>(tasks[i].m, mn, f64) > 0 && tasks[i].m > kMfmaMaxRows)
      throw std::invalid_argument("matrix-core products support at most 2^24 rows");
    tasks[i].tile0 = tiles;
    tiles += xtw_tiles(tasks[i].k, mn) * tasks[i].splits;
  }
  if (maxn) *maxn = mn;
  return tiles;
}

template<class T>
void launch_xw(const XwTask<T>* device_tasks, int count, long long tiles, int maxn,
               hipStream_t stream, bool f64) {
  if (count <= 0) return;
  if (count > 65535) throw std::invalid_argument("xw task list longer than 65535");
  dispatch_xw<T>(device_tasks, count, XwTask<T>{}, tiles, maxn, stream, f64);
}

template<class T>
void launch_xtw(const XtwTask<T>* device_tasks, int count, long long tiles, int maxn,
                hipStream_t stream, bool f64, bool stats) {
  if (count <= 0) return;
  if (count > 65535) throw std::invalid_argument("xtw task list longer than 65535");
  dispatch_xtw<T>(device_tasks, count, XtwTask<T>{}, tiles, maxn, stream, f64, stats);
}

template<class T>
void launch_xw_single(XwTask<T> task, hipStream_t stream, bool f64) {
  // FP32 bf16x6 launches take at most 32 right-hand columns: wider tasks run
  // as consecutive 32-column pieces (same stream, so the pieces may share the
  // operand workspace)
  const int step = (sizeof(T) == 4 && !f64 && task.N >= kMfmaMinN) ? kXwBf6MaxN : task.N;
  for (int c0 = 0; c0 < task.N; c0 += step) {
    XwTask<T> piece = task;
    piece.N = std::min(step, task.N - c0);
    piece.W = task.W + (long long)c0 * task.ldw;
    piece.out = task.out + (long long)c0 * task.ldo;
    int maxn = 0;
    const long long tiles = pack_xw(&piece, 1, &maxn, f64);
    dispatch_xw<T>(nullptr, 1, piece, tiles, maxn, stream, f64);
  }
}

template<class T>
void launch_xtw_single(XtwTask<T> task, hipStream_t stream, bool f64) {
  int maxn = 0;
  const long long tiles = pack_xtw(&task, 1, &maxn, f64);
  dispatch_xtw<T>(nullptr, 1, task, tiles, maxn, stream, f64, task.mode == kShiftStats);
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

void launch_sum_slabs_list(const SlabSum* device_list, int count, long long max_count,
                           hipStream_t stream) {
  if (count <= 0 || max_count <= 0) return;
  if (count > 65535) throw std::invalid_argument("slab-sum list longer than 65535");
  long long blocks = (max_count + kWG - 1) / kWG;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_sum_slabs_list, dim3(unsigned(blocks), unsigned(count)), dim3(kWG), 0,
                     stream, device_list);
  FPLS_LAUNCHED("k_sum_slabs_list");
}

#define FPLS_STREAM_INSTANCES(T)                                                          \
  template long long pack_xw<T>(XwTask<T>*, int, int*, bool);                              \
  template long long pack_xtw<T>(XtwTask<T>*, int, int*, bool, int);                       \
  template void launch_xw<T>(const XwTask<T>*, int, long long, int, hipStream_t, bool);   \
  template void launch_xtw<T>(const XtwTask<T>*, int, long long, int, hipStream_t, bool, bool); \
  template void launch_xw_single<T>(XwTask<T>, hipStream_t, bool);                         \
  template void launch_xtw_single<T>(XtwTask<T>, hipStream_t, bool);
FPLS_STREAM_INSTANCES(float)
FPLS_STREAM_INSTANCES(double)
#undef FPLS_STREAM_INSTANCES

}  // namespace fpls
