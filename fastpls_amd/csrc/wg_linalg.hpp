// FP64 dense linear algebra executed by ONE work-group of NT threads (NT a
// multiple of 64, kWG unless stated).
//
// Every routine must be called by all threads of the work-group with
// identical arguments; it synchronises internally and returns with the
// results visible to the whole work-group. Matrices are column-major and may
// live in global memory (L2-resident scratch) or LDS.
#pragma once

#include "common.hpp"

namespace fpls {
namespace wg {

// C(m x n) = alpha * op(A) * op(B) + beta * C ; thread per output element.
template<int NT = kWG>
__device__ inline void gemm(int m, int n, int k, double alpha, const double* A,
                            long lda, bool ta, const double* B, long ldb, bool tb,
                            double beta, double* C, long ldc) {
  for (int e = threadIdx.x; e < m * n; e += NT) {
    const int i = e % m, j = e / m;
    double acc = 0.0;
    for (int l = 0; l < k; ++l) {
      const double a = ta ? A[l + i * lda] : A[i + l * lda];
      const double b = tb ? B[j + l * ldb] : B[l + j * ldb];
      acc = fma(a, b, acc);
    }
    C[i + j * ldc] = alpha * acc + (beta != 0.0 ? beta * C[i + j * ldc] : 0.0);
  }
  __syncthreads();
}

// C(m x n) = Aᵀ B with A (k x m), B (k x n) and a long inner dimension k:
// one wave per output, lanes stride the inner dimension (coalesced), four
// independent partial sums so that eight loads per lane are in flight.
template<int NT = kWG>
__device__ __attribute__((noinline)) void gemm_tn_long(int m, int n, int k, const double* A, long lda,
                                    const double* B, long ldb, double* C, long ldc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int e = wave; e < m * n; e += (NT / 64)) {
    const int i = e % m, j = e / m;
    const double* a = A + (long)i * lda;
    const double* b = B + (long)j * ldb;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int l = lane;
    for (; l + 192 < k; l += 256) {
      s0 = fma(a[l], b[l], s0);
      s1 = fma(a[l + 64], b[l + 64], s1);
      s2 = fma(a[l + 128], b[l + 128], s2);
      s3 = fma(a[l + 192], b[l + 192], s3);
    }
    for (; l < k; l += 64) s0 = fma(a[l], b[l], s0);
    const double acc = wave_sum((s0 + s1) + (s2 + s3));
    if (lane == 0) C[i + j * ldc] = acc;
  }
  __syncthreads();
}

constexpr int kStageElems = 4096;
constexpr int kStageDoubles = kStageElems + 128;
typedef double wg_f64x4 __attribute__((ext_vector_type(4)));

// out (m) = Aᵀ b for a tall A (k x m, lda) and b (k), m <= MMAX, by one
// work-group of 256 threads: one thread per row of A (two rows per pass, b
// read once per row), m partial sums per thread, then per-wave sums and the
// four wave sums added in wave order (fixed order) through `red` (4 MMAX
// doubles of LDS). Ends with a barrier.
template<int MMAX>
__device__ __attribute__((noinline)) void gemm_tv_rows(int m, long k, const double* A, long lda,
                                                       const double* b, double* out,
                                                       double* red) {
  constexpr int NT = 256;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double FPLS_GLOBAL* Ag = (const double FPLS_GLOBAL*)A;
  const double FPLS_GLOBAL* bg = (const double FPLS_GLOBAL*)b;
  double acc[MMAX];
#pragma unroll
  for (int i = 0; i < MMAX; ++i) acc[i] = 0.0;
  for (long l0 = tid; l0 < k; l0 += 2 * NT) {
    const bool two = l0 + NT < k;
    const long l1 = two ? l0 + NT : l0;
    const double b0 = bg[l0];
    const double b1 = two ? bg[l1] : 0.0;
    double a0[MMAX], a1[MMAX];
#pragma unroll
    for (int i = 0; i < MMAX; ++i) {
      if (i < m) {
        a0[i] = Ag[l0 + (long)i * lda];
        a1[i] = Ag[l1 + (long)i * lda];
      }
    }
#pragma unroll
    for (int i = 0; i < MMAX; ++i)
      if (i < m) acc[i] = fma(a1[i], b1, fma(a0[i], b0, acc[i]));
  }
#pragma unroll
  for (int i = 0; i < MMAX; ++i) {
    if (i < m) {
      const double v = wave_sum(acc[i]);
      if (lane == 0) red[wave * MMAX + i] = v;
    }
  }
  __syncthreads();
  if (tid < m) out[tid] = ((red[tid] + red[MMAX + tid]) + red[2 * MMAX + tid]) + red[3 * MMAX + tid];
  __syncthreads();
}

// C (m x n, ldc) = Aᵀ B for A (k x m, lda) and B (k x n, ldb), m, n <= 64,
// with a long inner dimension k, on the FP64 matrix cores by one work-group of
// 256 threads (16x16x4 FP64 MFMA: lane l supplies column l % 16 of a 16 x 16
// tile at row l / 16 of each 4-row step; item r of lane l holds tile entry
// (l/16 + 4r, l % 16)). The 4-row steps of k are dealt to the four waves in
// turn and the operands are loaded straight from global memory, U steps at a
// time, the next U while the current ones multiply; TM x TN tiles cover
// m <= 16 TM, n <= 16 TN; SYM (A == B) loads one operand stream. The wave
// partials are added in wave order (fixed order) through `red` (4096 doubles
// of LDS). Ends with a barrier.
template<int TM, int TN, int U, bool SYM>
__device__ __attribute__((noinline)) void gemm_tn_direct(int m, int n, long k, const double* A,
                                                         long lda, const double* B, long ldb,
                                                         double* C, long ldc, double* red) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m16 = lane & 15, g = lane >> 4;
  // columns past m (n) read column 0: output (i, j) depends on columns i
  // and j only, and entries past m x n are not written
  const double FPLS_GLOBAL* pa[TM];
  const double FPLS_GLOBAL* pb[TN];
#pragma unroll
  for (int t = 0; t < TM; ++t) {
    const int c = t * 16 + m16;
    pa[t] = (const double FPLS_GLOBAL*)(A + (long)(c < m ? c : 0) * lda);
  }
#pragma unroll
  for (int t = 0; t < TN; ++t) {
    const int c = t * 16 + m16;
    pb[t] = (const double FPLS_GLOBAL*)(B + (long)(c < n ? c : 0) * ldb);
  }
  wg_f64x4 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = wg_f64x4{0.0, 0.0, 0.0, 0.0};
  // wave w owns a contiguous quarter of the 4-row steps (so its loads use
  // whole cache lines); an iteration covers U consecutive steps and the
  // operands of the next iteration are loaded before the matrix products of
  // the current one (two register sets)
  const long nsteps = (k + 3) / 4;
  const long quarter = (nsteps + 3) / 4;
  const long sb = wave * quarter;
  const long se = sb + quarter < nsteps ? sb + quarter : nsteps;
  double av[2][U][TM], bv[2][U][SYM ? 1 : TN];
  // A block of U steps whose rows all lie in the wave's range and in k is
  // loaded without conditions, so its loads stay in flight together while
  // the other set multiplies; only a wave's last block masks rows (a masked
  // load is waited for where it is selected).
  auto load = [&](int set, long s0) {
    if (s0 + U <= se && 4 * (s0 + U) <= k) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long row = 4 * (s0 + u) + g;
#pragma unroll
        for (int t = 0; t < TM; ++t) av[set][u][t] = pa[t][row];
        if constexpr (!SYM) {
#pragma unroll
          for (int t = 0; t < TN; ++t) bv[set][u][t] = pb[t][row];
        }
      }
      return;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long row = 4 * (s0 + u) + g;
      const bool ok = s0 + u < se && row < k;
      const long rr = ok ? row : 0;
#pragma unroll
      for (int t = 0; t < TM; ++t) {
        const double v = pa[t][rr];
        av[set][u][t] = ok ? v : 0.0;
      }
      if constexpr (!SYM) {
#pragma unroll
        for (int t = 0; t < TN; ++t) {
          const double v = pb[t][rr];
          bv[set][u][t] = ok ? v : 0.0;
        }
      }
    }
  };
  auto multiply = [&](int set) {
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) {
          const double bb = SYM ? av[set][u][b] : bv[set][u][SYM ? 0 : b];
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[set][u][a], bb, acc[a][b], 0, 0, 0);
        }
  };
  long s0 = sb;
  if (s0 < se) load(0, s0);
  for (; s0 < se; s0 += 2 * U) {
    if (s0 + U < se) load(1, s0 + U);
    multiply(0);
    if (s0 + U >= se) break;
    if (s0 + 2 * U < se) load(0, s0 + 2 * U);
    multiply(1);
  }
  for (int w2 = 0; w2 < 4; ++w2) {
    if (wave == w2) {
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = a * 16 + g + 4 * r, col = b * 16 + m16;
            if (w2 == 0) red[row + 64 * col] = acc[a][b][r];
            else red[row + 64 * col] += acc[a][b][r];
          }
    }
    __syncthreads();
  }
  for (int e = tid; e < m * n; e += 256) {
    const int row = e % m, col = e / m;
    C[row + (long)col * ldc] = red[row + 64 * col];
  }
  __syncthreads();
}

// Dispatch of gemm_tn_direct by tile counts (m, n <= 64).
// A Gram (A == B, m == n) loads each operand once.
__device__ inline void gemm_tn_fast(int m, int n, long k, const double* A, long lda,
                                    const double* B, long ldb, double* C, long ldc, double* red) {
  const int tm = (m + 15) / 16, tn = (n + 15) / 16;
  const bool sym = A == B && lda == ldb && m == n;
  if (sym) {
    if (tm <= 1) gemm_tn_direct<1, 1, 32, true>(m, n, k, A, lda, B, ldb, C, ldc, red);
    else if (tm <= 2) gemm_tn_direct<2, 2, 8, true>(m, n, k, A, lda, B, ldb, C, ldc, red);
    else gemm_tn_direct<4, 4, 4, true>(m, n, k, A, lda, B, ldb, C, ldc, red);
  } else if (tm <= 1 && tn <= 1) {
    gemm_tn_direct<1, 1, 24, false>(m, n, k, A, lda, B, ldb, C, ldc, red);
  } else if (tm <= 2 && tn <= 2) {
    gemm_tn_direct<2, 2, 8, false>(m, n, k, A, lda, B, ldb, C, ldc, red);
  } else {
    gemm_tn_direct<4, 4, 4, false>(m, n, k, A, lda, B, ldb, C, ldc, red);
  }
}

// Bs (k x NMAX, row-major: Bs[l NMAX + j]) = B (k x n, ldb) padded with zero
// columns, for gemm_rows_lds. Ends with a barrier.
template<int NMAX, int NT = kWG>
__device__ inline void stage_rows(const double* B, long ldb, int k, int n, double* Bs) {
  for (int e = threadIdx.x; e < k * NMAX; e += NT) {
    const int l = e / NMAX, j = e % NMAX;
    Bs[e] = j < n ? B[l + (long)j * ldb] : 0.0;
  }
  __syncthreads();
}

// C (m x n, ldc) = A B for a tall A (m x k, lda) in global memory and B
// staged by stage_rows in LDS (n <= NMAX): one thread per row of C (R rows
// per pass: 4 for NMAX <= 16, else 2), B read as LDS broadcasts, no
// data-dependent branches. A row is read in full before it is written, so C
// may be A itself (ldc == lda, n <= k). Ends with a barrier.
template<int NMAX, int NT = kWG>
__device__ __attribute__((noinline)) void gemm_rows_lds(long m, int n, int k, const double* A,
                                                        long lda, const double* Bs, double* C,
                                                        long ldc) {
  // rows past m are clamped to the pass's first row and not written, so
  // every row's loads are in flight together
  constexpr int R = NMAX <= 16 ? 4 : 2;
  const double FPLS_GLOBAL* Ag = (const double FPLS_GLOBAL*)A;
  for (long i0 = threadIdx.x; i0 < m; i0 += R * NT) {
    long ir[R];
#pragma unroll
    for (int r = 0; r < R; ++r) ir[r] = i0 + r * NT < m ? i0 + r * NT : i0;
    double acc[R][NMAX];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int j = 0; j < NMAX; ++j) acc[r][j] = 0.0;
    int l = 0;
    for (; l + 4 <= k; l += 4) {
      double a[R][4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < R; ++r) a[r][u] = Ag[ir[r] + (long)(l + u) * lda];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int j = 0; j < NMAX; ++j) {
          const double bv = Bs[(l + u) * NMAX + j];
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r][j] = fma(a[r][u], bv, acc[r][j]);
        }
    }
    for (; l < k; ++l) {
      double a[R];
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = Ag[ir[r] + (long)l * lda];
#pragma unroll
      for (int j = 0; j < NMAX; ++j) {
        const double bv = Bs[l * NMAX + j];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r][j] = fma(a[r], bv, acc[r][j]);
      }
    }
#pragma unroll
    for (int j = 0; j < NMAX; ++j)
      if (j < n) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (r == 0 || i0 + r * NT < m) C[ir[r] + (long)j * ldc] = acc[r][j];
      }
  }
  __syncthreads();
}

template<int NT = kWG>
__device__ inline void copy(int m, int n, const double* A, long lda, double* B, long ldb) {
  for (int e = threadIdx.x; e < m * n; e += NT) {
    const int i = e % m, j = e / m;
    B[i + j * ldb] = A[i + j * lda];
  }
  __syncthreads();
}

// Euclidean norm of a vector (all threads receive it).
template<int NT = kWG>
__device__ inline double norm(const double* v, long n, double* red) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  long i = threadIdx.x;
  for (; i + 3 * NT < n; i += 4 * NT) {
    s0 = fma(v[i], v[i], s0);
    s1 = fma(v[i + NT], v[i + NT], s1);
    s2 = fma(v[i + 2 * NT], v[i + 2 * NT], s2);
    s3 = fma(v[i + 3 * NT], v[i + 3 * NT], s3);
  }
  for (; i < n; i += NT) s0 = fma(v[i], v[i], s0);
  return sqrt(fmax(wg_sum_n<NT>((s0 + s1) + (s2 + s3), red), 0.0));
}

// d_c = Σ_{i=j..m-1} v_i B[i, c] for the columns c in [c0, n), with the
// Householder vector v of column j of V (v_j = 1, v_i = V[i, j] below).
// G threads share a column (G a power of two, G · (n - c0) <= NT): each sums
// a strided subset of the rows and the G partials are added in LDS, so no
// step waits on a chain of cross-lane shuffles. `part` is NT doubles of LDS.
template<int NT = kWG>
__device__ inline void reflector_dots(const double* V, long ldv, const double* B, long ldb,
                                      int m, int j, int c0, int n, double* dots,
                                      double* part) {
  const int tid = threadIdx.x;
  for (int base = c0; base < n; base += NT) {
    const int cols = n - base < NT ? n - base : NT;
    int lg = 0;
    while ((2 << lg) * cols <= NT && lg < 6) ++lg;
    const int G = 1 << lg;
    const int c = base + (tid >> lg), g = tid & (G - 1);
    double acc = 0.0;
    if (tid < G * cols) {
      const double* b = B + (long)c * ldb;
      const double* v = V + (long)j * ldv;
      int i = j + g;
      if (g == 0) {
        acc = b[j];
        i += G;
      }
      for (; i < m; i += G) acc = fma(v[i], b[i], acc);
    }
    part[tid] = acc;
    __syncthreads();
    if (tid < cols) {
      double s = 0.0;
      for (int t = 0; t < G; ++t) s += part[(tid << lg) + t];
      dots[base + tid] = s;
    }
    __syncthreads();
  }
}

// Householder economy QR of A (m x n, m >= n), overwritten by the reflectors;
// the explicit Q (m x n) goes to Q. tau and dots are LDS arrays of n doubles,
// part NT doubles of LDS. Returns false on a non-finite column norm.
template<int NT = kWG>
__device__ __attribute__((noinline)) bool householder_qr(double* A, long lda, int m, int n, double* Q,
                                      long ldq, double* tau, double* dots, double* red,
                                      double* part) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  bool finite = true;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = j + tid; i < m; i += NT) s = fma(A[i + j * lda], A[i + j * lda], s);
    const double norm2 = wg_sum_n<NT>(s, red);
    const double alpha = A[j + j * lda];
    const double x2 = norm2 - alpha * alpha;
    double tj = 0.0, beta = alpha, scale = 0.0;
    if (!isfinite(norm2)) finite = false;
    if (finite && x2 > 0.0) {
      const double nrm = sqrt(norm2);
      beta = alpha >= 0.0 ? -nrm : nrm;
      tj = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    __syncthreads();
    if (tj != 0.0) {
      for (int i = j + 1 + tid; i < m; i += NT) A[i + j * lda] *= scale;
      if (tid == 0) A[j + j * lda] = beta;
    }
    if (tid == 0) tau[j] = tj;
    __syncthreads();
    if (tj != 0.0 && j + 1 < n) {
      // dots of the trailing columns with the reflector (its unit head is
      // implied: the stored diagonal is beta)
      reflector_dots<NT>(A, lda, A, lda, m, j, j + 1, n, dots, part);
      // one wave per column, lanes over the rows (no index division)
      for (int c = j + 1 + wave; c < n; c += (NT / 64)) {
        const double d = dots[c] * tj;
        for (int i = j + lane; i < m; i += 64) {
          const double v = i == j ? 1.0 : A[i + j * lda];
          A[i + c * lda] -= d * v;
        }
      }
      __syncthreads();
    }
  }
  for (int c = wave; c < n; c += (NT / 64))
    for (int i = lane; i < m; i += 64) Q[i + c * ldq] = i == c ? 1.0 : 0.0;
  __syncthreads();
  for (int j = n - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj == 0.0) continue;
    reflector_dots<NT>(A, lda, Q, ldq, m, j, j, n, dots, part);
    for (int c = j + wave; c < n; c += (NT / 64)) {
      const double d = dots[c] * tj;
      for (int i = j + lane; i < m; i += 64) {
        const double v = i == j ? 1.0 : A[i + j * lda];
        Q[i + c * ldq] -= d * v;
      }
    }
    __syncthreads();
  }
  return finite;
}

// ------------------------------------------------------ one-wave routines
// Value of v in lane l (l uniform).
__device__ inline double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(b & 0xffffffffLL), l));
  const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(b >> 32), l));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Orders the LDS accesses of one wave: stores issued before it are seen by
// the loads of every lane after it. A wave's LDS operations execute in issue
// order, so only compiler reordering has to be prevented.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Cholesky factorisation of the symmetric w x w matrix G (w <= W = 32) by
// ONE wave, and the inverse of its upper factor: G = Rᵀ R, Rinv = R⁻¹ (upper
// triangular, leading dimension ldr), so that A R⁻¹ has orthonormal columns
// when G = AᵀA. The factorisation is G = L D Lᵀ (unit L, right-looking on
// unscaled columns), R = D^½ Lᵀ and R⁻¹ = L⁻ᵀ D^-½. Lane i keeps row i of the
// trailing matrix and row i of L⁻¹ in registers; step j eliminates column j
// from both (Gauss-Jordan on [G | I]), broadcasting column j of G and row j
// of L⁻¹ through LDS. A pivot that is non-finite or not above 1e-13 of the
// largest diagonal entry of G makes it return false (uniformly) without
// writing Rinv. lc is W·W doubles of LDS, xb W doubles; G may be in LDS or
// global memory and is read before anything is written.
template<int W>
__device__ __forceinline__ bool wave_chol_inv(const double* G, long ldg, int w, double* Rinv,
                                              long ldr, double* lc, double* xb) {
  static_assert(W == 32, "register rows of 32 entries");
  const int lane = threadIdx.x & 63;
  double a[W], b[W];
#pragma unroll
  for (int c = 0; c < W; ++c) {
    a[c] = (lane < w && c < w) ? G[lane + c * ldg] : 0.0;
    b[c] = c == lane ? 1.0 : 0.0;
  }
  double dg = lane < w ? G[lane + (long)lane * ldg] : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) dg = fmax(dg, __shfl_xor(dg, off, 64));
  const double floor = 1e-13 * dg;
  wave_lds_sync();
  // dn: this lane's diagonal entry after the previous step, formed from its
  // own registers (the same operation the trailing update applies), so the
  // next pivot does not wait for the LDS exchange.
  double dn = a[0], mine = 1.0;
  bool ok = true;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    if (j >= w || !ok) continue;
    const double d = readlane_d(dn, j);
    if (!isfinite(d) || !(d > floor)) {
      ok = false;
      continue;
    }
    if (lane == j) mine = d;
    const double l = a[j] * (1.0 / d);
    const double lb = lane > j ? l : 0.0;  // rows below the pivot only
    if (lane < W) lc[j * W + lane] = a[j];
    if (lane == j) {
#pragma unroll
      for (int c = 0; c <= j; ++c) xb[c] = b[c];
    }
    if (j + 1 < W) dn = fma(-l, a[j], a[j + 1]);
    wave_lds_sync();
#pragma unroll
    for (int c = j + 1; c < W; ++c) a[c] = fma(-l, lc[j * W + c], a[c]);
#pragma unroll
    for (int c = 0; c <= j; ++c) b[c] = fma(-lb, xb[c], b[c]);
  }
  if (!ok) return false;
  // lane i holds row i of L⁻¹: column i of R⁻¹ is that row times d_i^-½
  const double rs = 1.0 / sqrt(mine);
#pragma unroll
  for (int r = 0; r < W; ++r)
    if (r < w && lane < w) Rinv[r + lane * ldr] = b[r] * rs;
  return true;
}

// CholQR2 of a tall A (m x w, leading dimension m, w <= 32) in global memory
// by one work-group of 256 threads, into Q (m x w, leading dimension m, not
// aliasing A): each Gram on the matrix cores (gemm_tn_fast), its factor by one
// wave (wave_chol_inv), the rows transformed by gemm_rows_lds (the second
// pass in place). Returns false, with A untouched, when a Cholesky factor is
// unusable. LDS: stage (2 kStageDoubles), gm, ri, lc (32·32 each), xb (32),
// okflag.
__device__ __attribute__((noinline)) bool cholqr2_tall(const double* A, long m, int w, double* Q,
                                                       double* stage, double* gm, double* ri,
                                                       double* lc, double* xb, int* okflag) {
  for (int pass = 0; pass < 2; ++pass) {
    const double* X = pass == 0 ? A : Q;
    gemm_tn_fast(w, w, m, X, m, X, m, gm, 32, stage);
    if (threadIdx.x < 64) {
      const bool good = wave_chol_inv<32>(gm, 32, w, ri, 32, lc, xb);
      if (threadIdx.x == 0) *okflag = good ? 1 : 0;
    }
    __syncthreads();
    if (*okflag == 0) return false;
    // Q = X R⁻¹ with the narrowest staged width that holds w columns
    if (w <= 8) {
      stage_rows<8, kWG>(ri, 32, w, w, gm);
      gemm_rows_lds<8, kWG>(m, w, w, X, m, gm, Q, m);
    } else if (w <= 16) {
      stage_rows<16, kWG>(ri, 32, w, w, gm);
      gemm_rows_lds<16, kWG>(m, w, w, X, m, gm, Q, m);
    } else {
      stage_rows<32, kWG>(ri, 32, w, w, gm);
      gemm_rows_lds<32, kWG>(m, w, w, X, m, gm, Q, m);
    }
  }
  return true;
}

// Pair (a, b) of round r for k-th slot of the round-robin (circle) schedule
// over m (even) players.
template<int NT = kWG>
__device__ inline void rr_pair(int m, int r, int k, int& a, int& b) {
  auto at = [&](int pos) { return pos == 0 ? 0 : ((pos - 1 + r) % (m - 1)) + 1; };
  a = at(k);
  b = at(m - 1 - k);
  if (a > b) { const int t = a; a = b; b = t; }
}

// 1/x and 1/sqrt(x) for positive normal x: the hardware estimate refined by
// two Newton steps (a few ulp, shorter dependency chain than IEEE division
// and square root).
__device__ __forceinline__ double rcp_refined(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

__device__ __forceinline__ double rsq_refined(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double e = fma(-(h * y), y, 0.5);
  y = fma(y, e, y);
  e = fma(-(h * y), y, 0.5);
  return fma(y, e, y);
}

// Rotation (c, s) that annihilates a_pq of the symmetric 2 x 2 block
// [[a_pp, a_pq], [a_pq, a_qq]]: t = sign(θ) / (|θ| + sqrt(1 + θ²)) with
// θ = (a_qq − a_pp) / (2 a_pq), written as sign(θ)|e| / (|d| + sqrt(d² + e²))
// for d = a_qq − a_pp, e = 2 a_pq scaled by a power of two (no overflow or
// underflow); c = 1 / sqrt(1 + t²), s = t c. Returns c = 1, s = 0 when a_pq is
// zero.
__device__ __forceinline__ void jacobi_rotation(double app, double aqq, double apq, double& c,
                                                double& s) {
  c = 1.0;
  s = 0.0;
  if (apq == 0.0) return;
  double d = aqq - app, e = 2.0 * apq;
  const int k = ilogb(fmax(fabs(d), fabs(e)));
  d = ldexp(d, -k);
  e = ldexp(e, -k);
  const double h = fma(d, d, e * e);
  const double root = h * rsq_refined(h);
  const double sg = (d == 0.0 || (d > 0.0) == (e > 0.0)) ? 1.0 : -1.0;
  const double t = sg * fabs(e) * rcp_refined(fabs(d) + root);
  c = rsq_refined(fma(t, t, 1.0));
  s = t * c;
}

// Player at position pos of round r of the round-robin (circle) schedule
// over m (even) players: position 0 is fixed, the others rotate.
__device__ __forceinline__ int rr_at(int pos, int r, int m) {
  if (pos == 0) return 0;
  int x = pos - 1 + r;
  if (x >= m - 1) x -= m - 1;
  return x + 1;
}

// Parallel cyclic Jacobi eigen-decomposition of the symmetric n x n matrix A
// (leading dimension n). On return V holds the eigenvectors as columns and
// w the eigenvalues, ascending (the dsyevd contract). cs/sn/pa/pb are LDS
// scratch of n/2+1 entries each. Returns false on non-finite input.
// Each round of the round-robin schedule rotates n/2 disjoint planes: one
// thread per plane forms its rotation, then one pass applies Jᵀ A J block by
// block (2 x 2 blocks of the row and column planes, the upper ones mirrored,
// so A stays exactly symmetric) together with V J. Up to n = 64 (with 256
// threads) every thread keeps its block and V assignments in registers.
template<int NT = kWG>
__device__ __attribute__((noinline)) bool jacobi_eigen(double* A, double* V, int n, double* w, double* cs,
                                    double* sn, int* pa, int* pb, double* red,
                                    double* vtmp /* n x n scratch */) {
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += NT) V[e] = (e % n) == (e / n) ? 1.0 : 0.0;
  double fin = 0.0;
  for (int e = tid; e < n * n; e += NT) fin += isfinite(A[e]) ? 0.0 : 1.0;
  if (wg_sum_n<NT>(fin, red) != 0.0) return false;
  const int m = n + (n & 1);
  const int half = m / 2;
  // per-thread assignments: blocks (k <= l) and V row-plane items (l, i)
  constexpr int kMaxB = (32 * 33 / 2 + NT - 1) / NT, kMaxV = (32 * 64 + NT - 1) / NT;
  const bool cached = n <= 64;
  int bk[kMaxB], bl[kMaxB], vl[kMaxV], vi[kMaxV];
  int nbk = 0, nvi = 0;
  if (cached) {
    int k = 0, rem = tid;
#pragma unroll
    for (int u = 0; u < kMaxB; ++u) {
      while (k < half && rem >= half - k) { rem -= half - k; ++k; }
      bk[u] = k < half ? k : -1;
      bl[u] = k + rem;
      if (k < half) nbk = u + 1;
      rem += NT;
    }
#pragma unroll
    for (int u = 0; u < kMaxV; ++u) {
      const int e = tid + u * NT;
      vl[u] = e / (n > 0 ? n : 1);
      vi[u] = e % (n > 0 ? n : 1);
      if (e < half * n) nvi = u + 1;
    }
  }
  auto block = [&](int k, int l) {
    const int ak = pa[k], bkk = pb[k], al = pa[l], bll = pb[l];
    const double ck = cs[k], sk = sn[k], cl = cs[l], sl = sn[l];
    const double x00 = A[ak + al * n];
    const double x01 = bll >= 0 ? A[ak + bll * n] : 0.0;
    const double x10 = bkk >= 0 ? A[bkk + al * n] : 0.0;
    const double x11 = (bkk >= 0 && bll >= 0) ? A[bkk + bll * n] : 0.0;
    // columns: (x0, x1) -> (c x0 - s x1, s x0 + c x1), then rows
    const double y00 = cl * x00 - sl * x01, y01 = sl * x00 + cl * x01;
    const double y10 = cl * x10 - sl * x11, y11 = sl * x10 + cl * x11;
    const double z00 = ck * y00 - sk * y10, z10 = sk * y00 + ck * y10;
    const double z01 = ck * y01 - sk * y11, z11 = sk * y01 + ck * y11;
    A[ak + al * n] = z00;
    A[al + ak * n] = z00;
    if (bll >= 0) { A[ak + bll * n] = z01; A[bll + ak * n] = z01; }
    if (bkk >= 0) { A[bkk + al * n] = z10; A[al + bkk * n] = z10; }
    if (bkk >= 0 && bll >= 0) { A[bkk + bll * n] = z11; A[bll + bkk * n] = z11; }
  };
  auto vrot = [&](int l, int i) {
    const int a = pa[l], b = pb[l];
    if (b < 0) return;
    const double c = cs[l], s = sn[l];
    if (s == 0.0) return;
    const double va = V[i + a * n], vb = V[i + b * n];
    V[i + a * n] = c * va - s * vb;
    V[i + b * n] = s * va + c * vb;
  };
  for (int sweep = 0; sweep < 40 && n > 1; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int e = tid; e < n * n; e += NT) {
      const int i = int(e % n), j = int(e / n);
      if (i < j) off = fma(A[e], A[e], off);
      else if (i == j) dg = fma(A[e], A[e], dg);
    }
    off = wg_sum_n<NT>(off, red);
    dg = wg_sum_n<NT>(dg, red);
    if (off == 0.0 || off <= 1e-34 * dg) break;
    for (int r = 0; r < m - 1; ++r) {
      if (tid < half) {
        int a = rr_at(tid, r, m), b = rr_at(m - 1 - tid, r, m);
        if (a > b) { const int t = a; a = b; b = t; }
        double c = 1.0, s = 0.0;
        if (b < n) jacobi_rotation(A[a + a * n], A[b + b * n], A[a + b * n], c, s);
        else b = -1;
        cs[tid] = c; sn[tid] = s; pa[tid] = a; pb[tid] = b;
      }
      __syncthreads();
      if (cached) {
#pragma unroll
        for (int u = 0; u < kMaxB; ++u)
          if (u < nbk) block(bk[u], bl[u]);
#pragma unroll
        for (int u = 0; u < kMaxV; ++u)
          if (u < nvi) vrot(vl[u], vi[u]);
      } else {
        for (int e = tid; e < half * half; e += NT) {
          const int k = e % half, l = e / half;
          if (k <= l) block(k, l);
        }
        for (int e = tid; e < half * n; e += NT) vrot(int(e / n), int(e % n));
      }
      __syncthreads();
    }
  }
  // stable ascending order by rank
  for (int i = tid; i < n; i += NT) {
    const double di = A[i + i * n];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = A[j + j * n];
      rank += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    w[rank] = di;
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += NT) {
    const int i = int(e % n), j = int(e / n);
    const double dj = A[j + j * n];
    int rank = 0;
    for (int l = 0; l < n; ++l) {
      const double dl = A[l + l * n];
      rank += (dl < dj || (dl == dj && l < j)) ? 1 : 0;
    }
    vtmp[i + (long)rank * n] = V[i + (long)j * n];
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += NT) V[e] = vtmp[e];
  __syncthreads();
  return true;
}

// One-sided (Hestenes) Jacobi SVD of W (m x n, m >= n), leading dimension m.
// On return: s (n, descending), U (m x n) left vectors, Vm (n x n) right
// vectors as columns. W is destroyed. cs/sn/pa/pb: LDS, n/2+1 entries.
template<int NT = kWG>
__device__ __attribute__((noinline)) bool jacobi_svd(double* W, int m, int n, double* U, double* Vm,
                                  double* s, double* cs, double* sn, int* pa, int* pb,
                                  double* red, double* sc /* 3*(n/2+1) LDS */) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < n * n; e += NT) Vm[e] = (e % n) == (e / n) ? 1.0 : 0.0;
  const int mm = n + (n & 1);
  const int half = mm / 2;
  for (int sweep = 0; sweep < 40 && n > 1; ++sweep) {
    double rotated = 0.0;
    for (int r = 0; r < mm - 1; ++r) {
      // dot products per pair: one wave per pair
      for (int k = wave; k < half; k += (NT / 64)) {
        int a, b;
        rr_pair(mm, r, k, a, b);
        double al = 0.0, be = 0.0, ga = 0.0;
        if (b < n) {
          for (int i = lane; i < m; i += 64) {
            const double wa = W[i + (long)a * m], wb = W[i + (long)b * m];
            al = fma(wa, wa, al); be = fma(wb, wb, be); ga = fma(wa, wb, ga);
          }
        }
        al = wave_sum(al); be = wave_sum(be); ga = wave_sum(ga);
        if (lane == 0) {
          double c = 1.0, sv = 0.0;
          if (b < n && ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
            const double zeta = (be - al) / (2.0 * ga);
            const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            c = 1.0 / sqrt(1.0 + t * t);
            sv = c * t;
          }
          cs[k] = c; sn[k] = sv; pa[k] = a; pb[k] = b < n ? b : -1;
        }
      }
      __syncthreads();
      for (int k = 0; k < half; ++k) rotated += (sn[k] != 0.0) ? 1.0 : 0.0;
      for (int e = tid; e < half * m; e += NT) {
        const int k = int(e / m), i = int(e % m);
        const int a = pa[k], b = pb[k];
        if (b < 0 || sn[k] == 0.0) continue;
        const double c = cs[k], sv = sn[k];
        const double wa = W[i + (long)a * m], wb = W[i + (long)b * m];
        W[i + (long)a * m] = c * wa - sv * wb;
        W[i + (long)b * m] = sv * wa + c * wb;
      }
      for (int e = tid; e < half * n; e += NT) {
        const int k = int(e / n), i = int(e % n);
        const int a = pa[k], b = pb[k];
        if (b < 0 || sn[k] == 0.0) continue;
        const double c = cs[k], sv = sn[k];
        const double va = Vm[i + (long)a * n], vb = Vm[i + (long)b * n];
        Vm[i + (long)a * n] = c * va - sv * vb;
        Vm[i + (long)b * n] = sv * va + c * vb;
      }
      __syncthreads();
    }
    if (rotated == 0.0) break;
  }
  (void)sc;
  // column norms
  for (int j = wave; j < n; j += (NT / 64)) {
    double acc = 0.0;
    for (int i = lane; i < m; i += 64) acc = fma(W[i + (long)j * m], W[i + (long)j * m], acc);
    acc = wave_sum(acc);
    if (lane == 0) s[j] = sqrt(acc);
  }
  __syncthreads();
  // rank by descending norm (stable); write U and reorder Vm through U's
  // storage is not possible in place, so permute into U directly and keep a
  // rank table in red-sized chunks via recomputation.
  for (int e = tid; e < m * n; e += NT) {
    const int i = int(e % m), j = int(e / m);
    const double sj = s[j];
    int rank = 0;
    for (int l = 0; l < n; ++l) rank += (s[l] > sj || (s[l] == sj && l < j)) ? 1 : 0;
    U[i + (long)rank * m] = sj > 0.0 ? W[i + (long)j * m] / sj : 0.0;
  }
  __syncthreads();
  // reorder Vm columns using W as scratch (W no longer needed; m >= n)
  for (int e = tid; e < n * n; e += NT) {
    const int i = int(e % n), j = int(e / n);
    const double sj = s[j];
    int rank = 0;
    for (int l = 0; l < n; ++l) rank += (s[l] > sj || (s[l] == sj && l < j)) ? 1 : 0;
    W[i + (long)rank * n] = Vm[i + (long)j * n];
  }
  __syncthreads();
  double fin = 0.0;
  for (int e = tid; e < n * n; e += NT) { Vm[e] = W[e]; fin += isfinite(W[e]) ? 0.0 : 1.0; }
  __syncthreads();
  // sort singular values descending
  if (tid == 0) {
    for (int i = 1; i < n; ++i) {
      const double v = s[i];
      int j = i - 1;
      while (j >= 0 && s[j] < v) { s[j + 1] = s[j]; --j; }
      s[j + 1] = v;
    }
  }
  __syncthreads();
  return wg_sum_n<NT>(fin, red) == 0.0;
}

}  // namespace wg
}  // namespace fpls
