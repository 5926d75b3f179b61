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
// one wave per output, lanes stride the inner dimension (coalesced).
template<int NT = kWG>
__device__ inline void gemm_tn_long(int m, int n, int k, const double* A, long lda,
                                    const double* B, long ldb, double* C, long ldc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int e = wave; e < m * n; e += (NT / 64)) {
    const int i = e % m, j = e / m;
    double acc = 0.0;
    for (int l = lane; l < k; l += 64) acc = fma(A[l + i * lda], B[l + (long)j * ldb], acc);
    acc = wave_sum(acc);
    if (lane == 0) C[i + j * ldc] = acc;
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
  double s = 0.0;
  for (long i = threadIdx.x; i < n; i += NT) s = fma(v[i], v[i], s);
  return sqrt(fmax(wg_sum_n<NT>(s, red), 0.0));
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
__device__ inline bool householder_qr(double* A, long lda, int m, int n, double* Q,
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

// Pair (a, b) of round r for k-th slot of the round-robin (circle) schedule
// over m (even) players.
template<int NT = kWG>
__device__ inline void rr_pair(int m, int r, int k, int& a, int& b) {
  auto at = [&](int pos) { return pos == 0 ? 0 : ((pos - 1 + r) % (m - 1)) + 1; };
  a = at(k);
  b = at(m - 1 - k);
  if (a > b) { const int t = a; a = b; b = t; }
}

// Parallel cyclic Jacobi eigen-decomposition of the symmetric n x n matrix A
// (leading dimension n). On return V holds the eigenvectors as columns and
// w the eigenvalues, ascending (the dsyevd contract). cs/sn/pa/pb are LDS
// scratch of n/2+1 entries each. Returns false on non-finite input.
template<int NT = kWG>
__device__ inline bool jacobi_eigen(double* A, double* V, int n, double* w, double* cs,
                                    double* sn, int* pa, int* pb, double* red,
                                    double* vtmp /* n x n scratch */) {
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += NT) V[e] = (e % n) == (e / n) ? 1.0 : 0.0;
  double fin = 0.0;
  for (int e = tid; e < n * n; e += NT) fin += isfinite(A[e]) ? 0.0 : 1.0;
  if (wg_sum_n<NT>(fin, red) != 0.0) return false;
  const int m = n + (n & 1);
  const int half = m / 2;
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
        int a, b;
        rr_pair(m, r, tid, a, b);
        double c = 1.0, s = 0.0;
        if (b < n) {
          const double apq = A[a + b * n];
          if (apq != 0.0) {
            const double theta = (A[b + b * n] - A[a + a * n]) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        } else {
          b = -1;
        }
        cs[tid] = c; sn[tid] = s; pa[tid] = a; pb[tid] = b;
      }
      __syncthreads();
      // columns of A and V
      for (int e = tid; e < half * n; e += NT) {
        const int k = int(e / n), i = int(e % n);
        const int a = pa[k], b = pb[k];
        if (b < 0 || sn[k] == 0.0) continue;
        const double c = cs[k], s = sn[k];
        const double xa = A[i + a * n], xb = A[i + b * n];
        A[i + a * n] = c * xa - s * xb;
        A[i + b * n] = s * xa + c * xb;
        const double va = V[i + a * n], vb = V[i + b * n];
        V[i + a * n] = c * va - s * vb;
        V[i + b * n] = s * va + c * vb;
      }
      __syncthreads();
      // rows of A
      for (int e = tid; e < half * n; e += NT) {
        const int k = int(e / n), i = int(e % n);
        const int a = pa[k], b = pb[k];
        if (b < 0 || sn[k] == 0.0) continue;
        const double c = cs[k], s = sn[k];
        const double xa = A[a + i * n], xb = A[b + i * n];
        A[a + i * n] = c * xa - s * xb;
        A[b + i * n] = s * xa + c * xb;
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
__device__ inline bool jacobi_svd(double* W, int m, int n, double* U, double* Vm,
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
