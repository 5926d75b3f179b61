// Per-fit FP64 kernels: one work-group owns one fit (or one small matrix).
//
// These carry the sequential parts of randomized SIMPLS and PLS-SVD: the
// direction refresh (right-Gram randomized range finder, QR, eigen of the
// reduced Gram), the component update (normalisation, loading
// orthogonalisation, rank-one deflation of S and of the right Gram), the
// finalisation of a randomized SVD sketch, the PLS-SVD prefix solves, the LDA
// head and the variance-explained bookkeeping. The bulk passes over X run in
// stream_kernels.hip between these launches.
#include "kernels.hpp"
#include "wg_linalg.hpp"

#include <algorithm>
#include <cfloat>

namespace fpls {

namespace {

constexpr int kMaxW = 512;       // widest sketch / response count handled
constexpr int kMaxPairs = kMaxW / 2 + 1;

struct WgScratch {
  double tau[kMaxW];
  double dots[kMaxW];
  double cs[kMaxPairs];
  double sn[kMaxPairs];
  int pa[kMaxPairs];
  int pb[kMaxPairs];
  double red[kWaves];
  double part[kWG];   // partial sums of the Householder dots
};

// Scratch layout of one SIMPLS fit (doubles).
struct RefreshLayout {
  long long Z, Y2, sample, Qb, Bm, Gm, Ve, vtmp, evals, wvec, proj, ptot, row, total;
  __host__ __device__ RefreshLayout(int p, int q, int w, int A) {
    long long o = 0;
    Z = o; o += (long long)q * w;
    Y2 = o; o += (long long)q * w;
    sample = o; o += (long long)p * w;
    Qb = o; o += (long long)p * w;
    Bm = o; o += (long long)w * q;
    Gm = o; o += (long long)w * w;
    Ve = o; o += (long long)w * w;
    vtmp = o; o += (long long)w * w;
    evals = o; o += w;
    wvec = o; o += p;
    proj = o; o += (A > 0 ? A : 1);
    ptot = o; o += (A > 0 ? A : 1);
    row = o; o += q;
    total = o;
  }
};

__device__ inline void set_flag(int* state, int flag) {
  if (threadIdx.x == 0) atomicOr(&state[2], flag);
}

// dst[i] = src[i] * factor for i < len by one work-group, sixteen elements
// per thread loaded before any is stored (dst may alias src), so the loads of
// a pass are in flight together.
__device__ inline void scale_copy(double* dst, const double* src, long long len, double factor) {
  constexpr int U = 16;
  for (long long i0 = threadIdx.x; i0 < len; i0 += U * kWG) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = i0 + (long long)u * kWG;
      v[u] = src[i < len ? i : i0];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = i0 + (long long)u * kWG;
      if (i < len) dst[i] = v[u] * factor;
    }
  }
}

// C (m x n) = A B for a tall A (m x k) and a small B (k x n): B staged in LDS
// (`stage`, 2 kStageDoubles doubles) and one thread per row of C when n <= 32
// and k n fits, else the generic work-group product.
__device__ inline void tall_times_small(long m, int n, int k, const double* A, long lda,
                                        const double* B, long ldb, double* C, long ldc,
                                        double* stage) {
  constexpr int kCap = 2 * wg::kStageDoubles;
  if (n <= 8 && k * 8 <= kCap) {
    wg::stage_rows<8>(B, ldb, k, n, stage);
    wg::gemm_rows_lds<8>(m, n, k, A, lda, stage, C, ldc);
  } else if (n <= 16 && k * 16 <= kCap) {
    wg::stage_rows<16>(B, ldb, k, n, stage);
    wg::gemm_rows_lds<16>(m, n, k, A, lda, stage, C, ldc);
  } else if (n <= 32 && k * 32 <= kCap) {
    wg::stage_rows<32>(B, ldb, k, n, stage);
    wg::gemm_rows_lds<32>(m, n, k, A, lda, stage, C, ldc);
  } else {
    wg::gemm(int(m), n, k, 1.0, A, lda, false, B, ldb, false, 0.0, C, ldc);
  }
}

// C (m x n) = Aᵀ B for A (k x m), B (k x n) with a long k: on the matrix
// cores when m, n <= 64, one wave per output otherwise.
__device__ inline void tall_tn(int m, int n, long k, const double* A, long lda, const double* B,
                               long ldb, double* C, long ldc, double* stage) {
  if (m <= 64 && n <= 64) wg::gemm_tn_fast(m, n, k, A, lda, B, ldb, C, ldc, stage);
  else wg::gemm_tn_long(m, n, int(k), A, lda, B, ldb, C, ldc);
}

// LDS of the one-work-group tall CholQR2.
struct TallQrLds {
  double stage[2 * wg::kStageDoubles];
  double gm[32 * 32], ri[32 * 32], lc[32 * 32], xb[32];
  int ok;
};

// Orthonormal basis Qb of a tall sample (p x w): CholQR2 by one work-group
// when w <= 32, Householder when a Cholesky factor is unusable or w > 32.
// Returns false on a non-finite column norm (Householder).
__device__ inline bool tall_basis(double* sample, int p, int w, double* Qb, TallQrLds& t,
                                  WgScratch& sh) {
  if (w <= 32 && wg::cholqr2_tall(sample, p, w, Qb, t.stage, t.gm, t.ri, t.lc, t.xb, &t.ok))
    return true;
  return wg::householder_qr(sample, p, p, w, Qb, p, sh.tau, sh.dots, sh.red, sh.part);
}

// Whole CholQR2 of a small A (m x w, m >= w, m·w <= kSmallQr, w <= 32) by one
// work-group: A and the first-pass basis stay in LDS, wave 0 factors each
// Gram. When a Cholesky factor is unusable the Householder QR of the original
// A (kept in LDS) gives Q instead, as in the multi-kernel route.
constexpr int kSmallQr = 4096;

__device__ inline void lds_gram(const double* X, int m, int w, double* G, int ldg) {
  for (int e = threadIdx.x; e < w * w; e += kWG) {
    const int i = e % w, j = e / w;
    const double* xi = X + (long)i * m;
    const double* xj = X + (long)j * m;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int l = 0;
    for (; l + 4 <= m; l += 4) {
      s0 = fma(xi[l], xj[l], s0);
      s1 = fma(xi[l + 1], xj[l + 1], s1);
      s2 = fma(xi[l + 2], xj[l + 2], s2);
      s3 = fma(xi[l + 3], xj[l + 3], s3);
    }
    for (; l < m; ++l) s0 = fma(xi[l], xj[l], s0);
    G[i + j * ldg] = (s0 + s1) + (s2 + s3);
  }
}

// Y (m x w, ld m) = X Rinv for an upper-triangular Rinv (ld ldr).
__device__ inline void lds_apply_upper(const double* X, int m, int w, const double* Rinv,
                                       int ldr, double* Y) {
  for (int e = threadIdx.x; e < m * w; e += kWG) {
    const int l = e % m, c = e / m;
    double acc = 0.0;
    for (int k = 0; k <= c; ++k) acc = fma(X[l + (long)k * m], Rinv[k + c * ldr], acc);
    Y[e] = acc;
  }
}

// Q (m x w, ld m, global) = the orthonormal factor of A (m x w, ld m,
// global, not modified). la, lq: kSmallQr doubles of LDS each; gm, ri, lc:
// 32·32, xb 32 doubles, okflag one int of LDS. Returns false on a non-finite
// column norm (Householder route). Ends with a barrier.
__device__ __attribute__((noinline)) bool small_qr_lds(const double* A, int m, int w, double* Q,
                                                       double* la, double* lq, double* gm,
                                                       double* ri, double* lc, double* xb,
                                                       int* okflag, WgScratch& sh) {
  constexpr int W = 32;
  auto factor = [&](const double* X) __attribute__((always_inline)) -> bool {
    lds_gram(X, m, w, gm, W);
    __syncthreads();
    if (threadIdx.x < 64) {
      const bool good = wg::wave_chol_inv<W>(gm, W, w, ri, W, lc, xb);
      if (threadIdx.x == 0) *okflag = good ? 1 : 0;
    }
    __syncthreads();
    return *okflag != 0;
  };
  for (int e = threadIdx.x; e < m * w; e += kWG) la[e] = A[e];
  __syncthreads();
  if (factor(la)) {
    lds_apply_upper(la, m, w, ri, W, lq);
    __syncthreads();
    if (factor(lq)) {
      lds_apply_upper(lq, m, w, ri, W, Q);
      __syncthreads();
      return true;
    }
  }
  const bool fine = wg::householder_qr<kWG>(la, m, m, w, lq, m, sh.tau, sh.dots, sh.red, sh.part);
  for (int e = threadIdx.x; e < m * w; e += kWG) Q[e] = lq[e];
  __syncthreads();
  return fine;
}

__global__ __launch_bounds__(kWG) void k_cholqr2_small(const double* A, int m, int w, double* Q,
                                                       int* flags) {
  constexpr int W = 32;
  __shared__ double la[kSmallQr], lq[kSmallQr];
  __shared__ double gm[W * W], ri[W * W], lc[W * W], xb[W];
  __shared__ WgScratch sh;
  __shared__ int ok_sh;
  const bool fine = small_qr_lds(A, m, w, Q, la, lq, gm, ri, lc, xb, &ok_sh, sh);
  if (!fine && threadIdx.x == 0) atomicOr(flags, kFlagDecomposition);
}

// ----------------------------------------------------------- SIMPLS refresh
// A sketch as wide as the response count (w == q, so q <= p) spans ℝ^q
// whatever its values: every Z of refresh_directions (simpls.hpp:236-299) is
// square and orthogonal, range(S Z) = range(S), and the directions it returns
// (simpls.hpp:301-366) are the top-b left singular vectors of S. They are
// formed here from the eigen-decomposition of SᵀS computed from the current S
// (not from the deflated right Gram, whose rank-one downdates lose relative
// accuracy as S shrinks): candidate c = S v_c / |S v_c| for the eigenvector
// v_c of the c-th largest eigenvalue. Returns false, having written nothing
// the sketch route needs, when the eigen-decomposition fails or a candidate
// norm is zero or non-finite; the caller then runs the sketch route, which
// reports such data as the reference does. Uniform across the work-group.
__device__ bool square_sketch_refresh(const SimplsFitDev& f, int p, int q, int b, TallQrLds& tq,
                                      WgScratch& sh) {
  double* Gs = tq.gm;  // q x q, then the top-b eigenvectors (q x b)
  double* Vs = tq.ri;
  tall_tn(q, q, p, f.S, p, f.S, p, Gs, q, tq.stage);
  if (!wg::jacobi_eigen(Gs, Vs, q, tq.xb, sh.cs, sh.sn, sh.pa, sh.pb, sh.red, tq.lc)) return false;
  for (int e = threadIdx.x; e < q * b; e += kWG) {
    const int l = e % q, c = e / q;
    Gs[e] = Vs[l + (q - 1 - c) * q];
  }
  __syncthreads();
  tall_times_small(p, b, q, f.S, p, Gs, q, f.cand, p, tq.stage);
  for (int c = 0; c < b; ++c) {
    const double* r = f.cand + (long long)c * p;
    double part = 0.0;
    for (int i = threadIdx.x; i < p; i += kWG) part = fma(r[i], r[i], part);
    const double ss = wg_sum_n<kWG>(part, sh.red);
    if (!isfinite(ss) || !(ss > 0.0)) return false;
    if (threadIdx.x == 0) sh.dots[c] = 1.0 / sqrt(ss);
  }
  __syncthreads();
  for (int c = 0; c < b; ++c) {
    const double inv = sh.dots[c];
    double* r = f.cand + (long long)c * p;
    double* rs = f.cand_scaled + (long long)c * p;
    for (int i = threadIdx.x; i < p; i += kWG) {
      const double v = r[i] * inv;
      r[i] = v;
      rs[i] = v / f.sx[i];
    }
  }
  return true;
}

// simpls.hpp:226-368 (refresh_directions), right-Gram and plain branches.
__global__ __launch_bounds__(kWG) void k_simpls_refresh(const SimplsFitDev* fits, int count,
                                                        int step) {
  const SimplsFitDev f = load_const(fits, blockIdx.x);
  if (step >= f.nsteps) return;
  if (f.state[1] != 0) return;
  const RefreshStep st = f.steps[step];
  const int p = f.p, q = f.q, w = st.w, b = st.b;
  __shared__ WgScratch sh;
  __shared__ TallQrLds tq;
  const RefreshLayout L(p, q, w, f.A);
  double* base = f.scratch;
  double* Z = base + L.Z;
  double* Y2 = base + L.Y2;
  double* sample = base + L.sample;
  double* Qb = base + L.Qb;
  double* Bm = base + L.Bm;
  double* Gm = base + L.Gm;
  double* Ve = base + L.Ve;
  double* vtmp = base + L.vtmp;
  double* ev = base + L.evals;
  const double* omega = f.omega + st.omega_off;
  bool ok = true;
  if (w == q && q <= 32 && square_sketch_refresh(f, p, q, b, tq, sh)) return;

  // right-hand (q x w) QRs: whole CholQR2 in LDS when small, else Householder
  auto qside_qr = [&](double* Y, double* Zo) -> bool {
    if (w <= 32 && q * w <= kSmallQr)
      return small_qr_lds(Y, q, w, Zo, tq.stage, tq.stage + kSmallQr, tq.gm, tq.ri, tq.lc, tq.xb,
                          &tq.ok, sh);
    return wg::householder_qr(Y, q, q, w, Zo, q, sh.tau, sh.dots, sh.red, sh.part);
  };
  if (f.use_right_gram) {
    wg::copy(q, w, omega, q, Z, q);
    for (int it = 0; it < f.power; ++it) {
      wg::gemm(q, w, q, 1.0, f.G, q, false, Z, q, false, 0.0, Y2, q);
      ok = qside_qr(Y2, Z) && ok;
    }
    if (f.power == 0) {
      wg::copy(q, w, Z, q, Y2, q);
      ok = qside_qr(Y2, Z) && ok;
    }
    tall_times_small(p, w, q, f.S, p, Z, q, sample, p, tq.stage);
  } else {
    tall_times_small(p, w, q, f.S, p, omega, q, sample, p, tq.stage);
    for (int it = 0; it < f.power; ++it) {
      ok = tall_basis(sample, p, w, Qb, tq, sh) && ok;
      tall_tn(q, w, p, f.S, p, Qb, p, Y2, q, tq.stage);
      ok = qside_qr(Y2, Z) && ok;
      tall_times_small(p, w, q, f.S, p, Z, q, sample, p, tq.stage);
    }
  }
  ok = tall_basis(sample, p, w, Qb, tq, sh) && ok;
  tall_tn(w, q, p, Qb, p, f.S, p, Bm, w, tq.stage);
  // eigen of the projected Gram, in LDS when it fits
  const bool eig_lds = 3 * w * w + w <= 2 * wg::kStageDoubles;
  double* gE = eig_lds ? tq.stage : Gm;
  double* vE = eig_lds ? tq.stage + w * w : Ve;
  double* tE = eig_lds ? tq.stage + 2 * w * w : vtmp;
  double* eE = eig_lds ? tq.stage + 3 * w * w : ev;
  wg::gemm(w, w, q, 1.0, Bm, w, false, Bm, w, true, 0.0, gE, w);
  ok = wg::jacobi_eigen(gE, vE, w, eE, sh.cs, sh.sn, sh.pa, sh.pb, sh.red, tE) && ok;
  // candidates = Qb * (top-b eigenvectors, largest first)
  for (int c = 0; c < b; ++c) {
    const double* v = vE + (long long)(w - 1 - c) * w;
    for (int i = threadIdx.x; i < p; i += kWG) {
      double acc = 0.0;
#pragma unroll 8
      for (int l = 0; l < w; ++l) acc = fma(Qb[i + (long long)l * p], v[l], acc);
      f.cand[i + (long long)c * p] = acc;
      f.cand_scaled[i + (long long)c * p] = acc / f.sx[i];
    }
  }
  if (!ok) {
    set_flag(f.state, kFlagDecomposition);
    if (threadIdx.x == 0) f.state[1] = 1;
  }
}

// Streaming helpers of the update: one work-group walks p- or n-long vectors,
// so each thread keeps several independent loads in flight (indices past the
// end are clamped to a valid element and their values discarded).
constexpr int kStreamU = 32;

// Σ v[i]² over n doubles; every thread receives it.
__device__ inline double sumsq_stream(const double* v, long long n, double* red) {
  double s[kStreamU];
#pragma unroll
  for (int u = 0; u < kStreamU; ++u) s[u] = 0.0;
  for (long long i0 = threadIdx.x; i0 < n; i0 += kStreamU * kWG) {
    double x[kStreamU];
#pragma unroll
    for (int u = 0; u < kStreamU; ++u) {
      const long long i = i0 + (long long)u * kWG;
      x[u] = v[i < n ? i : i0];
    }
#pragma unroll
    for (int u = 0; u < kStreamU; ++u)
      if (i0 + (long long)u * kWG < n) s[u] = fma(x[u], x[u], s[u]);
  }
  double tot = 0.0;
#pragma unroll
  for (int u = 0; u < kStreamU; ++u) tot += s[u];
  return wg_sum_n<kWG>(tot, red);
}

// One pass over the rows of the loading v (p, in place): v -= V[:, :comp] proj
// (nothing when comp is 0; the sum over the earlier loadings in index order),
// returning this thread's partial Σ v² of the result when `norm` is set. Four
// rows per thread, sixteen loadings at a time.
__device__ __forceinline__ double orth_pass(double* v, int p, const double* V, const double* proj,
                                            int comp, bool norm) {
  constexpr int L = 16, R = 4;
  double ss = 0.0;
  for (int i0 = threadIdx.x; i0 < p; i0 += R * kWG) {
    int ir[R];
    double x[R], s[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      ir[r] = i0 + r * kWG < p ? i0 + r * kWG : i0;
      x[r] = v[ir[r]];
      s[r] = 0.0;
    }
    for (int l0 = 0; l0 < comp; l0 += L) {
      double a[R][L], pl[L];
#pragma unroll
      for (int u = 0; u < L; ++u) {
        const int l = l0 + u < comp ? l0 + u : comp - 1;
        pl[u] = l0 + u < comp ? proj[l] : 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) a[r][u] = V[ir[r] + (long long)l * p];
      }
#pragma unroll
      for (int u = 0; u < L; ++u)
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = fma(a[r][u], pl[u], s[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool in = i0 + r * kWG < p;
      if (comp > 0) {
        x[r] -= s[r];
        if (in) v[ir[r]] = x[r];
      }
      if (norm && in) ss = fma(x[r], x[r], ss);
    }
  }
  return ss;
}

// S (p x q, in place) -= (v vs) (row rs)ᵀ element-wise: sixteen entries per
// thread loaded (with their v and row values) before any is stored.
__device__ __forceinline__ void deflate_elementwise(double* S, int p, int q, const double* v,
                                                   double vs, const double* row) {
  constexpr int U = 16;
  const long long len = (long long)p * q;
  for (long long e0 = threadIdx.x; e0 < len; e0 += U * kWG) {
    double sv[U], vv[U], rr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      long long e = e0 + (long long)u * kWG;
      if (e >= len) e = e0;
      const int c = int(e / p);
      const int i = int(e - (long long)c * p);
      sv[u] = S[e];
      vv[u] = v[i];
      rr[u] = row[c];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long e = e0 + (long long)u * kWG;
      if (e < len) S[e] = sv[u] - (vv[u] * vs) * rr[u];
    }
  }
}

// ------------------------------------------------------------ SIMPLS update
// simpls.hpp:588-756: per candidate of the block, in order.
__global__ __launch_bounds__(kWG) void k_simpls_update(const SimplsFitDev* fits, int count,
                                                       int step) {
  const SimplsFitDev f = load_const(fits, blockIdx.x);
  if (step >= f.nsteps) return;
  if (f.state[1] != 0) return;
  const RefreshStep st = f.steps[step];
  const int n = f.n, p = f.p, q = f.q;
  __shared__ WgScratch sh;
  __shared__ double red[4 * 64];
  const RefreshLayout L(p, q, st.w, f.A);
  double* proj = f.scratch + L.proj;
  double* row = f.scratch + L.row;
  // out (m) = Xᵀ v for X (p x m): one thread per row of X, 32 columns of X
  // per pass (more partial sums per thread spill registers) when m <= 128
  // (len rows: p, or n for the scores)
  auto dots_len = [&](int m, int len, const double* X, const double* v, double* out) {
    if (m <= 16) {
      wg::gemm_tv_rows<16>(m, len, X, len, v, out, red);
    } else if (m <= 128) {
      for (int c0 = 0; c0 < m; c0 += 32)
        wg::gemm_tv_rows<32>(m - c0 < 32 ? m - c0 : 32, len, X + (long long)c0 * len, len, v,
                             out + c0, red);
    } else {
      wg::gemm_tn_long(m, 1, len, X, len, v, len, out, m);
    }
  };
  auto dots = [&](int m, const double* X, const double* v, double* out) {
    dots_len(m, p, X, v, out);
  };
  auto stop = [&](double x) {
    set_flag(f.state, isfinite(x) ? kFlagZeroNorm : kFlagNonFinite);
    if (threadIdx.x == 0) f.state[1] = 1;
  };

  for (int j = 0; j < st.b; ++j) {
    const int comp = st.a + j;
    if (comp >= f.A) break;
    double* t = f.t_raw + (long long)j * n;
    double* Rc = f.R + (long long)comp * p;
    double* Vc = f.V + (long long)comp * p;
    double* Qc = f.Q + (long long)comp * q;
    double inv = 0.0;  // 1 / |t|
    if (f.xgram) {
      // cached route (simpls.hpp:600-620): loading w = G r, |t|² = rᵀ w,
      // r and w divided by |t|
      const double* r = f.cand + (long long)j * p;
      for (int i = threadIdx.x; i < p; i += kWG) {
        double s0 = 0.0, s1 = 0.0;
        int l = 0;
        for (; l + 2 <= p; l += 2) {
          s0 = fma(f.xgram[i + (long long)l * p], r[l], s0);
          s1 = fma(f.xgram[i + (long long)(l + 1) * p], r[l + 1], s1);
        }
        if (l < p) s0 = fma(f.xgram[i + (long long)l * p], r[l], s0);
        Vc[i] = s0 + s1;
      }
      __syncthreads();
      double part = 0.0;
      for (int i = threadIdx.x; i < p; i += kWG) part = fma(r[i], Vc[i], part);
      const double ns = wg_sum_n<kWG>(part, sh.red);
      if (!(ns > 0.0) || !isfinite(ns)) return stop(ns);
      inv = 1.0 / sqrt(ns);
      if (f.T) scale_copy(f.T + (long long)comp * n, t, n, inv);
      scale_copy(Rc, r, p, inv);
      scale_copy(Vc, Vc, p, inv);
      __syncthreads();
    } else {
      // re-orthogonalising (simpls.hpp:161-200, 635-642): twice, the
      // projection of t on the earlier scores T is removed from t, and the
      // same combination of the earlier weights R from r; the loading
      // X_sᵀ t then is the block's loading minus P (the earlier loadings
      // X_sᵀ T) times the summed projections
      double* cand = f.cand + (long long)j * p;
      const bool rsb = f.reorth && comp > 0 && f.T != nullptr && f.P != nullptr;
      double* ptot = f.scratch + L.ptot;
      if (rsb) {
        for (int pass = 0; pass < 2; ++pass) {
          dots_len(comp, n, f.T, t, proj);
          __syncthreads();
          orth_pass(t, n, f.T, proj, comp, false);
          orth_pass(cand, p, f.R, proj, comp, false);
          for (int c = threadIdx.x; c < comp; c += kWG) ptot[c] = pass ? ptot[c] + proj[c] : proj[c];
          __syncthreads();
        }
      }
      const double tn = sqrt(fmax(sumsq_stream(t, n, sh.red), 0.0));
      if (!(tn > 0.0) || !isfinite(tn)) return stop(tn);
      inv = 1.0 / tn;
      if (f.T) scale_copy(f.T + (long long)comp * n, t, n, inv);
      // one pass over the rows: R = r / |t| and loading = (sum of the
      // row-split partials, in split order) / sx / |t| (the division by |t|
      // follows the correction when re-orthogonalising)
      const double lscale = rsb ? 1.0 : inv;
      const double* part = f.ls_part + (long long)j * p;
      const int splits = f.ls_splits;
      for (int i0 = threadIdx.x; i0 < p; i0 += 4 * kWG) {
        int ic[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) ic[u] = i0 + u * kWG < p ? i0 + u * kWG : i0;
        double acc[4] = {0.0, 0.0, 0.0, 0.0}, r[4], sx[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          r[u] = cand[ic[u]];
          sx[u] = f.sx[ic[u]];
        }
        for (int sp0 = 0; sp0 < splits; sp0 += 8) {
          double v[8][4];
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const long long off = (long long)(sp0 + e < splits ? sp0 + e : splits - 1) * f.ls_slab;
#pragma unroll
            for (int u = 0; u < 4; ++u) v[e][u] = part[ic[u] + off];
          }
#pragma unroll
          for (int e = 0; e < 8; ++e)
            if (sp0 + e < splits)
#pragma unroll
              for (int u = 0; u < 4; ++u) acc[u] += v[e][u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (i0 + u * kWG < p) {
            Vc[i0 + u * kWG] = acc[u] / sx[u] * lscale;
            Rc[i0 + u * kWG] = r[u] * inv;
          }
      }
      __syncthreads();
      if (rsb) {
        orth_pass(Vc, p, f.P, ptot, comp, false);
        __syncthreads();
        scale_copy(Vc, Vc, p, inv);
        __syncthreads();
      }
      if (f.P) {
        scale_copy(f.P + (long long)comp * p, Vc, p, 1.0);
        __syncthreads();
      }
    }
    // response loadings q_a = S0ᵀ r (= Ỹᵀ t on the matrix-free route)
    if (f.qraw) {
      // Ỹᵀ t; with the earlier scores projected out, minus Q (= Ỹᵀ T) times
      // the summed projections
      const bool rsb = !f.xgram && f.reorth && comp > 0 && f.T != nullptr && f.P != nullptr;
      const double* ptot = f.scratch + L.ptot;
      for (int c = threadIdx.x; c < q; c += kWG) {
        double v = f.qraw[c + (long long)j * q];
        if (rsb)
          for (int l = 0; l < comp; ++l) v -= f.Q[c + (long long)l * q] * ptot[l];
        Qc[c] = v * inv;
      }
      __syncthreads();
    } else {
      dots(q, f.S0, Rc, Qc);
    }
    // v = loading - V (Vᵀ loading), once (twice when re-orthogonalising); the
    // last pass also sums |v|²
    const int passes = comp > 0 && f.reorth ? 2 : 1;
    double ss = 0.0;
    for (int pass = 0; pass < passes; ++pass) {
      if (comp > 0) dots(comp, f.V, Vc, proj);
      const bool last = pass == passes - 1;
      ss = orth_pass(Vc, p, f.V, proj, comp, last);
      if (!last) __syncthreads();
    }
    const double vn = sqrt(fmax(wg_sum_n<kWG>(ss, sh.red), 0.0));
    if (!(vn > 0.0) || !isfinite(vn)) return stop(vn);
    const double vinv = 1.0 / vn;
    // deflation row = vᵀ S ; S -= v row (the matrix-free route keeps v in V
    // and projects it out of every product instead); the row is formed from
    // the unit v as (Sᵀ v_raw) / |v_raw|
    if (f.S) {
      dots(q, f.S, Vc, row);
      if (threadIdx.x < q) row[threadIdx.x] *= vinv;
      __syncthreads();
      deflate_elementwise(f.S, p, q, Vc, vinv, row);
      __syncthreads();
    }
    scale_copy(Vc, Vc, p, vinv);
    if (f.S && f.use_right_gram) {
      for (int e = threadIdx.x; e < q * q; e += kWG) {
        const int i = e % q, c = e / q;
        f.G[e] -= row[i] * row[c];
      }
      __syncthreads();
      for (int e = threadIdx.x; e < q * q; e += kWG) {
        const int i = e % q, c = e / q;
        if (i < c) {
          const double v = 0.5 * (f.G[i + c * q] + f.G[c + i * q]);
          f.G[i + c * q] = v;
          f.G[c + i * q] = v;
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) f.state[0] = comp + 1;
    __syncthreads();
  }
}

// ------------------------------------------- SIMPLS steps of wide fits
// The refresh and update of one SIMPLS fit whose p is large, as several
// launches whose products run through the multi-work-group xw / xtw passes
// (one work-group walking p-long vectors streams at tens of GB/s). These
// kernels are the steps between those products; they read the fit's scalars
// (|t|², |v|²) from device memory, so the host never waits, and do nothing
// once the fit has stopped (state[1]).

// Top-b eigenvectors (largest first) of Bm Bmᵀ for Bm (w x q): the eigen step
// of refresh_directions (simpls.hpp:322-368). With `gram` set, Bm is already
// the symmetric w x w matrix to decompose. vtop: w x b. scratch: 3 w² + w
// doubles (used when the problem does not fit in LDS).
__global__ __launch_bounds__(kWG) void k_refresh_eigen(const double* Bm, int w, int q, int b,
                                                       double* vtop, double* scratch,
                                                       int* state, int gram) {
  __shared__ WgScratch sh;
  __shared__ double stage[2 * wg::kStageDoubles];
  if (state[1] != 0) return;
  const bool lds = 3 * w * w + w <= 2 * wg::kStageDoubles;
  double* base = lds ? stage : scratch;
  double* gE = base;
  double* vE = base + w * w;
  double* tE = base + 2 * w * w;
  double* eE = base + 3 * w * w;
  if (gram) wg::copy(w, w, Bm, w, gE, w);
  else wg::gemm(w, w, q, 1.0, Bm, w, false, Bm, w, true, 0.0, gE, w);
  const bool ok = wg::jacobi_eigen(gE, vE, w, eE, sh.cs, sh.sn, sh.pa, sh.pb, sh.red, tE);
  for (int e = threadIdx.x; e < w * b; e += kWG) {
    const int i = e % w, c = e / w;
    vtop[e] = vE[i + (long long)(w - 1 - c) * w];
  }
  if (!ok) {
    set_flag(state, kFlagDecomposition);
    if (threadIdx.x == 0) state[1] = 1;
  }
}

// A QR of the refresh flagged a decomposition failure: the fit stops.
__global__ void k_refresh_gate(int* state) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && (state[2] & kFlagDecomposition) != 0) state[1] = 1;
}

// Component `comp` from its raw score t (n), candidate r (p) and loading
// split partials: with inv = 1 / |t| (|t|² at *ss), T[:, comp] = t inv,
// R[:, comp] = r inv, v = (sum of the splits) / sx · inv (simpls.hpp:600-660).
// A zero or non-finite |t| stops the fit (k_simpls_update's rule).
__global__ __launch_bounds__(kWG) void k_wide_component(long long n, int p, const double* t,
                                                        const double* r, const double* part,
                                                        int splits, long long slab,
                                                        const double* sx, const double* ss,
                                                        double* Tc, double* Rc, double* Vc,
                                                        int* state) {
  if (state[1] != 0) return;
  const double tn = sqrt(*ss);
  if (!(tn > 0.0) || !isfinite(tn)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      atomicOr(&state[2], isfinite(tn) ? kFlagZeroNorm : kFlagNonFinite);
      state[1] = 1;
    }
    return;
  }
  const double inv = 1.0 / tn;
  const long long stride = (long long)gridDim.x * kWG;
  for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < n; i += stride)
    if (Tc) Tc[i] = t[i] * inv;
  for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < p; i += stride) {
    double acc = 0.0;
    for (int sp = 0; sp < splits; ++sp) acc += part[i + (long long)sp * slab];
    Rc[i] = r[i] * inv;
    Vc[i] = acc / sx[i] * inv;
  }
}

// v -= d (the projection of v on the earlier loadings).
__global__ __launch_bounds__(kWG) void k_wide_subtract(double* v, const double* d, int p,
                                                       const int* state) {
  if (state[1] != 0) return;
  for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < p;
       i += (long long)gridDim.x * kWG)
    v[i] -= d[i];
}

// v /= |v| (|v|² at *vv); a zero or non-finite norm stops the fit.
__global__ __launch_bounds__(kWG) void k_wide_unit(double* v, int p, const double* vv,
                                                   int* state) {
  if (state[1] != 0) return;
  const double vn = sqrt(*vv);
  if (!(vn > 0.0) || !isfinite(vn)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      atomicOr(&state[2], isfinite(vn) ? kFlagZeroNorm : kFlagNonFinite);
      state[1] = 1;
    }
    return;
  }
  const double inv = 1.0 / vn;
  for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < p;
       i += (long long)gridDim.x * kWG)
    v[i] *= inv;
}

// S (p x q) -= v rowᵀ.
__global__ __launch_bounds__(kWG) void k_wide_deflate(double* S, int p, int q, const double* v,
                                                      const double* row, const int* state) {
  if (state[1] != 0) return;
  const long long total = (long long)p * q;
  for (long long e = (long long)blockIdx.x * kWG + threadIdx.x; e < total;
       e += (long long)gridDim.x * kWG) {
    const long long c = e / p, i = e - c * p;
    S[e] -= v[i] * row[c];
  }
}

// G (q x q) -= row rowᵀ, kept exactly symmetric, and component comp done.
__global__ __launch_bounds__(kWG) void k_wide_finish(double* G, int q, const double* row,
                                                     int comp, int* state) {
  if (state[1] != 0) return;
  if (G) {
    for (int e = threadIdx.x; e < q * q; e += kWG) {
      const int i = e % q, c = e / q;
      G[e] -= row[i] * row[c];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < q * q; e += kWG) {
      const int i = e % q, c = e / q;
      if (i < c) {
        const double v = 0.5 * (G[i + c * q] + G[c + i * q]);
        G[i + c * q] = v;
        G[c + i * q] = v;
      }
    }
  }
  if (threadIdx.x == 0) state[0] = comp + 1;
}

// --------------------------------------------------------------- tall QR
// Householder QR; a matrix of up to 6144 doubles is factorised in LDS.
template<int NT>
__global__ __launch_bounds__(NT) void k_qr(double* A, int m, int n, double* Q, int* flags) {
  __shared__ WgScratch sh;
  __shared__ double la[6144];
  __shared__ double lq[6144];
  bool ok;
  if ((long long)m * n <= 6144) {
    for (int e = threadIdx.x; e < m * n; e += NT) la[e] = A[e];
    __syncthreads();
    ok = wg::householder_qr<NT>(la, m, m, n, lq, m, sh.tau, sh.dots, sh.red, sh.part);
    for (int e = threadIdx.x; e < m * n; e += NT) Q[e] = lq[e];
  } else {
    ok = wg::householder_qr<NT>(A, m, m, n, Q, m, sh.tau, sh.dots, sh.red, sh.part);
  }
  if (!ok && threadIdx.x == 0) atomicOr(flags, kFlagDecomposition);
}

// ---------------------------------------------------------------- CholQR2
// Cholesky of a w x w Gram (w <= 64) and the inverse of its transposed
// factor, so that Q = A Rinv with R = Lᵀ. A pivot that is non-finite or
// below 1e-13 of the largest diagonal entry marks the factorisation as
// failed (*fail = 1) and leaves Rinv as the identity; the Householder
// fallback then recomputes Q. A factorisation already marked failed is
// skipped. Up to 32 columns one wave factors the Gram in registers
// (wave_chol_inv); wider Grams take the work-group LDS form below.
__global__ __launch_bounds__(64) void k_chol_inv_wave(const double* G, int w, double* Rinv,
                                                      int* fail) {
  constexpr int W = 32;
  __shared__ double lc[W * W], xb[W];
  const bool failed_before = *fail != 0;
  bool ok = false;
  if (!failed_before) ok = wg::wave_chol_inv<W>(G, w, w, Rinv, w, lc, xb);
  if (!ok) {
    if (threadIdx.x == 0) *fail = 1;
    for (int e = threadIdx.x; e < w * w; e += 64) Rinv[e] = (e % w) == (e / w) ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(kWG) void k_chol_inv(const double* G, int w, double* Rinv,
                                                  int* fail) {
  __shared__ double Lm[64 * 65];
  __shared__ double Ri[64 * 65];
  __shared__ double dmax;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  for (int e = tid; e < w * w; e += kWG) Lm[(e % w) + (e / w) * 65] = G[e];
  const bool failed_before = *fail != 0;
  __syncthreads();
  if (tid < 64) {
    double d = tid < w ? Lm[tid + tid * 65] : 0.0;
    for (int off = 32; off > 0; off >>= 1) d = fmax(d, __shfl_xor(d, off, 64));
    if (tid == 0) dmax = d;
  }
  __syncthreads();
  // Right-looking Cholesky on the unscaled columns: step j subtracts
  // a_ij a_cj / a_jj from the trailing lower triangle; the columns are scaled
  // by 1/sqrt(a_jj) afterwards. Every thread reads the same pivot, so the
  // failure test is uniform and one barrier per step suffices.
  bool bad = failed_before;
  for (int j = 0; j < w && !bad; ++j) {
    const double v = Lm[j + j * 65];
    if (!isfinite(v) || !(v > 1e-13 * dmax)) {
      bad = true;
      break;
    }
    const double inv = 1.0 / v;
    for (int i = j + 1 + ty; i < w; i += 16) {
      const double lij = Lm[i + j * 65] * inv;
      for (int c = j + 1 + tx; c <= i; c += 16) Lm[i + c * 65] -= lij * Lm[c + j * 65];
    }
    __syncthreads();
  }
  if (bad) {
    if (tid == 0) *fail = 1;
    for (int e = tid; e < w * w; e += kWG) Rinv[e] = (e % w) == (e / w) ? 1.0 : 0.0;
    return;
  }
  for (int c = ty; c < w; c += 16) {
    const double d = sqrt(Lm[c + c * 65]);
    for (int i = c + 1 + tx; i < w; i += 16) Lm[i + c * 65] /= d;
  }
  __syncthreads();
  for (int c = tid; c < w; c += kWG) Lm[c + c * 65] = sqrt(Lm[c + c * 65]);
  // Rinv = L^{-T} (upper triangular) by a sweep over the rows from the
  // bottom: at step k row k of X holds Lᵀ-eliminated values; rows i < k take
  // off l_ki x_kc / l_kk. Rows are divided by their pivots at the end.
  for (int e = tid; e < w * w; e += kWG) {
    const int i = e % w, c = e / w;
    Ri[i + c * 65] = i == c ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int k = w - 1; k > 0; --k) {
    const double inv = 1.0 / Lm[k + k * 65];
    for (int c = k + tx; c < w; c += 16) {
      const double xk = Ri[k + c * 65] * inv;
      for (int i = ty; i < k; i += 16) Ri[i + c * 65] -= Lm[k + i * 65] * xk;
    }
    __syncthreads();
  }
  for (int e = tid; e < w * w; e += kWG) {
    const int i = e % w, c = e / w;
    Rinv[e] = i <= c ? Ri[i + c * 65] / Lm[i + i * 65] : 0.0;
  }
}

// Householder recomputation of Q when CholQR2 failed; a no-op otherwise.
__global__ __launch_bounds__(kWG) void k_qr_fallback(double* A, int m, int n, double* Q,
                                                     const int* fail, int* flags) {
  if (*fail == 0) return;
  __shared__ WgScratch sh;
  const bool ok = wg::householder_qr(A, m, m, n, Q, m, sh.tau, sh.dots, sh.red, sh.part);
  if (!ok && threadIdx.x == 0) atomicOr(flags, kFlagDecomposition);
}

// ----------------------------------------------------- rSVD finalisation
// rsvd.hpp:85-165: eigen route on P Pᵀ with the rank tolerance, else dense SVD.
// The Gram of a sketch up to 64 wide stays in LDS.
template<int NT>
__global__ __launch_bounds__(NT) void k_rsvd_finalize(const double* P, int w, int q, int k,
                                                       double* Up, double* sigma, double* Vt,
                                                       int* usable_out, double* scratch,
                                                       int* flags) {
  __shared__ WgScratch sh;
  __shared__ int usable_sh;
  // the eigen route of a sketch up to 64 wide runs in LDS
  __shared__ double lg[64 * 64], lv[64 * 64], lt[64 * 64], lev[64];
  __shared__ double lp[4096];
  if (w * q <= 4096) {
    for (int e = threadIdx.x; e < w * q; e += NT) lp[e] = P[e];
    __syncthreads();
    P = lp;
  }
  const bool in_lds = w <= 64;
  double* Gm = in_lds ? lg : scratch;
  double* Ve = in_lds ? lv : scratch + (long long)w * w;
  double* vtmp = in_lds ? lt : scratch + 2LL * w * w;
  double* ev = in_lds ? lev : scratch + 3LL * w * w;
  double* Wc = scratch + 3LL * w * w + w + q;  // w x q copy (dense route)
  double* Uc = Wc + (long long)w * q;       // w x q or q x w
  double* Vm = Uc + (long long)w * q;       // up to max(w,q)^2
  double* sv = Vm + (long long)(w > q ? w : q) * (w > q ? w : q);
  int usable = 0;
  if (w <= q) {
    wg::gemm<NT>(w, w, q, 1.0, P, w, false, P, w, true, 0.0, Gm, w);
    if (wg::jacobi_eigen<NT>(Gm, Ve, w, ev, sh.cs, sh.sn, sh.pa, sh.pb, sh.red, vtmp)) {
      const double largest = fmax(ev[w - 1], 1.0);
      const double tol = DBL_EPSILON * double(w > q ? w : q) * largest;
      if (threadIdx.x == 0) {
        int u = 0;
        for (int idx = w; idx > 0; --idx) {
          if (ev[idx - 1] <= tol || u == k) break;
          ++u;
        }
        usable_sh = u;
      }
      __syncthreads();
      usable = usable_sh;
      if (usable > 0) {
        for (int e = threadIdx.x; e < w * usable; e += NT) {
          const int i = e % w, c = e / w;
          Up[i + c * w] = Ve[i + (w - 1 - c) * w];
        }
        for (int c = threadIdx.x; c < usable; c += NT) sigma[c] = sqrt(fmax(ev[w - 1 - c], 0.0));
        __syncthreads();
        for (int e = threadIdx.x; e < usable * q; e += NT) {
          const int c = e % usable, j = e / usable;
          double acc = 0.0;
          for (int l = 0; l < w; ++l) acc = fma(Up[l + c * w], P[l + j * w], acc);
          Vt[c + j * k] = acc / sigma[c];
        }
        __syncthreads();
        if (threadIdx.x == 0) *usable_out = usable;
        return;
      }
    }
  }
  // dense route: SVD of P (w x q)
  if (w >= q) {
    wg::copy<NT>(w, q, P, w, Wc, w);
    const bool ok = wg::jacobi_svd<NT>(Wc, w, q, Uc, Vm, sv, sh.cs, sh.sn, sh.pa, sh.pb, sh.red, nullptr);
    const int ret = k < q ? k : q;
    for (int e = threadIdx.x; e < w * ret; e += NT) Up[e % w + (e / w) * w] = Uc[e % w + (long long)(e / w) * w];
    for (int c = threadIdx.x; c < ret; c += NT) sigma[c] = sv[c];
    for (int e = threadIdx.x; e < ret * q; e += NT) {
      const int c = e % ret, j = e / ret;
      Vt[c + j * k] = Vm[j + c * q];
    }
    if (!ok && threadIdx.x == 0) atomicOr(flags, kFlagDecomposition);
    __syncthreads();
    if (threadIdx.x == 0) *usable_out = ret;
  } else {
    // wide P: decompose Pᵀ (q x w) = U' s V'ᵀ so P = V' s U'ᵀ
    for (int e = threadIdx.x; e < w * q; e += NT) {
      const int i = e % w, j = e / w;
      Wc[j + i * q] = P[i + j * w];
    }
    __syncthreads();
    const bool ok = wg::jacobi_svd<NT>(Wc, q, w, Uc, Vm, sv, sh.cs, sh.sn, sh.pa, sh.pb, sh.red, nullptr);
    const int ret = k < w ? k : w;
    for (int e = threadIdx.x; e < w * ret; e += NT) {
      const int i = e % w, c = e / w;
      Up[i + c * w] = Vm[i + c * w];
    }
    for (int c = threadIdx.x; c < ret; c += NT) sigma[c] = sv[c];
    for (int e = threadIdx.x; e < ret * q; e += NT) {
      const int c = e % ret, j = e / ret;
      Vt[c + j * k] = Uc[j + c * q];
    }
    if (!ok && threadIdx.x == 0) atomicOr(flags, kFlagDecomposition);
    __syncthreads();
    if (threadIdx.x == 0) *usable_out = ret;
  }
}

// ------------------------------------------------ operator finalisation
// Eigen route of the operator finalisation (operator_rsvd.hpp:64-168): the
// projected Gram G = P Pᵀ (w x w) of P = Qᵀ A (w x cols) is decomposed, the
// rank tolerance is (eps max(w, cols))² λmax with no unit floor, and the route
// is taken only when it yields all r directions. Writes the top-r
// eigenvectors (left), the same divided by σ (left_scaled), σ, and
// status[0] = 1 when taken, 0 when the dense route has to run.
__global__ __launch_bounds__(kWG) void k_operator_finalize(const double* gram, int w, int cols,
                                                           int r, double* left,
                                                           double* left_scaled, double* sigma,
                                                           double* scratch, int* status) {
  __shared__ WgScratch sh;
  __shared__ double lg[64 * 64], lv[64 * 64], lt[64 * 64], lev[64];
  __shared__ int usable_sh;
  const int tid = threadIdx.x;
  const bool in_lds = w <= 64;
  double* Gm = in_lds ? lg : scratch;
  double* Ve = in_lds ? lv : scratch + (long)w * w;
  double* vt = in_lds ? lt : scratch + 2L * w * w;
  double* ev = in_lds ? lev : scratch + 3L * w * w;
  for (int e = tid; e < w * w; e += kWG) Gm[e] = gram[e];
  __syncthreads();
  int done = 0;
  if (wg::jacobi_eigen(Gm, Ve, w, ev, sh.cs, sh.sn, sh.pa, sh.pb, sh.red, vt)) {
    if (tid == 0) {
      const double largest = ev[w - 1];
      const double rel = DBL_EPSILON * double(w > cols ? w : cols);
      const double tol = largest > 0.0 ? rel * rel * largest : 0.0;
      int u = 0;
      for (int idx = w; idx > 0; --idx) {
        if (ev[idx - 1] <= tol || u == r) break;
        ++u;
      }
      usable_sh = u;
    }
    __syncthreads();
    if (usable_sh == r) {
      for (int e = tid; e < w * r; e += kWG) {
        const int i = e % w, c = e / w;
        const double v = Ve[i + (w - 1 - c) * w];
        left[e] = v;
        left_scaled[e] = v / sqrt(fmax(ev[w - 1 - c], 0.0));
      }
      for (int c = tid; c < r; c += kWG) sigma[c] = sqrt(fmax(ev[w - 1 - c], 0.0));
      done = 1;
    }
  }
  if (tid == 0) status[0] = done;
}

// Dense SVD of M (m x n, m >= n), first k triplets.
__global__ __launch_bounds__(kWG) void k_small_svd(const double* M, int m, int n, int k,
                                                   double* U, double* sigma, double* V,
                                                   double* scratch, int* flags) {
  __shared__ WgScratch sh;
  double* Wc = scratch;
  double* Uc = Wc + (long long)m * n;
  double* Vm = Uc + (long long)m * n;
  double* sv = Vm + (long long)n * n;
  wg::copy(m, n, M, m, Wc, m);
  const bool ok = wg::jacobi_svd(Wc, m, n, Uc, Vm, sv, sh.cs, sh.sn, sh.pa, sh.pb, sh.red, nullptr);
  for (long long e = threadIdx.x; e < (long long)m * k; e += kWG) U[e] = Uc[e];
  for (long long e = threadIdx.x; e < (long long)n * k; e += kWG) V[e] = Vm[e];
  for (int c = threadIdx.x; c < k; c += kWG) sigma[c] = sv[c];
  if (!ok && threadIdx.x == 0) atomicOr(flags, kFlagDecomposition);
}

// ------------------------------------------------------ PLS-SVD weights
// plssvd.hpp:138-164 with linalg.hpp:57-177: prefix k (work-group k - 1)
// solves G_k L = diag(sigma_1..k) for the leading k x k block of TᵀT by
// Cholesky (linalg.hpp:57-100), or by pivoted elimination (linalg.hpp:104-154)
// when a pivot is not positive or the solution is not finite, and writes
// W_k = L Q_kᵀ (k x q, leading dimension k). Prefixes are independent.
__global__ __launch_bounds__(kWG) void k_plssvd_weights(const double* G, const double* sigma,
                                                        const double* Q, int A, int q,
                                                        double* W, int* flags) {
  __shared__ double F[64 * 65];   // factor (Cholesky or elimination)
  __shared__ double X[64 * 65];   // right-hand side / solution
  __shared__ int ok_sh;
  const int tid = threadIdx.x;
  const int k = blockIdx.x + 1;
  if (A > 64) {
    if (tid == 0 && blockIdx.x == 0) atomicOr(flags, kFlagDecomposition);
    return;
  }
  constexpr int L = 65;
  for (int e = tid; e < k * k; e += kWG) {
    const int i = e % k, j = e / k;
    F[i + j * L] = G[i + j * A];
    X[i + j * L] = i == j ? sigma[j] : 0.0;
  }
  if (tid == 0) ok_sh = 1;
  __syncthreads();
  // Cholesky, right-looking over columns (lower triangle; the update order
  // differs from the row-oriented reference only by rounding)
  for (int j = 0; j < k; ++j) {
    if (tid == 0) {
      const double v = F[j + j * L];
      if (!isfinite(v) || v <= 0.0) ok_sh = 0;
      else F[j + j * L] = sqrt(v);
    }
    __syncthreads();
    if (!ok_sh) break;
    const double djj = F[j + j * L];
    for (int i = j + 1 + tid; i < k; i += kWG) F[i + j * L] /= djj;
    __syncthreads();
    const int r = k - j - 1;
    for (int e = tid; e < r * r; e += kWG) {
      const int i = j + 1 + e % r, c = j + 1 + e / r;
      if (c <= i) F[i + c * L] -= F[i + j * L] * F[c + j * L];
    }
    __syncthreads();
  }
  bool solved = false;
  if (ok_sh) {
    // forward then backward substitution, one column of diag(sigma) per thread
    bool good = true;
    for (int c = tid; c < k; c += kWG) {
      for (int i = c; i < k; ++i) {   // rows above c of L⁻¹ e_c sigma_c are zero
        double v = i == c ? sigma[c] : 0.0;
        for (int l = c; l < i; ++l) v -= F[i + l * L] * X[l + c * L];
        X[i + c * L] = v / F[i + i * L];
      }
      for (int i = k - 1; i >= 0; --i) {
        double v = i >= c ? X[i + c * L] : 0.0;
        for (int l = i + 1; l < k; ++l) v -= F[l + i * L] * X[l + c * L];
        v /= F[i + i * L];
        X[i + c * L] = v;
        if (!isfinite(v)) good = false;
      }
    }
    solved = __syncthreads_and(good ? 1 : 0) != 0;
  }
  if (!solved) {
    // pivoted elimination (linalg.hpp:104-154, pivoted_solve), serial
    if (tid == 0) {
      double scale = 0.0;
      for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) {
          F[i + j * L] = G[i + j * A];
          scale = fmax(scale, fabs(F[i + j * L]));
          X[i + j * L] = i == j ? sigma[j] : 0.0;
        }
      const double tol = DBL_EPSILON * double(k) * fmax(scale, 1.0);
      bool good = true;
      for (int piv = 0; piv < k && good; ++piv) {
        int sel = piv;
        for (int i = piv + 1; i < k; ++i) if (fabs(F[i + piv * L]) > fabs(F[sel + piv * L])) sel = i;
        if (!isfinite(F[sel + piv * L]) || fabs(F[sel + piv * L]) <= tol) { good = false; break; }
        if (sel != piv) {
          for (int c = 0; c < k; ++c) {
            double t = F[piv + c * L]; F[piv + c * L] = F[sel + c * L]; F[sel + c * L] = t;
            t = X[piv + c * L]; X[piv + c * L] = X[sel + c * L]; X[sel + c * L] = t;
          }
        }
        for (int i = piv + 1; i < k; ++i) {
          const double mlt = F[i + piv * L] / F[piv + piv * L];
          F[i + piv * L] = 0.0;
          for (int c = piv + 1; c < k; ++c) F[i + c * L] -= mlt * F[piv + c * L];
          for (int c = 0; c < k; ++c) X[i + c * L] -= mlt * X[piv + c * L];
        }
      }
      for (int c = 0; c < k && good; ++c) {
        for (int i = k - 1; i >= 0; --i) {
          double v = X[i + c * L];
          for (int l = i + 1; l < k; ++l) v -= F[i + l * L] * X[l + c * L];
          v /= F[i + i * L];
          X[i + c * L] = v;
          if (!isfinite(v)) good = false;
        }
      }
      if (!good) atomicOr(flags, kFlagDecomposition);
    }
    __syncthreads();
  }
  // W_k = X (k x k) Q_kᵀ -> k x q, leading dimension k
  double* Wk = W + (long long)(k - 1) * A * q;
  for (int e = tid; e < k * q; e += kWG) {
    const int i = e % k, c = e / k;
    double acc = 0.0;
    for (int l = 0; l < k; ++l) acc = fma(X[i + l * L], Q[c + l * q], acc);
    Wk[i + c * k] = acc;
  }
}

// ------------------------------------------------------------------- LDA
// lda.hpp:86-152 / train_lda_prefixes_from_moments (lda.hpp:236-305).
__global__ __launch_bounds__(kWG) void k_lda_prepare(const double* gram, const double* sums,
                                                     const int* counts, long long n, int a,
                                                     int K, int k, double* linear,
                                                     double* constants, double* ridge_out,
                                                     double* scratch, int* flags) {
  __shared__ double cov[64 * 64];
  __shared__ double Lm[64 * 64];
  __shared__ int ok_sh;
  __shared__ double red[kWaves];
  const int tid = threadIdx.x;
  double* means = scratch;  // K x k
  if (k > 64 || k > a) {
    if (tid == 0) atomicOr(flags, kFlagDecomposition);
    return;
  }
  for (int e = tid; e < K * k; e += kWG) {
    const int c = e % K, j = e / K;
    means[c + j * K] = sums[c + j * K] / double(counts[c]);
  }
  __syncthreads();
  const double den = double(n > K ? n - K : 1);
  for (int e = tid; e < k * k; e += kWG) {
    const int r = e % k, c = e / k;
    double v = gram[r + c * a];
    for (int cl = 0; cl < K; ++cl) v -= double(counts[cl]) * means[cl + r * K] * means[cl + c * K];
    cov[r + c * 64] = v / den;
  }
  __syncthreads();
  double tr = 0.0;
  for (int j = tid; j < k; j += kWG) tr += cov[j + j * 64];
  double scale = wg_sum(tr, red) / double(k);
  if (!isfinite(scale) || scale <= 0.0) scale = 1.0;
  const double grid[6] = {1e-8, 1e-6, 1e-5, 1e-4, 1e-3, 1e-2};
  bool solved = false;
  for (int gi = 0; gi < 6 && !solved; ++gi) {
    const double ridge = grid[gi] * scale;
    if (tid == 0) {
      int good = 1;
      for (int i = 0; i < k && good; ++i) {
        for (int j = 0; j <= i; ++j) {
          double v = cov[i + j * 64];
          if (i == j) v += ridge;
          for (int l = 0; l < j; ++l) v -= Lm[i + l * 64] * Lm[j + l * 64];
          if (i == j) {
            if (!isfinite(v) || v <= 0.0) { good = 0; break; }
            Lm[i + j * 64] = sqrt(v);
          } else {
            const double dg = Lm[j + j * 64];
            if (!isfinite(dg) || dg <= 0.0) { good = 0; break; }
            Lm[i + j * 64] = v / dg;
          }
        }
      }
      ok_sh = good;
    }
    __syncthreads();
    const int factored = ok_sh;
    __syncthreads();  // every thread has read ok_sh before thread 0 rewrites it
    if (!factored) continue;
    // solve for each class (right-hand side = class mean row)
    int good = 1;
    for (int c = tid; c < K; c += kWG) {
      double* x = linear + c;  // row c of linear (K x k), stride K
      for (int i = 0; i < k; ++i) {
        double v = means[c + i * K];
        for (int l = 0; l < i; ++l) v -= Lm[i + l * 64] * x[(long long)l * K];
        x[(long long)i * K] = v / Lm[i + i * 64];
      }
      for (int i = k - 1; i >= 0; --i) {
        double v = x[(long long)i * K];
        for (int l = i + 1; l < k; ++l) v -= Lm[l + i * 64] * x[(long long)l * K];
        v /= Lm[i + i * 64];
        if (!isfinite(v)) good = 0;
        x[(long long)i * K] = v;
      }
    }
    solved = __syncthreads_and(good) != 0;
    if (solved && tid == 0) *ridge_out = ridge;
  }
  if (!solved) {
    if (tid == 0) atomicOr(flags, kFlagDecomposition);
    return;
  }
  for (int c = tid; c < K; c += kWG) {
    const double prior = double(counts[c]) / double(n);
    double prod = 0.0;
    for (int j = 0; j < k; ++j) prod += means[c + j * K] * linear[c + (long long)j * K];
    constants[c] = -0.5 * prod + log(fmax(prior, DBL_MIN));
  }
}

// --------------------------------------------------------------- variance
// cuda_resident_variance.cuh:10-57: loadings, sequential residual loadings,
// per-component sums of squares and the total predictor sum of squares.
__global__ __launch_bounds__(kWG) void k_variance(const double* cross, const double* gram,
                                                  const double* colss, int p, int a,
                                                  double* loadings, double* ss, double* seq) {
  __shared__ double red[kWaves];
  for (int f = threadIdx.x; f < p; f += kWG) {
    for (int j = 0; j < a; ++j) {
      const double den = gram[j + j * a];
      double v = cross[f + (long long)j * p];
      const bool good = isfinite(den) && den > 0.0;
      loadings[f + (long long)j * p] = good ? v / den : 0.0;
      for (int i = 0; i < j; ++i) v -= seq[f + (long long)i * p] * gram[i + j * a];
      seq[f + (long long)j * p] = good ? v / den : 0.0;
    }
  }
  __syncthreads();
  double tot = 0.0;
  for (int f = threadIdx.x; f < p; f += kWG) tot += colss[f];
  tot = wg_sum(tot, red);
  if (threadIdx.x == 0) ss[a] = tot;
  for (int j = 0; j < a; ++j) {
    double g = 0.0;
    for (int f = threadIdx.x; f < p; f += kWG) g = fma(seq[f + (long long)j * p], seq[f + (long long)j * p], g);
    g = wg_sum(g, red);
    if (threadIdx.x == 0) {
      const double gain = g * gram[j + j * a];
      ss[j] = (isfinite(gain) && gain > 0.0) ? gain : 0.0;
    }
  }
}

}  // namespace

long long simpls_scratch_doubles(int p, int q, int wmax) {
  return RefreshLayout(p, q, wmax, p).total;
}

void launch_simpls_refresh(const SimplsFitDev* fits, int count, int step, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_simpls_refresh, dim3(count), dim3(kWG), 0, s, fits, count, step);
  FPLS_LAUNCHED("k_simpls_refresh");
}

void launch_simpls_update(const SimplsFitDev* fits, int count, int step, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_simpls_update, dim3(count), dim3(kWG), 0, s, fits, count, step);
  FPLS_LAUNCHED("k_simpls_update");
}

void launch_refresh_eigen(const double* Bm, int w, int q, int b, double* vtop, double* scratch,
                          int* state, hipStream_t s, bool gram) {
  hipLaunchKernelGGL(k_refresh_eigen, dim3(1), dim3(kWG), 0, s, Bm, w, q, b, vtop, scratch,
                     state, gram ? 1 : 0);
  FPLS_LAUNCHED("k_refresh_eigen");
}

void launch_refresh_gate(int* state, hipStream_t s) {
  hipLaunchKernelGGL(k_refresh_gate, dim3(1), dim3(64), 0, s, state);
  FPLS_LAUNCHED("k_refresh_gate");
}

namespace {
unsigned wide_grid(long long count) {
  return unsigned(std::max<long long>(1, std::min<long long>(1024, (count + kWG - 1) / kWG)));
}
}  // namespace

void launch_wide_component(long long n, int p, const double* t, const double* r,
                           const double* part, int splits, long long slab, const double* sx,
                           const double* ss, double* Tc, double* Rc, double* Vc, int* state,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_wide_component, dim3(wide_grid(std::max<long long>(n, p))), dim3(kWG), 0,
                     s, n, p, t, r, part, splits, slab, sx, ss, Tc, Rc, Vc, state);
  FPLS_LAUNCHED("k_wide_component");
}

void launch_wide_subtract(double* v, const double* d, int p, const int* state, hipStream_t s) {
  hipLaunchKernelGGL(k_wide_subtract, dim3(wide_grid(p)), dim3(kWG), 0, s, v, d, p, state);
  FPLS_LAUNCHED("k_wide_subtract");
}

void launch_wide_unit(double* v, int p, const double* vv, int* state, hipStream_t s) {
  hipLaunchKernelGGL(k_wide_unit, dim3(wide_grid(p)), dim3(kWG), 0, s, v, p, vv, state);
  FPLS_LAUNCHED("k_wide_unit");
}

void launch_wide_deflate(double* S, int p, int q, const double* v, const double* row,
                         const int* state, hipStream_t s) {
  hipLaunchKernelGGL(k_wide_deflate, dim3(wide_grid((long long)p * q)), dim3(kWG), 0, s, S, p, q,
                     v, row, state);
  FPLS_LAUNCHED("k_wide_deflate");
}

void launch_wide_finish(double* G, int q, const double* row, int comp, int* state,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_wide_finish, dim3(1), dim3(kWG), 0, s, G, q, row, comp, state);
  FPLS_LAUNCHED("k_wide_finish");
}

void launch_qr(double* A, int m, int n, double* Q, int* flags, hipStream_t s) {
  if (n > kMaxW || m < n) throw std::invalid_argument("QR shape outside the supported range");
  hipLaunchKernelGGL((k_qr<kWG>), dim3(1), dim3(kWG), 0, s, A, m, n, Q, flags);
  FPLS_LAUNCHED("k_qr");
}

void launch_chol_inv(const double* G, int w, double* Rinv, int* fail, hipStream_t s) {
  if (w < 1 || w > 64) throw std::invalid_argument("CholQR2 handles 1 to 64 columns");
  if (w <= 32) hipLaunchKernelGGL(k_chol_inv_wave, dim3(1), dim3(64), 0, s, G, w, Rinv, fail);
  else hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(kWG), 0, s, G, w, Rinv, fail);
  FPLS_LAUNCHED("k_chol_inv");
}

bool small_qr_fits(int m, int w) { return w >= 1 && w <= 32 && m >= w && (long long)m * w <= kSmallQr; }

void launch_cholqr2_small(const double* A, int m, int w, double* Q, int* flags, hipStream_t s) {
  if (!small_qr_fits(m, w)) throw std::invalid_argument("small CholQR2 shape outside the supported range");
  hipLaunchKernelGGL(k_cholqr2_small, dim3(1), dim3(kWG), 0, s, A, m, w, Q, flags);
  FPLS_LAUNCHED("k_cholqr2_small");
}

void launch_qr_fallback(double* A, int m, int n, double* Q, const int* fail, int* flags,
                        hipStream_t s) {
  if (n > kMaxW || m < n) throw std::invalid_argument("QR shape outside the supported range");
  hipLaunchKernelGGL(k_qr_fallback, dim3(1), dim3(kWG), 0, s, A, m, n, Q, fail, flags);
  FPLS_LAUNCHED("k_qr_fallback");
}

void launch_rsvd_finalize(const double* P, int w, int q, int k, double* Up, double* sigma,
                          double* Vt, int* usable, double* scratch, int* flags,
                          hipStream_t s) {
  if (w > kMaxW || q > kMaxW) throw std::invalid_argument("sketch too wide");
  hipLaunchKernelGGL((k_rsvd_finalize<kWG>), dim3(1), dim3(kWG), 0, s, P, w, q, k, Up, sigma,
                     Vt, usable, scratch, flags);
  FPLS_LAUNCHED("k_rsvd_finalize");
}

void launch_operator_finalize(const double* gram, int w, int cols, int r, double* left,
                              double* left_scaled, double* sigma, double* scratch, int* status,
                              hipStream_t s) {
  if (w < 1 || w > kMaxW || r < 1 || r > w) throw std::invalid_argument("operator finalisation shape outside the supported range");
  hipLaunchKernelGGL(k_operator_finalize, dim3(1), dim3(kWG), 0, s, gram, w, cols, r, left,
                     left_scaled, sigma, scratch, status);
  FPLS_LAUNCHED("k_operator_finalize");
}

void launch_small_svd(const double* M, int m, int n, int k, double* U, double* sigma,
                      double* V, double* scratch, int* flags, hipStream_t s) {
  if (n > kMaxW || m < n) throw std::invalid_argument("small SVD shape outside the supported range");
  hipLaunchKernelGGL(k_small_svd, dim3(1), dim3(kWG), 0, s, M, m, n, k, U, sigma, V, scratch,
                     flags);
  FPLS_LAUNCHED("k_small_svd");
}

void launch_plssvd_weights(const double* G, const double* sigma, const double* Q, int A,
                           int q, double* W, int* flags, hipStream_t s) {
  hipLaunchKernelGGL(k_plssvd_weights, dim3(unsigned(A > 0 ? A : 1)), dim3(kWG), 0, s, G, sigma, Q, A, q, W, flags);
  FPLS_LAUNCHED("k_plssvd_weights");
}

void launch_lda_prepare(const double* gram, const double* class_sums, const int* counts,
                        long long n, int a, int K, int k, double* linear,
                        double* constants, double* ridge, double* scratch, int* flags,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_lda_prepare, dim3(1), dim3(kWG), 0, s, gram, class_sums, counts, n, a,
                     K, k, linear, constants, ridge, scratch, flags);
  FPLS_LAUNCHED("k_lda_prepare");
}

void launch_variance(const double* cross, const double* gram, const double* colss, int p,
                     int a, double* loadings, double* ss, double* scratch, hipStream_t s) {
  hipLaunchKernelGGL(k_variance, dim3(1), dim3(kWG), 0, s, cross, gram, colss, p, a,
                     loadings, ss, scratch);
  FPLS_LAUNCHED("k_variance");
}

}  // namespace fpls
