// Stream-ordered operator products and tall-skinny QR, shared by the
// resident model handles and the audited rSVD. Every launch goes to the
// owner's stream; the owner creates the stream, binds the buffers listed by
// ops_bufs() to it and releases them before destroying it.
#pragma once

#include "host_util.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

namespace fpls {

class StreamOps {
 protected:
  // Operand workspace of at least `bytes` (null when no launch needs one).
  // frag_ is bound to the stream, so the block it replaces when it grows is
  // handed out again only in stream order: no drain is needed.
  void* frag(size_t bytes) {
    if (bytes == 0) return nullptr;
    frag_.reserve(bytes);
    return frag_.as<void>();
  }

  void fill_zero_int(int* ptr, int count) {
    // A kernel (not hipMemsetAsync) so the stream stays graph-capturable.
    launch_fill_int(ptr, count, 0, stream_);
  }

  // --------------------------------------------------------- products
  // out (m x N, ld m) = (A - 1 cᵀ) W for A (m x k) of element U.
  template<class U>
  void product(const U* A, long long m, long long k, const double* center, const double* W,
               long long ldw, int N, double* out) {
    for (int c0 = 0; c0 < N; c0 += 64) {
      const int nb = std::min(64, N - c0);
      const bool f64 = f64_products_;
      const int splits = invariant_ ? xw_splits_by_depth(k)
                                    : xw_splits(m, k, nb, sizeof(U) == 4 && !f64);
      XwTask<U> t{};
      t.A = A; t.m = m; t.k = k; t.center = center; t.W = W + (long long)c0 * ldw; t.ldw = ldw;
      t.N = nb; t.splits = splits; t.frag = frag(xw_frag_bytes<U>(k, nb, f64));
      if (splits == 1) {
        t.out = out + (long long)c0 * m; t.ldo = m; t.slab = 0;
        launch_xw_single<U>(t, stream_, f64);
      } else {
        part_.reserve(sizeof(double) * (size_t)splits * m * nb);
        t.out = part_.d(); t.ldo = m; t.slab = m * nb;
        launch_xw_single<U>(t, stream_, f64);
        launch_sum_slabs(part_.d(), splits, m * nb, m * nb, out + (long long)c0 * m, stream_);
      }
    }
  }

  // out (k x N, ld k) = (A - 1 cᵀ)ᵀ W for A (m x k), W (m x N).
  template<class U>
  void tproduct(const U* A, long long m, long long k, const double* center, const double* W,
                long long ldw, int N, double* out) {
    for (int c0 = 0; c0 < N; c0 += 64) {
      const int nb = std::min(64, N - c0);
      const bool f64 = f64_products_;
      const int splits = xtw_splits(m, k, nb, sizeof(U) == 4 && !f64);
      XtwTask<U> t{};
      t.A = A; t.m = m; t.k = k; t.center = center; t.W = W + (long long)c0 * ldw; t.ldw = ldw;
      t.N = nb; t.mode = center ? kCentered : kPlain; t.splits = splits;
      t.frag = frag(xtw_frag_bytes<U>(m, nb, f64));
      if (splits == 1) {
        t.out = out + (long long)c0 * k; t.ldo = k; t.slab = 0;
        launch_xtw_single<U>(t, stream_, f64);
      } else {
        part_.reserve(sizeof(double) * (size_t)splits * k * nb);
        t.out = part_.d(); t.ldo = k; t.slab = k * nb;
        launch_xtw_single<U>(t, stream_, f64);
        launch_sum_slabs(part_.d(), splits, k * nb, k * nb, out + (long long)c0 * k, stream_);
      }
    }
  }

  // Tall-skinny QR: CholQR2 in FP64 (two Gram / Cholesky / triangular-apply
  // rounds through the multi-work-group products), with a device-side
  // Householder fallback when a Cholesky factor is unusable. A is destroyed
  // only by the fallback; Q receives the orthonormal basis. Small matrices
  // (w <= 32, m·w <= 4096) run the whole CholQR2 in one work-group; other
  // shapes with m < 4w or w > 64 take the Householder QR directly.
  void qr_tall(double* A, int m, int w, double* Q, int* flags) {
    if (small_qr_fits(m, w)) {
      launch_cholqr2_small(A, m, w, Q, flags, stream_);
      return;
    }
    if (w > 64 || m < 4 * w) {
      launch_qr(A, m, w, Q, flags, stream_);
      return;
    }
    qr_gram_.reserve(sizeof(double) * (size_t)w * w);
    qr_rinv_.reserve(sizeof(double) * (size_t)w * w);
    qr_tmp_.reserve(sizeof(double) * (size_t)m * w);
    qr_fail_.reserve(sizeof(int));
    fill_zero_int_async(qr_fail_.as<int>(), 1);
    tproduct<double>(A, m, w, nullptr, A, m, w, qr_gram_.d());
    launch_chol_inv(qr_gram_.d(), w, qr_rinv_.d(), qr_fail_.as<int>(), stream_);
    product<double>(A, m, w, nullptr, qr_rinv_.d(), w, w, qr_tmp_.d());
    tproduct<double>(qr_tmp_.d(), m, w, nullptr, qr_tmp_.d(), m, w, qr_gram_.d());
    launch_chol_inv(qr_gram_.d(), w, qr_rinv_.d(), qr_fail_.as<int>(), stream_);
    product<double>(qr_tmp_.d(), m, w, nullptr, qr_rinv_.d(), w, w, Q);
    launch_qr_fallback(A, m, w, Q, qr_fail_.as<int>(), flags, stream_);
  }

  void fill_zero_int_async(int* ptr, int count) {
    zero_ints_.reserve(sizeof(int) * 64);
    if (count > 64) throw std::invalid_argument("zero fill too large");
    if (!zero_ready_) {
      launch_fill_int(zero_ints_.as<int>(), 64, 0, stream_);
      zero_ready_ = true;
    }
    FPLS_HIP(hipMemcpyAsync(ptr, zero_ints_.as<int>(), sizeof(int) * count,
                            hipMemcpyDeviceToDevice, stream_));
  }

  // ----------------------------------------------------- operator rSVD
  // operator_rsvd.hpp:204-272 (randomized_operator_svd) for an operator A
  // (rows x cols) known through fwd(Z, w, out): out (rows x w) = A Z for Z
  // (cols x w), and adj(B, w, out): out (cols x w) = Aᵀ B for B (rows x w).
  // The Gaussian sketch is the host stream of `seed`. Writes U (rows x r),
  // s (r) and, unless left_only, V (cols x r), the right vectors. The eigen
  // route of the finalisation (squared rank tolerance, all r directions
  // required) runs on the projected Gram; otherwise the dense route factors
  // Pᵀ = Qr R and takes the SVD of R. One host synchronisation decides the
  // route. QR failures set kFlagDecomposition in *flags.
  template<class Fwd, class Adj>
  void operator_rsvd(Fwd&& fwd, Adj&& adj, long long rows, long long cols, int r,
                     int oversample, int power, unsigned int seed, bool left_only, double* U,
                     double* s, double* V, int* flags) {
    const long long maximum = std::min(rows, cols);
    const int w = int(std::min<long long>(maximum, (long long)r + std::max(oversample, 0)));
    if (r < 1 || r > w) throw std::invalid_argument("operator rSVD rank outside the sketch");
    if (w > kOpMaxW)
      throw std::runtime_error("sketch width " + std::to_string(w) +
                               " exceeds the supported 512 columns");
    Buf &omega = opw_[0], &sample = opw_[1], &fb = opw_[2], &rev = opw_[3], &rb = opw_[4],
        &gram = opw_[5], &left = opw_[6], &lscaled = opw_[7], &scratch = opw_[8],
        &status = opw_[9], &rcopy = opw_[10], &rm = opw_[11], &su = opw_[12], &ss = opw_[13],
        &sv = opw_[14];
    // the host sketch of the previous call must have been copied before it
    // is overwritten: calls that took no route decision leave a wait here
    if (op_upload_pending_) FPLS_HIP(hipStreamSynchronize(stream_));
    op_upload_pending_ = true;
    op_omega_host_.resize((size_t)cols * w);
    gaussian(seed, op_omega_host_.size(), op_omega_host_.data());
    omega.reserve(sizeof(double) * op_omega_host_.size());
    FPLS_HIP(hipMemcpyAsync(omega.d(), op_omega_host_.data(),
                            sizeof(double) * op_omega_host_.size(), hipMemcpyHostToDevice,
                            stream_));
    sample.reserve(sizeof(double) * (size_t)rows * w);
    fb.reserve(sizeof(double) * (size_t)rows * w);
    rev.reserve(sizeof(double) * (size_t)cols * w);
    rb.reserve(sizeof(double) * (size_t)cols * w);
    fwd(omega.d(), w, sample.d());
    for (int it = 0; it < power; ++it) {
      qr_tall(sample.d(), int(rows), w, fb.d(), flags);
      adj(fb.d(), w, rev.d());
      qr_tall(rev.d(), int(cols), w, rb.d(), flags);
      fwd(rb.d(), w, sample.d());
    }
    qr_tall(sample.d(), int(rows), w, fb.d(), flags);
    adj(fb.d(), w, rev.d());  // Pᵀ (cols x w)
    bool eigen = false;
    if (w >= 4) {
      gram.reserve(sizeof(double) * (size_t)w * w);
      left.reserve(sizeof(double) * (size_t)w * r);
      lscaled.reserve(sizeof(double) * (size_t)w * r);
      scratch.reserve(sizeof(double) * (3 * (size_t)w * w + w + 64));
      status.reserve(sizeof(int));
      tproduct<double>(rev.d(), cols, w, nullptr, rev.d(), cols, w, gram.d());
      launch_operator_finalize(gram.d(), w, int(cols), r, left.d(), lscaled.d(), s,
                               scratch.d(), status.as<int>(), stream_);
      int st = 0;
      FPLS_HIP(hipMemcpyAsync(&st, status.as<int>(), sizeof(int), hipMemcpyDeviceToHost,
                              stream_));
      FPLS_HIP(hipStreamSynchronize(stream_));
      op_upload_pending_ = false;
      eigen = st == 1;
    }
    if (eigen) {
      product<double>(fb.d(), rows, w, nullptr, left.d(), w, r, U);
      if (!left_only) product<double>(rev.d(), cols, w, nullptr, lscaled.d(), w, r, V);
      return;
    }
    // dense route: Pᵀ = Qr R, R = Ur σ Vrᵀ  =>  P = Vr σ (Qr Ur)ᵀ
    rcopy.reserve(sizeof(double) * (size_t)cols * w);
    rm.reserve(sizeof(double) * (size_t)w * w);
    su.reserve(sizeof(double) * (size_t)w * w);
    ss.reserve(sizeof(double) * (size_t)w);
    sv.reserve(sizeof(double) * (size_t)w * w);
    scratch.reserve(sizeof(double) * (3 * (size_t)w * w + w + 64));
    FPLS_HIP(hipMemcpyAsync(rcopy.d(), rev.d(), sizeof(double) * (size_t)cols * w,
                            hipMemcpyDeviceToDevice, stream_));
    qr_tall(rcopy.d(), int(cols), w, rb.d(), flags);
    tproduct<double>(rb.d(), cols, w, nullptr, rev.d(), cols, w, rm.d());
    launch_small_svd(rm.d(), w, w, w, su.d(), ss.d(), sv.d(), scratch.d(), flags, stream_);
    product<double>(fb.d(), rows, w, nullptr, sv.d(), w, r, U);
    if (!left_only) product<double>(rb.d(), cols, w, nullptr, su.d(), w, r, V);
    FPLS_HIP(hipMemcpyAsync(s, ss.d(), sizeof(double) * r, hipMemcpyDeviceToDevice, stream_));
  }

  std::vector<Buf*> ops_bufs() {
    std::vector<Buf*> v = {&frag_, &part_, &qr_gram_, &qr_rinv_, &qr_tmp_, &qr_fail_,
                           &zero_ints_};
    for (Buf& b : opw_) v.push_back(&b);
    return v;
  }

  hipStream_t stream_ = nullptr;
  // batch-invariant query splits: xw splits depend on the depth only
  bool invariant_ = false;
  // FP32 operands of product / tproduct run on the FP64 matrix cores
  // (widened on load) instead of in bf16x6 split precision
  bool f64_products_ = false;
  bool zero_ready_ = false;
  Buf frag_;  // matrix-core operand workspace of the streaming products
  Buf part_, qr_gram_, qr_rinv_, qr_tmp_, qr_fail_, zero_ints_;
  static constexpr int kOpMaxW = 512;   // widest operator sketch (one-work-group QR / SVD)
  Buf opw_[15];                          // operator rSVD workspaces
  std::vector<double> op_omega_host_;
  bool op_upload_pending_ = false;
};

}  // namespace fpls
