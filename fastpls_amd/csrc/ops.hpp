// Stream-ordered operator products and tall-skinny QR, shared by the
// resident model handles and the audited rSVD. Every launch goes to the
// owner's stream; the owner creates the stream, binds the buffers listed by
// ops_bufs() to it and releases them before destroying it.
#pragma once

#include "host_util.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <stdexcept>
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
      const int splits = invariant_ ? xw_splits_by_depth(k) : xw_splits(m, k, nb);
      XwTask<U> t{};
      t.A = A; t.m = m; t.k = k; t.center = center; t.W = W + (long long)c0 * ldw; t.ldw = ldw;
      t.N = nb; t.splits = splits; t.frag = frag(xw_frag_bytes<U>(k, nb));
      if (splits == 1) {
        t.out = out + (long long)c0 * m; t.ldo = m; t.slab = 0;
        launch_xw_single<U>(t, stream_);
      } else {
        part_.reserve(sizeof(double) * (size_t)splits * m * nb);
        t.out = part_.d(); t.ldo = m; t.slab = m * nb;
        launch_xw_single<U>(t, stream_);
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
      const int splits = xtw_splits(m, k, nb);
      XtwTask<U> t{};
      t.A = A; t.m = m; t.k = k; t.center = center; t.W = W + (long long)c0 * ldw; t.ldw = ldw;
      t.N = nb; t.mode = center ? kCentered : kPlain; t.splits = splits;
      t.frag = frag(xtw_frag_bytes<U>(m, nb));
      if (splits == 1) {
        t.out = out + (long long)c0 * k; t.ldo = k; t.slab = 0;
        launch_xtw_single<U>(t, stream_);
      } else {
        part_.reserve(sizeof(double) * (size_t)splits * k * nb);
        t.out = part_.d(); t.ldo = k; t.slab = k * nb;
        launch_xtw_single<U>(t, stream_);
        launch_sum_slabs(part_.d(), splits, k * nb, k * nb, out + (long long)c0 * k, stream_);
      }
    }
  }

  // Tall-skinny QR: CholQR2 in FP64 (two Gram / Cholesky / triangular-apply
  // rounds through the multi-work-group products), with a device-side
  // Householder fallback when a Cholesky factor is unusable. A is destroyed
  // only by the fallback; Q receives the orthonormal basis.
  void qr_tall(double* A, int m, int w, double* Q, int* flags) {
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

  std::vector<Buf*> ops_bufs() {
    return {&frag_, &part_, &qr_gram_, &qr_rinv_, &qr_tmp_, &qr_fail_, &zero_ints_};
  }

  hipStream_t stream_ = nullptr;
  // batch-invariant query splits: xw splits depend on the depth only
  bool invariant_ = false;
  bool zero_ready_ = false;
  Buf frag_;  // matrix-core operand workspace of the streaming products
  Buf part_, qr_gram_, qr_rinv_, qr_tmp_, qr_fail_, zero_ints_;
};

}  // namespace fpls
