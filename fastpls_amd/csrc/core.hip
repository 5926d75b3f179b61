// FP32 products for a host-resident float core (the reference's hybrid
// accelerator hooks, src/accelerator_core_backend.h:10-36, implemented for
// CUDA in src/core_cuda_backend.cu:38-293 and called from
// RoutedLinearAlgebraF32, src/r_api.cpp:421-660): op(A) op(B) and the
// self-Gram op(A) op(A)ᵀ of host matrices with a leading dimension, through a
// persistent device workspace (one stream, buffers that only grow).
//
// The products run on this library's streaming kernels: the right operand is
// widened to FP64 on the device (transposed when op(B) or the Gram's second
// factor needs it), the left operand stays FP32 and is read in place, and the
// output columns go in chunks of 64 (xw for op(A) = A, xtw for op(A) = Aᵀ).
// Chunks of 8 columns or more run on the matrix cores in bf16x6 split
// precision (FP32-accurate products, FP64 accumulation across steps), narrower
// ones on the FP64 VALU kernels; the FP64 result is rounded once to FP32.
#include "../../include/fastpls_hip.h"
#include "host_util.hpp"
#include "kernels.hpp"
#include "stream.hpp"

#include <algorithm>
#include <limits>
#include <stdexcept>
#include <string>

namespace fpls {

namespace {

// out = in widened to FP64 (rows x cols, ld rows), or its transpose (cols x
// rows, ld cols). 32 x 32 tiles through LDS keep both sides coalesced.
__global__ __launch_bounds__(256) void k_widen(const float* in, long long rows, long long cols,
                                               int transpose, double* out) {
  __shared__ double tile[32][33];
  const long long r0 = (long long)blockIdx.x * 32, c0 = (long long)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  for (int j = ty; j < 32; j += 8) {
    const long long r = r0 + tx, c = c0 + j;
    tile[j][tx] = (r < rows && c < cols) ? double(in[r + c * rows]) : 0.0;
  }
  __syncthreads();
  if (!transpose) {
    for (int j = ty; j < 32; j += 8) {
      const long long r = r0 + tx, c = c0 + j;
      if (r < rows && c < cols) out[r + c * rows] = tile[j][tx];
    }
  } else {
    // out(c, r) with c fastest: thread tx walks c
    for (int j = ty; j < 32; j += 8) {
      const long long c = c0 + tx, r = r0 + j;
      if (r < rows && c < cols) out[c + r * cols] = tile[tx][j];
    }
  }
}

void launch_widen(const float* in, long long rows, long long cols, bool transpose, double* out,
                  hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  const long long gx = (rows + 31) / 32, gy = (cols + 31) / 32;
  if (gx > 0x7fffffffLL || gy > 65535) throw std::invalid_argument("matrix too large to widen");
  hipLaunchKernelGGL(k_widen, dim3(unsigned(gx), unsigned(gy)), dim3(256), 0, s, in, rows, cols,
                     transpose ? 1 : 0, out);
  FPLS_LAUNCHED("k_widen");
}

int checked_int(size_t v, const char* what) {
  if (v > size_t(std::numeric_limits<int>::max()))
    throw std::overflow_error(std::string(what) + " exceeds HIP limits");
  return int(v);
}

// One host matrix view: rows x cols, column-major, leading dimension ld.
struct View {
  const float* data;
  size_t rows, cols, ld;
};

class CoreWorkspace {
 public:
  CoreWorkspace() : stream_(stream_take()) {
    for (Buf* b : {&a_, &b32_, &b64_, &out64_, &out32_, &part_, &frag_}) b->bind(stream_);
  }
  ~CoreWorkspace() {
    for (Buf* b : {&a_, &b32_, &b64_, &out64_, &out32_, &part_, &frag_}) b->release();
    stream_give(stream_);
  }

  // output (rows x cols, ld out_ld) = op(left) op(right)
  void gemm(View L, View R, bool ta, bool tb, float* out, size_t out_rows, size_t out_cols,
            size_t out_ld) {
    const size_t rows = ta ? L.cols : L.rows, inner = ta ? L.rows : L.cols;
    const size_t rinner = tb ? R.cols : R.rows, cols = tb ? R.rows : R.cols;
    if (inner != rinner || out_rows != rows || out_cols != cols)
      throw std::invalid_argument("HIP matrix dimensions are not conformable");
    if (rows == 0 || cols == 0 || inner == 0) return;
    if (!out || out_ld < rows) throw std::invalid_argument("invalid output view");
    upload(L, a_);
    upload(R, b32_);
    // W = op(right) as FP64, inner x cols
    b64_.reserve(sizeof(double) * inner * cols);
    launch_widen(b32_.as<float>(), (long long)R.rows, (long long)R.cols, tb, b64_.d(), stream_);
    product(a_.as<float>(), (long long)L.rows, (long long)L.cols, ta, b64_.d(), (long long)inner,
            (long long)rows, (long long)cols);
    download(out, rows, cols, out_ld);
  }

  // output (dim x dim) = A Aᵀ (transpose_input false) or Aᵀ A, lower triangle;
  // full_output mirrors it to the upper triangle, otherwise the upper triangle
  // is zero.
  void self_gram(View A, bool ta, float* out, size_t out_rows, size_t out_cols, size_t out_ld,
                 bool full) {
    const size_t dim = ta ? A.cols : A.rows, rank = ta ? A.rows : A.cols;
    if (out_rows != dim || out_cols != dim)
      throw std::invalid_argument("HIP float32 self-Gram dimensions are inconsistent");
    if (dim == 0 || rank == 0) return;
    if (!out || out_ld < dim) throw std::invalid_argument("invalid output view");
    upload(A, a_);
    // second factor as FP64, rank x dim: A (rows x cols) for AᵀA, Aᵀ for AAᵀ
    b64_.reserve(sizeof(double) * rank * dim);
    launch_widen(a_.as<float>(), (long long)A.rows, (long long)A.cols, !ta, b64_.d(), stream_);
    product(a_.as<float>(), (long long)A.rows, (long long)A.cols, ta, b64_.d(), (long long)rank,
            (long long)dim, (long long)dim);
    download(out, dim, dim, out_ld);
    for (size_t c = 0; c < dim; ++c)
      for (size_t r = 0; r < c; ++r) out[r + c * out_ld] = full ? out[c + r * out_ld] : 0.0f;
  }

 private:
  // host view -> contiguous device FP32 (rows x cols, ld rows)
  void upload(View v, Buf& dst) {
    if (!v.data || v.ld < v.rows) throw std::invalid_argument("invalid input view");
    dst.reserve(sizeof(float) * v.rows * v.cols);
    FPLS_HIP(hipMemcpy2DAsync(dst.as<float>(), v.rows * sizeof(float), v.data,
                              v.ld * sizeof(float), v.rows * sizeof(float), v.cols,
                              hipMemcpyHostToDevice, stream_));
  }

  // out64 (rows x cols) = op(A) W for device A (m x k FP32) and W (inner x
  // cols FP64), in chunks of 64 output columns
  void product(const float* A, long long m, long long k, bool ta, const double* W,
               long long inner, long long rows, long long cols) {
    out64_.reserve(sizeof(double) * size_t(rows) * size_t(cols));
    for (long long c0 = 0; c0 < cols; c0 += 64) {
      const int nb = int(std::min<long long>(64, cols - c0));
      double* dst = out64_.d() + c0 * rows;
      if (!ta) {
        const int splits = xw_splits(m, k, nb, true);
        frag_.reserve(xw_frag_bytes<float>(k, nb));
        XwTask<float> t{};
        t.A = A; t.m = m; t.k = k; t.center = nullptr; t.W = W + c0 * inner; t.ldw = inner;
        t.N = nb; t.splits = splits; t.frag = frag_.empty() ? nullptr : frag_.as<void>();
        t.ldo = m;
        run_split(t, splits, m * nb, dst);
      } else {
        const int splits = xtw_splits(m, k, nb, true);
        frag_.reserve(xtw_frag_bytes<float>(m, nb));
        XtwTask<float> t{};
        t.A = A; t.m = m; t.k = k; t.center = nullptr; t.mode = kPlain; t.W = W + c0 * inner;
        t.ldw = inner; t.N = nb; t.splits = splits;
        t.frag = frag_.empty() ? nullptr : frag_.as<void>();
        t.ldo = k;
        run_split(t, splits, k * nb, dst);
      }
    }
  }

  template<class Task>
  void run_split(Task& t, int splits, long long count, double* dst) {
    if (splits == 1) {
      t.out = dst; t.slab = 0;
      launch(t);
    } else {
      part_.reserve(sizeof(double) * size_t(splits) * size_t(count));
      t.out = part_.d(); t.slab = count;
      launch(t);
      launch_sum_slabs(part_.d(), splits, count, count, dst, stream_);
    }
  }
  void launch(const XwTask<float>& t) { launch_xw_single<float>(t, stream_); }
  void launch(const XtwTask<float>& t) { launch_xtw_single<float>(t, stream_); }

  void download(float* out, size_t rows, size_t cols, size_t ld) {
    out32_.reserve(sizeof(float) * rows * cols);
    launch_convert<float>(out64_.d(), out32_.as<float>(), (long long)(rows * cols), stream_);
    FPLS_HIP(hipMemcpy2DAsync(out, ld * sizeof(float), out32_.as<float>(), rows * sizeof(float),
                              rows * sizeof(float), cols, hipMemcpyDeviceToHost, stream_));
    FPLS_HIP(hipStreamSynchronize(stream_));
  }

  hipStream_t stream_;
  Buf a_, b32_, b64_, out64_, out32_, part_, frag_;
};

void require_device() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    throw std::runtime_error("HIP is unavailable; no CPU fallback is performed");
}

}  // namespace
}  // namespace fpls

using fpls::CoreWorkspace;
using fpls::View;

extern "C" {

void* fastpls_hip_core_workspace_create_f32(char* error, size_t error_capacity) {
  CoreWorkspace* w = nullptr;
  if (fpls::guarded(error, error_capacity, [&] {
        fpls::require_device();
        w = new CoreWorkspace();
      }) != 0)
    return nullptr;
  return w;
}

void fastpls_hip_core_workspace_destroy_f32(void* workspace) {
  delete static_cast<CoreWorkspace*>(workspace);
}

int fastpls_hip_core_gemm_into_f32(void* workspace, const float* left, size_t left_rows,
                                   size_t left_columns, size_t left_ld, const float* right,
                                   size_t right_rows, size_t right_columns, size_t right_ld,
                                   int transpose_left, int transpose_right, float* output,
                                   size_t output_rows, size_t output_columns, size_t output_ld,
                                   char* error, size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!workspace) throw std::invalid_argument("HIP core workspace is null");
    static_cast<CoreWorkspace*>(workspace)->gemm(
        View{left, left_rows, left_columns, left_ld}, View{right, right_rows, right_columns, right_ld},
        transpose_left != 0, transpose_right != 0, output, output_rows, output_columns, output_ld);
  });
}

int fastpls_hip_core_self_gram_into_f32(void* workspace, const float* input, size_t rows,
                                        size_t columns, size_t ld, int transpose_input,
                                        float* output, size_t output_rows, size_t output_columns,
                                        size_t output_ld, int full_output, char* error,
                                        size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    if (!workspace) throw std::invalid_argument("HIP core workspace is null");
    static_cast<CoreWorkspace*>(workspace)->self_gram(View{input, rows, columns, ld},
                                                      transpose_input != 0, output, output_rows,
                                                      output_columns, output_ld, full_output != 0);
  });
}

int fastpls_hip_core_gemm_f32(const float* left, size_t left_rows, size_t left_columns,
                              size_t left_ld, const float* right, size_t right_rows,
                              size_t right_columns, size_t right_ld, int transpose_left,
                              int transpose_right, float* output, char* error,
                              size_t error_capacity) {
  return fpls::guarded(error, error_capacity, [&] {
    fpls::require_device();
    const size_t rows = transpose_left ? left_columns : left_rows;
    const size_t cols = transpose_right ? right_rows : right_columns;
    CoreWorkspace w;
    w.gemm(View{left, left_rows, left_columns, left_ld},
           View{right, right_rows, right_columns, right_ld}, transpose_left != 0,
           transpose_right != 0, output, rows, cols, rows);
  });
}

}  // extern "C"
