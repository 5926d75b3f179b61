// Shared definitions for the fastPLS HIP backend (gfx950 / MI355X).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace fpls {

constexpr int kWG = 256;          // work-group size of every kernel
constexpr int kWaves = kWG / 64;  // wave64 count per work-group

inline void hip_check(hipError_t status, const char* what) {
  if (status != hipSuccess) {
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
  }
}

#define FPLS_HIP(call) ::fpls::hip_check((call), #call)
#define FPLS_LAUNCHED(name) ::fpls::hip_check(hipGetLastError(), name)

// ---------------------------------------------------------------- device
#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Work-group sum with a fixed combination order (deterministic). Every thread
// of the work-group must call it; every thread receives the result. `red`
// points at kWaves doubles of LDS.
__device__ inline double wg_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) s += red[w];
  return s;
}

#endif

}  // namespace fpls
