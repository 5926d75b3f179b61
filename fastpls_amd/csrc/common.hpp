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
// Address-space qualifiers exist for the device pass only; the host pass
// merely parses kernel bodies.
#if defined(__HIP_DEVICE_COMPILE__)
#define FPLS_GLOBAL __attribute__((address_space(1)))
#define FPLS_CONSTANT __attribute__((address_space(4)))
#else
#define FPLS_GLOBAL
#define FPLS_CONSTANT
#endif

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)

// Element `index` of a device array that no kernel writes while it runs,
// copied into registers through the constant address space (scalar loads), so
// later field reads are not generic memory accesses.
template<class T>
__device__ __forceinline__ T load_const(const T* array, long long index) {
  return ((const T FPLS_CONSTANT*)array)[index];
}

// p[0] + p[stride] + ... + p[(count - 1) stride], summed left to right (the
// order is fixed, so results are reproducible) with eight loads in flight.
template<class P>
__device__ __forceinline__ double sum_strided(P p, long long stride, int count) {
  double s = 0.0;
  int k = 0;
  for (; k + 8 <= count; k += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[(k + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; k < count; ++k) s += p[k * stride];
  return s;
}

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

// The same over a work-group of NT threads (NT a multiple of 64); one wave
// needs no LDS and no barrier.
template<int NT>
__device__ inline double wg_sum_n(double v, double* red) {
  if constexpr (NT == 64) {
    return wave_sum(v);
  } else {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) s += red[w];
    return s;
  }
}

#endif

}  // namespace fpls
