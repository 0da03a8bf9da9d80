// Device helpers: XCD-aware work mapping and wavefront reductions (gfx950).
#ifndef FSL_DEVICE_UTIL_H_
#define FSL_DEVICE_UTIL_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsl {

constexpr int kXcd = 8;  // MI355X: 8 XCDs, each with its own L2

// Workgroups are dispatched to XCDs round-robin (workgroup b runs on XCD
// b % 8). Launch 8 * chunk workgroups and give XCD x the contiguous work range
// [x*chunk, (x+1)*chunk): neighbouring leaves of the (spatially ordered) leaf
// list then share one L2, so their halo re-reads hit there.
inline int xcd_grid(int64_t nwork) {
  int64_t chunk = (nwork + kXcd - 1) / kXcd;
  return int(chunk * kXcd);
}
__device__ inline int64_t xcd_work(int64_t nwork) {
  const int64_t chunk = gridDim.x / kXcd;
  return int64_t(blockIdx.x % kXcd) * chunk + blockIdx.x / kXcd;
}

// Fused multiply-add in the operands' own precision (__builtin_fma is the
// double builtin: float operands would be promoted, multiplied and added in
// double, then rounded back).
__device__ inline float fmav(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmav(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Sum over a workgroup of `nthreads` (multiple of 64, at most 1024). The
// combination order is fixed, so the result is deterministic. Every thread
// gets the total.
template <typename T>
__device__ inline T block_sum(T v, T *lds /* >= 16 entries */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  T s = T(0);
  for (int w = 0; w < nw; ++w) s += lds[w];
  __syncthreads();
  return s;
}

// Neumaier-compensated accumulation (Neumaier 1974): running sum s with
// correction c; s + c is the compensated total.
__device__ inline void neumaier_add(double &s, double &c, double x) {
  const double t = s + x;
  if (fabs(s) >= fabs(x)) c += (s - t) + x;
  else c += (x - t) + s;
  s = t;
}

}  // namespace fsl

#endif  // FSL_DEVICE_UTIL_H_
