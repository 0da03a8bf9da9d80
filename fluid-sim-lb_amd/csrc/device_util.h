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

// XCD-aware walk over a leaf list of n entries (n read from *count when
// count is set): workgroup b runs on XCD x = b % 8 and takes entries
// x * chunk + j, j = b / 8, b / 8 + G / 8, ... of the XCD's contiguous chunk
// (chunk = ceil(n / 8), G the grid size). A grid of 8 * ceil(n / 8)
// workgroups gives each one entry (as xcd_work); a smaller grid makes the
// workgroups persistent. f(leaf) runs once per entry; every thread of the
// workgroup takes the same entries, so f may synchronise the workgroup.
template <typename F>
__device__ inline void for_each_leaf(const int32_t *list, const uint32_t *count, int64_t n,
                                     F &&f) {
  if (count) n = int64_t(*count);
  const int64_t chunk = (n + kXcd - 1) / kXcd;
  const int64_t beg = int64_t(blockIdx.x % kXcd) * chunk;
  const int64_t end = beg + chunk < n ? beg + chunk : n;
  const int64_t per = gridDim.x / kXcd;
  for (int64_t w = beg + blockIdx.x / kXcd; w < end; w += per) f(int64_t(list[w]));
}
// Workgroups for a walk over a leaf list: one per entry for a host-known
// count, at most kLeafGridCap (persistent) for a device count.
constexpr int kLeafGridCap = kXcd * 1024;
inline int leaf_grid(int64_t n, bool device_count) {
  const int g = xcd_grid(n);
  return device_count && g > kLeafGridCap ? kLeafGridCap : g;
}

// Fused multiply-add in the operands' own precision (__builtin_fma is the
// double builtin: float operands would be promoted, multiplied and added in
// double, then rounded back).
__device__ inline float fmav(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmav(double a, double b, double c) { return __builtin_fma(a, b, c); }

// DPP move of a 32- or 64-bit value (64-bit values move as two halves).
// Lanes whose row is off in ROW_MASK receive an unspecified value: with no
// defined old value the move needs no register initialised first (two
// fewer instructions per 64-bit step).
template <int CTRL, int ROW_MASK = 0xF, typename T>
__device__ inline T dpp_move(T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
  if constexpr (sizeof(T) == 4) {
    const int r = __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, ROW_MASK, 0xF, false);
    return __builtin_bit_cast(T, r);
  } else {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp(int(b), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(int(b >> 32), CTRL, ROW_MASK, 0xF, false);
    return __builtin_bit_cast(T, (long long)(unsigned(lo)) | ((long long)hi << 32));
  }
}
template <typename T>
__device__ inline T read_lane(T v, int lane) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
  } else {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane(int(b), lane);
    const int hi = __builtin_amdgcn_readlane(int(b >> 32), lane);
    return __builtin_bit_cast(T, (long long)(unsigned(lo)) | ((long long)hi << 32));
  }
}

// Sum over a 64-lane wavefront, returned to every lane. DPP steps: swaps
// within quads, rotations within rows of 16, then row 15 / lane 31
// broadcasts into the upper rows; lane 63 holds the total. The order of the
// additions is fixed, so the result is deterministic. Lanes must all be
// active. The broadcast steps leave unspecified values in the rows they do
// not write (rows 0 and 2, then rows 0 and 1); lane 63's chain never reads
// them: step 5 gives row 3 lane 47's sum of row 2, step 6 gives it lane
// 31's sum of rows 0 and 1 (row 1 was written in step 5).
template <typename T>
__device__ inline T wave_sum(T v) {
  v += dpp_move<0xB1>(v);        // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);        // quad_perm [2,3,0,1]
  v += dpp_move<0x124>(v);       // row_ror:4
  v += dpp_move<0x128>(v);       // row_ror:8
  v += dpp_move<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return read_lane(v, 63);
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
#pragma clang fp contract(off)
  const double t = s + x;
  if (fabs(s) >= fabs(x)) c += (s - t) + x;
  else c += (x - t) + s;
  s = t;
}

// Knuth's TwoSum: s = fl(a + b) and the exact error e, a + b = s + e.
__device__ inline void two_sum(double a, double b, double &s, double &e) {
#pragma clang fp contract(off)
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
// Adds the compensated pair (s2, c2) to (s, c).
__device__ inline void comp_add(double &s, double &c, double s2, double c2) {
#pragma clang fp contract(off)
  double t, e;
  two_sum(s, s2, t, e);
  s = t;
  c = (c + c2) + e;
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ inline void comp_dpp_step(double &s, double &c) {
  comp_add(s, c, dpp_move<CTRL, ROW_MASK>(s), dpp_move<CTRL, ROW_MASK>(c));
}
// Compensated sum of the pairs (s, c) over a 64-lane wavefront, in the DPP
// order of wave_sum: every combination is a TwoSum whose error joins the
// correction term. Returns the total pair to every lane (fold with s + c).
__device__ inline void wave_sum_c(double &s, double &c) {
  comp_dpp_step<0xB1>(s, c);
  comp_dpp_step<0x4E>(s, c);
  comp_dpp_step<0x124>(s, c);
  comp_dpp_step<0x128>(s, c);
  comp_dpp_step<0x142, 0xA>(s, c);
  comp_dpp_step<0x143, 0xC>(s, c);
  s = read_lane(s, 63);
  c = read_lane(c, 63);
}

}  // namespace fsl

#endif  // FSL_DEVICE_UTIL_H_
