// Device-wide exclusive prefix sum of uint32 arrays (counting-sort offsets,
// leaf compaction, seeding offsets). Three passes: per-tile reduce, a
// single-workgroup scan of the tile sums, per-tile scan with offsets.
// out[n] receives the total.
#include "device_util.h"
#include "fsl_internal.h"

namespace fsl {

namespace {

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Exclusive scan of one value per thread over the workgroup; returns the
// exclusive prefix and writes the workgroup total to *total.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *lds, uint32_t *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  uint32_t inc = wave_incl_scan(v);
  if (lane == 63) lds[wave] = inc;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < nw ? lds[lane] : 0u;
    uint32_t wi = wave_incl_scan(w);
    if (lane < nw) lds[32 + lane] = wi - w;
    if (lane == nw - 1) lds[31] = wi;
  }
  __syncthreads();
  uint32_t res = inc - v + lds[32 + wave];
  *total = lds[31];
  __syncthreads();
  return res;
}

// The kScanItems values of a thread: two 16-byte loads when all of them are
// in range and the array is 16-byte aligned, else one load each (0 past n).
__device__ inline void load_items(const uint32_t *in, int64_t base, int64_t n,
                                  uint32_t v[kScanItems]) {
  if (base + kScanItems <= n && (reinterpret_cast<uintptr_t>(in) & 15) == 0) {
    const uint4 a = *reinterpret_cast<const uint4 *>(in + base);
    const uint4 b = *reinterpret_cast<const uint4 *>(in + base + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) v[q] = base + q < n ? in[base + q] : 0u;
  }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *in, int64_t n,
                                                              uint32_t *tile_sum) {
  __shared__ uint32_t lds[64];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
  uint32_t vals[kScanItems];
  load_items(in, base, n, vals);
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) s += vals[q];
  uint32_t tot;
  block_excl_scan(s, lds, &tot);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(uint32_t *tile_sum, int64_t ntiles,
                                                             uint32_t *out_total) {
  __shared__ uint32_t lds[64];
  const int64_t per = (ntiles + kScanThreads - 1) / kScanThreads;
  const int64_t b = int64_t(threadIdx.x) * per;
  uint32_t s = 0;
  for (int64_t q = 0; q < per; ++q)
    if (b + q < ntiles) s += tile_sum[b + q];
  uint32_t tot;
  uint32_t run = block_excl_scan(s, lds, &tot);
  for (int64_t q = 0; q < per; ++q)
    if (b + q < ntiles) {
      uint32_t t = tile_sum[b + q];
      tile_sum[b + q] = run;
      run += t;
    }
  if (threadIdx.x == 0) *out_total = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t *in, uint32_t *out,
                                                             int64_t n, const uint32_t *tile_off) {
  __shared__ uint32_t lds[64];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
  uint32_t vals[kScanItems];
  load_items(in, base, n, vals);
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) s += vals[q];
  uint32_t tot;
  uint32_t run = block_excl_scan(s, lds, &tot) + tile_off[blockIdx.x];
  uint32_t res[kScanItems];
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    res[q] = run;
    run += vals[q];
  }
  if (base + kScanItems <= n && (reinterpret_cast<uintptr_t>(out) & 15) == 0) {
    *reinterpret_cast<uint4 *>(out + base) = make_uint4(res[0], res[1], res[2], res[3]);
    *reinterpret_cast<uint4 *>(out + base + 4) = make_uint4(res[4], res[5], res[6], res[7]);
  } else {
#pragma unroll
    for (int q = 0; q < kScanItems; ++q)
      if (base + q < n) out[base + q] = res[q];
  }
}

}  // namespace

int64_t scan_tmp_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

// in and out may alias.
hipError_t exclusive_scan(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *tmp,
                          hipStream_t s) {
  const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
  if (ntiles == 0) {
    return hipMemsetAsync(out, 0, sizeof(uint32_t), s);
  }
  k_scan_reduce<<<dim3(unsigned(ntiles)), dim3(kScanThreads), 0, s>>>(in, n, tmp);
  k_scan_tiles<<<dim3(1), dim3(kScanThreads), 0, s>>>(tmp, ntiles, out + n);
  k_scan_apply<<<dim3(unsigned(ntiles)), dim3(kScanThreads), 0, s>>>(in, out, n, tmp);
  return hipGetLastError();
}

}  // namespace fsl
