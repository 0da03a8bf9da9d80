// Bandwidth calibration for the single-reduction PCG kernel (k_cg): plain
// streaming passes over arrays of the 256^3 dam break's size (7680 leaves of
// 512 cells, random contents) with k_cg's algorithmic byte pattern per row
// (read r, p, label; write r', p'; read and write x every other launch: 21 B),
// and with parts of it, each timed over back-to-back launches with HIP events.
// It measures the bandwidth a kernel with that byte pattern and no halo
// gathers reaches at that size.
//
// Build and run (on the GPU host):
//   hipcc --offload-arch=gfx950 -O3 fluid-sim-lb_amd/bench/stream_calibration.hip -o stream_calibration
//   ./stream_calibration [leaves]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                   \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                             \
    }                                                                           \
  } while (0)

// M = 0: r -> r'; 1: r, p -> r', p'; 2: + labels; 3: + x every other launch.
template <int M>
__global__ __launch_bounds__(256) void k_stream(size_t n4, const float4 *rin, float4 *rout,
                                                const float4 *pin, float4 *pout, float4 *x,
                                                const unsigned *lab, float a, int dox) {
  for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n4; i += size_t(gridDim.x) * 256) {
    float4 r = rin[i];
    if (M >= 1) {
      float4 p = pin[i];
      r.x += a * p.x;
      r.y -= a * p.w;
      p.y += a * r.z;
      pout[i] = p;
    }
    if (M >= 2) r.z += float(lab[i] & 1);
    r.w = r.w * 0.999f + a;
    rout[i] = r;
    if (M >= 3 && dox) {
      float4 xv = x[i];
      xv.x += a * r.x;
      x[i] = xv;
    }
  }
}

__device__ inline unsigned mix(unsigned v) {
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  v ^= v >> 16;
  return v;
}
__global__ void k_random(float *a, size_t n, unsigned seed) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    a[i] = float(mix(unsigned(i) * 2654435761u + seed) >> 8) * (1.0f / 16777216.0f) - 0.5f;
}
__global__ void k_random_bytes(unsigned *a, size_t n, unsigned seed) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    a[i] = mix(unsigned(i) + seed);
}

int main(int argc, char **argv) {
  const size_t leaves = argc > 1 ? size_t(std::atoi(argv[1])) : 7680;
  const size_t n = leaves * 512;
  const int launches = 300, grid = 1024;
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  float *r[2], *p[2], *x;
  unsigned *lab;
  for (int i = 0; i < 2; ++i) {
    CK(hipMalloc(&r[i], n * 4));
    CK(hipMalloc(&p[i], n * 4));
    k_random<<<1024, 256>>>(r[i], n, 11 + i);
    k_random<<<1024, 256>>>(p[i], n, 21 + i);
  }
  CK(hipMalloc(&x, n * 4));
  CK(hipMalloc(&lab, n));
  k_random<<<1024, 256>>>(x, n, 5);
  k_random_bytes<<<1024, 256>>>(lab, n / 4, 7);
  CK(hipDeviceSynchronize());
  auto run = [&](auto kern, const char *name, double bytes_per_row) {
    for (int w = 0; w < 20; ++w)
      kern<<<grid, 256>>>(n / 4, (float4 *)r[0], (float4 *)r[1], (float4 *)p[0], (float4 *)p[1],
                          (float4 *)x, lab, 1e-3f, 0);
    CK(hipEventRecord(e0));
    for (int k = 0; k < launches; ++k) {
      const int b = k & 1;
      kern<<<grid, 256>>>(n / 4, (float4 *)r[b], (float4 *)r[b ^ 1], (float4 *)p[b],
                          (float4 *)p[b ^ 1], (float4 *)x, lab, 1e-3f, k & 1);
    }
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double us = 1e3 * ms / launches;
    std::printf("%-38s %7.2f us/launch %6.0f GB/s (%2.0f B/row, %zu rows)\n", name, us,
                bytes_per_row * n / (us * 1e-6) / 1e9, bytes_per_row, n);
  };
  for (int rep = 0; rep < 2; ++rep) {
    run(k_stream<3>, "r, p, label -> r', p'; x every other", 21);
    run(k_stream<2>, "r, p, label -> r', p'", 17);
    run(k_stream<1>, "r, p -> r', p'", 16);
    run(k_stream<0>, "r -> r'", 8);
  }
  // Empty launches of the same grid: the per-launch floor.
  for (int w = 0; w < 20; ++w) k_stream<0><<<grid, 256>>>(0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0.f, 0);
  CK(hipEventRecord(e0));
  for (int k = 0; k < launches; ++k)
    k_stream<0><<<grid, 256>>>(0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0.f, 0);
  CK(hipEventRecord(e1));
  CK(hipEventSynchronize(e1));
  float ms = 0;
  CK(hipEventElapsedTime(&ms, e0, e1));
  std::printf("%-38s %7.2f us/launch\n", "empty launch (same grid)", 1e3 * ms / launches);
  return 0;
}
