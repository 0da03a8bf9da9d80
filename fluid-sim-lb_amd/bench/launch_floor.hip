// Per-launch cost of empty kernels on one GPU: 256 to 2,048 workgroups of 256
// threads with 0, 17 or 34 KB of LDS, timed two ways: the average of 500
// back-to-back launches between two events, and start/stop events recorded
// with each launch (hipExtLaunchKernelGGL, as the library's timing hook
// records the PCG launches). It measures the dispatch floor that every
// launch of the single-reduction PCG kernel pays.
//
// Build and run (on the GPU host):
//   hipcc --offload-arch=gfx950 -O3 fluid-sim-lb_amd/bench/launch_floor.hip -o launch_floor
//   ./launch_floor
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdio>
template <int LDS_FLOATS>
__global__ __launch_bounds__(256) void k_empty(float *out, int flag) {
  if constexpr (LDS_FLOATS > 0) {
    __shared__ float s[LDS_FLOATS];
    if (flag == 12345) {  // never true: keeps the LDS allocation
      s[threadIdx.x] = float(threadIdx.x);
      __syncthreads();
      out[threadIdx.x] = s[(threadIdx.x + 1) % 256];
    }
  }
}
template <typename K>
static void time_it(const char *name, K kern, int grid, int block, float *buf) {
  hipEvent_t a, b;
  hipEventCreate(&a); hipEventCreate(&b);
  for (int w = 0; w < 50; ++w) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, 0, buf, 0);
  hipDeviceSynchronize();
  const int n = 500;
  hipEventRecord(a, 0);
  for (int i = 0; i < n; ++i) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, 0, buf, 0);
  hipEventRecord(b, 0);
  hipEventSynchronize(b);
  float ms = 0; hipEventElapsedTime(&ms, a, b);
  // per-launch events, as the library's timing hook records them
  const int m = 100;
  hipEvent_t e0[m], e1[m];
  for (int i = 0; i < m; ++i) { hipEventCreate(&e0[i]); hipEventCreate(&e1[i]); }
  for (int i = 0; i < m; ++i)
    hipExtLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, 0, e0[i], e1[i], 0, buf, 0);
  hipDeviceSynchronize();
  float tot = 0;
  for (int i = 0; i < m; ++i) { float t; hipEventElapsedTime(&t, e0[i], e1[i]); tot += t; }
  std::printf("%-34s grid %5d  back-to-back %6.2f us  per-launch events %6.2f us\n", name, grid,
              1e3f * ms / n, 1e3f * tot / m);
  for (int i = 0; i < m; ++i) { hipEventDestroy(e0[i]); hipEventDestroy(e1[i]); }
  hipEventDestroy(a); hipEventDestroy(b);
}
int main() {
  float *buf; hipMalloc(&buf, 1024 * sizeof(float));
  for (int g : {256, 512, 1024, 2048}) time_it("empty, no LDS", k_empty<0>, g, 256, buf);
  for (int g : {512, 1024}) time_it("empty, 17 KB LDS", k_empty<4352>, g, 256, buf);
  for (int g : {512, 1024}) time_it("empty, 34 KB LDS", k_empty<8704>, g, 256, buf);
  hipFree(buf);
  return 0;
}
