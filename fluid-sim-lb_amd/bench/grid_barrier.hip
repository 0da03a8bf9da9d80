// Cost of a software grid-wide barrier inside one kernel, the alternative to
// one launch per PCG iteration: a cooperative launch (the runtime refuses a
// grid that cannot be resident at once) of 256 to 1,024 workgroups of 256
// threads runs 1 or 201 barriers. Two forms: one agent-scope counter that
// every workgroup increments and polls, and a counter per XCD whose last
// arrival increments a top counter, whose last arrival stores a release
// flag that all workgroups poll. Every poll loop is bounded (it gives up and
// reports after 2^22 polls), so the kernel ends in every case.
//
// Build and run (on the GPU host):
//   hipcc --offload-arch=gfx950 -O3 fluid-sim-lb_amd/bench/grid_barrier.hip -o grid_barrier
//   ./grid_barrier
#include <hip/hip_runtime.h>
#include <cstdio>
constexpr unsigned kLimit = 1u << 22;  // polls before giving up (sets *err)
__device__ inline unsigned ld_acq(unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
// flat: every workgroup adds to one counter and polls it.
__device__ inline void bar_flat(unsigned *ctr, unsigned target, unsigned *err) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    unsigned n = 0;
    while (ld_acq(ctr) < target && ++n < kLimit) __builtin_amdgcn_s_sleep(1);
    if (n >= kLimit) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
}
// two-level: workgroups of XCD x add to counter x; the last arrival of each XCD adds to the top
// counter; all poll a release flag (the epoch) written by the last arrival at the top.
__device__ inline void bar_tree(unsigned *xc, unsigned *top, unsigned *flag, unsigned epoch,
                                unsigned per_xcd, unsigned *err) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned x = blockIdx.x % 8;
    const unsigned a = __hip_atomic_fetch_add(xc + 16 * x, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (a + 1 == per_xcd * (epoch + 1)) {
      const unsigned b = __hip_atomic_fetch_add(top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (b + 1 == 8 * (epoch + 1)) __hip_atomic_store(flag, epoch + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    unsigned n = 0;
    while (ld_acq(flag) < epoch + 1 && ++n < kLimit) __builtin_amdgcn_s_sleep(1);
    if (n >= kLimit) __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
}
__global__ __launch_bounds__(256) void k_flat(unsigned *ctr, int rounds, unsigned *err) {
  for (int r = 0; r < rounds; ++r) bar_flat(ctr, (r + 1) * gridDim.x, err);
}
__global__ __launch_bounds__(256) void k_tree(unsigned *xc, unsigned *top, unsigned *flag, int rounds,
                                              unsigned *err) {
  const unsigned per_xcd = gridDim.x / 8;
  for (int r = 0; r < rounds; ++r) bar_tree(xc, top, flag, r, per_xcd, err);
}
__global__ void k_nop(int) {}
int main() {
  unsigned *mem, *err;
  hipMalloc(&mem, 4096);
  hipMalloc(&err, 4);
  int dev = 0, ncu = 0, per = 0, coop = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_flat, 256, 0);
  std::printf("CUs %d, flat kernel blocks per CU %d, cooperative %d\n", ncu, per, coop);
  for (int grid : {256, 512, 1024}) {
    for (int kind = 0; kind < 2; ++kind) {
      for (int rounds : {1, 201}) {
        hipMemset(mem, 0, 4096);
        hipMemset(err, 0, 4);
        hipDeviceSynchronize();
        hipEvent_t a, b;
        hipEventCreate(&a); hipEventCreate(&b);
        unsigned *m0 = mem, *m1 = mem + 256, *m2 = mem + 512;
        int rr = rounds;
        void *args_flat[] = {&m0, &rr, &err};
        void *args_tree[] = {&m0, &m1, &m2, &rr, &err};
        hipEventRecord(a, 0);
        hipError_t e = kind == 0
            ? hipLaunchCooperativeKernel((const void *)k_flat, dim3(grid), dim3(256), args_flat, 0, 0)
            : hipLaunchCooperativeKernel((const void *)k_tree, dim3(grid), dim3(256), args_tree, 0, 0);
        hipEventRecord(b, 0);
        hipEventSynchronize(b);
        float ms = 0; hipEventElapsedTime(&ms, a, b);
        unsigned herr = 0; hipMemcpy(&herr, err, 4, hipMemcpyDeviceToHost);
        std::printf("%s grid %4d rounds %3d: %8.2f us total (launch err %d, spin limit hit %u)\n",
                    kind ? "tree" : "flat", grid, rounds, 1e3f * ms, int(e), herr);
        hipEventDestroy(a); hipEventDestroy(b);
      }
    }
  }
  hipFree(mem); hipFree(err);
  return 0;
}
