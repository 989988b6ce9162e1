// grid_barrier.hip -- what one launch per step would cost the ensemble step.
//
// The batched-ensemble step is a fixed sequence of kernels replayed as one
// captured HIP graph.  Folding it into a single persistent (cooperative)
// launch would put a grid-wide barrier between each pair of dependent
// phases (sort, scan, scatter, rank, move, list, statistics, forces).  This
// measures, on the GPU it runs on, the cost of such a barrier (release fence,
// one atomic per work-group, acquire spin) for 1, 2 and 4 work-groups of 256
// threads per CU, against 11 dependent empty kernels per step launched on a
// stream and replayed from a captured graph.  A spin bound ends the kernel
// (reported) instead of waiting forever if the grid were not co-resident.
//
//   make -C caviar-v2_amd probes && caviar-v2_amd/build/probes/grid_barrier
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e_), __LINE__); return 1; } } while (0)

__device__ unsigned g_count;
__device__ int g_fail;

// nb barriers; every wave of every work-group waits; a spin bound ends the
// kernel (with g_fail set) instead of hanging if the grid is not co-resident
__global__ void k_barriers(int nb, long long spin_max) {
  const unsigned nwg = gridDim.x;
  for (int b = 1; b <= nb; ++b) {
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();
      atomicAdd(&g_count, 1u);
      const unsigned target = (unsigned)b * nwg;
      long long s = 0;
      while (__hip_atomic_load(&g_count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++s > spin_max) { g_fail = 1; break; }
        __builtin_amdgcn_s_sleep(1);
      }
      __threadfence();
    }
    __syncthreads();
    if (g_fail) return;
  }
}

__global__ void k_empty(int *p) { if (p && threadIdx.x == 0 && blockIdx.x == 0) p[0] = 1; }

int main() {
  int dev = 0, ncu = 0;
  CK(hipGetDevice(&dev));
  CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  int per_cu = 0;
  CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_barriers, 256, 0));
  printf("CUs %d, co-resident 256-thread blocks per CU %d\n", ncu, per_cu);
  hipStream_t st;
  CK(hipStreamCreate(&st));
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  for (int wpc : {1, 2, 4}) {
    if (wpc > per_cu) break;
    const int grid = ncu * wpc;
    for (int nb : {10, 100}) {
      unsigned zero = 0;
      int zi = 0;
      CK(hipMemcpyToSymbol(HIP_SYMBOL(g_count), &zero, sizeof(zero)));
      CK(hipMemcpyToSymbol(HIP_SYMBOL(g_fail), &zi, sizeof(zi)));
      long long spin = 20000000LL;
      void *args[] = {&nb, &spin};
      CK(hipEventRecord(a, st));
      CK(hipLaunchCooperativeKernel((void *)k_barriers, dim3(grid), dim3(256), args, 0, st));
      CK(hipEventRecord(b, st));
      CK(hipEventSynchronize(b));
      float ms = 0;
      CK(hipEventElapsedTime(&ms, a, b));
      int fail = 0;
      CK(hipMemcpyFromSymbol(&fail, HIP_SYMBOL(g_fail), sizeof(fail)));
      printf("cooperative grid %d x 256: %d barriers in %.1f us (%.2f us per barrier)%s\n", grid, nb,
             ms * 1e3, ms * 1e3 / nb, fail ? "  SPIN LIMIT HIT" : "");
    }
  }
  // chains of empty kernels: 11 per step, 100 steps (stream, then a captured graph)
  for (int g : {1, 4096}) {
    for (int rep = 0; rep < 2; ++rep) {
      CK(hipEventRecord(a, st));
      for (int s = 0; s < 100; ++s)
        for (int k = 0; k < 11; ++k) k_empty<<<g, 256, 0, st>>>(nullptr);
      CK(hipEventRecord(b, st));
      CK(hipEventSynchronize(b));
      float ms = 0;
      CK(hipEventElapsedTime(&ms, a, b));
      if (rep) printf("stream: 11 empty kernels of %d blocks per step: %.1f us per step\n", g, ms * 1e3 / 100);
    }
    hipGraph_t graph;
    hipGraphExec_t ge;
    CK(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
    for (int k = 0; k < 11; ++k) k_empty<<<g, 256, 0, st>>>(nullptr);
    CK(hipStreamEndCapture(st, &graph));
    CK(hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0));
    for (int rep = 0; rep < 2; ++rep) {
      CK(hipEventRecord(a, st));
      for (int s = 0; s < 100; ++s) CK(hipGraphLaunch(ge, st));
      CK(hipEventRecord(b, st));
      CK(hipEventSynchronize(b));
      float ms = 0;
      CK(hipEventElapsedTime(&ms, a, b));
      if (rep) printf("graph: 11 empty kernels of %d blocks per step: %.1f us per step\n", g, ms * 1e3 / 100);
    }
    CK(hipGraphExecDestroy(ge));
    CK(hipGraphDestroy(graph));
  }
  return 0;
}
