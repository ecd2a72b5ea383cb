// Cost of a grid-wide barrier among G co-resident blocks on one GPU: every
// block's thread 0 reads the generation word, fences (agent scope), counts
// its arrival atomically; the last block to arrive clears the count and
// advances the generation, the others wait for it (s_sleep between reads,
// every wait bounded: a block that waits too long stops and the run reports
// the failure).  FENCE 0 leaves the fences out (the count and the flag alone).
// Prints microseconds per barrier for G = 1 ... 256 (200 barriers per
// launch, each after one store per thread).
//
//   hipcc --offload-arch=gfx950 -O3 -std=c++17 -Xclang -target-feature -Xclang -packed-fp32-ops \
//       tests/gpu_native/grid_barrier_cost.hip -o /tmp/grid_barrier_cost
//   /tmp/grid_barrier_cost
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e_), __LINE__); exit(1); } } while (0)

constexpr long SPIN_LIMIT = 2000000;   // about 0.1 s of s_sleep(1) per barrier at most

__device__ inline unsigned ald(unsigned *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// FENCE 0: no fences (counter only), 1: thread 0 fences at arrival and departure
template <int FENCE>
__global__ void barrier_k(unsigned *count, unsigned *gen, int rounds, unsigned nb, double *data, int *fail) {
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int r = 0; r < rounds; ++r) {
        // a little work: each block writes one value per thread
        data[(size_t)blockIdx.x * blockDim.x + threadIdx.x] += 1.0;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned g = ald(gen);
            if (FENCE) __threadfence();
            const unsigned arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (arrived == nb - 1) {
                __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (FENCE) __threadfence();
                __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                long spins = 0;
                while (ald(gen) == g && spins < SPIN_LIMIT) { __builtin_amdgcn_s_sleep(1); ++spins; }
                if (spins >= SPIN_LIMIT) bad = 1;
            }
            if (FENCE) __threadfence();
        }
        __syncthreads();
        if (bad) break;   // uniform in the block
    }
    if (threadIdx.x == 0 && bad) atomicAdd(fail, 1);
}

int main() {
    unsigned *count, *gen;
    int *fail;
    double *data;
    CK(hipMalloc(&count, 256));
    CK(hipMalloc(&gen, 256));
    CK(hipMalloc(&fail, 256));
    CK(hipMalloc(&data, (size_t)256 * 256 * 8));
    CK(hipMemset(count, 0, 256));
    CK(hipMemset(gen, 0, 256));
    CK(hipMemset(fail, 0, 256));
    CK(hipMemset(data, 0, (size_t)256 * 256 * 8));
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, 0));
    printf("CUs %d\n", prop.multiProcessorCount);
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int rounds = 200;
    for (int fence = 0; fence < 2; ++fence)
        for (int G : {1, 8, 16, 32, 64, 128, 256}) {
            if (G > prop.multiProcessorCount) continue;
            for (int rep = 0; rep < 2; ++rep) {
                CK(hipEventRecord(e0));
                if (fence) barrier_k<1><<<G, 256>>>(count, gen, rounds, G, data, fail);
                else barrier_k<0><<<G, 256>>>(count, gen, rounds, G, data, fail);
                CK(hipEventRecord(e1));
                CK(hipEventSynchronize(e1));
                CK(hipGetLastError());
                float ms;
                CK(hipEventElapsedTime(&ms, e0, e1));
                int hf = 0;
                CK(hipMemcpy(&hf, fail, sizeof(int), hipMemcpyDeviceToHost));
                if (rep == 1)
                    printf("fence %d blocks %3d: %.2f us per barrier (launch of %d rounds %.1f us) fail %d\n", fence, G,
                           1000.0 * ms / rounds, rounds, 1000.0 * ms, hf);
                fflush(stdout);
                if (hf) return 1;
            }
        }
    return 0;
}
