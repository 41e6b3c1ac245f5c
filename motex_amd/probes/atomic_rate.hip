// Global atomic throughput probe (sets the break-even of the pair-symmetric Hamming plan,
// whose text side costs two global atomics per occurrence): scattered 32-bit atomicAdd
// and atomicOr-with-return into arrays of 2^20 .. 2^26 words from 2048 workgroups.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>

__device__ __forceinline__ uint32_t mix(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

template <int KIND>
__global__ __launch_bounds__(256) void kern(uint32_t *arr, uint32_t mask, uint32_t n, uint32_t *sink)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t a = mix(tid * 0x9E3779B9u + i) & mask;   // a <= mask < words
        if (KIND == 0) atomicAdd(arr + a, 1u);
        else acc += atomicOr(arr + a, 1u << (i & 31));
    }
    if (KIND == 1 && acc == 0x12345678u) sink[0] = acc;
}

int main()
{
    int ncu = 0;
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0);
    const uint32_t grid = ncu * 8;  // 8 blocks of 256 threads per CU
    const uint32_t n = 256;
    uint32_t *sink;
    if (hipMalloc(&sink, 64) != hipSuccess) return 1;
    for (uint32_t logw : {20u, 22u, 24u, 26u}) {
        const uint32_t words = 1u << logw;
        uint32_t *arr;
        if (hipMalloc(&arr, (size_t)words * 4) != hipSuccess) return 1;
        hipMemset(arr, 0, (size_t)words * 4);
        hipEvent_t e0, e1;
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        for (int kind = 0; kind < 2; ++kind) {
            for (int rep = 0; rep < 2; ++rep) {
                hipEventRecord(e0, 0);
                if (kind == 0) hipLaunchKernelGGL(kern<0>, dim3(grid), dim3(256), 0, 0, arr, words - 1, n, sink);
                else hipLaunchKernelGGL(kern<1>, dim3(grid), dim3(256), 0, 0, arr, words - 1, n, sink);
                hipEventRecord(e1, 0);
                if (hipEventSynchronize(e1) != hipSuccess) return 2;
                float ms = 0;
                hipEventElapsedTime(&ms, e0, e1);
                const double ops = (double)grid * 256 * n;
                if (rep) printf("words 2^%u %s ms %.3f atomics/s %.3e\n", logw, kind ? "or+ret" : "add", ms, ops / (ms * 1e-3));
            }
        }
        hipFree(arr);
    }
    return 0;
}
