// VALU issue-rate probe (calibrates the kernels' rooflines): 12 independent chains of one
// instruction kind -- v_xor_b32, v_bitop3_b32, v_alignbit_b32 (the bit-sliced Hamming
// sweep), v_bcnt_u32_b32, v_min3_u32, v_lshl_or_b32, v_add_u32 (the edit recurrence's
// min* keys), v_bfi_b32, v_cndmask_b32 and DPP moves (the text-parallel edit step) -- at 8,
// 16 and 32 waves per CU.  No memory traffic apart from one
// store per thread at the end.  Prints wave-instructions per CU-cycle at 2.4 GHz and
// lane-ops/s.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <vector>

#define CHAINS 12
template <int KIND>
__global__ __launch_bounds__(256) void kern(uint32_t *out, uint32_t n, uint32_t seed)
{
    uint32_t a[CHAINS];
    for (int c = 0; c < CHAINS; ++c) a[c] = threadIdx.x * 2654435761u + c * 40503u + seed;
    uint32_t b = seed ^ 0x5bd1e995u, d = seed * 7u + 3u;
    const uint64_t sm = 0x5555555555555555ull ^ seed;
    for (uint32_t i = 0; i < n; ++i) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
#pragma unroll
            for (int c = 0; c < CHAINS; ++c) {
                if (KIND == 0) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a[c]) : "v"(b));
                else if (KIND == 1) asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96" : "+v"(a[c]) : "v"(b), "v"(d));
                else if (KIND == 2) asm volatile("v_alignbit_b32 %0, %0, %1, %2" : "+v"(a[c]) : "v"(b), "v"(d));
                else if (KIND == 3) asm volatile("v_bcnt_u32_b32 %0, %0, %1" : "+v"(a[c]) : "v"(b));
                else if (KIND == 4) asm volatile("v_min3_u32 %0, %0, %1, %2" : "+v"(a[c]) : "v"(b), "v"(d));
                else if (KIND == 5) asm volatile("v_lshl_or_b32 %0, %0, 3, %1" : "+v"(a[c]) : "v"(b));
                else if (KIND == 6) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[c]) : "v"(b));
                else if (KIND == 7) asm volatile("v_bfi_b32 %0, %0, %1, %2" : "+v"(a[c]) : "v"(b), "v"(d));
                else if (KIND == 8) asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a[c]) : "v"(b), "s"(sm));
                else asm volatile("v_mov_b32_dpp %0, %1 row_shr:1 row_mask:0xf bank_mask:0xf" : "+v"(a[c]) : "v"(a[(c + 1) % CHAINS]));
            }
        }
    }
    uint32_t x = 0;
    for (int c = 0; c < CHAINS; ++c) x ^= a[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = x;
}

int main()
{
    int dev = 0, ncu = 0, clk = 0;
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, dev);
    printf("cus %d clock_khz %d\n", ncu, clk);
    const uint32_t n = 4096;
    for (int wpc : {8, 16, 32}) {
        const uint32_t grid = ncu * wpc / 4;  // 4 waves per block
        uint32_t *out;
        if (hipMalloc(&out, (size_t)grid * 256 * 4) != hipSuccess) return 1;
        hipEvent_t e0, e1;
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        const char *names[] = {"v_xor_b32", "v_bitop3_b32", "v_alignbit_b32", "v_bcnt_u32_b32",
                               "v_min3_u32", "v_lshl_or_b32", "v_add_u32", "v_bfi_b32",
                               "v_cndmask_b32_e64", "v_mov_b32_dpp"};
        for (int kind = 0; kind < 10; ++kind) {
            for (int rep = 0; rep < 2; ++rep) {
                hipEventRecord(e0, 0);
                switch (kind) {
                case 0: hipLaunchKernelGGL(kern<0>, dim3(grid), dim3(256), 0, 0, out, n, 1u); break;
                case 1: hipLaunchKernelGGL(kern<1>, dim3(grid), dim3(256), 0, 0, out, n, 1u); break;
                case 2: hipLaunchKernelGGL(kern<2>, dim3(grid), dim3(256), 0, 0, out, n, 1u); break;
                case 3: hipLaunchKernelGGL(kern<3>, dim3(grid), dim3(256), 0, 0, out, n, 1u); break;
                case 4: hipLaunchKernelGGL(kern<4>, dim3(grid), dim3(256), 0, 0, out, n, 1u); break;
                case 5: hipLaunchKernelGGL(kern<5>, dim3(grid), dim3(256), 0, 0, out, n, 1u); break;
                case 6: hipLaunchKernelGGL(kern<6>, dim3(grid), dim3(256), 0, 0, out, n, 1u); break;
                case 7: hipLaunchKernelGGL(kern<7>, dim3(grid), dim3(256), 0, 0, out, n, 1u); break;
                case 8: hipLaunchKernelGGL(kern<8>, dim3(grid), dim3(256), 0, 0, out, n, 1u); break;
                default: hipLaunchKernelGGL(kern<9>, dim3(grid), dim3(256), 0, 0, out, n, 1u); break;
                }
                hipEventRecord(e1, 0);
                if (hipEventSynchronize(e1) != hipSuccess) return 2;
                float ms = 0;
                hipEventElapsedTime(&ms, e0, e1);
                double winstr = (double)grid * 4 * n * 8 * CHAINS;
                if (rep) printf("waves/cu %d %-15s ms %.3f wave-instr/s %.3e per CU-cycle@2.4GHz %.3f lane-ops/s %.3e\n",
                                wpc, names[kind], ms, winstr / (ms * 1e-3), winstr / (ms * 1e-3) / (ncu * 2.4e9),
                                winstr * 64 / (ms * 1e-3));
            }
        }
        hipFree(out);
    }
    return 0;
}
