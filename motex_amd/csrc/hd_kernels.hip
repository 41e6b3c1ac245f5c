// Hamming-distance FLASM kernels for gfx950: the mismatch-table build, the text-side u fold
// and fills, the generic instantiations of the sweep kernel (hd_sweep.h) and the dispatch
// over all of them.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include "motex_internal.h"
#include "hd_device.h"
#include "hd_sweep.h"

namespace motex {

// pad: bit offset of text position 0 (kPadBits for the Hamming kernel; structured runs,
// whose pair windows reach a whole sequence length to the left, use more)
__global__ void ne_build_kernel(const uint8_t *__restrict__ codes, uint32_t total,
                                uint32_t sigma, uint32_t W, uint32_t *__restrict__ ne, uint32_t pad)
{
    uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    int64_t p0 = (int64_t)w * 32 - (int64_t)pad;
    uint8_t cs[32];
#pragma unroll
    for (int t = 0; t < 32; ++t) {
        int64_t p = p0 + t;
        cs[t] = (p >= 0 && p < (int64_t)total) ? codes[p] : 0xFF;
    }
    for (uint32_t c = 0; c < sigma; ++c) {
        uint32_t bits = 0;
#pragma unroll
        for (int t = 0; t < 32; ++t) bits |= (uint32_t)(cs[t] != c) << t;
        ne[(size_t)c * W + w] = bits;
    }
}

// u of the text side of pair-symmetric runs: for every text position p, the number of x
// sequences whose bit (x, p) is set.  A workgroup takes 64 words of positions; its four
// waves each count a quarter of the x sequences, keeping the 32 counts of a word as
// bit-sliced planes (plane q = bit q of each count; adding one x's word is a ripple-carry
// increment that stops when the carry dies out), and the quarters are summed with a
// bit-sliced adder through LDS.
__global__ __launch_bounds__(256) void fold_fbits_kernel(const uint32_t *__restrict__ fbits, uint32_t nx,
                                                         uint32_t f_words, uint32_t total,
                                                         uint32_t *__restrict__ u)
{
    __shared__ uint32_t part[3][16][64];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t w = blockIdx.x * 64 + lane;
    uint32_t plane[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) plane[q] = 0;
    if (w < f_words) {
        const uint32_t x0 = (uint32_t)((uint64_t)nx * wv / 4), x1 = (uint32_t)((uint64_t)nx * (wv + 1) / 4);
        for (uint32_t x = x0; x < x1; ++x) {
            uint32_t carry = fbits[(size_t)x * f_words + w];
            for (int q = 0; q < 16; ++q) {  // stops where the carry dies out
                if (!carry) break;
                const uint32_t t = plane[q] & carry;
                plane[q] ^= carry;
                carry = t;
            }
        }
    }
    if (wv) {
#pragma unroll
        for (int q = 0; q < 16; ++q) part[wv - 1][q][lane] = plane[q];
    }
    __syncthreads();
    if (wv || w >= f_words) return;
    for (int o = 0; o < 3; ++o) {
        uint32_t c = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t a = plane[q], b = part[o][q][lane];
            plane[q] = a ^ b ^ c;
            c = (a & b) | (c & (a ^ b));
        }
    }
    uint32_t any = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) any |= plane[q];
    while (any) {
        const uint32_t b = __builtin_ctz(any);
        any &= any - 1;
        const uint32_t p = w * 32 + b;
        if (p >= total) break;
        uint32_t cnt = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) cnt |= ((plane[q] >> b) & 1u) << q;
        u[p] += cnt;
    }
}

__global__ void fill_u32_kernel(uint32_t *__restrict__ p, uint64_t n, uint32_t val)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = val;
}

// ------------------------------------------------------------------------------------
// Host-side launchers
// ------------------------------------------------------------------------------------

// Whether a four-block launch has a ring kernel for (k, e).  MOTEX_HD_RING=0 turns them off.
bool hd_has_ring(uint32_t k, uint32_t e)
{
    static const bool off = [] {
        const char *env = getenv("MOTEX_HD_RING");
        return env && env[0] == '0';
    }();
    return !off && e <= 3 && k >= kRingKMin && k <= kRingKMax;
}

uint32_t hd_hitq_cap(uint32_t k, uint32_t e, uint32_t blocks)
{
    return blocks == 4 && hd_has_ring(k, e) ? std::max<uint32_t>(kHitQueueCap, 64 * (k + 2)) : kHitQueueCap;
}

template <int B, int P>
static hipError_t launch_ph(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    switch (hd_low_planes(a.e)) {
    case 0: return launch_bp<B, P, 0>(a, grid, lds, s);
    case 1: return launch_bp<B, P, 1>(a, grid, lds, s);
    case 2: return launch_bp<B, P, 2>(a, grid, lds, s);
    case 3: return launch_bp<B, P, 3>(a, grid, lds, s);
    case 4: return launch_bp<B, P, 4>(a, grid, lds, s);
    case 5: return launch_bp<B, P, 5>(a, grid, lds, s);
    case 6: return launch_bp<B, P, 6>(a, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

template <int B>
static hipError_t launch_p(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    // the biased counter holds bias(e) + k: 2k at most, so 7 planes for k <= 64 and 11 for
    // k <= 1024 (one generic single-block kernel with the threshold plane set at run time)
    switch (hd_counter_planes(a.k, a.e)) {
    case 0: case 1: case 2: case 3: return launch_ph<B, 3>(a, grid, lds, s);
    case 4: return launch_ph<B, 4>(a, grid, lds, s);
    case 5: return launch_ph<B, 5>(a, grid, lds, s);
    case 6: return launch_ph<B, 6>(a, grid, lds, s);
    case 7: return launch_ph<B, 7>(a, grid, lds, s);
    case 8: case 9: case 10: case 11:
        if constexpr (B == 1) return launch_bp<1, 11, -1>(a, grid, lds, s);
        else return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

size_t hd_lds_bytes(uint32_t sigma, uint32_t wl, uint32_t rows_max)
{
    // NE window | v, sequence mask | range starts | row offsets (+2) | u
    return (size_t)sigma * wl * 4 + (size_t)rows_max * 8 + (kMaxRangeSeqs + 1) * 4 +
           ((size_t)rows_max + 2) * 4 + (size_t)rows_max * 4;
}

uint32_t hd_lds_words(uint32_t blocks, uint32_t rows_max)
{
    return 64 * blocks + (rows_max >> 5) + 3;
}

hipError_t launch_ne_build(const uint8_t *codes, uint32_t total, uint32_t sigma, uint32_t W,
                           uint32_t *ne, hipStream_t s, uint32_t pad)
{
    hipLaunchKernelGGL(ne_build_kernel, dim3((W + 255) / 256), dim3(256), 0, s, codes, total,
                       sigma, W, ne, pad);
    return hipGetLastError();
}

hipError_t launch_fold_fbits(const uint32_t *fbits, uint32_t nx, uint32_t f_words, uint32_t total,
                             uint32_t *u, hipStream_t s)
{
    if (nx > 65535) return hipErrorInvalidValue;  // counts are 16-bit planes
    hipLaunchKernelGGL(fold_fbits_kernel, dim3((f_words + 63) / 64), dim3(256), 0, s, fbits, nx, f_words,
                       total, u);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t val, hipStream_t s)
{
    if (!n) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, val);
    return hipGetLastError();
}

hipError_t launch_hd(const HdArgs &a, uint32_t blocks, uint32_t grid, hipStream_t s)
{
    size_t lds = hd_lds_bytes(a.sigma, a.wl, a.rows_max);
    switch (blocks) {
    case 1: return launch_p<1>(a, grid, lds, s);
    case 2: return launch_p<2>(a, grid, lds, s);
    default:
        // every work item has at least k rows (plan_hd_items), as the ring kernels assume
        if (hd_has_ring(a.k, a.e) && a.hitq_cap >= 64 * (a.k + 2)) {
            switch (a.k) {
            case 8: case 9: return launch_hd_ring_k8(a, grid, lds, s);
            case 10: case 11: return launch_hd_ring_k10(a, grid, lds, s);
            case 12: case 13: return launch_hd_ring_k12(a, grid, lds, s);
            case 14: case 15: return launch_hd_ring_k14(a, grid, lds, s);
            case 16: return launch_hd_ring_k16(a, grid, lds, s);
            default: break;
            }
        }
        return launch_p<4>(a, grid, lds, s);
    }
}

}  // namespace motex
