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

// Text side of pair-symmetric runs: the first occurrences that the sweep listed per range
// (sym_flush), as a histogram in LDS, then u and v of every position += its count.  Plain
// adds: every position belongs to one range, and the sweep kernel has finished.  A list
// starts at a multiple of 8 entries and is read 8 entries (16 bytes) per load, 1024 threads
// per range, so that many loads are in flight.
__global__ __launch_bounds__(1024) void sym_fold_kernel(const uint16_t *__restrict__ tlist,
                                                       const uint32_t *__restrict__ toff,
                                                       const uint32_t *__restrict__ tcnt,
                                                       const HdRange *__restrict__ ranges, uint32_t *u,
                                                       uint32_t *v)
{
    __shared__ uint32_t hist[kSymBitmapBits];
    const uint32_t ri = blockIdx.x, tid = threadIdx.x;
    for (uint32_t i = tid; i < kSymBitmapBits; i += 1024) hist[i] = 0;
    __syncthreads();
    const uint32_t off = toff[ri], n = min(tcnt[ri], toff[ri + 1] - off);
    const uint4 *l8 = reinterpret_cast<const uint4 *>(tlist + off);
    for (uint32_t g = tid; g < n / 8; g += 1024) {
        const uint4 q = l8[g];
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            atomicAdd(&hist[w[i] & 0xFFFFu], 1u);
            atomicAdd(&hist[w[i] >> 16], 1u);
        }
    }
    for (uint32_t i = (n & ~7u) + tid; i < n; i += 1024) atomicAdd(&hist[tlist[off + i]], 1u);
    __syncthreads();
    const HdRange R = ranges[ri];
    for (uint32_t bit = tid + 32; bit < R.ye - R.ys + 32; bit += 1024) {
        const uint32_t h = hist[bit];
        if (h) {
            const uint32_t pos = R.ys + bit - 32;
            u[pos] += h;
            v[pos] += h;
        }
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
    return blocks == 4 && hd_has_ring(k, e) ? std::max<uint32_t>(kHitQueueCap, 64 * (2 * k + 2)) : kHitQueueCap;
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

size_t hd_lds_bytes(uint32_t sigma, uint32_t wl, uint32_t rows_max, uint32_t sym_group)
{
    // NE window | v, sequence mask | range starts | row offsets (+2) | u | text-side bitmaps
    // (pair-symmetric launches) or the range's word-to-text table (ordered launches)
    return (size_t)sigma * wl * 4 + (size_t)rows_max * 8 + (kMaxRangeSeqs + 1) * 4 +
           ((size_t)rows_max + 2) * 4 + (size_t)rows_max * 4 + (sym_group ? (size_t)sym_group * 1024 : kTextTabBytes);
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

hipError_t launch_sym_fold(const uint16_t *tlist, const uint32_t *toff, const uint32_t *tcnt, const HdRange *ranges,
                           uint32_t nranges, uint32_t *u, uint32_t *v, hipStream_t s)
{
    if (!nranges) return hipSuccess;
    hipLaunchKernelGGL(sym_fold_kernel, dim3(nranges), dim3(1024), 0, s, tlist, toff, tcnt, ranges, u, v);
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
    size_t lds = hd_lds_bytes(a.sigma, a.wl, a.rows_max, a.sym_group);
    switch (blocks) {
    case 1: return launch_p<1>(a, grid, lds, s);
    case 2: return launch_p<2>(a, grid, lds, s);
    default:
        // every work item has at least k rows (plan_hd_items), as the ring kernels assume
        if (hd_has_ring(a.k, a.e) && a.hitq_cap >= 64 * (2 * a.k + 2)) {
            switch (a.k) {
            case 8: case 9: return launch_hd_ring_k8(a, grid, lds, s);
            case 10: case 11: return launch_hd_ring_k10(a, grid, lds, s);
            case 12: case 13: return launch_hd_ring_k12(a, grid, lds, s);
            case 14: case 15: return launch_hd_ring_k14(a, grid, lds, s);
            default: break;
            }
        }
        return launch_p<4>(a, grid, lds, s);
    }
}

}  // namespace motex
