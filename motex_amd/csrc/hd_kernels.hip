// Hamming-distance FLASM kernels for gfx950.
//
// What is computed (the per-pair operator motifs_extraction_hd, functions.cc:526-559, over
// libflasm::flasm_hd, libflasm.h:74): for x (length m) and text y (length n), every pair
// (pos_x, pos_t) with pos_x >= k-1, pos_t >= k-1 and
//     HD(x[pos_x-k+1 .. pos_x], y[pos_t-k+1 .. pos_t]) <= e
// is an occurrence; v[pos_x] counts them and u[pos_x] counts the texts with at least one.
//
// Layout.  All sequences are concatenated into one text; for each symbol c a bit table
// NE_c holds bit p = (text[p] != c).  A wave owns 64 lanes x 32 diagonals per block
// (q = pos_t - pos_x) and sweeps the rows pos_x of x.  For its 32 diagonals a lane keeps
// the mismatch count of the current k-window as P bit-sliced planes; one row step adds the
// entering mismatch bits (NE_{x[px]} read at diagonal offset px) and subtracts the leaving
// ones (NE_{x[px-k]} at offset px-k).  Both are funnel-shifted reads of two LDS words, so
// no counter ever moves between lanes.  A bit-sliced "count <= e" gives 32 hit flags per
// lane per row; hits are rare, so a wave-uniform slow path folds them into u and v
// (distinct texts per row for u, a global flag word per factor for texts split across
// work items).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "motex_internal.h"
#include "hd_device.h"

namespace motex {

__global__ void ne_build_kernel(const uint8_t *__restrict__ codes, uint32_t total,
                                uint32_t sigma, uint32_t W, uint32_t *__restrict__ ne)
{
    uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    int64_t p0 = (int64_t)w * 32 - (int64_t)kPadBits;
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

__global__ void fill_u32_kernel(uint32_t *__restrict__ p, uint64_t n, uint32_t val)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = val;
}

// Persistent batched kernel: one wave per workgroup pulls items from the queue head.
//
// Hit folding is lane-parallel.  A lane with hit flags in a row checks each flag (text
// range, complete window inside its sequence), finds the flag's sequence among the range's
// at most kMaxRangeSeqs starts (LDS binary search) and adds its count to the row's v with
// an LDS atomic add and its sequences to the row's sequence mask with an LDS atomic OR.
// When a range is done, each row's mask popcount (the distinct sequences of that range
// with an occurrence) is added to the row's u; a range that is part of a long sequence
// marks u once.  The item's rows are flushed to global u/v at the end.
template <int B, int P, int E>
__global__ __launch_bounds__(64) void hd_batched_kernel(HdArgs a)
{
    extern __shared__ uint32_t smem[];
    const uint32_t lane = threadIdx.x;
    const uint32_t wl = a.wl;
    // LDS: NE window (sigma + 1 rows, the last all zero) | per-row v, u, sequence mask |
    //      range sequence starts | x offsets
    uint32_t *ne_l = smem;
    uint32_t *acc_v = smem + (a.sigma + 1) * wl;
    uint32_t *acc_u = acc_v + a.rows_max;
    uint32_t *ymask = acc_u + a.rows_max;
    uint32_t *starts = ymask + a.rows_max;
    uint16_t *xoff = reinterpret_cast<uint16_t *>(starts + kMaxRangeSeqs + 1);
    const uint32_t zero_row = a.sigma * wl;
    for (uint32_t w = lane; w < wl; w += 64) ne_l[zero_row + w] = 0;

    for (uint32_t round = 0;; ++round) {
        uint32_t it;
        if (a.persistent) {
            it = 0;
            if (lane == 0) it = atomicAdd(a.queue, 1u);
            it = __builtin_amdgcn_readfirstlane(it);
        } else {
            it = round == 0 ? blockIdx.x : a.nitems;
        }
        if (it >= a.nitems) break;
        const HdItem I = a.items[it];
        const uint32_t nrows = I.r1 - I.px_start;
        __syncthreads();  // previous item finished with LDS
        for (uint32_t r = lane; r < nrows; r += 64) {
            xoff[r] = (uint16_t)(a.codes[I.x_start + I.px_start + r] * wl);
            acc_v[r] = 0;
            acc_u[r] = 0;
            ymask[r] = 0;
        }
        const uint32_t rep0 = I.r0 > a.k - 1 ? I.r0 : a.k - 1;
        const uint32_t pxs5 = I.px_start >> 5;
        for (uint32_t ri = I.rb; ri < I.re; ++ri) {
            const HdRange R = a.ranges[ri];
            // first diagonal (pt - px) of this range, a multiple of 32
            const int32_t q0 = ((int32_t)R.ys - (int32_t)I.r1 + 1) & ~31;
            const uint32_t wbase = (uint32_t)(((q0 + (int32_t)kPadBits) >> 5) + (int32_t)pxs5);
            __syncthreads();
            for (uint32_t c = 0; c < a.sigma; ++c) {
                const uint32_t *src = a.ne + (size_t)c * a.W + wbase;
                for (uint32_t w = lane; w < wl; w += 64) ne_l[c * wl + w] = src[w];
            }
            if (lane <= R.ny) starts[lane] = a.seq_start[R.j0 + lane];
            __syncthreads();

            uint32_t cnt[B][P];
#pragma unroll
            for (int b = 0; b < B; ++b)
#pragma unroll
                for (int p = 0; p < P; ++p) cnt[b][p] = 0;

            for (uint32_t r = 0; r < nrows; ++r) {
                const uint32_t px = I.px_start + r;
                const uint32_t wn = xoff[r] + (px >> 5) - pxs5 + lane;
                uint32_t wo, sho;
                if (r >= a.k) {
                    wo = xoff[r - a.k] + ((px - a.k) >> 5) - pxs5 + lane;
                    sho = (px - a.k) & 31;
                } else {
                    wo = zero_row + lane;
                    sho = 0;
                }
                const uint32_t shn = px & 31;
                uint32_t res[B];
                uint32_t any = 0;
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    uint32_t an = __builtin_amdgcn_alignbit(ne_l[wn + 64 * b + 1], ne_l[wn + 64 * b], shn);
                    uint32_t bo = __builtin_amdgcn_alignbit(ne_l[wo + 64 * b + 1], ne_l[wo + 64 * b], sho);
                    uint32_t inc = an & ~bo, dec = bo & ~an;
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        uint32_t cp = cnt[b][p];
                        cnt[b][p] = cp ^ (inc | dec);
                        uint32_t ni = cp & inc, nd = ~cp & dec;
                        inc = ni;
                        dec = nd;
                    }
                    res[b] = le_const<P, E>(cnt[b], a.e);
                    any |= res[b];
                }
                if (px >= rep0 && __ballot(any != 0) && any) {
                    uint32_t hits = 0, ym = 0;
#pragma unroll
                    for (int b = 0; b < B; ++b) {
                        uint32_t h = res[b];
                        while (h) {
                            uint32_t t = __builtin_ctz(h);
                            h &= h - 1;
                            int32_t pt = q0 + (int32_t)(kDiagPerBlock * b + 32 * lane + t + px);
                            if (pt < (int32_t)R.ys || pt >= (int32_t)R.ye) continue;
                            uint32_t y = 0;
                            if (R.ny > 1) {
                                uint32_t lo = 0, hi = R.ny - 1;
                                while (lo < hi) {
                                    uint32_t mid = (lo + hi + 1) >> 1;
                                    if (starts[mid] <= (uint32_t)pt) lo = mid; else hi = mid - 1;
                                }
                                y = lo;
                            }
                            if ((uint32_t)pt - starts[y] + 1 < a.k) continue;  // window starts before y
                            hits++;
                            ym |= 1u << y;
                        }
                    }
                    if (hits) {
                        atomicAdd(acc_v + r, hits);
                        atomicOr(ymask + r, ym);
                    }
                }
            }
            __syncthreads();
            // distinct sequences of this range per row
            for (uint32_t r = lane; r < nrows; r += 64) {
                uint32_t m = ymask[r];
                if (m) {
                    if (I.split != kNoSplit) acc_u[r] = 1;
                    else acc_u[r] += __builtin_popcount(m);
                    ymask[r] = 0;
                }
            }
        }
        __syncthreads();
        // flush the item's rows; a long text split over several items counts once per row
        for (uint32_t r = lane; r < nrows; r += 64) {
            uint32_t dv = acc_v[r];
            if (!dv) continue;
            uint32_t du = acc_u[r];
            const uint32_t px = I.px_start + r;
            if (I.split != kNoSplit && du) {
                uint32_t xf = I.xf_base + px;
                uint32_t bit = 1u << (xf & 31);
                uint32_t old = atomicOr(a.split_flags + (size_t)I.split * a.split_words + (xf >> 5), bit);
                du = (old & bit) ? 0u : 1u;
            }
            if (du) atomicAdd(a.u + I.uv_base + px, du);
            atomicAdd(a.v + I.uv_base + px, dv);
        }
    }
}

// ------------------------------------------------------------------------------------
// Host-side launchers
// ------------------------------------------------------------------------------------

template <int B, int P, int E>
static hipError_t launch_bp(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    hipLaunchKernelGGL((hd_batched_kernel<B, P, E>), dim3(grid), dim3(64), lds, s, a);
    return hipGetLastError();
}

template <int B, int P>
static hipError_t launch_e(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    switch (a.e) {
    case 0: return launch_bp<B, P, 0>(a, grid, lds, s);
    case 1: return launch_bp<B, P, 1>(a, grid, lds, s);
    case 2: return launch_bp<B, P, 2>(a, grid, lds, s);
    case 3: return launch_bp<B, P, 3>(a, grid, lds, s);
    default: return launch_bp<B, P, -1>(a, grid, lds, s);
    }
}

template <int B>
static hipError_t launch_p(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    uint32_t planes = 32 - __builtin_clz(a.k);  // bits to hold 0..k
    if (planes <= 3) return launch_e<B, 3>(a, grid, lds, s);
    if (planes == 4) return launch_e<B, 4>(a, grid, lds, s);
    if (planes == 5) return launch_e<B, 5>(a, grid, lds, s);
    if (planes == 6) return launch_e<B, 6>(a, grid, lds, s);
    return launch_e<B, 7>(a, grid, lds, s);
}

size_t hd_lds_bytes(uint32_t sigma, uint32_t wl, uint32_t rows_max)
{
    return (size_t)(sigma + 1) * wl * 4 + (size_t)rows_max * 12 + (kMaxRangeSeqs + 1) * 4 +
           (size_t)((rows_max + 1) & ~1u) * 2;
}

uint32_t hd_lds_words(uint32_t blocks, uint32_t rows_max)
{
    return 64 * blocks + (rows_max >> 5) + 3;
}

hipError_t launch_ne_build(const uint8_t *codes, uint32_t total, uint32_t sigma, uint32_t W,
                           uint32_t *ne, hipStream_t s)
{
    hipLaunchKernelGGL(ne_build_kernel, dim3((W + 255) / 256), dim3(256), 0, s, codes, total,
                       sigma, W, ne);
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
    default: return launch_p<4>(a, grid, lds, s);
    }
}

}  // namespace motex
