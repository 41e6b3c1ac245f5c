// Edit-distance FLASM kernels for gfx950 (MaxShift bit-vectors).
//
// What is computed (motifs_extraction_ed, functions.cc:734-768, over libflasm::flasm_ed,
// libflasm.h:71): the MaxShift edit recurrence of functions.cc:437-447 with bitminmax
// (functions.cc:923-967), for text rows i = 1..n and x columns j = 1..m:
//     j <= k: D = min*(shift(up)|1, shift(left)|1, shift(diag)|delta)
//     j >  k: D = min*(shiftc(left)|1, shift(up)|1, shiftc(diag)|delta)
// where shiftc(a) = (a & (2^(k-1)-1)) << 1, min* picks the word with the fewest set bits
// and, among those, the largest; row 0 holds min(j, k) ones and column 0 holds 0.  Every
// (i, j) with j >= k and popcount(D) <= e is an occurrence of the factor ending at j-1.
//
// Layout.  A wave owns a strip of 64*C columns (C adjacent columns per lane) and streams
// the text rows through it as a skewed wavefront: at step s lane L works on row s-L, so
// the left neighbour's word arrives by one DPP wave shift per step.  Strips of one x run
// one after another; lane 63's last column of every row is collected 64 rows at a time
// and becomes the next strip's left boundary.  Words are 32-bit for k <= 14 (every
// selected word is narrower than 2k+1 bits), 64-bit up to k = 32 and 128-bit up to k = 64
// (four 32-bit pieces per DPP shift).  On 32-bit words min* uses one unsigned key per
// candidate, popcount in the top bits above the complemented word, so a single v_min3
// selects it.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "motex_internal.h"
#include "ed_device.h"

namespace motex {

// Occurrence counts go straight to global u/v with one atomic pair per (strip, text,
// column with an occurrence): occurrences are rare, and keeping per-item accumulators in
// LDS (16 KB for 2,000-symbol x) limited the kernel to 10 waves per CU, 23% slower.
template <int C, int W>
__global__ __launch_bounds__(64) void ed_kernel(EdArgs a)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t k = a.k, e = a.e;
    uint32_t *buf0 = a.scratch + (size_t)blockIdx.x * 2 * a.stream_words;
    uint32_t *buf1 = buf0 + a.stream_words;
    const EdOut none{nullptr, 0};

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
        const EdItemDev I = a.items[it];
        for (uint32_t yj = I.j0; yj < I.j0 + I.ny; ++yj) {
            const uint32_t ys = a.seq_start[yj];
            const uint32_t n = a.seq_start[yj + 1] - ys;
            if (n < k) continue;  // the pair is skipped (motex.cc:536-540)
            ed_pair<C, W, false>(a.codes + I.x_start, I.m, a.codes + ys, n, k, e, buf0, buf1,
                                    a.u + I.uv_base, a.v + I.uv_base, none);
        }
    }
}

hipError_t launch_ed(const EdArgs &a, int width, uint32_t grid, hipStream_t s)
{
    if (width == 2)
        hipLaunchKernelGGL((ed_kernel<ed_cols(2), 2>), dim3(grid), dim3(64), 0, s, a);
    else if (width == 1)
        hipLaunchKernelGGL((ed_kernel<ed_cols(1), 1>), dim3(grid), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((ed_kernel<ed_cols(0), 0>), dim3(grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

uint32_t ed_strip_cols(int width) { return 64 * ed_cols(width); }
uint32_t ed_stream_row_words(int width) { return width == 0 ? 1 : (width == 1 ? 2 : 4); }

}  // namespace motex
