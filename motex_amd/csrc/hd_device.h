// Device-side helpers of the bit-sliced Hamming sweep (see hd_kernels.hip).  Shared by the
// Hamming and structured kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace motex {

// Index of the sequence containing text position pt among seq_start[j0 .. j0+ny-1].
__device__ __forceinline__ uint32_t find_seq(const uint32_t *__restrict__ seq_start,
                                             uint32_t j0, uint32_t ny, uint32_t pt)
{
    uint32_t lo = j0, hi = j0 + ny - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (seq_start[mid] <= pt) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// "count <= E" over P bit-sliced planes, least significant plane first.
template <int P, int E>
__device__ __forceinline__ uint32_t le_const(const uint32_t *c, uint32_t e)
{
    uint32_t res = ~0u;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        uint32_t eb;
        if constexpr (E >= 0) eb = ((E >> p) & 1) ? ~0u : 0u;
        else eb = ((e >> p) & 1) ? ~0u : 0u;
        res = (~c[p] & (res | eb)) | (res & eb);
    }
    return res;
}

}  // namespace motex
