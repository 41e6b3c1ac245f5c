// Device-side helpers of the bit-sliced Hamming sweep (see hd_kernels.hip).  Shared by the
// Hamming and structured kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace motex {

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
