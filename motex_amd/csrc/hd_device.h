// Device-side helpers of the bit-sliced Hamming sweep (see hd_kernels.hip).  Shared by the
// Hamming and structured kernels.
//
// A window's mismatch count lives in P bit-sliced planes (plane p = bit p of the 32 counts of
// a word), least significant plane first.  The counts are stored BIASED: a counter starts at
//     bias = 2^PH - 1 - e,    PH = bits(e) (0 for e = 0),
// so that "count <= e" is "biased count < 2^PH", i.e. planes PH..P-1 all zero: the threshold
// test is an OR of the high planes instead of a bit-serial compare.  The counter must hold
// bias + k without wrapping (P = bits(bias + k)).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace motex {

// Planes below the threshold plane for error bound e.
__host__ __device__ constexpr uint32_t hd_low_planes(uint32_t e)
{
    return e ? 32u - (uint32_t)__builtin_clz(e) : 0u;
}

__host__ __device__ constexpr uint32_t hd_bias(uint32_t e)
{
    return (1u << hd_low_planes(e)) - 1u - e;
}

// Planes a biased window counter of length k needs.
__host__ __device__ constexpr uint32_t hd_counter_planes(uint32_t k, uint32_t e)
{
    return 32u - (uint32_t)__builtin_clz(hd_bias(e) + k);
}

// Counter planes holding `bias` in every lane.
template <int P>
__device__ __forceinline__ void planes_set(uint32_t *c, uint32_t bias)
{
#pragma unroll
    for (int p = 0; p < P; ++p) c[p] = ((bias >> p) & 1u) ? ~0u : 0u;
}

// +1 in the lanes of t.
template <int P>
__device__ __forceinline__ void planes_inc(uint32_t *c, uint32_t t)
{
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const uint32_t cp = c[p];
        c[p] = cp ^ t;
        if (p + 1 < P) t &= cp;
    }
}

// Slides the window: +1 where a mismatch enters and none leaves, -1 where one leaves and none
// enters.  The lanes that change are t = enter ^ leave, and in those lanes the direction is
// the entering bit itself, so one carry chain serves both: an increment carries out of a
// plane that was 1, a decrement borrows out of a plane that was 0, i.e. the chain continues
// where the old plane equals the direction bit.
//
// Written as 3-input logic ops (v_bitop3_b32; the table's bit a*4 + b*2 + c is f(a, b, c)),
// two per plane: plane 0 takes enter and leave directly (its new bit is c ^ enter ^ leave,
// table 0x96, and its carry (enter ^ leave) & ~(c ^ enter), table 0x42), every later plane
// XORs the carry in and passes on carry & ~(c ^ enter) (table 0x90).
template <int P>
__device__ __forceinline__ void planes_slide(uint32_t *c, uint32_t enter, uint32_t leave)
{
    const uint32_t c0 = c[0];
    c[0] = __builtin_amdgcn_bitop3_b32(c0, enter, leave, 0x96);
    uint32_t t = __builtin_amdgcn_bitop3_b32(c0, enter, leave, 0x42);
#pragma unroll
    for (int p = 1; p < P; ++p) {
        const uint32_t cp = c[p];
        c[p] = cp ^ t;
        if (p + 1 < P) t = __builtin_amdgcn_bitop3_b32(t, cp, enter, 0x90);
    }
}

// OR of planes PH..P-1: bit clear <=> the window's count is within the error bound.
template <int P, int PH>
__device__ __forceinline__ uint32_t planes_high(const uint32_t *c)
{
    static_assert(PH < P, "the counter needs a plane above the threshold");
    uint32_t h = c[PH];
#pragma unroll
    for (int p = PH + 1; p < P; ++p) h |= c[p];
    return h;
}

// acc & planes_high(c): the last two high planes and acc in one 3-input op ((a | b) & c,
// table 0xA8).
template <int P, int PH>
__device__ __forceinline__ uint32_t planes_high_and(const uint32_t *c, uint32_t acc)
{
    static_assert(PH < P, "the counter needs a plane above the threshold");
    if constexpr (P - PH == 1) {
        return acc & c[PH];
    } else {
        uint32_t h = c[PH];
#pragma unroll
        for (int p = PH + 1; p < P - 1; ++p) h |= c[p];
        return __builtin_amdgcn_bitop3_b32(h, c[P - 1], acc, 0xA8);
    }
}

// Masks selecting planes ph..P-1 (wave-uniform), for planes_high_masked: the threshold
// plane chosen at run time.
template <int P>
__device__ __forceinline__ void planes_high_masks(uint32_t *hm, uint32_t ph)
{
#pragma unroll
    for (int p = 0; p < P; ++p) hm[p] = (uint32_t)p >= ph ? ~0u : 0u;
}

// planes_high with the threshold plane chosen at run time, from planes_high_masks: one
// 3-input logic op per plane, (c & m) | h (v_bitop3_b32 table 0xEA), in place of a select and
// an OR.
template <int P>
__device__ __forceinline__ uint32_t planes_high_masked(const uint32_t *c, const uint32_t *hm)
{
    uint32_t h = 0;
#pragma unroll
    for (int p = 0; p < P; ++p) h = __builtin_amdgcn_bitop3_b32(c[p], hm[p], h, 0xEA);
    return h;
}

}  // namespace motex
