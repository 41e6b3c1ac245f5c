// Structured range sweep (struct_range_sweep): one wave per unit, every distinct box shape's
// window counters down the rows of x, hits into the unit's slot.  Shared by the generic kernels
// (struct_range_kernels.hip) and the register-ring kernel (struct_range_ring.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hd_device.h"
#include "motex_internal.h"

namespace motex {

// First diagonal (pt - px) of a unit's window: the range's first position at the last row,
// rounded down to a multiple of 32.
__device__ __forceinline__ int32_t sr_window_q0(int32_t ys, uint32_t m)
{
    return (ys - (int32_t)m + 1) & ~31;
}

// ---------------------------------------------------------------------------------------
// Sweep
// ---------------------------------------------------------------------------------------

// Appends row px's hits (the clear bits of high) to the slot; false when they do not fit.
__device__ __forceinline__ bool sr_emit(const uint32_t *high, uint32_t px, uint32_t lane, uint32_t *hits,
                                        uint32_t hcap, uint32_t &nq)
{
    const uint32_t w0 = ~high[0], w1 = ~high[1], w2 = ~high[2], w3 = ~high[3];
    const uint32_t any = w0 | w1 | w2 | w3;
    const uint64_t mb = __ballot(any != 0);
    if (!mb) return true;
    const uint32_t nh = __builtin_popcount(w0) + __builtin_popcount(w1) + __builtin_popcount(w2) + __builtin_popcount(w3);
    uint64_t bt = __ballot(nh > 1);
    if (!bt) {
        // at most one hit a lane (the common case): slot = lanes with a hit below
        const uint32_t tot = (uint32_t)__builtin_popcountll(mb);
        if (nq + tot > hcap) return false;
        if (any) {
            const uint32_t b = ((w1 | w3) ? 1u : 0u) | ((w2 | w3) ? 2u : 0u);
            const uint32_t at = nq + __builtin_amdgcn_mbcnt_hi((uint32_t)(mb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mb, 0u));
            hits[at] = (px << 16) | (kDiagPerBlock * b + 32 * lane + __builtin_ctz(any));
        }
        nq += tot;
        return true;
    }
    // some lane has several: slots from the wave's exclusive prefix of the counts, one ballot
    // per count level
    uint32_t pre = 0, tot = 0;
    bt = mb;
    for (uint32_t t = 1; bt; ++t) {
        pre = __builtin_amdgcn_mbcnt_hi((uint32_t)(bt >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bt, pre));
        tot += (uint32_t)__builtin_popcountll(bt);
        bt = __ballot(nh > t);
    }
    if (nq + tot > hcap) return false;
    if (nh) {
        uint32_t at = nq + pre;
        const uint32_t tag = (px << 16) | (32 * lane);
#pragma unroll
        for (int b = 0; b < (int)kSrBlocks; ++b) {
            uint32_t hv = ~high[b];
            while (hv) {
                hits[at++] = tag + kDiagPerBlock * b + __builtin_ctz(hv);
                hv &= hv - 1;
            }
        }
    }
    nq += tot;
    return true;
}

// OR of the planes at and above the threshold plane PH (a clear bit: the window is within the
// error bound).  PH >= P only when the bound is at least the box length: every window passes.
template <int P, int PH>
__device__ __forceinline__ uint32_t sr_high(const uint32_t *c)
{
    if constexpr (PH >= P) return 0u;
    else return planes_high<P, PH>(c);
}

// One shape's sweep down the rows of x: P counter planes, threshold plane PH.  The first len
// rows fill the windows (no leaving symbol); returns true when the hits overflowed the slot.
template <int P, int PH>
__device__ bool sr_sweep_shape(const char *ne_b, const uint16_t *xoff, uint32_t m, uint32_t len, uint32_t err,
                               uint32_t *hits, uint32_t hcap, uint32_t lane, uint32_t &nq)
{
    uint32_t cnt[kSrBlocks][P];
#pragma unroll
    for (int b = 0; b < (int)kSrBlocks; ++b) planes_set<P>(cnt[b], hd_bias(err));
    uint32_t high[kSrBlocks];
    const uint32_t warm = min(len, m);
    for (uint32_t px = 0; px < warm; ++px) {
        const uint32_t *wn = reinterpret_cast<const uint32_t *>(ne_b + xoff[px]);
#pragma unroll
        for (int b = 0; b < (int)kSrBlocks; ++b) {
            planes_inc<P>(cnt[b], __builtin_amdgcn_alignbit(wn[64 * b + 1], wn[64 * b], px & 31));
            high[b] = sr_high<P, PH>(cnt[b]);
        }
    }
    if (len <= m && !sr_emit(high, len - 1, lane, hits, hcap, nq)) return true;
    for (uint32_t px = len; px < m; ++px) {
        const uint32_t *wn = reinterpret_cast<const uint32_t *>(ne_b + xoff[px]);
        const uint32_t *wo = reinterpret_cast<const uint32_t *>(ne_b + xoff[px - len]);
        const uint32_t sho = (px - len) & 31;
#pragma unroll
        for (int b = 0; b < (int)kSrBlocks; ++b) {
            const uint32_t an = __builtin_amdgcn_alignbit(wn[64 * b + 1], wn[64 * b], px & 31);
            const uint32_t bo = __builtin_amdgcn_alignbit(wo[64 * b + 1], wo[64 * b], sho);
            planes_slide<P>(cnt[b], an, bo);
            high[b] = sr_high<P, PH>(cnt[b]);
        }
        if (!sr_emit(high, px, lane, hits, hcap, nq)) return true;
    }
    return false;
}

// Entries of a ring sweep's row queue in LDS: row px | lane << 16, and the lane's four
// high-plane words (a clear bit: a hit).  A row with hits queues one entry per lane with a hit;
// the queue is drained into the slot once more than kSrQueue - 64 entries are held.
constexpr uint32_t kSrQueue = 128;

// Appends the hits of queue entries [0, n) to the slot in queue order (rows in order; a row's
// lanes in order; a lane's blocks and bits in order -- the order sr_emit writes), each lane
// taking one entry a pass; false when they do not fit.
__device__ __forceinline__ bool sr_drain(const uint32_t *qmeta, const uint4 *qw, uint32_t n, uint32_t lane,
                                         uint32_t *hits, uint32_t hcap, uint32_t &nq)
{
    __syncthreads();  // the queue stores of every lane are visible to the loads below
    for (uint32_t q0 = 0; q0 < n; q0 += 64) {
        const uint32_t q = q0 + lane;
        uint32_t w[4] = {0u, 0u, 0u, 0u}, meta = 0;
        if (q < n) {
            const uint4 v = qw[q];
            meta = qmeta[q];
            w[0] = ~v.x;
            w[1] = ~v.y;
            w[2] = ~v.z;
            w[3] = ~v.w;
        }
        const uint32_t c = __builtin_popcount(w[0]) + __builtin_popcount(w[1]) + __builtin_popcount(w[2]) +
                           __builtin_popcount(w[3]);
        uint32_t incl = c;  // inclusive prefix sum over the wave's lanes
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d, 64);
            if (lane >= (uint32_t)d) incl += o;
        }
        const uint32_t tot = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
        if (nq + tot > hcap) return false;
        uint32_t at = nq + incl - c;
        const uint32_t tag = ((meta & 0xFFFFu) << 16) | (32 * (meta >> 16));
#pragma unroll
        for (int b = 0; b < (int)kSrBlocks; ++b)
            for (uint32_t hv = w[b]; hv; hv &= hv - 1) hits[at++] = tag + kDiagPerBlock * b + __builtin_ctz(hv);
        nq += tot;
    }
    return true;
}

// Box lengths with a register-ring sweep (struct_range_ring_*.hip; longer rings spill).
constexpr uint32_t kSrRingLenMin = 5, kSrRingLenMax = 10;

// One shape of length L (L <= m): the column leaving at row px is the one that entered at row
// px - L, whose bits the lane computed then and keeps in a ring of L rows' entering words
// (slot px mod L: the loop runs in trips of L rows from row L -- 2 L for odd L -- so every slot
// index is a constant).  Rows are software-pipelined: while row px computes, the window words
// of row px + 1 load into the other buffer (row px uses buffer (px + L) & 1) and the offset of
// row px + 2 loads from LDS (xoff holds two spare entries past the last row).  A row with hits
// queues its lanes' high-plane words in LDS (qmeta, qw) instead of writing each hit, and the
// queue drains into the slot in order (sr_drain): the slot ends up as sr_emit writes it.
// Returns true when the hits overflowed the slot.
template <int P, int PH, int L>
__device__ bool sr_sweep_ring(const char *ne_b, const uint16_t *xoff, uint32_t m, uint32_t err, uint32_t *hits,
                              uint32_t hcap, uint32_t lane, uint32_t &nq, uint32_t *qmeta, uint4 *qw)
{
    constexpr int NB = (int)kSrBlocks;
    constexpr int TRIP = (L % 2 == 0) ? L : 2 * L;
    uint32_t cnt[NB][P];
#pragma unroll
    for (int b = 0; b < NB; ++b) planes_set<P>(cnt[b], hd_bias(err));
    uint32_t ring[L][NB];
    uint32_t wbuf[2][NB][2];
    auto load = [&](uint32_t off, uint32_t(&w)[NB][2]) {
        const uint32_t *wn = reinterpret_cast<const uint32_t *>(ne_b + off);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            w[b][0] = wn[64 * b];
            w[b][1] = wn[64 * b + 1];
        }
    };
    uint32_t nqs = 0;  // queued entries (wave-uniform)
    bool ovf = false;
    auto emit = [&](uint32_t px) {
        uint32_t high[NB], all = ~0u;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            high[b] = sr_high<P, PH>(cnt[b]);
            all &= high[b];
        }
        const uint64_t mq = __ballot(all != ~0u);
        if (mq) {
            if (all != ~0u) {
                const uint32_t at =
                    nqs + __builtin_amdgcn_mbcnt_hi((uint32_t)(mq >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mq, 0u));
                qw[at] = make_uint4(high[0], high[1], high[2], high[3]);
                qmeta[at] = px | (lane << 16);
            }
            nqs += (uint32_t)__builtin_popcountll(mq);
            if (nqs > kSrQueue - 64) {
                // (after an overflow the rows still run, but the unit's slot is not read)
                ovf = ovf || !sr_drain(qmeta, qw, nqs, lane, hits, hcap, nq);
                nqs = 0;
                __syncthreads();  // the drain's loads are done before the queue is written again
            }
        }
    };
    // the first L rows fill the windows
    load(xoff[0], wbuf[L & 1]);
    uint32_t x1 = xoff[1];
#pragma unroll
    for (int s = 0; s < L; ++s) {
        load(x1, wbuf[(s + 1 + L) & 1]);
        x1 = xoff[s + 2];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            ring[s][b] = __builtin_amdgcn_alignbit(wbuf[(s + L) & 1][b][1], wbuf[(s + L) & 1][b][0], (uint32_t)s);
            planes_inc<P>(cnt[b], ring[s][b]);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    emit(L - 1);
    auto row = [&](uint32_t px, int s) {
        load(x1, wbuf[(s + 1) & 1]);
        x1 = xoff[px + 2];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const uint32_t an = __builtin_amdgcn_alignbit(wbuf[s & 1][b][1], wbuf[s & 1][b][0], px & 31);
            planes_slide<P>(cnt[b], an, ring[s % L][b]);
            ring[s % L][b] = an;
        }
        emit(px);
    };
    uint32_t px = L;
    for (; px + TRIP <= m; px += TRIP) {
#pragma unroll
        for (int s = 0; s < TRIP; ++s) {
            row(px + s, s);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int s = 0; s < TRIP - 1; ++s)
        if (px + s < m) row(px + s, s);
    if (nqs) {
        ovf = ovf || !sr_drain(qmeta, qw, nqs, lane, hits, hcap, nq);
        __syncthreads();
    }
    return ovf;
}

// RL > 0: shapes of length RL whose threshold plane is RPH take the register-ring sweep (P = 4
// holds the counters of every such shape); the others, the generic one.
template <int P, int RL = 0, int RPH = 0>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RL > 0 ? 4 : 1))) void struct_range_sweep(SrSweepArgs a)
{
    static_assert(kSrBlocks == 4, "the hit emission names four blocks");
    extern __shared__ uint32_t smem[];
    const uint32_t lane = threadIdx.x;
    const uint32_t wl = a.wl;
    // LDS: NE window (sigma rows of wl words) | x table (u16 byte offsets of a row's window
    //      words, two spare entries of offset 0 past the last row) | one word to broadcast
    //      through
    uint32_t *ne_l = smem;
    uint16_t *xoff = reinterpret_cast<uint16_t *>(smem + a.sigma * wl);
    uint32_t *bcast = smem + a.sigma * wl + (a.m_max + 3) / 2;
    // (ring kernels: the row queue follows, 16-byte aligned, kSrQueue tags then kSrQueue entries
    // of four words)
    uint32_t *qmeta = smem + ((a.sigma * wl + (a.m_max + 3) / 2 + 1 + 3) & ~3u);
    uint4 *qw = reinterpret_cast<uint4 *>(qmeta + kSrQueue);
    const uint32_t slot_words = kSrHead + a.hcap;

    for (;;) {
        uint32_t it = 0;
        if (lane == 0) it = atomicAdd(a.queue, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= a.nitems) break;
        const SrItem I = a.items[a.first + it];
        const HdRange HR = a.ranges[I.r];
        const uint32_t m = I.m;
        uint32_t *slot = a.slots + (size_t)it * slot_words;
        uint32_t *hits = slot + kSrHead;
        const int32_t q0 = sr_window_q0((int32_t)HR.ys, m);
        const uint32_t wbase = (uint32_t)((q0 + (int32_t)a.pad_bits) >> 5);
        __syncthreads();  // the previous unit is done with LDS
        for (uint32_t r = lane; r < m + 2; r += 64)
            xoff[r] = r < m ? (uint16_t)(4 * (a.codes[I.x_start + r] * wl + (r >> 5))) : (uint16_t)0;
        for (uint32_t c = 0; c < a.sigma; ++c) {
            const uint32_t *src = a.ne + (size_t)c * a.W + wbase;
            for (uint32_t w = lane; w < wl; w += 64) ne_l[c * wl + w] = src[w];
        }
        __syncthreads();

        uint32_t nq = 0;  // hits so far (wave-uniform)
        bool overflow = false;
        const char *ne_b = reinterpret_cast<const char *>(ne_l) + 4 * lane;
        for (uint32_t d = 0; d < a.ndist && !overflow; ++d) {
            const uint32_t len = a.dlen[d], err = a.derr[d];
            // the threshold plane is fixed per shape: one loop nest per (P, PH)
            const uint32_t ph = min(hd_low_planes(err), (uint32_t)P);
            if (RL > 0 && len == (uint32_t)RL && ph == (uint32_t)RPH) {
                // (a shape longer than x has no window: no hits)
                if constexpr (RL > 0)
                    if (len <= m)
                        overflow = sr_sweep_ring<P, RPH, RL>(ne_b, xoff, m, err, hits, a.hcap, lane, nq, qmeta, qw);
            } else switch (ph) {
            case 0: overflow = sr_sweep_shape<P, 0>(ne_b, xoff, m, len, err, hits, a.hcap, lane, nq); break;
            case 1: overflow = sr_sweep_shape<P, 1>(ne_b, xoff, m, len, err, hits, a.hcap, lane, nq); break;
            case 2: overflow = sr_sweep_shape<P, 2>(ne_b, xoff, m, len, err, hits, a.hcap, lane, nq); break;
            case 3: overflow = sr_sweep_shape<P, 3>(ne_b, xoff, m, len, err, hits, a.hcap, lane, nq); break;
            case 4: overflow = sr_sweep_shape<P, (P > 4 ? 4 : P)>(ne_b, xoff, m, len, err, hits, a.hcap, lane, nq); break;
            case 5: overflow = sr_sweep_shape<P, (P > 5 ? 5 : P)>(ne_b, xoff, m, len, err, hits, a.hcap, lane, nq); break;
            case 6: overflow = sr_sweep_shape<P, (P > 6 ? 6 : P)>(ne_b, xoff, m, len, err, hits, a.hcap, lane, nq); break;
            default: overflow = sr_sweep_shape<P, P>(ne_b, xoff, m, len, err, hits, a.hcap, lane, nq); break;
            }
            if (lane == 0) slot[1 + d] = nq;
        }
        if (lane == 0) slot[0] = overflow ? 1u : 0u;
        if (overflow) {
            // hand the unit's pairs to the per-pair kernel
            // pair-symmetric units hand over the pairs with x and the texts after it
            const uint32_t j_first = a.sym && I.xi > HR.j0 ? I.xi : HR.j0;
            const uint32_t nhand = HR.j0 + HR.ny - j_first;
            if (lane == 0) {
                const uint32_t at = atomicAdd(a.ovf_count, nhand);
                if (at + nhand > a.ovf_cap) atomicOr(a.ovf_count + 1, 1u);
                bcast[0] = at;
                // x against itself goes with them: the fold must not credit its diagonal
                if (I.xi >= HR.j0 && I.xi < HR.j0 + HR.ny) a.self_ovf[I.xi] = 1;
            }
            __syncthreads();
            const uint32_t at = bcast[0];
            for (uint32_t y = lane; y < nhand; y += 64) {
                if (at + y >= a.ovf_cap) break;
                const uint32_t j = j_first + y;
                const uint32_t ts = a.seq_start[j], n = a.seq_start[j + 1] - ts;
                // pair-symmetric units: the pair (j, x) goes along
                a.ovf[at + y] = StructItem{I.x_start, m, ts, n, I.uv_base, a.sym && j != I.xi ? a.uvb[j] : kNoSplit};
            }
        }
    }
}


// The register-ring kernels for shapes of length len in [kSrRingLenMin, kSrRingLenMax] with
// threshold plane ph <= 3, at most 4 counter planes over the launch's shapes (the caller
// checks): lengths 5-7 in struct_range_ring_short.hip, 8-10 in struct_range_ring_long.hip
// (hipErrorInvalidValue for the other's lengths).  lds: struct_range_sweep_lds's bytes; the
// kernels add their row queue.
hipError_t launch_struct_range_sweep_ring_short(const SrSweepArgs &a, uint32_t len, uint32_t ph, uint32_t grid,
                                                size_t lds, hipStream_t s);
hipError_t launch_struct_range_sweep_ring_long(const SrSweepArgs &a, uint32_t len, uint32_t ph, uint32_t grid,
                                               size_t lds, hipStream_t s);

template <int L>
inline hipError_t launch_sr_ring(const SrSweepArgs &a, uint32_t ph, uint32_t grid, size_t lds0, hipStream_t s)
{
    const size_t lds = ((lds0 + 15) & ~(size_t)15) + (size_t)kSrQueue * 20;
    switch (ph) {
    case 0: hipLaunchKernelGGL((struct_range_sweep<4, L, 0>), dim3(grid), dim3(64), lds, s, a); break;
    case 1: hipLaunchKernelGGL((struct_range_sweep<4, L, 1>), dim3(grid), dim3(64), lds, s, a); break;
    case 2: hipLaunchKernelGGL((struct_range_sweep<4, L, 2>), dim3(grid), dim3(64), lds, s, a); break;
    case 3: hipLaunchKernelGGL((struct_range_sweep<4, L, 3>), dim3(grid), dim3(64), lds, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace motex
