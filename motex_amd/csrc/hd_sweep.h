// The bit-sliced Hamming sweep kernel of gfx950 and its launch templates, included by
// hd_kernels.hip (the generic instantiations and the dispatch) and hd_ring_kernels*.hip (the
// register-ring instantiations per factor length).
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
// the mismatch count of the current k-window as P bit-sliced planes, biased so that the
// count is within e exactly when planes PH..P-1 are zero (hd_device.h); one row step adds the
// entering mismatch bits (NE_{x[px]} read at diagonal offset px, a funnel-shifted read of two
// LDS words) and subtracts the leaving ones with one up/down carry chain, so no counter ever
// moves between lanes.  The leaving bits of row px are the entering bits of row px - k: the
// kernels for k = 8 .. 16 keep them in a register ring, the generic ones read them again
// from LDS.  The clear bits of the high planes' OR are the row's 32 hit flags per lane.
// Where they are rare, lanes with flags queue them and a lane-parallel drain folds them into
// u and v (distinct texts per row for u, a global flag word per factor for texts split
// across work items); in the interior of a long sequence, where every cell of a row is a
// valid pair of one text, each row counts its hits in place.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>
#include "motex_internal.h"
#include "hd_device.h"

namespace motex {

// Folds the queued hit words of the current text range into the per-row accumulators.
// Entry q of the wave's queue holds the tag r | lane << 16 (row r of the item, lane of the
// wave) and the lane's B high-plane words of that row; a clear bit t of word b flags text
// position pt = q0 + 2048 b + 32 lane + t + px.  Each lane takes whole entries.  A word is split
// at the sequence starts it spans (found once per word by binary search over the range's
// starts in LDS); within one sequence y the flags that count are those inside the text
// range whose window lies entirely in y (pt >= start(y) + k - 1).  Their number goes to
// the row's v (LDS atomic add) and y to the row's sequence mask (LDS atomic OR).
//
// Pair-symmetric items (sym): texts j < xi are dropped (their pairs are computed with the
// roles swapped), and an occurrence in a text j > xi is also credited to the text side:
// v of (j, text position) += 1, and u of (j, text position) += 1 for the first occurrence of
// x_xi there.  The item runs every row of x_xi against this range, so the range's bitmap tb
// (bit pt - ys + 32 for text position pt, in LDS) tells the first occurrence exactly; first
// occurrences stay in the bitmap until the item ends (sym_flush), later ones add to v here.
template <int B>
__device__ __forceinline__ void drain_hits(const HdArgs &a, const HdRange &R, int32_t q0,
                                           uint32_t px_start, uint32_t xi, bool sym,
                                           const uint32_t *hq_meta, const uint32_t *hq_res,
                                           uint32_t nq, const uint32_t *starts, uint32_t *acc_v,
                                           uint32_t *ymask, uint32_t *tb, const uint8_t *ttab)
{
    __syncthreads();  // the queue stores of every lane are visible to the loads below
    const int32_t ys = (int32_t)R.ys, ye = (int32_t)R.ye, km1 = (int32_t)a.k - 1;
    // entries hold the high-plane words (a clear bit is a hit); each lane's next entry loads
    // while it works on the current one
    auto load_entry = [&](uint32_t q, uint32_t &meta, uint32_t *w) {
        meta = hq_meta[q];
        if constexpr (B == 4) {
            const uint4 v4 = *reinterpret_cast<const uint4 *>(hq_res + q * 4);
            w[0] = v4.x; w[1] = v4.y; w[2] = v4.z; w[3] = v4.w;
        } else if constexpr (B == 2) {
            const uint2 v2 = *reinterpret_cast<const uint2 *>(hq_res + q * 2);
            w[0] = v2.x; w[1] = v2.y;
        } else {
            w[0] = hq_res[q];
        }
    };
    uint32_t meta_n = 0, w_n[B];
    if (threadIdx.x < nq) load_entry(threadIdx.x, meta_n, w_n);
    for (uint32_t q = threadIdx.x; q < nq; q += 64) {
        const uint32_t meta = meta_n;
        uint32_t w[B];
#pragma unroll
        for (int b = 0; b < B; ++b) w[b] = ~w_n[b];
        if (q + 64 < nq) load_entry(q + 64, meta_n, w_n);
        const uint32_t r = meta & 0xFFFFu;
        const int32_t base0 = q0 + (int32_t)(32 * (meta >> 16) + px_start + r);
        uint32_t hits = 0, ym = 0;
        // the entry's blocks with hits, one per trip (almost always one): a loop over all B
        // blocks would run the body B times for the lanes of different blocks in turn
        uint32_t bmask = 0;
#pragma unroll
        for (int b = 0; b < B; ++b) bmask |= (w[b] ? 1u : 0u) << b;
        while (bmask) {
            const uint32_t b = __builtin_ctz(bmask);
            bmask &= bmask - 1;
            uint32_t wb = w[0];
#pragma unroll
            for (int bb = 1; bb < B; ++bb)
                if (b == (uint32_t)bb) wb = w[bb];
            const int32_t base = base0 + (int32_t)(kDiagPerBlock * b);
            const int32_t first = base > ys ? base : ys;
            // the first text to look at: the one holding the word of `first` (from the range's
            // table in ordered launches; a text at or before first's own, which the walk below
            // steps over) or a binary search over the starts
            uint32_t y = 0;
            if (R.ny > 1) {
                if (ttab) {
                    y = ttab[(uint32_t)(first - ys + 32) >> 5];
                } else {
                    // the last text starting at or before first: fixed halving steps over
                    // ny <= 32 texts (first may lie past the range's end: the last text then)
#pragma unroll
                    for (uint32_t step = 16; step; step >>= 1) {
                        const uint32_t idx = min(y + step, R.ny - 1);
                        if ((int32_t)starts[idx] <= first) y = idx;
                    }
                }
            }
            for (;;) {
                int32_t s1 = (int32_t)starts[y + 1];
                if (s1 > ye) s1 = ye;
                int32_t lo = (int32_t)starts[y] + km1;
                if (lo < first) lo = first;
                int32_t hi = base + 32 < s1 ? base + 32 : s1;
                const uint32_t j = R.j0 + y;
                if (lo < hi && (!sym || j >= xi)) {
                    const uint64_t span = ((1ull << (hi - base)) - 1) & ~((1ull << (lo - base)) - 1);
                    const uint32_t hv = wb & (uint32_t)span;
                    if (hv) {
                        hits += __builtin_popcount(hv);
                        ym |= 1u << y;
                        if (sym && j > xi) {
                            // text side: one v per occurrence, one u per position that x_xi
                            // reaches for the first time (the bitmap's word offset is >= 0:
                            // base >= ys - 31)
                            const uint32_t woff = (uint32_t)(base - ys + 32);
                            const uint32_t w0 = woff >> 5, sh = woff & 31;
                            const uint32_t lo_bits = hv << sh;
                            const uint32_t hi_bits = sh ? hv >> (32 - sh) : 0u;
                            uint32_t seen = 0;  // positions x_xi had reached before
                            if (lo_bits) seen |= (lo_bits & atomicOr(tb + w0, lo_bits)) >> sh;
                            if (hi_bits) seen |= (hi_bits & atomicOr(tb + w0 + 1, hi_bits)) << (32 - sh);
                            for (uint32_t h2 = seen; h2;) {
                                const uint32_t t = __builtin_ctz(h2);
                                h2 &= h2 - 1;
                                atomicAdd(a.v + (uint32_t)(base + (int32_t)t), 1u);
                            }
                        }
                    }
                }
                if (s1 >= base + 32 || ++y >= R.ny) break;
            }
        }
        if (hits) {
            atomicAdd(acc_v + r, hits);
            atomicOr(ymask + r, ym);
        }
    }
}

// End of a pair-symmetric item: appends the first occurrences of x_xi in range ri (the set
// bits of its bitmap tb, offsets pt - ys + 32) to the range's list, whose fold after the
// kernel credits each one u += 1 and v += 1 (sym_fold_kernel).  One wave-wide prefix sum and
// one counter atomic per range; a list that is full takes the rest as direct atomics.
__device__ __forceinline__ void sym_flush(const HdArgs &a, uint32_t ri, uint32_t ys, const uint32_t *tb)
{
    const uint32_t lane = threadIdx.x;
    uint32_t wd[4], n = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        wd[i] = tb[4 * lane + i];
        n += __builtin_popcount(wd[i]);
    }
    uint32_t incl = n;  // inclusive prefix sum over the wave's lanes
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= (uint32_t)d) incl += o;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    if (!total) return;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(a.tcnt + ri, total);
    base = __builtin_amdgcn_readfirstlane(base);
    const uint32_t off = a.toff[ri], cap = a.toff[ri + 1] - off;
    uint32_t idx = base + incl - n;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        for (uint32_t h = wd[i]; h; h &= h - 1, ++idx) {
            const uint32_t bit = 32 * (4 * lane + i) + __builtin_ctz(h);
            if (idx < cap) {
                a.tlist[off + idx] = (uint16_t)bit;
            } else {
                const uint32_t pos = ys + bit - 32;
                atomicAdd(a.u + pos, 1u);
                atomicAdd(a.v + pos, 1u);
            }
        }
}

// Persistent batched kernel: one wave per workgroup pulls items from the queue head.
//
// Hit handling is deferred and lane-parallel.  In a row where some lane has hit flags,
// the lanes with flags append their B hit words to the wave's queue in global memory
// (slots from the row's ballot, so the append costs a few instructions whatever the
// number of flags); the queue is drained (drain_hits) when fewer than 64 slots remain and
// at the end of each text range.  Keeping the queue out of LDS keeps the LDS footprint
// per wave (which bounds the number of resident waves) to the NE window and the per-row
// accumulators.  When a range is done, each row's sequence-mask popcount (the distinct
// sequences of that range with an occurrence) is added to the row's u; a range that is
// part of a long sequence marks u once.  The item's rows are flushed to global u/v at the
// end.  The first k rows of an item have no leaving column, so they run a step without it.
// Register budget: the LDS window of a B = 4 wave (about 6 KB for DNA) already limits a CU
// to 26 waves, so those shapes may use up to 80 VGPRs, which lets the compiler keep a row's
// 16 window words in flight at once (measured faster than 64 or 72); the smaller windows of
// B = 1, 2 fit 8 waves per SIMD (64 VGPRs) where the counters allow.
template <int B, int P, int PH, int K>
constexpr int hd_waves_per_eu()
{
    if constexpr (PH < 0) return 1;
    else if constexpr (K > 0) return 4;
    else if constexpr (B == 4) return P <= 4 ? 6 : 1;
    else if constexpr (B == 2) return P + PH <= 11 ? 8 : 1;
    else return 8;
}

template <int B, int P, int PH, int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(hd_waves_per_eu<B, P, PH, K>())))
void hd_batched_kernel(HdArgs a)
{
    extern __shared__ uint32_t smem[];
    const uint32_t lane = threadIdx.x;
    const uint32_t wl = a.wl;
    // LDS: NE window (sigma rows; at offset 0 so that the block offsets fit the read
    //      instructions' immediates) | per-row v, sequence mask | range sequence starts |
    //      per-row NE byte offsets (2 spare entries for the prefetch) | per-row u |
    //      text-side bitmaps
    uint32_t *ne_l = smem;
    uint32_t *acc_v = ne_l + a.sigma * wl;
    uint32_t *ymask = acc_v + a.rows_max;
    uint32_t *starts = ymask + a.rows_max;
    uint32_t *xb = starts + kMaxRangeSeqs + 1;
    // u counts distinct texts per row over all of the item's ranges: 32 bits, since an item
    // may span more texts than 16 bits can count (short sequences in large inputs)
    uint32_t *acc_u = xb + a.rows_max + 2;
    // pair-symmetric launches: 256 words per range of an item (a.sym_group ranges at most);
    // ordered launches: the range's word-to-text table in their place
    uint32_t *tbits = acc_u + a.rows_max;
    uint8_t *ttab_l = reinterpret_cast<uint8_t *>(tbits);
    const uint8_t *ttab = a.sym_group ? nullptr : ttab_l;
    const uint32_t lane4 = lane * 4;
    // this wave's hit queue: tags, then B words per entry (16-byte aligned)
    uint32_t *hq_meta = a.hitq + (size_t)blockIdx.x * a.hitq_cap * (B + 1);
    uint32_t *hq_res = hq_meta + a.hitq_cap;

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
        const uint32_t bias = hd_bias(a.e);
        // PH < 0 (factors longer than 64): the threshold plane is chosen at run time
        uint32_t hmask[P];
#pragma unroll
        for (int p = 0; p < P; ++p)
            hmask[p] = __builtin_amdgcn_readfirstlane((uint32_t)p >= hd_low_planes(a.e) ? ~0u : 0u);
        __syncthreads();  // previous item finished with LDS
        // pair-symmetric items: the text-side first-occurrence bitmaps of their ranges
        if (I.sym)
            for (uint32_t w = lane; w < (I.re - I.rb) * 256; w += 64) tbits[w] = 0;
        // row tiles of tile_rows reported rows (one, except on pair-symmetric items, which take
        // every row of x)
        for (uint32_t tr0 = I.r0; tr0 < I.r1; tr0 += a.tile_rows) {
            const uint32_t tr1 = min(I.r1, tr0 + a.tile_rows);
            const uint32_t px_start = tr0 - (a.k - 1);  // first computed row (k - 1 warm-up rows)
            const uint32_t nrows = tr1 - px_start;
            const uint32_t pxs5 = px_start >> 5;
            __syncthreads();  // previous tile finished with the row arrays
            // row r reads its window of NE_{x[px]} at byte offset xb[r] (+ 4 lane + 256 block)
            for (uint32_t r = lane; r < nrows + 2; r += 64) {
                if (r < nrows) {
                    const uint32_t px = px_start + r;
                    xb[r] = (a.codes[I.x_start + px] * wl + (px >> 5) - pxs5) * 4;
                    acc_v[r] = 0;
                    acc_u[r] = 0;
                    ymask[r] = 0;
                } else {
                    xb[r] = 0;
                }
            }
            const uint32_t rep0 = tr0;  // first reported row (tiles start at k - 1 or later)
            for (uint32_t ri = I.rb; ri < I.re; ++ri) {
                uint32_t *tb = tbits + (ri - I.rb) * 256;
                const HdRange R = a.ranges[ri];
                // first diagonal (pt - px) of this range, a multiple of 32
                const int32_t q0 = (((int32_t)R.ys - (int32_t)tr1 + 1) & ~31) + (int32_t)R.qoff;
                const uint32_t wbase = (uint32_t)(((q0 + (int32_t)kPadBits) >> 5) + (int32_t)pxs5);
                // A block of a long sequence whose every cell of a reported row is a valid pair
                // (text window inside the sequence: start + k - 1 <= pt < end): its rows count
                // their hits in place (popcount, one LDS add) instead of queueing them.
                const bool interior = I.split != kNoSplit &&
                                      (int64_t)rep0 + q0 >= (int64_t)R.ys + a.k - 1 &&
                                      (int64_t)tr1 - 1 + q0 + (int64_t)kDiagPerBlock * B <= (int64_t)R.ye;
                __syncthreads();
                for (uint32_t c = 0; c < a.sigma; ++c) {
                    const uint32_t *src = a.ne + (size_t)c * a.W + wbase;
                    for (uint32_t w = lane; w < wl; w += 64) ne_l[c * wl + w] = src[w];
                }
                if (lane <= R.ny) starts[lane] = a.seq_start[R.j0 + lane];
                __syncthreads();
                if (ttab && R.ny > 1) {
                    // word w of the range's positions starts at ys - 32 + 32 w: the last text
                    // starting at or before it (the first text for words before the range)
                    for (uint32_t w = lane; w < kTextTabEntries; w += 64) {
                        const int32_t p0 = (int32_t)R.ys - 32 + 32 * (int32_t)w;
                        uint32_t lo = 0, hi = R.ny - 1;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi + 1) >> 1;
                            if ((int32_t)starts[mid] <= p0) lo = mid; else hi = mid - 1;
                        }
                        ttab_l[w] = (uint8_t)lo;
                    }
                    __syncthreads();
                }

                uint32_t nq = 0;  // queued (row, lane) entries, wave-uniform
                uint32_t cnt[B][P];
#pragma unroll
                for (int b = 0; b < B; ++b) planes_set<P>(cnt[b], bias);

                // The NE words a row reads: for each block, the two words under its window for the
                // entering column (en) and, in the generic form, for the column leaving at r - k (lv).
                struct Win {
                    uint32_t en[B][2], lv[B][2];
                };
                // (xn, xo: xb of rows r and r - k)
                auto load_win = [&](uint32_t xn, uint32_t xo, Win &w, auto leave) {
                    const char *pn = reinterpret_cast<const char *>(ne_l) + (xn + lane4);
                    const char *po = reinterpret_cast<const char *>(ne_l) + (xo + lane4);
#pragma unroll
                    for (int b = 0; b < B; ++b) {
                        const uint32_t *wn = reinterpret_cast<const uint32_t *>(pn + 256 * b);
                        w.en[b][0] = wn[0];
                        w.en[b][1] = wn[1];
                        if constexpr (decltype(leave)::value) {
                            const uint32_t *wo = reinterpret_cast<const uint32_t *>(po + 256 * b);
                            w.lv[b][0] = wo[0];
                            w.lv[b][1] = wo[1];
                        }
                    }
                };
                // entering mismatch bits of row r (px = px_start + r) from its loaded words
                auto enter_bits = [&](uint32_t r, const Win &w, uint32_t *an) {
                    const uint32_t shn = (px_start + r) & 31;
#pragma unroll
                    for (int b = 0; b < B; ++b) an[b] = __builtin_amdgcn_alignbit(w.en[b][1], w.en[b][0], shn);
                };
                // one row: window counters of row r from its entering bits (and the bits of the
                // column leaving at r - k when LEAVE), then the row's hit flags into the queue
                // (the ring form appends a whole trip of up to 2 K rows before it drains: the
                // queue then holds hitq_cap >= 64 (2 K + 2) entries, and it is drained once more
                // than hitq_cap - 64 x trip rows are queued)
                // (rep: row r is reported, px >= rep0 -- a bool, or std::true_type where that is
                // known at compile time; inner: the range is interior)
                auto step = [&](uint32_t r, const uint32_t *an, const uint32_t *bo, auto leave, auto rep,
                                auto inner) {
                    uint32_t all = ~0u;  // lanes' words with no hit in any block
#pragma unroll
                    for (int b = 0; b < B; ++b) {
                        if constexpr (decltype(leave)::value) planes_slide<P>(cnt[b], an[b], bo[b]);
                        else planes_inc<P>(cnt[b], an[b]);
                        if constexpr (PH >= 0) all = planes_high_and<P, PH>(cnt[b], all);
                    }
                    // the blocks' high-plane words (clear bit = hit), needed on rows with hits
                    auto highs = [&](uint32_t *high) {
#pragma unroll
                        for (int b = 0; b < B; ++b) {
                            if constexpr (PH >= 0) {
                                high[b] = planes_high<P, PH>(cnt[b]);
                            } else {
                                uint32_t h = 0;
#pragma unroll
                                for (int p = 0; p < P; ++p) h |= cnt[b][p] & hmask[p];
                                high[b] = h;
                            }
                        }
                    };
                    if constexpr (PH < 0) {
                        uint32_t high[B];
                        highs(high);
#pragma unroll
                        for (int b = 0; b < B; ++b) all &= high[b];
                    }
                    if (rep) {
                        // (the range's kind is tested on rows with a hit only)
                        const uint64_t m = __ballot(all != ~0u);
                        if (m) {
                            if (inner) {
                                if (all != ~0u) {
                                    uint32_t high[B];
                                    highs(high);
                                    uint32_t c = 0;  // set high-plane bits: the lane's cells with no hit
#pragma unroll
                                    for (int b = 0; b < B; ++b) c += __builtin_popcount(high[b]);
                                    atomicAdd(acc_v + r, 32u * B - c);
                                }
                            } else {
                                if (all != ~0u) {
                                    uint32_t high[B];
                                    highs(high);
                                    // the entry holds the high-plane words (clear bit = hit); 32-bit
                                    // byte offsets from the wave's (scalar) queue bases
                                    const uint32_t slot =
                                        nq + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                                    char *res_b = reinterpret_cast<char *>(hq_res) + slot * (4u * B);
                                    if constexpr (B == 4) {
                                        *reinterpret_cast<uint4 *>(res_b) =
                                            make_uint4(high[0], high[1], high[2], high[3]);
                                    } else if constexpr (B == 2) {
                                        *reinterpret_cast<uint2 *>(res_b) = make_uint2(high[0], high[1]);
                                    } else {
                                        *reinterpret_cast<uint32_t *>(res_b) = high[0];
                                    }
                                    // (the trips of odd k from 13, 2 k rows long, leave no register
                                    // for the lane number: it is recomputed here, where the compiler
                                    // can neither hoist it nor spill it; k = 13, e = 3 on 1200 x
                                    // 1000 nt: 17.2 -> 14.8 ms.  For k = 9 and 11 the register is
                                    // there.)
                                    uint32_t tag_lane = lane;
                                    if constexpr (K % 2 == 1 && K >= 13)
                                        asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0"
                                                     : "=v"(tag_lane));
                                    *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hq_meta) + slot * 4u) =
                                        r | (tag_lane << 16);
                                }
                                nq += (uint32_t)__builtin_popcountll(m);
                                if constexpr (K == 0) {
                                    if (nq > a.hitq_cap - 64) {
                                        drain_hits<B>(a, R, q0, px_start, I.xi, I.sym != 0, hq_meta, hq_res, nq,
                                                      starts, acc_v, ymask, tb, ttab);
                                        nq = 0;
                                    }
                                }
                            }
                        }
                    }
                };
                if constexpr (K > 0) {
                    // The column leaving at row r is the one that entered at row r - K, whose bits
                    // the lane computed then: a ring of the last K rows' entering words (slot
                    // r mod K; every item has at least K rows, and the loop runs in trips of K
                    // rows from row K on, so every slot index is a constant; rows from k - 1 on are
                    // reported, and the trips start past them).
                    uint32_t ring[K][B];
                    // Rows are software-pipelined: while row r computes, the words of row r + 1
                    // load into the other buffer (row r uses buffer (r + K) & 1, so that trips of
                    // TRIP rows -- K, or 2 K for odd K -- start on buffer 0) and the offset of
                    // row r + 2 loads from LDS.  The scheduling barrier after each row keeps the
                    // compiler from hoisting later rows' loads (a register per word and row).
                    // xb has two spare entries past the last row.
                    constexpr int TRIP = (K % 2 == 0) ? K : 2 * K;
                    Win wbuf[2];
                    load_win(xb[0], 0u, wbuf[K & 1], std::false_type{});
                    uint32_t x1 = xb[1];
#pragma unroll
                    for (int s = 0; s < K; ++s) {
                        load_win(x1, 0u, wbuf[(s + 1 + K) & 1], std::false_type{});
                        x1 = xb[s + 2];
                        enter_bits(s, wbuf[(s + K) & 1], ring[s]);
                        step(s, ring[s], ring[s], std::false_type{}, s >= K - 1, interior);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    auto trip_drain = [&] {
                        if (nq > a.hitq_cap - 64 * TRIP) {
                            drain_hits<B>(a, R, q0, px_start, I.xi, I.sym != 0, hq_meta, hq_res, nq, starts, acc_v,
                                          ymask, tb, ttab);
                            nq = 0;
                        }
                    };
                    trip_drain();
                    // (the ring's slot for row r is r mod K: trips start at multiples of K)
                    auto ring_row = [&](uint32_t r, int s) {
                        load_win(x1, 0u, wbuf[(s + 1) & 1], std::false_type{});
                        x1 = xb[r + 2];
                        uint32_t an[B];
                        enter_bits(r, wbuf[s & 1], an);
                        step(r, an, ring[s % K], std::true_type{}, std::true_type{}, interior);
#pragma unroll
                        for (int b = 0; b < B; ++b) ring[s % K][b] = an[b];
                    };
                    uint32_t r = K;
                    for (; r + TRIP <= nrows; r += TRIP) {
#pragma unroll
                        for (int s = 0; s < TRIP; ++s) {
                            ring_row(r + s, s);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                        trip_drain();
                    }
#pragma unroll
                    for (int s = 0; s < TRIP - 1; ++s)
                        if (r + s < nrows) ring_row(r + s, s);
                } else {
                    // generic k: the leaving column is read again from the NE window
                    auto row = [&](uint32_t r, const Win &w, auto leave) {
                        uint32_t an[B], bo[B];
                        enter_bits(r, w, an);
                        if constexpr (decltype(leave)::value) {
                            const uint32_t sho = (px_start + r - a.k) & 31;
#pragma unroll
                            for (int b = 0; b < B; ++b) bo[b] = __builtin_amdgcn_alignbit(w.lv[b][1], w.lv[b][0], sho);
                        }
                        step(r, an, bo, leave, px_start + r >= rep0, interior);
                    };
                    const uint32_t warm = nrows < a.k ? nrows : a.k;
                    // two rows per trip: the counter planes then stay in place across the back edge
                    // (one row per trip leaves register copies of the planes in the loop).  The row
                    // offsets are loaded a trip ahead, so a row waits for one LDS round trip.
                    uint32_t r = 0;
                    Win wa, wb;
                    for (; r + 1 < warm; r += 2) {
                        load_win(xb[r], 0u, wa, std::false_type{});
                        load_win(xb[r + 1], 0u, wb, std::false_type{});
                        row(r, wa, std::false_type{});
                        row(r + 1, wb, std::false_type{});
                    }
                    if (r < warm) {
                        load_win(xb[r], 0u, wa, std::false_type{});
                        row(r, wa, std::false_type{});
                        ++r;
                    }
                    if (r < nrows) {
                        uint32_t n0 = xb[r], n1 = xb[r + 1], o0 = xb[r - a.k], o1 = xb[r + 1 - a.k];
                        for (; r + 1 < nrows; r += 2) {
                            const uint32_t m0 = xb[r + 2], m1 = xb[r + 3];
                            const uint32_t p0 = xb[r + 2 - a.k], p1 = xb[r + 3 - a.k];
                            load_win(n0, o0, wa, std::true_type{});
                            row(r, wa, std::true_type{});
                            load_win(n1, o1, wb, std::true_type{});
                            row(r + 1, wb, std::true_type{});
                            n0 = m0;
                            n1 = m1;
                            o0 = p0;
                            o1 = p1;
                        }
                        load_win(n0, o0, wa, std::true_type{});
                        if (r < nrows) row(r, wa, std::true_type{});
                    }
                }
                if (nq)
                    drain_hits<B>(a, R, q0, px_start, I.xi, I.sym != 0, hq_meta, hq_res, nq, starts, acc_v, ymask, tb,
                                  ttab);
                __syncthreads();
                // distinct sequences of this range per row
                for (uint32_t r = lane; r < nrows; r += 64) {
                    const uint32_t m = ymask[r];
                    if (m) {
                        if (I.split != kNoSplit) acc_u[r] = 1;
                        else acc_u[r] += __builtin_popcount(m);
                        ymask[r] = 0;
                    }
                }
            }
            __syncthreads();
            // flush the tile's rows; a long text split over several items counts once per row
            // (every count of an item over a long sequence is that sequence's, so u is dv != 0)
            for (uint32_t r = lane; r < nrows; r += 64) {
                const uint32_t dv = acc_v[r];
                if (!dv) continue;
                uint32_t du = I.split != kNoSplit ? 1u : acc_u[r];
                const uint32_t px = px_start + r;
                if (I.split != kNoSplit && du) {
                    const uint32_t xf = I.xf_base + px;
                    const uint32_t bit = 1u << (xf & 31);
                    const uint32_t old = atomicOr(a.split_flags + (size_t)I.split * a.split_words + (xf >> 5), bit);
                    du = (old & bit) ? 0u : 1u;
                }
                if (du) atomicAdd(a.u + I.uv_base + px, du);
                atomicAdd(a.v + I.uv_base + px, dv);
            }
        }  // row tiles
        if (I.sym) {
            __syncthreads();  // the drains' bitmap updates are visible
            for (uint32_t ri = I.rb; ri < I.re; ++ri)
                sym_flush(a, ri, a.ranges[ri].ys, tbits + (ri - I.rb) * 256);
        }
    }
}

// ------------------------------------------------------------------------------------
// Launch templates
// ------------------------------------------------------------------------------------

template <int B, int P, int PH, int K = 0>
inline hipError_t launch_bp(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    if constexpr (PH < P) {
        hipLaunchKernelGGL((hd_batched_kernel<B, P, PH, K>), dim3(grid), dim3(64), lds, s, a);
        return hipGetLastError();
    } else {
        return hipErrorInvalidValue;
    }
}

// Four-block kernels with the leaving column kept in a register ring, for factor length K
// and error bound E (counter planes and threshold plane fixed by them).
template <int K, int E>
inline hipError_t launch_ring_e(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    return launch_bp<4, (int)hd_counter_planes(K, E), (int)hd_low_planes(E), K>(a, grid, lds, s);
}

template <int K>
inline hipError_t launch_ring(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    switch (a.e) {
    case 0: return launch_ring_e<K, 0>(a, grid, lds, s);
    case 1: return launch_ring_e<K, 1>(a, grid, lds, s);
    case 2: return launch_ring_e<K, 2>(a, grid, lds, s);
    case 3: return launch_ring_e<K, 3>(a, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

// Ring kernels of factor lengths kRingKMin .. kRingKMax, one translation unit per pair of
// lengths (hd_ring_kernels_k*.hip); each returns hipErrorInvalidValue outside its lengths.
// (k = 16 needs five counter planes beside a ring of 64 words, and its ring kernels spilled:
// 1200 x 1000 nt, e = 2: 11.4 ms against the generic kernel's 8.2)
constexpr uint32_t kRingKMin = 8, kRingKMax = 15;
hipError_t launch_hd_ring_k8(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s);
hipError_t launch_hd_ring_k10(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s);
hipError_t launch_hd_ring_k12(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s);
hipError_t launch_hd_ring_k14(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s);

}  // namespace motex
