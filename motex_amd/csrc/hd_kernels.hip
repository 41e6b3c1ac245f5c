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
// the mismatch count of the current k-window as P bit-sliced planes, biased so that the
// count is within e exactly when planes PH..P-1 are zero (hd_device.h); one row step adds the
// entering mismatch bits (NE_{x[px]} read at diagonal offset px) and subtracts the leaving
// ones (NE_{x[px-k]} at offset px-k) with one up/down carry chain.  Both are funnel-shifted
// reads of two LDS words, so no counter ever moves between lanes.  The clear bits of the
// high planes' OR are the row's 32 hit flags per lane; hits are rare, so a wave-uniform slow
// path folds them into u and v
// (distinct texts per row for u, a global flag word per factor for texts split across
// work items).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include "motex_internal.h"
#include "hd_device.h"

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
// v of (j, text position) += 1 and the (xi, text position) bit of fbits is set, from which
// hd_fold_fbits adds the distinct-x count to u after the kernel.
template <int B>
__device__ __forceinline__ void drain_hits(const HdArgs &a, const HdRange &R, int32_t q0,
                                           uint32_t px_start, uint32_t xi, bool sym,
                                           const uint32_t *hq_meta, const uint32_t *hq_res,
                                           uint32_t nq, const uint32_t *starts, uint32_t *acc_v,
                                           uint32_t *ymask)
{
    __syncthreads();  // the queue stores of every lane are visible to the loads below
    const int32_t ys = (int32_t)R.ys, ye = (int32_t)R.ye, km1 = (int32_t)a.k - 1;
    for (uint32_t q = threadIdx.x; q < nq; q += 64) {
        const uint32_t meta = hq_meta[q];
        uint32_t w[B];
        // entries hold the high-plane words: a clear bit is a hit
        if constexpr (B == 4) {
            const uint4 v4 = *reinterpret_cast<const uint4 *>(hq_res + q * 4);
            w[0] = ~v4.x; w[1] = ~v4.y; w[2] = ~v4.z; w[3] = ~v4.w;
        } else if constexpr (B == 2) {
            const uint2 v2 = *reinterpret_cast<const uint2 *>(hq_res + q * 2);
            w[0] = ~v2.x; w[1] = ~v2.y;
        } else {
            w[0] = ~hq_res[q];
        }
        const uint32_t r = meta & 0xFFFFu;
        const int32_t base0 = q0 + (int32_t)(32 * (meta >> 16) + px_start + r);
        uint32_t hits = 0, ym = 0;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if (!w[b]) continue;
            const int32_t base = base0 + (int32_t)kDiagPerBlock * b;
            const int32_t first = base > ys ? base : ys;
            uint32_t y = 0;
            if (R.ny > 1) {
                uint32_t lo = 0, hi = R.ny - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if ((int32_t)starts[mid] <= first) lo = mid; else hi = mid - 1;
                }
                y = lo;
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
                    const uint32_t hv = w[b] & (uint32_t)span;
                    if (hv) {
                        hits += __builtin_popcount(hv);
                        ym |= 1u << y;
                        if (sym && j > xi) {
                            // text side: one v per occurrence, one bit per (xi, position)
                            for (uint32_t h2 = hv; h2;) {
                                const uint32_t t = __builtin_ctz(h2);
                                h2 &= h2 - 1;
                                atomicAdd(a.v + (uint32_t)(base + (int32_t)t), 1u);
                            }
                            // the flags' positions (all >= 0) span words w0 and w0 + 1 of
                            // fbits; w0 is -1 when base < 0, and then holds none of them
                            const int32_t w0 = base >> 5;
                            const uint32_t sh = (uint32_t)base & 31;
                            const uint32_t lo_bits = hv << sh;
                            const uint32_t hi_bits = sh ? hv >> (32 - sh) : 0u;
                            uint32_t *frow = a.fbits + (size_t)xi * a.f_words;
                            if (lo_bits) atomicOr(frow + w0, lo_bits);
                            if (hi_bits) atomicOr(frow + (w0 + 1), hi_bits);
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
template <int B, int P, int PH>
constexpr int hd_waves_per_eu()
{
    if constexpr (PH < 0) return 1;
    else if constexpr (B == 4) return P <= 4 ? 6 : 1;
    else if constexpr (B == 2) return P + PH <= 11 ? 8 : 1;
    else return 8;
}

template <int B, int P, int PH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(hd_waves_per_eu<B, P, PH>())))
void hd_batched_kernel(HdArgs a)
{
    extern __shared__ uint32_t smem[];
    const uint32_t lane = threadIdx.x;
    const uint32_t wl = a.wl;
    // LDS: NE window (sigma rows; at offset 0 so that the block offsets fit the read
    //      instructions' immediates) | per-row v, sequence mask | range sequence starts |
    //      per-row NE byte offsets (2 spare entries for the prefetch) | per-row u
    uint32_t *ne_l = smem;
    uint32_t *acc_v = ne_l + a.sigma * wl;
    uint32_t *ymask = acc_v + a.rows_max;
    uint32_t *starts = ymask + a.rows_max;
    uint32_t *xb = starts + kMaxRangeSeqs + 1;
    // u counts distinct texts per row over all of the item's ranges: 32 bits, since an item
    // may span more texts than 16 bits can count (short sequences in large inputs)
    uint32_t *acc_u = xb + a.rows_max + 2;
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
        const uint32_t nrows = I.r1 - I.px_start;
        const uint32_t pxs5 = I.px_start >> 5;
        __syncthreads();  // previous item finished with LDS
        // row r reads its window of NE_{x[px]} at byte offset xb[r] (+ 4 lane + 256 block)
        for (uint32_t r = lane; r < nrows + 2; r += 64) {
            if (r < nrows) {
                const uint32_t px = I.px_start + r;
                xb[r] = (a.codes[I.x_start + px] * wl + (px >> 5) - pxs5) * 4;
                acc_v[r] = 0;
                acc_u[r] = 0;
                ymask[r] = 0;
            } else {
                xb[r] = 0;
            }
        }
        const uint32_t rep0 = I.r0 > a.k - 1 ? I.r0 : a.k - 1;
        const uint32_t bias = hd_bias(a.e);
        // PH < 0 (factors longer than 64): the threshold plane is chosen at run time
        uint32_t hmask[P];
#pragma unroll
        for (int p = 0; p < P; ++p)
            hmask[p] = __builtin_amdgcn_readfirstlane((uint32_t)p >= hd_low_planes(a.e) ? ~0u : 0u);
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

            uint32_t nq = 0;  // queued (row, lane) entries, wave-uniform
            uint32_t cnt[B][P];
#pragma unroll
            for (int b = 0; b < B; ++b) planes_set<P>(cnt[b], bias);

            // The NE words a row reads: for each block, the two words under its window for the
            // entering column (en) and for the column leaving at r - k (lv).
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
            // one row: window counters of row r (with the column leaving at r - k when
            // LEAVE) from its loaded words, then the row's hit flags into the queue
            auto row = [&](uint32_t r, const Win &w, auto leave) {
                const uint32_t px = I.px_start + r;
                const uint32_t shn = px & 31, sho = (px - a.k) & 31;
                uint32_t high[B];
                uint32_t all = ~0u;  // lanes' words with no hit in any block
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const uint32_t an = __builtin_amdgcn_alignbit(w.en[b][1], w.en[b][0], shn);
                    if constexpr (decltype(leave)::value) {
                        const uint32_t bo = __builtin_amdgcn_alignbit(w.lv[b][1], w.lv[b][0], sho);
                        planes_slide<P>(cnt[b], an, bo);
                    } else {
                        planes_inc<P>(cnt[b], an);
                    }
                    if constexpr (PH >= 0) {
                        high[b] = planes_high<P, PH>(cnt[b]);
                    } else {
                        uint32_t h = 0;
#pragma unroll
                        for (int p = 0; p < P; ++p) h |= cnt[b][p] & hmask[p];
                        high[b] = h;
                    }
                    all &= high[b];
                }
                if (px >= rep0) {
                    const uint64_t m = __ballot(all != ~0u);
                    if (m) {
                        if (all != ~0u) {
                            // the entry holds the high-plane words (clear bit = hit); 32-bit
                            // byte offsets from the wave's (scalar) queue bases
                            const uint32_t slot =
                                nq + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                            char *res_b = reinterpret_cast<char *>(hq_res) + slot * (4u * B);
                            if constexpr (B == 4) {
                                *reinterpret_cast<uint4 *>(res_b) = make_uint4(high[0], high[1], high[2], high[3]);
                            } else if constexpr (B == 2) {
                                *reinterpret_cast<uint2 *>(res_b) = make_uint2(high[0], high[1]);
                            } else {
                                *reinterpret_cast<uint32_t *>(res_b) = high[0];
                            }
                            *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hq_meta) + slot * 4u) = r | (lane << 16);
                        }
                        nq += (uint32_t)__builtin_popcountll(m);
                        if (nq > a.hitq_cap - 64) {
                            drain_hits<B>(a, R, q0, I.px_start, I.xi, I.sym != 0, hq_meta, hq_res, nq, starts, acc_v, ymask);
                            nq = 0;
                        }
                    }
                }
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
            if (nq) drain_hits<B>(a, R, q0, I.px_start, I.xi, I.sym != 0, hq_meta, hq_res, nq, starts, acc_v, ymask);
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
        // flush the item's rows; a long text split over several items counts once per row
        for (uint32_t r = lane; r < nrows; r += 64) {
            const uint32_t dv = acc_v[r];
            if (!dv) continue;
            uint32_t du = acc_u[r];
            const uint32_t px = I.px_start + r;
            if (I.split != kNoSplit && du) {
                const uint32_t xf = I.xf_base + px;
                const uint32_t bit = 1u << (xf & 31);
                const uint32_t old = atomicOr(a.split_flags + (size_t)I.split * a.split_words + (xf >> 5), bit);
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

template <int B, int P, int PH>
static hipError_t launch_bp(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    if constexpr (PH < P) {
        hipLaunchKernelGGL((hd_batched_kernel<B, P, PH>), dim3(grid), dim3(64), lds, s, a);
        return hipGetLastError();
    } else {
        return hipErrorInvalidValue;
    }
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
    default: return launch_p<4>(a, grid, lds, s);
    }
}

}  // namespace motex
