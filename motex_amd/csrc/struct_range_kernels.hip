// Structured motifs with Hamming boxes, range-batched (structured_motifs_extraction_hd,
// functions.cc:564-705).  A unit is every row of one x sequence against one text range: several
// whole texts packed into one 8192-diagonal window, as the Hamming kernel packs them.
//
// struct_range_sweep (one wave per unit, persistent): for every distinct box shape (length,
// errors) the bit-sliced window counters of hd_device.h run down the rows of x over 4 x 2048
// diagonals, the NE window in LDS; every hit goes to the unit's slot in HBM as
// px << 16 | (its diagonal - the window's first), shape by shape, rows in order.  Hits are rare
// (random DNA, 8 nt with one error: about 4 per row of 8192 diagonals), so a slot takes a few
// KB and the sweep needs no LDS beyond the NE window and the x table.
//
// struct_range_merge (one workgroup of 256 per unit), functions.cc:647-690: every shape's hits
// of the slot in LDS buckets of 16 text positions.  Each thread takes occurrences (px0, pt0) of
// box 0 (read from the slot) with the window inside one text j; box b's row is fixed by the
// x-side gap combination kx, px_b = px_{b-1} + S[kx][b-1] + len_b, and its occurrences are the
// hits of that row in pt_{b-1} + [gmin + len_b, gmax + len_b] inside text j (one chain per
// y-side combination), found in the buckets of that position window.  v[kx, px0] counts the
// chains; u[kx, px0] gets 1 per text j with a chain: the unit keeps the (kx, px0, j) already
// credited in an LDS key set.  With one gap, each box-1 hit of the window names its kx by its
// row, so every gap combination is found in one lookup.
//
// struct_range_fold: the merge's accumulators (v and u of a u/v index in one 64-bit word, for
// one atomic a chain) into u and v, adding x against itself's chains on its own diagonal.
//
// A unit whose hits do not fit its slot hands its pairs (x, j) to the per-pair kernel
// (struct_kernels.hip) through an overflow list, which a later launch consumes; its slot is
// marked and the merge skips it.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hd_device.h"
#include "motex_internal.h"
#include "struct_range_sweep.h"

#include <algorithm>
#include <climits>

namespace motex {

namespace {

// One unit as the merge reads it: the slot's hits (HBM) and every shape's hits in position
// buckets (LDS): entries px << 4 | (position & 15), bucket b of the running sequence over
// (shape, bucket of 16 positions) holding entries [C[b-1], C[b]).  Positions are pt - q0.
struct SrUnit {
    const uint32_t *hits;   // px << 16 | (pt - px - q0); shape d's hits in row order
    const uint32_t *ends;   // ends[d]: end of shape d's hits
    const uint16_t *bq;
    const uint32_t *C;
    uint32_t nbk;           // buckets per shape
    uint32_t m;             // rows of x
    uint32_t self;          // text of the range that is x itself, or kSrNoSelf
    int32_t diag;           // position - row on x's own diagonal (pt = x_start + px)
    const uint32_t *tuvb;   // per text of the range: its u/v index of (structure 0, position 0)
};

// First hit of shape d in row px or later.
__device__ uint32_t sr_row_lower(const SrUnit &U, uint32_t d, uint32_t px)
{
    uint32_t lo = d ? U.ends[d - 1] : 0u, hi = U.ends[d];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((U.hits[mid] >> 16) < px) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// f(px, position) for every hit of shape d at a position in [lo, hi].
template <class F>
__device__ __forceinline__ void sr_scan(const SrUnit &U, uint32_t d, int32_t lo, int32_t hi, F &&f)
{
    const uint32_t base = d * U.nbk;
    for (uint32_t bk = (uint32_t)lo >> 4; bk <= ((uint32_t)hi >> 4); ++bk) {
        const uint32_t b = base + bk;
        for (uint32_t e = b ? U.C[b - 1] : 0u, ee = U.C[b]; e < ee; ++e) {
            const uint32_t c = U.bq[e];
            const int32_t r = (int32_t)((bk << 4) | (c & 15u));
            if (r >= lo && r <= hi) f(c >> 4, r);
        }
    }
}

// Text of the range holding position r (starts[0 .. ny] are the range's text starts, ny <= 32,
// starts[0] <= r < starts[ny]): the last start at or before r, by five fixed halving steps
// (no loop control; an index past the last text reads the end, which is above r).
__device__ __forceinline__ uint32_t sr_text_of(const int32_t *starts, uint32_t ny, int32_t r)
{
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t step = 16; step; step >>= 1) {
        const uint32_t idx = min(lo + step, ny);
        if (starts[idx] <= r) lo = idx;
    }
    return lo;
}

// Complete chains of boxes B..nb_gaps for gap combination kx, given box B-1 ending at row pxp,
// position rp, in a text ending before rend: box B's row is fixed by kx, its occurrences lie in
// rp + [gmin + len_B, gmax + len_B] and inside the text.  Levels nest at compile time (one loop
// nest, no per-thread stack).
template <int B>
__device__ uint32_t sr_chains(const SrMergeArgs &a, const SrUnit &U, uint32_t kx, uint32_t pxp, int32_t rp,
                              int32_t rend)
{
    const uint32_t px = pxp + a.S[(size_t)kx * a.nb_gaps + B - 1] + a.blen[B];
    if (px >= U.m) return 0;
    const int32_t lo = rp + (int32_t)(a.gmin[B - 1] + a.blen[B]);
    const int32_t hi = min(rp + (int32_t)(a.gmax[B - 1] + a.blen[B]), rend - 1);
    if (lo > hi) return 0;
    uint32_t total = 0;
    sr_scan(U, a.dist_of[B], lo, hi, [&](uint32_t pxb, int32_t r) {
        if (pxb != px) return;
        if (B == (int)a.nb_gaps) {
            ++total;
        } else if constexpr (B < (int)kSrMaxGaps) {
            total += sr_chains<B + 1>(a, U, kx, px, r, rend);
        }
    });
    return total;
}

// u is credited once per (gap combination kx, row px0, text): the first chain of the three
// that reaches the unit's key set (LDS, open addressing, key (kx * m + px0) * 32 + text) takes
// it.  Returns 1 for a new key, 0 for a known one, 2 when the set is too full to say.
__device__ __forceinline__ uint32_t sr_key_insert(uint32_t *keys, uint32_t key)
{
    uint32_t h = (key * 2654435761u) >> (32 - kSrKeyBits);
    for (uint32_t probe = 0; probe < kSrKeyProbes; ++probe) {
        const uint32_t old = atomicCAS(keys + h, kSrNoKey, key);
        if (old == kSrNoKey) return 1;
        if (old == key) return 0;
        h = (h + 1) & ((1u << kSrKeyBits) - 1);
    }
    return 2;
}

// One gap: whether the chain (px0, r0) -> (px1, r1) is the least of its (kx, px0, text): no hit of
// row px1 before r1 completes the occurrence, and no occurrence of row px0 in the text with a
// smaller position has a chain for kx.  (The key set's fallback.)
__device__ bool sr_first_chain(const SrMergeArgs &a, const SrUnit &U, uint32_t kx, uint32_t px0, int32_t r0,
                               int32_t ts, int32_t te, uint32_t px1, int32_t r1);

template <int B>
__device__ uint32_t sr_chains(const SrMergeArgs &a, const SrUnit &U, uint32_t kx, uint32_t pxp, int32_t rp,
                              int32_t rend);

// Credits c chains of (kx, px0) from occurrence (px0, r0) in text ti: v always, u once per
// (kx, px0, text) -- the unit's key set, or its fallback when full.  One 64-bit atomic adds
// both (v in the low word).  x's own text takes no u here: the fold credits it, together with
// the chain on x's own diagonal.
__device__ __forceinline__ void sr_credit(const SrMergeArgs &a, const SrUnit &U, uint32_t *keys, uint32_t uv_base,
                                          uint32_t kx, uint32_t px0, int32_t r0, uint32_t ti, int32_t ts,
                                          int32_t te, uint32_t c, uint32_t px1, int32_t r1)
{
    unsigned long long add = c;
    if (ti != U.self) {
        const uint32_t k = sr_key_insert(keys, ((kx * U.m + px0) << 5) | ti);
        if (k == 1 || (k == 2 && sr_first_chain(a, U, kx, px0, r0, ts, te, px1, r1))) add += 1ull << 32;
    }
    if (add) atomicAdd(a.uv + (size_t)uv_base + (size_t)kx * U.m + px0, add);
}

// Every chain of boxes B..nb_gaps from box B-1 ending at (pxp, rp), inside a text ending before
// rend: the box-B hits of the position window rp + [gmin + len_B, gmax + len_B], from the
// buckets, whose rows lie in pxp + [gmin + len_B, gmax + len_B]; the row names box B's x-side
// gap and the position its text-side gap, which build the combination indexes kx and ky (last
// gap fastest).  f(kx, ky) for each complete chain but the one on x's own diagonal (dg: the
// chain so far lies on it).
template <int B, class F>
__device__ void sr_enum(const SrMergeArgs &a, const SrUnit &U, uint32_t pxp, int32_t rp, int32_t rend, uint32_t kxa,
                        uint32_t kya, bool dg, F &f)
{
    const uint32_t span = a.gmax[B - 1] - a.gmin[B - 1];
    const uint32_t glo = pxp + a.gmin[B - 1] + a.blen[B];
    const int32_t lo = rp + (int32_t)(a.gmin[B - 1] + a.blen[B]);
    const int32_t hi = min(rp + (int32_t)(a.gmax[B - 1] + a.blen[B]), rend - 1);
    if (lo > hi || glo >= U.m) return;
    sr_scan(U, a.dist_of[B], lo, hi, [&](uint32_t px, int32_t r) {
        const uint32_t g = px - glo;
        if (g > span) return;
        const uint32_t kx = kxa + g * a.gstride[B - 1];
        const uint32_t ky = kya + (uint32_t)(r - lo) * a.gstride[B - 1];
        const bool d = dg && r - (int32_t)px == U.diag;
        if (B == (int)a.nb_gaps) {
            if (!d) f(kx, ky);
        } else if constexpr (B < (int)kSrMaxGaps) {
            sr_enum<B + 1>(a, U, px, r, rend, kx, ky, d, f);
        }
    });
}

// Whether the t-th chain of occurrence (px0, r0) (enumeration order), of combination kx, is the
// first of its (kx, px0, text): no earlier chain of this occurrence has kx, and no occurrence of
// row px0 in the text with a smaller position has a chain for kx.  (The key set's fallback,
// several gaps.)
__device__ bool sr_first_chain_g(const SrMergeArgs &a, const SrUnit &U, uint32_t kx, uint32_t px0, int32_t r0,
                                 int32_t ts, int32_t te, uint32_t t)
{
    uint32_t seen = 0, before = 0;
    auto count = [&](uint32_t k, uint32_t) {
        if (seen++ < t && k == kx) ++before;
    };
    sr_enum<1>(a, U, px0, r0, te, 0u, 0u, false, count);
    if (before) return false;
    const uint32_t d0 = a.dist_of[0];
    const int32_t from = ts + (int32_t)a.blen[0] - 1;
    for (uint32_t h2 = sr_row_lower(U, d0, px0); h2 < U.ends[d0] && (U.hits[h2] >> 16) == px0; ++h2) {
        const int32_t r2 = (int32_t)(U.hits[h2] & 0xFFFFu) + (int32_t)px0;
        if (r2 < from || r2 >= r0) continue;
        if (sr_chains<1>(a, U, kx, px0, r2, te)) return false;
    }
    return true;
}

// The transposed pair's side of the same: whether the t-th chain of occurrence (px0, r0), of
// text-side combination ky, is the first of its (ky, r0): no earlier chain of this occurrence
// has ky, and no occurrence of box 0 at r0 in an earlier row has a chain with ky.
__device__ bool sr_first_chain_tg(const SrMergeArgs &a, const SrUnit &U, uint32_t ky, uint32_t px0, int32_t r0,
                                  int32_t te, uint32_t t)
{
    uint32_t seen = 0, before = 0;
    auto count = [&](uint32_t, uint32_t k) {
        if (seen++ < t && k == ky) ++before;
    };
    sr_enum<1>(a, U, px0, r0, te, 0u, 0u, false, count);
    if (before) return false;
    bool first = true;
    sr_scan(U, a.dist_of[0], r0, r0, [&](uint32_t q, int32_t) {
        if (!first || q >= px0 || q + 1 < a.blen[0]) return;
        auto any = [&](uint32_t, uint32_t k) {
            if (k == ky) first = false;
        };
        sr_enum<1>(a, U, q, r0, te, 0u, 0u, false, any);
    });
    return first;
}

// Occurrence of box 0 at row px0, position r0, in text ti = [ts, te) (positions), several gaps:
// each chain is one v of its combination; u once per (kx, px0, text) through the key set.  On
// pair-symmetric units a chain in a text j after x is also one of the transposed pair (j, x):
// box 0 at j's position r0 - ts, text-side combination ky; u of j once per (ky, r0).
template <bool SYM>
__device__ void sr_walk(const SrMergeArgs &a, const SrUnit &U, uint32_t *keys, uint32_t uv_base, uint32_t px0,
                        int32_t r0, uint32_t ti, int32_t ts, int32_t te)
{
    const bool diag = ti == U.self && r0 - (int32_t)px0 == U.diag;
    const bool both = SYM && ti != U.self;
    const uint32_t n = (uint32_t)(te - ts);
    uint32_t ord = 0;
    auto credit = [&](uint32_t kx, uint32_t ky) {
        const uint32_t t = ord++;
        unsigned long long add = 1;
        if (ti != U.self) {
            const uint32_t k = sr_key_insert(keys, ((kx * U.m + px0) << 5) | ti);
            if (k == 1 || (k == 2 && sr_first_chain_g(a, U, kx, px0, r0, ts, te, t))) add += 1ull << 32;
        }
        atomicAdd(a.uv + (size_t)uv_base + (size_t)kx * U.m + px0, add);
        if constexpr (SYM) {
            if (!both) return;
            const uint32_t k = sr_key_insert(keys, 0x80000000u | (ky << 14) | (uint32_t)r0);
            const unsigned long long tadd =
                1ull + ((k == 1 || (k == 2 && sr_first_chain_tg(a, U, ky, px0, r0, te, t))) ? 1ull << 32 : 0ull);
            atomicAdd(a.uv + (size_t)U.tuvb[ti] + (size_t)ky * n + (uint32_t)(r0 - ts), tadd);
        }
    };
    sr_enum<1>(a, U, px0, r0, te, 0u, 0u, diag, credit);
}

// Whether the chain (px0, r0) -> (px1, r1) is the first, in rows, of the transposed pair's
// (ky, r0): no occurrence of box 0 at r0 in an earlier row has a box-1 hit at r1 in its row
// window, and no box-1 hit at r1 lies in an earlier row of px0's window.  (The key set's
// fallback for the text side.)
__device__ bool sr_first_chain_t(const SrMergeArgs &a, const SrUnit &U, uint32_t px0, int32_t r0, uint32_t px1,
                                 int32_t r1)
{
    const uint32_t lo_g = a.blen[1] + a.gmin[0], hi_g = a.blen[1] + a.gmax[0];
    bool first = true;
    sr_scan(U, a.dist_of[1], r1, r1, [&](uint32_t p, int32_t) {
        if (p >= px0 + lo_g && p < px1) first = false;
    });
    sr_scan(U, a.dist_of[0], r0, r0, [&](uint32_t q, int32_t) {
        if (!first || q >= px0 || q + 1 < a.blen[0]) return;
        sr_scan(U, a.dist_of[1], r1, r1, [&](uint32_t p, int32_t) {
            if (p >= q + lo_g && p <= q + hi_g && p < U.m) first = false;
        });
    });
    return first;
}

// Occurrence of box 0, one gap: box 1 ends in row px1 = px0 + len1 + gmin + kx at a position
// in r0 + [gmin + len1, gmax + len1], inside the text.  The buckets of that window give every
// such hit, and each is one chain of kx (but the one on x's own diagonal).  With pair-symmetric
// units a chain in a text j after x is also one of the transposed pair (j, x): box 0 at j's
// position r0 - ts, ky = r1 - r0 - len1 - gmin; u of j once per (ky, r0) in the unit.
template <bool SYM>
__device__ void sr_join(const SrMergeArgs &a, const SrUnit &U, uint32_t *keys, uint32_t uv_base, uint32_t px0,
                        int32_t r0, uint32_t ti, int32_t ts, int32_t te)
{
    const uint32_t len1 = a.blen[1], gmn = a.gmin[0];
    const uint32_t span = a.gmax[0] - gmn;
    if (px0 + len1 + gmn >= U.m) return;
    const int32_t lo = r0 + (int32_t)(gmn + len1), hi = min(r0 + (int32_t)(a.gmax[0] + len1), te - 1);
    if (lo > hi) return;
    const bool diag = ti == U.self && r0 - (int32_t)px0 == U.diag;
    const bool both = SYM && ti != U.self;
    sr_scan(U, a.dist_of[1], lo, hi, [&](uint32_t px1, int32_t r1) {
        const uint32_t kx = px1 - px0 - len1 - gmn;
        if (kx > span || (diag && r1 - (int32_t)px1 == U.diag)) return;
        sr_credit(a, U, keys, uv_base, kx, px0, r0, ti, ts, te, 1, px1, r1);
        if (SYM && both) {
            const uint32_t ky = (uint32_t)(r1 - lo);
            const uint32_t n = (uint32_t)(te - ts);
            const uint32_t k = sr_key_insert(keys, 0x80000000u | (ky << 14) | (uint32_t)r0);
            const unsigned long long add =
                1ull + ((k == 1 || (k == 2 && sr_first_chain_t(a, U, px0, r0, px1, r1))) ? 1ull << 32 : 0ull);
            atomicAdd(a.uv + (size_t)U.tuvb[ti] + (size_t)ky * n + (uint32_t)(r0 - ts), add);
        }
    });
}

__device__ bool sr_first_chain(const SrMergeArgs &a, const SrUnit &U, uint32_t kx, uint32_t px0, int32_t r0,
                               int32_t ts, int32_t te, uint32_t px1, int32_t r1)
{
    (void)kx;
    const uint32_t d0 = a.dist_of[0];
    const int32_t from = ts + (int32_t)a.blen[0] - 1;
    const uint32_t len1 = a.blen[1];
    const int32_t off_lo = (int32_t)(a.gmin[0] + len1), off_hi = (int32_t)(a.gmax[0] + len1);
    bool first = true;
    sr_scan(U, a.dist_of[1], r0 + off_lo, r1 - 1, [&](uint32_t p, int32_t) {
        if (p == px1) first = false;
    });
    for (uint32_t h2 = sr_row_lower(U, d0, px0); first && h2 < U.ends[d0] && (U.hits[h2] >> 16) == px0; ++h2) {
        const int32_t r2 = (int32_t)(U.hits[h2] & 0xFFFFu) + (int32_t)px0;
        if (r2 < from || r2 >= r0) continue;
        sr_scan(U, a.dist_of[1], r2 + off_lo, min(r2 + off_hi, te - 1), [&](uint32_t p, int32_t) {
            if (p == px1) first = false;
        });
    }
    return first;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Merge
// ---------------------------------------------------------------------------------------

template <bool ONE_GAP, bool SYM>
__global__ __launch_bounds__(kSrMergeThreads) void struct_range_merge(SrMergeArgs a)
{
    extern __shared__ uint32_t smem[];
    const uint32_t tid = threadIdx.x;
    const uint32_t *slot = a.slots + (size_t)blockIdx.x * (kSrHead + a.hcap);
    if (slot[0]) return;  // overflowed: the per-pair kernel takes its pairs
    const SrItem I = a.items[a.first + blockIdx.x];
    const HdRange HR = a.ranges[I.r];
    const uint32_t m = I.m;
    const int32_t q0 = sr_window_q0((int32_t)HR.ys, m);
    // LDS: text starts (kMaxRangeSeqs + 1, as positions pt - q0) | texts' u/v bases
    //      (kMaxRangeSeqs) | shape ends (kSrMaxBoxes) | u key set (2^kSrKeyBits) | bucket ends
    //      (ndist x nbk) | bucket entries (u16, hcap)
    int32_t *starts = reinterpret_cast<int32_t *>(smem);
    uint32_t *tuvb = smem + kMaxRangeSeqs + 1;
    uint32_t *ends = tuvb + kMaxRangeSeqs;
    uint32_t *keys = ends + kSrMaxBoxes;
    uint32_t *C = keys + (1u << kSrKeyBits);
    const uint32_t nbc = a.ndist * a.nbk;
    uint16_t *bq = reinterpret_cast<uint16_t *>(C + nbc);
    const uint32_t *hits = slot + kSrHead;
    if (tid <= HR.ny) starts[tid] = (int32_t)a.seq_start[HR.j0 + tid] - q0;
    if (tid < HR.ny) tuvb[tid] = a.uvb[HR.j0 + tid];
    if (tid < a.ndist) ends[tid] = slot[1 + tid];
    for (uint32_t b = tid; b < nbc; b += kSrMergeThreads) C[b] = 0;
    for (uint32_t k = tid; k < (1u << kSrKeyBits); k += kSrMergeThreads) keys[k] = kSrNoKey;
    __syncthreads();
    const uint32_t nh = ends[a.ndist - 1];
    // every shape's hits into position buckets: count, scan, scatter
    for (uint32_t h = tid; h < nh; h += kSrMergeThreads) {
        const uint32_t hv = hits[h];
        uint32_t d = 0;
        while (h >= ends[d]) ++d;
        atomicAdd(C + d * a.nbk + (((hv & 0xFFFFu) + (hv >> 16)) >> 4), 1u);
    }
    __syncthreads();
    if (tid < 64) {
        const uint32_t per = (nbc + 63) / 64, c0 = min(tid * per, nbc), c1 = min(c0 + per, nbc);
        uint32_t sum = 0;
        for (uint32_t c = c0; c < c1; ++c) sum += C[c];
        uint32_t incl = sum;
#pragma unroll
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o);
            if (tid >= o) incl += t;
        }
        uint32_t run = incl - sum;
        for (uint32_t c = c0; c < c1; ++c) {
            const uint32_t n = C[c];
            C[c] = run;
            run += n;
        }
    }
    __syncthreads();
    for (uint32_t h = tid; h < nh; h += kSrMergeThreads) {
        const uint32_t hv = hits[h];
        uint32_t d = 0;
        while (h >= ends[d]) ++d;
        const uint32_t px = hv >> 16, r = (hv & 0xFFFFu) + px;
        bq[atomicAdd(C + d * a.nbk + (r >> 4), 1u)] = (uint16_t)((px << 4) | (r & 15u));
    }
    __syncthreads();

    // occurrences of box 0
    const uint32_t self = I.xi >= HR.j0 && I.xi < HR.j0 + HR.ny ? I.xi - HR.j0 : kSrNoSelf;
    const SrUnit U{hits, ends, bq, C, a.nbk, m, self, (int32_t)I.x_start - q0, tuvb};
    // pair-symmetric units: the texts before x are x's pairs with earlier sequences, whose
    // units hold them
    const uint32_t t_first = a.sym && I.xi > HR.j0 ? I.xi - HR.j0 : 0u;
    const int32_t ys = starts[0], ye = starts[HR.ny];
    const uint32_t len0 = a.blen[0];
    const uint32_t d0 = a.dist_of[0];
    for (uint32_t h = (d0 ? ends[d0 - 1] : 0u) + tid, he = ends[d0]; h < he; h += kSrMergeThreads) {
        const uint32_t hv = hits[h];
        const uint32_t px0 = hv >> 16;
        const int32_t r0 = (int32_t)(hv & 0xFFFFu) + (int32_t)px0;
        if (r0 < ys || r0 >= ye) continue;
        const uint32_t ti = sr_text_of(starts, HR.ny, r0);
        const int32_t ts = starts[ti], te = starts[ti + 1];
        if (ti < t_first || r0 - (int32_t)len0 + 1 < ts) continue;  // box 0's window inside the text
        if constexpr (ONE_GAP)
            sr_join<SYM>(a, U, keys, I.uv_base, px0, r0, ti, ts, te);
        else
            sr_walk<SYM>(a, U, keys, I.uv_base, px0, r0, ti, ts, te);
    }
}

// ---------------------------------------------------------------------------------------
// Fold: the merge's accumulators into u and v, with x against itself's all-diagonal chains
// ---------------------------------------------------------------------------------------

// x against itself has, for every gap combination kx and row px0 whose structure fits in x
// (box 0 ends at px0 >= len0 - 1, the last box at px0 + sum(S[kx][g] + len_{g+1}) < m), the
// chain that lies on the diagonal -- every box a 0-error occurrence of itself.  It is one chain
// of v and, x being one text, all of u for that text; the merge leaves both to this pass, unless
// the unit holding the pair overflowed (then the per-pair kernel computes the pair whole).
__global__ __launch_bounds__(256) void struct_range_fold(SrFoldArgs a)
{
    const SrX X = a.xs[blockIdx.x];
    const bool self = a.self_ovf[X.xi] == 0;
    const uint32_t n = a.nb_structs * X.m;
    for (uint32_t e = threadIdx.x; e < n; e += 256) {
        const uint32_t kx = e / X.m, px = e - kx * X.m;
        uint32_t end = px;
        for (uint32_t g = 0; g < a.nb_gaps; ++g) end += a.S[(size_t)kx * a.nb_gaps + g] + a.blen[g + 1];
        const uint32_t d = (self && px + 1 >= a.blen[0] && end < X.m) ? 1u : 0u;
        const size_t o = (size_t)X.uv_base + e;
        const unsigned long long w = a.uv[o];
        a.v[o] += (uint32_t)w + d;
        a.u[o] += (uint32_t)(w >> 32) + d;
    }
}

hipError_t launch_struct_range_fold(const SrFoldArgs &a, uint32_t nx, hipStream_t s)
{
    hipLaunchKernelGGL(struct_range_fold, dim3(nx), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_struct_range_sweep(const SrSweepArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    uint32_t planes = 1;
    for (uint32_t d = 0; d < a.ndist; ++d) {
        const uint32_t p = hd_counter_planes(a.dlen[d], a.derr[d]);
        planes = p > planes ? p : planes;
    }
    const char *env = getenv("MOTEX_SR_RING");  // "0": the generic sweep for every shape
    const bool ring_on = !(env && env[0] == '0');
    // the first shape with a length the ring kernels take sweeps in a register ring
    const bool ring_fits = ((lds + 15) & ~(size_t)15) + (size_t)kSrQueue * 20 <= 65536;  // with the row queue
    for (uint32_t d = 0; ring_on && ring_fits && planes <= 4 && d < a.ndist; ++d) {
        const uint32_t len = a.dlen[d], ph = std::min(hd_low_planes(a.derr[d]), 4u);
        if (len < kSrRingLenMin || len > kSrRingLenMax || ph > 3) continue;
        return len <= 7 ? launch_struct_range_sweep_ring_short(a, len, ph, grid, lds, s)
                        : launch_struct_range_sweep_ring_long(a, len, ph, grid, lds, s);
    }
    switch (planes <= 3 ? 3 : planes) {
    case 3: hipLaunchKernelGGL((struct_range_sweep<3>), dim3(grid), dim3(64), lds, s, a); break;
    case 4: hipLaunchKernelGGL((struct_range_sweep<4>), dim3(grid), dim3(64), lds, s, a); break;
    case 5: hipLaunchKernelGGL((struct_range_sweep<5>), dim3(grid), dim3(64), lds, s, a); break;
    case 6: hipLaunchKernelGGL((struct_range_sweep<6>), dim3(grid), dim3(64), lds, s, a); break;
    case 7: hipLaunchKernelGGL((struct_range_sweep<7>), dim3(grid), dim3(64), lds, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_struct_range_merge(const SrMergeArgs &a, uint32_t units, size_t lds, hipStream_t s)
{
    const dim3 grid(units), block(kSrMergeThreads);
    if (a.nb_gaps == 1 && a.sym)
        hipLaunchKernelGGL((struct_range_merge<true, true>), grid, block, lds, s, a);
    else if (a.nb_gaps == 1)
        hipLaunchKernelGGL((struct_range_merge<true, false>), grid, block, lds, s, a);
    else if (a.sym)
        hipLaunchKernelGGL((struct_range_merge<false, true>), grid, block, lds, s, a);
    else
        hipLaunchKernelGGL((struct_range_merge<false, false>), grid, block, lds, s, a);
    return hipGetLastError();
}

// LDS bytes of a sweep workgroup and of a merge workgroup (the layouts at the top of each).
size_t struct_range_sweep_lds(uint32_t sigma, uint32_t wl, uint32_t m_max)
{
    return ((size_t)sigma * wl + (m_max + 3) / 2 + 1) * 4;
}

size_t struct_range_merge_lds(uint32_t hcap, uint32_t ndist, uint32_t nbk)
{
    return ((size_t)2 * kMaxRangeSeqs + 1 + kSrMaxBoxes + (1u << kSrKeyBits) + (size_t)ndist * nbk + (hcap + 1) / 2) *
           4;
}

}  // namespace motex
