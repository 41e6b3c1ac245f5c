// Structured-motif kernels for gfx950 (structured_motifs_extraction_{hd,ed},
// functions.cc:564-705 and 773-915).
//
// One persistent wave per workgroup takes one (x, y) pair at a time:
//   1. per distinct box shape (length, errors) the FLASM hits of the pair go into a bitmap
//      in the wave's scratch, word-major: bm[(pos_t >> 5) * m + pos_x], bit pos_t & 31.
//      Hamming boxes use the diagonal sweep of the Hamming kernel (bit-sliced window
//      counts; hits are rare and set bits with atomic OR); edit boxes use the strip
//      wavefront of the edit kernel in bitmap mode (every word written once).  A box whose
//      error bound is at least m passes everywhere, as the reference's B matrices, which
//      hold m where FLASM returned no tuple (functions.cc:597-598), do.
//      Hamming hits also go to a per-shape occurrence list in scratch, so that the merge
//      needs no scan of the bitmaps and the pair clears only the words it set.
//   2. the merge of functions.cc:647-690: every box-0 hit (pos_t, pos_x) is taken from its
//      list (or, for edit boxes and overflowing lists, gathered from the bitmap into an LDS
//      list); for every x-side gap combination kx the y-side combinations that hit
//      every box are counted -- explicit loops over all but the last gap, a popcount of a
//      bit window of the last box's column for the last one.  v[kx*m + pos_x] adds the
//      count; u[kx*m + pos_x] adds 1 once per pair, deduplicated with a per-pair bitmap of
//      (kx, pos_x) entries already counted.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "ed_device.h"
#include "hd_device.h"
#include "motex_internal.h"

namespace motex {

__device__ __forceinline__ uint32_t bm_bit(const uint32_t *bm, uint32_t m, uint32_t px, uint32_t pt)
{
    return (bm[(size_t)(pt >> 5) * m + px] >> (pt & 31)) & 1u;
}

// Set bits of column px over text positions [lo, hi] (inclusive).
__device__ __forceinline__ uint32_t bm_window(const uint32_t *bm, uint32_t m, uint32_t px,
                                              uint32_t lo, uint32_t hi)
{
    uint32_t cnt = 0;
    for (uint32_t w = lo >> 5; w <= (hi >> 5); ++w) {
        uint32_t word = bm[(size_t)w * m + px];
        uint32_t b0 = w == (lo >> 5) ? (lo & 31) : 0;
        uint32_t b1 = w == (hi >> 5) ? (hi & 31) : 31;
        uint32_t mask = (b1 == 31 ? ~0u : ((1u << (b1 + 1)) - 1)) & ~((1u << b0) - 1);
        cnt += __builtin_popcount(word & mask);
    }
    return cnt;
}

// Number of y-side gap combinations that complete the structured motif whose box 0 ends
// at (pos_x = px0, pos_t = pt0) with x-side combination kx (functions.cc:653-685).
template <int NB>
__device__ uint32_t struct_walk(const StructArgs &a, uint32_t *const bm[NB], uint32_t m,
                                uint32_t n, uint32_t kx, uint32_t px0, uint32_t pt0)
{
    const uint32_t G = a.nb_gaps;
    uint32_t jx[NB - 1];
    uint32_t px = px0;
    for (uint32_t g = 0; g < G; ++g) {
        px += a.S[(size_t)kx * G + g] + a.blen[g + 1];
        if (px >= m) return 0;  // box g+1 does not fit in x
        jx[g] = px;
    }
    // odometer over the y-side gaps 0 .. G-2; the last gap is a window popcount
    uint32_t s[NB - 1], pt[NB];
    uint32_t total = 0;
    int g = 0;
    pt[0] = pt0;
    s[0] = a.gmin[0];
    if (G == 1) {
        uint32_t lo = pt0 + a.gmin[0] + a.blen[1], hi = pt0 + a.gmax[0] + a.blen[1];
        if (lo >= n) return 0;
        if (hi >= n) hi = n - 1;
        return bm_window(bm[1], m, jx[0], lo, hi);
    }
    for (;;) {
        if (s[g] > a.gmax[g]) {
            if (g == 0) break;
            --g;
            ++s[g];
            continue;
        }
        uint32_t p = pt[g] + s[g] + a.blen[g + 1];
        if (p >= n) {  // this and every longer gap run past the text
            if (g == 0) break;
            --g;
            ++s[g];
            continue;
        }
        if (!bm_bit(bm[g + 1], m, jx[g], p)) { ++s[g]; continue; }
        if (g + 2 == (int)G) {
            uint32_t lo = p + a.gmin[g + 1] + a.blen[g + 2], hi = p + a.gmax[g + 1] + a.blen[g + 2];
            if (lo < n) total += bm_window(bm[g + 2], m, jx[g + 1], lo, hi < n ? hi : n - 1);
            ++s[g];
            continue;
        }
        pt[g + 1] = p;
        ++g;
        s[g] = a.gmin[g];
    }
    return total;
}

// Hamming hits of one box shape for the pair, by the diagonal sweep, into its occurrence
// list and (WRITE_BM) into bm and bmT.  The sweep runs over blocks of 2048 diagonals and the
// rows of a block in order, so the list is a sequence of runs sorted by pos_x, one per
// block; when runs is not null, runs[0] receives their number and runs[1 + r] .. runs[2 + r]
// bound run r.
template <int P, bool WRITE_BM>
__device__ void struct_hd_box(const StructArgs &a, const StructItem &I, uint32_t len, uint32_t err,
                              uint32_t *bm, uint32_t *bmT, uint32_t *ne_l, const uint16_t *xoff,
                              uint32_t *list, uint32_t *lcount, uint32_t *runs)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t m = I.m, n = I.n, wl = a.wl;
    const uint32_t zero_row = a.sigma * wl;
    uint32_t hm[P];  // window counts are biased (hd_device.h): hits have planes ph.. clear
    planes_high_masks<P>(hm, hd_low_planes(err));
    int64_t q0 = (int64_t)I.y_start - (int64_t)(m - 1);
    q0 = q0 >= 0 ? (q0 & ~31LL) : -(((-q0) + 31) & ~31LL);
    const int64_t qend = (int64_t)I.y_start + n;  // diagonals below this can hold hits
    uint32_t nruns = 0;
    if (runs && lane == 0) runs[1] = 0;
    // list entries so far: in list-only mode a wave-uniform register (slots from ballots, no
    // LDS atomic per occurrence); with bitmaps, the LDS counter
    uint32_t nl = __builtin_amdgcn_readfirstlane(*lcount);
    for (int64_t qb = q0; qb < qend; qb += kDiagPerBlock) {
        const uint32_t wbase = (uint32_t)((qb + (int64_t)a.pad_bits) / 32);
        // text position of lane 0's flag 0 at row 0: within (-m - 32, n + 2048), so 32 bits
        const int32_t rb = (int32_t)(qb - (int64_t)I.y_start);
        __syncthreads();
        for (uint32_t c = 0; c < a.sigma; ++c) {
            const uint32_t *src = a.ne + (size_t)c * a.W + wbase;
            for (uint32_t w = lane; w < wl; w += 64) ne_l[c * wl + w] = src[w];
        }
        __syncthreads();
        uint32_t cnt[P];
        planes_set<P>(cnt, hd_bias(err));
        // xoff[px] is the byte offset of column px's first window word (its symbol's row, word
        // px / 32); lane L reads word L after it.
        const char *ne_b = reinterpret_cast<const char *>(ne_l) + 4 * lane;
        // Row px: the window gains column px and, from px = len on, loses column px - len (ob:
        // its byte offset as above; before len, the all-zero row: nothing leaves).  WARM: rows
        // before len, of which only the last can hold hits.
        auto row = [&](uint32_t px, uint32_t ob, uint32_t sho, auto warm) {
            const uint32_t *wn = reinterpret_cast<const uint32_t *>(ne_b + xoff[px]);
            const uint32_t *wo = reinterpret_cast<const uint32_t *>(ne_b + ob);
            uint32_t an = __builtin_amdgcn_alignbit(wn[1], wn[0], px & 31);
            uint32_t bo = __builtin_amdgcn_alignbit(wo[1], wo[0], sho);
            planes_slide<P>(cnt, an, bo);
            const uint32_t res = ~planes_high_masked<P>(cnt, hm);
            if ((!decltype(warm)::value || px + 1 >= len) && __ballot(res != 0)) {
                // flag t of this lane is text position pt = base + t; only [len - 1, n) counts
                // valid flags t in [tlo, thi), both clamped to [0, 32]; the 64-bit shifts give
                // the empty and full masks at the ends without selects
                const int32_t base = rb + (int32_t)(32 * lane + px);
                const uint32_t tlo = (uint32_t)min(max((int32_t)len - 1 - base, 0), 32);
                const uint32_t thi = (uint32_t)min(max((int32_t)n - base, 0), 32);
                uint32_t hv = res & (uint32_t)(~0ull << tlo) & ~(uint32_t)(~0ull << thi);
                if constexpr (WRITE_BM) {
                    while (hv) {
                        const uint32_t t = __builtin_ctz(hv);
                        hv &= hv - 1;
                        const uint32_t pt = (uint32_t)(base + t);
                        atomicOr(bm + (size_t)(pt >> 5) * m + px, 1u << (pt & 31));
                        if (bmT) atomicOr(bmT + (size_t)(px >> 5) * n + pt, 1u << (px & 31));
                        const uint32_t slot = atomicAdd(lcount, 1u);
                        if (slot < a.lcap) list[slot] = (px << 16) | pt;
                    }
                } else {
                    // one occurrence per lane per round; slots from the round's ballot
                    for (;;) {
                        const uint64_t mb = __ballot(hv != 0);
                        if (!mb) break;
                        if (hv) {
                            const uint32_t t = __builtin_ctz(hv);
                            hv &= hv - 1;
                            const uint32_t slot =
                                nl + __builtin_amdgcn_mbcnt_hi((uint32_t)(mb >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)mb, 0u));
                            if (slot < a.lcap) list[slot] = (px << 16) | (uint32_t)(base + t);
                        }
                        nl += (uint32_t)__builtin_popcountll(mb);
                    }
                }
            }
        };
        const uint32_t warm = len < m ? len : m;
        uint32_t px = 0;
        for (; px < warm; ++px) row(px, 4 * zero_row, 0u, std::true_type{});
        for (; px < m; ++px) row(px, xoff[px - len], (px - len) & 31, std::false_type{});
        ++nruns;
        if constexpr (WRITE_BM) nl = __builtin_amdgcn_readfirstlane(*lcount);
        if (runs && lane == 0) runs[1 + nruns] = min(nl, a.lcap);
    }
    if (runs && lane == 0) runs[0] = nruns;
    if (!WRITE_BM && lane == 0) *lcount = nl;
}

// DENSE: the register allocation is held to eight resident waves per SIMD (64 registers,
// about 180 B of spills per lane), for launches whose LDS footprint lets that many waves in
// and whose pairs take the list merge (one gap, Hamming boxes): there the pair sweeps are
// short and their set-up latency-bound, so more waves beat the spills (1000 x 600 nt, boxes
// 8+[10,30]+8: 18.2 ms at five waves, 16.6 at eight with the grid at 32 waves per CU).  Two
// gaps (the bitmap merge) ran 1.6% slower and a protein alphabet (a larger LDS window, so
// fewer waves fit anyway) 13% slower that way, so they keep the free allocation.
template <bool DENSE>
constexpr int struct_waves_per_eu() { return DENSE ? 8 : 1; }

// NB: boxes the per-pair arrays hold (5 up to four gaps, kMaxBoxes beyond): larger arrays
// of bitmap pointers and walk state live in scratch and slow the common shapes down.
template <int P, bool DENSE, int NB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(struct_waves_per_eu<DENSE>())))
void struct_kernel(StructArgs a)
{
    extern __shared__ uint32_t smem[];
    const uint32_t lane = threadIdx.x;
    const uint32_t wl = a.wl;
    // LDS: NE window (sigma+1 rows) | x window offsets (m_max u16, bytes; during the list
    // merge, box 1's pos_x keys) | hit list | count
    uint32_t *ne_l = smem;
    uint16_t *xoff = reinterpret_cast<uint16_t *>(smem + (a.sigma + 1) * wl);
    uint32_t *hits = smem + (a.sigma + 1) * wl + (a.m_max + 1) / 2;
    uint32_t *hcount = hits + a.hcap;
    uint32_t *lcount = hcount + 1;  // per box shape: occurrences listed this pair
    uint32_t *runs = lcount + kMaxBoxes;  // per box shape: sorted runs of its list (run_words)
    for (uint32_t w = lane; w < wl; w += 64) ne_l[a.sigma * wl + w] = 0;
    // scratch: box bitmaps | (Hamming) their transposes | two edit boundary streams |
    //          counted (kx, pos_x) bits | (Hamming) occurrence lists
    uint32_t *scr = a.scratch + (size_t)blockIdx.x * a.scratch_words;
    uint32_t *scrT = scr + (size_t)a.ndist * a.bm_words;
    uint32_t *buf0 = scrT + (a.distance == 0 ? (size_t)a.ndist * a.bm_words : 0);
    uint32_t *buf1 = buf0 + a.stream_words;
    uint32_t *counted = buf1 + a.stream_words;
    uint32_t *lists = counted + ((uint64_t)a.nb_structs * a.m_max + 31) / 32 + 1;
    // gathered box-0 hits of the bitmap-scan path: in LDS, or after the lists in scratch
    uint32_t *gather = a.gather_scratch ? lists + (size_t)a.ndist * a.lcap : hits;
    // Hamming boxes: the bitmaps are all zero between pairs (each pair clears the words it
    // set, through the occurrence lists, or all of them when a list overflowed); the first
    // pair of a launch clears them in full.
    bool clean = false;
    // the pairs a range-batched launch handed over: their count is known on the device only
    const uint32_t nitems = a.nitems_dev ? min(a.nitems, *a.nitems_dev) : a.nitems;

    for (uint32_t round = 0;; ++round) {
        uint32_t it;
        if (a.persistent) {
            it = 0;
            if (lane == 0) it = atomicAdd(a.queue, 1u);
            it = __builtin_amdgcn_readfirstlane(it);
        } else {
            it = round == 0 ? blockIdx.x : nitems;
        }
        if (it >= nitems) break;
        // The merge of functions.cc:647-690 for two boxes from their occurrence lists.  With
        // box 0 ending at (px0, pt0) and box 1 at (px1, pt1), the structured motif with x-side
        // gap gx and y-side gap gy occurs when px1 = px0 + gx + len1 and pt1 = pt0 + gy + len1:
        // every pair of occurrences with both differences in [gmin + len1, gmax + len1] is one
        // (kx, ky) combination, counted for x at (kx, px0) and, for a pair-symmetric item,
        // for y at (ky, pt0) (the transposed pair has the same occurrences, coordinates
        // swapped).  Box-1 occurrences are found by binary search in the sorted runs of its
        // list; u is deduplicated per (combination, position) with the counted bits.
        auto pair_lists = [&](const StructItem &I, bool sym) {
            const uint32_t m = I.m, n = I.n;
            const uint32_t d0 = a.dist_of[0], d1 = a.dist_of[1];
            const uint32_t *l0 = lists + (size_t)d0 * a.lcap, *l1 = lists + (size_t)d1 * a.lcap;
            const uint32_t c0 = __builtin_amdgcn_readfirstlane(lcount[d0]);
            const uint32_t *r1 = runs + a.run_words * d1;
            const uint32_t nr1 = __builtin_amdgcn_readfirstlane(r1[0]);
            const uint32_t lo = a.gmin[0] + a.blen[1], hi = a.gmax[0] + a.blen[1];
            // The binary searches probe box 1's pos_x keys from LDS when they fit where the x
            // table was (it is rebuilt for the next pair); the barrier below orders the copy.
            const uint32_t c1 = __builtin_amdgcn_readfirstlane(lcount[d1]);
            const bool lds_keys = c1 <= a.m_max;
            if (lds_keys)
                for (uint32_t i = lane; i < c1; i += 64) xoff[i] = (uint16_t)(l1[i] >> 16);
            for (int orient = 0; orient < (sym ? 2 : 1); ++orient) {
                const uint32_t mo = orient ? n : m;
                const uint32_t uvb = orient ? I.uv_base_t : I.uv_base;
                const uint32_t cw = (a.nb_structs * mo + 31) / 32;
                uint32_t *cnt_bits = (uint64_t)a.nb_structs * mo <= (uint64_t)a.hcap * 32 ? hits : counted;
                __syncthreads();
                for (uint32_t w = lane; w < cw; w += 64) cnt_bits[w] = 0;
                __threadfence_block();
                __syncthreads();
                for (uint32_t h = lane; h < c0; h += 64) {
                    const uint32_t e0 = l0[h];
                    const uint32_t px0 = e0 >> 16, pt0 = e0 & 0xFFFFu;
                    const uint32_t xlo = px0 + lo, xhi = px0 + hi;
                    for (uint32_t r = 0; r < nr1; ++r) {
                        uint32_t b = r1[1 + r], en = r1[2 + r];
                        while (b < en) {  // first entry of the run with pos_x >= xlo
                            const uint32_t mid = (b + en) >> 1;
                            const uint32_t key = lds_keys ? xoff[mid] : l1[mid] >> 16;
                            if (key < xlo) b = mid + 1; else en = mid;
                        }
                        for (const uint32_t end = r1[2 + r]; b < end; ++b) {
                            const uint32_t e1 = l1[b];
                            const uint32_t px1 = e1 >> 16, pt1 = e1 & 0xFFFFu;
                            if (px1 > xhi) break;
                            if (pt1 < pt0 + lo || pt1 > pt0 + hi) continue;
                            const uint32_t kx = orient ? pt1 - pt0 - lo : px1 - px0 - lo;
                            const uint32_t idx = kx * mo + (orient ? pt0 : px0), bit = 1u << (idx & 31);
                            const uint32_t old = atomicOr(cnt_bits + (idx >> 5), bit);
                            const size_t o = (size_t)uvb + idx;
                            if (!(old & bit)) atomicAdd(a.u + o, 1u);
                            atomicAdd(a.v + o, 1u);
                        }
                    }
                }
            }
            __threadfence_block();
        };
        // One ordered pair I (and, when sym, the transposed pair from the same box DPs).
        // Returns true when the transposed pair still has to be run on its own (its
        // occurrences did not fit the lists, or a box passes everywhere).
        auto pair = [&](const StructItem &I, bool sym) -> bool {
            const uint32_t m = I.m, n = I.n;
            const uint32_t wn = (n + 31) / 32;
            uint32_t *bmd[NB];
            for (uint32_t d = 0; d < a.ndist; ++d) bmd[d] = scr + (size_t)d * wn * m;
            // transposed bitmaps of a pair-symmetric item: bmT[(pos_x >> 5) * n + pos_t]
            const uint32_t wm = (m + 31) / 32;
            uint32_t *bmTd[NB], *bmT[NB];
            for (uint32_t d = 0; d < a.ndist; ++d) bmTd[d] = scrT + (size_t)d * wm * n;
            for (uint32_t b = 0; b <= a.nb_gaps; ++b) bmT[b] = bmTd[a.dist_of[b]];
            uint32_t *bm[NB];
            for (uint32_t b = 0; b <= a.nb_gaps; ++b) bm[b] = bmd[a.dist_of[b]];
            __syncthreads();
            if (a.distance == 0) {
                // below (sigma + 1) * wl * 4 <= the LDS size, checked on the host: fits 16 bits
                for (uint32_t r = lane; r < m; r += 64)
                    xoff[r] = (uint16_t)(4 * (a.codes[I.x_start + r] * wl + (r >> 5)));
            }
            if (a.distance == 0 && !clean) {
                for (size_t w = lane; w < (size_t)2 * a.ndist * a.bm_words; w += 64) scr[w] = 0;
                clean = true;
            }
            if (lane < a.ndist) lcount[lane] = 0;
            bool listed = a.distance == 0;  // every shape's occurrences are in its list
            // Two Hamming boxes, neither passing everywhere: the merge pairs the two boxes'
            // occurrence lists directly and needs no bitmaps (pair_lists below).
            bool pairlist = a.distance == 0 && a.nb_gaps == 1;
            for (uint32_t d = 0; d < a.ndist; ++d) pairlist = pairlist && a.derr[d] < m && a.derr[d] < n;
            // 1. box bitmaps
            for (uint32_t d = 0; d < a.ndist; ++d) {
                const uint32_t len = a.dlen[d], err = a.derr[d];
                uint32_t *out = bmd[d];
                if (err >= m) {
                    // passes everywhere inside the pair (functions.cc:597-598, 651, 669)
                    for (size_t w = lane; w < (size_t)wn * m; w += 64) {
                        uint32_t word = (uint32_t)(w / m);
                        uint32_t rem = n - 32 * word;
                        out[w] = rem >= 32 ? ~0u : ((1u << rem) - 1);
                    }
                    listed = false;
                    continue;
                }
                if (a.distance == 0) {
                    __syncthreads();
                    if (pairlist)
                        struct_hd_box<P, false>(a, I, len, err, out, sym ? bmTd[d] : nullptr, ne_l, xoff,
                                                lists + (size_t)d * a.lcap, lcount + d, runs + a.run_words * d);
                    else
                        struct_hd_box<P, true>(a, I, len, err, out, sym ? bmTd[d] : nullptr, ne_l, xoff,
                                               lists + (size_t)d * a.lcap, lcount + d, nullptr);
                } else {
                    EdOut eo{out, m};
                    if (len <= 14)
                        ed_pair<ed_cols(0), 0, true>(a.codes + I.x_start, m, a.codes + I.y_start, n, len, err,
                                                     buf0, buf1, nullptr, nullptr, eo);
                    else if (len <= 32)
                        ed_pair<ed_cols(1), 1, true>(a.codes + I.x_start, m, a.codes + I.y_start, n, len, err,
                                                     buf0, buf1, nullptr, nullptr, eo);
                    else
                        ed_pair<ed_cols(2), 2, true>(a.codes + I.x_start, m, a.codes + I.y_start, n, len, err,
                                                     buf0, buf1, nullptr, nullptr, eo);
                }
            }
            __threadfence_block();
            __syncthreads();
            if (pairlist) {
                bool fits = true;
                for (uint32_t d = 0; d < a.ndist; ++d)
                    if (__builtin_amdgcn_readfirstlane(lcount[d]) > a.lcap) fits = false;
                if (fits) {
                    pair_lists(I, sym);
                    return false;
                }
                // a list overflowed: redo the box DPs with bitmaps for the scan below
                __syncthreads();
                if (lane < a.ndist) lcount[lane] = 0;
                for (uint32_t d = 0; d < a.ndist; ++d) {
                    __syncthreads();
                    struct_hd_box<P, true>(a, I, a.dlen[d], a.derr[d], bmd[d], sym ? bmTd[d] : nullptr, ne_l,
                                           xoff, lists + (size_t)d * a.lcap, lcount + d, nullptr);
                }
                __threadfence_block();
                __syncthreads();
            }
            // 2. merge: gather box-0 hits word row by word row, then count per x combination
            const uint32_t cwords = (a.nb_structs * m + 31) / 32;
            for (uint32_t w = lane; w < cwords; w += 64) counted[w] = 0;
            __threadfence_block();
            if (lane == 0) *hcount = 0;
            __syncthreads();
            const uint32_t *b0 = bm[0];
            for (uint32_t d = 0; d < a.ndist && listed; ++d)
                if (__builtin_amdgcn_readfirstlane(lcount[d]) > a.lcap) listed = false;
            if (listed) {
                // walk box 0's listed occurrences (one per lane, all its x-side combinations in
                // turn, so a lane's bitmap reads for different combinations overlap); for a
                // pair-symmetric item walk them again with the roles of x and y swapped, over the
                // transposed bitmaps; then clear the bitmap words of every list.  The
                // (combination, pos_x) counted bits live in LDS (in the hit-list area, unused
                // here) when they fit.
                const uint32_t d0 = a.dist_of[0];
                const uint32_t *l0 = lists + (size_t)d0 * a.lcap;
                const uint32_t c0 = __builtin_amdgcn_readfirstlane(lcount[d0]);
                // a box that passes everywhere in (y, x) (errors >= n) has no bitmap here
                bool sym2 = sym;
                for (uint32_t d = 0; d < a.ndist; ++d) sym2 = sym2 && a.derr[d] < n;
                for (int orient = 0; orient < (sym2 ? 2 : 1); ++orient) {
                    const uint32_t mo = orient ? n : m, no = orient ? m : n;
                    uint32_t *const *bmo = orient ? bmT : bm;
                    const uint32_t uvb = orient ? I.uv_base_t : I.uv_base;
                    const uint32_t cw = (a.nb_structs * mo + 31) / 32;
                    const bool lds_counted = (uint64_t)a.nb_structs * mo <= (uint64_t)a.hcap * 32;
                    uint32_t *cnt_bits = lds_counted ? hits : counted;
                    __syncthreads();
                    for (uint32_t w = lane; w < cw; w += 64) cnt_bits[w] = 0;
                    __threadfence_block();
                    __syncthreads();
                    for (uint32_t h = lane; h < c0; h += 64) {
                        const uint32_t e = l0[h];
                        const uint32_t px0 = orient ? (e & 0xFFFFu) : (e >> 16);
                        const uint32_t pt0 = orient ? (e >> 16) : (e & 0xFFFFu);
    #pragma unroll 4
                        for (uint32_t kx = 0; kx < a.nb_structs; ++kx) {
                            const uint32_t c = struct_walk<NB>(a, bmo, mo, no, kx, px0, pt0);
                            if (c) {
                                const uint32_t idx = kx * mo + px0, bit = 1u << (idx & 31);
                                const uint32_t old = atomicOr(cnt_bits + (idx >> 5), bit);
                                const size_t o = (size_t)uvb + idx;
                                if (!(old & bit)) atomicAdd(a.u + o, 1u);
                                atomicAdd(a.v + o, c);
                            }
                        }
                    }
                }
                __threadfence_block();
                __syncthreads();
                for (uint32_t d = 0; d < a.ndist; ++d) {
                    const uint32_t *ld = lists + (size_t)d * a.lcap;
                    const uint32_t cd = __builtin_amdgcn_readfirstlane(lcount[d]);
                    for (uint32_t h = lane; h < cd; h += 64) {
                        const uint32_t e = ld[h];
                        const uint32_t px = e >> 16, pt = e & 0xFFFFu;
                        bmd[d][(size_t)(pt >> 5) * m + px] = 0;
                        if (sym) bmTd[d][(size_t)(px >> 5) * n + pt] = 0;
                    }
                }
                __threadfence_block();
                return sym && !sym2;
            }
            // the bitmaps are scanned; Hamming ones are cleared in full after the merge
            clean = a.distance != 0;
            // box-0 occurrences of the pair: count them first (every lane streams its share of
            // the bitmap with no barrier in between); when they fit the LDS list, gather them in
            // one more pass and walk them once.  Otherwise fall back to gathering chunk by
            // chunk, walking whenever the next chunk could overflow the list.
            uint32_t mine = 0;
            for (uint32_t w = 0; w < wn; ++w)
                for (uint32_t px = lane; px < m; px += 64) mine += __builtin_popcount(b0[(size_t)w * m + px]);
            if (lane == 0) *hcount = 0;
            __syncthreads();
            if (mine) atomicAdd(hcount, mine);
            __syncthreads();
            const uint32_t total = __builtin_amdgcn_readfirstlane(*hcount);
            __syncthreads();
            if (lane == 0) *hcount = 0;
            __syncthreads();
            auto walk_list = [&](uint32_t have) {
                for (uint32_t kx = 0; kx < a.nb_structs; ++kx) {
                    for (uint32_t h = lane; h < have; h += 64) {
                        uint32_t e = gather[h];
                        uint32_t px0 = e >> 16, pt0 = e & 0xFFFFu;
                        uint32_t c = struct_walk<NB>(a, bm, m, n, kx, px0, pt0);
                        if (c) {
                            uint32_t idx = kx * m + px0, bit = 1u << (idx & 31);
                            uint32_t old = atomicOr(counted + (idx >> 5), bit);
                            size_t o = (size_t)I.uv_base + idx;
                            if (!(old & bit)) atomicAdd(a.u + o, 1u);
                            atomicAdd(a.v + o, c);
                        }
                    }
                }
            };
            auto append = [&](uint32_t w, uint32_t px) {
                uint32_t word = px < m ? b0[(size_t)w * m + px] : 0u;
                if (word) {
                    uint32_t off = atomicAdd(hcount, (uint32_t)__builtin_popcount(word));
                    while (word) {
                        uint32_t t = __builtin_ctz(word);
                        word &= word - 1;
                        gather[off++] = (px << 16) | (32 * w + t);
                    }
                }
            };
            if (total <= a.gcap) {
                if (total) {
                    for (uint32_t w = 0; w < wn; ++w)
                        for (uint32_t px0 = 0; px0 < m; px0 += 64) append(w, px0 + lane);
                    __syncthreads();
                    walk_list(total);
                }
                return sym;
            }
            const uint32_t chunks = (m + 63) / 64;
            for (uint32_t wc = 0; wc <= wn * chunks; ++wc) {
                // walk before a chunk that could overflow the list, and at the end
                uint32_t have = __builtin_amdgcn_readfirstlane(*hcount);
                bool flush = (wc == wn * chunks) ? have > 0 : have + 64 * 32 > a.gcap;
                if (flush) {
                    walk_list(have);
                    __syncthreads();
                    if (lane == 0) *hcount = 0;
                    __syncthreads();
                }
                if (wc == wn * chunks) break;
                append(wc / chunks, (wc % chunks) * 64 + lane);
                __syncthreads();
            }
            return sym;
        };
        const StructItem I = a.items[it];
        const bool sym = I.uv_base_t != kNoSplit;
        if (pair(I, sym)) {
            clean = false;  // the transposed bitmaps were set: start the swapped run clean
            const StructItem I2{I.y_start, I.n, I.x_start, I.m, I.uv_base_t, kNoSplit};
            pair(I2, false);
        }
    }
}

template <bool DENSE>
static void launch_struct_p(const StructArgs &a, uint32_t planes, uint32_t grid, size_t lds, hipStream_t s)
{
    if (planes <= 3) hipLaunchKernelGGL((struct_kernel<3, DENSE, 5>), dim3(grid), dim3(64), lds, s, a);
    else if (planes == 4) hipLaunchKernelGGL((struct_kernel<4, DENSE, 5>), dim3(grid), dim3(64), lds, s, a);
    else if (planes == 5) hipLaunchKernelGGL((struct_kernel<5, DENSE, 5>), dim3(grid), dim3(64), lds, s, a);
    else if (planes == 6) hipLaunchKernelGGL((struct_kernel<6, DENSE, 5>), dim3(grid), dim3(64), lds, s, a);
    else hipLaunchKernelGGL((struct_kernel<7, DENSE, 5>), dim3(grid), dim3(64), lds, s, a);
}

hipError_t launch_struct(const StructArgs &a, uint32_t grid, size_t lds, bool dense, hipStream_t s)
{
    uint32_t planes = 1;  // biased Hamming window counters of every shape (hd_device.h)
    for (uint32_t d = 0; d < a.ndist; ++d) {
        uint32_t p = hd_counter_planes(a.dlen[d], a.derr[d]);
        planes = p > planes ? p : planes;
    }
    if (planes > 7 || a.nb_gaps > 4) {
        // boxes longer than 64 (up to 1024) or more than four gaps: one general form with
        // eleven planes, the threshold set at run time
        hipLaunchKernelGGL((struct_kernel<11, false, kMaxBoxes>), dim3(grid), dim3(64), lds, s, a);
    } else if (dense) {
        launch_struct_p<true>(a, planes, grid, lds, s);
    } else {
        launch_struct_p<false>(a, planes, grid, lds, s);
    }
    return hipGetLastError();
}


}  // namespace motex
