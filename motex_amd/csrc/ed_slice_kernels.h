// Text-parallel edit-distance kernel for gfx950: x against 32 texts at once.
//
// What is computed is the edit recurrence of ed_kernels.hip (motifs_extraction_ed,
// functions.cc:734-768, MaxShift functions.cc:437-447), for one x and a group of up to 32
// texts: bit t of every word belongs to text t of the group, and ed_slice.h advances one
// cell of all 32 texts with bitwise operations (a word is a stack of bit planes).
//
// Layout.  A wave owns a strip of 64 columns of x, one per lane, and streams the group's
// rows through it as a skewed wavefront: at step s lane L works on row s-L.  The left
// neighbour's cell arrives from lane L-1 by one DPP wave shift per word; lane 0 takes it from
// an LDS chunk of 64 rows holding the previous strip's last column (the boundary stream, in
// the wave's scratch).  Every lane reads its row's text planes from a ring of the group's
// last 128 rows in LDS, filled with the chunk (a read, not six DPP shifts, per step).  Lane 63
// writes its cells to the next strip's boundary stream.  Past the first strip every column
// has j > K, so only the low K-1 planes of a word travel (all a truncated left or diagonal
// candidate keeps) and the receiver recounts their popcount (slice_cell_masked).
// Per column, the texts with an occurrence (OR of the hit masks over rows) and the number
// of occurrences (popcount of the hit masks) go to u and v with one atomic each per strip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "ed_slice.h"
#include "motex_internal.h"

namespace motex {

// wave_shr:1 -- lane L receives lane L-1's x; lane 0 receives in0
__device__ __forceinline__ uint32_t shr1_from(uint32_t in0, uint32_t x)
{
    return __builtin_amdgcn_update_dpp(in0, x, 0x138, 0xF, 0xF, false);
}

// wave_shr:1 with lane 0 receiving zero (bound_ctrl: no old value to set up first)
__device__ __forceinline__ uint32_t shr1_zero(uint32_t x)
{
    return __builtin_amdgcn_update_dpp(0u, x, 0x138, 0xF, 0xF, true);
}

// Past the first strip every column has j > K, where the left and diagonal candidates keep
// only bits 0 .. K-2 of their source: those SP = K-1 planes are all a cell's right neighbour
// reads from it, so only they cross lanes and strips.
template <int K>
struct SliceLayout {
    static constexpr int W = SliceW<K>::W;
    static constexpr int SP = K - 1;               // planes a later strip's left neighbour sends
    static constexpr int SWS = (SP + 3) & ~3;      // boundary-stream row
    static constexpr int LW = SWS;                 // LDS chunk row: the stream row
};

// The group's text planes in LDS: rows r of the last 128 at (r mod 128) x 6 words.  A step's
// lanes read rows s-64 .. s-1 and a chunk fill writes rows s-1 .. s+62: 127 distinct slots.
constexpr uint32_t kTextRing = 128;
constexpr uint32_t kTextWords = 6;

// Text planes of row r (5 symbol-code planes and the valid-text mask), 8-byte reads.
__device__ __forceinline__ void slice_text(const uint32_t *tring, uint32_t r, uint32_t *t)
{
    const uint2 *src = reinterpret_cast<const uint2 *>(tring + (r & (kTextRing - 1)) * kTextWords);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const uint2 v = src[q];
        t[2 * q] = v.x;
        t[2 * q + 1] = v.y;
    }
}

// Lane rr fills row rr of the LDS chunk of 64 rows from r0: the previous strip's boundary row
// (when from_stream; zeros for the first strip, whose left edge is column 0) with 16-byte
// loads and stores, and the row's slot of the text ring with the group's 6 text words in
// 8-byte ones.  The stream rows (SWS words, a multiple of 4, from a 16-byte aligned base), the
// text rows (6 words) and the chunk rows (LW words, a multiple of 4) keep those accesses
// aligned.
template <int K>
__device__ __forceinline__ void slice_fill_row(uint32_t *chunk, uint32_t *tring, uint32_t r0, uint32_t nrows,
                                               bool from_stream, const uint32_t *sin, const uint32_t *gp,
                                               uint32_t rr)
{
    typedef SliceLayout<K> Lay;
    constexpr int SWS = Lay::SWS, LW = Lay::LW;
    const uint32_t row = r0 + rr;
    uint4 *dst = reinterpret_cast<uint4 *>(chunk + rr * LW);
    uint2 *dtext = reinterpret_cast<uint2 *>(tring + (row & (kTextRing - 1)) * kTextWords);
    if (row < nrows) {
        if (from_stream) {
            const uint4 *src = reinterpret_cast<const uint4 *>(sin + (size_t)row * SWS);
#pragma unroll
            for (int q = 0; q < SWS / 4; ++q) dst[q] = src[q];
        } else {
#pragma unroll
            for (int q = 0; q < SWS / 4; ++q) dst[q] = make_uint4(0, 0, 0, 0);
        }
        const uint2 *tsrc = reinterpret_cast<const uint2 *>(gp + (size_t)row * 6);
#pragma unroll
        for (int q = 0; q < 3; ++q) dtext[q] = tsrc[q];
    } else {
#pragma unroll
        for (int q = 0; q < SWS / 4; ++q) dst[q] = make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 3; ++q) dtext[q] = make_uint2(0, 0);
    }
}

// Loads the LDS chunk of 64 rows from r0: the previous strip's boundary cells (strip > 0)
// and the group's text planes.
template <int K>
__device__ __forceinline__ void slice_chunk(uint32_t *chunk, uint32_t *tring, uint32_t r0, uint32_t nrows,
                                            bool from_stream, const uint32_t *sin, const uint32_t *gp)
{
    __syncthreads();  // the previous chunk is no longer read
    slice_fill_row<K>(chunk, tring, r0, nrows, from_stream, sin, gp, threadIdx.x);
    __syncthreads();
}

// Lane 63 writes its cell of row s-63 (0-based s-64) to the next strip's boundary stream.
template <int K>
__device__ __forceinline__ void slice_emit(uint32_t *sout, uint32_t s, const uint32_t *d)
{
    typedef SliceLayout<K> Lay;
    constexpr int SP = Lay::SP, SWS = Lay::SWS;
    uint32_t w[SWS];
#pragma unroll
    for (int p = 0; p < SWS; ++p) w[p] = p < SP ? d[p] : 0u;
    uint4 *dst = reinterpret_cast<uint4 *>(sout + (size_t)(s - 64) * SWS);
#pragma unroll
    for (int p = 0; p < SWS / 4; ++p) dst[p] = make_uint4(w[4 * p], w[4 * p + 1], w[4 * p + 2], w[4 * p + 3]);
}

// The left neighbour's cell D[i][j-1]: lane L-1's, by one DPP wave shift per plane; lane 0
// takes column 0 (zero) in the first strip and the previous strip's boundary row cr later.
// The first strip moves all W planes and recounts pc and pm; later strips move the SP
// planes their cells read (the rest stay zero) and recount pm.
template <int K, bool FIRST>
__device__ __forceinline__ void slice_left(Slice<K> &L, const uint32_t *ud, const uint32_t *cr)
{
    typedef SliceLayout<K> Lay;
    constexpr int W = Lay::W, SP = Lay::SP;
    if constexpr (FIRST) {
#pragma unroll
        for (int p = 0; p < W; ++p) L.d[p] = shr1_zero(ud[p]);
        slice_popcount<W, slice_pcq<K>()>(L.d, L.pc);
    } else {
#pragma unroll
        for (int p = 0; p < SP; ++p) L.d[p] = shr1_from(cr[p], ud[p]);
#pragma unroll
        for (int p = SP; p < W; ++p) L.d[p] = 0u;
    }
    slice_pm<K>(L.d, L.pm);
}

// One strip of 64 columns over all rows of the group.  Only a cell's planes travel between
// lanes; the receiver recounts its popcounts.  FIRST: the first strip, whose left edge is
// column 0 and whose first K columns enter left and diag untruncated (slice_cell, needing
// pc and pm of left).  Later strips have only columns j > K (slice_cell_masked, needing pm).
template <int K, bool FIRST>
__device__ __forceinline__ void slice_strip(const EdSliceArgs &a, const EdSliceItem &I, uint32_t strip,
                                            uint32_t nrows, const uint32_t *gp, const uint32_t *sin,
                                            uint32_t *sout, bool last, uint32_t *chunk, uint32_t *tring)
{
    typedef SliceLayout<K> Lay;
    constexpr int LW = Lay::LW;
    const uint32_t lane = threadIdx.x;
    const uint32_t m = I.m;
    const uint32_t j = strip * 64 + lane + 1;  // 1-based column
    const uint32_t code = j <= m ? a.codes[I.x_start + j - 1] : 0u;
    uint32_t xb[5];
#pragma unroll
    for (int b = 0; b < 5; ++b) xb[b] = ((code >> b) & 1) ? ~0u : 0u;
    const bool report = j <= m && j >= (uint32_t)K;
    uint32_t eb[kSlicePc];  // the planes of e, wave-uniform
#pragma unroll
    for (int q = 0; q < kSlicePc; ++q) eb[q] = __builtin_amdgcn_readfirstlane(((a.e >> q) & 1) ? ~0u : 0u);
    const uint32_t full = j <= (uint32_t)K ? ~0u : 0u;
    // state: D[i-1][j] (what lane L+1 receives next) and D[i-1][j-1]
    Slice<K> U, G;
    slice_row0<K>(j, U);
    slice_row0<K>(j - 1, G);
    uint32_t vcnt = 0, anyh = 0;
    const uint32_t steps = nrows + 63;
    auto step = [&](uint32_t s, auto check) {
        const uint32_t r0 = s - 1;  // 0-based row lane 0 starts this step
        if ((r0 & 63) == 0) slice_chunk<K>(chunk, tring, r0, nrows, !FIRST, sin, gp);
        const uint32_t *cr = chunk + (r0 & 63) * LW;
        Slice<K> L;  // D[i][j-1]; only its planes travel, the receiver recounts pc and pm
        slice_left<K, FIRST>(L, U.d, cr);
        uint32_t t[kTextWords];  // row r0 - lane's text planes (garbage on lanes without a row)
        slice_text(tring, r0 - lane, t);
        bool active = true;
        if constexpr (decltype(check)::value) active = s > lane && s - lane <= nrows;
        if (active) {
            uint32_t ne = 0;
#pragma unroll
            for (int b = 0; b < 5; ++b) ne |= t[b] ^ xb[b];
            Slice<K> D;
            if constexpr (FIRST) {
                slice_cell<K>(U, L, G, ne, full, D);
            } else {
                slice_cell_masked<K>(U.d, U.hi, U.pc, L.d, L.pm, G.d, G.pm, ne, D.d, D.hi, D.pc);
            }
            if (report) {
                const uint32_t hit = slice_le_planes<slice_pcq<K>()>(D.pc, eb) & t[5];
                vcnt += __builtin_popcount(hit);
                anyh |= hit;
            }
            G = L;
            U = D;
            if (!last && lane == 63) slice_emit<K>(sout, s, D.d);
        }
    };
    // fill (s < 64), steady state (64 <= s <= nrows: every lane has a row; two steps per
    // trip, so the cell planes are not copied between registers on each back edge), drain
    uint32_t s = 1;
    const uint32_t fill_end = nrows < 63 ? nrows : 63;
    for (; s <= fill_end; ++s) step(s, std::true_type{});
    for (; s + 1 <= nrows; s += 2) {
        step(s, std::false_type{});
        step(s + 1, std::false_type{});
    }
    if (s <= nrows) {  // here s >= 64
        step(s, std::false_type{});
        ++s;
    }
    for (; s <= steps; ++s) step(s, std::true_type{});
    if (report) {
        if (anyh) atomicAdd(a.u + I.uv_base + j - 1, (uint32_t)__builtin_popcount(anyh));
        if (vcnt) atomicAdd(a.v + I.uv_base + j - 1, vcnt);
    }
    __threadfence_block();  // the boundary stream is read by this wave's next strip
    __syncthreads();
}

// Resident waves per SIMD the register allocation is held to.  From k = 21 the state needs
// more than 256 registers; holding two waves per SIMD anyway spills the excess to scratch
// (12 B per lane at k = 21, about 360 B at k = 30), which on MI355X runs about 1.6x faster
// (k = 22 to 30) than one wave per SIMD with the excess in AGPRs.  For k = 9 to 14 (175-225
// registers) three waves per SIMD with 20-250 B of spills run 6-16% faster than two; for
// k <= 8 (three waves without spills) four waves, and for k = 16-20 three, ran slower.
template <int K>
constexpr int slice_waves_per_eu() { return K >= 9 && K <= 14 ? 3 : 2; }

template <int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(slice_waves_per_eu<K>())))
void ed_slice_kernel(EdSliceArgs a)
{
    __shared__ __align__(16) uint32_t chunk[64 * SliceLayout<K>::LW];
    __shared__ __align__(16) uint32_t tring[kTextRing * kTextWords];
    const uint32_t lane = threadIdx.x;
    uint32_t *st0 = a.scratch + (size_t)blockIdx.x * 2 * a.stream_words;
    uint32_t *st1 = st0 + a.stream_words;

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
        const EdSliceItem I = a.items[it];
        const uint32_t g0 = a.goff[I.group];
        const uint32_t nrows = a.goff[I.group + 1] - g0;
        const uint32_t *gp = a.gplanes + (size_t)g0 * 6;
        const uint32_t nstrips = (I.m + 63) / 64;
        slice_strip<K, true>(a, I, 0, nrows, gp, st1, st0, nstrips == 1, chunk, tring);
        for (uint32_t strip = 1; strip < nstrips; ++strip) {
            const uint32_t *sin = (strip & 1) ? st0 : st1;
            uint32_t *sout = (strip & 1) ? st1 : st0;
            slice_strip<K, false>(a, I, strip, nrows, gp, sin, sout, strip + 1 == nstrips, chunk, tring);
        }
    }
}


// Fills an LDS chunk of 64 rows from r0 with the previous strip's boundary cells (when
// from_stream; zeros for the first strip, whose left edge is column 0) and the group's text
// planes.  Done by the 64 lanes of the wave that reads the chunk.
template <int K>
__device__ __forceinline__ void slice_chunk_fill(uint32_t *chunk, uint32_t *tring, uint32_t r0, uint32_t nrows,
                                                 bool from_stream, const uint32_t *sin, const uint32_t *gp,
                                                 uint32_t lane)
{
    slice_fill_row<K>(chunk, tring, r0, nrows, from_stream, sin, gp, lane);
}

// One lane's column j of a strip, carried from row to row: D[i-1][j] (what lane L+1 receives
// next), D[i-1][j-1], and the column's hit counts.
template <int K>
struct SliceCol {
    Slice<K> U, G;
    uint32_t xb[5];          // planes of x[j]'s code
    uint32_t eb[kSlicePc];   // planes of e, wave-uniform
    uint32_t vcnt, anyh, j, full;
    bool report;
};

template <int K>
__device__ __forceinline__ void slice_col_begin(SliceCol<K> &c, const EdSliceArgs &a, const EdSliceItem &I,
                                                uint32_t strip, uint32_t lane)
{
    c.j = strip * 64 + lane + 1;  // 1-based column
    const uint32_t code = c.j <= I.m ? a.codes[I.x_start + c.j - 1] : 0u;
#pragma unroll
    for (int b = 0; b < 5; ++b) c.xb[b] = ((code >> b) & 1) ? ~0u : 0u;
    c.report = c.j <= I.m && c.j >= (uint32_t)K;
#pragma unroll
    for (int q = 0; q < kSlicePc; ++q) c.eb[q] = __builtin_amdgcn_readfirstlane(((a.e >> q) & 1) ? ~0u : 0u);
    c.full = c.j <= (uint32_t)K ? ~0u : 0u;
    slice_row0<K>(c.j, c.U);
    slice_row0<K>(c.j - 1, c.G);
    c.vcnt = 0;
    c.anyh = 0;
}

// Step s of the strip's skewed wavefront: lane L computes row s-1-L.  Only a cell's planes
// travel between lanes; the receiver recounts its popcounts.  FIRST: the first strip, whose
// left edge is column 0 and whose first K columns enter left and diag untruncated
// (slice_cell, needing pc and pm of left).  Later strips have only columns j > K
// (slice_cell_masked, needing pm).  CHECK: some lane has no row at this step.
template <int K, bool FIRST, bool CHECK>
__device__ __forceinline__ void slice_step(SliceCol<K> &c, uint32_t s, const uint32_t *chunk, const uint32_t *tring,
                                           uint32_t nrows, uint32_t lane, uint32_t *sout, bool emit)
{
    typedef SliceLayout<K> Lay;
    constexpr int LW = Lay::LW;
    const uint32_t *cr = chunk + ((s - 1) & 63) * LW;
    Slice<K> L;  // D[i][j-1]
    slice_left<K, FIRST>(L, c.U.d, cr);
    uint32_t t[kTextWords];  // row s-1-lane's text planes (garbage on lanes without a row)
    slice_text(tring, s - 1 - lane, t);
    bool active = true;
    if constexpr (CHECK) active = s > lane && s - lane <= nrows;
    if (active) {
        uint32_t ne = 0;
#pragma unroll
        for (int b = 0; b < 5; ++b) ne |= t[b] ^ c.xb[b];
        Slice<K> D;
        if constexpr (FIRST) {
            slice_cell<K>(c.U, L, c.G, ne, c.full, D);
        } else {
            slice_cell_masked<K>(c.U.d, c.U.hi, c.U.pc, L.d, L.pm, c.G.d, c.G.pm, ne, D.d, D.hi, D.pc);
        }
        if (c.report) {
            const uint32_t hit = slice_le_planes<slice_pcq<K>()>(D.pc, c.eb) & t[5];
            c.vcnt += __builtin_popcount(hit);
            c.anyh |= hit;
        }
        c.G = L;
        c.U = D;
        if (emit && lane == 63) slice_emit<K>(sout, s, D.d);
    }
}

// The steps of chunk ch: 64 ch + 1 .. 64 ch + 64, up to nrows + 63.  With SPLIT, chunks
// where every lane has a row run two steps per trip without the row check, so the cell
// planes are not copied between registers on each back edge; without, every step checks
// (fewer registers: the full-cell form of the first strip).
template <int K, bool FIRST, bool SPLIT = true>
__device__ __forceinline__ void slice_run_chunk(SliceCol<K> &c, uint32_t ch, const uint32_t *chunk,
                                                const uint32_t *tring, uint32_t nrows, uint32_t lane, uint32_t *sout,
                                                bool emit)
{
    const uint32_t s0 = 64 * ch + 1, s1 = min(64 * ch + 64, nrows + 63);
    if (SPLIT && s0 >= 64 && s1 <= nrows) {
        uint32_t s = s0;
        for (; s + 1 <= s1; s += 2) {
            slice_step<K, FIRST, false>(c, s, chunk, tring, nrows, lane, sout, emit);
            slice_step<K, FIRST, false>(c, s + 1, chunk, tring, nrows, lane, sout, emit);
        }
        if (s <= s1) slice_step<K, FIRST, false>(c, s, chunk, tring, nrows, lane, sout, emit);
    } else {
        for (uint32_t s = s0; s <= s1; ++s)
            slice_step<K, FIRST, true>(c, s, chunk, tring, nrows, lane, sout, emit);
    }
}

// Per column, the texts with an occurrence and the number of occurrences go to u and v.
template <int K>
__device__ __forceinline__ void slice_col_end(const SliceCol<K> &c, const EdSliceArgs &a, const EdSliceItem &I)
{
    if (c.report) {
        if (c.anyh) atomicAdd(a.u + I.uv_base + c.j - 1, (uint32_t)__builtin_popcount(c.anyh));
        if (c.vcnt) atomicAdd(a.v + I.uv_base + c.j - 1, c.vcnt);
    }
}

// Chunks of 64 rows a strip runs through: its steps are 1 .. nrows + 63.
__device__ __forceinline__ uint32_t slice_chunks(uint32_t nrows) { return (nrows + 63 + 63) / 64; }

// Pipelined form for items that are few and long (kSlicePipe waves per item).  Wave w takes
// strips w, w + P, w + 2P, ... and all waves advance in ticks of one 64-row chunk, with a
// workgroup barrier between ticks.  Strip t starts two ticks after strip t - 1 (the rows
// lane 63 of strip t - 1 emits during its chunks c and c + 1 are the input of strip t's
// chunk c), and a wave starts its next strip when the previous one is done:
//     start(t) = (t / P) * Cp + 2 (t % P),  Cp = max(chunks per strip, 2P).
// Strip t writes boundary buffer t mod (P + 1) and strip t + 1 reads it; the next strip to
// write that buffer, t + P + 1, starts Cp >= chunks ticks after strip t + 1, when its reads
// are done.  An item of S strips then takes about (S / P) * Cp ticks instead of S * chunks.
template <int K>
__global__ __launch_bounds__(64 * kSlicePipe) __attribute__((amdgpu_waves_per_eu(slice_waves_per_eu<K>())))
void ed_slice_pipe_kernel(EdSliceArgs a)
{
    typedef SliceLayout<K> Lay;
    constexpr uint32_t P = kSlicePipe;
    __shared__ __align__(16) uint32_t chunks[P][64 * Lay::LW];
    __shared__ __align__(16) uint32_t trings[P][kTextRing * kTextWords];
    __shared__ uint32_t item_slot;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t *chunk = chunks[wave];
    uint32_t *tring = trings[wave];
    uint32_t *bufs = a.scratch + (size_t)blockIdx.x * (P + 1) * a.stream_words;

    for (;;) {
        __syncthreads();  // every wave has read item_slot
        if (threadIdx.x == 0) item_slot = atomicAdd(a.queue, 1u);
        __syncthreads();
        const uint32_t it = item_slot;
        if (it >= a.nitems) break;
        const EdSliceItem I = a.items[it];
        const uint32_t g0 = a.goff[I.group];
        const uint32_t nrows = a.goff[I.group + 1] - g0;
        const uint32_t *gp = a.gplanes + (size_t)g0 * 6;
        const uint32_t nstrips = (I.m + 63) / 64, nch = slice_chunks(nrows);
        const uint32_t cp = nch > 2 * P ? nch : 2 * P;
        const uint32_t ticks = ((nstrips - 1) / P) * cp + 2 * ((nstrips - 1) % P) + nch;
        SliceCol<K> col;
        uint32_t T = 0;
        // Strip 0, whose first K columns take the full cell, is wave 0's during ticks
        // [0, nch); its own loop keeps the full-cell registers out of the loop below.  Every
        // wave passes two barriers per tick in either loop.
        if (wave == 0) {
            slice_col_begin<K>(col, a, I, 0, lane);
            for (; T < nch; ++T) {
                __syncthreads();
                slice_chunk_fill<K>(chunk, tring, 64 * T, nrows, false, nullptr, gp, lane);
                __syncthreads();
                slice_run_chunk<K, true, false>(col, T, chunk, tring, nrows, lane, bufs, nstrips > 1);
            }
            slice_col_end<K>(col, a, I);
        }
        for (; T < ticks; ++T) {
            // this wave's strip (never strip 0 here) and chunk at tick T, wave-uniform
            uint32_t t = 0, ch = 0;
            bool act = false;
            if (T >= 2 * wave) {
                const uint32_t i = (T - 2 * wave) / cp;
                ch = (T - 2 * wave) - i * cp;
                t = wave + P * i;
                act = t > 0 && ch < nch && t < nstrips;
            }
            __syncthreads();  // boundary cells of the earlier ticks are written
            if (act) {
                const uint32_t *sin = bufs + (size_t)((t - 1) % (P + 1)) * a.stream_words;
                slice_chunk_fill<K>(chunk, tring, 64 * ch, nrows, true, sin, gp, lane);
            }
            __syncthreads();  // the chunk is in LDS
            if (act) {
                uint32_t *sout = bufs + (size_t)(t % (P + 1)) * a.stream_words;
                if (ch == 0) slice_col_begin<K>(col, a, I, t, lane);
                slice_run_chunk<K, false>(col, ch, chunk, tring, nrows, lane, sout, t + 1 < nstrips);
                if (ch + 1 == nch) slice_col_end<K>(col, a, I);
            }
        }
    }
}

// Launch and boundary-stream row size for K0 <= k <= K1.  The kernels are instantiated in
// one translation unit per range of K (ed_slice_kernels*.hip), so the ranges compile in
// parallel.
template <int K, int K1>
hipError_t launch_slice_range(const EdSliceArgs &a, uint32_t grid, bool pipe, hipStream_t s)
{
    if constexpr (K > K1) {
        return hipErrorInvalidValue;
    } else {
        if (a.k != (uint32_t)K) return launch_slice_range<K + 1, K1>(a, grid, pipe, s);
        if (pipe) hipLaunchKernelGGL(ed_slice_pipe_kernel<K>, dim3(grid), dim3(64 * kSlicePipe), 0, s, a);
        else hipLaunchKernelGGL(ed_slice_kernel<K>, dim3(grid), dim3(64), 0, s, a);
        return hipGetLastError();
    }
}

template <int K, int K1>
uint32_t slice_stream_row_words_range(uint32_t k)
{
    if constexpr (K > K1) return 0;
    else return k == (uint32_t)K ? (uint32_t)SliceLayout<K>::SWS : slice_stream_row_words_range<K + 1, K1>(k);
}

// The ranges instantiated outside ed_slice_kernels.hip.
hipError_t launch_ed_slice_k15(const EdSliceArgs &a, uint32_t grid, bool pipe, hipStream_t s);
uint32_t ed_slice_stream_row_words_k15(uint32_t k);
hipError_t launch_ed_slice_k23(const EdSliceArgs &a, uint32_t grid, bool pipe, hipStream_t s);
uint32_t ed_slice_stream_row_words_k23(uint32_t k);

}  // namespace motex
