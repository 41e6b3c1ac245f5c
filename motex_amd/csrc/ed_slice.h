// Text-parallel ("sliced") MaxShift edit recurrence: one cell of 32 texts at once.
//
// The scalar recurrence (functions.cc:437-447, bitminmax functions.cc:923-967) keeps one
// word D[i][j] per cell (row i of the text, column j of x):
//     j <= K: D = min*(shift(up)|1, shift(left)|1, shift(diag)|delta)
//     j >  K: D = min*(shift(up)|1, shiftc(left)|1, shiftc(diag)|delta)
// with shift(a) = a << 1, shiftc(a) = (a & (2^(K-1)-1)) << 1 and min* = fewest set bits,
// largest word on ties.  Here the same cell of 32 different texts (same x, same i and j) is
// computed at once: plane p of a Slice holds bit p of the 32 words, one text per bit, so a
// shift is a renaming of planes and every operation is a bitwise op over 32 texts.  Each
// Slice also carries its popcount pc and the popcount pm of its low K-1 bits (the bits
// shiftc keeps), as 5 bit-sliced planes each; the candidates' popcounts then follow from
// their sources by a bit-sliced increment, and only pm of the selected word is counted.
//
// Selected words stay below 2K+1 bits (width <= K + popcount, popcount <= K), so W = 2K
// planes hold every word; the candidates carry one more plane (shift(up) of a full-width
// up) for the comparison, which is never selected.  The functions are plain C++ on 32-bit
// masks, shared by the GPU kernel and the host check of this file against the scalar
// recurrence (tests/slice_check.cpp).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MOTEX_HD __host__ __device__ __forceinline__
#else
#define MOTEX_HD inline
#endif

namespace motex {

constexpr int kSlicePc = 5;  // popcount planes held (counts up to 31)

// Popcount planes a factor length K needs: a selected word's popcount is at most K (the left
// candidate holds it there, and the host check of this file asserts it), a candidate's at
// most K + 1, so bits(K + 1) planes carry them; the planes above stay zero and are not
// computed.
template <int K>
constexpr int slice_pcq()
{
    return K + 1 < 16 ? 4 : 5;
}

// Truth tables of the 3-input logic op (v_bitop3_b32: result bit = table bit number
// 4 src0 + 2 src1 + src2) used below; the host check of this file evaluates them against the
// plain expressions the host build uses.
constexpr uint32_t kBop3Mux = 0xE4;  // src2 ? src0 : src1
constexpr uint32_t kBop3And = 0xC0;  // src0 & src1
constexpr uint32_t kBop3Gt = 0xB2;   // majority(src0, ~src1, src2)
constexpr uint32_t kBop3Lt = 0x8E;   // majority(~src0, src1, src2)
constexpr uint32_t kBop3Or3 = 0xFE;  // src0 | src1 | src2
constexpr uint32_t kBop3AndNot = 0x30;   // src0 & ~src1
constexpr uint32_t kBop3AndNot2 = 0x10;  // src0 & ~src1 & ~src2
constexpr uint32_t kBop3AndNotUnless = 0xB0;  // src0 & ~(src1 & ~src2)

template <int K> struct SliceW { static constexpr int W = 2 * K; };

template <int K>
struct Slice {
    uint32_t d[SliceW<K>::W];
    uint32_t pc[kSlicePc];  // popcount of the word
    uint32_t pm[kSlicePc];  // popcount of bits 0 .. K-2
    uint32_t hi;            // any of bits K-1 .. W-1 (slice_cell_masked, which keeps only this)
};

// sel ? t : f, per bit (one 3-input logic op on gfx950: the compiler's own choice for this
// pattern, v_bfi_b32, issues at about 60% of the v_bitop3_b32 rate there)
MOTEX_HD uint32_t slice_mux(uint32_t sel, uint32_t t, uint32_t f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(t, f, sel, kBop3Mux);
#else
    return (t & sel) | (f & ~sel);
#endif
}

// x & m for a mask m that is all ones or all zeros per lane (kept a logic op, not a select)
MOTEX_HD uint32_t slice_and(uint32_t x, uint32_t m)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(x, m, x, kBop3And);
#else
    return x & m;
#endif
}

MOTEX_HD uint32_t slice_or3(uint32_t a, uint32_t b, uint32_t c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, kBop3Or3);
#else
    return a | b | c;
#endif
}

MOTEX_HD uint32_t slice_andnot(uint32_t a, uint32_t s)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, s, a, kBop3AndNot);
#else
    return a & ~s;
#endif
}

MOTEX_HD uint32_t slice_andnot2(uint32_t a, uint32_t s1, uint32_t s2)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, s1, s2, kBop3AndNot2);
#else
    return a & ~s1 & ~s2;
#endif
}

// a & ~(h & ~s): a, with the lanes of h cleared except where s is set
MOTEX_HD uint32_t slice_andnot_unless(uint32_t a, uint32_t h, uint32_t s)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, h, s, kBop3AndNotUnless);
#else
    return a & ~(h & ~s);
#endif
}

// out = in + carry (carry: one 0/1 per text) over the low NQ planes
template <int NQ = kSlicePc>
MOTEX_HD void slice_inc(const uint32_t in[kSlicePc], uint32_t carry, uint32_t out[kSlicePc])
{
    for (int q = 0; q < NQ; ++q) {
        out[q] = in[q] ^ carry;
        carry = in[q] & carry;
    }
    for (int q = NQ; q < kSlicePc; ++q) out[q] = 0;
}

// Bit-sliced popcount of N planes into kSlicePc planes: full adders (3 planes -> sum and
// carry, two 3-input logic ops each) at weight 1, then at weight 2, the weight-4 carries
// rippled into the top planes.
MOTEX_HD void slice_csa(uint32_t &s, uint32_t a, uint32_t b, uint32_t &carry)
{
    carry = (s & a) | (s & b) | (a & b);
    s ^= a ^ b;
}

MOTEX_HD void slice_ripple(uint32_t *h, int n, uint32_t c)
{
    for (int q = 0; q < n; ++q) {
        const uint32_t t = h[q] & c;
        h[q] ^= c;
        c = t;
    }
}

template <int N, int NQ = kSlicePc>
MOTEX_HD void slice_popcount(const uint32_t *x, uint32_t out[kSlicePc])
{
    uint32_t ones = 0, twos = 0, hi[kSlicePc - 2] = {0, 0, 0};
    uint32_t c2 = 0;          // a pending weight-2 carry
    bool have_c2 = false;
    for (int i = 0; i < N; i += 2) {
        uint32_t c;
        slice_csa(ones, x[i], i + 1 < N ? x[i + 1] : 0u, c);
        if (have_c2) {
            uint32_t c4;
            slice_csa(twos, c2, c, c4);
            slice_ripple(hi, kSlicePc - 2, c4);
            have_c2 = false;
        } else {
            c2 = c;
            have_c2 = true;
        }
    }
    if (have_c2) {
        const uint32_t c4 = twos & c2;
        twos ^= c2;
        slice_ripple(hi, kSlicePc - 2, c4);
    }
    out[0] = ones;
    out[1] = twos;
    for (int q = 2; q < kSlicePc; ++q) out[q] = q < NQ ? hi[q - 2] : 0u;
}

// Running comparison bits.  Scanning planes from the least significant up, "a > b so far"
// changes only where the planes differ: it becomes a's bit there.  That is
// majority(a, ~b, r), one 3-input logic op per plane (gfx950 v_bitop3_b32, kBop3Gt); for
// "a < b so far" it is majority(~a, b, r) (kBop3Lt).
MOTEX_HD uint32_t slice_gt_step(uint32_t a, uint32_t b, uint32_t r)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, r, kBop3Gt);
#else
    return (a & ~b) | (a & r) | (~b & r);
#endif
}

MOTEX_HD uint32_t slice_lt_step(uint32_t a, uint32_t b, uint32_t r)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, r, kBop3Lt);
#else
    return (~a & b) | (~a & r) | (b & r);
#endif
}

// Texts where (pa, va) is strictly better than (pb, vb) under min*: smaller popcount, or
// equal popcount and larger word.  va / vb have NV planes.  The key (popcount inverted, then
// word) is compared in one ripple from the word's lowest plane up through the popcount
// planes.
template <int NV, int NQ = kSlicePc>
MOTEX_HD uint32_t slice_better(const uint32_t pa[kSlicePc], const uint32_t *va,
                               const uint32_t pb[kSlicePc], const uint32_t *vb)
{
    uint32_t r = 0;
    for (int p = 0; p < NV; ++p) r = slice_gt_step(va[p], vb[p], r);
    for (int q = 0; q < NQ; ++q) r = slice_lt_step(pa[q], pb[q], r);
    return r;
}

// The planes of the constant e (all ones where bit q of e is set), for slice_le_planes.
MOTEX_HD void slice_e_planes(uint32_t e, uint32_t eb[kSlicePc])
{
    for (int q = 0; q < kSlicePc; ++q) eb[q] = ((e >> q) & 1) ? ~0u : 0u;
}

// "count <= e" over the low NQ planes (the counts and e fit them), e given as planes:
// deciding from the low bits up, "c <= e so far" is majority(~c, e, r).
template <int NQ = kSlicePc>
MOTEX_HD uint32_t slice_le_planes(const uint32_t c[kSlicePc], const uint32_t eb[kSlicePc])
{
    uint32_t res = ~0u;
    for (int q = 0; q < NQ; ++q) res = slice_lt_step(c[q], eb[q], res);
    return res;
}

// "count <= e" over kSlicePc planes.
MOTEX_HD uint32_t slice_le(const uint32_t c[kSlicePc], uint32_t e)
{
    uint32_t eb[kSlicePc];
    slice_e_planes(e, eb);
    return slice_le_planes(c, eb);
}

// One cell of 32 texts.  U = D[i-1][j], L = D[i][j-1], G = D[i-1][j-1]; delta = texts whose
// symbol at row i differs from x[j]; full = all ones when j <= K (no truncation) else 0.
// Returns D[i][j].
template <int K>
MOTEX_HD void slice_cell(const Slice<K> &U, const Slice<K> &L, const Slice<K> &G, uint32_t delta,
                         uint32_t full, Slice<K> &D, uint32_t *dropped = nullptr)
{
    constexpr int W = SliceW<K>::W;
    constexpr int NQ = slice_pcq<K>();
    // candidates, W+1 planes each (plane W only for the comparison)
    uint32_t a[W + 1], b[W + 1], c[W + 1];
    a[0] = ~0u;
    b[0] = ~0u;
    c[0] = delta;
    for (int p = 1; p <= W; ++p) {
        a[p] = U.d[p - 1];
        // shiftc keeps bits 0 .. K-2 of the source: planes 1 .. K-1 of the candidate
        b[p] = p <= K - 1 ? L.d[p - 1] : slice_and(L.d[p - 1], full);
        c[p] = p <= K - 1 ? G.d[p - 1] : slice_and(G.d[p - 1], full);
    }
    uint32_t lsrc[kSlicePc], gsrc[kSlicePc], pcc[kSlicePc];
    for (int q = 0; q < kSlicePc; ++q) {
        lsrc[q] = q < NQ ? slice_mux(full, L.pc[q], L.pm[q]) : 0u;
        gsrc[q] = q < NQ ? slice_mux(full, G.pc[q], G.pm[q]) : 0u;
    }
    slice_inc<NQ>(gsrc, delta, pcc);
    // min*: start from a, replace where b then c is strictly better.  a and b both set plane
    // 0 and add one to their source's popcount, so they compare as their sources do.
    uint32_t sel = slice_better<W, NQ>(lsrc, b + 1, U.pc, a + 1);
    uint32_t d[W + 1], pd[kSlicePc], ps[kSlicePc];
    d[0] = ~0u;
    for (int p = 1; p <= W; ++p) d[p] = slice_mux(sel, b[p], a[p]);
    for (int q = 0; q < kSlicePc; ++q) ps[q] = q < NQ ? slice_mux(sel, lsrc[q], U.pc[q]) : 0u;
    slice_inc<NQ>(ps, ~0u, pd);
    sel = slice_better<W + 1, NQ>(pcc, c, pd, d);
    for (int p = 0; p < W; ++p) D.d[p] = slice_mux(sel, c[p], d[p]);
    if (dropped) *dropped |= slice_mux(sel, c[W], d[W]);  // host check of the width bound
    for (int q = 0; q < kSlicePc; ++q) D.pc[q] = q < NQ ? slice_mux(sel, pcc[q], pd[q]) : 0u;
    slice_popcount<K - 1, NQ>(D.d, D.pm);
    D.hi = 0;  // kept by the masked cell only
}

// The cell for columns j > K, where left and diag enter truncated: their candidates have no
// plane at or above K, so only pm of left and diag and pc of up are needed, and the result's
// pm is left to the reader (slice_pm).  up: planes 0..K-2, uhi, pc; left, diag: planes
// 0..K-2, pm.  Returns planes 0..K-1, dhi and pc.
//
// Planes K..W of the candidates are up's alone (left and diag are zero there), so both
// comparisons decide those planes from one OR of them: in the ripple, a plane where the
// other side is zero turns "better so far" off wherever up's plane is set.  That OR is uhi,
// the OR of up's planes K-1..W-1, and it is all the cell needs of them: the selected word's
// planes K..W-1 are up's planes K-1..W-2 where neither left nor diag won, and up's plane W-1
// is never kept (selected words stay below 2K+1 bits), so the result's OR of planes
// K-1..W-1 is dhi = plane K-1 | (uhi where neither won).  Words past the first strip thus
// carry K planes and one OR instead of 2K planes.
template <int K>
MOTEX_HD void slice_cell_masked(const uint32_t *ud, uint32_t uhi, const uint32_t *upc, const uint32_t *ld,
                                const uint32_t *lpm, const uint32_t *gd, const uint32_t *gpm,
                                uint32_t delta, uint32_t *dd, uint32_t &dhi, uint32_t *dpc)
{
    constexpr int NQ = slice_pcq<K>();
    // candidate planes 0..K-1: a = shift(up) | 1, b = shiftc(left) | 1, c = shiftc(diag) | delta
    uint32_t a[K], b[K], c[K];
    a[0] = ~0u;
    b[0] = ~0u;
    c[0] = delta;
    for (int p = 1; p < K; ++p) {
        a[p] = ud[p - 1];
        b[p] = ld[p - 1];
        c[p] = gd[p - 1];
    }
    uint32_t pcc[kSlicePc];
    slice_inc<NQ>(gpm, delta, pcc);
    // min*: start from a, replace where b then c is strictly better.  a and b both set plane
    // 0 and add one to their source's popcount, so they compare as their sources do.
    uint32_t r = 0;
    for (int p = 1; p < K; ++p) r = slice_gt_step(b[p], a[p], r);
    r = slice_andnot(r, uhi);
    for (int q = 0; q < NQ; ++q) r = slice_lt_step(lpm[q], upc[q], r);
    const uint32_t sel = r;
    uint32_t d[K], pd[kSlicePc], ps[kSlicePc];
    d[0] = ~0u;
    for (int p = 1; p < K; ++p) d[p] = slice_mux(sel, b[p], a[p]);
    for (int q = 0; q < kSlicePc; ++q) ps[q] = q < NQ ? slice_mux(sel, lpm[q], upc[q]) : 0u;
    slice_inc<NQ>(ps, ~0u, pd);
    // c against d: d's planes K..W are a's where b was not chosen
    r = 0;
    for (int p = 0; p < K; ++p) r = slice_gt_step(c[p], d[p], r);
    r = slice_andnot_unless(r, uhi, sel);
    for (int q = 0; q < NQ; ++q) r = slice_lt_step(pcc[q], pd[q], r);
    const uint32_t sel2 = r;
    for (int p = 0; p < K; ++p) dd[p] = slice_mux(sel2, c[p], d[p]);
    dhi = dd[K - 1] | slice_andnot2(uhi, sel, sel2);
    for (int q = 0; q < kSlicePc; ++q) dpc[q] = q < NQ ? slice_mux(sel2, pcc[q], pd[q]) : 0u;
}

// pm of a word: the popcount of its low K-1 planes.
template <int K>
MOTEX_HD void slice_pm(const uint32_t *d, uint32_t pm[kSlicePc])
{
    slice_popcount<K - 1, slice_pcq<K>()>(d, pm);
}

// Row 0 of column j (1-based): min(j, K) ones in every text.
template <int K>
MOTEX_HD void slice_row0(uint32_t j, Slice<K> &D)
{
    constexpr int W = SliceW<K>::W;
    const uint32_t r = j < (uint32_t)K ? j : (uint32_t)K;
    for (int p = 0; p < W; ++p) D.d[p] = (uint32_t)p < r ? ~0u : 0u;
    const uint32_t rm = r < (uint32_t)(K - 1) ? r : (uint32_t)(K - 1);
    for (int q = 0; q < kSlicePc; ++q) {
        D.pc[q] = ((r >> q) & 1) ? ~0u : 0u;
        D.pm[q] = ((rm >> q) & 1) ? ~0u : 0u;
    }
    D.hi = r >= (uint32_t)K ? ~0u : 0u;
}

// Column 0: the zero word.
template <int K>
MOTEX_HD void slice_zero(Slice<K> &D)
{
    for (int p = 0; p < SliceW<K>::W; ++p) D.d[p] = 0;
    for (int q = 0; q < kSlicePc; ++q) D.pc[q] = D.pm[q] = 0;
}

}  // namespace motex
