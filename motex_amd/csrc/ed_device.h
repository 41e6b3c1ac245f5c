// Device-side building blocks of the MaxShift edit recurrence (see ed_kernels.hip for the
// recurrence and the wavefront layout).  Shared by the edit and structured kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace motex {

// Words of the recurrence: W = 0 (32-bit, k <= 14), 1 (64-bit, k <= 32), 2 (128-bit,
// k <= 64).  Every selected word is narrower than 2k+1 bits, so these widths are exact.
struct U128 {
    uint64_t lo, hi;
};
__device__ __forceinline__ U128 operator&(U128 a, U128 b) { return {a.lo & b.lo, a.hi & b.hi}; }
__device__ __forceinline__ U128 operator|(U128 a, U128 b) { return {a.lo | b.lo, a.hi | b.hi}; }
__device__ __forceinline__ bool operator>(U128 a, U128 b) { return a.hi > b.hi || (a.hi == b.hi && a.lo > b.lo); }

template <int W> struct EdWord;
template <> struct EdWord<0> {
    typedef uint32_t T;
    static constexpr int NW = 1;  // 32-bit pieces of a word
    static __device__ __forceinline__ uint32_t pc(T x) { return __builtin_popcount(x); }
    static __device__ __forceinline__ T shl1(T x) { return x << 1; }
    static __device__ __forceinline__ T bit(bool b) { return b ? 1u : 0u; }
    static __device__ __forceinline__ T ones(uint32_t r) { return r >= 32 ? ~0u : ((1u << r) - 1); }
    static __device__ __forceinline__ uint32_t piece(T x, int) { return x; }
    static __device__ __forceinline__ T from(const uint32_t *p) { return p[0]; }
};
template <> struct EdWord<1> {
    typedef uint64_t T;
    static constexpr int NW = 2;
    static __device__ __forceinline__ uint32_t pc(T x) { return __builtin_popcountll(x); }
    static __device__ __forceinline__ T shl1(T x) { return x << 1; }
    static __device__ __forceinline__ T bit(bool b) { return b ? 1ull : 0ull; }
    static __device__ __forceinline__ T ones(uint32_t r) { return r >= 64 ? ~0ull : ((1ull << r) - 1); }
    static __device__ __forceinline__ uint32_t piece(T x, int i) { return (uint32_t)(x >> (32 * i)); }
    static __device__ __forceinline__ T from(const uint32_t *p) { return ((uint64_t)p[1] << 32) | p[0]; }
};
template <> struct EdWord<2> {
    typedef U128 T;
    static constexpr int NW = 4;
    static __device__ __forceinline__ uint32_t pc(T x)
    {
        return __builtin_popcountll(x.lo) + __builtin_popcountll(x.hi);
    }
    static __device__ __forceinline__ T shl1(T x) { return {x.lo << 1, (x.hi << 1) | (x.lo >> 63)}; }
    static __device__ __forceinline__ T bit(bool b) { return {b ? 1ull : 0ull, 0ull}; }
    static __device__ __forceinline__ T ones(uint32_t r)
    {
        if (r >= 128) return {~0ull, ~0ull};
        if (r >= 64) return {~0ull, r == 64 ? 0ull : ((1ull << (r - 64)) - 1)};
        return {r == 0 ? 0ull : ((1ull << r) - 1), 0ull};
    }
    static __device__ __forceinline__ uint32_t piece(T x, int i)
    {
        return (uint32_t)((i < 2 ? x.lo : x.hi) >> (32 * (i & 1)));
    }
    static __device__ __forceinline__ T from(const uint32_t *p)
    {
        return {((uint64_t)p[1] << 32) | p[0], ((uint64_t)p[3] << 32) | p[2]};
    }
};

// Word width of the recurrence for factors of length k.
__host__ __device__ constexpr int ed_width(uint32_t k) { return k <= 14 ? 0 : (k <= 32 ? 1 : 2); }
// Columns per lane of the one-text-per-lane kernels for each width.
__host__ __device__ constexpr int ed_cols(int w) { return w == 0 ? 4 : (w == 1 ? 2 : 1); }

// (bound_ctrl: the lane without a source receives zero, so no old value is set up first)
__device__ __forceinline__ uint32_t dpp_shr1(uint32_t x)
{
    // wave_shr:1 (lane L receives lane L-1; lane 0 receives 0)
    return __builtin_amdgcn_update_dpp(0u, x, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t dpp_shl1(uint32_t x)
{
    // wave_shl:1 (lane L receives lane L+1; lane 63 receives 0)
    return __builtin_amdgcn_update_dpp(0u, x, 0x130, 0xF, 0xF, true);
}

// min*: among a, b, c the word with the fewest set bits, largest on ties; returns its
// popcount in *pcout.
template <int W>
__device__ __forceinline__ typename EdWord<W>::T minstar(typename EdWord<W>::T a, typename EdWord<W>::T b,
                                                         typename EdWord<W>::T c, uint32_t *pcout)
{
    typedef EdWord<W> E;
    if constexpr (W == 0) {
        // one key per candidate, popcount above the complemented word: one v_min3 selects
        const uint32_t M = 0x0FFFFFFFu;
        uint32_t ka = (__builtin_popcount(a) << 28) | (M & ~a);
        uint32_t kb = (__builtin_popcount(b) << 28) | (M & ~b);
        uint32_t kc = (__builtin_popcount(c) << 28) | (M & ~c);
        uint32_t km = min(min(ka, kb), kc);
        *pcout = km >> 28;
        return M & ~km;
    } else {
        const uint32_t pa = E::pc(a), pb = E::pc(b), pcc = E::pc(c);
        const uint32_t mn = min(min(pa, pb), pcc);
        typename E::T r = E::bit(false);
        if (pa == mn) r = a;
        if (pb == mn && b > r) r = b;
        if (pcc == mn && c > r) r = c;
        *pcout = mn;
        return r;
    }
}

template <int C, int W>
struct EdState {
    typedef typename EdWord<W>::T T;
    T up[C], diag0;          // diag of column 0 (the others are the previous column's up)
    T ym[C];                 // truncation mask per column (all ones while j <= k)
    uint32_t pcode[C];       // x symbol per column
    uint32_t cnt[C];         // occurrences of this column against the current text
    bool any[C];             // this column has an occurrence in the current text
    bool ok[C];              // this column is a reportable factor (k <= j <= m)
    T pubD;                  // this lane's last-column word (read by lane L+1 next step)
    uint32_t pubT;           // text symbol of this lane's current row
    uint32_t hw[C];          // bitmap output: hit bits of the current 32 rows per column
};

// Where a pair's occurrences go: per-column counts (u/v folding) or a hit bitmap laid out
// word-major, bm[(pos_t >> 5) * m + pos_x], bit pos_t & 31.
struct EdOut {
    uint32_t *bm;    // bitmap base (bitmap mode)
    uint32_t m;      // columns of the bitmap
};

// One wavefront step.  ACTIVE_CHECK selects the fill/drain form in which lanes outside
// rows 1..n keep their state.
template <int C, int W, bool ACTIVE_CHECK, bool BITMAP>
__device__ __forceinline__ void ed_step(EdState<C, W> &st, uint32_t s, uint32_t n, uint32_t e,
                                        uint32_t inT, typename EdWord<W>::T inD,
                                        const EdOut &out, uint32_t col0)
{
    typedef EdWord<W> E;
    typedef typename E::T T;
    const uint32_t lane = threadIdx.x;
    // left neighbour's word and text symbol of this lane's row
    uint32_t tc = dpp_shr1(st.pubT);
    uint32_t lp[E::NW];
#pragma unroll
    for (int i = 0; i < E::NW; ++i) {
        const uint32_t l = dpp_shr1(E::piece(st.pubD, i));
        lp[i] = lane == 0 ? E::piece(inD, i) : l;
    }
    const T left = E::from(lp);
    tc = lane == 0 ? inT : tc;
    bool active = true;
    if constexpr (ACTIVE_CHECK) active = (s > lane) && (s - lane <= n);
    if (active) {
        T lft = left, dg = st.diag0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            T ca = E::shl1(lft & st.ym[c]) | E::bit(true);
            T cb = E::shl1(st.up[c]) | E::bit(true);
            T cc = E::shl1(dg & st.ym[c]) | E::bit(tc != st.pcode[c]);
            uint32_t p;
            T d = minstar<W>(ca, cb, cc, &p);
            dg = st.up[c];       // D[i-1][j] is the next column's diag
            st.up[c] = d;
            lft = d;
            bool hit = (p <= e) && st.ok[c];
            if constexpr (BITMAP) {
                const uint32_t r = s - lane - 1;  // 0-based row of this lane
                st.hw[c] |= (hit ? 1u : 0u) << (r & 31);
                if ((r & 31) == 31 || r + 1 == n) {
                    const uint32_t j = col0 + c;  // 0-based column
                    if (j < out.m) out.bm[(size_t)(r >> 5) * out.m + j] = st.hw[c];
                    st.hw[c] = 0;
                }
            } else {
                st.any[c] |= hit;
                st.cnt[c] += hit ? 1u : 0u;
            }
        }
        st.diag0 = left;
        st.pubD = lft;
        st.pubT = tc;
    }
}

// One (x, text) pair: the strips of x one after another over the whole text.  In count
// mode each reportable column's occurrences go straight to u/v in global memory (gu, gv);
// in bitmap mode they go to out.bm.  The boundary streams hold EdWord<W>::NW words per row.
template <int C, int W, bool BITMAP>
__device__ void ed_pair(const uint8_t *__restrict__ xcodes, uint32_t m,
                        const uint8_t *__restrict__ text, uint32_t n, uint32_t k, uint32_t e,
                        uint32_t *buf0, uint32_t *buf1, uint32_t *gu, uint32_t *gv,
                        const EdOut &out)
{
    typedef EdWord<W> E;
    typedef typename E::T T;
    constexpr int NW = E::NW;
    const uint32_t lane = threadIdx.x;
    const T ymask = E::ones(k >= 2 ? k - 1 : 0);
    const uint32_t strip_cols = 64 * C;
    const uint32_t nstrips = (m + strip_cols - 1) / strip_cols;
    for (uint32_t strip = 0; strip < nstrips; ++strip) {
        uint32_t *bin = (strip & 1) ? buf1 : buf0;
        uint32_t *bout = (strip & 1) ? buf0 : buf1;
        const bool last = strip + 1 == nstrips;
        EdState<C, W> st;
        const uint32_t col0 = strip * strip_cols + lane * C;  // 0-based column of c = 0
#pragma unroll
        for (int c = 0; c < C; ++c) {
            uint32_t j = col0 + c + 1;  // 1-based
            uint32_t r = j < k ? j : k;
            st.up[c] = E::ones(r);
            st.ym[c] = j <= k ? E::ones(128) : ymask;
            st.pcode[c] = j <= m ? xcodes[j - 1] : 0xFFFFu;
            st.cnt[c] = 0;
            st.any[c] = false;
            st.ok[c] = j >= k && j <= m;
            st.hw[c] = 0;
        }
        {
            uint32_t j = col0;  // column left of c = 0, row 0
            uint32_t r = j < k ? j : k;
            st.diag0 = E::ones(r);
        }
        st.pubD = st.up[C - 1];
        st.pubT = 0;
        T coll = E::bit(false);
        uint32_t coll_n = 0;
        uint32_t out_row = 0;  // next row index (0-based) the collector writes
        const uint32_t steps = n + 63;
        uint32_t vinT = 0;
        T vinD = E::bit(false);
        // one wavefront step s; CHECK: fill/drain form (rows outside 1..n keep their state)
        auto step = [&](uint32_t s, auto check) {
            const uint32_t r0 = s - 1;  // 0-based row lane 0 starts this step
            if ((r0 & 63) == 0) {
                uint32_t rr = r0 + lane;
                vinT = rr < n ? text[rr] : 0u;
                uint32_t vp[NW];
#pragma unroll
                for (int i = 0; i < NW; ++i) vp[i] = (strip != 0 && rr < n) ? bin[(size_t)NW * rr + i] : 0u;
                vinD = E::from(vp);
            }
            const uint32_t sl = r0 & 63;
            uint32_t inT = __builtin_amdgcn_readlane(vinT, sl);
            uint32_t ip[NW];
#pragma unroll
            for (int i = 0; i < NW; ++i) ip[i] = __builtin_amdgcn_readlane(E::piece(vinD, i), sl);
            const T inD = E::from(ip);
            ed_step<C, W, decltype(check)::value, BITMAP>(st, s, n, e, inT, inD, out, col0);
            if (!last && s >= 64) {
                // collect lane 63's last-column word of row s-63
                uint32_t cp[NW];
#pragma unroll
                for (int i = 0; i < NW; ++i) {
                    const uint32_t c0 = dpp_shl1(E::piece(coll, i));
                    cp[i] = lane == 63 ? E::piece(st.pubD, i) : c0;
                }
                coll = E::from(cp);
                if (++coll_n == 64 || s == steps) {
                    // lanes 64-coll_n .. 63 hold rows out_row .. out_row+coll_n-1
                    int rr = (int)out_row + (int)lane - (int)(64 - coll_n);
                    if (rr >= (int)out_row && rr < (int)n) {
#pragma unroll
                        for (int i = 0; i < NW; ++i) bout[(size_t)NW * rr + i] = cp[i];
                    }
                    out_row += coll_n;
                    coll_n = 0;
                }
            }
        };
        // fill (s < 64), steady state (64 <= s <= n, two steps per trip: with one, the
        // column words are copied between registers on every back edge), drain (s > n)
        uint32_t s = 1;
        const uint32_t fill_end = n < 63 ? n : 63;
        for (; s <= fill_end; ++s) step(s, std::true_type{});
        for (; s + 1 <= n; s += 2) {
            step(s, std::false_type{});
            step(s + 1, std::false_type{});
        }
        if (s <= n) {  // here s >= 64: a steady-state step
            step(s, std::false_type{});
            ++s;
        }
        for (; s <= steps; ++s) step(s, std::true_type{});
        if constexpr (!BITMAP) {
            // fold this strip's counts for the current text
#pragma unroll
            for (int c = 0; c < C; ++c) {
                uint32_t j = col0 + c + 1;
                uint32_t anyb = st.any[c] ? 1u : 0u;
                if (j >= k && j <= m && anyb) {
                    atomicAdd(gu + j - 1, 1u);
                    atomicAdd(gv + j - 1, st.cnt[c]);
                }
            }
        }
        __threadfence_block();  // boundary stream visible to the next strip's loads (same wave)
    }
}

}  // namespace motex
