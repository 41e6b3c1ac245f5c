// Device-side building blocks of the MaxShift edit recurrence (see ed_kernels.hip for the
// recurrence and the wavefront layout).  Shared by the edit and structured kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace motex {

template <bool WIDE> struct EdWord;
template <> struct EdWord<false> {
    typedef uint32_t T;
    static __device__ __forceinline__ uint32_t pc(T x) { return __builtin_popcount(x); }
};
template <> struct EdWord<true> {
    typedef uint64_t T;
    static __device__ __forceinline__ uint32_t pc(T x) { return __builtin_popcountll(x); }
};

__device__ __forceinline__ uint32_t dpp_shr1(uint32_t x)
{
    // wave_shr:1 (lane L receives lane L-1; lane 0 receives 0)
    return __builtin_amdgcn_update_dpp(0u, x, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t dpp_shl1(uint32_t x)
{
    // wave_shl:1 (lane L receives lane L+1; lane 63 receives 0)
    return __builtin_amdgcn_update_dpp(0u, x, 0x130, 0xF, 0xF, false);
}

// min*: among a, b, c the word with the fewest set bits, largest on ties; returns its
// popcount in *pcout.
template <bool WIDE>
__device__ __forceinline__ typename EdWord<WIDE>::T minstar(typename EdWord<WIDE>::T a,
                                                            typename EdWord<WIDE>::T b,
                                                            typename EdWord<WIDE>::T c,
                                                            uint32_t *pcout)
{
    if constexpr (!WIDE) {
        const uint32_t M = 0x0FFFFFFFu;
        uint32_t ka = (__builtin_popcount(a) << 28) | (M & ~a);
        uint32_t kb = (__builtin_popcount(b) << 28) | (M & ~b);
        uint32_t kc = (__builtin_popcount(c) << 28) | (M & ~c);
        uint32_t km = min(min(ka, kb), kc);
        *pcout = km >> 28;
        return M & ~km;
    } else {
        uint32_t pa = __builtin_popcountll(a), pb = __builtin_popcountll(b), pcc = __builtin_popcountll(c);
        uint32_t mn = min(min(pa, pb), pcc);
        uint64_t sa = pa == mn ? a : 0, sb = pb == mn ? b : 0, sc = pcc == mn ? c : 0;
        uint64_t r = sa > sb ? sa : sb;
        r = r > sc ? r : sc;
        *pcout = mn;
        return r;
    }
}

__device__ __forceinline__ uint32_t lo32(uint64_t x) { return (uint32_t)x; }
__device__ __forceinline__ uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }

template <int C, bool WIDE>
struct EdState {
    typedef typename EdWord<WIDE>::T T;
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
template <int C, bool WIDE, bool ACTIVE_CHECK, bool BITMAP>
__device__ __forceinline__ void ed_step(EdState<C, WIDE> &st, uint32_t s, uint32_t n, uint32_t e,
                                        uint32_t inT, typename EdWord<WIDE>::T inD,
                                        const EdOut &out, uint32_t col0)
{
    typedef typename EdWord<WIDE>::T T;
    const uint32_t lane = threadIdx.x;
    // left neighbour's word and text symbol of this lane's row
    uint32_t tc = dpp_shr1(st.pubT);
    T left;
    if constexpr (WIDE) {
        uint32_t l0 = dpp_shr1(lo32(st.pubD)), l1 = dpp_shr1(hi32(st.pubD));
        l0 = lane == 0 ? lo32(inD) : l0;
        l1 = lane == 0 ? hi32(inD) : l1;
        left = ((uint64_t)l1 << 32) | l0;
    } else {
        uint32_t l0 = dpp_shr1(st.pubD);
        left = lane == 0 ? inD : l0;
    }
    tc = lane == 0 ? inT : tc;
    bool active = true;
    if constexpr (ACTIVE_CHECK) active = (s > lane) && (s - lane <= n);
    if (active) {
        T lft = left, dg = st.diag0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            T ca = ((lft & st.ym[c]) << 1) | (T)1;
            T cb = (st.up[c] << 1) | (T)1;
            T cc = ((dg & st.ym[c]) << 1) | (T)(tc != st.pcode[c]);
            uint32_t p;
            T d = minstar<WIDE>(ca, cb, cc, &p);
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
// mode each reportable column's occurrences are added to acc (LDS, indexed by column) or,
// when acc is null, straight to u/v in global memory; in bitmap mode they go to out.bm.
template <int C, bool WIDE, bool BITMAP>
__device__ void ed_pair(const uint8_t *__restrict__ xcodes, uint32_t m,
                        const uint8_t *__restrict__ text, uint32_t n, uint32_t k, uint32_t e,
                        uint32_t *buf0, uint32_t *buf1, uint32_t *gu, uint32_t *gv,
                        const EdOut &out)
{
    typedef typename EdWord<WIDE>::T T;
    const uint32_t lane = threadIdx.x;
    const T ymask = (k >= 2) ? (((T)1 << (k - 1)) - 1) : (T)0;
    const uint32_t strip_cols = 64 * C;
    const uint32_t nstrips = (m + strip_cols - 1) / strip_cols;
    for (uint32_t strip = 0; strip < nstrips; ++strip) {
        uint32_t *bin = (strip & 1) ? buf1 : buf0;
        uint32_t *bout = (strip & 1) ? buf0 : buf1;
        const bool last = strip + 1 == nstrips;
        EdState<C, WIDE> st;
        const uint32_t col0 = strip * strip_cols + lane * C;  // 0-based column of c = 0
#pragma unroll
        for (int c = 0; c < C; ++c) {
            uint32_t j = col0 + c + 1;  // 1-based
            uint32_t r = j < k ? j : k;
            st.up[c] = r >= 8 * sizeof(T) ? ~(T)0 : (((T)1 << r) - 1);
            st.ym[c] = j <= k ? ~(T)0 : ymask;
            st.pcode[c] = j <= m ? xcodes[j - 1] : 0xFFFFu;
            st.cnt[c] = 0;
            st.any[c] = false;
            st.ok[c] = j >= k && j <= m;
            st.hw[c] = 0;
        }
        {
            uint32_t j = col0;  // column left of c = 0, row 0
            uint32_t r = j < k ? j : k;
            st.diag0 = r >= 8 * sizeof(T) ? ~(T)0 : (((T)1 << r) - 1);
        }
        st.pubD = st.up[C - 1];
        st.pubT = 0;
        T coll = 0;
        uint32_t coll_n = 0;
        uint32_t out_row = 0;  // next row index (0-based) the collector writes
        const uint32_t steps = n + 63;
        uint32_t vinT = 0;
        T vinD = 0;
        // one wavefront step s; CHECK: fill/drain form (rows outside 1..n keep their state)
        auto step = [&](uint32_t s, auto check) {
            const uint32_t r0 = s - 1;  // 0-based row lane 0 starts this step
            if ((r0 & 63) == 0) {
                uint32_t rr = r0 + lane;
                vinT = rr < n ? text[rr] : 0u;
                if (strip == 0) {
                    vinD = 0;
                } else if constexpr (WIDE) {
                    vinD = rr < n ? (((uint64_t)bin[2 * rr + 1] << 32) | bin[2 * rr]) : 0;
                } else {
                    vinD = rr < n ? bin[rr] : 0;
                }
            }
            const uint32_t sl = r0 & 63;
            uint32_t inT = __builtin_amdgcn_readlane(vinT, sl);
            T inD;
            if constexpr (WIDE) {
                uint32_t lo = __builtin_amdgcn_readlane(lo32(vinD), sl);
                uint32_t hi = __builtin_amdgcn_readlane(hi32(vinD), sl);
                inD = ((uint64_t)hi << 32) | lo;
            } else {
                inD = __builtin_amdgcn_readlane(vinD, sl);
            }
            ed_step<C, WIDE, decltype(check)::value, BITMAP>(st, s, n, e, inT, inD, out, col0);
            if (!last && s >= 64) {
                // collect lane 63's last-column word of row s-63
                if constexpr (WIDE) {
                    uint32_t c0 = dpp_shl1(lo32(coll)), c1 = dpp_shl1(hi32(coll));
                    c0 = lane == 63 ? lo32(st.pubD) : c0;
                    c1 = lane == 63 ? hi32(st.pubD) : c1;
                    coll = ((uint64_t)c1 << 32) | c0;
                } else {
                    uint32_t c0 = dpp_shl1(coll);
                    coll = lane == 63 ? st.pubD : c0;
                }
                if (++coll_n == 64 || s == steps) {
                    // lanes 64-coll_n .. 63 hold rows out_row .. out_row+coll_n-1
                    int rr = (int)out_row + (int)lane - (int)(64 - coll_n);
                    if (rr >= (int)out_row && rr < (int)n) {
                        if constexpr (WIDE) {
                            bout[2 * rr] = lo32(coll);
                            bout[2 * rr + 1] = hi32(coll);
                        } else {
                            bout[rr] = coll;
                        }
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
