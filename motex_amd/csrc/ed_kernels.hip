// Edit-distance FLASM kernels for gfx950 (MaxShift bit-vectors).
//
// What is computed (motifs_extraction_ed, functions.cc:734-768, over libflasm::flasm_ed,
// libflasm.h:71): the MaxShift edit recurrence of functions.cc:437-447 with bitminmax
// (functions.cc:923-967), for text rows i = 1..n and x columns j = 1..m:
//     j <= k: D = min*(shift(up)|1, shift(left)|1, shift(diag)|delta)
//     j >  k: D = min*(shiftc(left)|1, shift(up)|1, shiftc(diag)|delta)
// where shiftc(a) = (a & (2^(k-1)-1)) << 1, min* picks the word with the fewest set bits
// and, among those, the largest; row 0 holds min(j, k) ones and column 0 holds 0.  Every
// (i, j) with j >= k and popcount(D) <= e is an occurrence of the factor ending at j-1.
//
// Layout.  A wave owns a strip of 64*C columns (C adjacent columns per lane) and streams
// the text rows through it as a skewed wavefront: at step s lane L works on row s-L, so
// the left neighbour's word arrives by one DPP wave shift per step.  Strips of one x run
// one after another; lane 63's last column of every row is collected 64 rows at a time
// and becomes the next strip's left boundary.  Words are 32-bit for k <= 14 (every
// selected word is narrower than 2k+1 bits) and 64-bit up to k = 32.  min* uses one
// unsigned key per candidate, popcount in the top bits above the complemented word, so a
// single v_min3 selects it.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "motex_internal.h"

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
};

// One wavefront step.  ACTIVE_CHECK selects the fill/drain form in which lanes outside
// rows 1..n keep their state.
template <int C, bool WIDE, bool ACTIVE_CHECK>
__device__ __forceinline__ void ed_step(EdState<C, WIDE> &st, uint32_t s, uint32_t n, uint32_t e,
                                        uint32_t inT, typename EdWord<WIDE>::T inD)
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
            st.any[c] |= hit;
            st.cnt[c] += hit ? 1u : 0u;
        }
        st.diag0 = left;
        st.pubD = lft;
        st.pubT = tc;
    }
}

template <int C, bool WIDE>
__global__ __launch_bounds__(64) void ed_kernel(EdArgs a)
{
    typedef typename EdWord<WIDE>::T T;
    extern __shared__ uint32_t smem[];
    uint32_t *acc_u = smem;
    uint32_t *acc_v = smem + a.lds_cols;
    const uint32_t lane = threadIdx.x;
    const uint32_t k = a.k, e = a.e;
    const T ymask = (k >= 2) ? (((T)1 << (k - 1)) - 1) : (T)0;
    const uint32_t words_per = WIDE ? 2 : 1;
    uint32_t *buf0 = a.scratch + (size_t)blockIdx.x * 2 * a.stream_words;
    uint32_t *buf1 = buf0 + a.stream_words;
    const uint32_t strip_cols = 64 * C;

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
        const EdItemDev I = a.items[it];
        const uint32_t m = I.m;
        const bool use_lds = m <= a.lds_cols;
        __syncthreads();
        if (use_lds)
            for (uint32_t c = lane; c < m; c += 64) { acc_u[c] = 0; acc_v[c] = 0; }
        __syncthreads();
        const uint32_t nstrips = (m + strip_cols - 1) / strip_cols;
        for (uint32_t yj = I.j0; yj < I.j0 + I.ny; ++yj) {
            const uint32_t ys = a.seq_start[yj];
            const uint32_t n = a.seq_start[yj + 1] - ys;
            if (n < k) continue;  // the pair is skipped (motex.cc:536-540)
            const uint8_t *text = a.codes + ys;
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
                    st.pcode[c] = j <= m ? a.codes[I.x_start + j - 1] : 0xFFFFu;
                    st.cnt[c] = 0;
                    st.any[c] = false;
                    st.ok[c] = j >= k && j <= m;
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
                for (uint32_t s = 1; s <= steps; ++s) {
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
                    if (s >= 64 && s <= n)
                        ed_step<C, WIDE, false>(st, s, n, e, inT, inD);
                    else
                        ed_step<C, WIDE, true>(st, s, n, e, inT, inD);
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
                }
                // fold this strip's counts for the current text
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    uint32_t j = col0 + c + 1;
                    uint32_t anyb = st.any[c] ? 1u : 0u;
                    if (j >= k && j <= m && (anyb || st.cnt[c])) {
                        if (use_lds) {
                            acc_u[j - 1] += anyb;
                            acc_v[j - 1] += st.cnt[c];
                        } else {
                            if (anyb) atomicAdd(a.u + I.uv_base + j - 1, anyb);
                            atomicAdd(a.v + I.uv_base + j - 1, st.cnt[c]);
                        }
                    }
                }
                __threadfence();  // boundary stream visible to the next strip's loads
            }
        }
        __syncthreads();
        if (use_lds) {
            for (uint32_t c = lane; c < m; c += 64) {
                uint32_t cu = acc_u[c], cv = acc_v[c];
                if (cu) atomicAdd(a.u + I.uv_base + c, cu);
                if (cv) atomicAdd(a.v + I.uv_base + c, cv);
            }
        }
    }
    (void)words_per;
}

hipError_t launch_ed(const EdArgs &a, bool wide, uint32_t grid, size_t lds, hipStream_t s)
{
    if (wide)
        hipLaunchKernelGGL((ed_kernel<2, true>), dim3(grid), dim3(64), lds, s, a);
    else
        hipLaunchKernelGGL((ed_kernel<4, false>), dim3(grid), dim3(64), lds, s, a);
    return hipGetLastError();
}

uint32_t ed_strip_cols(bool wide) { return wide ? 128 : 256; }

}  // namespace motex
