// Host check of the text-parallel edit step (motex_amd/csrc/ed_slice.h): runs the sliced
// recurrence over one x and a group of up to 32 texts with plain loops and compares the
// per-column u, v summed over the group with the oracle's per-pair edit operator
// (orc_ed_pair).  Also checks that no selected word needs the extra candidate plane.
// Usage: slice_check seed k e m nmin nmax sigma ntexts
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

#include "ed_slice.h"

extern "C" int orc_ed_pair(const char *x, uint32_t m, const char *y, uint32_t n, uint32_t k,
                           uint32_t e, uint32_t *u, uint32_t *v);

static uint64_t rng_state;
static uint32_t rnd(uint32_t n)
{
    rng_state += 0x9E3779B97F4A7C15ull;
    uint64_t z = rng_state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z % n);
}

template <int K>
static int run(uint32_t e, uint32_t m, uint32_t nmin, uint32_t nmax, uint32_t sigma, uint32_t nt)
{
    using namespace motex;
    std::vector<std::string> texts(nt);
    std::string x(m, 'A');
    for (auto &c : x) c = (char)('A' + rnd(sigma));
    uint32_t nlong = 0;
    for (auto &t : texts) {
        t.resize(nmin + rnd(nmax - nmin + 1));
        for (auto &c : t) c = (char)('A' + rnd(sigma));
        // low-complexity runs make wide words and ties
        if (rnd(3) == 0)
            for (size_t i = 0; i < t.size(); ++i) if (rnd(2)) t[i] = x[i % m];
        nlong = t.size() > nlong ? (uint32_t)t.size() : nlong;
    }
    // reference: per pair, summed over the texts (pairs with n < k are skipped, as the driver does)
    std::vector<uint32_t> ru(m, 0), rv(m, 0);
    for (auto &t : texts)
        if (t.size() >= K) orc_ed_pair(x.data(), m, t.data(), (uint32_t)t.size(), K, e, ru.data(), rv.data());
    // sliced: column by column over all rows, one bit per text
    std::vector<Slice<K>> prev(m + 1), cur(m + 1);
    for (uint32_t j = 0; j <= m; ++j) {
        if (j == 0) slice_zero<K>(prev[0]);
        else slice_row0<K>(j, prev[j]);
    }
    std::vector<uint32_t> anyhit(m, 0), su(m, 0), sv(m, 0);
    int wide = 0, masked_bad = 0, pc_bad = 0;
    for (uint32_t i = 1; i <= nlong; ++i) {
        uint32_t valid = 0;
        for (uint32_t t = 0; t < nt; ++t)
            if (texts[t].size() >= K && i <= texts[t].size()) valid |= 1u << t;
        slice_zero<K>(cur[0]);
        for (uint32_t j = 1; j <= m; ++j) {
            uint32_t delta = 0;
            for (uint32_t t = 0; t < nt; ++t)
                if (i <= texts[t].size() && texts[t][i - 1] != x[j - 1]) delta |= 1u << t;
            const uint32_t full = j <= (uint32_t)K ? ~0u : 0u;
            // width: the candidates' extra plane must never be selected
            uint32_t dropped = 0;
            slice_cell<K>(prev[j], cur[j - 1], prev[j - 1], delta, full, cur[j], &dropped);
            wide |= (dropped & valid) != 0;
            // the carried popcount planes must equal the word's popcount, which stays <= K
            for (uint32_t t = 0; t < nt; ++t) {
                if (!((valid >> t) & 1)) continue;
                uint32_t cnt = 0, held = 0;
                for (int p = 0; p < SliceW<K>::W; ++p) cnt += (cur[j].d[p] >> t) & 1;
                for (int q = 0; q < kSlicePc; ++q) held |= ((cur[j].pc[q] >> t) & 1) << q;
                pc_bad |= cnt != held || cnt > (uint32_t)K;
            }
            if (j > (uint32_t)K) {
                // the truncated-column form must give the same low planes, popcounts and OR of
                // planes K-1..W-1, from up's low planes and that OR alone
                constexpr int W = SliceW<K>::W;
                uint32_t uhi = 0, dhi_full = 0;
                for (int p = K - 1; p < W; ++p) {
                    uhi |= prev[j].d[p];
                    dhi_full |= cur[j].d[p];
                }
                Slice<K> alt;
                slice_cell_masked<K>(prev[j].d, uhi, prev[j].pc, cur[j - 1].d, cur[j - 1].pm, prev[j - 1].d,
                                     prev[j - 1].pm, delta, alt.d, alt.hi, alt.pc);
                slice_pm<K>(alt.d, alt.pm);
                for (int p = 0; p < K; ++p) masked_bad |= alt.d[p] != cur[j].d[p];
                masked_bad |= alt.hi != dhi_full;
                for (int q = 0; q < kSlicePc; ++q)
                    masked_bad |= alt.pc[q] != cur[j].pc[q] || alt.pm[q] != cur[j].pm[q];
            }
            if (j >= (uint32_t)K) {
                uint32_t hit = slice_le(cur[j].pc, e) & valid;
                anyhit[j - 1] |= hit;
                sv[j - 1] += __builtin_popcount(hit);
            }
        }
        std::swap(prev, cur);
    }
    for (uint32_t j = 0; j < m; ++j) su[j] = __builtin_popcount(anyhit[j]);
    int bad = 0;
    for (uint32_t j = 0; j < m; ++j)
        if (su[j] != ru[j] || sv[j] != rv[j]) {
            if (bad < 5) printf("col %u: sliced u=%u v=%u oracle u=%u v=%u\n", j, su[j], sv[j], ru[j], rv[j]);
            ++bad;
        }
    uint64_t tv = 0;
    for (uint32_t j = 0; j < m; ++j) tv += rv[j];
    printf("k=%d e=%u m=%u texts=%u mismatched_columns=%d occurrences=%llu wide_state=%d masked_form_differs=%d "
           "popcount_differs=%d\n",
           K, e, m, nt, bad, (unsigned long long)tv, wide, masked_bad, pc_bad);
    return bad || wide || masked_bad || pc_bad ? 1 : 0;
}

// The device build computes slice_mux, slice_and, slice_gt_step and slice_lt_step with
// v_bitop3_b32 and the kBop3* tables; the host build with the plain expressions.  Evaluate
// each table as the instruction does (bit 4 src0 + 2 src1 + src2 of the table) and compare.
static uint32_t bitop3(uint32_t a, uint32_t b, uint32_t c, uint32_t table)
{
    uint32_t r = 0;
    for (int t = 0; t < 32; ++t) {
        const uint32_t idx = (((a >> t) & 1) << 2) | (((b >> t) & 1) << 1) | ((c >> t) & 1);
        r |= ((table >> idx) & 1u) << t;
    }
    return r;
}

static int check_tables()
{
    // the 8 input combinations, one per bit position
    const uint32_t a = 0xF0u, b = 0xCCu, c = 0xAAu;
    int bad = 0;
    bad += bitop3(a, b, c, motex::kBop3Mux) != motex::slice_mux(c, a, b);
    bad += bitop3(a, b, a, motex::kBop3And) != motex::slice_and(a, b);
    bad += bitop3(a, b, c, motex::kBop3Gt) != motex::slice_gt_step(a, b, c);
    bad += bitop3(a, b, c, motex::kBop3Lt) != motex::slice_lt_step(a, b, c);
    bad += bitop3(a, b, c, motex::kBop3Or3) != motex::slice_or3(a, b, c);
    bad += bitop3(a, b, a, motex::kBop3AndNot) != motex::slice_andnot(a, b);
    bad += bitop3(a, b, c, motex::kBop3AndNot2) != motex::slice_andnot2(a, b, c);
    bad += bitop3(a, b, c, motex::kBop3AndNotUnless) != motex::slice_andnot_unless(a, b, c);
    // and the comparisons themselves on small numbers, exhaustively
    for (uint32_t x = 0; x < 64; ++x)
        for (uint32_t y = 0; y < 64; ++y) {
            uint32_t gt = 0, lt = 0;
            for (int p = 0; p < 6; ++p) {
                const uint32_t xp = (x >> p) & 1 ? ~0u : 0u, yp = (y >> p) & 1 ? ~0u : 0u;
                gt = bitop3(xp, yp, gt, motex::kBop3Gt);
                lt = bitop3(xp, yp, lt, motex::kBop3Lt);
            }
            bad += (gt != 0) != (x > y) || (lt != 0) != (x < y);
        }
    printf("tables: %d mismatches\n", bad);
    return bad ? 1 : 0;
}

int main(int argc, char **argv)
{
    if (argc == 2 && !strcmp(argv[1], "tables")) return check_tables();
    if (argc < 9) return 2;
    rng_state = strtoull(argv[1], 0, 10);
    const int k = atoi(argv[2]);
    const uint32_t e = atoi(argv[3]), m = atoi(argv[4]), nmin = atoi(argv[5]), nmax = atoi(argv[6]);
    const uint32_t sigma = atoi(argv[7]), nt = atoi(argv[8]);
    switch (k) {
    case 1: return run<1>(e, m, nmin, nmax, sigma, nt);
    case 2: return run<2>(e, m, nmin, nmax, sigma, nt);
    case 3: return run<3>(e, m, nmin, nmax, sigma, nt);
    case 4: return run<4>(e, m, nmin, nmax, sigma, nt);
    case 5: return run<5>(e, m, nmin, nmax, sigma, nt);
    case 6: return run<6>(e, m, nmin, nmax, sigma, nt);
    case 8: return run<8>(e, m, nmin, nmax, sigma, nt);
    case 10: return run<10>(e, m, nmin, nmax, sigma, nt);
    case 12: return run<12>(e, m, nmin, nmax, sigma, nt);
    case 14: return run<14>(e, m, nmin, nmax, sigma, nt);
    case 15: return run<15>(e, m, nmin, nmax, sigma, nt);
    case 20: return run<20>(e, m, nmin, nmax, sigma, nt);
    case 30: return run<30>(e, m, nmin, nmax, sigma, nt);
    default: return 2;
    }
}
