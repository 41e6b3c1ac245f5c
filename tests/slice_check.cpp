// Host check of the text-parallel edit step (motex_amd/csrc/ed_slice.h): runs the sliced
// recurrence over one x and a group of up to 32 texts with plain loops and compares the
// per-column u, v summed over the group with the oracle's per-pair edit operator
// (orc_ed_pair).  Also checks that no selected word needs the extra candidate plane.
// Usage: slice_check seed k e m nmin nmax sigma ntexts
#include <cstdint>
#include <cstdio>
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
    int wide = 0, masked_bad = 0;
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
            if (j > (uint32_t)K) {
                // the truncated-column form must give the same word and popcounts
                Slice<K> alt;
                uint32_t dr2 = 0;
                slice_cell_masked<K>(prev[j].d, prev[j].pc, cur[j - 1].d, cur[j - 1].pm, prev[j - 1].d,
                                     prev[j - 1].pm, delta, alt.d, alt.pc, &dr2);
                slice_pm<K>(alt.d, alt.pm);
                for (int p = 0; p < SliceW<K>::W; ++p) masked_bad |= alt.d[p] != cur[j].d[p];
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
    printf("k=%d e=%u m=%u texts=%u mismatched_columns=%d occurrences=%llu wide_state=%d masked_form_differs=%d\n",
           K, e, m, nt, bad, (unsigned long long)tv, wide, masked_bad);
    return bad || wide || masked_bad ? 1 : 0;
}

int main(int argc, char **argv)
{
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
    default: return 2;
    }
}
