// Structured-motif batches: pair work queue, scratch sizing and launches of
// struct_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "runtime.h"

namespace motex {

hipError_t launch_struct(const StructArgs &a, uint32_t grid, size_t lds, bool dense, hipStream_t s);

static const uint32_t kStructHitCap = 2048;  // >= one chunk: 64 lanes x 32 bits
// Two Hamming boxes merge from their occurrence lists, which needs only the counted bits in
// LDS; the gather list of the fallback path then lives in scratch, and the smaller LDS
// footprint lets more waves run (1000 x 600 nt, boxes 8+[10,30]+8: about 14 -> 27 waves per CU).
static const uint32_t kStructPairListLds = 512;
static const uint32_t kStructListCap = 8192; // Hamming occurrences listed per pair and box shape

static int struct_alloc_scratch(motex_batch *b);
static int struct_fill_args(motex_batch *b, StructArgs &a);

// NE-table bit offset of text position 0 for sequences of up to maxlen symbols: the sweep of a
// pair starts at diagonal pos_t - pos_x = -(m - 1), rounded down to a multiple of 32.
static uint32_t struct_pad_bits(uint32_t maxlen)
{
    return std::max<uint32_t>(kPadBits, (maxlen + 63) / 32 * 32);
}

static size_t struct_lds_bytes(const StructArgs &a)
{
    return ((size_t)(a.sigma + 1) * a.wl + (a.m_max + 1) / 2 + a.hcap + 1 + kMaxBoxes +
            (size_t)a.run_words * kMaxBoxes) * 4;
}

// Fills the shape fields of StructArgs from the parameters.
static int struct_shape(const motex_batch *b, StructArgs &a)
{
    const motex_params &p = b->params;
    if (p.nb_gaps > kMaxGaps)
        return set_error(MOTEX_EUNSUPPORTED, "structured motifs with more than %u gaps", kMaxGaps);
    a.nb_gaps = p.nb_gaps;
    a.nb_structs = b->nb_structs;
    a.distance = p.distance;
    a.blen[0] = p.k;
    a.berr[0] = p.e;
    for (uint32_t g = 0; g < p.nb_gaps; ++g) {
        a.gmin[g] = p.gap_min[g];
        a.gmax[g] = p.gap_max[g];
        a.blen[g + 1] = p.box_len[g];
        a.berr[g + 1] = p.box_err[g];
    }
    a.ndist = 0;
    for (uint32_t bx = 0; bx <= p.nb_gaps; ++bx) {
        uint32_t d = 0;
        while (d < a.ndist && !(a.dlen[d] == a.blen[bx] && a.derr[d] == a.berr[bx])) ++d;
        if (d == a.ndist) { a.dlen[d] = a.blen[bx]; a.derr[d] = a.berr[bx]; a.ndist++; }
        a.dist_of[bx] = d;
    }
    return MOTEX_OK;
}

struct StructPlan {
    std::vector<StructItem> items;
};

// Items: ordered pairs, or (sym) unordered pairs i <= j whose box DPs serve both (i, j) and
// (j, i): Hamming boxes are symmetric, so the transposed bitmaps are those of (j, i).
static void plan_pairs(const SeqSet &ss, uint32_t total_mot, const std::vector<std::pair<uint32_t, uint32_t>> *only,
                       std::vector<StructItem> &out, bool sym = false)
{
    out.clear();
    auto add = [&](uint32_t i, uint32_t j, uint32_t uvt) {
        if (ss.len[i] < total_mot || ss.len[j] < total_mot) return;
        out.push_back({ss.start[i], ss.len[i], ss.start[j], ss.len[j], (uint32_t)ss.uv_off[i], uvt});
    };
    if (only) {
        for (auto &pr : *only) add(pr.first, pr.second, kNoSplit);
    } else if (sym) {
        for (uint32_t i = 0; i < ss.num; ++i)
            for (uint32_t j = i; j < ss.num; ++j) add(i, j, j == i ? kNoSplit : (uint32_t)ss.uv_off[j]);
    } else {
        for (uint32_t i = 0; i < ss.num; ++i)
            for (uint32_t j = 0; j < ss.num; ++j) add(i, j, kNoSplit);
    }
}

// Pair-symmetric structured plan: Hamming boxes only; MOTEX_SYM=0 turns it off.
static bool struct_symmetric(const motex_batch *b)
{
    const char *env = getenv("MOTEX_SYM");
    if (env && env[0] == '0') return false;
    return b->params.distance == 0;
}

int struct_prepare(motex_batch *b)
{
    const SeqSet &ss = b->ss;
    StructArgs a{};
    int rc = struct_shape(b, a);
    if (rc) return rc;
    uint32_t maxlen = 0;
    for (uint32_t i = 0; i < ss.num; ++i) maxlen = std::max(maxlen, ss.len[i]);
    if (maxlen > kStructMaxLen)
        return set_error(MOTEX_EUNSUPPORTED, "structured motifs: sequences longer than %u", kStructMaxLen);

    hipError_t e;
    if (b->params.distance == 0) {
        // NE tables over the whole text; every pair's diagonal window must fit inside
        uint32_t wl = 64 + ((maxlen - 1) >> 5) + 3;
        // a pair's first block starts up to m - 1 + 31 positions left of the text's start
        const uint32_t pad = struct_pad_bits(maxlen);
        uint64_t W = (ss.total + pad) / 32 + kDiagPerBlock / 32 + wl + 8;
        if ((uint64_t)(ss.sigma + 1) * wl >= 65536 || ss.sigma * wl >= 65536)
            return set_error(MOTEX_EUNSUPPORTED, "alphabet too large for the structured kernel");
        b->hd.W = (uint32_t)W;
        b->hd.wl = wl;
        if ((e = b->d_ne.alloc(W * ss.sigma * 4)) != hipSuccess) return hip_fail(e, "alloc NE tables");
        if ((e = launch_ne_build((const uint8_t *)b->d_codes.p, (uint32_t)ss.total, ss.sigma, (uint32_t)W,
                                 (uint32_t *)b->d_ne.p, bstream(b), pad)) != hipSuccess)
            return hip_fail(e, "NE build");
    }
    {
        StructArgs t{};
        if ((rc = struct_fill_args(b, t))) return rc;
        if (struct_lds_bytes(t) > 65536)
            return set_error(MOTEX_EUNSUPPORTED,
                             "structured motifs: sequences of %u symbols over %u letters exceed the kernel's LDS window",
                             maxlen, ss.sigma);
    }
    std::vector<StructItem> items;
    b->stats.symmetric = struct_symmetric(b) ? 1u : 0u;
    plan_pairs(ss, total_motif_length(&b->params), nullptr, items, b->stats.symmetric != 0);
    {
        std::vector<uint64_t> cost(items.size());
        for (size_t t = 0; t < items.size(); ++t) cost[t] = (uint64_t)items[t].m * items[t].n;
        const std::vector<uint32_t> order = order_by_cost_desc(cost);
        std::vector<StructItem> sorted(order.size());
        for (size_t t = 0; t < order.size(); ++t) sorted[t] = items[order[t]];
        items.swap(sorted);
    }
    if ((e = b->d_items.alloc(std::max<size_t>(items.size(), 1) * sizeof(StructItem))) != hipSuccess)
        return hip_fail(e, "alloc items");
    if (!items.empty() &&
        (e = hipMemcpyAsync(b->d_items.p, items.data(), items.size() * sizeof(StructItem), hipMemcpyHostToDevice, bstream(b))) != hipSuccess)
        return hip_fail(e, "copy items");
    b->ed_nitems = (uint32_t)items.size();
    if ((e = b->d_aux2.alloc(std::max<size_t>(b->S.size(), 1) * 4)) != hipSuccess) return hip_fail(e, "alloc S");
    if (!b->S.empty() &&
        (e = hipMemcpyAsync(b->d_aux2.p, b->S.data(), b->S.size() * 4, hipMemcpyHostToDevice, bstream(b))) != hipSuccess)
        return hip_fail(e, "copy S");
    b->tile = choose_tile(b->ctx, b->params, ss);
    // cells: one DP per distinct box shape and pair
    double cells = 0;
    for (const StructItem &it : items)
        for (uint32_t d = 0; d < a.ndist; ++d) {
            if (it.m >= a.dlen[d]) cells += (double)(it.m - a.dlen[d] + 1) * it.n;
            if (it.uv_base_t != kNoSplit && it.n >= a.dlen[d]) cells += (double)(it.n - a.dlen[d] + 1) * it.m;
        }
    b->stats.cells = cells;
    if ((rc = struct_alloc_scratch(b))) return rc;
    return hipStreamSynchronize(bstream(b)) == hipSuccess ? MOTEX_OK : set_error(MOTEX_EDEVICE, "NE build");
}

// Fills every StructArgs field that does not depend on the launch (items, grid).
static int struct_fill_args(motex_batch *b, StructArgs &a)
{
    const SeqSet &ss = b->ss;
    int rc = struct_shape(b, a);
    if (rc) return rc;
    uint32_t maxlen = 0;
    for (uint32_t i = 0; i < ss.num; ++i) maxlen = std::max(maxlen, ss.len[i]);
    a.ne = (const uint32_t *)b->d_ne.p;
    a.codes = (const uint8_t *)b->d_codes.p;
    a.S = (const uint32_t *)b->d_aux2.p;
    a.u = (uint32_t *)b->d_u.p;
    a.v = (uint32_t *)b->d_v.p;
    a.queue = (uint32_t *)b->d_queue.p;
    a.W = b->hd.W;
    a.sigma = ss.sigma;
    a.wl = b->params.distance == 0 ? b->hd.wl : 0;
    a.m_max = std::max<uint32_t>(maxlen, 1);
    a.pad_bits = struct_pad_bits(maxlen);
    const bool pair_lists = b->params.distance == 0 && b->params.nb_gaps == 1;
    // list merge: LDS keeps the counted bits (nb_structs x m_max), up to kStructPairListLds
    // words; larger ones go to scratch
    a.hcap = pair_lists ? (uint32_t)std::min<uint64_t>(kStructPairListLds, ((uint64_t)b->nb_structs * a.m_max + 31) / 32)
                        : kStructHitCap;
    a.gcap = kStructHitCap;
    a.gather_scratch = pair_lists ? 1u : 0u;
    a.bm_words = (uint64_t)((maxlen + 31) / 32) * maxlen;
    a.stream_words = (maxlen + 128) * 4;  // up to four words per row (128-bit edit words)
    a.lcap = b->params.distance == 0 ? kStructListCap : 0;
    // a pair's diagonals q = pt - px run over (-m, n), from a multiple of 32 below: at most
    // (m + n + 31) / 2048 + 1 blocks
    a.run_words = 2 + (2 * a.m_max + 31) / kDiagPerBlock + 1;
    // bitmaps (Hamming: and their transposes), edit streams, counted bits, lists, gather list
    a.scratch_words = (b->params.distance == 0 ? 2ull : 1ull) * a.ndist * a.bm_words + 2ull * a.stream_words +
                      ((uint64_t)a.nb_structs * a.m_max + 31) / 32 + 1 + (uint64_t)a.ndist * a.lcap +
                      (a.gather_scratch ? a.gcap : 0u);
    return MOTEX_OK;
}

static int struct_launch(motex_batch *b, const void *items, uint32_t nitems, uint32_t grid,
                         uint32_t persistent, hipStream_t s)
{
    StructArgs a{};
    int rc = struct_fill_args(b, a);
    if (rc) return rc;
    a.items = (const StructItem *)items;
    a.nitems = nitems;
    a.persistent = persistent;
    uint64_t need = (uint64_t)grid * a.scratch_words * 4;
    hipError_t e;
    if (b->d_aux.bytes < need && (e = batch_realloc_scratch(b, need)) != hipSuccess) return hip_fail(e, "alloc structured scratch");
    a.scratch = (uint32_t *)b->d_aux.p;
    size_t lds = struct_lds_bytes(a);
    if (lds > 65536) return set_error(MOTEX_EUNSUPPORTED, "structured kernel LDS window too large");
    // eight waves per SIMD for the list merge when the LDS lets at least 24 waves per CU in
    const bool dense = a.distance == 0 && a.nb_gaps == 1 && lds * 24 <= 160 * 1024;
    if ((e = launch_struct(a, grid, lds, dense, s)) != hipSuccess) return hip_fail(e, "launch structured kernel");
    return MOTEX_OK;
}

static uint32_t struct_grid(motex_batch *b, uint32_t nitems)
{
    // bound the scratch to about 2 GiB
    const SeqSet &ss = b->ss;
    uint32_t maxlen = 0;
    for (uint32_t i = 0; i < ss.num; ++i) maxlen = std::max(maxlen, ss.len[i]);
    uint64_t per = ((uint64_t)(maxlen + 31) / 32 * maxlen * (b->params.nb_gaps + 1) * 2 + 2ull * (maxlen + 128) * 2 +
                    ((uint64_t)b->nb_structs * maxlen + 31) / 32 + 1 +
                    (uint64_t)(b->params.nb_gaps + 1) * kStructListCap) * 4;
    uint64_t cap = std::max<uint64_t>(1, (2ull << 30) / std::max<uint64_t>(per, 1));
    uint64_t g = std::min<uint64_t>((uint64_t)b->ctx->num_cu * b->tile.waves_per_cu, nitems);
    return (uint32_t)std::max<uint64_t>(1, std::min(g, cap));
}

// Scratch of the persistent grid, allocated at batch creation so that runs allocate nothing.
static int struct_alloc_scratch(motex_batch *b)
{
    StructArgs a{};
    int rc = struct_fill_args(b, a);
    if (rc) return rc;
    const uint64_t need = (uint64_t)struct_grid(b, std::max<uint32_t>(b->ed_nitems, 1)) * a.scratch_words * 4;
    hipError_t e = batch_realloc_scratch(b, need);
    return e == hipSuccess ? MOTEX_OK : hip_fail(e, "alloc structured scratch");
}

int struct_launch_all(motex_batch *b)
{
    if (!b->ed_nitems) return MOTEX_OK;
    uint32_t grid = struct_grid(b, b->ed_nitems);
    if (b->grid_cap) grid = std::min(grid, b->grid_cap);
    return struct_launch(b, b->d_items.p, b->ed_nitems, grid, 1, bstream(b));
}

static int struct_launch_pairs(motex_batch *b, const std::vector<std::pair<uint32_t, uint32_t>> &pairs, double *ms_out)
{
    hipStream_t s = bstream(b);
    hipError_t e;
    std::vector<StructItem> items;
    plan_pairs(b->ss, total_motif_length(&b->params), &pairs, items);
    if (ms_out) *ms_out = 0;
    if (items.empty()) return MOTEX_OK;
    DevBuf d;
    if ((e = d.alloc(items.size() * sizeof(StructItem))) != hipSuccess) return hip_fail(e, "alloc items");
    if ((e = hipMemcpyAsync(d.p, items.data(), items.size() * sizeof(StructItem), hipMemcpyHostToDevice, bstream(b))) != hipSuccess)
        return hip_fail(e, "copy items");
    if ((e = hipEventRecord(b->ev0, s)) != hipSuccess) return hip_fail(e, "record plain start");
    for (size_t t = 0; t < items.size(); ++t) {
        int rc = struct_launch(b, (const StructItem *)d.p + t, 1, 1, 0, s);
        if (rc) return rc;
    }
    if ((e = hipEventRecord(b->ev1, s)) != hipSuccess) return hip_fail(e, "record plain end");
    if ((e = hipEventSynchronize(b->ev1)) != hipSuccess) return hip_fail(e, "plain structured kernel");
    float ms = 0;
    if ((e = hipEventElapsedTime(&ms, b->ev0, b->ev1)) != hipSuccess) return hip_fail(e, "plain timing");
    if (ms_out) *ms_out = ms;
    return MOTEX_OK;
}

int struct_run_plain(motex_batch *b, uint32_t first_x, uint32_t num_x)
{
    const SeqSet &ss = b->ss;
    double total = 0;
    for (uint32_t i = first_x; i < first_x + num_x && i < ss.num; ++i) {
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        for (uint32_t j = 0; j < ss.num; ++j) pairs.push_back({i, j});
        double ms = 0;
        int rc = struct_launch_pairs(b, pairs, &ms);
        if (rc) return rc;
        total += ms;
    }
    b->stats.kernel_ms = total;
    b->stats.main_kernel_ms = total;
    return MOTEX_OK;
}

int struct_pair_run(motex_batch *b, uint32_t i, uint32_t j)
{
    std::vector<std::pair<uint32_t, uint32_t>> pairs{{i, j}};
    double ms = 0;
    int rc = struct_launch_pairs(b, pairs, &ms);
    b->stats.kernel_ms = ms;
    b->stats.main_kernel_ms = ms;
    return rc;
}

}  // namespace motex
