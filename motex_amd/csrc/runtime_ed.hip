// Edit-distance batches: work-queue planning and launches for ed_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "runtime.h"

namespace motex {

hipError_t launch_ed(const EdArgs &a, bool wide, uint32_t grid, hipStream_t s);
uint32_t ed_strip_cols(bool wide);

static const uint64_t kEdTextPerItem = 8192;  // text symbols per work item

static bool ed_wide(const motex_params &p)
{
    uint32_t longest = p.k;
    for (uint32_t g = 0; g < p.nb_gaps; ++g) longest = std::max(longest, p.box_len[g]);
    return longest > 14;
}

// Items: (x, run of consecutive texts of about kEdTextPerItem symbols).
static void plan_ed(const SeqSet &ss, uint32_t k, const std::vector<uint32_t> &xs,
                    const std::vector<uint32_t> *only_y, std::vector<EdItemDev> &out)
{
    out.clear();
    struct R { uint32_t j0, ny; };
    std::vector<R> ranges;
    if (only_y) {
        for (uint32_t j : *only_y) ranges.push_back({j, 1});
    } else {
        uint32_t j = 0;
        while (j < ss.num) {
            uint32_t j0 = j;
            uint64_t acc = 0;
            while (j < ss.num && (acc == 0 || acc + ss.len[j] <= kEdTextPerItem)) { acc += ss.len[j]; ++j; }
            ranges.push_back({j0, j - j0});
        }
    }
    for (uint32_t i : xs) {
        if (ss.len[i] < k) continue;
        for (const R &r : ranges)
            out.push_back({ss.start[i], ss.len[i], (uint32_t)ss.uv_off[i], r.j0, r.ny, 0});
    }
}

static uint64_t ed_item_cost(const SeqSet &ss, const EdItemDev &it, bool wide)
{
    uint32_t sc = ed_strip_cols(wide);
    uint64_t strips = (it.m + sc - 1) / sc, steps = 0;
    for (uint32_t j = it.j0; j < it.j0 + it.ny; ++j) steps += ss.len[j] + 63;
    return strips * steps;
}

// Boundary-stream words per wave (two streams of this many words each).
static uint32_t ed_stream_words(const motex_batch *b)
{
    uint32_t maxn = 0;
    for (uint32_t j = 0; j < b->ss.num; ++j) maxn = std::max(maxn, b->ss.len[j]);
    return (maxn + 128) * (ed_wide(b->params) ? 2 : 1);
}

static uint32_t ed_grid(const motex_batch *b)
{
    return (uint32_t)std::min<uint64_t>((uint64_t)b->ctx->num_cu * b->tile.waves_per_cu, b->ed_nitems);
}

int ed_prepare(motex_batch *b)
{
    const SeqSet &ss = b->ss;
    std::vector<uint32_t> xs(ss.num);
    for (uint32_t i = 0; i < ss.num; ++i) xs[i] = i;
    std::vector<EdItemDev> items;
    plan_ed(ss, b->params.k, xs, nullptr, items);
    bool wide = ed_wide(b->params);
    std::vector<uint64_t> cost(items.size());
    std::vector<size_t> order(items.size());
    for (size_t t = 0; t < items.size(); ++t) { cost[t] = ed_item_cost(ss, items[t], wide); order[t] = t; }
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t z) { return cost[a] > cost[z]; });
    std::vector<EdItemDev> sorted(items.size());
    for (size_t t = 0; t < items.size(); ++t) sorted[t] = items[order[t]];
    b->ed.items.resize(sorted.size());
    for (size_t t = 0; t < sorted.size(); ++t) b->ed.items[t] = {sorted[t].x_start, sorted[t].j0, sorted[t].ny};
    b->tile = choose_tile(b->ctx, b->params, ss);
    b->tile.variant = wide ? 1 : 0;
    hipError_t e;
    uint64_t ni = std::max<size_t>(sorted.size(), 1);
    if ((e = b->d_items.alloc(ni * sizeof(EdItemDev))) != hipSuccess) return hip_fail(e, "alloc items");
    if (!sorted.empty() &&
        (e = hipMemcpyAsync(b->d_items.p, sorted.data(), sorted.size() * sizeof(EdItemDev), hipMemcpyHostToDevice, bstream(b))) != hipSuccess)
        return hip_fail(e, "copy items");
    b->ed_nitems = (uint32_t)sorted.size();
    // boundary streams of the persistent grid, allocated here so that runs allocate nothing
    const uint64_t need = (uint64_t)std::max<uint32_t>(ed_grid(b), 1) * 2 * ed_stream_words(b) * 4;
    if ((e = batch_realloc_scratch(b, need)) != hipSuccess) return hip_fail(e, "alloc edit streams");
    if ((e = hipStreamSynchronize(bstream(b))) != hipSuccess) return hip_fail(e, "upload items");
    return MOTEX_OK;
}

static int ed_launch(motex_batch *b, const void *items, uint32_t nitems, uint32_t grid,
                     uint32_t persistent, hipStream_t s)
{
    hipError_t e;
    const SeqSet &ss = b->ss;
    bool wide = ed_wide(b->params);
    uint32_t stream_words = ed_stream_words(b);
    uint64_t need = (uint64_t)grid * 2 * stream_words * 4;
    if (b->d_aux.bytes < need && (e = batch_realloc_scratch(b, need)) != hipSuccess) return hip_fail(e, "alloc edit streams");
    EdArgs a{};
    a.codes = (const uint8_t *)b->d_codes.p;
    a.seq_start = (const uint32_t *)b->d_seq_start.p;
    a.items = (const EdItemDev *)items;
    a.u = (uint32_t *)b->d_u.p;
    a.v = (uint32_t *)b->d_v.p;
    a.queue = (uint32_t *)b->d_queue.p;
    a.scratch = (uint32_t *)b->d_aux.p;
    a.stream_words = stream_words;
    a.nitems = nitems;
    a.k = b->params.k;
    a.e = b->params.e;
    a.persistent = persistent;
    if ((e = launch_ed(a, wide, grid, s)) != hipSuccess) return hip_fail(e, "launch edit kernel");
    return MOTEX_OK;
}

int ed_launch_all(motex_batch *b)
{
    if (!b->ed_nitems) return MOTEX_OK;
    uint32_t grid = ed_grid(b);
    if (b->grid_cap) grid = std::min(grid, b->grid_cap);
    return ed_launch(b, b->d_items.p, b->ed_nitems, grid, 1, bstream(b));
}

// Plain form: one launch per pair (one work item, one wave).
static int ed_launch_pairs(motex_batch *b, const std::vector<std::pair<uint32_t, uint32_t>> &pairs, double *ms_out)
{
    motex_ctx *c = b->ctx;
    hipStream_t s = bstream(b);
    hipError_t e;
    const SeqSet &ss = b->ss;
    std::vector<EdItemDev> all;
    for (auto &pr : pairs) {
        if (ss.len[pr.first] < b->params.k || ss.len[pr.second] < b->params.k) continue;
        all.push_back({ss.start[pr.first], ss.len[pr.first], (uint32_t)ss.uv_off[pr.first], pr.second, 1, 0});
    }
    if (ms_out) *ms_out = 0;
    if (all.empty()) return MOTEX_OK;
    DevBuf items;
    if ((e = items.alloc(all.size() * sizeof(EdItemDev))) != hipSuccess) return hip_fail(e, "alloc items");
    if ((e = hipMemcpyAsync(items.p, all.data(), all.size() * sizeof(EdItemDev), hipMemcpyHostToDevice, bstream(b))) != hipSuccess)
        return hip_fail(e, "copy items");
    hipEventRecord(b->ev0, s);
    for (size_t t = 0; t < all.size(); ++t) {
        int rc = ed_launch(b, (const EdItemDev *)items.p + t, 1, 1, 0, s);
        if (rc) return rc;
    }
    hipEventRecord(b->ev1, s);
    if ((e = hipEventSynchronize(b->ev1)) != hipSuccess) return hip_fail(e, "plain edit kernel");
    float ms = 0;
    hipEventElapsedTime(&ms, b->ev0, b->ev1);
    if (ms_out) *ms_out = ms;
    return MOTEX_OK;
}

int ed_run_plain(motex_batch *b, uint32_t first_x, uint32_t num_x)
{
    const SeqSet &ss = b->ss;
    double total = 0;
    for (uint32_t i = first_x; i < first_x + num_x && i < ss.num; ++i) {
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        for (uint32_t j = 0; j < ss.num; ++j) pairs.push_back({i, j});
        double ms = 0;
        int rc = ed_launch_pairs(b, pairs, &ms);
        if (rc) return rc;
        total += ms;
    }
    b->stats.kernel_ms = total;
    b->stats.main_kernel_ms = total;
    return MOTEX_OK;
}

int ed_pair_run(motex_batch *b, uint32_t i, uint32_t j)
{
    std::vector<std::pair<uint32_t, uint32_t>> pairs{{i, j}};
    double ms = 0;
    int rc = ed_launch_pairs(b, pairs, &ms);
    b->stats.kernel_ms = ms;
    b->stats.main_kernel_ms = ms;
    return rc;
}

}  // namespace motex
