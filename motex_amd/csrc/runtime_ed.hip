// Edit-distance batches: work-queue planning and launches for ed_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "ed_device.h"
#include "runtime.h"

namespace motex {

hipError_t launch_ed_slice(const EdSliceArgs &a, uint32_t grid, bool pipe, hipStream_t s);
uint32_t ed_slice_stream_row_words(uint32_t k);

static const uint64_t kEdTextPerItem = 8192;  // text symbols per work item, at most

// Word width of the one-text-per-lane kernel: 32, 64 or 128 bits (ed_device.h).
static int ed_word_width(const motex_params &p)
{
    uint32_t longest = p.k;
    for (uint32_t g = 0; g < p.nb_gaps; ++g) longest = std::max(longest, p.box_len[g]);
    return ed_width(longest);
}

// Items: (x, run of consecutive texts of about text_per_item symbols).
static void plan_ed(const SeqSet &ss, uint32_t k, const std::vector<uint32_t> &xs,
                    const std::vector<uint32_t> *only_y, uint64_t text_per_item,
                    std::vector<EdItemDev> &out)
{
    out.clear();
    struct R { uint32_t j0, ny; };
    std::vector<R> ranges;
    if (only_y) {
        for (uint32_t j : *only_y) ranges.push_back({j, 1});
    } else {
        uint32_t j = ss.y0;
        while (j < ss.num) {
            uint32_t j0 = j;
            uint64_t acc = 0;
            while (j < ss.num && (acc == 0 || acc + ss.len[j] <= text_per_item)) { acc += ss.len[j]; ++j; }
            ranges.push_back({j0, j - j0});
        }
    }
    for (uint32_t i : xs) {
        if (ss.len[i] < k) continue;
        for (const R &r : ranges)
            out.push_back({ss.start[i], ss.len[i], (uint32_t)ss.uv_off[i], r.j0, r.ny, 0});
    }
}

static uint64_t ed_item_cost(const SeqSet &ss, const EdItemDev &it, int width)
{
    uint32_t sc = ed_strip_cols(width);
    uint64_t strips = (it.m + sc - 1) / sc, steps = 0;
    for (uint32_t j = it.j0; j < it.j0 + it.ny; ++j) steps += ss.len[j] + 63;
    return strips * steps;
}

// Boundary-stream words per wave (two streams of this many words each).
static uint32_t ed_stream_words(const motex_batch *b)
{
    uint32_t maxn = 0;
    for (uint32_t j = 0; j < b->ss.num; ++j) maxn = std::max(maxn, b->ss.len[j]);
    return (maxn + 128) * ed_stream_row_words(ed_word_width(b->params));
}

static uint32_t ed_grid(const motex_batch *b)
{
    return (uint32_t)std::min<uint64_t>((uint64_t)b->ctx->num_cu * b->tile.waves_per_cu, b->ed_nitems);
}

// Workgroups of the text-parallel kernel (one wave each, or kSlicePipe when pipelined) and
// the boundary-stream words they hold (two buffers per wave, or kSlicePipe + 1 per item).
static uint32_t ed_slice_grid(const motex_batch *b)
{
    uint64_t waves = (uint64_t)b->ctx->num_cu * b->tile.waves_per_cu;
    if (b->grid_cap) waves = std::min<uint64_t>(waves, b->grid_cap);
    const uint64_t groups = b->ed_pipe ? waves / kSlicePipe : waves;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(groups, b->ed_nitems));
}

static uint64_t ed_slice_scratch_bytes(const motex_batch *b, uint32_t grid)
{
    return (uint64_t)grid * (b->ed_pipe ? kSlicePipe + 1 : 2) * b->ed_slice_stream_words * 4;
}

// Text-parallel plan (ed_slice_kernels.h): the texts that can hold an occurrence
// (length >= k), longest first, in groups of 32; per group row, the 5 bit planes of the 32
// texts' symbol codes and the mask of texts still running.  Items are (x, group), most
// expensive first.  Used for k in 4..kMaxSliceK and alphabets of at most 32 symbols (5 planes);
// MOTEX_ED_SLICE=0 keeps the one-text-per-lane kernel.
static bool ed_slice_ok(const motex_batch *b)
{
    const char *env = getenv("MOTEX_ED_SLICE");
    if (env && env[0] == '0') return false;
    const uint32_t k = b->params.k;
    if (k < 4 || k > kMaxSliceK || b->ss.sigma > 32 || ed_slice_stream_row_words(k) == 0) return false;
    if (env && env[0] == '1') return true;
    // An item (x, group of 32 texts) is a sequential sweep of 64-column strips, so its
    // latency is long; with very few items the one-text-per-lane kernel, whose items shrink
    // to (x, one text), finishes first (40-100 protein sequences: 4.5-9 ms against 10-16 ms
    // pipelined and 31-51 ms with one wave per item).
    uint64_t nx = 0, ny = 0;
    for (uint32_t i = 0; i < b->ss.nx; ++i) nx += b->ss.len[i] >= k;
    for (uint32_t j = b->ss.y0; j < b->ss.num; ++j) ny += b->ss.len[j] >= k;
    const uint64_t items = nx * ((ny + 31) / 32);
    return items >= (uint64_t)b->ctx->num_cu * kSliceItemsPerCu;
}

static int ed_prepare_slice(motex_batch *b)
{
    const SeqSet &ss = b->ss;
    const uint32_t k = b->params.k;
    std::vector<uint32_t> texts;
    for (uint32_t j = ss.y0; j < ss.num; ++j)
        if (ss.len[j] >= k) texts.push_back(j);
    std::stable_sort(texts.begin(), texts.end(), [&](uint32_t x, uint32_t z) { return ss.len[x] > ss.len[z]; });
    const uint32_t ngroups = (uint32_t)((texts.size() + 31) / 32);
    std::vector<uint32_t> goff(ngroups + 1, 0);
    for (uint32_t g = 0; g < ngroups; ++g) goff[g + 1] = goff[g] + ss.len[texts[32 * g]];
    std::vector<uint32_t> planes((size_t)goff[ngroups] * 6, 0);
    for (uint32_t g = 0; g < ngroups; ++g) {
        for (uint32_t t = 0; t < 32 && 32 * g + t < texts.size(); ++t) {
            const uint32_t j = texts[32 * g + t];
            const uint8_t *code = ss.codes.data() + ss.start[j];
            for (uint32_t r = 0; r < ss.len[j]; ++r) {
                uint32_t *row = planes.data() + (size_t)(goff[g] + r) * 6;
                for (int bit = 0; bit < 5; ++bit) row[bit] |= (uint32_t)((code[r] >> bit) & 1) << t;
                row[5] |= 1u << t;
            }
        }
    }
    std::vector<EdSliceItem> items;
    for (uint32_t i = 0; i < ss.nx; ++i) {
        if (ss.len[i] < k) continue;
        for (uint32_t g = 0; g < ngroups; ++g)
            items.push_back({ss.start[i], ss.len[i], (uint32_t)ss.uv_off[i], g});
    }
    // an item's cost: its sequential steps, strips x (rows + the 63-step skew)
    uint64_t total_cost = 0, max_cost = 0;
    {
        std::vector<uint64_t> cost(items.size());
        for (size_t t = 0; t < items.size(); ++t) {
            cost[t] = (uint64_t)((items[t].m + 63) / 64) * (goff[items[t].group + 1] - goff[items[t].group] + 63);
            total_cost += cost[t];
            max_cost = std::max(max_cost, cost[t]);
        }
        const std::vector<uint32_t> order = order_by_cost_desc(cost);
        std::vector<EdSliceItem> sorted(order.size());
        for (size_t t = 0; t < order.size(); ++t) sorted[t] = items[order[t]];
        items.swap(sorted);
    }
    b->tile = choose_tile(b->ctx, b->params, ss);
    b->ed_slice = true;
    b->ed_nitems = (uint32_t)items.size();
    // One wave per item takes about max(0.41 us x the longest item's steps, 0.61 us x a
    // wave's share of all steps); the pipelined form (four waves per item, barriers between
    // chunks) about max(0.41 us x longest / 3.5, 0.82 us x the share) (k=10 protein sweeps of
    // 200-500 sequences, MI355X).  So the pipelined form wins when the share is below about
    // half the longest item: there one item's sequential sweep is the critical path.
    const char *pipe_env = getenv("MOTEX_ED_PIPE");
    if (pipe_env && (pipe_env[0] == '0' || pipe_env[0] == '1')) {
        b->ed_pipe = pipe_env[0] == '1';
    } else if (b->grid_share > 1) {
        // a serving job shares the device, where other jobs fill what its critical path
        // leaves idle: it takes the higher-throughput one-wave form unless its items are too
        // few to fill its share
        b->ed_pipe = (uint64_t)items.size() < (uint64_t)b->ctx->num_cu * kSlicePipeItemsPerCu;
    } else {
        const uint64_t waves = std::max<uint64_t>((uint64_t)b->ctx->num_cu * std::max<uint32_t>(b->tile.waves_per_cu, 1), 1);
        b->ed_pipe = (double)total_cost < kSlicePipeShare * (double)max_cost * (double)waves;
    }
    b->tile.variant = b->ed_pipe ? 3 : 2;
    uint32_t rows_max = 0;
    for (uint32_t g = 0; g < ngroups; ++g) rows_max = std::max(rows_max, goff[g + 1] - goff[g]);
    b->ed_slice_stream_words = std::max<uint32_t>(rows_max, 1) * ed_slice_stream_row_words(k);
    hipError_t e;
    hipStream_t s = bstream(b);
    if ((e = b->d_items.alloc(std::max<size_t>(items.size(), 1) * sizeof(EdSliceItem))) != hipSuccess)
        return hip_fail(e, "alloc items");
    if (!items.empty() &&
        (e = hipMemcpyAsync(b->d_items.p, items.data(), items.size() * sizeof(EdSliceItem), hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(e, "copy items");
    if ((e = b->d_gplanes.alloc(std::max<size_t>(planes.size(), 1) * 4)) != hipSuccess) return hip_fail(e, "alloc text planes");
    if (!planes.empty() &&
        (e = hipMemcpyAsync(b->d_gplanes.p, planes.data(), planes.size() * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(e, "copy text planes");
    if ((e = b->d_goff.alloc(goff.size() * 4)) != hipSuccess) return hip_fail(e, "alloc group offsets");
    if ((e = hipMemcpyAsync(b->d_goff.p, goff.data(), goff.size() * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(e, "copy group offsets");
    // boundary streams of the persistent grid, allocated here so that runs allocate nothing
    const uint64_t need = ed_slice_scratch_bytes(b, ed_slice_grid(b));
    if ((e = batch_realloc_scratch(b, need)) != hipSuccess) return hip_fail(e, "alloc edit streams");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "upload items");
    return MOTEX_OK;
}

// Text symbols per item: an item is one wave's sequential wavefront, so a small job cut into
// few long items is bound by the longest one.  Runs are shortened (down to one text) until
// the items fill the persistent grid about twice; large inputs keep kEdTextPerItem.
static uint64_t ed_text_per_item(const motex_batch *b, uint64_t nx)
{
    const uint64_t target = (uint64_t)b->ctx->num_cu * std::max<uint32_t>(b->tile.waves_per_cu, 1) * 2;
    const uint64_t t = (uint64_t)b->ss.total * nx / target;
    return std::min<uint64_t>(std::max<uint64_t>(t, 1), kEdTextPerItem);
}

int ed_prepare(motex_batch *b)
{
    if (ed_slice_ok(b)) return ed_prepare_slice(b);
    const SeqSet &ss = b->ss;
    std::vector<uint32_t> xs(ss.nx);
    uint64_t nx = 0;
    for (uint32_t i = 0; i < ss.nx; ++i) {
        xs[i] = i;
        nx += ss.len[i] >= b->params.k;
    }
    b->tile = choose_tile(b->ctx, b->params, ss);
    std::vector<EdItemDev> items;
    plan_ed(ss, b->params.k, xs, nullptr, ed_text_per_item(b, nx), items);
    const int width = ed_word_width(b->params);
    std::vector<uint64_t> cost(items.size());
    for (size_t t = 0; t < items.size(); ++t) cost[t] = ed_item_cost(ss, items[t], width);
    const std::vector<uint32_t> order = order_by_cost_desc(cost);
    std::vector<EdItemDev> sorted(items.size());
    for (size_t t = 0; t < items.size(); ++t) sorted[t] = items[order[t]];
    b->tile.variant = width == 0 ? 0 : (width == 1 ? 1 : 4);
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
    const int width = ed_word_width(b->params);
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
    if ((e = launch_ed(a, width, grid, s)) != hipSuccess) return hip_fail(e, "launch edit kernel");
    return MOTEX_OK;
}

int ed_launch_all(motex_batch *b)
{
    if (!b->ed_nitems) return MOTEX_OK;
    if (!b->ed_slice) {
        uint32_t grid = ed_grid(b);
        if (b->grid_cap) grid = std::min(grid, b->grid_cap);
        return ed_launch(b, b->d_items.p, b->ed_nitems, grid, 1, bstream(b));
    }
    const uint32_t grid = ed_slice_grid(b);
    if (ed_slice_scratch_bytes(b, grid) > b->d_aux.bytes)
        return set_error(MOTEX_EINVAL, "internal: edit grid of %u beyond its boundary streams", grid);
    EdSliceArgs a{};
    a.codes = (const uint8_t *)b->d_codes.p;
    a.items = (const EdSliceItem *)b->d_items.p;
    a.gplanes = (const uint32_t *)b->d_gplanes.p;
    a.goff = (const uint32_t *)b->d_goff.p;
    a.u = (uint32_t *)b->d_u.p;
    a.v = (uint32_t *)b->d_v.p;
    a.queue = (uint32_t *)b->d_queue.p;
    a.scratch = (uint32_t *)b->d_aux.p;
    a.stream_words = b->ed_slice_stream_words;
    a.nitems = b->ed_nitems;
    a.k = b->params.k;
    a.e = b->params.e;
    a.persistent = 1;
    hipError_t e = launch_ed_slice(a, grid, b->ed_pipe, bstream(b));
    return e == hipSuccess ? MOTEX_OK : hip_fail(e, "launch text-parallel edit kernel");
}

// Plain form: one launch per pair (one work item, one wave).
static int ed_launch_pairs(motex_batch *b, const std::vector<std::pair<uint32_t, uint32_t>> &pairs, double *ms_out)
{
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
    if ((e = hipEventRecord(b->ev0, s)) != hipSuccess) return hip_fail(e, "record plain start");
    for (size_t t = 0; t < all.size(); ++t) {
        int rc = ed_launch(b, (const EdItemDev *)items.p + t, 1, 1, 0, s);
        if (rc) return rc;
    }
    if ((e = hipEventRecord(b->ev1, s)) != hipSuccess) return hip_fail(e, "record plain end");
    if ((e = hipEventSynchronize(b->ev1)) != hipSuccess) return hip_fail(e, "plain edit kernel");
    float ms = 0;
    if ((e = hipEventElapsedTime(&ms, b->ev0, b->ev1)) != hipSuccess) return hip_fail(e, "plain timing");
    if (ms_out) *ms_out = ms;
    return MOTEX_OK;
}

int ed_run_plain(motex_batch *b, uint32_t first_x, uint32_t num_x)
{
    const SeqSet &ss = b->ss;
    double total = 0;
    for (uint32_t i = first_x; i < first_x + num_x && i < ss.nx; ++i) {
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        for (uint32_t j = ss.y0; j < ss.num; ++j) pairs.push_back({i, j});
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
