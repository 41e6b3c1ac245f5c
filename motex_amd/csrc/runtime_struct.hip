// Structured-motif batches: pair work queue, scratch sizing and launches of
// struct_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "hd_device.h"
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

// ---------------------------------------------------------------------------------------
// Range-batched plan (struct_range_kernels.hip)
// ---------------------------------------------------------------------------------------

hipError_t launch_struct_range_sweep(const SrSweepArgs &a, uint32_t grid, size_t lds, hipStream_t s);
hipError_t launch_struct_range_merge(const SrMergeArgs &a, uint32_t units, size_t lds, hipStream_t s);
size_t struct_range_sweep_lds(uint32_t sigma, uint32_t wl, uint32_t m_max);
size_t struct_range_merge_lds(uint32_t hcap, uint32_t ndist, uint32_t nbk);
hipError_t launch_struct_range_fold(const SrFoldArgs &a, uint32_t nx, hipStream_t s);

static const uint32_t kSrWindow = kDiagPerBlock * kSrBlocks;  // diagonals of a range window
static const uint32_t kSrHcapMin = 1024;
static const uint32_t kSrHcapMax = 65472;             // hit indexes are 16-bit
static const uint32_t kSrOvfCap = 1u << 22;           // pairs the overflow list holds
static const uint64_t kSrSlotBytes = 2ull << 30;      // unit slots of one launch pair

// Waves per CU the sweep kernel's registers allow, by counter planes.
static uint32_t sr_waves_vgpr(uint32_t planes)
{
    return planes <= 3 ? 32 : planes == 4 ? 28 : planes == 5 ? 24 : 20;
}

// NE window words per symbol of the sweep: a row reads words (px >> 5) + 64 b + lane and the
// next one, for blocks b < 4.
static uint32_t sr_window_words(uint32_t maxlen)
{
    return kSrBlocks * 64 + 2 + (maxlen + 31) / 32;
}

// Chance that two windows of len symbols are within err mismatches, letters drawn
// independently with match probability pm.
static double sr_hit_probability(double pm, uint32_t len, uint32_t err)
{
    if (err >= len) return 1.0;
    double p = 0, c = 1;  // c = C(len, i)
    for (uint32_t i = 0; i <= err; ++i) {
        p += c * std::pow(1 - pm, (double)i) * std::pow(pm, (double)(len - i));
        c = c * (len - i) / (i + 1);
    }
    return std::min(1.0, p);
}

// Takes the range-batched plan when its kernels can run the batch: Hamming boxes, at most
// kSrMaxGaps gaps, counters of at most 7 planes, every text and the longest x inside one
// window, the NE window and a unit's hits in LDS -- and, unless MOTEX_STRUCT_RANGE=1 asks for
// it regardless, few enough expected hits per unit that units rarely overflow their slots.
// MOTEX_STRUCT_RANGE=0 turns it off.  Plans units, slots and the overflow list; *taken says
// whether it did.
static int struct_range_prepare(motex_batch *b, uint32_t maxlen, bool *taken)
{
    *taken = false;
    const char *env = getenv("MOTEX_STRUCT_RANGE");
    const bool off = env && env[0] == '0', force = env && env[0] == '1';
    const SeqSet &ss = b->ss;
    const uint32_t G = b->params.nb_gaps;
    if (off || b->params.distance != 0 || ss.nx != ss.num || G > kSrMaxGaps) return MOTEX_OK;
    StructArgs t{};
    int rc = struct_shape(b, t);
    if (rc) return rc;
    uint32_t planes = 3;
    for (uint32_t d = 0; d < t.ndist; ++d) planes = std::max(planes, hd_counter_planes(t.dlen[d], t.derr[d]));
    if (planes > 7 || 2ull * maxlen + 32 > kSrWindow) return MOTEX_OK;
    const uint32_t m_max = std::max<uint32_t>(maxlen, 1);
    // the merge's u keys: (kx * m + px) * 32 + text below 2^31 (the text side's keys set bit 31)
    if ((uint64_t)t.nb_structs * m_max >= (1ull << 26)) return MOTEX_OK;
    const uint32_t wl = sr_window_words(maxlen);
    const size_t sweep_lds = struct_range_sweep_lds(ss.sigma, wl, m_max);
    if (sweep_lds > 65536) return MOTEX_OK;
    const uint32_t nbk = (kSrWindow + m_max + 15) / 16;
    // a unit's hits take 2 bytes each in the merge's LDS buckets
    const size_t fixed = struct_range_merge_lds(0, t.ndist, nbk);
    if (fixed + 2 * kSrHcapMin > 65536) return MOTEX_OK;
    const uint32_t hcap_lds = (uint32_t)std::min<size_t>(kSrHcapMax, (65536 - fixed) / 2 / 64 * 64);

    // expected hits of a unit: every shape's rows x window x hit chance
    std::vector<uint64_t> hist(256, 0);
    for (uint8_t c : ss.codes) hist[c]++;
    double pm = 0;
    for (uint64_t h : hist) pm += ss.total ? ((double)h / ss.total) * ((double)h / ss.total) : 0;
    double expect = 0;
    for (uint32_t d = 0; d < t.ndist; ++d) {
        if (maxlen < t.dlen[d]) continue;
        expect += (double)(maxlen - t.dlen[d] + 1) * kSrWindow * sr_hit_probability(pm, t.dlen[d], t.derr[d]);
    }
    const double want = std::ceil(std::max<double>(kSrHcapMin, 2.5 * expect + 512) / 64) * 64;
    if (want > hcap_lds && !force) return MOTEX_OK;
    uint32_t hcap = (uint32_t)std::min<double>(hcap_lds, want);
    // test knobs: smaller slots and overflow list (units that overflow, a full list)
    if (const char *q = getenv("MOTEX_SR_HCAP")) hcap = std::max(64u, std::min(hcap, (uint32_t)atoi(q) / 64 * 64));

    // ranges: consecutive texts that hold the whole structure, packed up to cap positions
    const uint32_t tot = total_motif_length(&b->params);
    const uint32_t cap = kSrWindow - m_max - 32;
    std::vector<HdRange> ranges;
    uint64_t texts = 0, text_syms = 0;
    for (uint32_t j = 0; j < ss.num;) {
        if (ss.len[j] < tot) { ++j; continue; }
        HdRange r{ss.start[j], ss.start[j] + ss.len[j], j, 1};
        uint32_t jj = j + 1;
        while (jj < ss.num && ss.len[jj] >= tot && jj - j < kMaxRangeSeqs &&
               (uint64_t)ss.start[jj] + ss.len[jj] - r.ys <= cap) {
            r.ye = ss.start[jj] + ss.len[jj];
            r.ny++;
            ++jj;
        }
        ranges.push_back(r);
        texts += r.ny;
        text_syms += r.ye - r.ys;
        j = jj;
    }
    // units: every x against every range (pair-symmetric: the ranges from the one holding x
    // on, each unordered pair once), longest x first
    // With several gaps the merge, not the sweep, bounds the run, and the text side's credits
    // share its key set: 500 x 600 nt, 8/1 boxes, two gaps 4.8 -> 4.5 ms, three gaps 7.6 ->
    // 9.5 ms.  So those take the plan only when MOTEX_SYM=1 asks for it.  (Text-side keys hold
    // the combination in 17 bits.)
    const char *sym_env = getenv("MOTEX_SYM");
    const bool sym = struct_symmetric(b) && (G == 1 || (sym_env && sym_env[0] == '1')) && t.nb_structs < (1u << 17);
    std::vector<uint32_t> range_of(ss.num, 0);
    for (uint32_t r = 0; r < ranges.size(); ++r)
        for (uint32_t j = ranges[r].j0; j < ranges[r].j0 + ranges[r].ny; ++j) range_of[j] = r;
    std::vector<uint32_t> xs;
    for (uint32_t i = 0; i < ss.num; ++i)
        if (ss.len[i] >= tot) xs.push_back(i);
    std::stable_sort(xs.begin(), xs.end(), [&](uint32_t p, uint32_t q) { return ss.len[p] > ss.len[q]; });
    std::vector<SrItem> items;
    items.reserve(xs.size() * ranges.size());
    for (uint32_t i : xs)
        for (uint32_t r = sym ? range_of[i] : 0; r < ranges.size(); ++r)
            items.push_back({ss.start[i], ss.len[i], (uint32_t)ss.uv_off[i], r, i, 0});
    std::vector<SrX> xlist;
    for (uint32_t i : xs) xlist.push_back({i, ss.len[i], (uint32_t)ss.uv_off[i], 0});
    const uint64_t slot_bytes = (uint64_t)(kSrHead + hcap) * 4;
    uint64_t slot_budget = kSrSlotBytes;
    if (const char *sb = getenv("MOTEX_SR_SLOT_BYTES")) slot_budget = std::max<uint64_t>(1, strtoull(sb, nullptr, 10));
    const uint32_t chunk =
        (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(items.size(), slot_budget / slot_bytes));
    const uint32_t nchunks = items.empty() ? 0 : (uint32_t)((items.size() + chunk - 1) / chunk);

    uint32_t wpc = (uint32_t)std::max<size_t>(1, std::min<size_t>(sr_waves_vgpr(planes), (160 * 1024) / sweep_lds));
    if (const char *w = getenv("MOTEX_SR_WAVES")) wpc = std::max(1, std::min(32, atoi(w)));

    hipError_t e;
    hipStream_t s = bstream(b);
    if ((e = b->d_ranges.alloc(std::max<size_t>(ranges.size(), 1) * sizeof(HdRange))) != hipSuccess)
        return hip_fail(e, "alloc ranges");
    if (!ranges.empty() &&
        (e = hipMemcpyAsync(b->d_ranges.p, ranges.data(), ranges.size() * sizeof(HdRange), hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(e, "copy ranges");
    if ((e = b->d_sr_items.alloc(std::max<size_t>(items.size(), 1) * sizeof(SrItem))) != hipSuccess)
        return hip_fail(e, "alloc units");
    if (!items.empty() &&
        (e = hipMemcpyAsync(b->d_sr_items.p, items.data(), items.size() * sizeof(SrItem), hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(e, "copy units");
    if ((e = b->d_sr_slots.alloc(std::max<uint64_t>(1, (uint64_t)chunk * slot_bytes))) != hipSuccess)
        return hip_fail(e, "alloc unit slots");
    if ((e = b->d_sr_queue.alloc(std::max<size_t>(nchunks, 1) * 4)) != hipSuccess) return hip_fail(e, "alloc queues");
    if ((e = b->d_sr_uv.alloc(std::max<uint64_t>(ss.uv_off[ss.num], 1) * 8)) != hipSuccess)
        return hip_fail(e, "alloc accumulators");
    if ((e = b->d_sr_self.alloc(std::max<size_t>(ss.num, 1) * 4)) != hipSuccess) return hip_fail(e, "alloc self flags");
    if ((e = b->d_sr_xs.alloc(std::max<size_t>(xlist.size(), 1) * sizeof(SrX))) != hipSuccess)
        return hip_fail(e, "alloc x list");
    if (!xlist.empty() &&
        (e = hipMemcpyAsync(b->d_sr_xs.p, xlist.data(), xlist.size() * sizeof(SrX), hipMemcpyHostToDevice, s)) != hipSuccess)
        return hip_fail(e, "copy x list");
    b->sr_nx = (uint32_t)xlist.size();
    {
        std::vector<uint32_t> uvb(ss.num);
        for (uint32_t i = 0; i < ss.num; ++i) uvb[i] = (uint32_t)ss.uv_off[i];
        if ((e = b->d_sr_uvb.alloc(std::max<size_t>(ss.num, 1) * 4)) != hipSuccess) return hip_fail(e, "alloc u/v bases");
        if (ss.num && (e = hipMemcpyAsync(b->d_sr_uvb.p, uvb.data(), ss.num * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
            return hip_fail(e, "copy u/v bases");
        // the host vectors of these copies end with this function
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "range plan upload");
    }
    b->sr_sym = sym;
    const uint64_t pairs = (uint64_t)xs.size() * texts;
    b->sr_ovf_cap = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(pairs, kSrOvfCap));
    if (const char *o = getenv("MOTEX_SR_OVF_CAP")) b->sr_ovf_cap = std::max(1u, std::min(b->sr_ovf_cap, (uint32_t)atoi(o)));
    if ((e = b->d_ovf.alloc((uint64_t)b->sr_ovf_cap * sizeof(StructItem))) != hipSuccess)
        return hip_fail(e, "alloc overflow list");
    b->struct_range = true;
    b->sr_nitems = (uint32_t)items.size();
    b->sr_chunk = chunk;
    b->sr_nchunks = nchunks;
    b->sr_grid = (uint32_t)std::min<uint64_t>((uint64_t)b->ctx->num_cu * wpc, std::max<uint32_t>(chunk, 1));
    b->sr_hcap = hcap;
    b->sr_wl = wl;
    b->sr_nbk = nbk;
    b->sr_m_max = m_max;
    // the per-pair kernel over handed-over pairs: a small persistent grid
    b->sr_ovf_grid = std::min<uint32_t>(b->ctx->num_cu * 4, b->sr_ovf_cap);
    b->tile.variant = 1;
    b->tile.waves_per_cu = wpc;
    b->stats.symmetric = sym ? 1u : 0u;
    // cells: one DP per distinct box shape and ordered pair; swept: every unit's window
    double cells = 0, swept = 0;
    for (uint32_t d = 0; d < t.ndist; ++d) {
        double rows = 0;
        for (uint32_t i : xs) rows += ss.len[i] >= t.dlen[d] ? ss.len[i] - t.dlen[d] + 1 : 0;
        cells += rows * (double)text_syms;
    }
    for (const SrItem &it : items) swept += (double)it.m * kSrWindow * t.ndist;
    b->stats.cells = cells;
    b->stats.swept_cells = swept;
    *taken = true;
    return MOTEX_OK;
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
        // (and the range kernel's four-block window)
        uint64_t W = (ss.total + pad) / 32 + std::max<uint32_t>(kDiagPerBlock / 32 + wl, sr_window_words(maxlen)) + 8;
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
    if ((e = b->d_aux2.alloc(std::max<size_t>(b->S.size(), 1) * 4)) != hipSuccess) return hip_fail(e, "alloc S");
    if (!b->S.empty() &&
        (e = hipMemcpyAsync(b->d_aux2.p, b->S.data(), b->S.size() * 4, hipMemcpyHostToDevice, bstream(b))) != hipSuccess)
        return hip_fail(e, "copy S");
    b->tile = choose_tile(b->ctx, b->params, ss);
    bool range = false;
    if ((rc = struct_range_prepare(b, maxlen, &range))) return rc;
    if (range) {
        // the per-pair kernel's scratch, for the pairs of ranges that overflow
        if ((rc = struct_alloc_scratch(b))) return rc;
        return hipStreamSynchronize(bstream(b)) == hipSuccess ? MOTEX_OK : set_error(MOTEX_EDEVICE, "NE build");
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
    const uint32_t grid = b->struct_range ? b->sr_ovf_grid : struct_grid(b, std::max<uint32_t>(b->ed_nitems, 1));
    const uint64_t need = (uint64_t)grid * a.scratch_words * 4;
    hipError_t e = batch_realloc_scratch(b, need);
    return e == hipSuccess ? MOTEX_OK : hip_fail(e, "alloc structured scratch");
}

// Range-batched run: per launch pair, the sweep over a chunk of units into their slots and the
// merge of those slots; then the per-pair kernel over the pairs of units whose hits overflowed
// their slots (a count the last launch reads on the device).
static int struct_range_launch(motex_batch *b)
{
    hipStream_t s = bstream(b);
    hipError_t e;
    StructArgs t{};
    int rc = struct_fill_args(b, t);
    if (rc) return rc;
    uint32_t *q = (uint32_t *)b->d_queue.p;   // [1] handed pairs' counter, [2..3] their list
    uint32_t *sq = (uint32_t *)b->d_sr_queue.p;
    if (b->sr_nchunks && (e = launch_fill_u32(sq, b->sr_nchunks, 0, s)) != hipSuccess) return hip_fail(e, "clear queues");
    if ((e = launch_fill_u32((uint32_t *)b->d_sr_uv.p, b->d_sr_uv.bytes / 4, 0, s)) != hipSuccess ||
        (e = launch_fill_u32((uint32_t *)b->d_sr_self.p, b->d_sr_self.bytes / 4, 0, s)) != hipSuccess)
        return hip_fail(e, "clear accumulators");
    SrSweepArgs sa{};
    sa.ne = t.ne;
    sa.codes = t.codes;
    sa.seq_start = (const uint32_t *)b->d_seq_start.p;
    sa.items = (const SrItem *)b->d_sr_items.p;
    sa.ranges = (const HdRange *)b->d_ranges.p;
    sa.slots = (uint32_t *)b->d_sr_slots.p;
    sa.ovf = (StructItem *)b->d_ovf.p;
    sa.ovf_count = q + 2;
    sa.ovf_cap = b->sr_ovf_cap;
    sa.W = t.W;
    sa.sigma = t.sigma;
    sa.pad_bits = t.pad_bits;
    sa.wl = b->sr_wl;
    sa.m_max = b->sr_m_max;
    sa.hcap = b->sr_hcap;
    sa.ndist = t.ndist;
    SrMergeArgs ma{};
    ma.slots = sa.slots;
    ma.items = sa.items;
    ma.ranges = sa.ranges;
    ma.seq_start = sa.seq_start;
    ma.S = t.S;
    ma.uv = (unsigned long long *)b->d_sr_uv.p;
    sa.self_ovf = (uint32_t *)b->d_sr_self.p;
    sa.uvb = ma.uvb = (const uint32_t *)b->d_sr_uvb.p;
    sa.sym = ma.sym = b->sr_sym ? 1u : 0u;
    ma.m_max = b->sr_m_max;
    ma.hcap = b->sr_hcap;
    ma.ndist = t.ndist;
    ma.nbk = b->sr_nbk;
    ma.nb_gaps = t.nb_gaps;
    ma.nb_structs = t.nb_structs;
    for (uint32_t x = 0; x < kSrMaxBoxes; ++x) {
        sa.dlen[x] = t.dlen[x];
        sa.derr[x] = t.derr[x];
        ma.blen[x] = t.blen[x];
        ma.dist_of[x] = t.dist_of[x];
    }
    for (uint32_t g = 0; g < kSrMaxGaps; ++g) {
        ma.gmin[g] = t.gmin[g];
        ma.gmax[g] = t.gmax[g];
    }
    for (uint32_t g = t.nb_gaps, st = 1; g-- > 0;) {
        ma.gstride[g] = st;
        st *= t.gmax[g] - t.gmin[g] + 1;
    }
    const size_t sweep_lds = struct_range_sweep_lds(t.sigma, sa.wl, sa.m_max);
    const size_t merge_lds = struct_range_merge_lds(ma.hcap, ma.ndist, ma.nbk);
    for (uint32_t c = 0; c < b->sr_nchunks; ++c) {
        const uint32_t first = c * b->sr_chunk;
        const uint32_t n = std::min(b->sr_chunk, b->sr_nitems - first);
        sa.first = first;
        sa.nitems = n;
        sa.queue = sq + c;
        uint32_t grid = std::min(b->sr_grid, n);
        if (b->grid_cap) grid = std::min(grid, b->grid_cap);
        if ((e = launch_struct_range_sweep(sa, grid, sweep_lds, s)) != hipSuccess)
            return hip_fail(e, "launch structured range sweep");
        ma.first = first;
        if ((e = launch_struct_range_merge(ma, n, merge_lds, s)) != hipSuccess)
            return hip_fail(e, "launch structured range merge");
    }
    // the accumulators into u and v, with every x against itself's diagonal chains
    SrFoldArgs fa{};
    fa.xs = (const SrX *)b->d_sr_xs.p;
    fa.self_ovf = sa.self_ovf;
    fa.S = t.S;
    fa.uv = ma.uv;
    fa.u = t.u;
    fa.v = t.v;
    fa.nb_gaps = t.nb_gaps;
    fa.nb_structs = t.nb_structs;
    for (uint32_t x = 0; x < kSrMaxBoxes; ++x) fa.blen[x] = t.blen[x];
    if (b->sr_nx && (e = launch_struct_range_fold(fa, b->sr_nx, s)) != hipSuccess)
        return hip_fail(e, "launch structured range fold");
    // the handed-over pairs
    t.items = (const StructItem *)b->d_ovf.p;
    t.nitems = b->sr_ovf_cap;
    t.nitems_dev = q + 2;
    t.queue = q + 1;
    t.persistent = 1;
    uint32_t og = b->sr_ovf_grid;
    if (b->grid_cap) og = std::min(og, b->grid_cap);
    if ((uint64_t)og * t.scratch_words * 4 > b->d_aux.bytes)
        return set_error(MOTEX_EINVAL, "internal: overflow grid of %u beyond its scratch", og);
    t.scratch = (uint32_t *)b->d_aux.p;
    const size_t lds = struct_lds_bytes(t);
    const bool dense = t.nb_gaps == 1 && lds * 24 <= 160 * 1024;
    if ((e = launch_struct(t, og, lds, dense, s)) != hipSuccess) return hip_fail(e, "launch structured kernel");
    return MOTEX_OK;
}

// After a range-batched run: pairs handed to the per-pair kernel, and whether its list held them.
int struct_range_finish(motex_batch *b)
{
    if (!b->struct_range) return MOTEX_OK;
    uint32_t w[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(w, (uint32_t *)b->d_queue.p + 2, sizeof w, hipMemcpyDeviceToHost, bstream(b));
    if (e == hipSuccess) e = hipStreamSynchronize(bstream(b));
    if (e != hipSuccess) return hip_fail(e, "read overflow count");
    b->stats.handed_pairs = w[0];
    if (w[1])
        return set_error(MOTEX_EDEVICE,
                         "structured motifs: %u pairs overflowed the range kernel's queues, beyond the %u the "
                         "per-pair list holds (MOTEX_STRUCT_RANGE=0 runs every pair on the per-pair kernel)",
                         w[0], b->sr_ovf_cap);
    return MOTEX_OK;
}

int struct_launch_all(motex_batch *b)
{
    if (b->struct_range) return struct_range_launch(b);
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
