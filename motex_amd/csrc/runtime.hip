// Host runtime behind include/motex.h: contexts, device-resident batches, work-queue
// planning, the plain per-pair form and the C ABI.
//
// The all-pairs loop of motex.cc:494-583 becomes a plan of work items (x row tile x text
// range) executed by one persistent kernel; the reference's per-pair u/v fold
// (functions.cc:542-555) happens on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

#include "motex.h"
#include "motex_internal.h"
#include "runtime.h"

namespace motex {

// ---------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------
static thread_local std::string g_err;
thread_local hipStream_t g_alloc_stream = nullptr;

int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return set_error(MOTEX_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

// ---------------------------------------------------------------------------------------
// parameter validation (the extraction-relevant checks of motex.cc:98-201)
// ---------------------------------------------------------------------------------------
std::vector<uint32_t> order_by_cost_desc(const std::vector<uint64_t> &cost)
{
    const size_t n = cost.size();
    uint64_t hi = 0;
    for (uint64_t c : cost) hi = std::max(hi, c);
    // keys hi - cost ascending = cost descending; stable passes keep index order on ties
    std::vector<uint64_t> k0(n), k1(n);
    std::vector<uint32_t> a(n), b(n);
    for (size_t i = 0; i < n; ++i) {
        k0[i] = hi - cost[i];
        a[i] = (uint32_t)i;
    }
    constexpr int kBits = 11;
    uint32_t cnt[1u << kBits];
    for (int shift = 0; shift < 64 && (hi >> shift); shift += kBits) {
        std::fill(cnt, cnt + (1u << kBits), 0u);
        for (size_t i = 0; i < n; ++i) ++cnt[(k0[i] >> shift) & ((1u << kBits) - 1)];
        uint32_t sum = 0;
        for (uint32_t &c : cnt) {
            const uint32_t t = c;
            c = sum;
            sum += t;
        }
        for (size_t i = 0; i < n; ++i) {
            const uint32_t d = cnt[(k0[i] >> shift) & ((1u << kBits) - 1)]++;
            k1[d] = k0[i];
            b[d] = a[i];
        }
        k0.swap(k1);
        a.swap(b);
    }
    return a;
}

static double wall_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

PhaseTrace::PhaseTrace(const char *scope_) : on(false), t(0), scope(scope_)
{
    const char *env = getenv("MOTEX_TRACE");
    on = env && env[0] == '1';
    if (on) t = wall_ms();
}

void PhaseTrace::mark(const char *phase)
{
    if (!on) return;
    const double now = wall_ms();
    fprintf(stderr, "[motex %s] %s %.3f ms\n", scope, phase, now - t);
    t = now;
}

int validate_params(const motex_params *p)
{
    if (!p) return set_error(MOTEX_EINVAL, "params is NULL");
    if (p->k == 0) return set_error(MOTEX_EINVAL, "motif length k must be positive");
    if (p->e >= p->k)
        return set_error(MOTEX_EINVAL,
                         "error threshold e must be less than the fixed-length of motifs k");
    if (p->distance > 1)
        return set_error(MOTEX_EINVAL,
                         "distance d should be 0 (Hamming) or 1 (edit distance)");
    if (p->alphabet > MOTEX_ALPHA_USR) return set_error(MOTEX_EINVAL, "unknown alphabet");
    if (p->nb_gaps && (!p->gap_min || !p->gap_max || !p->box_len || !p->box_err))
        return set_error(MOTEX_EINVAL, "structure arrays missing");
    for (uint32_t g = 0; g < p->nb_gaps; ++g) {
        if (p->gap_max[g] < p->gap_min[g])
            return set_error(MOTEX_EINVAL, "gap %u: max below min", g + 1);
        if (p->box_len[g] == 0) return set_error(MOTEX_EINVAL, "box %u has length 0", g + 1);
    }
    const uint32_t kmax = p->distance ? kMaxEdK : kMaxHdLongK;
    uint32_t longest = p->k;
    for (uint32_t g = 0; g < p->nb_gaps; ++g) longest = std::max(longest, p->box_len[g]);
    if (longest > kmax)
        return set_error(MOTEX_EUNSUPPORTED, "box length %u exceeds the %s limit of %u", longest,
                         p->distance == 0 ? "Hamming" : "edit-distance", kmax);
    return MOTEX_OK;
}

// ---------------------------------------------------------------------------------------
// sequence set: symbol codes of the concatenated text
// ---------------------------------------------------------------------------------------
int SeqSet::build(const char *const *seqs, const uint32_t *lens, uint32_t n, uint32_t nb_structs)
{
    num = n;
    len.assign(lens, lens + n);
    start.resize(n + 1);
    uint64_t tot = 0;
    for (uint32_t i = 0; i < n; ++i) {
        start[i] = (uint32_t)tot;
        tot += lens[i];
        if (lens[i] && !seqs[i]) return set_error(MOTEX_EINVAL, "sequence %u is NULL", i);
    }
    if (tot + (1u << 20) >= (1ull << 31))
        return set_error(MOTEX_EUNSUPPORTED, "total sequence length %llu too large",
                         (unsigned long long)tot);
    start[n] = (uint32_t)tot;
    total = tot;
    nx = n;
    y0 = 0;
    int map[256];
    for (int c = 0; c < 256; ++c) map[c] = -1;
    sigma = 0;
    codes.resize(tot);
    for (uint32_t i = 0; i < n; ++i) {
        const unsigned char *s = (const unsigned char *)seqs[i];
        uint8_t *dst = codes.data() + start[i];
        for (uint32_t p = 0; p < lens[i]; ++p) {
            int c = map[s[p]];
            if (c < 0) { c = map[s[p]] = (int)sigma++; }
            dst[p] = (uint8_t)c;
        }
    }
    if (sigma > 200) return set_error(MOTEX_EUNSUPPORTED, "more than 200 distinct symbols");
    if (sigma == 0) sigma = 1;
    uv_off.resize(n + 1);
    uint64_t o = 0;
    for (uint32_t i = 0; i < n; ++i) { uv_off[i] = o; o += (uint64_t)nb_structs * lens[i]; }
    uv_off[n] = o;
    return MOTEX_OK;
}

// Gap combinations in the order of functions.cc:710-729 (last gap varies fastest).
std::vector<uint32_t> enumerate_structs(const motex_params *p, uint32_t *nb_structs)
{
    uint32_t tot = 1;
    for (uint32_t g = 0; g < p->nb_gaps; ++g) tot *= p->gap_max[g] - p->gap_min[g] + 1;
    std::vector<uint32_t> S((size_t)tot * p->nb_gaps);
    for (uint32_t s = 0; s < tot; ++s) {
        uint32_t rem = s;
        for (int g = (int)p->nb_gaps - 1; g >= 0; --g) {
            uint32_t span = p->gap_max[g] - p->gap_min[g] + 1;
            S[(size_t)s * p->nb_gaps + g] = p->gap_min[g] + rem % span;
            rem /= span;
        }
    }
    *nb_structs = tot;
    return S;
}

uint32_t total_motif_length(const motex_params *p)
{
    uint32_t t = p->k;
    for (uint32_t g = 0; g < p->nb_gaps; ++g) t += p->gap_min[g] + p->box_len[g];
    return t;
}

// ---------------------------------------------------------------------------------------
// Hamming work-queue planning
// ---------------------------------------------------------------------------------------

void plan_hd_items(const SeqSet &ss, uint32_t k, uint32_t blocks, uint32_t rows,
                   const std::vector<uint32_t> &xs, const std::vector<int32_t> *only_y,
                   HdPlan &plan, uint32_t group_max, bool sym)
{
    const int64_t cap = (int64_t)kDiagPerBlock * blocks - rows - 32;
    plan.items.clear();
    plan.ranges.clear();
    plan.blocks = blocks;
    plan.rows = rows;
    plan.rows_max = rows + k - 1;
    plan.split_slots = 0;
    // pair-symmetric items take every row of x against at most kMaxSymGroup ranges
    if (sym) group_max = std::min(group_max, kMaxSymGroup);
    plan.sym_group = sym ? group_max : 0;
    // text ranges: whole sequences packed up to cap positions, long ones split in parts
    std::vector<uint32_t> split_of;  // per range: kNoSplit or slot
    auto add_seq_ranges = [&](uint32_t j) {
        uint32_t n = ss.len[j];
        if (n < k) return;
        if ((int64_t)n <= cap) {
            plan.ranges.push_back({ss.start[j], ss.start[j] + n, j, 1});
            split_of.push_back(kNoSplit);
            return;
        }
        // blocks of D diagonals: a row tile [px_start, r1) meets diagonals from
        // (start - r1 + 1) rounded down (>= start - r1 - 30) to end - 1 - px_start
        uint32_t slot = plan.split_slots++;
        const uint32_t D = kDiagPerBlock * blocks;
        const uint64_t nb = ((uint64_t)n + plan.rows_max + 30 + D - 1) / D;
        for (uint64_t i = 0; i < nb; ++i) {
            plan.ranges.push_back({ss.start[j], ss.start[j] + n, j, 1, (uint32_t)(i * D)});
            split_of.push_back(slot);
        }
    };
    if (only_y) {
        for (int32_t j : *only_y) add_seq_ranges((uint32_t)j);
    } else {
        uint32_t j = ss.y0;
        while (j < ss.num) {
            if (ss.len[j] < k) { ++j; continue; }
            if ((int64_t)ss.len[j] > cap) { add_seq_ranges(j); ++j; continue; }
            HdRange r{ss.start[j], ss.start[j] + ss.len[j], j, 1};
            uint32_t jj = j + 1;
            while (jj < ss.num && (int64_t)ss.len[jj] <= cap && jj - j < kMaxRangeSeqs &&
                   (int64_t)(ss.start[jj] + ss.len[jj]) - r.ys <= cap) {
                if (ss.len[jj] >= k) { r.ye = ss.start[jj] + ss.len[jj]; r.ny = jj - j + 1; }
                ++jj;
            }
            plan.ranges.push_back(r);
            split_of.push_back(kNoSplit);
            j = r.j0 + r.ny;
            while (j < jj && ss.len[j] < k) ++j;
        }
    }
    // groups of consecutive ranges: whole texts only, or parts of one long text
    struct Group { uint32_t rb, re, split; };
    std::vector<Group> groups;
    for (uint32_t r = 0; r < plan.ranges.size();) {
        uint32_t e = r + 1;
        while (e < plan.ranges.size() && e - r < group_max && split_of[e] == split_of[r]) ++e;
        groups.push_back({r, e, split_of[r]});
        r = e;
    }
    for (uint32_t i : xs) {
        uint32_t m = ss.len[i];
        if (m < k) continue;
        // pair-symmetric plan: only the ranges from the one holding text i onwards
        uint32_t r_first = 0;
        if (sym)
            while (r_first < plan.ranges.size() &&
                   plan.ranges[r_first].j0 + plan.ranges[r_first].ny - 1 < i)
                ++r_first;
        auto add = [&](uint32_t r0, uint32_t r1, uint32_t rb, uint32_t re, uint32_t split) {
            HdItem it{};
            it.x_start = ss.start[i];
            it.uv_base = (uint32_t)ss.uv_off[i];
            it.xf_base = (uint32_t)ss.uv_off[i];
            it.r0 = r0;
            it.r1 = r1;
            it.px_start = r0 - (k - 1);
            it.rb = rb;
            it.re = re;
            it.split = split;
            it.xi = i;
            it.sym = sym ? 1u : 0u;
            plan.items.push_back(it);
        };
        if (sym) {
            // every row of x against group_max ranges at a time, except the last sixth of its
            // ranges, two at a time: sorted by cost, the queue then ends on small items (the
            // rows are flushed once per item, so the large ones keep the flushes few; 1200 x
            // 1000 nt, at an earlier kernel: 8.36 ms and 0.44 GiB written with a tail of single
            // ranges, against 8.55 ms and 0.73 GiB with every item one range and 8.85 ms and
            // 0.39 GiB with every item three; pairs in the tail write 10% less at the same time)
            const uint32_t nr = (uint32_t)plan.ranges.size();
            const uint32_t tail = (nr - r_first + 5) / 6;
            uint32_t rb = r_first;
            for (; rb + group_max <= nr - tail; rb += group_max) add(k - 1, m, rb, rb + group_max, kNoSplit);
            // (at most group_max ranges per item: the kernel holds that many bitmaps)
            const uint32_t g_tail = std::min<uint32_t>(2, group_max);
            for (; rb < nr; rb += g_tail) add(k - 1, m, rb, std::min(nr, rb + g_tail), kNoSplit);
            continue;
        }
        for (uint32_t r0 = k - 1; r0 < m; r0 += rows) {
            const uint32_t r1 = std::min(m, r0 + rows);
            for (const Group &g : groups)
                if (g.re > r_first) add(r0, r1, std::max(g.rb, r_first), g.re, g.split);
        }
    }
}

// First NE word a row tile [r0, r1) of an item reads for a range; mirrors the kernel.
static int64_t hd_wbase(const HdItem &it, uint32_t r0, uint32_t r1, const HdRange &R)
{
    (void)it;
    int64_t q0 = ((int64_t)R.ys - (int64_t)r1 + 1);
    q0 = q0 >= 0 ? (q0 & ~31LL) : -(((-q0) + 31) & ~31LL);
    q0 += R.qoff;
    int64_t qp = q0 + (int64_t)kPadBits;
    int64_t w = qp >= 0 ? qp / 32 : -((-qp + 31) / 32);
    return w + ((r0 - (it.r0 - it.px_start)) >> 5);
}

// Words per NE table row that every window the plan's items read fits in.
int hd_table_words(const SeqSet &ss, const HdPlan &plan, uint64_t *Wout)
{
    uint64_t W = (ss.total + kPadBits) / 32 + 2;
    for (const HdItem &it : plan.items)
        for (uint32_t t0 = it.r0; t0 < it.r1; t0 += plan.rows)
            for (uint32_t r = it.rb; r < it.re; ++r) {
                int64_t wb = hd_wbase(it, t0, std::min(it.r1, t0 + plan.rows), plan.ranges[r]);
                if (wb < 0) return set_error(MOTEX_EINVAL, "internal: negative table window");
                W = std::max<uint64_t>(W, (uint64_t)wb + plan.wl + 1);
            }
    *Wout = W;
    return MOTEX_OK;
}

// Upload ranges + items; size the NE tables for every window the items read.
static int hd_upload_plan(motex_batch *b, HdPlan &plan, DevBuf &d_items, DevBuf &d_ranges, uint64_t *Wout)
{
    hipError_t e;
    if (int rc = hd_table_words(b->ss, plan, Wout)) return rc;
    if ((e = d_items.alloc(std::max<size_t>(plan.items.size(), 1) * sizeof(HdItem))) != hipSuccess)
        return hip_fail(e, "alloc items");
    if (!plan.items.empty() &&
        (e = hipMemcpyAsync(d_items.p, plan.items.data(), plan.items.size() * sizeof(HdItem), hipMemcpyHostToDevice, bstream(b))) != hipSuccess)
        return hip_fail(e, "copy items");
    if ((e = d_ranges.alloc(std::max<size_t>(plan.ranges.size(), 1) * sizeof(HdRange))) != hipSuccess)
        return hip_fail(e, "alloc ranges");
    if (!plan.ranges.empty() &&
        (e = hipMemcpyAsync(d_ranges.p, plan.ranges.data(), plan.ranges.size() * sizeof(HdRange), hipMemcpyHostToDevice, bstream(b))) != hipSuccess)
        return hip_fail(e, "copy ranges");
    return MOTEX_OK;
}

// ---------------------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------------------
}  // namespace motex

using namespace motex;

extern "C" const char *motex_last_error(void) { return g_err.c_str(); }

extern "C" int motex_ctx_create(int device, size_t budget, motex_ctx **out)
{
    if (!out) return set_error(MOTEX_EINVAL, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return set_error(MOTEX_EDEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return set_error(MOTEX_EINVAL, "device %d out of range", device);
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    std::unique_ptr<motex_ctx> c(new motex_ctx());
    c->device = device;
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    c->num_cu = prop.multiProcessorCount;
    c->arch = prop.gcnArchName;
    size_t freeb = 0, totb = 0;
    if ((e = hipMemGetInfo(&freeb, &totb)) != hipSuccess) return hip_fail(e, "hipMemGetInfo");
    c->budget = budget ? budget : (size_t)(freeb * 0.9);
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "hipStreamCreate");
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if ((e = hipEventCreate(&c->ev1)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if ((e = hipEventCreate(&c->ev2)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if (int rc = noalloc_init(device)) return rc;
    *out = c.release();
    return MOTEX_OK;
}

extern "C" void motex_ctx_destroy(motex_ctx *c)
{
    if (!c) return;
    serving_shutdown(c);
    // destroy: nothing to report to
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    delete c;
}

// ---------------------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------------------
namespace motex {

int batch_upload_common(motex_batch *b)
{
    hipError_t e;
    const SeqSet &ss = b->ss;
    if ((e = b->d_codes.alloc(std::max<uint64_t>(ss.total, 1) + 64)) != hipSuccess) return hip_fail(e, "alloc codes");
    if (ss.total && (e = hipMemcpyAsync(b->d_codes.p, ss.codes.data(), ss.total, hipMemcpyHostToDevice, bstream(b))) != hipSuccess)
        return hip_fail(e, "copy codes");
    if ((e = b->d_seq_start.alloc((ss.num + 1) * 4)) != hipSuccess) return hip_fail(e, "alloc starts");
    if ((e = hipMemcpyAsync(b->d_seq_start.p, ss.start.data(), (ss.num + 1) * 4, hipMemcpyHostToDevice, bstream(b))) != hipSuccess)
        return hip_fail(e, "copy starts");
    uint64_t uvn = std::max<uint64_t>(ss.uv_off[ss.num], 1);
    if ((e = b->d_u.alloc(uvn * 4)) != hipSuccess) return hip_fail(e, "alloc u");
    if ((e = b->d_v.alloc(uvn * 4)) != hipSuccess) return hip_fail(e, "alloc v");
    if ((e = b->d_queue.alloc(256)) != hipSuccess) return hip_fail(e, "alloc queue");
    return MOTEX_OK;
}

// Pair symmetry of the Hamming path: the occurrences of ordered pair (j, i) are those of
// (i, j) with the roles of factor and text position swapped, so each unordered pair can be
// computed once and credited to both sides.  The text side's u needs, per (x sequence,
// text position), whether any factor of x occurs there: the pair-symmetric items take every
// row of x, so a first-occurrence bitmap per range in LDS answers it.  Worth it when
// occurrences are rare (each costs global atomics on the text side) and no text is split
// across work items.  MOTEX_SYM=0 / 1 forces it off / on (where allowed).
// Expected occurrences per cell, from the symbol frequencies of the text (letters drawn
// independently).
static double hd_cell_hit_rate(const SeqSet &ss, uint32_t k, uint32_t e)
{
    std::vector<uint64_t> freq(ss.sigma, 0);
    for (uint8_t c : ss.codes) freq[c]++;
    double pm = 0;
    for (uint64_t f : freq) pm += ((double)f / ss.total) * ((double)f / ss.total);
    double p = 0, binom = 1;
    for (uint32_t d = 0; d <= e && d <= k; ++d) {
        p += binom * std::pow(1 - pm, d) * std::pow(pm, k - d);
        binom = binom * (k - d) / (d + 1);
    }
    return p;
}

static bool hd_symmetric(motex_batch *b, const HdPlan &plan)
{
    const SeqSet &ss = b->ss;
    if (plan.split_slots || b->nb_structs != 1 || ss.num < 2 || ss.nx != ss.num) return false;
    // its items hold at most kMaxSymGroup ranges, about N x ranges / (2 kMaxSymGroup) of them:
    // many short texts (hundreds of ranges per x) would make the work queue huge
    if ((double)ss.num * plan.ranges.size() / (2.0 * kMaxSymGroup) > (double)kMaxSymItems) return false;
    if (const char *env = getenv("MOTEX_SYM")) return env[0] == '1';
    return hd_cell_hit_rate(ss, b->params.k, b->params.e) < kSymMaxRate;
}

// Capacity of every range's first-occurrence list (pair-symmetric runs), plus one full
// bitmap; a list that fills up takes the rest as atomics.  x with r rows reaches a text
// position first with chance 1 - (1 - p)^r <= p r for the cell rate p, and at most once, so a
// range of span s gets s x the sum of min(1, 3 p r) over the x sequences before its last text:
// three times the expectation, but never more than one entry per position and x.
static std::vector<uint32_t> hd_sym_list_offsets(const SeqSet &ss, const HdPlan &plan, uint32_t k, uint32_t e)
{
    const double p = hd_cell_hit_rate(ss, k, e);
    // MOTEX_SYM_LIST_CAP=n: n entries per range (tests of the lists' atomic fallback)
    uint64_t forced = 0;
    if (const char *env = getenv("MOTEX_SYM_LIST_CAP")) forced = strtoull(env, nullptr, 10);
    std::vector<double> reach_upto(ss.num + 1, 0);  // sum of min(1, p r) over sequences 0 .. j-1
    for (uint32_t j = 0; j < ss.num; ++j)
        reach_upto[j + 1] = reach_upto[j] + std::min(1.0, 3.0 * p * (double)(ss.len[j] >= k ? ss.len[j] - k + 1 : 0));
    std::vector<uint32_t> off(plan.ranges.size() + 1, 0);
    uint64_t total = 0;
    for (size_t r = 0; r < plan.ranges.size(); ++r) {
        const HdRange &R = plan.ranges[r];
        const double reach = reach_upto[R.j0 + R.ny - 1];
        const double cap = forced ? (double)forced : reach * (R.ye - R.ys) + kSymBitmapBits;
        total += ((uint64_t)std::min(cap, (double)(1u << 28)) + 7) & ~7ull;  // lists start at multiples of 8
        off[r + 1] = (uint32_t)std::min<uint64_t>(total, 0xFFFFFFF8u);
    }
    return off;
}

int hd_make_plan(motex_ctx *c, const SeqSet &ss, const motex_params &p, const TileShape &ts,
                 const std::function<bool(const HdPlan &)> &want_sym, HdPlan &plan, bool *sym_out)
{
    PhaseTrace tr("hd_plan");
    std::vector<uint32_t> xs(ss.nx);
    for (uint32_t i = 0; i < ss.nx; ++i) xs[i] = i;
    plan_hd_items(ss, p.k, ts.blocks, ts.rows, xs, nullptr, plan, 1u << 30);
    tr.mark("ranges");
    const bool sym = want_sym(plan);
    tr.mark("symmetry estimate");
    // group size: aim at ~16 items per resident wave so the tail of the queue is short
    uint64_t pairs_units = 0;  // items' row tiles x ranges (pair-symmetric items hold all tiles)
    {
        size_t r_first = 0;
        for (uint32_t i = 0; i < ss.nx; ++i) {
            if (ss.len[i] < p.k) continue;
            uint64_t ntiles = sym ? 1 : (ss.len[i] - p.k + 1 + ts.rows - 1) / ts.rows;
            if (sym)
                while (r_first < plan.ranges.size() &&
                       plan.ranges[r_first].j0 + plan.ranges[r_first].ny - 1 < i)
                    ++r_first;
            pairs_units += ntiles * (plan.ranges.size() - r_first);
        }
    }
    // (a pair-symmetric item holds every row tile of x, several of the ordered plan's items:
    // ~3 per wave keep the schedule's tail short, and its rows are flushed once per item)
    uint64_t target_items = (uint64_t)c->num_cu * ts.waves_per_cu * (sym ? 3 : 16);
    uint32_t gmax = (uint32_t)std::max<uint64_t>(1, pairs_units / std::max<uint64_t>(target_items, 1));
    plan_hd_items(ss, p.k, ts.blocks, ts.rows, xs, nullptr, plan, gmax, sym);
    for (const HdItem &it : plan.items)  // the kernel holds sym_group first-occurrence bitmaps
        if (it.sym && it.re - it.rb > plan.sym_group)
            return set_error(MOTEX_EINVAL, "internal: pair-symmetric item of %u ranges beyond %u bitmaps", it.re - it.rb,
                             plan.sym_group);
    tr.mark("items");
    // longest items first: they set the tail of the persistent schedule
    std::vector<uint64_t> span(plan.ranges.size() + 1, 0);
    for (size_t r = 0; r < plan.ranges.size(); ++r) span[r + 1] = span[r] + (plan.ranges[r].ye - plan.ranges[r].ys);
    {
        std::vector<uint64_t> cost(plan.items.size());
        for (size_t t = 0; t < cost.size(); ++t) {
            const HdItem &x = plan.items[t];
            const uint64_t tiles = (x.r1 - x.r0 + plan.rows - 1) / plan.rows;
            const uint64_t rows = (x.r1 - x.r0) + tiles * (x.r0 - x.px_start);
            cost[t] = rows * (span[x.re] - span[x.rb] + (x.re - x.rb) * std::min<uint64_t>(x.r1 - x.r0, plan.rows));
        }
        const std::vector<uint32_t> order = order_by_cost_desc(cost);
        std::vector<HdItem> sorted(order.size());
        for (size_t t = 0; t < order.size(); ++t) sorted[t] = plan.items[order[t]];
        plan.items.swap(sorted);
    }
    plan.wl = hd_lds_words(ts.blocks, plan.rows_max);
    if (hd_lds_bytes(ss.sigma, plan.wl, plan.rows_max, plan.sym_group) > 65536 || (uint64_t)ss.sigma * plan.wl >= 65536)
        return set_error(MOTEX_EUNSUPPORTED, "alphabet of %u symbols does not fit the kernel's LDS window", ss.sigma);
    tr.mark("sort");
    *sym_out = sym;
    return MOTEX_OK;
}

int hd_prepare(motex_batch *b)
{
    motex_ctx *c = b->ctx;
    const SeqSet &ss = b->ss;
    PhaseTrace tr("hd_prepare");
    TileShape ts = choose_tile(c, b->params, ss);
    b->tile = ts;
    bool sym = false;
    int rc = hd_make_plan(c, ss, b->params, ts, [b](const HdPlan &pl) { return hd_symmetric(b, pl); },
                          b->hd, &sym);
    if (rc) return rc;
    tr.mark("plan");
    uint64_t W = 0;
    rc = hd_upload_plan(b, b->hd, b->d_items, b->d_ranges, &W);
    if (rc) return rc;
    tr.mark("upload plan");
    // the plain form's windows (one block, short row tiles) must fit as well
    W = std::max<uint64_t>(W, (ss.total + kPadBits) / 32 + 4 * 64 + 64);
    b->hd.W = (uint32_t)W;
    hipError_t e;
    if ((e = b->d_ne.alloc((uint64_t)W * ss.sigma * 4)) != hipSuccess) return hip_fail(e, "alloc NE tables");
    if ((e = launch_ne_build((const uint8_t *)b->d_codes.p, (uint32_t)ss.total, ss.sigma, (uint32_t)W,
                             (uint32_t *)b->d_ne.p, bstream(b))) != hipSuccess)
        return hip_fail(e, "NE build");
    // hit queues of the persistent grid (one per workgroup)
    const uint64_t grid = std::min<uint64_t>((uint64_t)c->num_cu * ts.waves_per_cu, b->hd.items.size());
    tr.mark("NE tables");
    if ((e = hd_alloc_hitq(b->d_hitq, grid, ts.blocks, hd_hitq_cap(b->params.k, b->params.e, 4))) != hipSuccess)
        return hip_fail(e, "alloc hit queues");
    tr.mark("hit queues");
    b->stats.symmetric = sym ? 1u : 0u;
    b->stats.swept_cells = hd_swept_cells(b->hd);
    if (sym) {
        const std::vector<uint32_t> off = hd_sym_list_offsets(ss, b->hd, b->params.k, b->params.e);
        if ((e = b->d_toff.alloc(off.size() * 4)) != hipSuccess) return hip_fail(e, "alloc text-side list offsets");
        if ((e = hipMemcpyAsync(b->d_toff.p, off.data(), off.size() * 4, hipMemcpyHostToDevice, bstream(b))) != hipSuccess)
            return hip_fail(e, "copy text-side list offsets");
        if ((e = b->d_tlist.alloc(std::max<uint64_t>(off.back(), 1) * 2)) != hipSuccess)
            return hip_fail(e, "alloc text-side lists");
        if ((e = b->d_tcnt.alloc(std::max<size_t>(b->hd.ranges.size(), 1) * 4)) != hipSuccess)
            return hip_fail(e, "alloc text-side counts");
    }
    b->split_words = (uint32_t)((ss.uv_off[ss.num] + 31) / 32 + 1);
    uint64_t sf = (uint64_t)std::max<uint32_t>(b->hd.split_slots, 1) * b->split_words;
    if ((e = b->d_split.alloc(sf * 4)) != hipSuccess) return hip_fail(e, "alloc split flags");
    tr.mark("bits and flags");
    if ((e = hipStreamSynchronize(bstream(b))) != hipSuccess) return hip_fail(e, "NE build sync");
    tr.mark("sync");
    return MOTEX_OK;
}

double hd_swept_cells(const HdPlan &plan)
{
    double s = 0;
    for (const HdItem &it : plan.items) {
        const uint64_t tiles = (it.r1 - it.r0 + plan.rows - 1) / plan.rows;
        const uint64_t rows = (it.r1 - it.r0) + tiles * (it.r0 - it.px_start);
        s += (double)rows * (it.re - it.rb) * kDiagPerBlock * plan.blocks;
    }
    return s;
}

hipError_t hd_alloc_hitq(DevBuf &buf, uint64_t grid, uint32_t blocks, uint32_t cap)
{
    const uint64_t need = std::max<uint64_t>(grid, 1) * cap * (blocks + 1) * 4;
    if (buf.p && buf.bytes >= need) return hipSuccess;
    return buf.alloc(need);
}

HdArgs hd_args(motex_batch *b, const HdPlan &plan, const DevBuf &items, uint32_t persistent)
{
    HdArgs a{};
    a.ne = (const uint32_t *)b->d_ne.p;
    a.codes = (const uint8_t *)b->d_codes.p;
    a.seq_start = (const uint32_t *)b->d_seq_start.p;
    a.items = (const HdItem *)items.p;
    a.ranges = (const HdRange *)b->d_ranges.p;
    a.u = (uint32_t *)b->d_u.p;
    a.v = (uint32_t *)b->d_v.p;
    a.split_flags = (uint32_t *)b->d_split.p;
    a.queue = (uint32_t *)b->d_queue.p;
    a.hitq = (uint32_t *)b->d_hitq.p;
    a.hitq_cap = hd_hitq_cap(b->params.k, b->params.e, 4);
    a.W = b->hd.W;
    a.nitems = (uint32_t)plan.items.size();
    a.sigma = b->ss.sigma;
    a.k = b->params.k;
    a.e = b->params.e;
    a.wl = plan.wl;
    a.split_words = b->split_words;
    a.rows_max = plan.rows_max;
    a.tile_rows = plan.rows;
    a.sym_group = plan.sym_group;
    a.tlist = (uint16_t *)b->d_tlist.p;
    a.toff = (const uint32_t *)b->d_toff.p;
    a.tcnt = (uint32_t *)b->d_tcnt.p;
    a.persistent = persistent;
    return a;
}

int batch_clear(motex_batch *b)
{
    hipError_t e;
    hipStream_t s = bstream(b);
    uint64_t uvn = std::max<uint64_t>(b->ss.uv_off[b->ss.num], 1);
    if ((e = launch_fill_u32((uint32_t *)b->d_u.p, uvn, 0, s)) != hipSuccess) return hip_fail(e, "clear u");
    if ((e = launch_fill_u32((uint32_t *)b->d_v.p, uvn, 0, s)) != hipSuccess) return hip_fail(e, "clear v");
    if (b->d_split.p && (e = launch_fill_u32((uint32_t *)b->d_split.p, b->d_split.bytes / 4, 0, s)) != hipSuccess)
        return hip_fail(e, "clear flags");
    if ((e = launch_fill_u32((uint32_t *)b->d_queue.p, 64, 0, s)) != hipSuccess) return hip_fail(e, "clear queue");
    if (b->d_tcnt.p && (e = launch_fill_u32((uint32_t *)b->d_tcnt.p, b->d_tcnt.bytes / 4, 0, s)) != hipSuccess)
        return hip_fail(e, "clear text-side counts");
    return MOTEX_OK;
}

int hd_launch(motex_batch *b)
{
    motex_ctx *c = b->ctx;
    hipStream_t s = bstream(b);
    hipError_t e;
    if (!b->hd.items.empty()) {
        HdArgs a = hd_args(b, b->hd, b->d_items, 1);
        uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)c->num_cu * b->tile.waves_per_cu, b->hd.items.size());
        if (b->grid_cap) grid = std::min(grid, b->grid_cap);
        if ((uint64_t)grid * a.hitq_cap * (b->tile.blocks + 1) * 4 > b->d_hitq.bytes)
            return set_error(MOTEX_EINVAL, "internal: Hamming grid of %u beyond its hit queues", grid);
        if ((e = launch_hd(a, b->tile.blocks, grid, s)) != hipSuccess) return hip_fail(e, "launch Hamming kernel");
        if (b->hd.sym_group &&
            (e = launch_sym_fold(a.tlist, a.toff, a.tcnt, a.ranges, (uint32_t)b->hd.ranges.size(), a.u, a.v, s)) != hipSuccess)
            return hip_fail(e, "fold text side");
    }
    return MOTEX_OK;
}

hipError_t batch_realloc_scratch(motex_batch *b, uint64_t bytes)
{
    if (b->graph) {  // the captured graph holds the old scratch address
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return b->d_aux.alloc(bytes);
}

static int batch_kernels(motex_batch *b)
{
    if (b->params.nb_gaps) return struct_launch_all(b);
    if (b->params.distance == 0) return hd_launch(b);
    return ed_launch_all(b);
}

int batch_launch(motex_batch *b)
{
    hipStream_t s = bstream(b);
    AllocStreamGuard guard(b->stream);  // scratch allocated at launch follows the batch's stream
    hipError_t e;
    if ((e = hipEventRecord(b->ev0, s)) != hipSuccess) return hip_fail(e, "event");
    int rc = batch_clear(b);
    if (rc) return rc;
    if ((e = hipEventRecord(b->ev2, s)) != hipSuccess) return hip_fail(e, "event");
    if ((rc = batch_kernels(b))) return rc;
    if ((e = hipEventRecord(b->ev1, s)) != hipSuccess) return hip_fail(e, "event");
    b->runs++;
    b->stats.graph_replay = 0;
    return MOTEX_OK;
}

// Runs of a batch replay a HIP graph of the run's work: the clears (fill kernels, no
// memset nodes) and the batched kernel(s).  The graph is captured on the first run; every
// buffer it names was allocated at batch creation, so capture records no allocation.
// MOTEX_GRAPH=0 keeps every run eager.
static bool graphs_enabled()
{
    const char *env = getenv("MOTEX_GRAPH");
    return !(env && env[0] == '0');
}

int batch_launch_graph(motex_batch *b)
{
    if (!graphs_enabled()) return batch_launch(b);
    hipStream_t s = bstream(b);
    hipError_t e;
    if (!b->graph) {
        AllocStreamGuard guard(b->stream);
        if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return hip_fail(e, "begin capture");
        int rc = batch_clear(b);
        if (!rc) rc = batch_kernels(b);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "end capture");
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            b->graph = nullptr;
            return hip_fail(e, "instantiate graph");
        }
    }
    if ((e = hipEventRecord(b->ev0, s)) != hipSuccess) return hip_fail(e, "event");
    if ((e = hipEventRecord(b->ev2, s)) != hipSuccess) return hip_fail(e, "event");
    if ((e = hipGraphLaunch(b->graph, s)) != hipSuccess) return hip_fail(e, "graph launch");
    if ((e = hipEventRecord(b->ev1, s)) != hipSuccess) return hip_fail(e, "event");
    b->runs++;
    b->stats.graph_replay = 1;
    return MOTEX_OK;
}

int batch_finish(motex_batch *b)
{
    hipError_t e;
    if ((e = hipEventSynchronize(b->ev1)) != hipSuccess) return hip_fail(e, "extraction kernel");
    float ms = 0, kms = 0;
    if (hipEventElapsedTime(&ms, b->ev0, b->ev1) != hipSuccess) ms = 0;
    if (hipEventElapsedTime(&kms, b->ev2, b->ev1) != hipSuccess) kms = 0;
    b->stats.kernel_ms = ms;
    b->stats.main_kernel_ms = kms;
    return b->params.nb_gaps ? struct_range_finish(b) : MOTEX_OK;
}

// Plain per-pair form: one launch per (x, y) pair.  The launch covers the pair with a few
// row tiles (one wave each) and the smallest block count that holds the text.
static const uint32_t kPlainRows = 128;

static uint32_t plain_blocks(uint32_t n)
{
    uint32_t need = n + kPlainRows + 32;
    return need <= kDiagPerBlock ? 1 : (need <= 2 * kDiagPerBlock ? 2 : 4);
}

int hd_launch_pairs(motex_batch *b, const std::vector<std::pair<uint32_t, uint32_t>> &pairs, double *ms_out)
{
    hipStream_t s = bstream(b);
    hipError_t e;
    const SeqSet &ss = b->ss;
    const uint32_t k = b->params.k;
    struct Launch { uint32_t first, count, blocks; };
    HdPlan all, pp;
    std::vector<Launch> launches;
    uint32_t slots = 0;
    for (auto &pr : pairs) {
        uint32_t i = pr.first, j = pr.second;
        if (ss.len[i] < k || ss.len[j] < k) continue;
        std::vector<int32_t> oy{(int32_t)j};
        std::vector<uint32_t> xs{i};
        uint32_t blocks = plain_blocks(ss.len[j]);
        plan_hd_items(ss, k, blocks, kPlainRows, xs, &oy, pp, 1u << 30);
        uint32_t roff = (uint32_t)all.ranges.size();
        for (HdItem &it : pp.items) {
            it.rb += roff;
            it.re += roff;
            if (it.split != kNoSplit) it.split = slots;  // one slot per launch
        }
        if (pp.split_slots) slots++;
        launches.push_back({(uint32_t)all.items.size(), (uint32_t)pp.items.size(), blocks});
        all.items.insert(all.items.end(), pp.items.begin(), pp.items.end());
        all.ranges.insert(all.ranges.end(), pp.ranges.begin(), pp.ranges.end());
    }
    if (ms_out) *ms_out = 0;
    if (all.items.empty()) return MOTEX_OK;
    all.rows_max = kPlainRows + k - 1;
    all.wl = hd_lds_words(4, all.rows_max);
    DevBuf items, ranges, flags, hitq;
    uint64_t W = 0;
    int rc = hd_upload_plan(b, all, items, ranges, &W);
    if (rc) return rc;
    if (W > b->hd.W) return set_error(MOTEX_EINVAL, "internal: plain window beyond the NE tables");
    if ((e = flags.alloc((uint64_t)std::max<uint32_t>(slots, 1) * b->split_words * 4)) != hipSuccess)
        return hip_fail(e, "alloc plain flags");
    if ((e = launch_fill_u32((uint32_t *)flags.p, flags.bytes / 4, 0, s)) != hipSuccess) return hip_fail(e, "clear flags");
    uint32_t max_count = 0;
    for (const Launch &L : launches) max_count = std::max(max_count, L.count);
    if ((e = hd_alloc_hitq(hitq, max_count, 4, hd_hitq_cap(k, b->params.e, 4))) != hipSuccess)
        return hip_fail(e, "alloc plain hit queues");
    HdPlan shape;
    shape.rows = kPlainRows;
    shape.rows_max = all.rows_max;
    if ((e = hipEventRecord(b->ev0, s)) != hipSuccess) return hip_fail(e, "record plain start");
    for (const Launch &L : launches) {
        shape.wl = hd_lds_words(L.blocks, shape.rows_max);
        HdArgs a = hd_args(b, shape, items, 0);
        a.items = (const HdItem *)items.p + L.first;
        a.ranges = (const HdRange *)ranges.p;
        a.nitems = L.count;
        a.split_flags = (uint32_t *)flags.p;
        a.hitq = (uint32_t *)hitq.p;
        if ((e = launch_hd(a, L.blocks, L.count, s)) != hipSuccess) return hip_fail(e, "launch plain kernel");
    }
    if ((e = hipEventRecord(b->ev1, s)) != hipSuccess) return hip_fail(e, "record plain end");
    if ((e = hipEventSynchronize(b->ev1)) != hipSuccess) return hip_fail(e, "plain kernel");
    float ms = 0;
    if ((e = hipEventElapsedTime(&ms, b->ev0, b->ev1)) != hipSuccess) return hip_fail(e, "plain timing");
    if (ms_out) *ms_out = ms;
    return MOTEX_OK;
}

int hd_run_plain(motex_batch *b, uint32_t first_x, uint32_t num_x)
{
    const SeqSet &ss = b->ss;
    double total = 0;
    for (uint32_t i = first_x; i < first_x + num_x && i < ss.nx; ++i) {
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        for (uint32_t j = ss.y0; j < ss.num; ++j) pairs.push_back({i, j});
        double ms = 0;
        int rc = hd_launch_pairs(b, pairs, &ms);
        if (rc) return rc;
        total += ms;
    }
    b->stats.kernel_ms = total;
    b->stats.main_kernel_ms = total;
    return MOTEX_OK;
}

int pair_run(motex_batch *b, uint32_t i, uint32_t j)
{
    int rc = batch_clear(b);
    if (rc) return rc;
    if (b->params.nb_gaps) return struct_pair_run(b, i, j);
    if (b->params.distance == 1) return ed_pair_run(b, i, j);
    std::vector<std::pair<uint32_t, uint32_t>> pairs{{i, j}};
    double ms = 0;
    rc = hd_launch_pairs(b, pairs, &ms);
    b->stats.kernel_ms = ms;
    b->stats.main_kernel_ms = ms;
    return rc;
}

double pair_cells(const SeqSet &ss, const motex_params &p, uint64_t *pairs, uint64_t *skipped)
{
    uint32_t tot = total_motif_length(&p);
    double sm = 0, sn = 0;
    uint64_t okx = 0, oky = 0;
    for (uint32_t i = 0; i < ss.nx; ++i)
        if (ss.len[i] >= tot) { sm += ss.len[i] - p.k + 1; ++okx; }
    for (uint32_t j = ss.y0; j < ss.num; ++j)
        if (ss.len[j] >= tot) { sn += ss.len[j]; ++oky; }
    *pairs = okx * oky;
    *skipped = (uint64_t)ss.nx * (ss.num - ss.y0) - okx * oky;
    return sm * sn;
}

}  // namespace motex

extern "C" int motex_batch_create(motex_ctx *c, const char *const *seqs, const uint32_t *lens,
                                  uint32_t num, const motex_params *p, motex_batch **out)
{
    return batch_create(c, seqs, lens, num, p, nullptr, out);
}

int motex::batch_create(motex_ctx *c, const char *const *seqs, const uint32_t *lens, uint32_t num,
                        const motex_params *p, hipStream_t stream, motex_batch **out, uint32_t grid_share,
                        uint32_t num_x)
{
    if (!c || !out) return set_error(MOTEX_EINVAL, "ctx/out is NULL");
    *out = nullptr;
    int rc = validate_params(p);
    if (rc) return rc;
    if (num && (!seqs || !lens)) return set_error(MOTEX_EINVAL, "seqs/lens is NULL");
    hipError_t e;
    if ((e = hipSetDevice(c->device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    std::unique_ptr<motex_batch> b(new motex_batch());
    b->ctx = c;
    b->stream = stream;
    b->grid_share = std::max<uint32_t>(grid_share, 1);
    AllocStreamGuard guard(stream);
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess ||
        (e = hipEventCreate(&b->ev2)) != hipSuccess)
        return hip_fail(e, "hipEventCreate");
    b->params = *p;
    b->gmin.assign(p->gap_min, p->gap_min + p->nb_gaps);
    b->gmax.assign(p->gap_max, p->gap_max + p->nb_gaps);
    b->blen.assign(p->box_len, p->box_len + p->nb_gaps);
    b->berr.assign(p->box_err, p->box_err + p->nb_gaps);
    b->params.gap_min = b->gmin.data();
    b->params.gap_max = b->gmax.data();
    b->params.box_len = b->blen.data();
    b->params.box_err = b->berr.data();
    PhaseTrace tr("batch_create");
    b->S = enumerate_structs(&b->params, &b->nb_structs);
    if ((rc = b->ss.build(seqs, lens, num, b->nb_structs))) return rc;
    if (num_x < num) {
        if (p->nb_gaps) return set_error(MOTEX_EUNSUPPORTED, "cross runs take single motifs only");
        b->ss.nx = b->ss.y0 = num_x;
    }
    b->stats.cells = pair_cells(b->ss, b->params, &b->stats.pairs, &b->stats.skipped_pairs);
    tr.mark("pack");
    if ((rc = batch_upload_common(b.get()))) return rc;
    tr.mark("upload");
    if (p->nb_gaps) rc = struct_prepare(b.get());
    else if (p->distance == 0) rc = hd_prepare(b.get());
    else rc = ed_prepare(b.get());
    if (rc) return rc;
    tr.mark("prepare");
    if ((e = hipStreamSynchronize(bstream(b.get()))) != hipSuccess) return hip_fail(e, "batch upload");
    tr.mark("sync");
    if (b->stats.swept_cells == 0) b->stats.swept_cells = b->stats.cells;  // edit, structured
    b->stats.tile_blocks = b->tile.blocks;
    b->stats.tile_rows = b->tile.rows;
    b->stats.waves_per_cu = b->tile.waves_per_cu;
    b->stats.kernel_variant = b->tile.variant;
    *out = b.release();
    return MOTEX_OK;
}

extern "C" int motex_batch_run(motex_batch *b)
{
    if (!b) return set_error(MOTEX_EINVAL, "batch is NULL");
    if (hipError_t e = hipSetDevice(b->ctx->device); e != hipSuccess) return hip_fail(e, "hipSetDevice");
    int rc = batch_launch_graph(b);
    return rc ? rc : batch_finish(b);
}

extern "C" int motex_batch_clear(motex_batch *b)
{
    if (!b) return set_error(MOTEX_EINVAL, "batch is NULL");
    if (hipError_t e = hipSetDevice(b->ctx->device); e != hipSuccess) return hip_fail(e, "hipSetDevice");
    int rc = batch_clear(b);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(bstream(b));
    return e == hipSuccess ? MOTEX_OK : hip_fail(e, "clear");
}

extern "C" int motex_batch_run_plain(motex_batch *b, uint32_t first_x, uint32_t num_x)
{
    if (!b) return set_error(MOTEX_EINVAL, "batch is NULL");
    if (hipError_t e = hipSetDevice(b->ctx->device); e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if (b->params.nb_gaps) return struct_run_plain(b, first_x, num_x);
    if (b->params.distance == 0) return hd_run_plain(b, first_x, num_x);
    return ed_run_plain(b, first_x, num_x);
}

extern "C" int motex_batch_download(motex_batch *b, uint32_t *const *u, uint32_t *const *v)
{
    if (!b || !u || !v) return set_error(MOTEX_EINVAL, "NULL argument");
    if (hipError_t e = hipSetDevice(b->ctx->device); e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const SeqSet &ss = b->ss;
    uint64_t n = ss.uv_off[ss.num];
    std::vector<uint32_t> hu(n), hv(n);
    hipError_t e;
    if (n) {
        if ((e = hipMemcpy(hu.data(), b->d_u.p, n * 4, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "copy u");
        if ((e = hipMemcpy(hv.data(), b->d_v.p, n * 4, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "copy v");
    }
    for (uint32_t i = 0; i < ss.num; ++i) {
        uint64_t o = ss.uv_off[i], cnt = ss.uv_off[i + 1] - o;
        if (!cnt) continue;
        if (!u[i] || !v[i]) return set_error(MOTEX_EINVAL, "u/v row %u is NULL", i);
        for (uint64_t t = 0; t < cnt; ++t) { u[i][t] += hu[o + t]; v[i][t] += hv[o + t]; }
    }
    return MOTEX_OK;
}

extern "C" int motex_batch_stats(const motex_batch *b, motex_stats *out)
{
    if (!b || !out) return set_error(MOTEX_EINVAL, "NULL argument");
    *out = b->stats;
    out->work_items = b->work_items();
    return MOTEX_OK;
}

extern "C" void motex_batch_destroy(motex_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);  // destroy: nothing to report to
    delete b;
}

extern "C" int motex_extract_all_pairs(motex_ctx *c, const char *const *seqs, const uint32_t *lens,
                                       uint32_t num, const motex_params *p, uint32_t *const *u,
                                       uint32_t *const *v)
{
    motex_batch *b = nullptr;
    int rc = motex_batch_create(c, seqs, lens, num, p, &b);
    if (rc) return rc;
    rc = motex_batch_run(b);
    if (!rc) rc = motex_batch_download(b, u, v);
    motex_batch_destroy(b);
    return rc;
}

extern "C" int motex_extract_cross(motex_ctx *c, const char *const *xs, const uint32_t *xlens, uint32_t num_x,
                                   const char *const *ys, const uint32_t *ylens, uint32_t num_y,
                                   const motex_params *p, uint32_t *const *u, uint32_t *const *v)
{
    if (!c) return set_error(MOTEX_EINVAL, "ctx is NULL");
    int rc = validate_params(p);
    if (rc) return rc;
    if (p->nb_gaps) return set_error(MOTEX_EINVAL, "cross runs take single motifs only (no structure)");
    if ((num_x && (!xs || !xlens || !u || !v)) || (num_y && (!ys || !ylens)))
        return set_error(MOTEX_EINVAL, "NULL argument");
    if (!num_x) return MOTEX_OK;
    std::vector<const char *> all(xs, xs + num_x);
    std::vector<uint32_t> lens(xlens, xlens + num_x);
    all.insert(all.end(), ys, ys + num_y);
    lens.insert(lens.end(), ylens, ylens + num_y);
    motex_batch *b = nullptr;
    if ((rc = batch_create(c, all.data(), lens.data(), num_x + num_y, p, nullptr, &b, 1, num_x))) return rc;
    std::unique_ptr<motex_batch, void (*)(motex_batch *)> guard(b, motex_batch_destroy);
    if ((rc = motex_batch_run(b))) return rc;
    // the texts' rows stay zero in a cross run; they are downloaded into scratch rows
    std::vector<std::vector<uint32_t>> yu(num_y), yv(num_y);
    std::vector<uint32_t *> up(u, u + num_x), vp(v, v + num_x);
    for (uint32_t j = 0; j < num_y; ++j) {
        yu[j].assign(std::max<uint32_t>(ylens[j], 1), 0);
        yv[j].assign(std::max<uint32_t>(ylens[j], 1), 0);
        up.push_back(yu[j].data());
        vp.push_back(yv[j].data());
    }
    return motex_batch_download(b, up.data(), vp.data());
}

extern "C" int motex_extract_pair(motex_ctx *c, const char *x, uint32_t m, const char *y, uint32_t n,
                                  const motex_params *p, uint32_t *u, uint32_t *v)
{
    if (!c) return set_error(MOTEX_EINVAL, "ctx is NULL");
    int rc = validate_params(p);
    if (rc) return rc;
    const char *seqs[2] = {x, y};
    uint32_t lens[2] = {m, n};
    motex_batch *b = nullptr;
    if ((rc = motex_batch_create(c, seqs, lens, 2, p, &b))) return rc;
    rc = pair_run(b, 0, 1);
    if (!rc) {
        uint32_t dummy_u = 0, dummy_v = 0;
        std::vector<uint32_t> yu((size_t)b->nb_structs * n + 1), yv((size_t)b->nb_structs * n + 1);
        uint32_t *us[2] = {u, yu.data()}, *vs[2] = {v, yv.data()};
        (void)dummy_u; (void)dummy_v;
        rc = motex_batch_download(b, us, vs);
    }
    motex_batch_destroy(b);
    return rc;
}
