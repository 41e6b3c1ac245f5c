// All-pairs Hamming extraction that allocates no device memory (motex_extract_all_pairs_noalloc).
//
// Every device buffer of the run -- packed symbol codes, sequence starts, the work queue
// (ranges, items, head counter), the NE tables, the split-text flags, the per-wave hit queues
// and the u/v accumulators -- is carved out of one file-scope device array, `scratch`, whose
// per-device address is looked up once with hipGetSymbolAddress when a context is created on
// that device.  The host code passes that address (never the array's name, which in host code
// is only a stand-in for the symbol) to the copies and kernels.
//
// The run follows motex_extract_all_pairs' Hamming path (ordered plan: the pair-symmetric
// plan's occurrence bitmap would not fit a modest array) with shorter hit queues, and
// returns MOTEX_ENOMEM when the input needs more than the array holds.  The array persists
// per device until the process ends and keeps the last call's contents; everything a run
// reads before writing is cleared by fill kernels first.  Calls on one device share the array:
// this file serialises them within the process (a per-device mutex held until the run's
// stream has drained); calls from several processes on one device would race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "motex.h"
#include "runtime.h"

namespace motex {

static constexpr uint64_t kScratchBytes = MOTEX_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// Hit-queue entries per wave: the kernel's minimum, which keeps the queues of a full
// persistent grid (CUs x 32 waves x 128 entries x 5 words) at about 20 MB.
static constexpr uint32_t kNoallocHitqCap = 128;

struct DeviceScratch {
    void *ptr = nullptr;  // device address of `scratch` on this device
    std::mutex run_mu;    // held for a whole run: runs on one device never overlap
};
static std::mutex g_scratch_mu;
static std::map<int, DeviceScratch *> g_scratch;

// Set-up, called by motex_ctx_create with `device` current: looks up the array's address on
// this device once.
int noalloc_init(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    DeviceScratch *&ds = g_scratch[device];
    if (!ds) ds = new DeviceScratch();
    if (ds->ptr) return MOTEX_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
    ds->ptr = p;
    return MOTEX_OK;
}

static DeviceScratch *scratch_of(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(device);
    return it == g_scratch.end() ? nullptr : it->second;
}

// Byte layout of one run inside the array (offsets 256-byte aligned).
struct NoallocLayout {
    uint64_t queue, codes, starts, ranges, items, ne, split, hitq, u, v, end;
};

static uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

static NoallocLayout noalloc_layout(const SeqSet &ss, const HdPlan &plan, uint64_t W, uint32_t split_words,
                                    uint64_t grid)
{
    NoallocLayout L{};
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) { const uint64_t at = o; o = align256(o + std::max<uint64_t>(bytes, 4)); return at; };
    const uint64_t uvn = std::max<uint64_t>(ss.uv_off[ss.num], 1);
    L.queue = take(256);
    L.codes = take(ss.total + 64);
    L.starts = take((uint64_t)(ss.num + 1) * 4);
    L.ranges = take(plan.ranges.size() * sizeof(HdRange));
    L.items = take(plan.items.size() * sizeof(HdItem));
    L.ne = take(W * ss.sigma * 4);
    L.split = take((uint64_t)std::max<uint32_t>(plan.split_slots, 1) * split_words * 4);
    L.hitq = take(std::max<uint64_t>(grid, 1) * kNoallocHitqCap * (plan.blocks + 1) * 4);
    L.u = take(uvn * 4);
    L.v = take(uvn * 4);
    L.end = o;
    return L;
}

static int noalloc_run(motex_ctx *c, uint8_t *base, const char *const *seqs, const uint32_t *lens,
                       uint32_t num, const motex_params *p, uint32_t *const *u, uint32_t *const *v)
{
    SeqSet ss;
    int rc = ss.build(seqs, lens, num, 1);
    if (rc) return rc;
    const TileShape ts = choose_tile(c, *p, ss);
    HdPlan plan;
    bool sym = false;
    if ((rc = hd_make_plan(c, ss, *p, ts, [](const HdPlan &) { return false; }, plan, &sym))) return rc;
    if (plan.items.empty()) return MOTEX_OK;  // every pair skipped: u, v unchanged
    uint64_t W = 0;
    if ((rc = hd_table_words(ss, plan, &W))) return rc;
    const uint32_t split_words = (uint32_t)((ss.uv_off[ss.num] + 31) / 32 + 1);
    const uint64_t grid = std::min<uint64_t>((uint64_t)c->num_cu * ts.waves_per_cu, plan.items.size());
    const NoallocLayout L = noalloc_layout(ss, plan, W, split_words, grid);
    if (L.end > kScratchBytes)
        return set_error(MOTEX_ENOMEM,
                         "input needs %llu bytes of device scratch, more than the %llu bytes of the "
                         "no-allocation entry point",
                         (unsigned long long)L.end, (unsigned long long)kScratchBytes);
    hipStream_t s = c->stream;
    hipError_t e;
    auto up = [&](uint64_t off, const void *src, uint64_t bytes) {
        return bytes ? hipMemcpyAsync(base + off, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
    };
    if ((e = up(L.codes, ss.codes.data(), ss.total)) != hipSuccess ||
        (e = up(L.starts, ss.start.data(), (uint64_t)(ss.num + 1) * 4)) != hipSuccess ||
        (e = up(L.ranges, plan.ranges.data(), plan.ranges.size() * sizeof(HdRange))) != hipSuccess ||
        (e = up(L.items, plan.items.data(), plan.items.size() * sizeof(HdItem))) != hipSuccess)
        return hip_fail(e, "copy into scratch");
    const uint64_t uvn = std::max<uint64_t>(ss.uv_off[ss.num], 1);
    // the array keeps the previous run's contents: clear what this run reads before writing
    if ((e = launch_fill_u32((uint32_t *)(base + L.queue), 64, 0, s)) != hipSuccess ||
        (e = launch_fill_u32((uint32_t *)(base + L.split),
                             (uint64_t)std::max<uint32_t>(plan.split_slots, 1) * split_words, 0, s)) != hipSuccess ||
        (e = launch_fill_u32((uint32_t *)(base + L.u), uvn, 0, s)) != hipSuccess ||
        (e = launch_fill_u32((uint32_t *)(base + L.v), uvn, 0, s)) != hipSuccess)
        return hip_fail(e, "clear scratch");
    if ((e = launch_ne_build(base + L.codes, (uint32_t)ss.total, ss.sigma, (uint32_t)W, (uint32_t *)(base + L.ne),
                             s)) != hipSuccess)
        return hip_fail(e, "NE build");
    HdArgs a{};
    a.ne = (const uint32_t *)(base + L.ne);
    a.codes = base + L.codes;
    a.seq_start = (const uint32_t *)(base + L.starts);
    a.items = (const HdItem *)(base + L.items);
    a.ranges = (const HdRange *)(base + L.ranges);
    a.u = (uint32_t *)(base + L.u);
    a.v = (uint32_t *)(base + L.v);
    a.split_flags = (uint32_t *)(base + L.split);
    a.queue = (uint32_t *)(base + L.queue);
    a.hitq = (uint32_t *)(base + L.hitq);
    a.hitq_cap = kNoallocHitqCap;
    a.W = (uint32_t)W;
    a.nitems = (uint32_t)plan.items.size();
    a.sigma = ss.sigma;
    a.k = p->k;
    a.e = p->e;
    a.wl = plan.wl;
    a.split_words = split_words;
    a.rows_max = plan.rows_max;
    a.tile_rows = plan.rows;
    a.sym_group = 0;  // ordered plan
    a.persistent = 1;
    if ((e = launch_hd(a, plan.blocks, (uint32_t)grid, s)) != hipSuccess) return hip_fail(e, "launch Hamming kernel");
    std::vector<uint32_t> hu(uvn), hv(uvn);
    if ((e = hipMemcpyAsync(hu.data(), base + L.u, uvn * 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipMemcpyAsync(hv.data(), base + L.v, uvn * 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return hip_fail(e, "copy u/v");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "extraction kernel");
    for (uint32_t i = 0; i < ss.num; ++i) {
        const uint64_t o = ss.uv_off[i], cnt = ss.uv_off[i + 1] - o;
        if (!cnt) continue;
        if (!u[i] || !v[i]) return set_error(MOTEX_EINVAL, "u/v row %u is NULL", i);
        for (uint64_t t = 0; t < cnt; ++t) {
            u[i][t] += hu[o + t];
            v[i][t] += hv[o + t];
        }
    }
    return MOTEX_OK;
}

}  // namespace motex

using namespace motex;

extern "C" int motex_extract_all_pairs_noalloc(motex_ctx *c, const char *const *seqs, const uint32_t *lens,
                                               uint32_t num, const motex_params *p, uint32_t *const *u,
                                               uint32_t *const *v)
{
    if (!c) return set_error(MOTEX_EINVAL, "ctx is NULL");
    int rc = validate_params(p);
    if (rc) return rc;
    if (p->distance != 0 || p->nb_gaps)
        return set_error(MOTEX_EUNSUPPORTED, "the no-allocation entry point runs Hamming single motifs only");
    if (num && (!seqs || !lens || !u || !v)) return set_error(MOTEX_EINVAL, "NULL argument");
    hipError_t e;
    if ((e = hipSetDevice(c->device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    DeviceScratch *ds = scratch_of(c->device);
    if (!ds || !ds->ptr) return set_error(MOTEX_EDEVICE, "no scratch address for device %d", c->device);
    std::lock_guard<std::mutex> lk(ds->run_mu);
    return noalloc_run(c, (uint8_t *)ds->ptr, seqs, lens, num, p, u, v);
}
