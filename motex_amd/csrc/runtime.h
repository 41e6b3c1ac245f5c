// Host-side runtime structures (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "motex.h"
#include "motex_internal.h"

namespace motex {

int set_error(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
int validate_params(const motex_params *p);

// Indices 0..n-1 ordered by cost, largest first, equal costs in index order (an LSD radix
// sort on 16-bit digits: linear in n, for work queues of 10^5-10^6 items).
std::vector<uint32_t> order_by_cost_desc(const std::vector<uint64_t> &cost);

// Host-side phase timer for batch set-up: with MOTEX_TRACE=1 in the environment, mark()
// prints the milliseconds since the previous mark to stderr; otherwise it does nothing.
struct PhaseTrace {
    bool on;
    double t;
    const char *scope;
    explicit PhaseTrace(const char *scope_);
    void mark(const char *phase);
};

// Stream whose stream-ordered allocator DevBuf uses on this thread (serving jobs), or null
// for plain hipMalloc/hipFree.
extern thread_local hipStream_t g_alloc_stream;

struct AllocStreamGuard {
    hipStream_t prev;
    explicit AllocStreamGuard(hipStream_t s) : prev(g_alloc_stream) { g_alloc_stream = s; }
    ~AllocStreamGuard() { g_alloc_stream = prev; }
};

// Device allocation owned by a batch or job.
struct DevBuf {
    void *p = nullptr;
    uint64_t bytes = 0;
    hipStream_t s = nullptr;  // allocated stream-ordered on this stream
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    hipError_t alloc(uint64_t n)
    {
        release();
        s = g_alloc_stream;
        hipError_t e = s ? hipMallocAsync(&p, n ? n : 4, s) : hipMalloc(&p, n ? n : 4);
        if (e != hipSuccess) { p = nullptr; return e; }
        bytes = n;
        return hipSuccess;
    }
    void release()
    {
        if (p) {
            if (s) (void)hipFreeAsync(p, s);
            else (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
    }
};

// Concatenated symbol codes of a sequence set.
struct SeqSet {
    uint32_t num = 0;
    uint64_t total = 0;
    uint32_t sigma = 0;
    std::vector<uint32_t> len, start;   // start has num+1 entries
    std::vector<uint8_t> codes;
    std::vector<uint64_t> uv_off;       // num+1 entries, nb_structs * len prefix sums
    // Sequences [0, nx) are the x side (their factors are searched) and [y0, num) the texts.
    // All-pairs runs: nx = num, y0 = 0.  Cross runs (x = motifs, y = input sequences):
    // nx = y0 = the number of x sequences.
    uint32_t nx = 0, y0 = 0;
    int build(const char *const *seqs, const uint32_t *lens, uint32_t n, uint32_t nb_structs);
};

// Tile shape of the batched kernels (picked by the autotuner per alphabet/d/k/e).
struct TileShape {
    uint32_t blocks = 4;         // Hamming: 2048-diagonal blocks per wave
    uint32_t rows = 256;         // Hamming: x rows per work item
    uint32_t waves_per_cu = 24;  // persistent grid = num_cu * waves_per_cu
    uint32_t variant = 0;        // edit kernels: 0 = 32-bit keys, 1 = 64-bit words
    uint32_t ytile = 0;          // edit kernels: text symbols per work item
};

struct HdPlan {
    std::vector<HdItem> items;
    std::vector<HdRange> ranges;
    uint32_t blocks = 4, rows = 512, rows_max = 0, wl = 0, W = 0, split_slots = 0;
    uint32_t sym_group = 0;  // pair-symmetric plans: ranges per item at most
};

struct Job;

}  // namespace motex

struct motex_ctx {
    int device = 0;
    int num_cu = 0;
    std::string arch;
    size_t budget = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    std::mutex tune_mu;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t>, motex::TileShape> tune_cache;
    // serving
    std::mutex mu;
    std::condition_variable cv;
    std::thread worker;
    bool stop = false;
    bool worker_started = false;
    uint64_t next_id = 1;
    size_t resident_bytes = 0;
    std::map<uint64_t, std::shared_ptr<motex::Job>> jobs;
    std::deque<std::shared_ptr<motex::Job>> pending;
};

struct motex_batch {
    motex_ctx *ctx = nullptr;
    motex_params params{};
    std::vector<uint32_t> gmin, gmax, blen, berr, S;
    uint32_t nb_structs = 1;
    motex::SeqSet ss;
    motex::TileShape tile;
    motex::HdPlan hd;
    uint32_t split_words = 0;
    uint32_t ed_nitems = 0;
    uint32_t grid_cap = 0;           // serving: bound on the persistent grid (0 = none)
    uint32_t grid_share = 1;         // serving: the grid the plan can count on is 1/grid_share
    motex_stats stats{};
    hipStream_t stream = nullptr;   // overrides ctx->stream when set (serving)
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;  // run start, end, main kernel start
    hipGraphExec_t graph = nullptr;  // replay of clear + batched kernel(s), captured on run 2
    uint32_t runs = 0;
    ~motex_batch()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (ev2) (void)hipEventDestroy(ev2);
    }
    motex::DevBuf d_codes, d_seq_start, d_u, d_v, d_queue, d_ne, d_items, d_split, d_aux, d_aux2, d_ranges, d_hitq;
    // pair-symmetric Hamming runs: per range, the text side's first-occurrence list
    // (d_tlist from d_toff[r]) and its fill count (d_tcnt)
    motex::DevBuf d_tlist, d_toff, d_tcnt;
    bool ed_slice = false;           // edit runs on the text-parallel kernel
    bool ed_pipe = false;            // ... in its pipelined form (kSlicePipe waves per item)
    motex::DevBuf d_gplanes, d_goff; // its text groups
    // range-batched structured plan (Hamming boxes): its units (ranges in d_ranges), the unit
    // slots of one launch pair, per-launch unit counters, the list of pairs it hands to the
    // per-pair kernel (whose count and full flag are words 2 and 3 of d_queue), the merge's
    // 64-bit u/v accumulators, the x list of the fold and per-sequence self-pair flags
    bool struct_range = false;
    motex::DevBuf d_sr_items, d_sr_slots, d_sr_queue, d_ovf, d_sr_uv, d_sr_xs, d_sr_self, d_sr_uvb;
    uint32_t sr_nx = 0;
    bool sr_sym = false;
    uint32_t sr_nitems = 0, sr_chunk = 0, sr_nchunks = 0, sr_grid = 0, sr_hcap = 0, sr_ovf_cap = 0, sr_ovf_grid = 0;
    uint32_t sr_wl = 0, sr_nbk = 0, sr_m_max = 0;
    uint32_t ed_slice_stream_words = 0;
    uint64_t work_items() const { return hd.items.size() + ed_nitems + sr_nitems; }
};

namespace motex {

inline hipStream_t bstream(const motex_batch *b) { return b->stream ? b->stream : b->ctx->stream; }
// num_x < num: a cross run, sequences [0, num_x) against the texts [num_x, num) only.
int batch_create(motex_ctx *c, const char *const *seqs, const uint32_t *lens, uint32_t num,
                 const motex_params *p, hipStream_t stream, motex_batch **out, uint32_t grid_share = 1,
                 uint32_t num_x = ~0u);
int batch_launch(motex_batch *b);   // enqueue clear + batched kernel, no wait
int batch_launch_graph(motex_batch *b);  // same, replayed from a HIP graph after the first run
hipError_t batch_realloc_scratch(motex_batch *b, uint64_t bytes);  // grows d_aux, drops the graph
int batch_finish(motex_batch *b);   // wait for the last launch, record its timings

std::vector<uint32_t> enumerate_structs(const motex_params *p, uint32_t *nb_structs);
uint32_t total_motif_length(const motex_params *p);
void plan_hd_items(const SeqSet &ss, uint32_t k, uint32_t blocks, uint32_t rows,
                   const std::vector<uint32_t> &xs, const std::vector<int32_t> *only_y,
                   HdPlan &plan, uint32_t group_max, bool sym = false);
TileShape choose_tile(motex_ctx *c, const motex_params &p, const SeqSet &ss);

// kernels (hd_kernels.hip, ed_kernels.hip)
hipError_t launch_ne_build(const uint8_t *codes, uint32_t total, uint32_t sigma, uint32_t W,
                           uint32_t *ne, hipStream_t s, uint32_t pad = kPadBits);
hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t val, hipStream_t s);
hipError_t launch_ed(const EdArgs &a, int width, uint32_t grid, hipStream_t s);
uint32_t ed_strip_cols(int width);          // columns per strip of the one-text-per-lane kernel
uint32_t ed_stream_row_words(int width);    // boundary-stream words per row
hipError_t launch_hd(const HdArgs &a, uint32_t blocks, uint32_t grid, hipStream_t s);
size_t hd_lds_bytes(uint32_t sigma, uint32_t wl, uint32_t rows_max, uint32_t sym_group);
hipError_t launch_sym_fold(const uint16_t *tlist, const uint32_t *toff, const uint32_t *tcnt, const HdRange *ranges,
                           uint32_t nranges, uint32_t *u, uint32_t *v, hipStream_t s);
uint32_t hd_lds_words(uint32_t blocks, uint32_t rows_max);

// per-mode entry points (runtime.hip, ed_runtime.hip, struct_runtime.hip)
int batch_upload_common(motex_batch *b);
int batch_clear(motex_batch *b);
int hd_prepare(motex_batch *b);
// Hamming work-queue plan of an all-pairs run (ranges, items grouped for about 16 items per
// resident wave, longest first, LDS window words); want_sym decides the pair-symmetric plan
// from the first ranges pass.
int hd_make_plan(motex_ctx *c, const SeqSet &ss, const motex_params &p, const TileShape &ts,
                 const std::function<bool(const HdPlan &)> &want_sym, HdPlan &plan, bool *sym_out);
int hd_table_words(const SeqSet &ss, const HdPlan &plan, uint64_t *Wout);
double hd_swept_cells(const HdPlan &plan);
int hd_launch(motex_batch *b);
int hd_run_plain(motex_batch *b, uint32_t first_x, uint32_t num_x);
int ed_prepare(motex_batch *b);
int ed_launch_all(motex_batch *b);
int ed_run_plain(motex_batch *b, uint32_t first_x, uint32_t num_x);
int struct_prepare(motex_batch *b);
int struct_launch_all(motex_batch *b);
int struct_range_finish(motex_batch *b);
int struct_run_plain(motex_batch *b, uint32_t first_x, uint32_t num_x);
int pair_run(motex_batch *b, uint32_t i, uint32_t j);
int ed_pair_run(motex_batch *b, uint32_t i, uint32_t j);
int struct_pair_run(motex_batch *b, uint32_t i, uint32_t j);
HdArgs hd_args(motex_batch *b, const HdPlan &plan, const DevBuf &items, uint32_t persistent);
hipError_t hd_alloc_hitq(DevBuf &buf, uint64_t grid, uint32_t blocks, uint32_t cap);
// Whether four-block Hamming launches for (k, e) run the register-ring kernels (which hold
// 4 waves per SIMD).
bool hd_has_ring(uint32_t k, uint32_t e);
// Entries per wave's Hamming hit queue for (k, e) and a launch of `blocks` blocks.
uint32_t hd_hitq_cap(uint32_t k, uint32_t e, uint32_t blocks);
void serving_shutdown(motex_ctx *c);
int noalloc_init(int device);  // looks up the no-allocation scratch array's address on `device`

}  // namespace motex
