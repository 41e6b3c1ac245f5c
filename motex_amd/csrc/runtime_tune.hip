// Tile-shape selection for the batched kernels: built-in defaults per (alphabet size,
// distance), a per-context cache of shapes picked by timing (motex_autotune), and the
// MOTEX_TILE override for tuning runs.
//
// The tunable parameters:
//   Hamming     blocks (2048-diagonal blocks per wave: 1, 2, 4), rows (x rows per work
//               item) and waves_per_cu (persistent grid).  blocks x alphabet size sets the
//               NE window in LDS and rows the per-row accumulators; together they bound the
//               waves resident per CU, which the Hamming sweep is sensitive to.
//   edit        waves_per_cu.
//   structured  waves_per_cu.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hd_device.h"
#include "runtime.h"

namespace motex {

// Shape the autotuner is timing on this thread (overrides the cache and the defaults).
static thread_local const TileShape *t_force_tile = nullptr;

static std::tuple<uint32_t, uint32_t, uint32_t, uint32_t> tune_key(const motex_params &p, const SeqSet &ss)
{
    const uint32_t kind = p.nb_gaps ? 2u : p.distance;
    return {kind, ss.sigma, p.k, p.e};
}

// Defaults from sweeps on MI355X: the Hamming kernel runs best with the smallest LDS
// footprint that still amortises the k-1 warm-up rows of an item (128 rows), at full
// occupancy (32 waves per CU), with the NE window narrowed for large alphabets.
static TileShape default_tile(const motex_params &p, const SeqSet &ss)
{
    TileShape t;
    if (!p.nb_gaps && p.distance == 0) {
        t.blocks = ss.sigma <= 8 ? 4 : (ss.sigma <= 24 ? 2 : 1);
        // (144 rows with four blocks: the warm-up rows a smaller share, 1000-nt texts still 8
        // to a range; 1200 x 1000 nt, generic kernels: k=16 e=2 8.35 -> 8.26 ms, k=6 e=0
        // 10.16 -> 9.90, k=20 e=4 8.13 -> 7.92)
        t.rows = t.blocks == 4 ? 144 : 128;
        t.waves_per_cu = 32;
        if (ss.sigma <= 24 && hd_has_ring(p.k, p.e)) {
            // the register-ring kernels take four blocks up to 24 symbols too, though the NE
            // window then holds a CU to 5-6 waves (500 proteins, k=8 e=1: 3.44 ms against
            // 4.54 with the generic two-block kernels at 32 waves per CU; k=12 e=3: 3.38
            // against 4.38).  They hold 4 waves per SIMD (LDS and registers), and their
            // tiles take 144 rows: 1000-nt texts still pack 8 to a range, and the k - 1 warm-up
            // rows are a smaller share (1200 x 1000 nt: 6.40 ms at 128 rows, 6.31 at 144,
            // 6.36 at 160; H. sapiens genes 6.98, 6.83, 6.84).  Texts longer than a window are
            // cut into blocks of diagonals that do not depend on the row tile, so when every
            // text is, longer tiles amortise the warm-up rows further (one 4-Mnt sequence:
            // 168 ms at 128 rows, 162.5 at 256; 1000-nt texts lose a text per range above 160)
            t.blocks = 4;
            t.waves_per_cu = 16;
            t.rows = 144;
            bool all_long = true;
            for (uint32_t j = 0; j < ss.num && all_long; ++j)
                all_long = ss.len[j] < p.k || ss.len[j] > kDiagPerBlock * 4 - 256 - 32;
            if (all_long) t.rows = 256;
        }
    } else if (!p.nb_gaps && p.k >= 4 && p.k <= kMaxSliceK && ss.sigma <= 32) {
        // text-parallel edit kernel: 100-200 VGPRs up to k = 14 (2-4 waves per SIMD
        // resident), 2 waves per SIMD beyond
        t.waves_per_cu = p.k <= 14 ? 12 : 8;
    } else if (p.nb_gaps == 1 && p.distance == 0 && ss.sigma <= 8) {
        // structured list merge on a small alphabet: the kernel is held to 8 waves per SIMD
        t.blocks = 4;
        t.rows = 256;
        t.waves_per_cu = 32;
    } else {
        t.blocks = 4;
        t.rows = 256;
        t.waves_per_cu = 24;
    }
    return t;
}

static bool hd_fits(const SeqSet &ss, uint32_t k, uint32_t blocks, uint32_t rows)
{
    const uint32_t rows_max = rows + k;
    const uint32_t wl = hd_lds_words(blocks, rows_max);
    return hd_lds_bytes(ss.sigma, wl, rows_max, kMaxSymGroup) <= 65536 && (uint64_t)ss.sigma * wl < 65536;
}

TileShape choose_tile(motex_ctx *c, const motex_params &p, const SeqSet &ss)
{
    TileShape t = default_tile(p, ss);
    {
        std::lock_guard<std::mutex> lk(c->tune_mu);
        auto it = c->tune_cache.find(tune_key(p, ss));
        if (it != c->tune_cache.end()) t = it->second;
    }
    if (t_force_tile) t = *t_force_tile;
    if (const char *env = getenv("MOTEX_TILE")) {  // "blocks,rows,waves_per_cu" (tuning runs)
        unsigned b = 0, r = 0, w = 0;
        if (sscanf(env, "%u,%u,%u", &b, &r, &w) == 3 && (b == 1 || b == 2 || b == 4) && r >= 32 && w >= 1) {
            t.blocks = b;
            t.rows = r;
            t.waves_per_cu = w;
        }
    }
    if (!p.nb_gaps && p.distance == 0 && hd_counter_planes(p.k, p.e) > 7) {
        // factors longer than 64: the generic single-block kernel, with at least k rows per
        // item so that the k - 1 warm-up rows are at most half of them
        t.blocks = 1;
        t.rows = std::max<uint32_t>(t.rows, (p.k + 63) & ~63u);
    }
    // keep the Hamming kernel's LDS window under 64 KB
    while (t.blocks > 1 && !hd_fits(ss, p.k, t.blocks, t.rows)) t.blocks /= 2;
    while (t.rows > 64 && !hd_fits(ss, p.k, t.blocks, t.rows)) t.rows /= 2;
    return t;
}

// Times one candidate shape on the caller's sequences: one untimed run, then the best of
// two timed runs of the batched kernel (HIP events on the context stream).
static int time_shape(motex_ctx *c, const char *const *seqs, const uint32_t *lens, uint32_t num,
                      const motex_params *p, const TileShape &shape, double *ms)
{
    t_force_tile = &shape;
    motex_batch *b = nullptr;
    int rc = batch_create(c, seqs, lens, num, p, nullptr, &b);
    t_force_tile = nullptr;
    if (rc) return rc;
    double best = 1e30;
    for (int rep = 0; rep < 3 && !rc; ++rep) {
        rc = batch_launch(b);
        if (!rc) rc = batch_finish(b);
        if (!rc && rep > 0) best = std::min(best, b->stats.main_kernel_ms);
    }
    motex_batch_destroy(b);
    *ms = best;
    return rc;
}

}  // namespace motex

using namespace motex;

extern "C" int motex_autotune(motex_ctx *c, const char *const *seqs, const uint32_t *lens, uint32_t num_seqs,
                              const motex_params *p, motex_tile *best)
{
    if (!c) return set_error(MOTEX_EINVAL, "ctx is NULL");
    if (!seqs || !lens || !num_seqs) return set_error(MOTEX_EINVAL, "autotune needs a sample of sequences");
    int rc = validate_params(p);
    if (rc) return rc;
    if (hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(e, "hipSetDevice");
    uint32_t nb_structs = 1;
    enumerate_structs(p, &nb_structs);
    SeqSet ss;
    if ((rc = ss.build(seqs, lens, num_seqs, nb_structs))) return rc;
    std::vector<TileShape> cand;
    if (!p->nb_gaps && p->distance == 0) {
        for (uint32_t blocks : {1u, 2u, 4u})
            for (uint32_t rows : {64u, 128u, 192u, 256u})
                for (uint32_t waves : {16u, 24u, 32u}) {
                    if (!hd_fits(ss, p->k, blocks, rows)) continue;
                    TileShape t;
                    t.blocks = blocks;
                    t.rows = rows;
                    t.waves_per_cu = waves;
                    cand.push_back(t);
                }
    } else {
        for (uint32_t waves : {8u, 16u, 24u, 32u}) {
            TileShape t = default_tile(*p, ss);
            t.waves_per_cu = waves;
            cand.push_back(t);
        }
    }
    if (cand.empty()) return set_error(MOTEX_EUNSUPPORTED, "no tile shape fits this alphabet");
    TileShape win = cand[0];
    double win_ms = 1e30;
    for (const TileShape &t : cand) {
        double ms = 0;
        if ((rc = time_shape(c, seqs, lens, num_seqs, p, t, &ms))) return rc;
        if (ms < win_ms) {
            win_ms = ms;
            win = t;
        }
    }
    {
        std::lock_guard<std::mutex> lk(c->tune_mu);
        c->tune_cache[tune_key(*p, ss)] = win;
    }
    if (best) {
        best->blocks = win.blocks;
        best->rows = win.rows;
        best->waves_per_cu = win.waves_per_cu;
        best->candidates = (uint32_t)cand.size();
        best->kernel_ms = win_ms;
    }
    return MOTEX_OK;
}
