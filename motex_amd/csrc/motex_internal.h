// Internal structures shared by the host runtime and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace motex {

// Bit offset of text position 0 inside the per-symbol mismatch tables ("NE" tables).
// Positions before 0 and after the end are padding whose bits are never reported.
constexpr uint32_t kPadBits = 4096 + 2048;
constexpr uint32_t kDiagPerBlock = 2048;     // 64 lanes x 32 diagonals
constexpr uint32_t kMaxHdLongK = 1024;       // Hamming factors: 11 counter planes, single-block tiles
constexpr uint32_t kMaxEdK = 64;             // edit kernels: bit-vectors stay below 128 bits
// Text-parallel edit kernel: k up to 30, so that a candidate's popcount (at most k + 1) fits
// its 5 bit-sliced planes.
constexpr uint32_t kMaxSliceK = 30;
constexpr uint32_t kNoSplit = 0xFFFFFFFFu;
constexpr uint32_t kMaxRangeSeqs = 32;       // sequences per Hamming text range (one mask word)
constexpr uint32_t kHitQueueCap = 512;       // Hamming hit queue entries per wave (global)
// Text ranges per pair-symmetric Hamming item at most: each holds a 1 KB first-occurrence
// bitmap in LDS.
constexpr uint32_t kMaxSymGroup = 3;
// Ordered Hamming launches: per range, the text holding each 32-position word of the range's
// positions (word w starts at ys - 32 + 32 w; words up to (8192 + 63) / 32), one byte each.
constexpr uint32_t kTextTabEntries = 258;
constexpr uint32_t kTextTabBytes = 264;
constexpr uint32_t kSymBitmapBits = 8192;    // a range's first-occurrence bitmap (256 words)
constexpr uint32_t kMaxSymItems = 1u << 21;  // pair-symmetric Hamming plans: work items at most
// Pair-symmetric Hamming plans below this occurrence rate per cell: each occurrence in a later
// text also credits the text side (first ones through the range's bitmap and list, repeats by a
// global atomic), which beyond it costs more than the halved sweep saves (1200 x 1000 nt DNA:
// k=8 e=1 at 4e-4 10.5 ms against 17.2 ordered, k=9 e=2 at 1.6e-3 20.0 against 24.1; k=12
// e=4 at 3.4e-3 53.9 against 37.8, k=8 e=2 at 4.2e-3 95 against 50)
constexpr double kSymMaxRate = 2e-3;
constexpr uint32_t kSlicePipe = 4;           // waves per item of the pipelined text-parallel edit kernel
// Edit kernel choice, from sweeps of 40-500 protein and DNA sequences of U[100,2000], k 8-14:
// below kSliceItemsPerCu (x, group of 32 texts) items per CU the one-text-per-lane kernel
// (it and the pipelined kernel cross at about 0.9 items per CU: 7.4 against 6.3 ms at 0.94,
// 9.0 against 6.4 ms at 1.56); above, the text-parallel kernel, pipelined or one wave per
// item by kSlicePipeShare.
constexpr uint32_t kSliceItemsPerCu = 1;
// Pipelined text-parallel kernel when a wave's share of all sweep steps is below this
// fraction of the longest item's steps (runtime_ed.hip).
constexpr double kSlicePipeShare = 0.5;
// Serving jobs (sharing the device): pipelined below this many items per CU.
constexpr uint32_t kSlicePipeItemsPerCu = 12;

// A text range of the Hamming work queue: whole sequences j0 .. j0+ny-1 covering text
// positions [ys, ye), or one block of diagonals of a long sequence j0 (ny = 1, [ys, ye) the
// whole sequence).  A work item's window starts at diagonal ((ys - r1 + 1) rounded down to a
// multiple of 32) + qoff: whole-sequence ranges have qoff = 0; the parts of a long sequence
// have qoff = 0, D, 2D, ... (D = diagonals per window), so that their windows tile the
// diagonals of every row tile exactly.
struct HdRange {
    uint32_t ys, ye, j0, ny;
    uint32_t qoff;
};

// One unit of the Hamming work queue: rows [r0, r1) of sequence x, in row tiles of tile_rows,
// against the text ranges [rb, re), processed in order by one wave that keeps per-row u/v in
// LDS.  Items of the ordered plan hold one row tile; pair-symmetric items every row of x.
struct HdItem {
    uint32_t x_start;   // text position of x's first symbol
    uint32_t uv_base;   // u/v index of x's pos_x = 0
    uint32_t xf_base;   // global factor index of (x, pos_x = 0): row flags for split texts
    uint32_t r0, r1;    // reported rows (positions in x)
    uint32_t px_start;  // first computed row of the first tile = r0 - k + 1
    uint32_t rb, re;    // text ranges
    uint32_t split;     // kNoSplit (whole texts only) or the split slot of the one long text
                        // every range of the item is a part of
    uint32_t xi;        // sequence index of x
    uint32_t sym;       // 1: pair-symmetric item (texts j < xi dropped, j > xi also
                        //    credited to the text side)
    uint32_t pad;
};

struct HdArgs {
    const uint32_t *ne;        // sigma rows x W words
    const uint8_t *codes;      // symbol codes of the concatenated text
    const uint32_t *seq_start; // num_seqs + 1 entries
    const HdItem *items;
    const HdRange *ranges;
    uint32_t *u;
    uint32_t *v;
    uint32_t *split_flags;     // split slots x split_words
    uint32_t *queue;           // work-queue head (one counter)
    uint32_t *hitq;            // per-workgroup hit queues: hitq_cap tags + hitq_cap x B words
    uint32_t hitq_cap;         // entries per hit queue (a multiple of 4, at least 128)
    uint32_t W;                // words per NE row
    uint32_t nitems;
    uint32_t sigma;
    uint32_t k, e;
    uint32_t wl;               // LDS words per symbol row
    uint32_t split_words;
    uint16_t *tlist;           // pair-symmetric runs: per range, first text-side occurrences
    const uint32_t *toff;      //   (bitmap offsets pt - ys + 32) at tlist[toff[r] ..
    uint32_t *tcnt;            //   toff[r + 1]), tcnt[r] of them appended
    uint32_t rows_max;         // largest row tile + k - 1 (LDS row arrays)
    uint32_t tile_rows;        // reported rows per row tile
    uint32_t sym_group;        // ranges per pair-symmetric item at most (LDS text-side bitmaps)
    uint32_t persistent;       // 1: pull items from *queue; 0: item = blockIdx.x
};

// One unit of the edit-distance work queue: x against texts j0 .. j0+ny-1.
struct EdItemDev {
    uint32_t x_start;   // text position of x
    uint32_t m;         // length of x
    uint32_t uv_base;   // u/v index of x's pos_x = 0
    uint32_t j0, ny;    // texts j0 .. j0+ny-1
    uint32_t pad;
};

// Text-parallel edit kernel: one x against a group of up to 32 texts (one bit each).
struct EdSliceItem {
    uint32_t x_start;   // text position of x
    uint32_t m;         // length of x
    uint32_t uv_base;   // u/v index of x's pos_x = 0
    uint32_t group;     // text group
};
struct EdSliceArgs {
    const uint8_t *codes;
    const EdSliceItem *items;
    const uint32_t *gplanes;  // per group row: 5 symbol-code planes + the valid-text mask
    const uint32_t *goff;     // groups + 1 entries: first row of each group in gplanes
    uint32_t *u, *v;
    uint32_t *queue;
    uint32_t *scratch;        // per workgroup: two boundary streams of stream_words words
    uint32_t stream_words;
    uint32_t nitems;
    uint32_t k, e;
    uint32_t persistent;
};

struct EdArgs {
    const uint8_t *codes;
    const uint32_t *seq_start;
    const EdItemDev *items;
    uint32_t *u, *v;
    uint32_t *queue;
    uint32_t *scratch;        // per workgroup: two boundary streams of stream_words words
    uint32_t stream_words;
    uint32_t nitems;
    uint32_t k, e;
    uint32_t persistent;      // 1: pull items from *queue; 0: item = blockIdx.x
};

constexpr uint32_t kMaxGaps = 8;
constexpr uint32_t kMaxBoxes = kMaxGaps + 1;
// Longest sequence of the structured kernels: occurrence lists pack (pos_x, pos_t) as two
// 16-bit fields (the LDS window and x table must also fit, checked at batch creation).
constexpr uint32_t kStructMaxLen = 65535;

// One (x, y) pair of the structured-motif work queue.
struct StructItem {
    uint32_t x_start, m;    // x: text position and length
    uint32_t y_start, n;    // y: text position and length
    uint32_t uv_base;       // u/v index of (structure 0, pos_x 0) of x
    uint32_t uv_base_t;     // kNoSplit, or (pair-symmetric Hamming item) u/v index of
                            // (structure 0, pos 0) of y: the pair (y, x) is credited too
};

struct StructArgs {
    const uint32_t *ne;       // Hamming NE tables (as HdArgs)
    const uint8_t *codes;
    const StructItem *items;
    const uint32_t *S;        // nb_structs x nb_gaps gap lengths (last gap fastest)
    uint32_t *u, *v;
    uint32_t *queue;
    uint32_t *scratch;        // per workgroup: box bitmaps, two edit boundary streams, the
                              // counted bits, then (Hamming) one occurrence list per shape
    uint64_t scratch_words;   // per workgroup
    uint64_t bm_words;        // one box bitmap: ceil(n_max/32) * m_max words
    uint32_t stream_words;
    uint32_t W, sigma, wl;
    uint32_t nitems, persistent;
    uint32_t distance;
    uint32_t nb_gaps, nb_structs;
    uint32_t blen[kMaxBoxes], berr[kMaxBoxes];  // box b (box 0 is (k, e))
    uint32_t gmin[kMaxGaps], gmax[kMaxGaps];
    uint32_t ndist;                             // distinct box shapes
    uint32_t dist_of[kMaxBoxes];                // box b -> distinct shape
    uint32_t dlen[kMaxBoxes], derr[kMaxBoxes];
    uint32_t m_max, hcap;                       // hcap: LDS words of the hit list / counted bits
    uint32_t lcap;                              // Hamming boxes: occurrence list entries per shape
    uint32_t gcap, gather_scratch;              // gathered box-0 hits: entries, and whether they
                                                // live in scratch (after the lists) or in LDS
    uint32_t pad_bits;                          // bit offset of text position 0 in the NE tables
    uint32_t run_words;                         // LDS words per shape for its list's sorted runs
                                                // (count + bounds: 2 + the 2048-diagonal blocks
                                                // of the longest pair)
    const uint32_t *nitems_dev;                 // when set, the item count is read from here
                                                // (at most nitems): the pairs the range kernel
                                                // handed over
};

// Range-batched structured extraction (Hamming boxes).  A unit is every row of one x sequence
// against one text range -- whole texts packed into one window, as for the Hamming kernel.  The
// sweep kernel writes a unit's box hits to its slot; the merge kernel joins them into chains.
struct SrItem {
    uint32_t x_start, m;    // x: text position and length
    uint32_t uv_base;       // u/v index of (structure 0, pos_x 0) of x
    uint32_t r;             // text range
    uint32_t xi;            // sequence index of x
    uint32_t pad;
};

// One x of a range-batched run, for the fold of its accumulators.
struct SrX {
    uint32_t xi, m;
    uint32_t uv_base;
    uint32_t pad;
};

constexpr uint32_t kSrBlocks = 4;                 // 2048-diagonal blocks of a window (8192)
constexpr uint32_t kSrMaxGaps = 4;                // gaps the merge nests (more: per-pair kernel)
constexpr uint32_t kSrMaxBoxes = kSrMaxGaps + 1;
constexpr uint32_t kSrHead = 8;                   // slot header: [0] 1 if the unit overflowed,
                                                  // [1 + d] end of shape d's hits
constexpr uint32_t kSrMergeThreads = 256;
constexpr uint32_t kSrNoSelf = 0xFFFFFFFFu;

struct SrFoldArgs {
    const SrX *xs;
    const uint32_t *self_ovf;
    const uint32_t *S;
    const unsigned long long *uv;
    uint32_t *u, *v;
    uint32_t nb_gaps, nb_structs;
    uint32_t blen[kSrMaxBoxes];
};
constexpr uint32_t kSrKeyBits = 11;               // u key set of a unit: 2048 slots in LDS
constexpr uint32_t kSrKeyProbes = 32;
constexpr uint32_t kSrNoKey = 0xFFFFFFFFu;

struct SrSweepArgs {
    const uint32_t *ne;         // Hamming NE tables, text position 0 at bit pad_bits
    const uint8_t *codes;
    const uint32_t *seq_start;  // num_seqs + 1
    const SrItem *items;
    const HdRange *ranges;
    uint32_t *slots;            // per unit of the launch: kSrHead + hcap words; hits are
                                // px << 16 | (pt - px - q0), shape by shape, rows in order
    uint32_t *queue;            // unit counter
    StructItem *ovf;            // pairs handed to the per-pair kernel (a unit that overflowed)
    uint32_t *ovf_count;        // [0] pairs handed over, [1] set when the list was full
    uint32_t *self_ovf;         // per sequence: 1 if the unit holding its own pair overflowed
    const uint32_t *uvb;        // per sequence: u/v index of (structure 0, position 0)
    uint32_t sym;               // pair-symmetric units: texts before x are skipped, the pairs
                                // (x, j > x) credited both ways
    uint32_t ovf_cap;
    uint32_t W, sigma, pad_bits;
    uint32_t wl;                // NE window words per symbol: 4 blocks of a row's longest span
    uint32_t first, nitems;     // units [first, first + nitems); unit first + i uses slot i
    uint32_t m_max, hcap, ndist;
    uint32_t dlen[kSrMaxBoxes], derr[kSrMaxBoxes];
};

struct SrMergeArgs {
    const uint32_t *slots;
    const SrItem *items;
    const HdRange *ranges;
    const uint32_t *seq_start;
    const uint32_t *S;          // nb_structs x nb_gaps gap lengths
    unsigned long long *uv;     // per u/v index: v in the low word, u in the high word
    const uint32_t *uvb;        // per sequence: u/v index of (structure 0, position 0)
    uint32_t sym;               // pair-symmetric units (one gap)
    uint32_t first;             // unit of slot 0 (one workgroup per slot)
    uint32_t m_max, hcap, ndist, nbk;  // nbk: position buckets of 16, (window + m_max + 15) / 16
    uint32_t nb_gaps, nb_structs;
    uint32_t blen[kSrMaxBoxes], dist_of[kSrMaxBoxes];
    uint32_t gmin[kSrMaxGaps], gmax[kSrMaxGaps];
    uint32_t gstride[kSrMaxGaps];  // combinations per step of gap g (last gap fastest)
};

}  // namespace motex
