// Tile-shape selection for the batched kernels.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "runtime.h"

namespace motex {

TileShape choose_tile(motex_ctx *c, const motex_params &p, const SeqSet &ss)
{
    TileShape t;
    t.blocks = ss.sigma <= 8 ? 4 : 2;
    t.rows = 256;
    t.waves_per_cu = 24;
    if (const char *env = getenv("MOTEX_TILE")) {  // "blocks,rows,waves_per_cu" (tuning runs)
        unsigned b = 0, r = 0, w = 0;
        if (sscanf(env, "%u,%u,%u", &b, &r, &w) == 3 && (b == 1 || b == 2 || b == 4) && r >= 32 && w >= 1) {
            t.blocks = b;
            t.rows = r;
            t.waves_per_cu = w;
        }
    }
    // keep the Hamming kernel's LDS window under 64 KB
    while (t.blocks > 1 &&
           hd_lds_bytes(ss.sigma, hd_lds_words(t.blocks, t.rows + p.k), t.rows + p.k) > 65536)
        t.blocks /= 2;
    while (t.rows > 64 &&
           hd_lds_bytes(ss.sigma, hd_lds_words(t.blocks, t.rows + p.k), t.rows + p.k) > 65536)
        t.rows /= 2;
    (void)c;
    return t;
}

}  // namespace motex
