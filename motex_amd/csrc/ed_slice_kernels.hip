// Text-parallel edit-distance kernels for k = 4 .. 14 (the kernels: ed_slice_kernels.h), and
// the dispatch over every k the text-parallel path takes (4 .. kMaxSliceK).
#include "ed_slice_kernels.h"

namespace motex {

uint32_t ed_slice_stream_row_words(uint32_t k)
{
    if (k >= 23) return ed_slice_stream_row_words_k23(k);
    if (k >= 15) return ed_slice_stream_row_words_k15(k);
    return slice_stream_row_words_range<4, 14>(k);
}

// grid: workgroups (one wave each, or kSlicePipe waves each when pipe)
hipError_t launch_ed_slice(const EdSliceArgs &a, uint32_t grid, bool pipe, hipStream_t s)
{
    if (a.k >= 23) return launch_ed_slice_k23(a, grid, pipe, s);
    if (a.k >= 15) return launch_ed_slice_k15(a, grid, pipe, s);
    return launch_slice_range<4, 14>(a, grid, pipe, s);
}

}  // namespace motex
