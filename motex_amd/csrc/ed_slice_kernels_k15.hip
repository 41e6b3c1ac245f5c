// Text-parallel edit-distance kernels for k = 15 .. 22 (ed_slice_kernels.h).
#include "ed_slice_kernels.h"

namespace motex {

hipError_t launch_ed_slice_k15(const EdSliceArgs &a, uint32_t grid, bool pipe, hipStream_t s)
{
    return launch_slice_range<15, 22>(a, grid, pipe, s);
}

uint32_t ed_slice_stream_row_words_k15(uint32_t k) { return slice_stream_row_words_range<15, 22>(k); }

}  // namespace motex
