// Text-parallel edit-distance kernels for k = 23 .. kMaxSliceK (ed_slice_kernels.h).  From
// k = 23 a wave's state exceeds 256 VGPRs and the compiler keeps the rest in AGPRs.
#include "ed_slice_kernels.h"

namespace motex {

hipError_t launch_ed_slice_k23(const EdSliceArgs &a, uint32_t grid, bool pipe, hipStream_t s)
{
    return launch_slice_range<23, (int)kMaxSliceK>(a, grid, pipe, s);
}

uint32_t ed_slice_stream_row_words_k23(uint32_t k) { return slice_stream_row_words_range<23, (int)kMaxSliceK>(k); }

}  // namespace motex
