// Structured range sweep kernels with the leaving column kept in a register ring, box lengths
// 5-7 (one kernel per length and threshold plane).
#include "struct_range_sweep.h"

namespace motex {

hipError_t launch_struct_range_sweep_ring_short(const SrSweepArgs &a, uint32_t len, uint32_t ph, uint32_t grid,
                                                size_t lds, hipStream_t s)
{
    switch (len) {
    case 5: return launch_sr_ring<5>(a, ph, grid, lds, s);
    case 6: return launch_sr_ring<6>(a, ph, grid, lds, s);
    case 7: return launch_sr_ring<7>(a, ph, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace motex
