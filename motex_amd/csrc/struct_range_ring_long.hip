// Structured range sweep kernels with the leaving column kept in a register ring, box lengths
// 8-10 (one kernel per length and threshold plane).
#include "struct_range_sweep.h"

namespace motex {

hipError_t launch_struct_range_sweep_ring_long(const SrSweepArgs &a, uint32_t len, uint32_t ph, uint32_t grid,
                                               size_t lds, hipStream_t s)
{
    switch (len) {
    case 8: return launch_sr_ring<8>(a, ph, grid, lds, s);
    case 9: return launch_sr_ring<9>(a, ph, grid, lds, s);
    case 10: return launch_sr_ring<10>(a, ph, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace motex
