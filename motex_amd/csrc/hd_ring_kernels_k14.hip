// Hamming sweep kernels with the leaving column in a register ring (hd_sweep.h), factor
// lengths 14 and 15, error bounds 0 .. 3.
#include "hd_sweep.h"

namespace motex {

hipError_t launch_hd_ring_k14(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    switch (a.k) {
    case 14: return launch_ring<14>(a, grid, lds, s);
    case 15: return launch_ring<15>(a, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace motex
