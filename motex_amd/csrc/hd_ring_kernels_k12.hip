// Hamming sweep kernels with the leaving column in a register ring (hd_sweep.h), factor
// lengths 12 and 13, error bounds 0 .. 3.
#include "hd_sweep.h"

namespace motex {

hipError_t launch_hd_ring_k12(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    switch (a.k) {
    case 12: return launch_ring<12>(a, grid, lds, s);
    case 13: return launch_ring<13>(a, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace motex
