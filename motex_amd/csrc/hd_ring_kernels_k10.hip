// Hamming sweep kernels with the leaving column in a register ring (hd_sweep.h), factor
// lengths 10 and 11, error bounds 0 .. 3.
#include "hd_sweep.h"

namespace motex {

hipError_t launch_hd_ring_k10(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    switch (a.k) {
    case 10: return launch_ring<10>(a, grid, lds, s);
    case 11: return launch_ring<11>(a, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace motex
