// Hamming sweep kernels with the leaving column in a register ring (hd_sweep.h), factor
// lengths 8 and 9, error bounds 0 .. 3.
#include "hd_sweep.h"

namespace motex {

hipError_t launch_hd_ring_k8(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    switch (a.k) {
    case 8: return launch_ring<8>(a, grid, lds, s);
    case 9: return launch_ring<9>(a, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace motex
