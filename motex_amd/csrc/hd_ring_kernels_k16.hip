// Hamming sweep kernels with the leaving column in a register ring (hd_sweep.h), factor
// length 16, error bounds 0 .. 3.
#include "hd_sweep.h"

namespace motex {

hipError_t launch_hd_ring_k16(const HdArgs &a, uint32_t grid, size_t lds, hipStream_t s)
{
    switch (a.k) {
    case 16: return launch_ring<16>(a, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace motex
