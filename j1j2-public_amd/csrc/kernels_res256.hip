// kernels_res256.hip -- the LDS-resident sweep kernels with 256 threads per
// block (device code in resident.h), one translation unit per block size.
#include "resident.h"

namespace j1j2 {

template <>
ResKernel resident_kernels<256>(int32_t L, bool shadow, bool cubic, bool flips)
{
    return shadow ? resident_kernel<256, true>(L, cubic, flips) : resident_kernel<256, false>(L, cubic, flips);
}

}  // namespace j1j2
