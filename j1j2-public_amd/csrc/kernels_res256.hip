// kernels_res256.hip -- the LDS-resident sweep kernels with 256 threads per
// block (device code in resident.h), one translation unit per block size.
#include "resident.h"

namespace j1j2 {

template <>
ResKernel resident_kernels<256>(int32_t L, bool shadow, bool cubic)
{
    return shadow ? resident_kernel<256, true>(L, cubic) : resident_kernel<256, false>(L, cubic);
}

}  // namespace j1j2
