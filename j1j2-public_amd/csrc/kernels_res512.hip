// kernels_res512.hip -- the LDS-resident sweep kernels with 512 threads per
// block (device code in resident.h), one translation unit per block size.
#include "resident.h"

namespace j1j2 {

template <>
ResKernel resident_kernels<512>(int32_t L, bool shadow, bool cubic)
{
    return shadow ? resident_kernel<512, true>(L, cubic) : resident_kernel<512, false>(L, cubic);
}

}  // namespace j1j2
