// kernels_res576.hip -- the LDS-resident sweep kernels with 576 threads per
// block (device code in resident.h), one translation unit per block size.
#include "resident.h"

namespace j1j2 {

template <>
ResKernel resident_kernels<576>(int32_t L, bool shadow, bool cubic)
{
    return shadow ? resident_kernel<576, true>(L, cubic) : resident_kernel<576, false>(L, cubic);
}

}  // namespace j1j2
