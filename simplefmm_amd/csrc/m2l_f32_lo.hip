// M2L MFMA kernels, float stages, orders 1, 2, 3, 4, 5, 6, 7, 8.
#include "m2l_mfma.hpp"

namespace sfmm {

bool m2l_f32_lo(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s)
{
    return m2l_dispatch<float, 1, 2, 3, 4, 5, 6, 7, 8>(a, p, nc, M, L, s);
}

}  // namespace sfmm
