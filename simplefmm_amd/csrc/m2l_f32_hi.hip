// M2L MFMA kernels, float stages, orders 9, 10, 11, 12.
#include "m2l_mfma.hpp"

namespace sfmm {

bool m2l_f32_hi(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s)
{
    return m2l_dispatch<float, 9, 10, 11, 12>(a, p, nc, M, L, s);
}

}  // namespace sfmm
