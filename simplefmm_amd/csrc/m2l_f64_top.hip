// M2L MFMA kernels, double stages, orders 13, 14, 15, 16.
#include "m2l_mfma.hpp"

namespace sfmm {

bool m2l_f64_top(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s)
{
    return m2l_dispatch<double, 13, 14, 15, 16>(a, p, nc, M, L, s);
}

}  // namespace sfmm
