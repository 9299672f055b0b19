// Mixed-precision M2L kernels on the f16 matrix cores, orders 1 to 8.
#include "m2l_dense.hpp"

namespace sfmm {

bool m2l_dense_lo(const DenseArgs& a, int p, int nc, const float* M32, double* L, hipStream_t s)
{
    switch (p) {
    case 1: launch_m2l_dense_p<1>(a, nc, M32, L, s); return true;
    case 2: launch_m2l_dense_p<2>(a, nc, M32, L, s); return true;
    case 3: launch_m2l_dense_p<3>(a, nc, M32, L, s); return true;
    case 4: launch_m2l_dense_p<4>(a, nc, M32, L, s); return true;
    case 5: launch_m2l_dense_p<5>(a, nc, M32, L, s); return true;
    case 6: launch_m2l_dense_p<6>(a, nc, M32, L, s); return true;
    case 7: launch_m2l_dense_p<7>(a, nc, M32, L, s); return true;
    case 8: launch_m2l_dense_p<8>(a, nc, M32, L, s); return true;
    default: return false;
    }
}

}  // namespace sfmm
