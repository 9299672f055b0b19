// Mixed-precision M2L kernels on the f16 matrix cores, orders 9 to 12.
#include "m2l_dense.hpp"

namespace sfmm {

bool m2l_dense_hi(const DenseArgs& a, int p, int nc, const float* M32, double* L, hipStream_t s)
{
    switch (p) {
    case 9: launch_m2l_dense_p<9>(a, nc, M32, L, s); return true;
    case 10: launch_m2l_dense_p<10>(a, nc, M32, L, s); return true;
    case 11: launch_m2l_dense_p<11>(a, nc, M32, L, s); return true;
    case 12: launch_m2l_dense_p<12>(a, nc, M32, L, s); return true;
    default: return false;
    }
}

}  // namespace sfmm
