// L2L kernels on matrix cores, orders 1..16.
#include "translate_mfma.hpp"

namespace sfmm {

bool translate_upper(const TransArgs& a, LevelSpan b, int p, int nc, double* io, hipStream_t s)
{
    return translate_dispatch<true, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16>(a, b, p, nc, io, s);
}

}  // namespace sfmm
