// Dispatch of the offset-grouped MFMA M2L (kernels in m2l_mfma.hpp).
#include "m2l_mfma.hpp"

namespace sfmm {

void launch_m2l_mfma(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s)
{
    if (a.n_block == 0) return;
    if (a.fp32) {
        // orders 13..16; the mixed path of lower orders runs m2l_dense.hpp
        (void)m2l_f32_top(a, p, nc, M, L, s);
    } else {
        if (!m2l_f64_lo(a, p, nc, M, L, s) && !m2l_f64_hi(a, p, nc, M, L, s)) m2l_f64_top(a, p, nc, M, L, s);
    }
}

}  // namespace sfmm
