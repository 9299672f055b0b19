// Dispatch of the mixed-precision M2L on the f16 matrix cores (m2l_dense.hpp),
// the FP32 copy of the multipoles it reads and their per-component range, by
// which it scales.
#include "m2l_dense.hpp"

#include <algorithm>

namespace sfmm {

bool m2l_dense_lo(const DenseArgs& a, int p, int nc, const float* M32, double* L, hipStream_t s);
bool m2l_dense_hi(const DenseArgs& a, int p, int nc, const float* M32, double* L, hipStream_t s);

void launch_m2l_dense(const DenseArgs& a, int p, int nc, const float* M32, double* L, hipStream_t s)
{
    if (a.n_block == 0) return;
    if (!m2l_dense_lo(a, p, nc, M32, L, s)) (void)m2l_dense_hi(a, p, nc, M32, L, s);
}

namespace {

// Slot -> real-form row of the FP32 layout (-1 for a padding slot).
template <int P>
struct F32Rows {
    int r[F32LayoutOf<P>::L.row] = {};
    constexpr F32Rows()
    {
        for (int j = 0; j < F32LayoutOf<P>::L.row; ++j) r[j] = -1;
        for (int k = 0; k < (P + 1) * (P + 1); ++k) r[F32LayoutOf<P>::L.pos[k]] = k;
    }
};
template <int P>
struct F32RowsOf {
    static constexpr F32Rows<P> R{};
};

// One wavefront (workgroup) per node: its nc x (P+1)^2 FP64 multipole
// coefficients are read coalesced into LDS, then the FP32 copy in the layout
// the mixed M2L reads (padding zeroed) is written slot by slot (coalesced), and,
// when nodemax is given, the largest |value| per component is kept.
template <int P>
__global__ __launch_bounds__(64) void fp32_copy_kernel(const double* __restrict__ M, int64_t count, int nc,
                                                       float* __restrict__ M32, float* __restrict__ nodemax)
{
    constexpr int K = (P + 1) * (P + 1);
    constexpr int ROW = F32LayoutOf<P>::L.row;
    __shared__ double xs[4 * K];
    __shared__ int rows[ROW];  // slot -> row, once per workgroup
    const int lane = threadIdx.x;
    for (int j = lane; j < ROW; j += 64) rows[j] = F32RowsOf<P>::R.r[j];
    for (int64_t node = blockIdx.x; node < count; node += gridDim.x) {
        const double* x = M + node * nc * K;
        float* y = M32 + node * nc * ROW;
        __syncthreads();  // the previous node's reads of xs are done
        if ((nc * K) % 2 == 0) {  // 16-byte loads (the node's coefficients start 16-byte aligned)
            const double2* x2 = reinterpret_cast<const double2*>(x);
            for (int i = lane; i < nc * K / 2; i += 64) {
                const double2 v = x2[i];
                xs[2 * i] = v.x;
                xs[2 * i + 1] = v.y;
            }
        } else {
            for (int i = lane; i < nc * K; i += 64) xs[i] = x[i];
        }
        __syncthreads();
        float mx[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (ROW % 4 == 0) {  // four slots per lane, 16-byte stores
            float4* y4 = reinterpret_cast<float4*>(y);
            for (int i = lane; i < nc * ROW / 4; i += 64) {
                const int c = (4 * i) / ROW, j = 4 * i - c * ROW;
                float v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = rows[j + u];
                    v[u] = r >= 0 ? static_cast<float>(xs[c * K + r]) : 0.0f;
                    mx[c] = fmaxf(mx[c], fabsf(v[u]));
                }
                y4[i] = make_float4(v[0], v[1], v[2], v[3]);
            }
        } else {
            for (int i = lane; i < nc * ROW; i += 64) {
                const int c = i / ROW, r = rows[i - c * ROW];
                const float v = r >= 0 ? static_cast<float>(xs[c * K + r]) : 0.0f;
                y[i] = v;
                mx[c] = fmaxf(mx[c], fabsf(v));
            }
        }
        if (!nodemax) continue;
        for (int c = 0; c < nc; ++c)
            for (int o = 32; o > 0; o >>= 1) mx[c] = fmaxf(mx[c], __shfl_xor(mx[c], o));
        if (lane < nc) {
            float v = mx[0];
            for (int c = 1; c < nc; ++c) v = lane == c ? mx[c] : v;
            nodemax[node * nc + lane] = v;
        }
    }
}

template <int P>
bool fp32_copy_p(int p, const double* M, int64_t count, int nc, float* M32, float* nodemax, hipStream_t s)
{
    if constexpr (P > 16) {
        return false;
    } else {
        if (p != P) return fp32_copy_p<P + 1>(p, M, count, nc, M32, nodemax, s);
        const unsigned grid = static_cast<unsigned>(std::min<int64_t>(32768, count));
        hipLaunchKernelGGL(fp32_copy_kernel<P>, dim3(grid), dim3(64), 0, s, M, count, nc, M32, nodemax);
        return true;
    }
}

}  // namespace

void launch_fp32_copy(const double* M, int64_t count, int nc, int p, float* M32, float* nodemax, hipStream_t s)
{
    if (count > 0) (void)fp32_copy_p<1>(p, M, count, nc, M32, nodemax, s);
}

}  // namespace sfmm
