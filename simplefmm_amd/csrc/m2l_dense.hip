// Dispatch of the mixed-precision M2L on the f16 matrix cores (m2l_dense.hpp),
// and the per-component range of the FP32 multipoles it scales by.
#include "m2l_dense.hpp"

namespace sfmm {

bool m2l_dense_lo(const DenseArgs& a, int p, int nc, const float* M32, double* L, hipStream_t s);
bool m2l_dense_hi(const DenseArgs& a, int p, int nc, const float* M32, double* L, hipStream_t s);

void launch_m2l_dense(const DenseArgs& a, int p, int nc, const float* M32, double* L, hipStream_t s)
{
    if (a.n_block == 0) return;
    if (!m2l_dense_lo(a, p, nc, M32, L, s)) (void)m2l_dense_hi(a, p, nc, M32, L, s);
}

namespace {

// One wavefront per node: lanes stride over its nc x kd floats four at a time
// (kd is a multiple of 64, so every row is 16-byte aligned), keep the largest
// |value| per component, and reduce across the wavefront.
__global__ __launch_bounds__(256) void nodemax_kernel(const float* __restrict__ M32, int64_t count, int nc, int kd,
                                                      float* __restrict__ nodemax)
{
    const int lane = threadIdx.x & 63;
    const int kd4 = kd / 4;
    for (int64_t node = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / 64; node < count;
         node += static_cast<int64_t>(gridDim.x) * blockDim.x / 64) {
        const float4* x = reinterpret_cast<const float4*>(M32 + node * nc * kd);
        float mx[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int i = lane; i < nc * kd4; i += 64) {
            const int c = i / kd4;
            const float4 v = x[i];
            const float m = fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w)));
            mx[c] = fmaxf(mx[c], m);
        }
        for (int c = 0; c < nc; ++c)
            for (int o = 32; o > 0; o >>= 1) mx[c] = fmaxf(mx[c], __shfl_xor(mx[c], o));
        if (lane < nc) {
            float v = mx[0];
            for (int c = 1; c < nc; ++c) v = lane == c ? mx[c] : v;
            nodemax[node * nc + lane] = v;
        }
    }
}

}  // namespace

void launch_dense_nodemax(const float* M32, int64_t count, int nc, int kd, float* nodemax, hipStream_t s)
{
    if (count <= 0) return;
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>(8192, (count + 3) / 4));
    hipLaunchKernelGGL(nodemax_kernel, dim3(grid), dim3(256), 0, s, M32, count, nc, kd, nodemax);
}

}  // namespace sfmm
