// M2L on matrix cores (gfx950 MFMA), grouped by interaction offset.
//
// Targets of one level that share a child octant have the same set of (at most
// 189) interaction offsets.  A workgroup takes a block of such targets, stacks
// (target, component) as 16 columns and walks the offsets in order; for each
// offset it gathers the 16 source columns, applies the azimuthal rotation, and
// runs the three block-diagonal products of rotate-translate-rotate
// (fmm.hpp:620-688) as 16x16x4 MFMA tiles through LDS.  The back rotation of
// every offset is rotated in azimuth and accumulated in FP64 registers; each
// target's local expansion is written once at the end (this kernel is the
// first writer of L at every level >= 2).
//
// T = double: FP64 MFMA throughout.  T = float: FP32 MFMA stages with FP32
// operands and FP64 accumulation across offsets (mixed precision).
#pragma once

#include <type_traits>

#include "kernels.hpp"
#include "m2l_sched.hpp"

namespace sfmm {

namespace {

__host__ __device__ constexpr int rfc2(int n, int m) { return m == 0 ? n * n : n * n + 2 * m - 1; }
__host__ __device__ constexpr int rfs2(int n, int m) { return n * n + 2 * m; }

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
    typedef double V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V4 run(double a, double b, V4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of D held in register r of lane l (16x16x4 f64 layout)
    static __device__ __forceinline__ int drow(int lane, int r) { return 4 * r + (lane >> 4); }
};

template <>
struct Mfma<float> {
    typedef float V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V4 run(float a, float b, V4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

template <int P>
__host__ __device__ constexpr M2LSchedule<P> schedule()
{
    return M2LSchedule<P>{};
}

constexpr int kLdsW = 17;  // LDS row stride: 16 columns + 1 to spread banks
constexpr int kPadRows = 8;

template <typename T, int P, int NC, int W>
__device__ __forceinline__ void m2l_wave(const M2LArgs& a, int64_t blk, const double* __restrict__ M,
                                         double* __restrict__ L, T* Xs, T* Ys, T* cs_s)
{
    using F = Mfma<T>;
    using V4 = typename F::V4;
    constexpr int K = (P + 1) * (P + 1);
    constexpr int NT = (P + 1) * (P + 2) / 2;
    constexpr int TB = 16 / NC;
    constexpr int MT = rot_mts(P);
    constexpr int NRF = rot_nfrag(P);
    constexpr int NCF = coax_nfrag(P);
    constexpr auto S = schedule<P>();
    constexpr int NJ = S.s3.count[W];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int col = lane & 15;

    double acc[NJ > 0 ? NJ : 1][MT][2][4];
#pragma unroll
    for (int q = 0; q < (NJ > 0 ? NJ : 1); ++q)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[q][mt][0][r] = acc[q][mt][1][r] = 0.0;

    const int64_t e0 = a.blk_row[blk], e1 = a.blk_row[blk + 1];
    for (int64_t e = e0; e < e1; ++e) {
        const int off = a.ent_off[e];
        const int cls = a.off_class[off];
        const int d2 = a.off_d2[off];
        const double* csg = a.off_cs + static_cast<int64_t>(off) * (P + 1) * 2;
        // ---- gather the 16 source columns with the azimuthal rotation e^{i m phi}
        for (int it = tid; it < 16 * NT; it += 256) {
            const int c16 = it / NT, pr = it - c16 * NT;
            int n = 0;
            while ((n + 1) * (n + 2) / 2 <= pr) ++n;
            const int m = pr - n * (n + 1) / 2;
            const int t = c16 / NC, c = c16 - t * NC;
            double xr = 0.0, yr = 0.0;
            if (t < TB) {
                const int32_t src = a.ent_src[e * TB + t];
                if (src >= 0) {
                    const double* Ms = M + (static_cast<int64_t>(src) * NC + c) * K;
                    const double x = Ms[rfc2(n, m)];
                    if (m > 0) {
                        const double y = Ms[rfs2(n, m)];
                        const double cm = csg[2 * m], sm = csg[2 * m + 1];
                        xr = x * cm - y * sm;
                        yr = x * sm + y * cm;
                    } else {
                        xr = x;
                    }
                }
            }
            Xs[ps_row(n, m, 0) * kLdsW + c16] = static_cast<T>(xr);
            if (m > 0) Xs[ps_row(n, m, 1) * kLdsW + c16] = static_cast<T>(yr);
        }
        for (int i = tid; i < 2 * (P + 1); i += 256) cs_s[i] = static_cast<T>(csg[i]);
        __syncthreads();

        // ---- stage 1: polar rotation, degree-major X -> order-major Y
        {
            const T* A1 = static_cast<const T*>(a.rotA) + (static_cast<int64_t>(cls) * 2 + 0) * NRF * 64;
            static_for<0, S.s1.count[W]>([&](auto qc) {
                constexpr int item = S.s1.item[W][decltype(qc)::value];
                constexpr int n = item >> 8, part = (item >> 4) & 15, mt = item & 15;
                constexpr int nks = rot_nks(n, part);
                constexpr int frag = rot_frag(n, part, mt);
                constexpr int brow = part == 0 ? n * n : n * n + n + 1;
                V4 d = {0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < nks; ++ks) {
                    const T av = A1[(frag + ks) * 64 + lane];
                    const T bv = Xs[(brow + 4 * ks + (lane >> 4)) * kLdsW + col];
                    d = F::run(av, bv, d);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * mt + F::drow(lane, r);
                    if (row <= n && (part == 0 || row > 0)) Ys[om_row(P, row, n, part) * kLdsW + col] = d[r];
                }
            });
        }
        __syncthreads();

        // ---- stage 2: coaxial translation per order, order-major Y -> degree-major X
        {
            const T* A2 = static_cast<const T*>(a.coaxA) + static_cast<int64_t>(d2) * NCF * 64;
            static_for<0, S.s2.count[W]>([&](auto qc) {
                constexpr int item = S.s2.item[W][decltype(qc)::value];
                constexpr int k = item >> 8, part = (item >> 4) & 15, mt = item & 15;
                constexpr int nks = coax_nks(P, k);
                constexpr int frag = coax_frag(P, k, mt);
                constexpr int brow = om_row(P, k, k, part);
                V4 d = {0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < nks; ++ks) {
                    const T av = A2[(frag + ks) * 64 + lane];
                    const T bv = Ys[(brow + 4 * ks + (lane >> 4)) * kLdsW + col];
                    d = F::run(av, bv, d);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = k + 16 * mt + F::drow(lane, r);
                    if (j <= P) Xs[ps_row(j, k, part) * kLdsW + col] = d[r];
                }
            });
        }
        __syncthreads();

        // ---- stage 3: back rotation per degree, azimuth e^{-i m phi}, FP64 accumulation
        {
            const T* A3 = static_cast<const T*>(a.rotA) + (static_cast<int64_t>(cls) * 2 + 1) * NRF * 64;
            static_for<0, NJ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int j = S.s3.item[W][q] >> 8;
                static_for<0, rot_mts(j)>([&](auto mc) {
                    constexpr int mt = decltype(mc)::value;
                    constexpr int nkc = rot_nks(j, 0), nks = rot_nks(j, 1);
                    constexpr int fc = rot_frag(j, 0, mt), fs = rot_frag(j, 1, mt);
                    V4 dc = {0, 0, 0, 0}, ds = {0, 0, 0, 0};
#pragma unroll
                    for (int ks = 0; ks < nkc; ++ks) {
                        const T av = A3[(fc + ks) * 64 + lane];
                        const T bv = Xs[(j * j + 4 * ks + (lane >> 4)) * kLdsW + col];
                        dc = F::run(av, bv, dc);
                    }
#pragma unroll
                    for (int ks = 0; ks < nks; ++ks) {
                        const T av = A3[(fs + ks) * 64 + lane];
                        const T bv = Xs[(j * j + j + 1 + 4 * ks + (lane >> 4)) * kLdsW + col];
                        ds = F::run(av, bv, ds);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * mt + F::drow(lane, r);
                        if (row <= j) {
                            const double cm = cs_s[2 * row], sm = cs_s[2 * row + 1];
                            const double lc = dc[r], ls = ds[r];
                            acc[q][mt][0][r] += lc * cm + ls * sm;
                            acc[q][mt][1][r] += ls * cm - lc * sm;
                        }
                    }
                });
            });
        }
        __syncthreads();
    }

    const int t = col / NC, c = col - (col / NC) * NC;
    if (t < TB) {
        const int32_t tg = a.blk_tgt[blk * TB + t];
        if (tg >= 0) {
            double* Lt = L + (static_cast<int64_t>(tg) * NC + c) * K;
            static_for<0, NJ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int j = S.s3.item[W][q] >> 8;
#pragma unroll
                for (int mt = 0; mt < rot_mts(j); ++mt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * mt + F::drow(lane, r);
                        if (row <= j) {
                            Lt[rfc2(j, row)] = acc[q][mt][0][r];
                            if (row > 0) Lt[rfs2(j, row)] = acc[q][mt][1][r];
                        }
                    }
            });
        }
    }
}

template <typename T, int P, int NC>
__global__ __launch_bounds__(256) void m2l_mfma_kernel(M2LArgs a, const double* __restrict__ M,
                                                       double* __restrict__ L)
{
    constexpr int K = (P + 1) * (P + 1);
    __shared__ T Xs[(K + kPadRows) * kLdsW];
    __shared__ T Ys[(K + kPadRows) * kLdsW];
    __shared__ T cs_s[2 * (P + 1)];
    // XCD-aware order: consecutive blocks (one octant of one level, in Morton
    // order) run on the same XCD and share its L2 for sources and operators.
    const int64_t per = (a.n_block + 7) / 8;
    const int64_t blk = static_cast<int64_t>(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (blk >= a.n_block) return;
    for (int i = threadIdx.x; i < kPadRows * kLdsW; i += 256) {
        Xs[K * kLdsW + i] = T(0);
        Ys[K * kLdsW + i] = T(0);
    }
    __syncthreads();
    switch (threadIdx.x >> 6) {
    case 0: m2l_wave<T, P, NC, 0>(a, blk, M, L, Xs, Ys, cs_s); break;
    case 1: m2l_wave<T, P, NC, 1>(a, blk, M, L, Xs, Ys, cs_s); break;
    case 2: m2l_wave<T, P, NC, 2>(a, blk, M, L, Xs, Ys, cs_s); break;
    default: m2l_wave<T, P, NC, 3>(a, blk, M, L, Xs, Ys, cs_s); break;
    }
}

template <typename T, int P>
void launch_nc(const M2LArgs& a, int nc, const double* M, double* L, hipStream_t s)
{
    const unsigned grid = static_cast<unsigned>(((a.n_block + 7) / 8) * 8);
    switch (nc) {
    case 1: hipLaunchKernelGGL((m2l_mfma_kernel<T, P, 1>), dim3(grid), dim3(256), 0, s, a, M, L); break;
    case 2: hipLaunchKernelGGL((m2l_mfma_kernel<T, P, 2>), dim3(grid), dim3(256), 0, s, a, M, L); break;
    case 3: hipLaunchKernelGGL((m2l_mfma_kernel<T, P, 3>), dim3(grid), dim3(256), 0, s, a, M, L); break;
    default: hipLaunchKernelGGL((m2l_mfma_kernel<T, P, 4>), dim3(grid), dim3(256), 0, s, a, M, L); break;
    }
}

}  // namespace

// Launches the kernel when p is one of Ps; false otherwise.
template <typename T, int... Ps>
bool m2l_dispatch(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s)
{
    bool done = false;
    ((p == Ps ? (launch_nc<T, Ps>(a, nc, M, L, s), done = true) : false), ...);
    return done;
}

// one translation unit per (type, order range) keeps the builds parallel
bool m2l_f64_lo(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s);
bool m2l_f64_hi(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s);
bool m2l_f64_top(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s);
bool m2l_f32_lo(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s);
bool m2l_f32_hi(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s);
bool m2l_f32_top(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s);

}  // namespace sfmm
