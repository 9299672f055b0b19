// M2M and L2L on matrix cores.  Both shifts are level-independent dense
// operators in the scaled real form, one per child octant (tables.cpp):
//   M2M (fmm.hpp:558-590): output degree j reads input degrees <= j ("lower")
//   L2L (fmm.hpp:714-745): output degree j reads input degrees >= j ("upper")
// A workgroup stacks 16 columns (16/nc parents x nc components) and has one
// entry per child octant present, holding the children of that octant.
//   M2M: for every entry it gathers the children into LDS and adds the
//        octant's operator times them into FP64 MFMA accumulators, then writes
//        the parents once.
//   L2L: it gathers the parents into LDS once; for every entry it multiplies
//        them by the octant's operator and adds the result to the children.
#pragma once

#include <type_traits>

#include "kernels.hpp"
#include "m2l_dense.hpp"
#include "m2l_sched.hpp"

namespace sfmm {

__host__ __device__ constexpr int tr_K(int P) { return (P + 1) * (P + 1); }
__host__ __device__ constexpr int tr_deg(int r)  // degree of real-form row r
{
    int j = 0;
    while ((j + 1) * (j + 1) <= r) ++j;
    return j;
}
__host__ __device__ constexpr int tr_mts(int P) { return ceil_div(tr_K(P), 16); }
// k-step range [kb, ke) of m-tile mt: the input rows its output rows can read
__host__ __device__ constexpr int tr_kb(int P, bool upper, int mt)
{
    return upper ? (tr_deg(16 * mt) * tr_deg(16 * mt)) / 4 : 0;
}
__host__ __device__ constexpr int tr_ke(int P, bool upper, int mt)
{
    const int last = 16 * mt + 15 < tr_K(P) - 1 ? 16 * mt + 15 : tr_K(P) - 1;
    return upper ? ceil_div(tr_K(P), 4) : ceil_div((tr_deg(last) + 1) * (tr_deg(last) + 1), 4);
}
__host__ __device__ constexpr int tr_frag(int P, bool upper, int mt)
{
    int f = 0;
    for (int q = 0; q < mt; ++q) f += tr_ke(P, upper, q) - tr_kb(P, upper, q);
    return f;
}
__host__ __device__ constexpr int tr_nfrag(int P, bool upper) { return tr_frag(P, upper, tr_mts(P)); }

constexpr int kTrWaves = 4;

struct TrTiles {
    int count[kTrWaves] = {};
    int mt[kTrWaves][24] = {};
};

// m-tiles over the waves: longest k-range first, each to the least loaded wave
template <int P, bool UPPER>
__host__ __device__ constexpr TrTiles tr_schedule()
{
    TrTiles s;
    int load[kTrWaves] = {};
    bool done[64] = {};
    for (int it = 0; it < tr_mts(P); ++it) {
        int best = -1, cost = -1;
        for (int mt = 0; mt < tr_mts(P); ++mt)
            if (!done[mt] && tr_ke(P, UPPER, mt) - tr_kb(P, UPPER, mt) > cost) {
                best = mt;
                cost = tr_ke(P, UPPER, mt) - tr_kb(P, UPPER, mt);
            }
        done[best] = true;
        int w = 0;
        for (int v = 1; v < kTrWaves; ++v)
            if (load[v] < load[w]) w = v;
        load[w] += cost;
        s.mt[w][s.count[w]++] = best;
    }
    return s;
}

namespace {

constexpr int kTrLdsW = 17;  // LDS row stride: 16 columns + 1 to spread banks
constexpr int kTrPad = 4;    // zero rows read by the last k-step

template <int I, int N, class F>
__device__ __forceinline__ void tr_static_for(F&& f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        tr_static_for<I + 1, N>(f);
    }
}

// longest k-range among wave W's m-tiles
template <int P, bool UPPER, int W>
__host__ __device__ constexpr int tr_maxk()
{
    constexpr TrTiles S = tr_schedule<P, UPPER>();
    int m = 0;
    for (int q = 0; q < S.count[W]; ++q) {
        const int k = tr_ke(P, UPPER, S.mt[W][q]) - tr_kb(P, UPPER, S.mt[W][q]);
        m = k > m ? k : m;
    }
    return m;
}

// Operator fragments are requested this many k-steps ahead of their MFMAs:
// an L2 hit takes several MFMA durations.
constexpr int kTrDepth = 4;

template <int P, int NC, bool UPPER, int W>
__device__ __forceinline__ void translate_wave(const TransArgs& a, int64_t blk, double* __restrict__ io, double* Xs)
{
    typedef double V4 __attribute__((ext_vector_type(4)));
    constexpr int K = tr_K(P);
    constexpr int TB = 16 / NC;
    constexpr int NF = tr_nfrag(P, UPPER);
    constexpr TrTiles S = tr_schedule<P, UPPER>();
    constexpr int NT = S.count[W];
    constexpr int NQ = NT > 0 ? NT : 1;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int col = lane & 15;
    const int t = col / NC, c = col - (col / NC) * NC;  // this lane's output node slot and component
    V4 acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = V4{0, 0, 0, 0};

    // The inputs (16 columns x K rows of the nodes in[0..TB)) into LDS: every
    // load of a thread is issued before the first is used, without branches
    // (absent columns read node 0 and are zeroed), so that their latencies
    // overlap.  The first barrier waits for the previous entry's MFMAs.
    auto gather = [&](const int32_t* in) {
        constexpr int NX = ceil_div(16 * K, 64 * kTrWaves);
        double xv[NX];
#pragma unroll
        for (int j = 0; j < NX; ++j) {
            const int it = tid + 64 * kTrWaves * j;
            const int c16 = it / K, r = it - (it / K) * K;
            const int tt = c16 / NC, cc = c16 - (c16 / NC) * NC;
            const bool col_in = it < 16 * K && tt < TB;
            const int32_t src = in[col_in ? tt : 0];
            const bool have = col_in && src >= 0;
            const int64_t at = (static_cast<int64_t>(have ? src : 0) * NC + (col_in ? cc : 0)) * K + (col_in ? r : 0);
            xv[j] = io[at];
            if (!have) xv[j] = 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NX; ++j) {
            const int it = tid + 64 * kTrWaves * j;
            const int c16 = it / K, r = it - (it / K) * K;
            if (it < 16 * K) Xs[r * kTrLdsW + c16] = xv[j];
        }
        __syncthreads();
    };

    constexpr int KM = tr_maxk<P, UPPER, W>();
    constexpr int DK = kTrDepth;
    double ab[DK][NQ];  // requested fragments: k-step kk of tile q in ab[kk % DK][q]
    // the first k-steps' fragments of octant oct's operator
    auto request = [&](int oct) {
        const double* A = a.frag + static_cast<int64_t>(oct) * NF * 64 + lane;
        tr_static_for<0, DK>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            tr_static_for<0, NT>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int mt = S.mt[W][q];
                if constexpr (j < tr_ke(P, UPPER, mt) - tr_kb(P, UPPER, mt))
                    ab[j][q] = A[(tr_frag(P, UPPER, mt) + j) * 64];
            });
        });
        return A;
    };
    // k-step by k-step across the wave's tiles, so that consecutive MFMAs are
    // independent (each tile's chain keeps its order); every step is unrolled
    // (a branch makes the compiler wait for all loads in flight) and each
    // operator fragment is requested kTrDepth k-steps ahead.
    auto multiply = [&](const double* A) {
        tr_static_for<0, KM + DK>([&](auto kc) {
            constexpr int kk = decltype(kc)::value;
            tr_static_for<0, NT>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int mt = S.mt[W][q];
                constexpr int kb = tr_kb(P, UPPER, mt), len = tr_ke(P, UPPER, mt) - kb;
                constexpr int f0 = tr_frag(P, UPPER, mt);
                if constexpr (kk < len)
                    acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        ab[kk % DK][q], Xs[(4 * (kb + kk) + (lane >> 4)) * kTrLdsW + col], acc[q], 0, 0, 0);
                if constexpr (kk + DK < len) ab[kk % DK][q] = A[(f0 + kk + DK) * 64];
            });
            __builtin_amdgcn_sched_barrier(0);
        });
    };

    const int64_t e0 = a.blk_row[blk], e1 = a.blk_row[blk + 1];
    if constexpr (UPPER) {
        gather(a.blk_node + blk * TB);  // the parents
        for (int64_t e = e0; e < e1; ++e) {
            const double* A = request(a.ent_oct[e]);
            // this lane's rows of the children of the entry, read before the
            // MFMAs (absent children read node 0's rows and are not written)
            const int32_t dst = t < TB ? a.ent_node[e * TB + t] : -1;
            double* out = io + (static_cast<int64_t>(dst >= 0 ? dst : 0) * NC + c) * K;
            double old[NQ][4];
            tr_static_for<0, NT>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int mt = S.mt[W][q];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * mt + 4 * r + (lane >> 4);  // f64 16x16x4 D layout
                    old[q][r] = out[row < K ? row : K - 1];
                }
            });
            multiply(A);
            tr_static_for<0, NT>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int mt = S.mt[W][q];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * mt + 4 * r + (lane >> 4);
                    if (dst >= 0 && row < K) out[row] = old[q][r] + acc[q][r];
                }
                acc[q] = V4{0, 0, 0, 0};
            });
        }
    } else {
        for (int64_t e = e0; e < e1; ++e) {
            const double* A = request(a.ent_oct[e]);
            gather(a.ent_node + e * TB);  // the children of the entry's octant
            multiply(A);
        }
        if (t >= TB) return;
        const int32_t dst = a.blk_node[blk * TB + t];
        if (dst < 0) return;
        double* out = io + (static_cast<int64_t>(dst) * NC + c) * K;
        tr_static_for<0, NT>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int mt = S.mt[W][q];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * mt + 4 * r + (lane >> 4);  // f64 16x16x4 D layout
                if (row < K) out[row] = acc[q][r];
            }
        });
    }
}

template <int P, int NC, bool UPPER>
__global__ __launch_bounds__(64 * kTrWaves) void translate_kernel(TransArgs a, int64_t blk0, double* __restrict__ io)
{
    constexpr int K = tr_K(P);
    __shared__ double Xs[(K + kTrPad) * kTrLdsW];
    const int64_t blk = blk0 + blockIdx.x;
    for (int i = threadIdx.x; i < kTrPad * kTrLdsW; i += 64 * kTrWaves) Xs[K * kTrLdsW + i] = 0.0;
    const int w = threadIdx.x >> 6;
    tr_static_for<0, kTrWaves>([&](auto wc) {
        if (w == decltype(wc)::value) translate_wave<P, NC, UPPER, decltype(wc)::value>(a, blk, io, Xs);
    });
}

template <int P, bool UPPER>
void translate_nc(const TransArgs& a, LevelSpan b, int nc, double* io, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(b.count)), blk(64 * kTrWaves);
    switch (nc) {
    case 1: hipLaunchKernelGGL((translate_kernel<P, 1, UPPER>), grid, blk, 0, s, a, b.first, io); break;
    case 2: hipLaunchKernelGGL((translate_kernel<P, 2, UPPER>), grid, blk, 0, s, a, b.first, io); break;
    case 3: hipLaunchKernelGGL((translate_kernel<P, 3, UPPER>), grid, blk, 0, s, a, b.first, io); break;
    default: hipLaunchKernelGGL((translate_kernel<P, 4, UPPER>), grid, blk, 0, s, a, b.first, io); break;
    }
}

template <bool UPPER, int... Ps>
bool translate_dispatch(const TransArgs& a, LevelSpan b, int p, int nc, double* io, hipStream_t s)
{
    bool done = false;
    ((p == Ps ? (translate_nc<Ps, UPPER>(a, b, nc, io, s), done = true) : false), ...);
    return done;
}

}  // namespace

}  // namespace sfmm
