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

constexpr int kWaves = 8;         // wavefronts per workgroup
constexpr int kThreads = 64 * kWaves;

template <int P>
__host__ __device__ constexpr M2LSchedule<P, kWaves> schedule()
{
    return M2LSchedule<P, kWaves>{};
}

constexpr int kLdsW = 17;  // LDS row stride: 16 columns + 1 to spread banks
constexpr int kPadRows = 8;

// k-step offsets of a wave's tiles inside its register copy of the A operands
template <int P, int W>
__host__ __device__ constexpr int s1_koff(int q)
{
    const auto S = schedule<P>();
    int o = 0;
    for (int i = 0; i < q; ++i) o += rot_nks(S.s1.item[W][i] >> 8, (S.s1.item[W][i] >> 4) & 15);
    return o;
}
template <int P, int W>
__host__ __device__ constexpr int s2_koff(int q)
{
    const auto S = schedule<P>();
    int o = 0;
    for (int i = 0; i < q; ++i) o += coax_nks(P, S.s2.item[W][i] >> 8);
    return o;
}
template <int P, int W>
__host__ __device__ constexpr int s3_koff(int q)
{
    const auto S = schedule<P>();
    int o = 0;
    for (int i = 0; i < q; ++i) {
        const int j = S.s3.item[W][i] >> 8;
        o += rot_mts(j) * (rot_nks(j, 0) + rot_nks(j, 1));
    }
    return o;
}

template <typename T, int P, int W>
__device__ __forceinline__ void load_rot_fwd(T* a1, const T* A1, int lane)
{
    constexpr auto S = schedule<P>();
    static_for<0, S.s1.count[W]>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int item = S.s1.item[W][q];
        constexpr int n = item >> 8, part = (item >> 4) & 15, mt = item & 15;
        constexpr int frag = rot_frag(n, part, mt), ko = s1_koff<P, W>(q);
#pragma unroll
        for (int ks = 0; ks < rot_nks(n, part); ++ks) a1[ko + ks] = A1[(frag + ks) * 64 + lane];
    });
}

template <typename T, int P, int W>
__device__ __forceinline__ void load_coax(T* a2, const T* A2, int lane)
{
    constexpr auto S = schedule<P>();
    static_for<0, S.s2.count[W]>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int item = S.s2.item[W][q];
        constexpr int k = item >> 8, mt = item & 15;
        constexpr int frag = coax_frag(P, k, mt), ko = s2_koff<P, W>(q);
#pragma unroll
        for (int ks = 0; ks < coax_nks(P, k); ++ks) a2[ko + ks] = A2[(frag + ks) * 64 + lane];
    });
}

template <typename T, int P, int W>
__device__ __forceinline__ void load_rot_bwd(T* a3, const T* A3, int lane)
{
    constexpr auto S = schedule<P>();
    static_for<0, S.s3.count[W]>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int j = S.s3.item[W][q] >> 8;
        constexpr int ko = s3_koff<P, W>(q);
        constexpr int nkc = rot_nks(j, 0), nks = rot_nks(j, 1);
#pragma unroll
        for (int mt = 0; mt < rot_mts(j); ++mt) {
#pragma unroll
            for (int ks = 0; ks < nkc; ++ks) a3[ko + mt * (nkc + nks) + ks] = A3[(rot_frag(j, 0, mt) + ks) * 64 + lane];
#pragma unroll
            for (int ks = 0; ks < nks; ++ks)
                a3[ko + mt * (nkc + nks) + nkc + ks] = A3[(rot_frag(j, 1, mt) + ks) * 64 + lane];
        }
    });
}

// Entries per block: a target's interaction offsets are the 6x6x6 children
// of its parent's neighbours minus its own 3x3x3 neighbourhood, the same 189
// for every target of one child octant.
constexpr int kMaxEntries = 189;

// Packed entry word: offset id | polar class << 9 | squared-distance class << 16.
__device__ __forceinline__ int ent_offset(int w) { return w & 511; }
__device__ __forceinline__ int ent_class(int w) { return (w >> 9) & 127; }
__device__ __forceinline__ int ent_dist(int w) { return w >> 16; }

// One wavefront's part of a block.  Per entry (offset): gathered columns into
// X (LDS), stage 1 X -> Y, stage 2 Y -> X, stage 3 from X into FP64 registers.
// Every global load an entry needs (source columns, azimuth cos/sin, operators
// when the polar or distance class changes) is issued at the top of the
// previous entry, with no branch between the load and the end of that entry,
// so its latency overlaps a whole entry of matrix work.  The entry list of the
// block sits in LDS.
template <typename T, int P, int NC, int W>
__device__ __forceinline__ void m2l_wave(const M2LArgs& a, int64_t blk, int ne, const double* __restrict__ M,
                                         double* __restrict__ L, T* Xs, T* Ys, T* cs_s, const int* ent_w,
                                         const int* ent_s)
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
    constexpr int NK1 = s1_koff<P, W>(S.s1.count[W]);
    constexpr int NK2 = s2_koff<P, W>(S.s2.count[W]);
    constexpr int NK3 = s3_koff<P, W>(NJ);
    constexpr int CPT = kThreads / 16;        // threads per gathered column
    constexpr int GI = (NT + CPT - 1) / CPT;  // gather items per thread
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int col = lane & 15;
    const T* rotA = static_cast<const T*>(a.rotA);
    const T* coaxA = static_cast<const T*>(a.coaxA);

    double acc[NJ > 0 ? NJ : 1][MT][2][4];
#pragma unroll
    for (int q = 0; q < (NJ > 0 ? NJ : 1); ++q)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[q][mt][0][r] = acc[q][mt][1][r] = 0.0;

    // gather role: column gc = (target gt, component gcc); coefficients pr = gl + CPT*g
    const int gc = tid / CPT, gl = tid - (tid / CPT) * CPT;
    const int gt = gc / NC, gcc = gc - (gc / NC) * NC;
    const bool gon = gt < TB;
    const int gts = gon ? gt : 0;
    int gnm[GI];  // packed n | m << 8 (always a valid coefficient); gv: written
    bool gv[GI];
#pragma unroll
    for (int g = 0; g < GI; ++g) {
        const int pr = gl + CPT * g;
        int n = 0;
        while ((n + 1) * (n + 2) / 2 <= pr) ++n;
        gv[g] = gon && pr < NT;
        gnm[g] = gv[g] ? (n | ((pr - n * (n + 1) / 2) << 8)) : 0;
    }
    const double2* cs2 = reinterpret_cast<const double2*>(a.off_cs);
    const int tcs = tid < 2 * (P + 1) ? tid : 0;

    double gx[GI], gy[GI], gcm[GI], gsm[GI], c3;
    T a1[NK1 > 0 ? NK1 : 1], a2[NK2 > 0 ? NK2 : 1], a3[NK3 > 0 ? NK3 : 1];
    T b1[NK1 > 0 ? NK1 : 1], b2[NK2 > 0 ? NK2 : 1], b3[NK3 > 0 ? NK3 : 1];

    // loads of entry i (all unconditional except the operator sets)
    auto issue = [&](int i, int w_prev, bool first) {
        const int w = ent_w[i];
        const int off = ent_offset(w);
        const int32_t src = ent_s[i * TB + gts];
        const double* Ms = src < 0 ? a.zero + gcc * K : M + (static_cast<int64_t>(src) * NC + gcc) * K;
#pragma unroll
        for (int g = 0; g < GI; ++g) {
            const int n = gnm[g] & 0xff, m = gnm[g] >> 8;
            gx[g] = Ms[rfc2(n, m)];
            gy[g] = Ms[m > 0 ? rfs2(n, m) : rfc2(n, m)];
            const double2 cs = cs2[off * (P + 1) + m];
            gcm[g] = cs.x;
            gsm[g] = cs.y;
        }
        c3 = a.off_cs[off * (P + 1) * 2 + tcs];
        const int cls = __builtin_amdgcn_readfirstlane(ent_class(w));
        const int d2 = __builtin_amdgcn_readfirstlane(ent_dist(w));
        if (first || cls != ent_class(w_prev)) {
            load_rot_fwd<T, P, W>(b1, rotA + (static_cast<int64_t>(cls) * 2 + 0) * NRF * 64, lane);
            load_rot_bwd<T, P, W>(b3, rotA + (static_cast<int64_t>(cls) * 2 + 1) * NRF * 64, lane);
        }
        if (first || d2 != ent_dist(w_prev))
            load_coax<T, P, W>(b2, coaxA + static_cast<int64_t>(d2) * NCF * 64, lane);
    };

    issue(0, 0, true);
    for (int e = 0; e < ne; ++e) {
        // ---- columns of entry e rotated by e^{i m phi} into X; operators of e
#pragma unroll
        for (int g = 0; g < GI; ++g) {
            if (!gv[g]) continue;
            const int n = gnm[g] & 0xff, m = gnm[g] >> 8;
            if (m > 0) {
                Xs[ps_row(n, m, 0) * kLdsW + gc] = static_cast<T>(gx[g] * gcm[g] - gy[g] * gsm[g]);
                Xs[ps_row(n, m, 1) * kLdsW + gc] = static_cast<T>(gx[g] * gsm[g] + gy[g] * gcm[g]);
            } else {
                Xs[ps_row(n, 0, 0) * kLdsW + gc] = static_cast<T>(gx[g]);
            }
        }
        if (tid < 2 * (P + 1)) cs_s[tid] = static_cast<T>(c3);
#pragma unroll
        for (int i = 0; i < (NK1 > 0 ? NK1 : 1); ++i) a1[i] = b1[i];
#pragma unroll
        for (int i = 0; i < (NK2 > 0 ? NK2 : 1); ++i) a2[i] = b2[i];
#pragma unroll
        for (int i = 0; i < (NK3 > 0 ? NK3 : 1); ++i) a3[i] = b3[i];
        // ---- everything entry e+1 needs (the last entry reloads itself, harmlessly)
        const int en = e + 1 < ne ? e + 1 : e;
        issue(en, ent_w[e], false);
        __syncthreads();

        // ---- stage 1: polar rotation, degree-major X -> order-major Y
        static_for<0, S.s1.count[W]>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int item = S.s1.item[W][q];
            constexpr int n = item >> 8, part = (item >> 4) & 15, mt = item & 15;
            constexpr int nks = rot_nks(n, part);
            constexpr int ko = s1_koff<P, W>(q);
            constexpr int brow = part == 0 ? n * n : n * n + n + 1;
            T bv[nks];
#pragma unroll
            for (int ks = 0; ks < nks; ++ks) bv[ks] = Xs[(brow + 4 * ks + (lane >> 4)) * kLdsW + col];
            V4 d = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < nks; ++ks) d = F::run(a1[ko + ks], bv[ks], d);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * mt + F::drow(lane, r);
                if (row <= n && (part == 0 || row > 0)) Ys[om_row(P, row, n, part) * kLdsW + col] = d[r];
            }
        });
        __syncthreads();

        // ---- stage 2: coaxial translation per order, order-major Y -> degree-major X
        static_for<0, S.s2.count[W]>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int item = S.s2.item[W][q];
            constexpr int k = item >> 8, part = (item >> 4) & 15, mt = item & 15;
            constexpr int nks = coax_nks(P, k);
            constexpr int ko = s2_koff<P, W>(q);
            constexpr int brow = om_row(P, k, k, part);
            T bv[nks];
#pragma unroll
            for (int ks = 0; ks < nks; ++ks) bv[ks] = Ys[(brow + 4 * ks + (lane >> 4)) * kLdsW + col];
            V4 d = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < nks; ++ks) d = F::run(a2[ko + ks], bv[ks], d);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = k + 16 * mt + F::drow(lane, r);
                if (j <= P) Xs[ps_row(j, k, part) * kLdsW + col] = d[r];
            }
        });
        __syncthreads();

        // ---- stage 3: back rotation per degree, azimuth e^{-i m phi}, FP64 accumulation
        static_for<0, NJ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int j = S.s3.item[W][q] >> 8;
            constexpr int ko = s3_koff<P, W>(q);
            constexpr int nkc = rot_nks(j, 0), nks = rot_nks(j, 1);
            T bc[nkc], bs[nks > 0 ? nks : 1];
#pragma unroll
            for (int ks = 0; ks < nkc; ++ks) bc[ks] = Xs[(j * j + 4 * ks + (lane >> 4)) * kLdsW + col];
#pragma unroll
            for (int ks = 0; ks < nks; ++ks) bs[ks] = Xs[(j * j + j + 1 + 4 * ks + (lane >> 4)) * kLdsW + col];
            static_for<0, rot_mts(j)>([&](auto mc) {
                constexpr int mt = decltype(mc)::value;
                constexpr int kb = ko + mt * (nkc + nks);
                V4 dc = {0, 0, 0, 0}, ds = {0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < nkc; ++ks) dc = F::run(a3[kb + ks], bc[ks], dc);
#pragma unroll
                for (int ks = 0; ks < nks; ++ks) ds = F::run(a3[kb + nkc + ks], bs[ks], ds);
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

// Targets with an empty interaction list get L = 0.
template <int P, int NC>
__device__ __forceinline__ void m2l_zero(const M2LArgs& a, int64_t blk, double* __restrict__ L)
{
    constexpr int K = (P + 1) * (P + 1);
    constexpr int TB = 16 / NC;
    for (int i = threadIdx.x; i < TB * NC * K; i += kThreads) {
        const int t = i / (NC * K);
        const int32_t tg = a.blk_tgt[blk * TB + t];
        if (tg >= 0) L[static_cast<int64_t>(tg) * NC * K + (i - t * NC * K)] = 0.0;
    }
}

template <typename T, int P, int NC>
__global__ __launch_bounds__(kThreads, 1) void m2l_mfma_kernel(M2LArgs a, const double* __restrict__ M,
                                                              double* __restrict__ L)
{
    constexpr int K = (P + 1) * (P + 1);
    constexpr int TB = 16 / NC;
    __shared__ T Xs[(K + kPadRows) * kLdsW];
    __shared__ T Ys[(K + kPadRows) * kLdsW];
    __shared__ T cs_s[2 * (P + 1)];
    __shared__ int ent_w[kMaxEntries];
    __shared__ int ent_s[kMaxEntries * TB];
    // XCD-aware order: consecutive blocks (children of nearby parents, all
    // octants) run on the same XCD and share its L2 for sources and operators.
    const int64_t per = (a.n_block + 7) / 8;
    const int64_t blk = static_cast<int64_t>(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (blk >= a.n_block) return;
    const int64_t e0 = a.blk_row[blk];
    const int ne = static_cast<int>(a.blk_row[blk + 1] - e0);  // <= kMaxEntries (host-checked)
    if (ne == 0) {
        m2l_zero<P, NC>(a, blk, L);
        return;
    }
    for (int i = threadIdx.x; i < kPadRows * kLdsW; i += kThreads) {
        Xs[K * kLdsW + i] = T(0);
        Ys[K * kLdsW + i] = T(0);
    }
    for (int i = threadIdx.x; i < ne; i += kThreads) ent_w[i] = a.ent_meta[e0 + i];
    for (int i = threadIdx.x; i < ne * TB; i += kThreads) ent_s[i] = a.ent_src[e0 * TB + i];
    __syncthreads();
    switch (threadIdx.x >> 6) {
    case 0: m2l_wave<T, P, NC, 0>(a, blk, ne, M, L, Xs, Ys, cs_s, ent_w, ent_s); break;
    case 1: m2l_wave<T, P, NC, 1>(a, blk, ne, M, L, Xs, Ys, cs_s, ent_w, ent_s); break;
    case 2: m2l_wave<T, P, NC, 2>(a, blk, ne, M, L, Xs, Ys, cs_s, ent_w, ent_s); break;
    case 3: m2l_wave<T, P, NC, 3>(a, blk, ne, M, L, Xs, Ys, cs_s, ent_w, ent_s); break;
    case 4: m2l_wave<T, P, NC, 4>(a, blk, ne, M, L, Xs, Ys, cs_s, ent_w, ent_s); break;
    case 5: m2l_wave<T, P, NC, 5>(a, blk, ne, M, L, Xs, Ys, cs_s, ent_w, ent_s); break;
    case 6: m2l_wave<T, P, NC, 6>(a, blk, ne, M, L, Xs, Ys, cs_s, ent_w, ent_s); break;
    default: m2l_wave<T, P, NC, 7>(a, blk, ne, M, L, Xs, Ys, cs_s, ent_w, ent_s); break;
    }
}

template <typename T, int P>
void launch_nc(const M2LArgs& a, int nc, const double* M, double* L, hipStream_t s)
{
    const unsigned grid = static_cast<unsigned>(((a.n_block + 7) / 8) * 8);
    switch (nc) {
    case 1: hipLaunchKernelGGL((m2l_mfma_kernel<T, P, 1>), dim3(grid), dim3(kThreads), 0, s, a, M, L); break;
    case 2: hipLaunchKernelGGL((m2l_mfma_kernel<T, P, 2>), dim3(grid), dim3(kThreads), 0, s, a, M, L); break;
    case 3: hipLaunchKernelGGL((m2l_mfma_kernel<T, P, 3>), dim3(grid), dim3(kThreads), 0, s, a, M, L); break;
    default: hipLaunchKernelGGL((m2l_mfma_kernel<T, P, 4>), dim3(grid), dim3(kThreads), 0, s, a, M, L); break;
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
