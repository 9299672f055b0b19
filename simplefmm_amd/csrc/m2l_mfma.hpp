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
    // The FP32 operator fragments hold operator row 4 (i % 4) + i / 4 in MFMA
    // row i, so the expansion row held in register r of lane l is 4 r + l / 16,
    // as for FP64: register r covers rows [4r, 4r + 4) in both.
    static __device__ __forceinline__ int drow(int lane, int r) { return 4 * r + (lane >> 4); }
};

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// Schedule configuration SC = 2 * (wavefronts per workgroup) + (1 if the
// three stages share a phase, pipelined kernel).
__host__ __device__ constexpr int sc_waves(int SC) { return SC / 2; }

template <int P, int SC>
__host__ __device__ constexpr M2LSchedule<P, SC / 2, SC % 2 == 1> schedule()
{
    return M2LSchedule<P, SC / 2, SC % 2 == 1>{};
}

// Calls f(integral_constant<W>) for the wave index w in [0, NW).
template <int W, int NW, class F>
__device__ __forceinline__ void for_wave(int w, F&& f)
{
    if constexpr (W < NW) {
        if (w == W) f(std::integral_constant<int, W>{});
        else for_wave<W + 1, NW>(w, f);
    }
}

constexpr int kLdsW = 17;  // LDS row stride: 16 columns + 1 to spread banks
constexpr int kPadRows = 8;

// k-step offsets of a wave's tiles inside its register copy of the A operands
template <int P, int SC, int W>
__host__ __device__ constexpr int s1_koff(int q)
{
    const auto S = schedule<P, SC>();
    int o = 0;
    for (int i = 0; i < q; ++i) o += Packed<P>::s1.g[S.s1.item[W][i]].nks();
    return o;
}
template <int P, int SC, int W>
__host__ __device__ constexpr int s2_koff(int q)
{
    const auto S = schedule<P, SC>();
    int o = 0;
    for (int i = 0; i < q; ++i) o += Packed<P>::s2.g[S.s2.item[W][i]].nks();
    return o;
}

// LDS rows of the packed groups (m2l_sched.hpp).  Stage 1 reads degree-major
// X and writes order-major Y; stage 2 reads Y and writes degree-major X.
// Columns past a group's pieces read the zero row K; rows past them write the
// trash row.
template <int P>
__host__ __device__ constexpr int s1_in_row(int gi, int f)
{
    const Packing& pk = Packed<P>::s1;
    const int q = pk.piece_of_col(gi, f);
    if (q < 0) return (P + 1) * (P + 1);
    const Piece& pc = pk.g[gi].pc[q];
    const int c = f - pk.g[gi].col0[q];
    return pc.part == 0 ? pc.n * pc.n + c : pc.n * pc.n + pc.n + 1 + c;
}
template <int P>
__host__ __device__ constexpr int s1_out_row(int gi, int i);
template <int P>
__host__ __device__ constexpr int s2_in_row(int gi, int f)
{
    const Packing& pk = Packed<P>::s2;
    const int q = pk.piece_of_col(gi, f);
    if (q < 0) return (P + 1) * (P + 1);
    const Piece& pc = pk.g[gi].pc[q];
    return om_row(P, pc.n, pc.n + f - pk.g[gi].col0[q], pc.part);
}
template <int P>
__host__ __device__ constexpr int s2_out_row(int gi, int i);

// Stage 3 input row of packed column f of group gi's cos (part 0) or sin (part 1)
// tile; degree j reads X rows j^2 + b (cos) and j^2 + j + b (sin, b >= 1).
template <int P>
__host__ __device__ constexpr int s3_in_row(int gi, int part, int f)
{
    const Packing& pk = Packed<P>::s3;
    const int q = part == 0 ? pk.piece_of_col(gi, f) : pk.piece_of_col_s(gi, f);
    if (q < 0) return (P + 1) * (P + 1);
    const int j = pk.g[gi].pc[q].n;
    return part == 0 ? j * j + f - pk.g[gi].col0[q] : j * j + j + 1 + f - pk.g[gi].col0s[q];
}
// Degree (or -1) and order of packed row i of stage-3 group gi.
template <int P>
__host__ __device__ constexpr int s3_row_j(int gi, int i)
{
    const int q = Packed<P>::s3.piece_of_row(gi, i);
    return q < 0 ? -1 : Packed<P>::s3.g[gi].pc[q].n;
}
template <int P>
__host__ __device__ constexpr int s3_row_a(int gi, int i)
{
    const int q = Packed<P>::s3.piece_of_row(gi, i);
    return q < 0 ? 0 : Packed<P>::s3.g[gi].pc[q].r0 + i - Packed<P>::s3.g[gi].row0[q];
}

// One of four compile-time rows by h = lane / 16 (a run of four when possible).
template <int R0, int R1, int R2, int R3>
__device__ __forceinline__ int sel_row(int h)
{
    if constexpr (R1 == R0 + 1 && R2 == R0 + 2 && R3 == R0 + 3) return R0 + h;
    else return h < 2 ? (h == 0 ? R0 : R1) : (h == 2 ? R2 : R3);
}
template <int P, int SC, int W>
__host__ __device__ constexpr int s3_koff(int q)
{
    const auto S = schedule<P, SC>();
    int o = 0;
    for (int i = 0; i < q; ++i) {
        const Group& G = Packed<P>::s3.g[S.s3.item[W][i]];
        o += G.nks() + G.nks_s();
    }
    return o;
}

template <typename T, int P, int SC, int W>
__device__ __forceinline__ void load_rot_fwd(T* a1, const T* A1, int lane)
{
    constexpr auto S = schedule<P, SC>();
    static_for<0, S.s1.count[W]>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int g = S.s1.item[W][q];
        constexpr int frag = Packed<P>::s1.frag(g), ko = s1_koff<P, SC, W>(q), nk = Packed<P>::s1.g[g].nks();
#pragma unroll
        for (int ks = 0; ks < nk; ++ks) a1[ko + ks] = A1[(frag + ks) * 64 + lane];
    });
}

template <typename T, int P, int SC, int W>
__device__ __forceinline__ void load_coax(T* a2, const T* A2, int lane)
{
    constexpr auto S = schedule<P, SC>();
    static_for<0, S.s2.count[W]>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int g = S.s2.item[W][q];
        constexpr int frag = Packed<P>::s2.frag(g), ko = s2_koff<P, SC, W>(q), nk = Packed<P>::s2.g[g].nks();
#pragma unroll
        for (int ks = 0; ks < nk; ++ks) a2[ko + ks] = A2[(frag + ks) * 64 + lane];
    });
}

template <typename T, int P, int SC, int W>
__device__ __forceinline__ void load_rot_bwd(T* a3, const T* A3, int lane)
{
    constexpr auto S = schedule<P, SC>();
    static_for<0, S.s3.count[W]>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int g = S.s3.item[W][q];
        constexpr int ko = s3_koff<P, SC, W>(q), frag = Packed<P>::s3.frag3(g);
        constexpr int nk = Packed<P>::s3.g[g].nks() + Packed<P>::s3.g[g].nks_s();
#pragma unroll
        for (int ks = 0; ks < nk; ++ks) a3[ko + ks] = A3[(frag + ks) * 64 + lane];
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

// ---------------------------------------------------------------- stages
// Each wave runs its share (schedule<P>, wave W) of a stage over the 16
// columns held in LDS.  A stage first reads every B operand it needs, then runs
// its MFMA chains, then writes; rows that fall outside the expansion go to the
// trash row (kTrashRow) instead of being masked, so the whole stage is one
// basic block the compiler can schedule freely.
constexpr int kTrashRow = 0;  // relative to the end of the padded rows

template <int P>
__host__ __device__ constexpr int s1_out_row(int gi, int i)
{
    const Packing& pk = Packed<P>::s1;
    const int q = pk.piece_of_row(gi, i);
    if (q < 0) return (P + 1) * (P + 1) + kPadRows + kTrashRow;
    const Piece& pc = pk.g[gi].pc[q];
    const int a = pc.r0 + i - pk.g[gi].row0[q];  // sin rows start at order 1
    return om_row(P, pc.part == 0 ? a : a + 1, pc.n, pc.part);
}
template <int P>
__host__ __device__ constexpr int s2_out_row(int gi, int i)
{
    const Packing& pk = Packed<P>::s2;
    const int q = pk.piece_of_row(gi, i);
    if (q < 0) return (P + 1) * (P + 1) + kPadRows + kTrashRow;
    const Piece& pc = pk.g[gi].pc[q];
    return ps_row(pc.n + pc.r0 + i - pk.g[gi].row0[q], pc.n, pc.part);
}

// (cos, sin) pair of the azimuth table, read with one LDS access
template <typename T>
struct alignas(2 * sizeof(T)) CsPair {
    T c, s;
};

// B operands of stage 1 / 2 / 3 for wave W, in the same k-order as the A registers.
template <int P, int SC, int W>
__host__ __device__ constexpr int s3_boff(int q)
{
    return s3_koff<P, SC, W>(q);
}

template <typename T, int P, int SC, int W>
struct StageIn {
    static constexpr auto S = schedule<P, SC>();
    static constexpr int N1 = s1_koff<P, SC, W>(S.s1.count[W]);
    static constexpr int N2 = s2_koff<P, SC, W>(S.s2.count[W]);
    static constexpr int N3 = s3_boff<P, SC, W>(S.s3.count[W]);
    T b1[N1 > 0 ? N1 : 1], b2[N2 > 0 ? N2 : 1], b3[N3 > 0 ? N3 : 1];
};

template <typename T, int P, int SC, int W>
__device__ __forceinline__ void read_rot_fwd(T* b, const T* X, int lane)
{
    constexpr auto S = schedule<P, SC>();
    const int col = lane & 15, h = lane >> 4;
    static_for<0, S.s1.count[W]>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int g = S.s1.item[W][q];
        constexpr int ko = s1_koff<P, SC, W>(q);
        static_for<0, Packed<P>::s1.g[g].nks()>([&](auto kc) {
            constexpr int ks = decltype(kc)::value;
            const int row = sel_row<s1_in_row<P>(g, 4 * ks), s1_in_row<P>(g, 4 * ks + 1), s1_in_row<P>(g, 4 * ks + 2),
                                    s1_in_row<P>(g, 4 * ks + 3)>(h);
            b[ko + ks] = X[row * kLdsW + col];
        });
    });
}

template <typename T, int P, int SC, int W>
__device__ __forceinline__ void read_coax(T* b, const T* Y, int lane)
{
    constexpr auto S = schedule<P, SC>();
    const int col = lane & 15, h = lane >> 4;
    static_for<0, S.s2.count[W]>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int g = S.s2.item[W][q];
        constexpr int ko = s2_koff<P, SC, W>(q);
        static_for<0, Packed<P>::s2.g[g].nks()>([&](auto kc) {
            constexpr int ks = decltype(kc)::value;
            const int row = sel_row<s2_in_row<P>(g, 4 * ks), s2_in_row<P>(g, 4 * ks + 1), s2_in_row<P>(g, 4 * ks + 2),
                                    s2_in_row<P>(g, 4 * ks + 3)>(h);
            b[ko + ks] = Y[row * kLdsW + col];
        });
    });
}

template <typename T, int P, int SC, int W>
__device__ __forceinline__ void read_rot_back(T* b, const T* X, int lane)
{
    constexpr auto S = schedule<P, SC>();
    const int col = lane & 15, h = lane >> 4;
    static_for<0, S.s3.count[W]>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int g = S.s3.item[W][q];
        constexpr int bo = s3_boff<P, SC, W>(q);
        constexpr int nkc = Packed<P>::s3.g[g].nks();
        static_for<0, nkc>([&](auto kc) {
            constexpr int ks = decltype(kc)::value;
            const int row = sel_row<s3_in_row<P>(g, 0, 4 * ks), s3_in_row<P>(g, 0, 4 * ks + 1),
                                    s3_in_row<P>(g, 0, 4 * ks + 2), s3_in_row<P>(g, 0, 4 * ks + 3)>(h);
            b[bo + ks] = X[row * kLdsW + col];
        });
        static_for<0, Packed<P>::s3.g[g].nks_s()>([&](auto kc) {
            constexpr int ks = decltype(kc)::value;
            const int row = sel_row<s3_in_row<P>(g, 1, 4 * ks), s3_in_row<P>(g, 1, 4 * ks + 1),
                                    s3_in_row<P>(g, 1, 4 * ks + 2), s3_in_row<P>(g, 1, 4 * ks + 3)>(h);
            b[bo + nkc + ks] = X[row * kLdsW + col];
        });
    });
}

// Longest k-chain among a wave's tiles of each stage.
template <int P, int SC, int W>
__host__ __device__ constexpr int s1_maxks()
{
    const auto S = schedule<P, SC>();
    int m = 0;
    for (int q = 0; q < S.s1.count[W]; ++q) {
        const int nk = Packed<P>::s1.g[S.s1.item[W][q]].nks();
        m = nk > m ? nk : m;
    }
    return m;
}
template <int P, int SC, int W>
__host__ __device__ constexpr int s2_maxks()
{
    const auto S = schedule<P, SC>();
    int m = 0;
    for (int q = 0; q < S.s2.count[W]; ++q) {
        const int nk = Packed<P>::s2.g[S.s2.item[W][q]].nks();
        m = nk > m ? nk : m;
    }
    return m;
}
template <int P, int SC, int W>
__host__ __device__ constexpr int s3_maxks()
{
    const auto S = schedule<P, SC>();
    int m = 0;
    for (int q = 0; q < S.s3.count[W]; ++q) {
        const int nk = Packed<P>::s3.g[S.s3.item[W][q]].nks();
        m = nk > m ? nk : m;
    }
    return m;
}

// The MFMA chains of a stage run k-step by k-step across its tiles, so that
// consecutive MFMAs are independent.

// Stage 1: polar rotation per degree, degree-major X -> order-major Y.
template <typename T, int P, int SC, int W>
__device__ __forceinline__ void stage_rot_fwd(const T* a1, const T* b, T* Y, int lane)
{
    using F = Mfma<T>;
    using V4 = typename F::V4;
    constexpr auto S = schedule<P, SC>();
    constexpr int NQ = S.s1.count[W];
    const int col = lane & 15;
    V4 d[NQ > 0 ? NQ : 1];
#pragma unroll
    for (int q = 0; q < NQ; ++q) d[q] = V4{0, 0, 0, 0};
    static_for<0, s1_maxks<P, SC, W>()>([&](auto kc) {
        constexpr int ks = decltype(kc)::value;
        static_for<0, NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int g = S.s1.item[W][q];
            constexpr int ko = s1_koff<P, SC, W>(q);
            if constexpr (ks < Packed<P>::s1.g[g].nks()) d[q] = F::run(a1[ko + ks], b[ko + ks], d[q]);
        });
    });
    const int h = lane >> 4;  // register r holds packed rows 4r + h
    static_for<0, NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int g = S.s1.item[W][q];
        static_for<0, 4>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if constexpr (4 * r < Packed<P>::s1.g[g].rows) {
                const int dst = sel_row<s1_out_row<P>(g, 4 * r), s1_out_row<P>(g, 4 * r + 1), s1_out_row<P>(g, 4 * r + 2),
                                        s1_out_row<P>(g, 4 * r + 3)>(h);
                Y[dst * kLdsW + col] = d[q][r];
            }
        });
    });
}

// Stage 2: coaxial translation per order, order-major Y -> degree-major X.
template <typename T, int P, int SC, int W>
__device__ __forceinline__ void stage_coax(const T* a2, const T* b, T* X, int lane)
{
    using F = Mfma<T>;
    using V4 = typename F::V4;
    constexpr auto S = schedule<P, SC>();
    constexpr int NQ = S.s2.count[W];
    const int col = lane & 15;
    V4 d[NQ > 0 ? NQ : 1];
#pragma unroll
    for (int q = 0; q < NQ; ++q) d[q] = V4{0, 0, 0, 0};
    static_for<0, s2_maxks<P, SC, W>()>([&](auto kc) {
        constexpr int ks = decltype(kc)::value;
        static_for<0, NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int g = S.s2.item[W][q];
            constexpr int ko = s2_koff<P, SC, W>(q);
            if constexpr (ks < Packed<P>::s2.g[g].nks()) d[q] = F::run(a2[ko + ks], b[ko + ks], d[q]);
        });
    });
    const int h = lane >> 4;  // register r holds packed rows 4r + h
    static_for<0, NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int g = S.s2.item[W][q];
        static_for<0, 4>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if constexpr (4 * r < Packed<P>::s2.g[g].rows) {
                const int dst = sel_row<s2_out_row<P>(g, 4 * r), s2_out_row<P>(g, 4 * r + 1), s2_out_row<P>(g, 4 * r + 2),
                                        s2_out_row<P>(g, 4 * r + 3)>(h);
                X[dst * kLdsW + col] = d[q][r];
            }
        });
    });
}

// Stage 3: back rotation per degree, azimuth e^{-i m phi} from cs, FP64
// accumulation.  Rows beyond a degree have all-zero operator rows, so their
// results are zero and add nothing (cs is read at a clamped, valid index).
template <typename T, int P, int SC, int W, class Acc>
__device__ __forceinline__ void stage_rot_back(const T* a3, const T* b, const T* cs, Acc& acc, int lane)
{
    using F = Mfma<T>;
    using V4 = typename F::V4;
    constexpr auto S = schedule<P, SC>();
    constexpr int NQ = S.s3.count[W];
    V4 dc[NQ > 0 ? NQ : 1], ds[NQ > 0 ? NQ : 1];
#pragma unroll
    for (int q = 0; q < NQ; ++q) dc[q] = ds[q] = V4{0, 0, 0, 0};
    static_for<0, s3_maxks<P, SC, W>()>([&](auto kc) {
        constexpr int ks = decltype(kc)::value;
        static_for<0, NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int g = S.s3.item[W][q];
            constexpr int ko = s3_koff<P, SC, W>(q);
            constexpr int nkc = Packed<P>::s3.g[g].nks(), nks = Packed<P>::s3.g[g].nks_s();
            if constexpr (ks < nkc) dc[q] = F::run(a3[ko + ks], b[ko + ks], dc[q]);
            if constexpr (ks < nks) ds[q] = F::run(a3[ko + nkc + ks], b[ko + nkc + ks], ds[q]);
        });
    });
    const int h = lane >> 4;  // register r holds packed rows 4r + h
    static_for<0, NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int g = S.s3.item[W][q];
        static_for<0, 4>([&](auto rc_) {
            constexpr int r = decltype(rc_)::value;
            if constexpr (4 * r < Packed<P>::s3.g[g].rows) {  // rows past the pieces hold zeros
                const int m = sel_row<s3_row_a<P>(g, 4 * r), s3_row_a<P>(g, 4 * r + 1), s3_row_a<P>(g, 4 * r + 2),
                                      s3_row_a<P>(g, 4 * r + 3)>(h);
                const CsPair<T> e = reinterpret_cast<const CsPair<T>*>(cs)[m];
                const T lc = dc[q][r], ls = ds[q][r];
                acc[q][0][0][r] += static_cast<double>(lc * e.c + ls * e.s);
                acc[q][0][1][r] += static_cast<double>(ls * e.c - lc * e.s);
            }
        });
    });
}

// Source columns of one entry as a thread gathers them: for each of its
// items (n, m) the pair (Re, Im) of coefficient (n, m), or (c(n,0), next) for m = 0.
template <int GI, typename S>
struct Cols {
    S x[GI], y[GI];
};


// Gather role and operator registers of a wave.
template <typename T, int P, int NC, int SC, int W>
struct WaveState {
    static constexpr int K = (P + 1) * (P + 1);
    static constexpr int NT = (P + 1) * (P + 2) / 2;
    static constexpr int TB = 16 / NC;
    static constexpr int NRF = rot_nfrag(P);
    static constexpr int NCF = Packed<P>::s2.nfrag();  // packed stage-2 fragments per distance class
    static constexpr auto S = schedule<P, SC>();
    static constexpr int NJ = S.s3.count[W];
    static constexpr int NK1 = s1_koff<P, SC, W>(S.s1.count[W]);
    static constexpr int NK2 = s2_koff<P, SC, W>(S.s2.count[W]);
    static constexpr int NK3 = s3_koff<P, SC, W>(NJ);
    static constexpr int CPT = 4 * sc_waves(SC);     // threads per gathered column (64 NW / 16)
    static constexpr int GI = (NT + CPT - 1) / CPT;  // gather items per thread

    int gc, gcc, gts;
    int gnm[GI];  // packed n | m << 8 (always a valid coefficient)
    bool gv[GI];  // item written
    T a1[NK1 > 0 ? NK1 : 1], a2[NK2 > 0 ? NK2 : 1], a3[NK3 > 0 ? NK3 : 1];  // operators in use
    T b1[NK1 > 0 ? NK1 : 1], b2[NK2 > 0 ? NK2 : 1], b3[NK3 > 0 ? NK3 : 1];  // operators loading

    __device__ __forceinline__ void init(int tid)
    {
        gc = tid / CPT;
        const int gl = tid - gc * CPT;
        const int gt = gc / NC;
        gcc = gc - gt * NC;
        const bool gon = gt < TB;
        gts = gon ? gt : 0;
#pragma unroll
        for (int g = 0; g < GI; ++g) {
            const int pr = gl + CPT * g;
            int n = 0;
            while ((n + 1) * (n + 2) / 2 <= pr) ++n;
            gv[g] = gon && pr < NT;
            gnm[g] = gv[g] ? (n | ((pr - n * (n + 1) / 2) << 8)) : 0;
        }
    }

    // Columns of the entry with source row srow: one 16-byte load per item
    // (the real-form pair of (n, m) is adjacent; for m = 0 the second value is
    // the next coefficient, or the zero padding after the last one).
    // S: element type of the source expansions (double, or the FP32 copy).
    template <typename S>
    __device__ __forceinline__ void load_cols(Cols<GI, S>& c, const M2LArgs& a, const S* __restrict__ M,
                                              const int* srow) const
    {
        struct alignas(sizeof(S)) P2 {
            S x, y;
        };
        const int32_t src = srow[gts];
        const S* Ms = src < 0 ? reinterpret_cast<const S*>(a.zero) + gcc * K
                              : M + (static_cast<int64_t>(src) * NC + gcc) * K;
#pragma unroll
        for (int g = 0; g < GI; ++g) {
            const int n = gnm[g] & 0xff, m = gnm[g] >> 8;
            const P2 v = *reinterpret_cast<const P2*>(Ms + rfc2(n, m));
            c.x[g] = v.x;
            c.y[g] = v.y;
        }
    }

    // Columns rotated by e^{i m phi} (cos, sin from cs) into X (degree-major);
    // items this thread does not own, and the sine half of m = 0, go to the trash row.
    // (the rotation is computed in the stage precision T)
    template <typename S>
    __device__ __forceinline__ void store_cols(T* X, const Cols<GI, S>& c, const T* cs) const
    {
        constexpr int trash = K + kPadRows + kTrashRow;
#pragma unroll
        for (int g = 0; g < GI; ++g) {
            const int n = gnm[g] & 0xff, m = gnm[g] >> 8;
            const CsPair<T> e = reinterpret_cast<const CsPair<T>*>(cs)[m];
            const T x = static_cast<T>(c.x[g]);
            const T y = m > 0 ? static_cast<T>(c.y[g]) : T(0);
            const int rc = gv[g] ? ps_row(n, m, 0) : trash;
            const int rs = gv[g] && m > 0 ? ps_row(n, m, 1) : trash;
            X[rc * kLdsW + gc] = x * e.c - y * e.s;
            X[rs * kLdsW + gc] = x * e.s + y * e.c;
        }
    }

    __device__ __forceinline__ void use_loaded()
    {
#pragma unroll
        for (int i = 0; i < (NK1 > 0 ? NK1 : 1); ++i) a1[i] = b1[i];
#pragma unroll
        for (int i = 0; i < (NK2 > 0 ? NK2 : 1); ++i) a2[i] = b2[i];
#pragma unroll
        for (int i = 0; i < (NK3 > 0 ? NK3 : 1); ++i) a3[i] = b3[i];
    }

    __device__ __forceinline__ void load_fwd(const M2LArgs& a, int cls, int lane)
    {
        load_rot_fwd<T, P, SC, W>(b1, static_cast<const T*>(a.rotA) + (static_cast<int64_t>(cls) * 2 + 0) * NRF * 64, lane);
    }
    __device__ __forceinline__ void load_back(const M2LArgs& a, int cls, int lane)
    {
        load_rot_bwd<T, P, SC, W>(b3, static_cast<const T*>(a.rotA) + (static_cast<int64_t>(cls) * 2 + 1) * NRF * 64, lane);
    }
    __device__ __forceinline__ void load_cx(const M2LArgs& a, int d2, int lane)
    {
        load_coax<T, P, SC, W>(b2, static_cast<const T*>(a.coaxA) + static_cast<int64_t>(d2) * NCF * 64, lane);
    }
};

// cos, sin of m * azimuth of the entry with word w, m = lane / 2 (lanes past
// 2(p+1) read a valid element and are not stored).
template <int P>
__device__ __forceinline__ double load_cs(const M2LArgs& a, int w, int lane)
{
    return a.off_cs[ent_offset(w) * (P + 1) * 2 + (lane < 2 * (P + 1) ? lane : 0)];
}

// FP64 accumulators of a wave's stage-3 degrees: [item][m-tile][cos/sin][4]
template <int NJ, int MT>
struct Acc {
    double v[NJ > 0 ? NJ : 1][MT][2][4];
    __device__ __forceinline__ auto& operator[](int q) { return v[q]; }
    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int q = 0; q < (NJ > 0 ? NJ : 1); ++q)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) v[q][mt][0][r] = v[q][mt][1][r] = 0.0;
    }
};

template <typename T, int P, int NC, int SC, int W, class A>
__device__ __forceinline__ void write_locals(const M2LArgs& a, int64_t blk, double* __restrict__ L, A& acc, int lane)
{
    constexpr int K = (P + 1) * (P + 1);
    constexpr int TB = 16 / NC;
    constexpr auto S = schedule<P, SC>();
    const int col = lane & 15, h = lane >> 4;
    const int t = col / NC, c = col - (col / NC) * NC;
    if (t >= TB) return;
    const int32_t tg = a.blk_tgt[blk * TB + t];
    if (tg < 0) return;
    double* Lt = L + (static_cast<int64_t>(tg) * NC + c) * K;
    static_for<0, S.s3.count[W]>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int g = S.s3.item[W][q];
        static_for<0, 4>([&](auto rc_) {
            constexpr int r = decltype(rc_)::value;
            if constexpr (4 * r < Packed<P>::s3.g[g].rows) {
                const int j = sel_row<s3_row_j<P>(g, 4 * r), s3_row_j<P>(g, 4 * r + 1), s3_row_j<P>(g, 4 * r + 2),
                                      s3_row_j<P>(g, 4 * r + 3)>(h);
                const int m = sel_row<s3_row_a<P>(g, 4 * r), s3_row_a<P>(g, 4 * r + 1), s3_row_a<P>(g, 4 * r + 2),
                                      s3_row_a<P>(g, 4 * r + 3)>(h);
                if (j >= 0) {
                    Lt[rfc2(j, m)] = acc[q][0][0][r];
                    if (m > 0) Lt[rfs2(j, m)] = acc[q][0][1][r];
                }
            }
        });
    });
}

// ---------------------------------------------------------------- pipelined kernel
// Entries flow through a three-deep software pipeline: in phase i a wave
// stores the columns of entry i, runs stage 1 on entry i-1, stage 2 on entry
// i-2 and stage 3 on entry i-3; one barrier ends the phase.  Stage buffers
// alternate by entry parity (G: stored columns, Y: after stage 1, Z: after
// stage 2), so no two concurrent steps touch the same buffer.  Loads run ahead:
// source columns kAhead<T> entries (with two, two register sets alternate by
// phase parity), operators and the azimuth table one entry.  Entries outside [0, ne) run on
// zeroed or repeated data with zero azimuth factors, so they add nothing.
constexpr int kCsSlots = 8;  // azimuth table slots in LDS, by entry index mod 8

// Column prefetch distance: two entries for FP32 stages (which gather from
// the FP32 copy of the multipoles), one for FP64 (a second set of FP64
// columns would spill at p = 12).
template <typename T>
constexpr int kAhead = sizeof(T) == 4 ? 2 : 1;

// Source expansions the kernel gathers from: M itself, or its FP32 copy.
template <typename T>
__device__ __forceinline__ const T* m2l_source(const M2LArgs& a, const double* M);
template <>
__device__ __forceinline__ const double* m2l_source<double>(const M2LArgs&, const double* M)
{
    return M;
}
template <>
__device__ __forceinline__ const float* m2l_source<float>(const M2LArgs& a, const double*)
{
    return a.M32;
}

template <typename T, int P, int NC>
__host__ __device__ constexpr size_t pipe_lds_bytes()
{
    return 6 * static_cast<size_t>((P + 1) * (P + 1) + kPadRows + 1) * kLdsW * sizeof(T) +
           (kCsSlots * 2 * (P + 1) + 64) * sizeof(T) + static_cast<size_t>(kMaxEntries) * (1 + 16 / NC) * sizeof(int);
}

template <typename T, int P, int NC>
__host__ __device__ constexpr bool pipe_fits()
{
    return pipe_lds_bytes<T, P, NC>() <= 160 * 1024;
}

template <typename T, int P, int NC, int SC, int W>
__device__ __forceinline__ void m2l_pipe_wave(const M2LArgs& a, int64_t blk, int ne, const double* __restrict__ M,
                                              double* __restrict__ L, T* buf, T* cst, const int* ent_w,
                                              const int* ent_s)
{
    using WS = WaveState<T, P, NC, SC, W>;
    constexpr int K = (P + 1) * (P + 1);
    constexpr int BS = (K + kPadRows + 1) * kLdsW;
    constexpr int TB = 16 / NC;
    constexpr int NCS = 2 * (P + 1);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    T* G = buf;
    T* Y = buf + 2 * BS;
    T* Z = buf + 4 * BS;
    auto ce = [&](int k) { return k < 0 ? 0 : (k >= ne ? ne - 1 : k); };
    auto cs_slot = [&](int k) { return cst + (k & (kCsSlots - 1)) * NCS; };

    Acc<WS::NJ, 1> acc;
    acc.zero();
    WS ws;
    ws.init(tid);
    constexpr int AH = kAhead<T>;
    const T* Msrc = m2l_source<T>(a, M);
    Cols<WS::GI, T> c0, c1, c2;  // columns of the entries ahead (sets beyond AH unused)
    ws.load_cols(c0, a, Msrc, ent_s + ce(0) * TB);
    if constexpr (AH >= 2) ws.load_cols(c1, a, Msrc, ent_s + ce(1) * TB);
    if constexpr (AH >= 3) ws.load_cols(c2, a, Msrc, ent_s + ce(2) * TB);
    double cs_next = 0.0;
    if constexpr (W == 0) {
        const double v = load_cs<P>(a, ent_w[0], lane);
        cs_next = load_cs<P>(a, ent_w[ce(1)], lane);
        if (lane < NCS) cs_slot(0)[lane] = static_cast<T>(v);
    }
    {
        const int w = ent_w[0];
        const int cls = __builtin_amdgcn_readfirstlane(ent_class(w));
        const int d2 = __builtin_amdgcn_readfirstlane(ent_dist(w));
        ws.load_fwd(a, cls, lane);
        ws.load_cx(a, d2, lane);
        ws.load_back(a, cls, lane);
    }
    __syncthreads();

    auto phase = [&](int i, Cols<WS::GI, T>& cols) {
        // ---- columns of entry i (loaded two phases ago); azimuth table of entry i+1
        ws.store_cols(G + (i & 1) * BS, cols, cs_slot(i));
        if constexpr (W == 0) {  // lanes past the table write the trash tail
            T* dst = lane < NCS ? cs_slot(i + 1) + lane : cst + kCsSlots * NCS + lane;
            *dst = i + 1 < ne ? static_cast<T>(cs_next) : T(0);
            cs_next = load_cs<P>(a, ent_w[ce(i + 2)], lane);
        }
        ws.use_loaded();
        // ---- loads: columns of entry i+AH, operators of entries i, i-1, i-2
        ws.load_cols(cols, a, Msrc, ent_s + ce(i + AH) * TB);
        const int w0 = ent_w[ce(i)], w1 = ent_w[ce(i - 1)], w2 = ent_w[ce(i - 2)], w3 = ent_w[ce(i - 3)];
        const int k0 = __builtin_amdgcn_readfirstlane(ent_class(w0));
        const int k1 = __builtin_amdgcn_readfirstlane(ent_class(w1));
        const int k2 = __builtin_amdgcn_readfirstlane(ent_class(w2));
        const int k3 = __builtin_amdgcn_readfirstlane(ent_class(w3));
        const int d1 = __builtin_amdgcn_readfirstlane(ent_dist(w1));
        const int d2 = __builtin_amdgcn_readfirstlane(ent_dist(w2));
        if (k0 != k1) ws.load_fwd(a, k0, lane);
        if (d1 != d2) ws.load_cx(a, d1, lane);
        if (k2 != k3) ws.load_back(a, k2, lane);
        // ---- the three stages on entries i-1, i-2, i-3
        StageIn<T, P, SC, W> in;
        read_rot_fwd<T, P, SC, W>(in.b1, G + ((i - 1) & 1) * BS, lane);
        read_coax<T, P, SC, W>(in.b2, Y + (i & 1) * BS, lane);
        read_rot_back<T, P, SC, W>(in.b3, Z + ((i - 1) & 1) * BS, lane);
        stage_rot_back<T, P, SC, W>(ws.a3, in.b3, cs_slot(i - 3), acc, lane);
        stage_rot_fwd<T, P, SC, W>(ws.a1, in.b1, Y + ((i - 1) & 1) * BS, lane);
        stage_coax<T, P, SC, W>(ws.a2, in.b2, Z + (i & 1) * BS, lane);
        __syncthreads();
    };
    // ne + 3 phases drain the pipeline; with AH column sets, up to AH - 1
    // more phases (a multiple of AH) add nothing
    if constexpr (AH == 3) {
        for (int i = 0; i < ne + 3; i += 3) {
            phase(i, c0);
            phase(i + 1, c1);
            phase(i + 2, c2);
        }
    } else if constexpr (AH == 2) {
        for (int i = 0; i < ne + 3; i += 2) {
            phase(i, c0);
            phase(i + 1, c1);
        }
    } else {
        for (int i = 0; i < ne + 3; ++i) phase(i, c0);
    }
    write_locals<T, P, NC, SC, W>(a, blk, L, acc, lane);
}

// Targets with an empty interaction list get L = 0.
template <int P, int NC>
__device__ __forceinline__ void m2l_zero(const M2LArgs& a, int64_t blk, double* __restrict__ L)
{
    constexpr int K = (P + 1) * (P + 1);
    constexpr int TB = 16 / NC;
    for (int i = threadIdx.x; i < TB * NC * K; i += blockDim.x) {
        const int t = i / (NC * K);
        const int32_t tg = a.blk_tgt[blk * TB + t];
        if (tg >= 0) L[static_cast<int64_t>(tg) * NC * K + (i - t * NC * K)] = 0.0;
    }
}

// Block index in XCD-aware order: consecutive blocks (children of nearby
// parents, all octants) run on the same XCD and share its L2 for sources and
// operators.
__device__ __forceinline__ int64_t xcd_block(int64_t n_block)
{
    const int64_t per = (n_block + 7) / 8;
    return static_cast<int64_t>(blockIdx.x % 8) * per + blockIdx.x / 8;
}

// Wavefronts per workgroup: 12 (3 per SIMD), measured best at p = 12 for both
// stage precisions (FP32 against 8 and 16, FP64 against 8).  The sequential
// kernel uses 8.
template <typename T>
__host__ __device__ constexpr int pipe_waves()
{
    return 12;
}
constexpr int kSeqWaves = 8;

template <typename T, int P, int NC>
__global__ __launch_bounds__(64 * pipe_waves<T>(), 1) void m2l_pipe_kernel(M2LArgs a, const double* __restrict__ M,
                                                                          double* __restrict__ L)
{
    constexpr int NW = pipe_waves<T>();
    constexpr int NTH = 64 * NW;
    constexpr int K = (P + 1) * (P + 1);
    constexpr int TB = 16 / NC;
    constexpr int BS = (K + kPadRows + 1) * kLdsW;
    constexpr int NCST = kCsSlots * 2 * (P + 1) + 64;
    __shared__ T buf[pipe_fits<T, P, NC>() ? 6 * BS : 1];
    __shared__ alignas(16) T cst[NCST];
    __shared__ int ent_w[kMaxEntries];
    __shared__ int ent_s[kMaxEntries * TB];
    if constexpr (pipe_fits<T, P, NC>()) {
        const int64_t blk = xcd_block(a.n_block);
        if (blk >= a.n_block) return;
        const int64_t e0 = a.blk_row[blk];
        const int ne = static_cast<int>(a.blk_row[blk + 1] - e0);  // <= kMaxEntries (host-checked)
        if (ne == 0) {
            m2l_zero<P, NC>(a, blk, L);
            return;
        }
        for (int i = threadIdx.x; i < 6 * BS; i += NTH) buf[i] = T(0);
        for (int i = threadIdx.x; i < NCST; i += NTH) cst[i] = T(0);
        for (int i = threadIdx.x; i < ne; i += NTH) ent_w[i] = a.ent_meta[e0 + i];
        for (int i = threadIdx.x; i < ne * TB; i += NTH) ent_s[i] = a.ent_src[e0 * TB + i];
        __syncthreads();
        for_wave<0, NW>(threadIdx.x >> 6, [&](auto wc) {
            m2l_pipe_wave<T, P, NC, 2 * NW + 1, decltype(wc)::value>(a, blk, ne, M, L, buf, cst, ent_w, ent_s);
        });
    }
}

// ---------------------------------------------------------------- sequential kernel
// For orders whose six pipeline buffers exceed LDS (FP64, p > 12): one entry
// at a time, four barriers per entry, loads issued one entry ahead.
template <typename T, int P, int NC, int SC, int W>
__device__ __forceinline__ void m2l_seq_wave(const M2LArgs& a, int64_t blk, int ne, const double* __restrict__ M,
                                             double* __restrict__ L, T* Xs, T* Ys, T* cst, const int* ent_w,
                                             const int* ent_s)
{
    using WS = WaveState<T, P, NC, SC, W>;
    constexpr int TB = 16 / NC;
    constexpr int NCS = 2 * (P + 1);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    Acc<WS::NJ, 1> acc;
    acc.zero();
    WS ws;
    ws.init(tid);
    const T* Msrc = m2l_source<T>(a, M);
    Cols<WS::GI, T> cols;
    ws.load_cols(cols, a, Msrc, ent_s);
    double cs_next = 0.0;
    if constexpr (W == 0) {
        const double v = load_cs<P>(a, ent_w[0], lane);
        cs_next = load_cs<P>(a, ent_w[ne > 1 ? 1 : 0], lane);
        if (lane < NCS) cst[lane] = static_cast<T>(v);
    }
    {
        const int w = ent_w[0];
        const int cls = __builtin_amdgcn_readfirstlane(ent_class(w));
        const int d2 = __builtin_amdgcn_readfirstlane(ent_dist(w));
        ws.load_fwd(a, cls, lane);
        ws.load_cx(a, d2, lane);
        ws.load_back(a, cls, lane);
    }
    __syncthreads();
    for (int e = 0; e < ne; ++e) {
        // azimuth table: slot e & 1 holds entry e, slot (e + 1) & 1 receives entry e + 1
        const T* cs = cst + (e & 1) * NCS;
        ws.store_cols(Xs, cols, cs);
        if constexpr (W == 0) {
            cst[lane < NCS ? ((e + 1) & 1) * NCS + lane : 2 * NCS + lane] = static_cast<T>(cs_next);
            cs_next = load_cs<P>(a, ent_w[e + 2 < ne ? e + 2 : ne - 1], lane);
        }
        ws.use_loaded();
        const int en = e + 1 < ne ? e + 1 : e;  // the last entry reloads itself, harmlessly
        ws.load_cols(cols, a, Msrc, ent_s + en * TB);
        const int wn = ent_w[en], we = ent_w[e];
        const int cn = __builtin_amdgcn_readfirstlane(ent_class(wn));
        const int dn = __builtin_amdgcn_readfirstlane(ent_dist(wn));
        const int c_e = __builtin_amdgcn_readfirstlane(ent_class(we));
        const int d_e = __builtin_amdgcn_readfirstlane(ent_dist(we));
        if (cn != c_e) {
            ws.load_fwd(a, cn, lane);
            ws.load_back(a, cn, lane);
        }
        if (dn != d_e) ws.load_cx(a, dn, lane);
        __syncthreads();
        StageIn<T, P, SC, W> in;
        read_rot_fwd<T, P, SC, W>(in.b1, Xs, lane);
        stage_rot_fwd<T, P, SC, W>(ws.a1, in.b1, Ys, lane);
        __syncthreads();
        read_coax<T, P, SC, W>(in.b2, Ys, lane);
        stage_coax<T, P, SC, W>(ws.a2, in.b2, Xs, lane);
        __syncthreads();
        read_rot_back<T, P, SC, W>(in.b3, Xs, lane);
        stage_rot_back<T, P, SC, W>(ws.a3, in.b3, cs, acc, lane);
        __syncthreads();
    }
    write_locals<T, P, NC, SC, W>(a, blk, L, acc, lane);
}

template <typename T, int P, int NC>
__global__ __launch_bounds__(64 * kSeqWaves, 1) void m2l_seq_kernel(M2LArgs a, const double* __restrict__ M,
                                                                   double* __restrict__ L)
{
    constexpr int NTH = 64 * kSeqWaves;
    constexpr int K = (P + 1) * (P + 1);
    constexpr int TB = 16 / NC;
    __shared__ T Xs[(K + kPadRows + 1) * kLdsW];
    __shared__ T Ys[(K + kPadRows + 1) * kLdsW];
    __shared__ alignas(16) T cst[4 * (P + 1) + 64];
    __shared__ int ent_w[kMaxEntries];
    __shared__ int ent_s[kMaxEntries * TB];
    const int64_t blk = xcd_block(a.n_block);
    if (blk >= a.n_block) return;
    const int64_t e0 = a.blk_row[blk];
    const int ne = static_cast<int>(a.blk_row[blk + 1] - e0);  // <= kMaxEntries (host-checked)
    if (ne == 0) {
        m2l_zero<P, NC>(a, blk, L);
        return;
    }
    for (int i = threadIdx.x; i < kPadRows * kLdsW; i += NTH) {
        Xs[K * kLdsW + i] = T(0);
        Ys[K * kLdsW + i] = T(0);
    }
    for (int i = threadIdx.x; i < ne; i += NTH) ent_w[i] = a.ent_meta[e0 + i];
    for (int i = threadIdx.x; i < ne * TB; i += NTH) ent_s[i] = a.ent_src[e0 * TB + i];
    __syncthreads();
    for_wave<0, kSeqWaves>(threadIdx.x >> 6, [&](auto wc) {
        m2l_seq_wave<T, P, NC, 2 * kSeqWaves, decltype(wc)::value>(a, blk, ne, M, L, Xs, Ys, cst, ent_w, ent_s);
    });
}

template <typename T, int P, int NC>
void launch_one(const M2LArgs& a, const double* M, double* L, hipStream_t s)
{
    const unsigned grid = static_cast<unsigned>(((a.n_block + 7) / 8) * 8);
    if constexpr (pipe_fits<T, P, NC>())
        hipLaunchKernelGGL((m2l_pipe_kernel<T, P, NC>), dim3(grid), dim3(64 * pipe_waves<T>()), 0, s, a, M, L);
    else
        hipLaunchKernelGGL((m2l_seq_kernel<T, P, NC>), dim3(grid), dim3(64 * kSeqWaves), 0, s, a, M, L);
}

template <typename T, int P>
void launch_nc(const M2LArgs& a, int nc, const double* M, double* L, hipStream_t s)
{
    switch (nc) {
    case 1: launch_one<T, P, 1>(a, M, L, s); break;
    case 2: launch_one<T, P, 2>(a, M, L, s); break;
    case 3: launch_one<T, P, 3>(a, M, L, s); break;
    default: launch_one<T, P, 4>(a, M, L, s); break;
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
bool m2l_f32_top(const M2LArgs& a, int p, int nc, const double* M, double* L, hipStream_t s);

}  // namespace sfmm
