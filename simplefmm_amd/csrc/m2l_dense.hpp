// M2L of the mixed-precision path on the f16 matrix cores, with hi/lo split
// products (Ootomo & Yokota, "Recovering single precision accuracy from Tensor
// Cores while surpassing the FP32 theoretical peak performance", IJHPCA 2022).
//
// The rotate-translate-rotate operator of fmm.hpp:620-688 for an offset d with
// azimuth phi, polar class c and distance |d| factors as
//     L~ += Az(-phi) D_c Az(phi) M~,
// where Az(phi) multiplies each coefficient of order m by e^{i m phi} and D_c =
// R_c^T T_|d| R_c (polar rotation, coaxial translation, back rotation) depends on
// the polar class only (|d| is fixed by it).  In the real form D_c is block
// diagonal: a cos block on the Re parts (m >= 0) and a sin block on the Im parts
// (m >= 1).  This kernel stores D_c densely (70 classes, level independent) and
// applies it as one matrix product per interaction offset:
//   * a workgroup (8 wavefronts, two per SIMD, or 4) takes a block of targets of one
//     level and child octant (they share their 189 candidate offsets); wave w
//     holds component w % NCW of a group of 16 targets as the 16 columns of
//     16x16x32 f16 MFMA tiles;
//   * the offsets present in the block are walked grouped by azimuth: the
//     sources' columns are rotated by e^{i m phi} in FP32 registers and scaled
//     (the largest value over the wavefront's sources to below 2^15), split as
//     x = H + L with H = f16(x), L = f16(x - H), and C += A_hi H + A_lo H +
//     A_hi L runs on the f16 matrix cores with FP32 accumulation; D_c, scaled
//     to the f16 range and split the same way, is copied into LDS by the LDS
//     DMA one offset ahead (two buffers);
//   * at the end of an azimuth group C is rotated back by e^{-i m phi} and
//     added into the running FP32 sum S; each target's L~ is written once in
//     FP64 at the end (the first writer of L at levels >= 2).  (Rotating the
//     accumulators from group to group instead, without S, saves registers but
//     rounds each earlier contribution once per later group: 3x the error.)
// Orders 1..12 (the class operator of order 12 fills 66 KiB).
#pragma once

#if !defined(__HIPCC__) && !defined(__host__)
#define __host__
#define __device__
#endif

#include <cstdint>

namespace sfmm {

constexpr int kDenseMaxP = 12;
constexpr int kDenseWaves = 8;         // most wavefronts per workgroup (two per SIMD); 4 also runs
constexpr int kDenseMaxEntries = 192;  // interaction offsets of a block (at most 189)
constexpr int kDenseMaxSlots = 32;
constexpr int kDenseMaxRows = 112;

__host__ __device__ constexpr int dceil(int a, int b) { return (a + b - 1) / b; }

// Index orders of the dense operator.
//
// Columns (K): the B operand of a 16x16x32 MFMA gives lane l the K values
// 8 (l / 16) + 0..7 of column l % 16.  Coefficients are placed in "slots" of
// four: element q of slot s sits at K = 32 (s / 8) + 8 q + s % 8, so the four
// lane groups hold the four elements of a slot in the same register, and the
// azimuth factor e^{i m phi} of a slot is one wave-uniform value when its
// elements share the order m.  Slots take the orders m >= 1, four degrees at a
// time; the orders m = 0 (no azimuth factor) fill the free places of those
// slots (their sin places stay empty), then slots of their own.  The cos and
// sin parts share the K order, so lane and register of Re c(n,m) and Im c(n,m)
// coincide and the rotation is in-lane.
// Rows: the (n, m >= 1) pairs (cos and sin rows aligned), then the cos rows of
// order 0; padded to 16-row tiles.
struct DenseOrder {
    int p = 0;
    int ns = 0;           // slots
    int nks = 0;          // 32-wide K-steps (both parts)
    int kc = 0;           // K extent of each part: 32 nks
    int npair = 0;        // (n, m >= 1) coefficients
    int mtc = 0, mts = 0; // 16-row tiles of the cos / sin part
    int kd = 0;           // floats per (node, component) in the FP32 copy: cos kc, then sin kc
    int nfrag = 0;        // 16x32 A fragments per class
    int chunks = 0;       // 1 KiB copy chunks per class (hi and lo of each fragment), padded so
                          // that every wavefront of the workgroup copies the same number
    int el[4][kDenseMaxSlots] = {};  // element (32 n + m) at (q, slot), -1 = empty
    int row[kDenseMaxRows] = {};     // element of each row, -1 = padding
};

__host__ __device__ constexpr DenseOrder dense_order(int p)
{
    DenseOrder o{};
    o.p = p;
    for (int q = 0; q < 4; ++q)
        for (int s = 0; s < kDenseMaxSlots; ++s) o.el[q][s] = -1;
    for (int r = 0; r < kDenseMaxRows; ++r) o.row[r] = -1;
    int s = 0, nr = 0;
    for (int m = 1; m <= p; ++m)
        for (int n = m; n <= p; n += 4, ++s)
            for (int q = 0; q < 4 && n + q <= p; ++q) {
                o.el[q][s] = 32 * (n + q) + m;
                o.row[nr++] = 32 * (n + q) + m;
            }
    o.npair = nr;
    // orders 0 into the free places, then slots of their own
    int n0 = 0;
    for (int t = 0; t < s && n0 <= p; ++t)
        for (int q = 0; q < 4 && n0 <= p; ++q)
            if (o.el[q][t] < 0) o.el[q][t] = 32 * n0++;
    for (; n0 <= p; ++s)
        for (int q = 0; q < 4 && n0 <= p; ++q) o.el[q][s] = 32 * n0++;
    for (int n = 0; n <= p; ++n) o.row[nr++] = 32 * n;
    o.ns = s;
    o.nks = dceil(o.ns, 8);
    o.kc = 32 * o.nks;
    o.mtc = dceil(o.npair + p + 1, 16);
    o.mts = dceil(o.npair, 16);
    o.kd = 2 * o.kc;
    o.nfrag = (o.mtc + o.mts) * o.nks;
    o.chunks = dceil(2 * o.nfrag, kDenseWaves) * kDenseWaves;
    return o;
}

// K index of element q of slot s, and back.
__host__ __device__ constexpr int dense_kidx(int s, int q) { return 32 * (s / 8) + 8 * q + s % 8; }
__host__ __device__ constexpr int dense_kslot(int k) { return 8 * (k / 32) + k % 8; }
__host__ __device__ constexpr int dense_kq(int k) { return (k % 32) / 8; }

// Element (32 n + m) at K index k of the cos (part 0) or sin (part 1) part, -1 if empty.
__host__ __device__ constexpr int dense_elem(const DenseOrder& o, int part, int k)
{
    const int s = dense_kslot(k);
    if (k < 0 || k >= o.kc || s >= o.ns) return -1;
    const int e = o.el[dense_kq(k)][s];
    return (part == 1 && e % 32 == 0) ? -1 : e;
}
// Element of row r of a part, -1 for padding.
__host__ __device__ constexpr int dense_row_elem(const DenseOrder& o, int part, int r)
{
    if (r < 0 || r >= kDenseMaxRows) return -1;
    if (part == 1 && r >= o.npair) return -1;
    return o.row[r];
}

// Real-form index (degree n at [n^2, (n+1)^2): Re c(n,0), Re c(n,1), Im c(n,1),
// ...) of element e in part `part`.
__host__ __device__ constexpr int dense_rf(int e, int part)
{
    const int n = e / 32, m = e % 32;
    if (part == 1) return n * n + 2 * m;
    return m == 0 ? n * n : n * n + 2 * m - 1;
}

// Position in the FP32 copy ([cos kc][sin kc] per node and component) of the
// real-form coefficient r.
__host__ __device__ constexpr int dense_pos(const DenseOrder& o, int r)
{
    int n = 0;
    while ((n + 1) * (n + 1) <= r) ++n;
    const int qq = r - n * n;
    const int m = (qq + 1) / 2;
    const bool sin_part = qq > 0 && qq % 2 == 0;
    const int want = 32 * n + m;
    for (int s = 0; s < o.ns; ++s)
        for (int q = 0; q < 4; ++q)
            if (o.el[q][s] == want) return (sin_part ? o.kc : 0) + dense_kidx(s, q);
    return -1;
}

// The order shared by the elements of order m >= 1 of slot s (0 if it holds
// none), and the lane groups q whose element there is of order 0 (bit q):
// their azimuth factor is 1.
__host__ __device__ constexpr int dense_slot_m(const DenseOrder& o, int s)
{
    for (int q = 0; q < 4; ++q)
        if (s < o.ns && o.el[q][s] >= 0 && o.el[q][s] % 32 != 0) return o.el[q][s] % 32;
    return 0;
}
__host__ __device__ constexpr int dense_slot_zero(const DenseOrder& o, int s)
{
    int b = 0;
    for (int q = 0; q < 4; ++q)
        if (s < o.ns && o.el[q][s] >= 0 && o.el[q][s] % 32 == 0) b |= 1 << q;
    return b;
}

// Layout of the FP32 copy of the multipoles that P2M and M2M write beside M for
// the mixed-precision M2L: per (node, component) `row` floats, the real-form
// coefficient r at pos[r] (dense order for p <= 12, unpermuted above).
template <int P>
struct F32Layout {
    static constexpr bool dense = P <= kDenseMaxP;
    static constexpr int row = dense ? dense_order(P).kd : (P + 1) * (P + 1);
    int pos[(P + 1) * (P + 1)] = {};
    constexpr F32Layout()
    {
        const DenseOrder o = dense_order(dense ? P : 1);
        for (int r = 0; r < (P + 1) * (P + 1); ++r) pos[r] = dense ? dense_pos(o, r) : r;
    }
};

template <int P>
struct F32LayoutOf {
    static constexpr F32Layout<P> L{};
};

// Azimuth groups: interaction offsets (di, dj, dk) sharing the direction of
// (di, dj) (so the same e^{i m phi}); (0, 0) joins the group of (1, 0).
constexpr int kAzGroups = 32;

}  // namespace sfmm

// ------------------------------------------------------------------ device code
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hpp"

namespace sfmm {

typedef _Float16 dh8 __attribute__((ext_vector_type(8)));
typedef float df4 __attribute__((ext_vector_type(4)));
typedef float df8 __attribute__((ext_vector_type(8)));

template <int I, int N, class F>
__device__ __forceinline__ void dense_for(F&& f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        dense_for<I + 1, N>(f);
    }
}

// x = hi + lo in f16 to 22 bits (for |x| >= 2^-3; below, to 2^-25 absolute).
// neg1 = -1 held where the compiler cannot see it: lo = f16(hi * neg1 + x) then
// stays one mixed-precision FMA (f16 operand, f32 operand, f16 result) instead
// of a conversion, a subtraction and a conversion.
__device__ __forceinline__ void dense_split(float x, _Float16& hi, _Float16& lo, float neg1)
{
    hi = static_cast<_Float16>(x);
    lo = static_cast<_Float16>(__builtin_fmaf(static_cast<float>(hi), neg1, x));
}
// -1.0f in a scalar register, opaque to constant folding.
__device__ __forceinline__ float dense_neg1()
{
    float v;
    asm volatile("s_mov_b32 %0, -1.0" : "=s"(v));
    return v;
}

// C += A_hi B_hi + A_lo B_hi + A_hi B_lo (f16 MFMA, FP32 accumulation; the
// small products first).
__device__ __forceinline__ void dense_mma(const dh8& ahi, const dh8& alo, const dh8& bhi, const dh8& blo, df4& c)
{
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, blo, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, bhi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bhi, c, 0, 0, 0);
}

// One 16-byte chunk per lane from global memory into LDS at dst + 16 * lane
// (LDS DMA; dst is wave-uniform).
__device__ __forceinline__ void dense_lds_copy(const dh8* src, dh8* dst)
{
    __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
}
// Waits until at most N vector memory loads of the wave are in flight (LDS DMA
// included; they complete in issue order): with at least N loads issued after
// an operator copy, the copy is done.
template <int N>
__device__ __forceinline__ void dense_wait_loads()
{
    static_assert(N >= 0 && N <= 63, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}
// Keeps the compiler from moving memory operations across this point.
__device__ __forceinline__ void dense_fence() { asm volatile("" ::: "memory"); }

// Fragment of tile t (cos tiles, then sin tiles) in K-step ks.
template <int NKS>
__host__ __device__ constexpr int ks0_frag(int t, int ks) { return t * NKS + ks; }

// The order of P as a static member (usable in constant expressions inside the
// kernel's lambdas without capturing a local).
template <int P>
struct DenseCfg {
    static constexpr DenseOrder O = dense_order(P);
};

// Adds the accumulators of one azimuth group, rotated back by e^{-i m phi}
// ((c, s) = az[2 m], az[2 m + 1]; rowm: the order of each row), into the
// running sums and clears them: (sc + i ss) += (lc + i ls) e^{-i m phi}.
// Lane group q holds rows 16 mt + 4 q + 0..3 of each tile.
template <int MTC, int MTS>
__device__ __forceinline__ void dense_flush(df4* cc, df4* cs, df4* sumc, df4* sums, const float* az,
                                            const int8_t* rowm, int q)
{
    dense_for<0, MTC>([&](auto mc) {
        constexpr int mt = decltype(mc)::value;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float lc = cc[mt][r];
            if constexpr (mt < MTS) {
                const int m = rowm[16 * mt + 4 * q + r];
                const float c = az[2 * m], sn = az[2 * m + 1], ls = cs[mt][r];
                sumc[mt][r] += lc * c + ls * sn;
                sums[mt][r] += ls * c - lc * sn;
            } else {
                sumc[mt][r] += lc;  // rows of order 0 (and padding)
            }
        }
        cc[mt] = df4{};
        if constexpr (mt < MTS) cs[mt] = df4{};
    });
}

template <int P, int NC, int NW>
__global__ __launch_bounds__(64 * NW) void m2l_dense_kernel(DenseArgs a, const float* __restrict__ M32,
                                                           double* __restrict__ L)
{
    using Cfg = DenseCfg<P>;
    constexpr int K = (P + 1) * (P + 1);
    constexpr int NCW = NC == 3 ? 4 : NC;  // waves per target group (NC = 3: the fourth only copies)
    constexpr int G = NW / NCW;            // target groups per block
    constexpr int TB = 16 * G;             // targets per block
    constexpr int CH = Cfg::O.chunks;      // LDS copy chunks of one class (64 x 16 bytes each)
    constexpr int NKS = Cfg::O.nks, MTC = Cfg::O.mtc, MTS = Cfg::O.mts, KC = Cfg::O.kc, KD = Cfg::O.kd;
    static_assert(MTS <= MTC && 16 * MTC <= kDenseMaxRows, "row tiles");
    static_assert(CH % NW == 0, "copy chunks per wavefront");
    // the operator of the current entry and of the next one, each [frag][hi 64 | lo 64],
    // the azimuth factors, the block's entry descriptors, and per row its order
    // and real-form indices (cos, sin)
    __shared__ dh8 opbuf0[CH * 64];
    __shared__ dh8 opbuf1[CH * 64];
    __shared__ float saz[kAzGroups * (P + 1) * 2];
    __shared__ int32_t smeta[kDenseMaxEntries];
    __shared__ int8_t srow_m[16 * MTC];
    __shared__ int16_t srow_rf[2][16 * MTC];

    const int64_t per = (a.n_block + 7) / 8;
    const int64_t blk = static_cast<int64_t>(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (blk >= a.n_block) return;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, q = lane >> 4;
    const int comp = w % NCW, grp = w / NCW;
    const bool active = comp < NC;
    const int ti = 16 * grp + (lane & 15);
    const int32_t tg = a.blk_tgt[blk * TB + ti];
    const int64_t e0 = a.blk_row[blk];
    const int ne = static_cast<int>(a.blk_row[blk + 1] - e0);

    // scale of this wavefront's columns (its component of 16 neighbouring
    // targets): the largest |value| over their sources maps below 2^15, so
    // rotated values stay below 2^15.5 < 65504 and keep 22 bits down to 2^-18
    // of the largest; one value per wavefront keeps the azimuth factors scalar
    float mx = 0.0f;
    if (active && tg >= 0 && q == 0) {
        for (int i0 = 0; i0 < ne; i0 += 8) {
            int32_t sn[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) sn[u] = i0 + u < ne ? a.ent_src[(e0 + i0 + u) * TB + ti] : -1;
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (sn[u] >= 0) mx = fmaxf(mx, a.nodemax[static_cast<int64_t>(sn[u]) * NC + comp]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    int sexp = 0;
    if (mx > 0.0f) {
        int ex;
        (void)frexpf(mx, &ex);  // mx < 2^ex
        sexp = 15 - ex;
    }
    sexp = __builtin_amdgcn_readfirstlane(sexp);
    const float sc = ldexpf(1.0f, sexp);
    const float neg1 = dense_neg1();
    // a wave-uniform float held in a scalar register
    auto uniform = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };

    // copies the operator of class cls into an LDS buffer: CH / NW chunks of 1 KiB
    // per wavefront, a fixed count so that the wait below can count them
    auto dma = [&](int cls, dh8* __restrict__ buf) {
        const dh8* src = static_cast<const dh8*>(a.ops) + static_cast<int64_t>(cls) * CH * 64;
#pragma unroll
        for (int c = 0; c < CH / NW; ++c)
            dense_lds_copy(src + (w + NW * c) * 64 + lane, buf + (w + NW * c) * 64);
    };
    auto src_ptr = [&](int32_t s) -> const float* {
        const float* real = M32 + (static_cast<int64_t>(s < 0 ? 0 : s) * NC + comp) * KD;
        return (active && s >= 0) ? real : a.zero;
    };
    for (int k = tid; k < kAzGroups * (P + 1) * 2; k += 64 * NW) saz[k] = a.az[k];
    for (int k = tid; k < ne; k += 64 * NW) smeta[k] = a.ent_meta[e0 + k];
    for (int k = tid; k < 16 * MTC; k += 64 * NW) {
        const int ec = dense_row_elem(Cfg::O, 0, k), es = dense_row_elem(Cfg::O, 1, k);
        srow_m[k] = static_cast<int8_t>(ec < 0 ? 0 : ec % 32);
        srow_rf[0][k] = static_cast<int16_t>(ec < 0 ? -1 : dense_rf(ec, 0));
        srow_rf[1][k] = static_cast<int16_t>(es < 0 ? -1 : dense_rf(es, 1));
    }

    // accumulators of the current azimuth group, and the running sums (unrotated frame)
    df4 cc[MTC], cs[MTS], sumc[MTC], sums[MTS];
    for (int i = 0; i < MTC; ++i) cc[i] = sumc[i] = df4{};
    for (int i = 0; i < MTS; ++i) cs[i] = sums[i] = df4{};
    // the source columns of the next K-step, loaded one K-step ahead (lane group q
    // holds K = 8 q + 0..7 of each 32-wide K-step, cos then sin)
    df8 xr = df8{}, xi = df8{};
    const float* csp = a.zero;  // this lane's source of the current entry
    int32_t src1 = -1;          // and of the next one
    if (ne > 0) {
        dma(a.ent_meta[e0] & 0xff, opbuf0);
        csp = src_ptr(a.ent_src[e0 * TB + ti]);
        src1 = a.ent_src[(e0 + (ne > 1 ? 1 : 0)) * TB + ti];
        if (ne < 2) src1 = -1;
        xr = *reinterpret_cast<const df8*>(csp + 8 * q);
        xi = *reinterpret_cast<const df8*>(csp + KC + 8 * q);
    }
    __syncthreads();  // saz, smeta, srow_*
    float csx[P + 1], csy[P + 1];  // cos, sin of m phi of the current azimuth group, times sc
    int i = 0, g = ne > 0 ? (smeta[0] >> 8) & 0xff : -1;
    while (i < ne) {
        for (int m = 0; m <= P; ++m) {
            csx[m] = uniform(saz[(g * (P + 1) + m) * 2] * sc);
            csy[m] = uniform(saz[(g * (P + 1) + m) * 2 + 1] * sc);
        }
        int gn = -1;  // group of the entry after this group's run
        // one entry; even entries read opbuf0 and fill opbuf1, odd ones the reverse
        // (two arrays, so the reads are seen not to alias the copy in flight)
        auto entry = [&](const dh8* __restrict__ cur, dh8* __restrict__ nxt) {
            const bool more = i + 1 < ne;
            // the lane group, re-read each entry, so that the per-lane azimuth
            // factors of the slots shared by orders m and 0 are selected where they
            // are used instead of being held in registers across the loop (which
            // spilled the running sums)
            int qv = q;
            asm volatile("" : "+v"(qv));
            const int nmeta = more ? smeta[i + 1] : -1;
            // the operator of entry i is in LDS and every wave is done with entry i - 1.
            // After its copy (issued at the top of entry i - 1) an active wave issued a
            // source index and 4 x 16 bytes of source columns per K-step (loads complete
            // in order); an idle one issued nothing
            if (i == 0 || !active) dense_wait_loads<0>();
            else dense_wait_loads<4 * NKS>();
            __syncthreads();
            if (more) dma(nmeta & 0xff, nxt);
            dense_fence();
            const float* nsp = src_ptr(src1);
            src1 = i + 2 < ne ? a.ent_src[(e0 + i + 2) * TB + ti] : -1;
            if (active) {
                // A fragments of the 11-ish tiles of each K-step, one step ahead (the
                // fences keep the compiler from loading further ahead, which would
                // take the registers of the accumulators); tile t of K-step ks is
                // fragment t * NKS + ks (cos tiles, then sin tiles)
                constexpr int NT = MTC + MTS;
                dh8 ah = cur[ks0_frag<NKS>(0, 0) * 128 + lane], al = cur[ks0_frag<NKS>(0, 0) * 128 + 64 + lane];
                dense_for<0, NKS>([&](auto kc) {
                    constexpr int ks = decltype(kc)::value;
                    // nothing of this K-step moves into the previous one's MFMAs: its
                    // columns arrive during them
                    __builtin_amdgcn_sched_barrier(0);
                    // B fragments of this K-step: rotate by e^{i m phi}, scale, split
                    dh8 bch, bcl, bsh, bsl;
                    dense_for<0, 8>([&](auto jc) {
                        constexpr int j = decltype(jc)::value;
                        constexpr int s = 8 * ks + j;
                        constexpr int m = dense_slot_m(Cfg::O, s), z = dense_slot_zero(Cfg::O, s);
                        float cm = csx[m], sm = csy[m];
                        if constexpr (z != 0 && m != 0) {  // lane groups whose element is of order 0
                            const bool zq = (z >> qv) & 1;
                            cm = zq ? sc : cm;
                            sm = zq ? 0.0f : sm;
                        }
                        const float xrj = xr[j], xij = xi[j];
                        _Float16 hh, ll;
                        dense_split(xrj * cm - xij * sm, hh, ll, neg1);
                        bch[j] = hh;
                        bcl[j] = ll;
                        dense_split(xrj * sm + xij * cm, hh, ll, neg1);
                        bsh[j] = hh;
                        bsl[j] = ll;
                    });
                    // the columns of the next K-step, of this entry or of the next one (zeros
                    // after the last entry)
                    {
                        const float* np = ks + 1 < NKS ? csp + 32 * (ks + 1) : nsp;
                        xr = *reinterpret_cast<const df8*>(np + 8 * q);
                        xi = *reinterpret_cast<const df8*>(np + KC + 8 * q);
                    }
                    dense_for<0, NT>([&](auto tc) {
                        constexpr int t = decltype(tc)::value;
                        constexpr int tn = (t + 1) % NT, kn = t + 1 < NT ? ks : ks + 1;
                        dh8 nh, nl;
                        dense_fence();
                        if constexpr (kn < NKS) {
                            nh = cur[ks0_frag<NKS>(tn, kn) * 128 + lane];
                            nl = cur[ks0_frag<NKS>(tn, kn) * 128 + 64 + lane];
                        }
                        if constexpr (t < MTC) dense_mma(ah, al, bch, bcl, cc[t]);
                        else dense_mma(ah, al, bsh, bsl, cs[t - MTC]);
                        if constexpr (kn < NKS) {
                            ah = nh;
                            al = nl;
                        }
                    });
                });
            }
            csp = nsp;
            return more ? (nmeta >> 8) & 0xff : -1;
        };
        do {
            gn = (i & 1) == 0 ? entry(opbuf0, opbuf1) : entry(opbuf1, opbuf0);
            ++i;
        } while (gn == g);
        // the group's contribution, rotated back, into the running sums
        if (active) dense_flush<MTC, MTS>(cc, cs, sumc, sums, saz + g * (P + 1) * 2, srow_m, q);
        g = gn;
    }
    // L~ of the block's targets (first writer at its level), unscaled, in FP64
    if (!active || tg < 0) return;
    const double un = ldexp(1.0, -a.op_exp - sexp);
    double* Lt = L + (static_cast<int64_t>(tg) * NC + comp) * K;
    dense_for<0, MTC>([&](auto mc) {
        constexpr int mt = decltype(mc)::value;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * mt + 4 * q + r;
            const int rc = srow_rf[0][row];
            if (rc >= 0) Lt[rc] = static_cast<double>(sumc[mt][r]) * un;
            if constexpr (mt < MTS) {
                const int rs = srow_rf[1][row];
                if (rs >= 0) Lt[rs] = static_cast<double>(sums[mt][r]) * un;
            }
        }
    });
}

template <int P, int NW>
void launch_m2l_dense_pw(const DenseArgs& a, int nc, const float* M32, double* L, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(8 * ((a.n_block + 7) / 8))), blk(64 * NW);
    switch (nc) {
    case 1: hipLaunchKernelGGL((m2l_dense_kernel<P, 1, NW>), grid, blk, 0, s, a, M32, L); break;
    case 2: hipLaunchKernelGGL((m2l_dense_kernel<P, 2, NW>), grid, blk, 0, s, a, M32, L); break;
    case 3: hipLaunchKernelGGL((m2l_dense_kernel<P, 3, NW>), grid, blk, 0, s, a, M32, L); break;
    default: hipLaunchKernelGGL((m2l_dense_kernel<P, 4, NW>), grid, blk, 0, s, a, M32, L); break;
    }
}

// a.waves (4 or 8) wavefronts per workgroup: the tile shape, which fixes the
// targets per block the work lists were cut for (16 per wavefront group).
template <int P>
void launch_m2l_dense_p(const DenseArgs& a, int nc, const float* M32, double* L, hipStream_t s)
{
    if (a.waves == 4) launch_m2l_dense_pw<P, 4>(a, nc, M32, L, s);
    else launch_m2l_dense_pw<P, 8>(a, nc, M32, L, s);
}

}  // namespace sfmm
#endif
