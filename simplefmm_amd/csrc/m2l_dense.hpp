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
//   * a workgroup takes a block of targets of one level and child octant (they
//     share their 189 candidate offsets); wave w holds component w % NCW of a
//     group of 32 targets as the 32 columns of 32x32x16 f16 MFMA tiles;
//   * the offsets present in the block are walked grouped by azimuth: the
//     sources' columns are rotated by e^{i m phi} in FP32 registers and scaled
//     per component (largest value 2^4) and by 2^11, Y = H + L with H = f16(Y)
//     and L = f16(Y - H); C += A_hi H + A_lo H + A_hi L runs on the f16 matrix
//     cores with FP32 accumulation (the 2^11 keeps the residual L normal down to
//     2^-18 of the component's largest value); D_c, scaled to the f16 range and
//     split the same way, is copied into LDS by the LDS DMA one offset ahead;
//   * at the end of an azimuth group C is rotated back by e^{-i m phi} and
//     added into the running FP32 sum S; each target's L~ is written once in
//     FP64 at the end (the first writer of L at levels >= 2).  (Rotating the
//     accumulators from group to group instead, without S, saves registers but
//     rounds each earlier contribution once per later group: 3x the error.)
// Orders 1..12 (the class operator of order 12 fills 66 KiB; two of them are
// double-buffered in LDS).
#pragma once

#if !defined(__HIPCC__) && !defined(__host__)
#define __host__
#define __device__
#endif

#include <cstdint>

namespace sfmm {

constexpr int kDenseMaxP = 12;
constexpr int kDenseWaves = 4;  // wavefronts per workgroup (one per SIMD)
constexpr int kDenseMaxEntries = 192;  // interaction offsets of a block (at most 189)
constexpr int kDenseMaxSlots = 64;

__host__ __device__ constexpr int dceil(int a, int b) { return (a + b - 1) / b; }

// Index order of the dense operator's rows and columns.  Coefficients are
// placed in "slots" of two: slot s holds element h = 0 at K index
// 16 (s / 8) + s % 8 and element h = 1 at 16 (s / 8) + 8 + s % 8, so that in
// the B operand of a 32x32x16 MFMA (lane l holds K = 8 (l / 32) + 0..7) the
// element j of lanes 0..31 and 32..63 share its order m wherever possible and
// the azimuth factor is one wave-uniform value.  The sin part (orders m >= 1)
// takes the first nss slots; the cos part uses the same slots for the same
// (n, m >= 1), then the orders m = 0.  Rows use the same order as columns.
struct DenseOrder {
    int p = 0;
    int nss = 0, ncs = 0;  // slots of the sin part, of the cos part
    int ksc = 0, kss = 0;  // 16-wide K-steps of the cos / sin part
    int kc = 0, ks = 0;    // K extents (rows and columns) of the cos / sin part
    int mtc = 0, mts = 0;  // 32-row tiles of the cos / sin part
    int kd = 0;            // floats per (node, component) in the FP32 copy: kc + ks
    int nfrag = 0;         // 32x16 A fragments per class
    int chunks = 0;        // 1 KiB copy chunks per class (hi and lo of each fragment), padded so
                           // that every wavefront of the workgroup copies the same number
    int el[2][kDenseMaxSlots] = {};  // element (h, slot): 32 n + m, -1 = padding
};

__host__ __device__ constexpr DenseOrder dense_order(int p)
{
    DenseOrder o{};
    o.p = p;
    for (int h = 0; h < 2; ++h)
        for (int s = 0; s < kDenseMaxSlots; ++s) o.el[h][s] = -1;
    int s = 0;
    int left[kDenseMaxP + 1] = {};
    int nleft = 0;
    for (int m = 1; m <= p; ++m) {
        int n = m;
        for (; n + 1 <= p; n += 2) {
            o.el[0][s] = 32 * n + m;
            o.el[1][s] = 32 * (n + 1) + m;
            ++s;
        }
        if (n == p) left[nleft++] = 32 * p + m;
    }
    for (int i = 0; i < nleft; i += 2) {
        o.el[0][s] = left[i];
        o.el[1][s] = i + 1 < nleft ? left[i + 1] : -1;
        ++s;
    }
    o.nss = s;
    for (int n = 0; n <= p; n += 2) {
        o.el[0][s] = 32 * n;
        o.el[1][s] = n + 1 <= p ? 32 * (n + 1) : -1;
        ++s;
    }
    o.ncs = s;
    o.ksc = dceil(o.ncs, 8);
    o.kss = dceil(o.nss, 8);
    o.kc = 16 * o.ksc;
    o.ks = 16 * o.kss;
    o.mtc = dceil(o.kc, 32);
    o.mts = dceil(o.ks, 32);
    o.kd = o.kc + o.ks;
    o.nfrag = o.mtc * o.ksc + o.mts * o.kss;
    o.chunks = dceil(2 * o.nfrag, kDenseWaves) * kDenseWaves;
    return o;
}

// K index (and row index) of element h of slot s, and back.
__host__ __device__ constexpr int dense_k(int s, int h) { return 16 * (s / 8) + 8 * h + s % 8; }
__host__ __device__ constexpr int dense_slot(int k) { return 8 * (k / 16) + k % 8; }
__host__ __device__ constexpr int dense_half(int k) { return (k / 8) % 2; }

// Element (32 n + m) at K index k of the cos (part 0) or sin (part 1) part, -1 if padding.
__host__ __device__ constexpr int dense_elem(const DenseOrder& o, int part, int k)
{
    const int s = dense_slot(k);
    if (part == 1 && s >= o.nss) return -1;
    if (s >= o.ncs) return -1;
    return o.el[dense_half(k)][s];
}

// Position in the FP32 copy ([cos kc][sin ks] per node and component) of the
// real-form coefficient r (degree n at [n^2, (n+1)^2): Re c(n,0), Re c(n,1),
// Im c(n,1), ...).
__host__ __device__ constexpr int dense_pos(const DenseOrder& o, int r)
{
    int n = 0;
    while ((n + 1) * (n + 1) <= r) ++n;
    const int q = r - n * n;
    const int m = (q + 1) / 2;
    const bool sin_part = q > 0 && q % 2 == 0;
    const int want = 32 * n + m;
    const int s0 = (m == 0) ? o.nss : 0, s1 = (m == 0) ? o.ncs : o.nss;
    for (int s = s0; s < s1; ++s)
        for (int h = 0; h < 2; ++h)
            if (o.el[h][s] == want) return sin_part ? o.kc + dense_k(s, h) : dense_k(s, h);
    return -1;
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
typedef float df16 __attribute__((ext_vector_type(16)));
typedef float df8 __attribute__((ext_vector_type(8)));

template <int I, int N, class F>
__device__ __forceinline__ void dense_for(F&& f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        dense_for<I + 1, N>(f);
    }
}

// Order m of element h of slot s (padding takes the order of the slot's other
// element, so that a slot with one coefficient needs no per-lane choice).
__host__ __device__ constexpr int dense_m(const DenseOrder& o, int s, int h)
{
    if (s >= o.ncs) return 0;
    const int e = o.el[h][s] >= 0 ? o.el[h][s] : o.el[1 - h][s];
    return e >= 0 ? e % 32 : 0;
}
// Order m of cos row k, or of row k2 where k is padding (rows use the column order).
__host__ __device__ constexpr int dense_row_m(const DenseOrder& o, int k, int k2)
{
    const int e = dense_elem(o, 0, k) >= 0 ? dense_elem(o, 0, k) : dense_elem(o, 0, k2);
    return e < 0 ? 0 : e % 32;
}
// Real-form index (degree n at [n^2, (n+1)^2)) of row k of a part, -1 for padding.
__host__ __device__ constexpr int dense_row_rf(const DenseOrder& o, int part, int k)
{
    const int e = dense_elem(o, part, k);
    if (e < 0) return -1;
    const int n = e / 32, m = e % 32;
    if (part == 1) return n * n + 2 * m;
    return m == 0 ? n * n : n * n + 2 * m - 1;
}

// x = hi + lo in f16 to 22 bits (for |x| >= 2^-3; below, to 2^-25 absolute).
__device__ __forceinline__ void dense_split(float x, _Float16& hi, _Float16& lo)
{
    hi = static_cast<_Float16>(x);
    lo = static_cast<_Float16>(__builtin_fmaf(static_cast<float>(hi), -1.0f, x));
}

// C += A_hi B_hi + A_lo B_hi + A_hi B_lo (f16 MFMA, FP32 accumulation; the
// small products first).
__device__ __forceinline__ void dense_mma(const dh8& ahi, const dh8& alo, const dh8& bhi, const dh8& blo, df16& c)
{
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, blo, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bhi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bhi, c, 0, 0, 0);
}

// One 16-byte chunk per lane from global memory into LDS at dst + 16 * lane
// (LDS DMA; dst is wave-uniform).
__device__ __forceinline__ void dense_lds_copy(const dh8* src, dh8* dst)
{
    __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
}
// Waits until at most N vector memory loads of the wave are in flight (LDS DMA
// included): with the N loads issued after an operator copy, the copy is done.
template <int N>
__device__ __forceinline__ void dense_wait_loads()
{
    static_assert(N >= 0 && N <= 63, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}
// Keeps the compiler from moving memory operations across this point.
__device__ __forceinline__ void dense_fence() { asm volatile("" ::: "memory"); }

// The order of P as a static member (usable in constant expressions inside the
// kernel's lambdas without capturing a local).
template <int P>
struct DenseCfg {
    static constexpr DenseOrder O = dense_order(P);
};

// Adds the accumulators of one azimuth group, rotated back by e^{-i m phi}
// ((c, s) = az[m]), into the running sums and clears them:
// (sc + i ss) += (lc + i ls) e^{-i m phi}.
template <int P, int MTC, int MTS>
__device__ __forceinline__ void dense_flush(df16* cc, df16* cs, df16* sumc, df16* sums, const float* az, int h)
{
    using Cfg = DenseCfg<P>;
    dense_for<0, MTC>([&](auto mc) {
        constexpr int mt = decltype(mc)::value;
        dense_for<0, 16>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            constexpr int k0 = 32 * mt + 8 * (r / 4) + r % 4;
            constexpr int m0 = dense_row_m(Cfg::O, k0, k0 + 4), m1 = dense_row_m(Cfg::O, k0 + 4, k0);
            const float lc = cc[mt][r];
            if constexpr (mt >= MTS || (m0 == 0 && m1 == 0)) {
                sumc[mt][r] += lc;  // order 0 (its sin row is padding)
                if constexpr (mt < MTS) sums[mt][r] += cs[mt][r];
            } else {
                float c, sn;
                if constexpr (m0 == m1) {
                    c = az[2 * m0];
                    sn = az[2 * m0 + 1];
                } else {
                    c = h ? az[2 * m1] : az[2 * m0];
                    sn = h ? az[2 * m1 + 1] : az[2 * m0 + 1];
                }
                const float ls = cs[mt][r];
                sumc[mt][r] += lc * c + ls * sn;
                sums[mt][r] += ls * c - lc * sn;
            }
        });
        cc[mt] = df16{};
        if constexpr (mt < MTS) cs[mt] = df16{};
    });
}

template <int P, int NC>
__global__ __launch_bounds__(64 * kDenseWaves, 1) void m2l_dense_kernel(DenseArgs a, const float* __restrict__ M32,
                                                                       double* __restrict__ L)
{
    using Cfg = DenseCfg<P>;
    constexpr int K = (P + 1) * (P + 1);
    constexpr int NW = kDenseWaves;
    constexpr int NCW = NC == 3 ? 4 : NC;  // waves per target group (NC = 3: the fourth only copies)
    constexpr int G = NW / NCW;            // target groups per block
    constexpr int TB = 32 * G;             // targets per block
    constexpr int CH = Cfg::O.chunks;  // LDS copy chunks of one class (64 x 16 bytes each)
    constexpr int KSC = Cfg::O.ksc, KSS = Cfg::O.kss, MTC = Cfg::O.mtc, MTS = Cfg::O.mts, KC = Cfg::O.kc,
                  KD = Cfg::O.kd;
    static_assert(MTS <= MTC, "sin rows fit the cos tiles");
    // the operator of the current entry and of the next one, each [frag][hi 64 | lo 64],
    // the azimuth factors and the block's entry descriptors
    __shared__ dh8 opbuf0[CH * 64];
    __shared__ dh8 opbuf1[CH * 64];
    __shared__ float saz[kAzGroups * (P + 1) * 2];
    __shared__ int32_t smeta[kDenseMaxEntries];

    const int64_t per = (a.n_block + 7) / 8;
    const int64_t blk = static_cast<int64_t>(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (blk >= a.n_block) return;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, h = lane >> 5;
    const int comp = w % NCW, grp = w / NCW;
    const bool active = comp < NC;
    const int ti = 32 * grp + (lane & 31);
    const int32_t tg = a.blk_tgt[blk * TB + ti];
    const int64_t e0 = a.blk_row[blk];
    const int ne = static_cast<int>(a.blk_row[blk + 1] - e0);

    // scale of this wavefront's columns (its component of 32 neighbouring
    // targets): the largest |value| over their sources maps below 2^15, so
    // rotated values stay below 2^15.5 < 65504 and keep 22 bits down to 2^-18
    // of the largest; one value per wavefront keeps the azimuth factors scalar
    float mx = 0.0f;
    if (active && tg >= 0) {
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
    // a wave-uniform float held in a scalar register
    auto uniform = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };

    // copies the operator of class cls into LDS buffer b: CH / NW chunks of 1 KiB
    // per wavefront, a fixed count so that the wait below can count them
    auto dma = [&](int cls, dh8* __restrict__ buf) {
        const dh8* src = static_cast<const dh8*>(a.ops) + static_cast<int64_t>(cls) * CH * 64;
#pragma unroll
        for (int c = 0; c < CH / NW; ++c)
            dense_lds_copy(src + (w + NW * c) * 64 + lane, buf + (w + NW * c) * 64);
    };
    for (int k = tid; k < kAzGroups * (P + 1) * 2; k += 64 * NW) saz[k] = a.az[k];
    for (int k = tid; k < ne; k += 64 * NW) smeta[k] = a.ent_meta[e0 + k];
    auto src_ptr = [&](int32_t s) -> const float* {
        const float* real = M32 + (static_cast<int64_t>(s < 0 ? 0 : s) * NC + comp) * KD;
        return (active && s >= 0) ? real : a.zero;
    };

    // accumulators of the current azimuth group, and the running sums (unrotated frame)
    df16 cc[MTC], cs[MTS], sumc[MTC], sums[MTS];
    for (int i = 0; i < MTC; ++i) cc[i] = sumc[i] = df16{};
    for (int i = 0; i < MTS; ++i) cs[i] = sums[i] = df16{};
    // the sources' columns in flight: LA K-steps ahead (K-step ks in slot ks % LA;
    // half an entry ahead at order 12, to hold the registers)
    constexpr int LA = (KSC % 2 == 0 && KSC >= 6) ? KSC / 2 : KSC;
    df8 xr[LA], xi[LA];
    const float* csp = a.zero;  // this lane's source of the current entry
    int32_t src1 = -1;          // and of the next one (loaded an entry ahead)

    if (ne > 0) {
        dma(a.ent_meta[e0] & 0xff, opbuf0);
        csp = src_ptr(a.ent_src[e0 * TB + ti]);
        src1 = ne > 1 ? a.ent_src[(e0 + 1) * TB + ti] : -1;
        dense_for<0, LA>([&](auto kc) {
            constexpr int ks = decltype(kc)::value;
            xr[ks] = *reinterpret_cast<const df8*>(csp + 16 * ks + 8 * h);
            if constexpr (ks < KSS) xi[ks] = *reinterpret_cast<const df8*>(csp + KC + 16 * ks + 8 * h);
        });
    }
    float csx[P + 1], csy[P + 1];  // cos, sin of m phi of the current azimuth group, times sc
    __syncthreads();  // saz, smeta
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
            const int64_t e = e0 + i;
            const bool more = i + 1 < ne;
            const int nmeta = more ? smeta[i + 1] : -1;
            // the operator of entry i is in LDS and every wave is done with entry i - 1:
            // after its copy (issued at the top of entry i - 1) the wave issued the
            // source index and the refills, 2 x 16 bytes per K-step and part
            if (i == 0) dense_wait_loads<0>();  // the prologue's copy
            else dense_wait_loads<1 + 2 * (KSC + KSS)>();
            __syncthreads();
            if (more) dma(nmeta & 0xff, nxt);
            dense_fence();
            const float* nsp = src_ptr(src1);
            src1 = i + 2 < ne ? a.ent_src[(e + 2) * TB + ti] : -1;
            const dh8* A = cur;
            if (active) dense_for<0, KSC>([&](auto kc) {
                constexpr int ks = decltype(kc)::value;
                // B fragments of this K-step: rotate by e^{i m phi}, scale, split
                dh8 bch, bcl, bsh, bsl;
                dense_for<0, 8>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    constexpr int s = 8 * ks + j;
                    constexpr int m0 = dense_m(Cfg::O, s, 0), m1 = dense_m(Cfg::O, s, 1);
                    float cm, sm;
                    if constexpr (m0 == m1) {
                        cm = csx[m0];
                        sm = csy[m0];
                    } else {
                        cm = h ? csx[m1] : csx[m0];
                        sm = h ? csy[m1] : csy[m0];
                    }
                    const float xrj = xr[ks % LA][j];
                    const float xij = ks < KSS ? xi[ks % LA][j] : 0.0f;
                    const float vc = xrj * cm - xij * sm;
                    _Float16 hh, ll;
                    dense_split(vc, hh, ll);
                    bch[j] = hh;
                    bcl[j] = ll;
                    if constexpr (ks < KSS) {
                        const float vs = xrj * sm + xij * cm;
                        dense_split(vs, hh, ll);
                        bsh[j] = hh;
                        bsl[j] = ll;
                    }
                });
                // refill this slot: K-step ks + LA of this entry, or of the next one
                {
                    constexpr int kn = (ks + LA) % KSC;
                    const float* np = ks + LA < KSC ? csp : nsp;
                    if (ks + LA < KSC || more) {
                        xr[ks % LA] = *reinterpret_cast<const df8*>(np + 16 * kn + 8 * h);
                        if constexpr (kn < KSS) xi[ks % LA] = *reinterpret_cast<const df8*>(np + KC + 16 * kn + 8 * h);
                    }
                }
                dense_for<0, MTC>([&](auto mc) {
                    constexpr int mt = decltype(mc)::value;
                    constexpr int f = mt * KSC + ks;
                    const dh8 ahi = A[f * 128 + lane], alo = A[f * 128 + 64 + lane];
                    dense_mma(ahi, alo, bch, bcl, cc[mt]);
                });
                if constexpr (ks < KSS) {
                    dense_for<0, MTS>([&](auto mc) {
                        constexpr int mt = decltype(mc)::value;
                        constexpr int f = MTC * KSC + mt * KSS + ks;
                        const dh8 ahi = A[f * 128 + lane], alo = A[f * 128 + 64 + lane];
                        dense_mma(ahi, alo, bsh, bsl, cs[mt]);
                    });
                }
            });
            csp = nsp;
            return more ? (nmeta >> 8) & 0xff : -1;
        };
        do {
            gn = (i & 1) == 0 ? entry(opbuf0, opbuf1) : entry(opbuf1, opbuf0);
            ++i;
        } while (gn == g);
        // the group's contribution, rotated back, into the running sums
        if (active) dense_flush<P, MTC, MTS>(cc, cs, sumc, sums, saz + g * (P + 1) * 2, h);
        g = gn;
    }
    // L~ of the block's targets (first writer at its level), unscaled, in FP64
    if (!active || tg < 0) return;
    const double un = ldexp(1.0, -a.op_exp - sexp);
    double* Lt = L + (static_cast<int64_t>(tg) * NC + comp) * K;
    dense_for<0, MTC>([&](auto mc) {
        constexpr int mt = decltype(mc)::value;
        dense_for<0, 16>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            constexpr int k0 = 32 * mt + 8 * (r / 4) + r % 4;
            constexpr int rc0 = dense_row_rf(Cfg::O, 0, k0), rc1 = dense_row_rf(Cfg::O, 0, k0 + 4);
            const int rcv = h ? rc1 : rc0;
            if (rcv >= 0) Lt[rcv] = static_cast<double>(sumc[mt][r]) * un;
            if constexpr (mt < MTS) {
                constexpr int rs0 = dense_row_rf(Cfg::O, 1, k0), rs1 = dense_row_rf(Cfg::O, 1, k0 + 4);
                const int rsv = h ? rs1 : rs0;
                if (rsv >= 0) Lt[rsv] = static_cast<double>(sums[mt][r]) * un;
            }
        });
    });
}

template <int P>
void launch_m2l_dense_p(const DenseArgs& a, int nc, const float* M32, double* L, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(8 * ((a.n_block + 7) / 8))), blk(64 * kDenseWaves);
    switch (nc) {
    case 1: hipLaunchKernelGGL((m2l_dense_kernel<P, 1>), grid, blk, 0, s, a, M32, L); break;
    case 2: hipLaunchKernelGGL((m2l_dense_kernel<P, 2>), grid, blk, 0, s, a, M32, L); break;
    case 3: hipLaunchKernelGGL((m2l_dense_kernel<P, 3>), grid, blk, 0, s, a, M32, L); break;
    default: hipLaunchKernelGGL((m2l_dense_kernel<P, 4>), grid, blk, 0, s, a, M32, L); break;
    }
}

}  // namespace sfmm
#endif
