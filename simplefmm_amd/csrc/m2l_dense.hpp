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
//     FP64 at the end (the first writer of L at levels >= 2).
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
    lo = static_cast<_Float16>(x - static_cast<float>(hi));
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
// Waits for every vector memory load of the wave, LDS DMA included.
__device__ __forceinline__ void dense_wait_loads() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// The order of P as a static member (usable in constant expressions inside the
// kernel's lambdas without capturing a local).
template <int P>
struct DenseCfg {
    static constexpr DenseOrder O = dense_order(P);
};

template <int P, int NC>
__global__ __launch_bounds__(256, 1) void m2l_dense_kernel(DenseArgs a, const float* __restrict__ M32,
                                                            double* __restrict__ L)
{
    using Cfg = DenseCfg<P>;
    constexpr int K = (P + 1) * (P + 1);
    constexpr int NCW = NC == 3 ? 4 : NC;  // waves per target group (NC = 3: wave 3 only copies)
    constexpr int G = 4 / NCW;             // target groups per block
    constexpr int TB = 32 * G;             // targets per block
    constexpr int FR = Cfg::O.nfrag;
    constexpr int KSC = Cfg::O.ksc, KSS = Cfg::O.kss, MTC = Cfg::O.mtc, MTS = Cfg::O.mts, KC = Cfg::O.kc,
                  KD = Cfg::O.kd;
    // the operator of the current entry and of the next one, each [frag][hi 64 | lo 64]
    __shared__ dh8 opbuf[2 * FR * 128];

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

    // scale of this lane's column (target, component): the largest |value| over
    // the target's sources maps below 2^15, so rotated values stay below
    // 2^15.5 < 65504 and keep 22 bits down to 2^-18 of the largest
    int sexp = 0;
    if (active && tg >= 0) {
        float mx = 0.0f;
        for (int i = 0; i < ne; ++i) {
            const int32_t sn = a.ent_src[(e0 + i) * TB + ti];
            if (sn >= 0) mx = fmaxf(mx, a.nodemax[static_cast<int64_t>(sn) * NC + comp]);
        }
        if (mx > 0.0f) {
            int ex;
            (void)frexpf(mx, &ex);  // mx < 2^ex
            sexp = 15 - ex;
        }
    }
    const float sc = ldexpf(1.0f, sexp);

    // copies the operator of class cls into LDS buffer b (one 1 KiB chunk per wave instruction)
    auto dma = [&](int cls, int b) {
        const dh8* src = static_cast<const dh8*>(a.ops) + static_cast<int64_t>(cls) * FR * 128;
        for (int i = w; i < 2 * FR; i += 4)
            dense_lds_copy(src + i * 64 + lane, opbuf + b * FR * 128 + i * 64);
    };
    auto src_ptr = [&](int32_t s) -> const float* {
        return (active && s >= 0) ? M32 + (static_cast<int64_t>(s) * NC + comp) * KD : a.zero;
    };

    df16 cc[MTC], cs[MTS], sumc[MTC], sums[MTS];
    for (int i = 0; i < MTC; ++i) cc[i] = sumc[i] = df16{};
    for (int i = 0; i < MTS; ++i) cs[i] = sums[i] = df16{};
    df8 xr[KSC], xi[KSS];

    if (ne > 0) {
        const int meta0 = a.ent_meta[e0];
        dma(meta0 & 0xff, 0);
        const float* sp = src_ptr(a.ent_src[e0 * TB + ti]);
        dense_for<0, KSC>([&](auto kc) {
            constexpr int ks = decltype(kc)::value;
            xr[ks] = *reinterpret_cast<const df8*>(sp + 16 * ks + 8 * h);
            if constexpr (ks < KSS) xi[ks] = *reinterpret_cast<const df8*>(sp + KC + 16 * ks + 8 * h);
        });
    }
    int grp_cur = -1;
    float csx[P + 1], csy[P + 1];  // cos, sin of m phi of the current azimuth group, times sc
    for (int i = 0; i < ne; ++i) {
        const int64_t e = e0 + i;
        const int meta = a.ent_meta[e];
        const int g = (meta >> 8) & 0xff;
        const bool more = i + 1 < ne;
        const int nmeta = more ? a.ent_meta[e + 1] : -1;
        if (g != grp_cur) {
            grp_cur = g;
            for (int m = 0; m <= P; ++m) {
                csx[m] = a.az[(g * (P + 1) + m) * 2] * sc;
                csy[m] = a.az[(g * (P + 1) + m) * 2 + 1] * sc;
            }
        }
        // the operator of entry i is in LDS and every wave is done with entry i - 1
        dense_wait_loads();
        __syncthreads();
        if (more) dma(nmeta & 0xff, (i + 1) & 1);
        const float* nsp = src_ptr(more ? a.ent_src[(e + 1) * TB + ti] : -1);
        const dh8* A = opbuf + (i & 1) * FR * 128;
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
                const float xrj = xr[ks][j];
                const float xij = ks < KSS ? xi[ks < KSS ? ks : 0][j] : 0.0f;
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
            // the next entry's columns of this K-step
            if (more) {
                xr[ks] = *reinterpret_cast<const df8*>(nsp + 16 * ks + 8 * h);
                if constexpr (ks < KSS) xi[ks] = *reinterpret_cast<const df8*>(nsp + KC + 16 * ks + 8 * h);
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
        // end of an azimuth group: rotate back by e^{-i m phi} into the running sum
        if (!more || ((nmeta >> 8) & 0xff) != g) {
            dense_for<0, MTC>([&](auto mc) {
                constexpr int mt = decltype(mc)::value;
                dense_for<0, 16>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    constexpr int k0 = 32 * mt + 8 * (r / 4) + r % 4;
                    constexpr int m0 = dense_row_m(Cfg::O, k0, k0 + 4), m1 = dense_row_m(Cfg::O, k0 + 4, k0);
                    const float* az0 = a.az + (g * (P + 1) + m0) * 2;
                    const float* az1 = a.az + (g * (P + 1) + m1) * 2;
                    float cm, sm;
                    if constexpr (m0 == m1) {
                        cm = az0[0];
                        sm = az0[1];
                    } else {
                        cm = h ? az1[0] : az0[0];
                        sm = h ? az1[1] : az0[1];
                    }
                    const float oc = cc[mt][r];
                    if constexpr (mt < MTS) {
                        const float os = cs[mt][r];
                        sumc[mt][r] += oc * cm + os * sm;
                        sums[mt][r] += os * cm - oc * sm;
                    } else {
                        sumc[mt][r] += oc;  // rows of order 0 only
                    }
                });
                cc[mt] = df16{};
                if constexpr (mt < MTS) cs[mt] = df16{};
            });
        }
    }
    // L~ of the block's targets (first writer at its level)
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
    const dim3 grid(static_cast<unsigned>(8 * ((a.n_block + 7) / 8))), blk(256);
    switch (nc) {
    case 1: hipLaunchKernelGGL((m2l_dense_kernel<P, 1>), grid, blk, 0, s, a, M32, L); break;
    case 2: hipLaunchKernelGGL((m2l_dense_kernel<P, 2>), grid, blk, 0, s, a, M32, L); break;
    case 3: hipLaunchKernelGGL((m2l_dense_kernel<P, 3>), grid, blk, 0, s, a, M32, L); break;
    default: hipLaunchKernelGGL((m2l_dense_kernel<P, 4>), grid, blk, 0, s, a, M32, L); break;
    }
}

}  // namespace sfmm
#endif
