// HIP kernels of the FMM evaluation for gfx950 (64-lane wavefronts).
//
// Expansions are stored per node as nc components x (p+1)^2 reals in the
// level-scaled real form described in tables.hpp.  Harmonics are generated on
// the fly from Cartesian recurrences; for u = (x, y, z) and r2 = |u|^2,
//     R(m,m)   = d_m (-x + i y) R(m-1,m-1),           R(0,0) = 1
//     R(n,m)   = a(n,m) z R(n-1,m) - b(n,m) r2 R(n-2,m)
// with d_m = sqrt((2m-1)/(2m)), a = (2n-1)/sqrt(n^2-m^2),
// b = sqrt((n-1)^2-m^2)/sqrt(n^2-m^2).  R(n,m) equals the reference's
// rho^n Z(n,m,theta) e^{-i m phi} (fmm.hpp:107-111,226-232,537-548).
#include "kernels.hpp"
#include "m2l_dense.hpp"

#include <algorithm>
#include <type_traits>

namespace sfmm {

namespace {

constexpr int kRecStride = 21;  // recurrence tables cover n, m <= 20
__device__ __forceinline__ double rec_a(const double* rec, int n, int m) { return rec[n * kRecStride + m]; }
__device__ __forceinline__ double rec_b(const double* rec, int n, int m)
{
    return rec[kRecStride * kRecStride + n * kRecStride + m];
}
__device__ __forceinline__ double rec_d(const double* rec, int m) { return rec[2 * kRecStride * kRecStride + m]; }

__host__ __device__ constexpr int rfc(int n, int m) { return m == 0 ? n * n : n * n + 2 * m - 1; }
__host__ __device__ constexpr int rfs(int n, int m) { return n * n + 2 * m; }
__host__ __device__ constexpr int tri(int n, int m) { return n * (n + 1) / 2 + m; }  // m >= 0 triangle

inline int grid_for(int64_t n, int b) { return static_cast<int>((n + b - 1) / b); }

// ------------------------------------------------------------------ charges
__global__ void gather_charges_kernel(int64_t n, int qdim, const int32_t* __restrict__ perm,
                                      const double* __restrict__ Q, double* __restrict__ q)
{
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t o = perm[s];
    for (int c = 0; c < qdim; ++c) q[c * n + s] = Q[o * qdim + c];
}

// The sorted outputs are interleaved, out[slot * nout + c], so each particle's
// values are one contiguous read; the caller's array is walked in its own
// order (coalesced read-modify-write) and the sorted side gathered through the
// inverse permutation.
__global__ void scatter_outputs_kernel(int64_t n, int nout, const int32_t* __restrict__ iperm,
                                       const double* __restrict__ out, double* __restrict__ U)
{
    const int64_t o = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const int64_t s = iperm[o];
    for (int c = 0; c < nout; ++c) U[o * nout + c] += out[s * nout + c];
}

__global__ void fill_kernel(double* x, int64_t count, double v)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) x[i] = v;
}

// ------------------------------------------------------------------ P2M
// One wavefront per leaf.  Lane (u, g) = (lane % NU, lane / NU) carries the
// azimuths ma = u and mb = P - u (one azimuth when they coincide) through the
// degree recurrence and sums q * R(n,m) for n = m..P over the particles g,
// g+G, ... of the leaf, so no cross-lane reduction is needed until the G
// partial sums are added.  Pairing m with P - m gives every lane the same P+2
// recurrence steps (one azimuth per lane would leave lane m idle for m of its
// P+1 steps).  The multipole components are then formed from these plain
// moments S_c:
//   potential:        M_c = S_c
//   potential+grad:   M_0 = S, M_d = -(1/w) D_d[S]
//   curl:             M = (1/w) curl-like combination D x S (fmm.hpp:549-552)
// where D_d is the ladder form of d/dx_d acting on degree n-1 moments:
//   (dx + i dy) R(n,m) = -sqrt((n+m)(n+m-1)) R(n-1,m-1)   (m >= 1)
//   (dx + i dy) R(n,0) =  sqrt(n(n-1)) conj R(n-1,1)
//   (dx - i dy) R(n,m) =  sqrt((n-m)(n-m-1)) R(n-1,m+1)
//    dz         R(n,m) =  sqrt((n-m)(n+m))   R(n-1,m)
template <int P, int NCQ>
__global__ __launch_bounds__(64) void p2m_kernel(DevPlan d, int mode, int nc, const double* __restrict__ q,
                                                 double* __restrict__ M)
{
    constexpr int K = (P + 1) * (P + 1);
    constexpr int NT = (P + 1) * (P + 2) / 2;  // complex coefficients with m >= 0
    constexpr int NU = P / 2 + 1;              // azimuth pairs
    constexpr int G = 64 / NU;                 // particle groups
    constexpr int NS = P + 2;                  // recurrence steps per lane
    constexpr int ncq = NCQ;
    __shared__ double part[G][NCQ][2 * NT];
    __shared__ double S[NCQ][2 * NT];

    const int64_t f = blockIdx.x;
    const int tid = threadIdx.x;
    const int32_t gnode = d.fl_gnode[f];
    const int64_t b = d.fl_begin[f], e = d.fl_end[f];
    const double invw = d.inv_width[d.fl_level[f]];
    const int u = tid % NU, g = tid / NU;
    const int ma = u, mb = P - u;
    const int la = P - ma + 1;             // steps of azimuth ma; then those of mb
    const int lb = mb > ma ? ma + 1 : 0;   // (none when the two coincide)
    const double* rec = d.rec;
    // per step: degree n, azimuth m, whether it is one of this lane's terms,
    // and the recurrence coefficients a(n, m), b(n, m) (0 at an azimuth's first step)
    auto step_nm = [&](int t, int& n, int& m) {
        if (t < la) {
            n = ma + t;
            m = ma;
            return true;
        }
        n = mb + (t - la);
        m = mb;
        return t - la < lb;
    };
    double ra[NS], rb[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        int n, m;
        const bool use = step_nm(t, n, m) && t != 0 && t != la;
        ra[t] = use ? rec_a(rec, n, m) : 0.0;
        rb[t] = use ? rec_b(rec, n, m) : 0.0;
    }

    for (int cq = 0; cq < ncq; ++cq) {
        double ar[NS], ai[NS];
#pragma unroll
        for (int t = 0; t < NS; ++t) ar[t] = ai[t] = 0.0;
        if (g < G) {
            for (int64_t i = b + g; i < e; i += G) {
                const double x = d.ux[i], y = d.uy[i], z = d.uz[i];
                const double qi = q[cq * d.n + i];
                const double r2 = x * x + y * y + z * z;
                // R(m,m) = d_m (-x + i y) R(m-1,m-1) up to mb, keeping R(ma,ma)
                double dr = 1.0, di = 0.0, sar = 1.0, sai = 0.0;
#pragma unroll
                for (int k = 1; k <= P; ++k) {
                    if (k <= mb) {
                        const double c = rec_d(rec, k);
                        const double tr = c * (-x * dr - y * di);
                        const double ti = c * (y * dr - x * di);
                        dr = tr;
                        di = ti;
                    }
                    if (k == ma) {
                        sar = dr;
                        sai = di;
                    }
                }
                // R(n+1,m) = a z R(n,m) - b r2 R(n-1,m); b(m+1,m) = 0, so the
                // value carried over from the previous azimuth drops out
                double p1r = sar, p1i = sai, p2r = 0.0, p2i = 0.0;
                ar[0] += qi * sar;
                ai[0] += qi * sai;
#pragma unroll
                for (int t = 1; t < NS; ++t) {
                    const double a = ra[t] * z, bb = rb[t] * r2;
                    double rr = a * p1r - bb * p2r;
                    double ri = a * p1i - bb * p2i;
                    if (t == la) {  // first step of azimuth mb
                        rr = dr;
                        ri = di;
                    }
                    p2r = p1r;
                    p2i = p1i;
                    p1r = rr;
                    p1i = ri;
                    ar[t] += qi * rr;
                    ai[t] += qi * ri;
                }
            }
#pragma unroll
            for (int t = 0; t < NS; ++t) {
                int n, m;
                if (step_nm(t, n, m)) {
                    part[g][cq][2 * tri(n, m)] = ar[t];
                    part[g][cq][2 * tri(n, m) + 1] = ai[t];
                }
            }
        }
    }
    __syncthreads();
    for (int v = tid; v < ncq * 2 * NT; v += 64) {
        const int cq = v / (2 * NT), k = v % (2 * NT);
        double s = 0.0;
        for (int gg = 0; gg < G; ++gg) s += part[gg][cq][k];
        S[cq][k] = s;
    }
    __syncthreads();

    double* Mg = M + static_cast<int64_t>(gnode) * nc * K;
    auto put = [&](int c, int r, double v) { Mg[c * K + r] = v; };
    for (int t = tid; t < NT; t += 64) {
        // decode (n, mm) of triangle index t
        int n = 0;
        while ((n + 1) * (n + 2) / 2 <= t) ++n;
        const int mm = t - n * (n + 1) / 2;
        if (mode == MODE_POT) {
            for (int c = 0; c < nc; ++c) {
                put(c, rfc(n, mm), S[c][2 * t]);
                if (mm > 0) put(c, rfs(n, mm), S[c][2 * t + 1]);
            }
            continue;
        }
        // ladder gradients of each moment component needed
        double gx[NCQ < 3 ? 3 : NCQ][2], gy[NCQ < 3 ? 3 : NCQ][2], gz[NCQ < 3 ? 3 : NCQ][2];
        for (int c = 0; c < ncq; ++c) {
            const double* Sc = S[c];
            double ppr = 0, ppi = 0, mmr = 0, mmi = 0, zr = 0, zi = 0;
            if (n >= 1) {
                if (mm >= 1) {
                    const double f = -sqrt(static_cast<double>((n + mm) * (n + mm - 1)));
                    ppr = f * Sc[2 * tri(n - 1, mm - 1)];
                    ppi = f * Sc[2 * tri(n - 1, mm - 1) + 1];
                } else if (n >= 2) {
                    const double f = sqrt(static_cast<double>(n * (n - 1)));
                    ppr = f * Sc[2 * tri(n - 1, 1)];
                    ppi = -f * Sc[2 * tri(n - 1, 1) + 1];
                }
                if (mm + 1 <= n - 1) {
                    const double f = sqrt(static_cast<double>((n - mm) * (n - mm - 1)));
                    mmr = f * Sc[2 * tri(n - 1, mm + 1)];
                    mmi = f * Sc[2 * tri(n - 1, mm + 1) + 1];
                }
                if (mm <= n - 1) {
                    const double f = sqrt(static_cast<double>((n - mm) * (n + mm)));
                    zr = f * Sc[2 * tri(n - 1, mm)];
                    zi = f * Sc[2 * tri(n - 1, mm) + 1];
                }
            }
            gx[c][0] = 0.5 * (ppr + mmr);
            gx[c][1] = 0.5 * (ppi + mmi);
            gy[c][0] = 0.5 * (ppi - mmi);
            gy[c][1] = -0.5 * (ppr - mmr);
            gz[c][0] = zr;
            gz[c][1] = zi;
        }
        double out[4][2];
        int nout = 0;
        if (mode == MODE_POTGRAD) {
            out[0][0] = S[0][2 * t];
            out[0][1] = S[0][2 * t + 1];
            for (int k = 0; k < 2; ++k) {
                out[1][k] = -invw * gx[0][k];
                out[2][k] = -invw * gy[0][k];
                out[3][k] = -invw * gz[0][k];
            }
            nout = 4;
        } else {  // MODE_CURL: (Q x grad R)_c summed = (D x S)
            for (int k = 0; k < 2; ++k) {
                out[0][k] = invw * (gz[1][k] - gy[2][k]);
                out[1][k] = invw * (gx[2][k] - gz[0][k]);
                out[2][k] = invw * (gy[0][k] - gx[1][k]);
            }
            nout = 3;
        }
        for (int c = 0; c < nout; ++c) {
            put(c, rfc(n, mm), out[c][0]);
            if (mm > 0) put(c, rfs(n, mm), out[c][1]);
        }
    }
}

// ------------------------------------------------------------------ L2P
// One thread per particle of the far-field leaves (slot_fleaf maps a sorted
// slot to its leaf): U_c += (1/w) sum_{n,m} Re(L~ conj R(n,m)), the m > 0
// terms doubled for the conjugate half (fmm.hpp:747-771).  The harmonics are
// advanced degree by degree with every azimuth's recurrence carried along, so
// that the leaf's coefficients are read row by row in memory order (each cache
// line in one stretch; azimuth-major order revisits every line P+1 times and
// is 1.5x slower); the factor 2 enters once, in R(1,1), and is exact.  The
// coefficients reach each lane through the vector L1, whose return rate bounds
// the kernel.  FP64 on both precision paths (FP32 harmonics here cost 3x in
// potential error for a few percent of the stage).
template <int P, int NC>
__global__ __launch_bounds__(256, P <= 12 ? 4 : 2) void l2p_kernel(DevPlan d, const double* __restrict__ L,
                                                                   double* __restrict__ out)
{
    constexpr int K = (P + 1) * (P + 1);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= d.n) return;
    const int32_t f = d.slot_fleaf[i];
    if (f < 0) return;  // near-field-only leaf (level < 2)
    const double invw = d.inv_width[d.fl_level[f]];
    const double* Lg = L + static_cast<int64_t>(d.fl_gnode[f]) * NC * K;
    const double* rec = d.rec;
    const double x = d.ux[i], y = d.uy[i], z = d.uz[i];
    const double r2 = x * x + y * y + z * z;
    // R(n-1, m) and R(n-2, m) of the degree reached, per azimuth m
    double cr[P + 1], ci[P + 1], pr[P + 1], pi[P + 1];
#pragma unroll
    for (int m = 0; m <= P; ++m) cr[m] = ci[m] = pr[m] = pi[m] = 0.0;
    cr[0] = 1.0;
    double U[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) U[c] = Lg[c * K];
#pragma unroll
    for (int n = 1; n <= P; ++n) {
        const double tr = cr[n - 1], ti = ci[n - 1];
#pragma unroll
        for (int m = 0; m < n; ++m) {
            const double a = rec_a(rec, n, m) * z, bb = rec_b(rec, n, m) * r2;
            const double nr = a * cr[m] - bb * pr[m];
            const double ni = a * ci[m] - bb * pi[m];
            pr[m] = cr[m];
            pi[m] = ci[m];
            cr[m] = nr;
            ci[m] = ni;
        }
        const double dd = (n == 1 ? 2.0 : 1.0) * rec_d(rec, n);
        cr[n] = dd * (-x * tr - y * ti);
        ci[n] = dd * (y * tr - x * ti);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
#pragma unroll
            for (int m = 0; m <= n; ++m) {
                U[c] = fma(Lg[c * K + rfc(n, m)], cr[m], U[c]);
                if (m > 0) U[c] = fma(Lg[c * K + rfs(n, m)], ci[m], U[c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) out[i * NC + c] += invw * U[c];
}

// ------------------------------------------------------------------ M2M / L2L
// One workgroup per node, thread = output row; dense level-independent
// real-form operators stored transposed ([oct][col][row]).
__global__ __launch_bounds__(256) void m2m_kernel(DevPlan d, int64_t first, int nc, double* __restrict__ M)
{
    extern __shared__ double X[];  // nc * K
    const int K = d.K;
    const int64_t pg = first + blockIdx.x;
    const int32_t c0 = d.child0[pg];
    const int nch = d.nchild[pg];
    if (nch == 0) return;  // leaves keep their P2M multipole
    double acc[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    for (int ch = 0; ch < nch; ++ch) {
        const int64_t cg = c0 + ch;
        const double* T = d.m2mT + static_cast<int64_t>(d.octant[cg]) * K * K;
        __syncthreads();
        for (int v = threadIdx.x; v < nc * K; v += 256) X[v] = M[cg * nc * K + v];
        __syncthreads();
        for (int rr = 0; rr < 2; ++rr) {
            const int r = threadIdx.x + 256 * rr;
            if (r >= K) break;
            int j = 0;
            while ((j + 1) * (j + 1) <= r) ++j;
            const int cols = (j + 1) * (j + 1);
            for (int col = 0; col < cols; ++col) {
                const double t = T[static_cast<int64_t>(col) * K + r];
                for (int c = 0; c < nc; ++c) acc[rr][c] += t * X[c * K + col];
            }
        }
    }
    for (int rr = 0; rr < 2; ++rr) {
        const int r = threadIdx.x + 256 * rr;
        if (r >= K) break;
        for (int c = 0; c < nc; ++c) M[pg * nc * K + c * K + r] = acc[rr][c];
    }
}

__global__ __launch_bounds__(256) void l2l_kernel(DevPlan d, int64_t first, int nc, double* __restrict__ L)
{
    extern __shared__ double X[];
    const int K = d.K;
    const int64_t cg = first + blockIdx.x;
    const int64_t pg = d.parent[cg];
    const double* T = d.l2lT + static_cast<int64_t>(d.octant[cg]) * K * K;
    for (int v = threadIdx.x; v < nc * K; v += 256) X[v] = L[pg * nc * K + v];
    __syncthreads();
    for (int rr = 0; rr < 2; ++rr) {
        const int r = threadIdx.x + 256 * rr;
        if (r >= K) break;
        int j = 0;
        while ((j + 1) * (j + 1) <= r) ++j;
        double acc[4] = {0, 0, 0, 0};
        for (int col = j * j; col < K; ++col) {
            const double t = T[static_cast<int64_t>(col) * K + r];
            for (int c = 0; c < nc; ++c) acc[c] += t * X[c * K + col];
        }
        for (int c = 0; c < nc; ++c) L[cg * nc * K + c * K + r] += acc[c];
    }
}

// ------------------------------------------------------------------ M2L
// One workgroup per target node; every interaction is rotated to the z axis,
// translated along it and rotated back (fmm.hpp:620-688), in real form:
//   1. azimuth: X = e^{i m phi} M~src           (2x2 per order)
//   2. polar rotation per degree: cos block (n+1)^2, sin block n^2 -> Y (order-major)
//   3. coaxial translation per order k: (P+1-k)^2 block             -> Z
//   4. back rotation per degree, azimuth e^{-i m phi}, accumulate into L~
template <int P>
__global__ __launch_bounds__(256) void m2l_kernel(DevPlan d, int nc, const double* __restrict__ M,
                                                  double* __restrict__ L)
{
    constexpr int K = (P + 1) * (P + 1);
    constexpr int NT = (P + 1) * (P + 2) / 2;
    __shared__ double X[4][K];
    __shared__ double Y[4][K];
    const int tid = threadIdx.x;
    const int64_t ti = blockIdx.x;
    const int32_t tg = d.tgt_gnode[ti];

    // stage-4 ownership: (j, a) pair, a = 0..j
    int jj = 0;
    while ((jj + 1) * (jj + 2) / 2 <= tid) ++jj;
    const int aa = tid - jj * (jj + 1) / 2;
    double accr[4] = {0, 0, 0, 0}, acci[4] = {0, 0, 0, 0};

    for (int64_t e = d.il_row[ti]; e < d.il_row[ti + 1]; ++e) {
        const int32_t src = d.il_src[e];
        const int off = d.il_off[e];
        const int cls = d.off_class[off];
        const int d2 = d.off_d2[off];
        const double* cs = d.off_cs + static_cast<int64_t>(off) * (P + 1) * 2;
        const double* Ms = M + static_cast<int64_t>(src) * nc * K;
        __syncthreads();
        for (int v = tid; v < nc * K; v += 256) {
            const int c = v / K, r = v % K;
            int n = 0;
            while ((n + 1) * (n + 1) <= r) ++n;
            const int qq = r - n * n;
            const int m = (qq + 1) / 2;
            double val;
            if (m == 0) {
                val = Ms[c * K + r];
            } else {
                const double xr = Ms[c * K + rfc(n, m)], xi = Ms[c * K + rfs(n, m)];
                const double cm = cs[2 * m], sm = cs[2 * m + 1];
                val = (qq & 1) ? (xr * cm - xi * sm) : (xr * sm + xi * cm);
            }
            X[c][r] = val;
        }
        __syncthreads();
        const double* fwd = d.rot + static_cast<int64_t>(cls) * 2 * d.rot_size;
        for (int r = tid; r < K; r += 256) {
            int n = 0;
            while ((n + 1) * (n + 1) <= r) ++n;
            const int qq = r - n * n;
            const int a = (qq + 1) / 2;
            double s[4] = {0, 0, 0, 0};
            if (qq == 0 || (qq & 1)) {
                const double* row = fwd + d.rot_off[n] + a * (n + 1);
                for (int bb = 0; bb <= n; ++bb) {
                    const double t = row[bb];
                    for (int c = 0; c < nc; ++c) s[c] += t * X[c][rfc(n, bb)];
                }
            } else {
                const double* row = fwd + d.rot_off[n] + (n + 1) * (n + 1) + (a - 1) * n;
                for (int bb = 1; bb <= n; ++bb) {
                    const double t = row[bb - 1];
                    for (int c = 0; c < nc; ++c) s[c] += t * X[c][rfs(n, bb)];
                }
            }
            const int o = d.om[r];
            for (int c = 0; c < nc; ++c) Y[c][o] = s[c];
        }
        __syncthreads();
        const double* C = d.coax + static_cast<int64_t>(d2) * d.coax_size;
        for (int qrow = tid; qrow < K; qrow += 256) {
            const int pk = d.om_rows[qrow];
            const int k = pk & 0xff, j = (pk >> 8) & 0xff, partv = pk >> 16;
            const int w = P + 1 - k;
            const int base = (k == 0 ? 0 : (P + 1) + 2 * ((k - 1) * (P + 1) - (k - 1) * k / 2)) + partv * w;
            const double* row = C + d.coax_off[k] + (j - k) * w;
            double s[4] = {0, 0, 0, 0};
            for (int nn = 0; nn < w; ++nn) {
                const double t = row[nn];
                for (int c = 0; c < nc; ++c) s[c] += t * Y[c][base + nn];
            }
            const int r = partv ? rfs(j, k) : rfc(j, k);
            for (int c = 0; c < nc; ++c) X[c][r] = s[c];
        }
        __syncthreads();
        if (tid < NT) {
            const double* bwd = fwd + d.rot_size;
            const double* rowc = bwd + d.rot_off[jj] + aa * (jj + 1);
            double lc[4] = {0, 0, 0, 0}, ls[4] = {0, 0, 0, 0};
            for (int bb = 0; bb <= jj; ++bb) {
                const double t = rowc[bb];
                for (int c = 0; c < nc; ++c) lc[c] += t * X[c][rfc(jj, bb)];
            }
            if (aa >= 1) {
                const double* rows = bwd + d.rot_off[jj] + (jj + 1) * (jj + 1) + (aa - 1) * jj;
                for (int bb = 1; bb <= jj; ++bb) {
                    const double t = rows[bb - 1];
                    for (int c = 0; c < nc; ++c) ls[c] += t * X[c][rfs(jj, bb)];
                }
            }
            const double cm = cs[2 * aa], sm = cs[2 * aa + 1];
            for (int c = 0; c < nc; ++c) {
                accr[c] += lc[c] * cm + ls[c] * sm;
                acci[c] += ls[c] * cm - lc[c] * sm;
            }
        }
    }
    if (tid < NT) {
        double* Lt = L + static_cast<int64_t>(tg) * nc * K;
        for (int c = 0; c < nc; ++c) {
            Lt[c * K + rfc(jj, aa)] = accr[c];
            if (aa > 0) Lt[c * K + rfs(jj, aa)] = acci[c];
        }
    }
}

// ------------------------------------------------------------------ P2P
// One wavefront per target leaf.  Lanes are (target, split): a leaf with fewer
// than 64 particles spreads each target over 64/npow2(count) lanes that take
// every S-th source; sources are staged through LDS in tiles of 64.
// Near field of the reference N2N (fmm.hpp:773-809): every particle of the
// near list except the target itself.
template <int MODE, int NCQ>
__device__ __forceinline__ void pair_kernel(double dx, double dy, double dz, const double* qj, double* acc)
{
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double rinv = rsqrt(r2);
    if constexpr (MODE == MODE_POT) {
#pragma unroll
        for (int c = 0; c < NCQ; ++c) acc[c] += qj[c] * rinv;
    } else if constexpr (MODE == MODE_CURL) {
        const double r3 = rinv * rinv * rinv;
        acc[0] += (qj[1] * dz - qj[2] * dy) * r3;
        acc[1] += (qj[2] * dx - qj[0] * dz) * r3;
        acc[2] += (qj[0] * dy - qj[1] * dx) * r3;
    } else {
        const double qr = qj[0] * rinv;
        const double qr3 = qr * rinv * rinv;
        acc[0] += qr;
        acc[1] -= qr3 * dx;
        acc[2] -= qr3 * dy;
        acc[3] -= qr3 * dz;
    }
}

// 1/sqrt(r2) in FP64 from the hardware estimate (v_rsq_f64, relative error
// below 6e-8 on gfx950) and one Newton step y <- y + y e / 2, e = 1 - r2 y^2,
// which squares it: below 5e-15 (measured 4.3e-15 over 2^22 values spanning
// 2^-58..2^58), a few ulp of the correctly rounded value.
__device__ __forceinline__ double rsqrt_nr(double r2)
{
    const double y = __builtin_amdgcn_rsq(r2);
    const double e = __builtin_fma(-r2 * y, y, 1.0);
    return __builtin_fma(0.5 * y, e, y);
}

// pair_kernel with rsqrt_nr (the near field's FP64 pairs).
template <int MODE, int NCQ>
__device__ __forceinline__ void pair_kernel_nr(double dx, double dy, double dz, const double* qj, double* acc)
{
    const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
    const double rinv = rsqrt_nr(r2);
    if constexpr (MODE == MODE_POT) {
#pragma unroll
        for (int c = 0; c < NCQ; ++c) acc[c] = __builtin_fma(qj[c], rinv, acc[c]);
    } else if constexpr (MODE == MODE_CURL) {
        const double r3 = rinv * rinv * rinv;
        acc[0] += (qj[1] * dz - qj[2] * dy) * r3;
        acc[1] += (qj[2] * dx - qj[0] * dz) * r3;
        acc[2] += (qj[0] * dy - qj[1] * dx) * r3;
    } else {
        const double qr = qj[0] * rinv;
        const double qr3 = qr * rinv * rinv;
        acc[0] += qr;
        acc[1] = __builtin_fma(-qr3, dx, acc[1]);
        acc[2] = __builtin_fma(-qr3, dy, acc[2]);
        acc[3] = __builtin_fma(-qr3, dz, acc[3]);
    }
}

// Mixed-precision near field: the pair's coordinate differences arrive in
// FP32 (taken in FP64), the pair arithmetic is FP32 (v_rsq_f32), and each
// target's sum over a source tile of 64 is added to an FP64 accumulator.
template <int MODE, int NCQ>
__device__ __forceinline__ void pair_kernel_f32(float dx, float dy, float dz, const float* qj, float* acc)
{
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float rinv = rsqrtf(r2);
    if constexpr (MODE == MODE_POT) {
#pragma unroll
        for (int c = 0; c < NCQ; ++c) acc[c] += qj[c] * rinv;
    } else if constexpr (MODE == MODE_CURL) {
        const float r3 = rinv * rinv * rinv;
        acc[0] += (qj[1] * dz - qj[2] * dy) * r3;
        acc[1] += (qj[2] * dx - qj[0] * dz) * r3;
        acc[2] += (qj[0] * dy - qj[1] * dx) * r3;
    } else {
        const float qr = qj[0] * rinv;
        const float qr3 = qr * rinv * rinv;
        acc[0] += qr;
        acc[1] -= qr3 * dx;
        acc[2] -= qr3 * dy;
        acc[3] -= qr3 * dz;
    }
}

// Near field (N2N, fmm.hpp:773-809), one wavefront per work item.  An item is
// a target leaf and a run of at most a few thousand sources of its near list
// (the concatenation of its merged source ranges): leaves whose near list is
// long (coarse leaves next to deep subtrees) are split over several items so
// that no wavefront carries a long tail.  Each lane holds T targets (T = 1, 2
// or 4, the tile shape: every source read from LDS serves T independent
// pairs); the U = ceil(nt / T) target sets of a leaf are spread over
// S = floor(64 / U) lanes each, which take every S-th source.  Sources are
// staged through LDS 64 at a time, fetched one tile ahead; only tiles that
// hold particles of the target leaf itself test for the pair (i, i) (the
// reference's default is_self, fmm.hpp:875-881).  An unsplit item adds into
// the output; the items of a split leaf write partial sums that
// p2p_reduce_kernel adds in a fixed order, so results are bitwise repeatable
// for a given T.
// MIXED: coordinate differences in FP64 rounded to FP32 (so close pairs keep
// their relative precision), FP32 pair arithmetic, FP32 sums per source tile
// added into FP64 accumulators.
template <int MODE, int NCQ, int NOUT, bool MIXED, int T>
__global__ __launch_bounds__(64) void p2p_kernel(DevPlan d, P2PItems w, const double* __restrict__ q,
                                                 double* __restrict__ out, double* __restrict__ partial)
{
    using R = typename std::conditional<MIXED, float, double>::type;
    __shared__ double sx[64], sy[64], sz[64];
    __shared__ R sq[NCQ][64];
    __shared__ int32_t sidx[64];
    __shared__ double red[T][NOUT][64];
    const int64_t it = blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t g = w.leaf[it];
    const int64_t u0 = w.src_begin[it], u1 = u0 + w.src_count[it];
    const int64_t pslot = w.partial[it];
    const int64_t tb = d.lf_begin[g], te = d.lf_end[g];
    const int64_t nt = te - tb;
    const int64_t r0 = d.near_row[g], r1 = d.near_row[g + 1];
    const int ntp = nt < 64 * T ? static_cast<int>(nt) : 64 * T;  // targets per pass
    const int U = (ntp + T - 1) / T;                              // lanes' target sets per pass
    const int S = 64 / U;
    const int t = tid % U, s = tid / U;
    const bool lane_on = s < S;
    for (int64_t t0 = tb; t0 < te; t0 += ntp) {
        // this lane's targets: t0 + t + U h, h < T
        int64_t slot[T];
        bool val[T];
        double xt[T], yt[T], zt[T];
        int32_t it32[T];
#pragma unroll
        for (int h = 0; h < T; ++h) {
            slot[h] = t0 + t + static_cast<int64_t>(U) * h;
            val[h] = lane_on && slot[h] < te;
            xt[h] = yt[h] = zt[h] = 0.0;
            if (val[h]) {
                xt[h] = d.px[slot[h]];
                yt[h] = d.py[slot[h]];
                zt[h] = d.pz[slot[h]];
            }
            // waited for here, not inside the pair loop, where in-order load
            // counting would also wait for the next tile's prefetch
            asm volatile("" ::"v"(xt[h]), "v"(yt[h]), "v"(zt[h]));
            it32[h] = static_cast<int32_t>(slot[h]);
        }
        double acc[T][NOUT];
#pragma unroll
        for (int h = 0; h < T; ++h)
#pragma unroll
            for (int c = 0; c < NOUT; ++c) acc[h][c] = 0.0;
        int64_t rc = r0;  // range holding source u of this lane, advanced monotonically
        // this lane's source of a tile, loaded one tile ahead of its use
        double fx = 0, fy = 0, fz = 0;
        R fq[NCQ];
        int64_t fj = -1;
        auto fetch = [&](int64_t ut) {
            const int64_t u = ut + tid;
            fj = -1;
            if (u < u1) {
                while (rc + 1 < r1 && d.near_pref[rc + 1] <= u) ++rc;
                fj = d.near_begin[rc] + (u - d.near_pref[rc]);
                fx = d.px[fj];
                fy = d.py[fj];
                fz = d.pz[fj];
#pragma unroll
                for (int c = 0; c < NCQ; ++c) fq[c] = static_cast<R>(q[c * d.n + fj]);
            }
        };
        if (u0 < u1) fetch(u0);
        for (int64_t ut = u0; ut < u1; ut += 64) {
            __syncthreads();
            const bool mine = fj >= tb && fj < te;
            if (fj >= 0) {
                sidx[tid] = static_cast<int32_t>(fj);
                sx[tid] = fx;
                sy[tid] = fy;
                sz[tid] = fz;
#pragma unroll
                for (int c = 0; c < NCQ; ++c) sq[c][tid] = fq[c];
            }
            const bool own = __any(mine);  // the tile holds particles of the target leaf
            __syncthreads();
            if (ut + 64 < u1) fetch(ut + 64);
            const int cnt = static_cast<int>(u1 - ut < 64 ? u1 - ut : 64);
            if (val[0]) {
                R a[T][NOUT];
#pragma unroll
                for (int h = 0; h < T; ++h)
#pragma unroll
                    for (int c = 0; c < NOUT; ++c) a[h][c] = 0;
                auto pairs = [&](auto checked) {
                    for (int k = s; k < cnt; k += S) {
                        R qs[NCQ];
#pragma unroll
                        for (int c = 0; c < NCQ; ++c) qs[c] = sq[c][k];
                        const double xj = sx[k], yj = sy[k], zj = sz[k];
                        const int32_t sj = decltype(checked)::value ? sidx[k] : -1;
#pragma unroll
                        for (int h = 0; h < T; ++h) {
                            R qh[NCQ];
#pragma unroll
                            for (int c = 0; c < NCQ; ++c) qh[c] = qs[c];
                            double dx = xt[h] - xj;
                            const double dy = yt[h] - yj, dz = zt[h] - zj;
                            if constexpr (decltype(checked)::value) {
                                // the pair (i, i) contributes nothing: unit distance, zero charge
                                if (sj == it32[h]) {
                                    dx = 1.0;
#pragma unroll
                                    for (int c = 0; c < NCQ; ++c) qh[c] = 0;
                                }
                            }
                            if constexpr (MIXED)
                                pair_kernel_f32<MODE, NCQ>(static_cast<float>(dx), static_cast<float>(dy),
                                                           static_cast<float>(dz), qh, a[h]);
                            else
                                pair_kernel_nr<MODE, NCQ>(dx, dy, dz, qh, a[h]);
                        }
                    }
                };
                if (own) pairs(std::true_type{});
                else pairs(std::false_type{});
#pragma unroll
                for (int h = 0; h < T; ++h)
#pragma unroll
                    for (int c = 0; c < NOUT; ++c) acc[h][c] += static_cast<double>(a[h][c]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < T; ++h)
#pragma unroll
            for (int c = 0; c < NOUT; ++c) red[h][c][tid] = acc[h][c];
        __syncthreads();
        if (s == 0) {
#pragma unroll
            for (int h = 0; h < T; ++h) {
                if (slot[h] >= te) continue;
#pragma unroll
                for (int c = 0; c < NOUT; ++c) {
                    double v = 0.0;
                    for (int ss = 0; ss < S; ++ss) v += red[h][c][t + ss * U];
                    if (pslot < 0) out[slot[h] * NOUT + c] += v;
                    else partial[(pslot + (slot[h] - tb)) * NOUT + c] = v;
                }
            }
        }
    }
}

// Self exclusions (fmm.hpp:786-807): the near field above skips only j == i,
// so a listed source is subtracted when the reference would not have added it:
// far sources always (the far field included them, _slist), near sources when
// the exclusion predicate is the list itself.  FP64.
template <int MODE, int NCQ, int NOUT>
__global__ __launch_bounds__(64) void self_correction_kernel(DevPlan d, SelfCorr c, int use_near,
                                                             const double* __restrict__ q, double* __restrict__ out)
{
    const int64_t r = static_cast<int64_t>(blockIdx.x) * 64 + threadIdx.x;
    if (r >= c.n_row) return;
    const int64_t t = c.tgt[r];
    const double xi = d.px[t], yi = d.py[t], zi = d.pz[t];
    double a[NOUT];
#pragma unroll
    for (int k = 0; k < NOUT; ++k) a[k] = 0.0;
    for (int64_t k = c.row[r]; k < c.row[r + 1]; ++k) {
        if (c.near[k] && !use_near) continue;
        const int64_t j = c.src[k];
        double qj[NCQ];
#pragma unroll
        for (int cc = 0; cc < NCQ; ++cc) qj[cc] = q[cc * d.n + j];
        pair_kernel<MODE, NCQ>(xi - d.px[j], yi - d.py[j], zi - d.pz[j], qj, a);
    }
#pragma unroll
    for (int k = 0; k < NOUT; ++k) out[t * NOUT + k] -= a[k];
}

// Adds the partial sums of split leaves in item order.
__global__ void p2p_reduce_kernel(DevPlan d, P2PItems w, int nout, const double* __restrict__ partial,
                                  double* __restrict__ out)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= w.n_split_targets) return;
    const int64_t slot = w.split_slot[k];
    const int64_t first = w.split_first[k], nitem = w.split_nitem[k], stride = w.split_stride[k];
    for (int c = 0; c < nout; ++c) {
        double v = 0.0;
        for (int64_t i = 0; i < nitem; ++i) v += partial[(first + i * stride) * nout + c];
        out[slot * nout + c] += v;
    }
}

// ------------------------------------------------------------------ direct sum
// FP64 O(N * ntarget) sum over all sources j != target (main.cpp:70-96), the
// error reference.  Block (target chunk of 64, source split); partial sums are
// reduced in a fixed order by direct_reduce_kernel.  Charges are read in slot
// order, q[c * n + slot], or with qperm through the permutation from the
// caller's original-order array, q[qperm[slot] * NCQ + c] (the same values in
// the same order, so the same sums).  A target outside [0, n) gives NaN.
template <int MODE, int NCQ, int NOUT>
__global__ __launch_bounds__(64) void direct_kernel(DevPlan d, const double* __restrict__ q,
                                                    const int32_t* __restrict__ qperm, int64_t ntarget,
                                                    const int64_t* __restrict__ targets, int nsplit,
                                                    double* __restrict__ partial)
{
    __shared__ double sx[64], sy[64], sz[64], sq[NCQ][64];
    const int tid = threadIdx.x;
    const int64_t k = static_cast<int64_t>(blockIdx.x) * 64 + tid;
    const int split = blockIdx.y;
    const bool valid = k < ntarget;
    int64_t slot = -1;
    bool bad = false;
    double xi = 0, yi = 0, zi = 0;
    if (valid) {
        const int64_t t = targets[k];
        bad = t < 0 || t >= d.n;
        if (!bad) {
            slot = d.iperm[t];
            xi = d.px[slot];
            yi = d.py[slot];
            zi = d.pz[slot];
        }
    }
    const int64_t j0s = d.n * split / nsplit, j1s = d.n * (split + 1) / nsplit;
    double acc[NOUT];
#pragma unroll
    for (int c = 0; c < NOUT; ++c) acc[c] = 0.0;
    for (int64_t j0 = j0s; j0 < j1s; j0 += 64) {
        __syncthreads();
        const int64_t jl = j0 + tid;
        if (jl < j1s) {
            sx[tid] = d.px[jl];
            sy[tid] = d.py[jl];
            sz[tid] = d.pz[jl];
            if (qperm) {
                const int64_t o = static_cast<int64_t>(qperm[jl]) * NCQ;
#pragma unroll
                for (int c = 0; c < NCQ; ++c) sq[c][tid] = q[o + c];
            } else {
#pragma unroll
                for (int c = 0; c < NCQ; ++c) sq[c][tid] = q[c * d.n + jl];
            }
        }
        __syncthreads();
        const int cnt = static_cast<int>(j1s - j0 < 64 ? j1s - j0 : 64);
        if (valid && !bad) {
            for (int kk = 0; kk < cnt; ++kk) {
                if (j0 + kk == slot) continue;
                double qj[NCQ];
#pragma unroll
                for (int c = 0; c < NCQ; ++c) qj[c] = sq[c][kk];
                pair_kernel<MODE, NCQ>(xi - sx[kk], yi - sy[kk], zi - sz[kk], qj, acc);
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int c = 0; c < NOUT; ++c)
            partial[(static_cast<int64_t>(split) * NOUT + c) * ntarget + k] = bad ? __builtin_nan("") : acc[c];
    }
}

__global__ void direct_reduce_kernel(int64_t ntarget, int nout, int nsplit, const double* __restrict__ partial,
                                     double* __restrict__ out)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= ntarget) return;
    for (int c = 0; c < nout; ++c) {
        double v = 0.0;
        for (int s = 0; s < nsplit; ++s) v += partial[(static_cast<int64_t>(s) * nout + c) * ntarget + k];
        out[k * nout + c] = v;
    }
}

// ------------------------------------------------------------------ dispatch on p
template <template <int> class F, typename... A>
void with_p(int p, A&&... args)
{
    switch (p) {
#define SFMM_P(X) \
    case X: F<X>::run(args...); break;
        SFMM_P(1) SFMM_P(2) SFMM_P(3) SFMM_P(4) SFMM_P(5) SFMM_P(6) SFMM_P(7) SFMM_P(8) SFMM_P(9) SFMM_P(10)
        SFMM_P(11) SFMM_P(12) SFMM_P(13) SFMM_P(14) SFMM_P(15) SFMM_P(16)
#undef SFMM_P
    default: break;
    }
}

template <int P>
struct P2MRun {
    static void run(const DevPlan& d, int mode, int ncq, int nc, const double* q, double* M,
                    hipStream_t s)
    {
        const dim3 grid(static_cast<unsigned>(d.n_far_leaf)), blk(64);
        switch (ncq) {
        case 1: hipLaunchKernelGGL((p2m_kernel<P, 1>), grid, blk, 0, s, d, mode, nc, q, M); break;
        case 2: hipLaunchKernelGGL((p2m_kernel<P, 2>), grid, blk, 0, s, d, mode, nc, q, M); break;
        case 3: hipLaunchKernelGGL((p2m_kernel<P, 3>), grid, blk, 0, s, d, mode, nc, q, M); break;
        default: hipLaunchKernelGGL((p2m_kernel<P, 4>), grid, blk, 0, s, d, mode, nc, q, M); break;
        }
    }
};
template <int P>
struct L2PRun {
    static void run(const DevPlan& d, int nc, const double* L, double* out, hipStream_t s)
    {
        const dim3 grid(grid_for(d.n, 256)), blk(256);
        switch (nc) {
        case 1: hipLaunchKernelGGL((l2p_kernel<P, 1>), grid, blk, 0, s, d, L, out); break;
        case 2: hipLaunchKernelGGL((l2p_kernel<P, 2>), grid, blk, 0, s, d, L, out); break;
        case 3: hipLaunchKernelGGL((l2p_kernel<P, 3>), grid, blk, 0, s, d, L, out); break;
        default: hipLaunchKernelGGL((l2p_kernel<P, 4>), grid, blk, 0, s, d, L, out); break;
        }
    }
};
template <int P>
struct M2LRun {
    static void run(const DevPlan& d, int nc, const double* M, double* L, hipStream_t s)
    {
        hipLaunchKernelGGL(m2l_kernel<P>, dim3(static_cast<unsigned>(d.n_target)), dim3(256), 0, s, d, nc, M, L);
    }
};

}  // namespace

void launch_gather_charges(const DevPlan& d, int qdim, const double* Q, double* q, hipStream_t s)
{
    hipLaunchKernelGGL(gather_charges_kernel, dim3(grid_for(d.n, 256)), dim3(256), 0, s, d.n, qdim, d.perm, Q, q);
}

void launch_scatter_outputs(const DevPlan& d, int nout, const double* out, double* U, hipStream_t s)
{
    hipLaunchKernelGGL(scatter_outputs_kernel, dim3(grid_for(d.n, 256)), dim3(256), 0, s, d.n, nout, d.iperm, out, U);
}

void launch_fill(double* x, int64_t count, double v, hipStream_t s)
{
    if (count > 0) hipLaunchKernelGGL(fill_kernel, dim3(grid_for(count, 256)), dim3(256), 0, s, x, count, v);
}

void launch_p2m(const DevPlan& d, Mode mode, int nc, const double* q, double* M, hipStream_t s)
{
    if (d.n_far_leaf == 0) return;
    const int ncq = mode == MODE_POT ? nc : (mode == MODE_CURL ? 3 : 1);
    with_p<P2MRun>(d.p, d, static_cast<int>(mode), ncq, nc, q, M, s);
}

void launch_l2p(const DevPlan& d, int nc, const double* L, double* out, hipStream_t s)
{
    if (d.n_far_leaf == 0) return;
    with_p<L2PRun>(d.p, d, nc, L, out, s);
}

void launch_m2l(const DevPlan& d, int nc, const double* M, double* L, hipStream_t s)
{
    if (d.n_target == 0) return;
    with_p<M2LRun>(d.p, d, nc, M, L, s);
}

void launch_m2m(const DevPlan& d, LevelSpan parents, int nc, double* M, hipStream_t s)
{
    if (parents.count == 0) return;
    hipLaunchKernelGGL(m2m_kernel, dim3(static_cast<unsigned>(parents.count)), dim3(256),
                       sizeof(double) * nc * d.K, s, d, parents.first, nc, M);
}

void launch_l2l(const DevPlan& d, LevelSpan children, int nc, double* L, hipStream_t s)
{
    if (children.count == 0) return;
    hipLaunchKernelGGL(l2l_kernel, dim3(static_cast<unsigned>(children.count)), dim3(256),
                       sizeof(double) * nc * d.K, s, d, children.first, nc, L);
}

template <bool MIXED, int T>
void launch_p2p_t(const DevPlan& d, const P2PItems& w, Mode mode, int nc, const double* q, double* out,
                  double* partial, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(w.n_item)), blk(64);
    int nout = nc;
    if (mode == MODE_POTGRAD) {
        hipLaunchKernelGGL((p2p_kernel<MODE_POTGRAD, 1, 4, MIXED, T>), grid, blk, 0, s, d, w, q, out, partial);
        nout = 4;
    } else if (mode == MODE_CURL) {
        hipLaunchKernelGGL((p2p_kernel<MODE_CURL, 3, 3, MIXED, T>), grid, blk, 0, s, d, w, q, out, partial);
        nout = 3;
    } else {
        switch (nc) {
        case 1: hipLaunchKernelGGL((p2p_kernel<MODE_POT, 1, 1, MIXED, T>), grid, blk, 0, s, d, w, q, out, partial); break;
        case 2: hipLaunchKernelGGL((p2p_kernel<MODE_POT, 2, 2, MIXED, T>), grid, blk, 0, s, d, w, q, out, partial); break;
        case 3: hipLaunchKernelGGL((p2p_kernel<MODE_POT, 3, 3, MIXED, T>), grid, blk, 0, s, d, w, q, out, partial); break;
        default: hipLaunchKernelGGL((p2p_kernel<MODE_POT, 4, 4, MIXED, T>), grid, blk, 0, s, d, w, q, out, partial); break;
        }
    }
    if (w.n_split_targets > 0)
        hipLaunchKernelGGL(p2p_reduce_kernel, dim3(grid_for(w.n_split_targets, 256)), dim3(256), 0, s, d, w, nout,
                           partial, out);
}

void launch_p2p(const DevPlan& d, const P2PItems& w, Mode mode, int nc, const double* q, double* out,
                double* partial, hipStream_t s, bool mixed, int tpl)
{
    if (w.n_item == 0) return;
    switch (tpl) {
    case 1: mixed ? launch_p2p_t<true, 1>(d, w, mode, nc, q, out, partial, s)
                  : launch_p2p_t<false, 1>(d, w, mode, nc, q, out, partial, s); break;
    case 4: mixed ? launch_p2p_t<true, 4>(d, w, mode, nc, q, out, partial, s)
                  : launch_p2p_t<false, 4>(d, w, mode, nc, q, out, partial, s); break;
    default: mixed ? launch_p2p_t<true, 2>(d, w, mode, nc, q, out, partial, s)
                   : launch_p2p_t<false, 2>(d, w, mode, nc, q, out, partial, s); break;
    }
}

void launch_self_correction(const DevPlan& d, const SelfCorr& c, Mode mode, int nc, bool use_near, const double* q,
                            double* out, hipStream_t s)
{
    if (c.n_row == 0) return;
    const dim3 grid(grid_for(c.n_row, 64)), blk(64);
    const int un = use_near ? 1 : 0;
    if (mode == MODE_POTGRAD) {
        hipLaunchKernelGGL((self_correction_kernel<MODE_POTGRAD, 1, 4>), grid, blk, 0, s, d, c, un, q, out);
    } else if (mode == MODE_CURL) {
        hipLaunchKernelGGL((self_correction_kernel<MODE_CURL, 3, 3>), grid, blk, 0, s, d, c, un, q, out);
    } else {
        switch (nc) {
        case 1: hipLaunchKernelGGL((self_correction_kernel<MODE_POT, 1, 1>), grid, blk, 0, s, d, c, un, q, out); break;
        case 2: hipLaunchKernelGGL((self_correction_kernel<MODE_POT, 2, 2>), grid, blk, 0, s, d, c, un, q, out); break;
        case 3: hipLaunchKernelGGL((self_correction_kernel<MODE_POT, 3, 3>), grid, blk, 0, s, d, c, un, q, out); break;
        default: hipLaunchKernelGGL((self_correction_kernel<MODE_POT, 4, 4>), grid, blk, 0, s, d, c, un, q, out); break;
        }
    }
}

int64_t direct_scratch_size(int64_t ntarget, int nout, int nsplit) { return ntarget * nout * nsplit; }

int direct_nsplit(int64_t ntarget, int64_t n)
{
    const int64_t nblk = (ntarget + 63) / 64;
    const int64_t ns = std::max<int64_t>(1, std::min<int64_t>(512, 8192 / std::max<int64_t>(nblk, 1)));
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ns, n / 64 + 1)));
}

void launch_direct(const DevPlan& d, Mode mode, int nc, const double* q, int64_t ntarget, const int64_t* targets,
                   double* out, double* partial, int nsplit, hipStream_t s, const int32_t* qperm)
{
    if (ntarget == 0) return;
    const dim3 grid(grid_for(ntarget, 64), nsplit), blk(64);
    int nout = nc;
    if (mode == MODE_POTGRAD) {
        hipLaunchKernelGGL((direct_kernel<MODE_POTGRAD, 1, 4>), grid, blk, 0, s, d, q, qperm, ntarget, targets, nsplit,
                           partial);
        nout = 4;
    } else if (mode == MODE_CURL) {
        hipLaunchKernelGGL((direct_kernel<MODE_CURL, 3, 3>), grid, blk, 0, s, d, q, qperm, ntarget, targets, nsplit, partial);
        nout = 3;
    } else {
        switch (nc) {
        case 1: hipLaunchKernelGGL((direct_kernel<MODE_POT, 1, 1>), grid, blk, 0, s, d, q, qperm, ntarget, targets, nsplit, partial); break;
        case 2: hipLaunchKernelGGL((direct_kernel<MODE_POT, 2, 2>), grid, blk, 0, s, d, q, qperm, ntarget, targets, nsplit, partial); break;
        case 3: hipLaunchKernelGGL((direct_kernel<MODE_POT, 3, 3>), grid, blk, 0, s, d, q, qperm, ntarget, targets, nsplit, partial); break;
        default: hipLaunchKernelGGL((direct_kernel<MODE_POT, 4, 4>), grid, blk, 0, s, d, q, qperm, ntarget, targets, nsplit, partial); break;
        }
    }
    hipLaunchKernelGGL(direct_reduce_kernel, dim3(grid_for(ntarget, 256)), dim3(256), 0, s, ntarget, nout, nsplit,
                       partial, out);
}

}  // namespace sfmm
